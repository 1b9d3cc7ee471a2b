"""ctypes binding of libadfdtd.so (include/adfdtd.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object has not been built, and every call that needs a GPU
raises when the HIP runtime reports none.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libadfdtd.so")

ADF_F32, ADF_F64, ADF_F32_REF = 1, 2, 3
ADF_REC_P, ADF_REC_VX, ADF_REC_VY, ADF_REC_VZ, ADF_VN_REC = 0, 1, 2, 3, 4
ADF_SRC_P, ADF_SRC_VX, ADF_SRC_VY, ADF_SRC_VZ, ADF_VN_SRC = 5, 6, 7, 8, 9
ADF_GF_NONE, ADF_GF_PV_FROM_PIN, ADF_GF_PV_FROM_PV, ADF_GF_P_FROM_PIN = 0, 1, 2, 3

# Every symbol the public headers declare (checked by the CPU test suite).
EXPORTS = (
    "adf_version", "adf_last_error", "adf_ctx_create", "adf_ctx_destroy", "adf_ctx_synchronize",
    "adf_solver_create", "adf_solver_destroy", "adf_solver_set_field", "adf_solver_get_field",
    "adf_solver_reset", "adf_solver_set_cpml_memory", "adf_solver_get_cpml_memory",
    "adf_solver_set_mask", "adf_solver_build_mask", "adf_solver_get_mask",
    "adf_solver_set_transceivers", "adf_solver_get_trace", "adf_solver_set_trace",
    "adf_solver_set_gf_map", "adf_solver_run", "adf_solver_get_pressure_slice",
    "adf_solver_last_step_ms", "adf_solver_set_option", "adf_solver_sweep_stats", "adf_solver_next_step",
    "adf_batch_create", "adf_batch_destroy", "adf_batch_add", "adf_batch_remove", "adf_batch_size",
    "adf_batch_run", "adf_batch_set_option", "adf_batch_last_step_ms",
    "adf_server_create", "adf_server_destroy", "adf_server_set_option", "adf_server_submit", "adf_server_poll",
    "adf_server_wait", "adf_server_get_trace", "adf_server_get_field", "adf_server_request_error",
    "adf_server_release", "adf_server_stats", "adf_kh_analytic_accumulate", "adf_kh_last_ms",
    "adf_solver_gf_resample", "adf_resample_f32", "adf_resample_f32_noalloc", "adf_bandwidth_probe",
)
ADF_REQ_QUEUED, ADF_REQ_RUNNING, ADF_REQ_DONE, ADF_REQ_FAILED = 0, 1, 2, -1
ADF_SERVER_CHUNK, ADF_SERVER_BUDGET = 1, 2
ADF_OPT_PATH, ADF_OPT_TILE_Y, ADF_OPT_BLOCK_Z, ADF_OPT_TIME_SWEEP, ADF_OPT_WAVES, ADF_OPT_BLOCK_Z_GEN = 1, 2, 3, 4, 5, 6
ADF_OPT_PIPELINE, ADF_OPT_SWEEP_LAUNCH, ADF_OPT_STORE_NT, ADF_OPT_TX_LAUNCH, ADF_OPT_CELLS_PER_LANE = 7, 8, 9, 10, 11
ADF_SWEEP_ONE_LAUNCH, ADF_SWEEP_TWO_STREAMS, ADF_SWEEP_TWO_LAUNCHES = 0, 1, 2
ADF_BATCH_PERSISTENT, ADF_BATCH_GRAPH_STEPS = 1, 2
ADF_BW_COPY_BUF, ADF_BW_COPY_GLOBAL, ADF_BW_READ, ADF_BW_WRITE, ADF_BW_COPY_NT = 0, 1, 2, 3, 4
ADF_BW_ZMARCH, ADF_BW_ZMARCH_NT = 5, 6


class AdfDomain(C.Structure):
    """include/adfdtd.h adf_domain == reference Domain (src/simulation.jl:25-44)."""

    _fields_ = [
        ("tmax", C.c_double), ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("c0", C.c_double), ("r0", C.c_double), ("cf", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("fact_m0", C.c_double), ("fact_m1", C.c_double), ("dt", C.c_double),
        ("nt", C.c_int64), ("z0", C.c_double),
    ]


class AdfCpml(C.Structure):
    _fields_ = [("npml", C.c_int64)] + [(n, C.c_void_p) for n in
                                        ("a_l", "a_r", "a_hl", "a_hr", "b_l", "b_r", "b_hl", "b_hr")]


class AdfRequest(C.Structure):
    """include/adfdtd.h adf_request."""

    _fields_ = [
        ("dom", AdfDomain), ("dtype", C.c_int), ("cpml", AdfCpml), ("mask_shape", C.c_int),
        ("mask_params", C.c_double * 4), ("xc", C.c_void_p), ("yc", C.c_void_p), ("zc", C.c_void_p),
        ("mask4", C.c_void_p), ("ntx", C.c_int64), ("idx", C.c_void_p), ("w", C.c_void_p), ("src", C.c_void_p),
        ("dir", C.c_void_p), ("vn_src", C.c_void_p), ("field", C.c_void_p), ("start_step", C.c_int64),
        ("end_step", C.c_int64), ("want", C.c_uint32), ("keep_field", C.c_int),
    ]


class AdfServerInfo(C.Structure):
    """include/adfdtd.h adf_server_info."""

    _fields_ = [
        ("completed", C.c_int64), ("failed", C.c_int64), ("running", C.c_int64), ("queued", C.c_int64),
        ("bytes_in_use", C.c_uint64), ("peak_bytes", C.c_uint64), ("budget_bytes", C.c_uint64),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("solves_per_s", C.c_double),
        ("busy_s", C.c_double), ("glups", C.c_double),
    ]


class AdfError(RuntimeError):
    pass


_lock = threading.Lock()
_lib = None


def load():
    """Load libadfdtd.so; raise if it is missing (no fallback exists)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise AdfError(f"{LIB_PATH} is not built; run `make -C {PKG_DIR}` or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        p, i64, i32, d = C.c_void_p, C.c_int64, C.c_int, C.c_double
        sig = {
            "adf_version": ([], i32),
            "adf_last_error": ([], C.c_char_p),
            "adf_ctx_create": ([i32, C.POINTER(p)], i32),
            "adf_ctx_destroy": ([p], i32),
            "adf_ctx_synchronize": ([p], i32),
            "adf_solver_create": ([p, C.POINTER(AdfDomain), i32, C.POINTER(AdfCpml), C.POINTER(p)], i32),
            "adf_solver_destroy": ([p], i32),
            "adf_solver_set_field": ([p, p], i32),
            "adf_solver_get_field": ([p, p], i32),
            "adf_solver_reset": ([p], i32),
            "adf_solver_set_cpml_memory": ([p, C.POINTER(p), C.POINTER(p)], i32),
            "adf_solver_get_cpml_memory": ([p, C.POINTER(p), C.POINTER(p)], i32),
            "adf_solver_set_mask": ([p, p], i32),
            "adf_solver_build_mask": ([p, i32, p, p, p, p], i32),
            "adf_solver_get_mask": ([p, p], i32),
            "adf_solver_set_transceivers": ([p, i64, p, p, p, p, p], i32),
            "adf_solver_get_trace": ([p, i32, p], i32),
            "adf_solver_set_trace": ([p, i32, p], i32),
            "adf_solver_set_gf_map": ([p, i32, C.POINTER(p), i64, i64, i64, p, i64, p, i64, i32], i32),
            "adf_solver_run": ([p, i64, i64], i32),
            "adf_solver_get_pressure_slice": ([p, i32, i64, p], i32),
            "adf_solver_last_step_ms": ([p, C.POINTER(d)], i32),
            "adf_solver_set_option": ([p, i32, i64], i32),
            "adf_solver_sweep_stats": ([p, C.POINTER(d), C.POINTER(i64)], i32),
            "adf_solver_next_step": ([p, C.POINTER(i64)], i32),
            "adf_batch_create": ([p, C.POINTER(p)], i32),
            "adf_batch_destroy": ([p], i32),
            "adf_batch_add": ([p, p], i32),
            "adf_batch_remove": ([p, p], i32),
            "adf_batch_size": ([p, C.POINTER(i64)], i32),
            "adf_batch_run": ([p, i64], i32),
            "adf_batch_set_option": ([p, i32, i64], i32),
            "adf_batch_last_step_ms": ([p, C.POINTER(d)], i32),
            "adf_server_create": ([i32, C.c_uint64, C.POINTER(p)], i32),
            "adf_server_destroy": ([p], i32),
            "adf_server_set_option": ([p, i32, i64], i32),
            "adf_server_submit": ([p, C.POINTER(AdfRequest), C.POINTER(i64)], i32),
            "adf_server_poll": ([p, i64, C.POINTER(i32)], i32),
            "adf_server_wait": ([p, i64, d, C.POINTER(i32)], i32),
            "adf_server_get_trace": ([p, i64, i32, p], i32),
            "adf_server_get_field": ([p, i64, p], i32),
            "adf_server_request_error": ([p, i64], C.c_char_p),
            "adf_server_release": ([p, i64], i32),
            "adf_server_stats": ([p, C.POINTER(AdfServerInfo)], i32),
            "adf_kh_analytic_accumulate": ([p, i32, i32, i64, p, p, i64, p, p, p, i64, d, d, d, i32, d, i64, p, p],
                                           i32),
            "adf_kh_last_ms": ([C.POINTER(d)], i32),
            "adf_solver_gf_resample": ([p, i64, i64, d, i64, p, p], i32),
            "adf_resample_f32": ([p, p, i64, i64, d, d, d, d, i64, d, i32, p], i32),
            "adf_resample_f32_noalloc": ([p, p, i64, i64, d, d, d, d, i64, d, i32, p], i32),
            "adf_bandwidth_probe": ([p, i32, i64, i32, p], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
        return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().adf_last_error().decode(errors="replace")
        raise AdfError(f"{what}: {msg} (code {rc})")


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)
