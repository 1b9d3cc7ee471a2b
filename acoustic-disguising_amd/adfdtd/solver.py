"""Device-resident solver: one `adf_solver` of libadfdtd.so.

This is the persistent form behind `run_fdtd`: field, CPML memory, mask and
packed transceivers stay in HBM between calls; host arrays are copied in on
set_* and out on get_*.
"""
from __future__ import annotations

import ctypes as C
import threading
from typing import Optional

import numpy as np

from . import _lib as L
from .simulation import Cpml, Domain, TransceiverGrid, coords

_ctx_lock = threading.Lock()
_ctxs: dict = {}


def context(device: int = 0) -> C.c_void_p:
    """One adf_ctx (device + HIP stream) per device, created on first use."""
    with _ctx_lock:
        if device not in _ctxs:
            lib = L.load()
            h = C.c_void_p()
            L.check(lib.adf_ctx_create(device, C.byref(h)), "adf_ctx_create")
            _ctxs[device] = h
        return _ctxs[device]


def dtype_code(dtype, precision: str = "fast") -> int:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return L.ADF_F64
    if dt == np.float32:
        if precision == "ref":
            return L.ADF_F32_REF
        if precision == "fast":
            return L.ADF_F32
    raise ValueError(f"unsupported dtype/precision {dt}/{precision}")


def np_dtype(code: int):
    return np.float64 if code == L.ADF_F64 else np.float32


def to_c_domain(d: Domain) -> L.AdfDomain:
    return L.AdfDomain(d.tmax, d.xmax, d.ymax, d.zmax, d.nx, d.ny, d.nz, d.c0, d.r0, d.cf,
                       d.dx, d.dy, d.dz, d.fact_m0, d.fact_m1, d.dt, d.nt, d.z0)


SHAPES = {"none": 0, "sphere": 1, "cube": 2, "cross": 3, "ellipsoid": 4}


class Solver:
    """Wraps adf_solver_* (include/adfdtd.h)."""

    _cache: dict = {}
    _cache_lock = threading.Lock()

    def __init__(self, dom: Domain, code: int, cpml: Optional[Cpml] = None, device: int = 0):
        self.lib = L.load()
        self.dom = dom
        self.code = code
        self.T = np_dtype(code)
        self.device = device
        self.npml = cpml.n if cpml is not None else 0
        keep = []
        cp = L.AdfCpml()
        cp.npml = self.npml
        if cpml is not None:
            for k in ("a_l", "a_r", "a_hl", "a_hr", "b_l", "b_r", "b_hl", "b_hr"):
                arr = np.ascontiguousarray(cpml.c[k], np.float64)
                keep.append(arr)
                setattr(cp, k, arr.ctypes.data if arr.size else None)
        self._cdom = to_c_domain(dom)
        h = C.c_void_p()
        L.check(self.lib.adf_solver_create(context(device), C.byref(self._cdom), code, C.byref(cp), C.byref(h)),
                "adf_solver_create")
        self.h = h
        self.ntx = 0
        self._cpml = cpml
        self._masked = False

    @classmethod
    def cached(cls, dom: Domain, code: int, cpml: Optional[Cpml], device: int = 0) -> "Solver":
        """Reuse device allocations across run_fdtd calls of one sweep."""
        key = (dom, code, None if cpml is None else (cpml.n, tuple(np.concatenate(
            [cpml.c[k] for k in sorted(cpml.c)]).tolist())), device)
        with cls._cache_lock:
            s = cls._cache.get(key)
            if s is None:
                if len(cls._cache) >= 4:
                    cls._cache.clear()
                s = cls(dom, code, cpml, device)
                cls._cache[key] = s
            return s

    def __del__(self):
        try:
            if getattr(self, "h", None) and self.h.value:
                self.lib.adf_solver_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # -- state ---------------------------------------------------------
    def _arr(self, a, shape, dtype=None):
        dt = self.T if dtype is None else dtype
        a = np.ascontiguousarray(a, dt)
        if a.shape != tuple(shape):
            raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
        return a

    def set_field(self, field) -> None:
        d = self.dom
        f = self._arr(field, (4, d.nz, d.ny, d.nx))
        L.check(self.lib.adf_solver_set_field(self.h, L.ptr(f)), "set_field")

    def get_field(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        d = self.dom
        buf = np.empty((4, d.nz, d.ny, d.nx), self.T)
        L.check(self.lib.adf_solver_get_field(self.h, L.ptr(buf)), "get_field")
        if out is not None:
            out[...] = buf
            return out
        return buf

    def reset(self) -> None:
        L.check(self.lib.adf_solver_reset(self.h), "reset")

    def set_cpml_memory(self, cpml: Cpml) -> None:
        if cpml is None or cpml.n == 0:
            return
        ps = [np.ascontiguousarray(a, self.T) for a in cpml.p]
        vs = [np.ascontiguousarray(a, self.T) for a in cpml.v]
        PA = C.c_void_p * 3
        L.check(self.lib.adf_solver_set_cpml_memory(self.h, PA(*[a.ctypes.data for a in ps]),
                                                    PA(*[a.ctypes.data for a in vs])), "set_cpml_memory")

    def get_cpml_memory(self, cpml: Cpml) -> None:
        if cpml is None or cpml.n == 0:
            return
        ps = [np.empty(a.shape, self.T) for a in cpml.p]
        vs = [np.empty(a.shape, self.T) for a in cpml.v]
        PA = C.c_void_p * 3
        L.check(self.lib.adf_solver_get_cpml_memory(self.h, PA(*[a.ctypes.data for a in ps]),
                                                    PA(*[a.ctypes.data for a in vs])), "get_cpml_memory")
        for dst, src in zip(cpml.p + cpml.v, ps + vs):
            dst[...] = src

    def set_mask(self, va) -> None:
        if va is None:
            L.check(self.lib.adf_solver_set_mask(self.h, None), "set_mask")
            self._masked = False
            return
        d = self.dom
        m = self._arr(va, (4, d.nz, d.ny, d.nx), np.uint8)
        L.check(self.lib.adf_solver_set_mask(self.h, L.ptr(m)), "set_mask")
        self._masked = True

    def build_mask(self, shape: str, params) -> None:
        """Scatterer mask built on the device (src/geometry.jl:86-280)."""
        cs = coords(self.dom)
        prm = np.zeros(4)
        pv = np.atleast_1d(np.asarray(params, np.float64))
        prm[: len(pv)] = pv
        xs = [np.ascontiguousarray(cs[a]["p"]) for a in ("x", "y", "z")]
        L.check(self.lib.adf_solver_build_mask(self.h, SHAPES[shape], L.ptr(prm), *[L.ptr(v) for v in xs]),
                "build_mask")
        self._masked = shape != "none"

    def get_mask(self) -> np.ndarray:
        d = self.dom
        m = np.empty((4, d.nz, d.ny, d.nx), np.uint8)
        L.check(self.lib.adf_solver_get_mask(self.h, L.ptr(m)), "get_mask")
        return m

    def set_transceivers(self, txs: TransceiverGrid) -> None:
        d = self.dom
        n = txs.ntx
        idx = self._arr(txs.idx, (4, n, 3), np.int64)
        w = self._arr(txs.w, (4, n, 8))
        src = self._arr(txs.src, (4, n, d.nt))
        dr = self._arr(txs.direction, (n, 4))
        vs = self._arr(txs.vn_src, (n, d.nt))
        L.check(self.lib.adf_solver_set_transceivers(self.h, n, L.ptr(idx), L.ptr(w), L.ptr(src), L.ptr(dr),
                                                     L.ptr(vs)), "set_transceivers")
        self.ntx = n

    def get_trace(self, which: int) -> np.ndarray:
        out = np.empty((self.ntx, self.dom.nt), self.T)
        if self.ntx:
            L.check(self.lib.adf_solver_get_trace(self.h, which, L.ptr(out)), "get_trace")
        return out

    def set_trace(self, which: int, arr) -> None:
        a = self._arr(arr, (self.ntx, self.dom.nt))
        L.check(self.lib.adf_solver_set_trace(self.h, which, L.ptr(a)), "set_trace")

    def get_transceivers(self, txs: TransceiverGrid) -> None:
        """Copy rec, vn_rec and the (vn2v!-mutated) src buffers back."""
        if txs.ntx == 0:
            return
        for c in range(4):
            txs.rec[c][...] = self.get_trace(L.ADF_REC_P + c)
            txs.src[c][...] = self.get_trace(L.ADF_SRC_P + c)
        txs.vn_rec[...] = self.get_trace(L.ADF_VN_REC)
        txs.vn_src[...] = self.get_trace(L.ADF_VN_SRC)

    def set_gf_map(self, gf) -> None:
        """gf: None or a GF map of adfdtd.extrapolation (PVFromPin, PVFromPV,
        PFromPin) with iRec/iSrc set."""
        if gf is None:
            L.check(self.lib.adf_solver_set_gf_map(self.h, 0, None, 0, 0, 0, None, 0, None, 0, 1), "set_gf_map")
            return
        kind, tensors = gf.kind_and_tensors()
        ts = [np.ascontiguousarray(t_, np.float64) for t_ in tensors]
        d3, d2, ng = ts[0].shape  # numpy layout of the Julia (nt_gf, d2, d3)
        PA = C.c_void_p * 4
        arr = PA(*([t_.ctypes.data for t_ in ts] + [None] * (4 - len(ts))))
        ir = np.ascontiguousarray(gf.iRec, np.int64)
        isr = np.ascontiguousarray(gf.iSrc, np.int64)
        L.check(self.lib.adf_solver_set_gf_map(self.h, kind, arr, ng, d2, d3, L.ptr(ir), len(ir), L.ptr(isr),
                                               len(isr), 1 if gf.writeback else 0), "set_gf_map")

    # -- stepping ------------------------------------------------------
    def run(self, start_step: int = 1, end_step: Optional[int] = None) -> None:
        if end_step is None:
            end_step = self.dom.nt
        L.check(self.lib.adf_solver_run(self.h, start_step, end_step), "run")

    def pressure_slice(self, axis: int, index: int) -> np.ndarray:
        d = self.dom
        shp = [(d.nz, d.ny), (d.nz, d.nx), (d.ny, d.nx)][axis]
        out = np.empty(shp, self.T)
        L.check(self.lib.adf_solver_get_pressure_slice(self.h, axis, index, L.ptr(out)), "pressure_slice")
        return out

    def set_option(self, option: int, value: int) -> None:
        """Kernel path / tiling knob (ADF_OPT_*); results do not depend on it."""
        L.check(self.lib.adf_solver_set_option(self.h, option, int(value)), "set_option")

    def set_path(self, path: str = "auto", tile_y: int = 0, block_z: int = 0, waves: int = 0,
                 block_z_gen: int = 0, pipeline: int = 1, launch: int = 0, store_nt: int = -1,
                 tx_launch: int = 2, cells_per_lane: int = 0) -> None:
        self.set_option(L.ADF_OPT_PATH, {"auto": 0, "two-pass": 1, "fused": 2}[path])
        self.set_option(L.ADF_OPT_TILE_Y, tile_y)
        self.set_option(L.ADF_OPT_BLOCK_Z, block_z)
        self.set_option(L.ADF_OPT_WAVES, waves)
        self.set_option(L.ADF_OPT_BLOCK_Z_GEN, block_z_gen)
        self.set_option(L.ADF_OPT_PIPELINE, pipeline)
        self.set_option(L.ADF_OPT_SWEEP_LAUNCH, launch)
        self.set_option(L.ADF_OPT_STORE_NT, store_nt)
        self.set_option(L.ADF_OPT_TX_LAUNCH, tx_launch)
        self.set_option(L.ADF_OPT_CELLS_PER_LANE, cells_per_lane)

    def gf_resample(self, first: int, n_rec: int, timeshift: float, nt_gf: int):
        """The per-run tail of impulsive_gfs (src/greens.jl:166-191) on the
        device: p and vn traces of transceivers first..first+n_rec-1
        (0-based), denoised and resampled onto range(0, tmax; length =
        nt_gf) from range(-timeshift, tmax - timeshift; length = nt) ->
        two float32 arrays (n_rec, nt_gf) (Julia (nt_gf, n_rec))."""
        out_p = np.empty((n_rec, nt_gf), np.float32)
        out_vn = np.empty((n_rec, nt_gf), np.float32)
        L.check(self.lib.adf_solver_gf_resample(self.h, first, n_rec, float(timeshift), nt_gf, L.ptr(out_p),
                                                L.ptr(out_vn)), "gf_resample")
        return out_p, out_vn

    def sweep_stats(self):
        """(mean ms of one field-update launch, launches) of the last run with
        set_option(ADF_OPT_TIME_SWEEP, 1)."""
        v = C.c_double()
        n = C.c_int64()
        L.check(self.lib.adf_solver_sweep_stats(self.h, C.byref(v), C.byref(n)), "sweep_stats")
        return v.value, n.value

    def last_step_ms(self) -> float:
        v = C.c_double()
        L.check(self.lib.adf_solver_last_step_ms(self.h, C.byref(v)), "last_step_ms")
        return v.value

    def next_step(self) -> int:
        """The 1-based step this solver takes next (1 after create/reset)."""
        v = C.c_int64()
        L.check(self.lib.adf_solver_next_step(self.h, C.byref(v)), "next_step")
        return v.value
