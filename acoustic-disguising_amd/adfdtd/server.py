"""Served form: a request loop with continuous batching (adf_server_*).

Requests are `run_fdtd!` solves (src/simulation.jl:542-557) as
`impulsive_gfs` (src/greens.jl:143-164) or `build_hologram`
(src/hologram.jl:215-248) set them up, without a GF map.  The server admits
them in arrival order while their device footprint fits its memory budget,
steps every live solve together with the others, retires the finished ones
(copying their traces to host memory) and reports latency percentiles.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Union

import numpy as np

from . import _lib as L
from .simulation import Cpml, Domain, TransceiverGrid, coords
from .solver import SHAPES, np_dtype, to_c_domain

STATUS = {L.ADF_REQ_QUEUED: "queued", L.ADF_REQ_RUNNING: "running", L.ADF_REQ_DONE: "done", L.ADF_REQ_FAILED: "failed"}


class Server:
    """Wraps adf_server_* (include/adfdtd.h)."""

    def __init__(self, device: int = 0, budget_bytes: int = 0, chunk: int = 64):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.adf_server_create(device, int(budget_bytes), C.byref(h)), "adf_server_create")
        self.h = h
        self.meta = {}
        self.set_option(L.ADF_SERVER_CHUNK, chunk)

    def close(self) -> None:
        if getattr(self, "h", None) and self.h.value:
            self.lib.adf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, option: int, value: int) -> None:
        L.check(self.lib.adf_server_set_option(self.h, option, int(value)), "server_set_option")

    def submit(self, dom: Domain, code: int, cpml: Optional[Cpml], txs: Optional[TransceiverGrid],
               mask: Union[None, np.ndarray, tuple] = None, field: Optional[np.ndarray] = None,
               start_step: int = 1, end_step: Optional[int] = None, want: int = 0x1F,
               keep_field: bool = False) -> int:
        """Queue one solve; returns its id.  `mask`: None, an update mask
        (uint8 (4,nz,ny,nx)) or (shape name, params) built on the device."""
        T = np_dtype(code)
        keep = []
        r = L.AdfRequest()
        r.dom = to_c_domain(dom)
        r.dtype = code
        r.cpml.npml = cpml.n if cpml is not None else 0
        if cpml is not None and cpml.n:
            for k in ("a_l", "a_r", "a_hl", "a_hr", "b_l", "b_r", "b_hl", "b_hr"):
                a = np.ascontiguousarray(cpml.c[k], np.float64)
                keep.append(a)
                setattr(r.cpml, k, a.ctypes.data)
        r.mask_shape = 0
        if isinstance(mask, tuple):
            shape, params = mask
            r.mask_shape = SHAPES[shape]
            pv = np.atleast_1d(np.asarray(params, np.float64))
            for i, v in enumerate(pv[:4]):
                r.mask_params[i] = float(v)
            cs = coords(dom)
            xs = [np.ascontiguousarray(cs[a]["p"], np.float64) for a in ("x", "y", "z")]
            keep += xs
            r.xc, r.yc, r.zc = (x.ctypes.data for x in xs)
        elif mask is not None:
            m = np.ascontiguousarray(mask, np.uint8)
            if m.shape != (4, dom.nz, dom.ny, dom.nx):
                raise ValueError("mask must be (4, nz, ny, nx)")
            keep.append(m)
            r.mask_shape = -1
            r.mask4 = m.ctypes.data
        n = txs.ntx if txs is not None else 0
        r.ntx = n
        if n:
            arrs = [np.ascontiguousarray(txs.idx, np.int64), np.ascontiguousarray(txs.w, T),
                    np.ascontiguousarray(txs.src, T), np.ascontiguousarray(txs.direction, T),
                    np.ascontiguousarray(txs.vn_src, T)]
            keep += arrs
            r.idx, r.w, r.src, r.dir, r.vn_src = (a.ctypes.data for a in arrs)
        if field is not None:
            f = np.ascontiguousarray(field, T)
            keep.append(f)
            r.field = f.ctypes.data
        r.start_step = start_step
        r.end_step = dom.nt if end_step is None else end_step
        r.want = want
        r.keep_field = 1 if keep_field else 0
        rid = C.c_int64()
        L.check(self.lib.adf_server_submit(self.h, C.byref(r), C.byref(rid)), "server_submit")
        self.meta[rid.value] = (dom, T, n)
        return rid.value

    def poll(self, rid: int) -> str:
        st = C.c_int()
        L.check(self.lib.adf_server_poll(self.h, rid, C.byref(st)), "server_poll")
        return STATUS[st.value]

    def wait(self, rid: int, timeout: float = -1.0) -> str:
        st = C.c_int()
        L.check(self.lib.adf_server_wait(self.h, rid, float(timeout), C.byref(st)), "server_wait")
        s = STATUS[st.value]
        if s == "failed":
            raise L.AdfError(f"request {rid} failed: {self.error(rid)}")
        return s

    def trace(self, rid: int, which: int) -> np.ndarray:
        dom, T, n = self.meta[rid]
        out = np.empty((n, dom.nt), T)
        if n:
            L.check(self.lib.adf_server_get_trace(self.h, rid, which, L.ptr(out)), "server_get_trace")
        return out

    def field(self, rid: int) -> np.ndarray:
        dom, T, _ = self.meta[rid]
        out = np.empty((4, dom.nz, dom.ny, dom.nx), T)
        L.check(self.lib.adf_server_get_field(self.h, rid, L.ptr(out)), "server_get_field")
        return out

    def error(self, rid: int) -> str:
        return self.lib.adf_server_request_error(self.h, rid).decode(errors="replace")

    def release(self, rid: int) -> None:
        L.check(self.lib.adf_server_release(self.h, rid), "server_release")
        self.meta.pop(rid, None)

    def stats(self) -> dict:
        st = L.AdfServerInfo()
        L.check(self.lib.adf_server_stats(self.h, C.byref(st)), "server_stats")
        return {k: getattr(st, k) for k, _ in L.AdfServerInfo._fields_}
