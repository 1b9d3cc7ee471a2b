"""Post-hoc Kirchhoff-Helmholtz extrapolation on the device.

`kh_analytic_accumulate` replaces the reference's
`eval_one_kernel` + `eval_kernel_pedestal` + `direct_kh_accumulate!`
(src/kh_kernels.jl:106-157,174-203,256-296): the analytic kernel is
evaluated inside the convolution kernel on its arrival window, never stored.
Arrays use the numpy layouts of the Julia matrices (axes reversed):
points/normals (n, 3), inp (n_outer, nt), out (n_inner, nt).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .solver import context

KERNELS = {"p_p": 0, "p_v": 1, "v_p": 2, "v_v": 3}


def kh_analytic_accumulate(out: np.ndarray, which: str, outer_pts, outer_nrm, inner_pts, inner_nrm, t_grid,
                           sigma_t: float, c0: float, rho0: float, inp, alpha: float = 1.0,
                           pedestal: bool = False, dtype=np.float64, device: int = 0) -> np.ndarray:
    """out[s, t] += alpha * sum_r sum_{tt <= min(t, ntK)} K(tt; r, s) inp[r, t-tt+1]
    (1-based t, tt), in place; K is the named kernel on t_grid (minus its
    pedestal, plus the pedestal times cumsum(inp), when pedestal=True)."""
    lib = L.load()
    if out.dtype != np.float64 or not out.flags.c_contiguous:
        raise ValueError("out must be a C-contiguous float64 array (n_inner, nt)")
    op = np.ascontiguousarray(np.asarray(outer_pts, np.float64).T)
    on = np.ascontiguousarray(np.asarray(outer_nrm, np.float64).T)
    ip = np.ascontiguousarray(np.asarray(inner_pts, np.float64).T)
    inn = np.ascontiguousarray(np.asarray(inner_nrm, np.float64).T)
    tg = np.ascontiguousarray(t_grid, np.float64)
    x = np.ascontiguousarray(inp, np.float64)
    no, ni, nt = op.shape[1], ip.shape[1], out.shape[1]
    if x.shape != (no, nt) or out.shape[0] != ni:
        raise ValueError("shape mismatch: inp (n_outer, nt), out (n_inner, nt)")
    code = L.ADF_F64 if np.dtype(dtype) == np.float64 else L.ADF_F32
    L.check(lib.adf_kh_analytic_accumulate(context(device), KERNELS[which], code, no, L.ptr(op), L.ptr(on), ni,
                                           L.ptr(ip), L.ptr(inn), L.ptr(tg), len(tg), float(sigma_t), float(c0),
                                           float(rho0), 1 if pedestal else 0, float(alpha), nt, L.ptr(x),
                                           L.ptr(out)), "kh_analytic_accumulate")
    return out


def kh_last_ms() -> float:
    """Device time of this thread's last kh_analytic_accumulate (ms)."""
    v = C.c_double()
    L.check(L.load().adf_kh_last_ms(C.byref(v)), "kh_last_ms")
    return v.value
