"""ORACLE — test infrastructure only.

ctypes front end of ``oracle/liboracle.so``, the plain-C restatement of the
reference FDTD (``oracle/fdtd_ref.c``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this module, and only as the checker / the timed CPU baseline.  The product
package never imports it.

Array layouts are the reference's Julia column-major layouts written as
C-order numpy shapes with the axes reversed:

  field   Julia (nx,ny,nz,4)       -> numpy (4, nz, ny, nx)
  mask    Julia Bool (nx,ny,nz,4)  -> numpy uint8 (4, nz, ny, nx)
  mv[0]   Julia (p,ny,nz,2)        -> numpy (2, nz, ny, p)      (mp: p+1)
  mv[1]   Julia (nx,p,nz,2)        -> numpy (2, nz, p, nx)
  mv[2]   Julia (nx,ny,p,2)        -> numpy (2, p, ny, nx)
  a, b    Julia (p,2)              -> numpy (2, p)              (ah, bh: p+1)
  idx     Julia 4 x (3,ntx)        -> numpy int64 (4, ntx, 3)   1-based
  w       Julia 4 x (2,2,2,ntx)    -> numpy (4, ntx, 8)
  src/rec Julia 4 x (nt,ntx)       -> numpy (4, ntx, nt)        channel p,vx,vy,vz
  dir     Julia (4,ntx)            -> numpy (ntx, 4)
  vn_*    Julia (nt,ntx)           -> numpy (ntx, nt)

Parity pins for every function are in ``oracle/README.md``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lock = threading.Lock()
_lib = None


class RefDomain(C.Structure):
    """simulation.jl:25-44 field order."""

    _fields_ = [
        ("tmax", C.c_double), ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("c0", C.c_double), ("r0", C.c_double), ("cf", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("fact_m0", C.c_double), ("fact_m1", C.c_double), ("dt", C.c_double),
        ("nt", C.c_int64), ("z0", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RefGFMap(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("writeback", C.c_int32),
        ("nt_gf", C.c_int64), ("nsrc", C.c_int64), ("nrec", C.c_int64),
        ("g", C.c_void_p * 4),
        ("iRec", C.c_void_p), ("iSrc", C.c_void_p),
        ("n_irec", C.c_int64), ("n_isrc", C.c_int64),
    ]


def build() -> str:
    """Compile liboracle.so with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                build()
            L = C.CDLL(LIB_PATH)
            d = C.c_double
            i64 = C.c_int64
            p = C.c_void_p
            L.ref_domain_make.argtypes = [C.POINTER(RefDomain), d, d, d, d, i64, i64, i64, d, d, d]
            L.ref_coord.restype = d
            L.ref_coord.argtypes = [d, i64, i64, d, C.c_int32]
            L.ref_time.restype = d
            L.ref_time.argtypes = [d, i64, i64]
            L.ref_cpml_coefficients.argtypes = [d, d, i64, d, d, d] + [p] * 8
            L.ref_trilinear.restype = C.c_int
            L.ref_trilinear.argtypes = [C.POINTER(RefDomain), p, p, p]
            L.ref_build_update_mask.argtypes = [C.POINTER(RefDomain), C.c_int, p, p]
            L.ref_fibonacci_sphere.argtypes = [i64, d, p, p]
            L.ref_impulsive_wavelet.restype = d
            L.ref_impulsive_wavelet.argtypes = [C.POINTER(RefDomain), d, p]
            L.ref_ricker_wavelet.argtypes = [C.POINTER(RefDomain), d, C.c_int32, p]
            L.ref_initial_field.argtypes = [C.POINTER(RefDomain), d, d, p]
            L.ref_c_for_n.argtypes = [i64, d, d, d, d, d, p]
            L.ref_kh_erf.restype = d
            L.ref_kh_erf.argtypes = [d]
            L.ref_eval_one_kernel.argtypes = [C.c_int, i64, p, p, i64, p, p, i64, p, d, d, d, p]
            L.ref_eval_kernel_pedestal.argtypes = [C.c_int, i64, p, p, i64, p, p, d, p]
            L.ref_direct_kh_accumulate.restype = C.c_int
            L.ref_direct_kh_accumulate.argtypes = [p, i64, i64, p, i64, i64, p, d, p, p]
            run_args = [C.POINTER(RefDomain), p, i64] + [p] * 11 + [i64] + [p] * 7 + [i64, i64, p]
            for suf in ("_f64", "_r32"):
                f = getattr(L, "ref_run_fdtd" + suf)
                f.restype = C.c_int
                f.argtypes = run_args
            _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------- setup


def domain(tmax, xmax, ymax, zmax, nx, ny, nz, c0=1500.0, r0=1000.0, cf=0.25) -> RefDomain:
    """simulation.jl:46-79."""
    d = RefDomain()
    lib().ref_domain_make(C.byref(d), tmax, xmax, ymax, zmax, nx, ny, nz, c0, r0, cf)
    return d


def axis_coords(xmax, n, dx, staggered=False):
    """simulation.jl:89-100 for one axis."""
    L = lib()
    return np.array([L.ref_coord(xmax, n, i, dx, 1 if staggered else 0) for i in range(1, n + 1)])


def time_axis(d):
    """simulation.jl:86 t(d)."""
    L = lib()
    return np.array([L.ref_time(d.tmax, d.nt, i) for i in range(1, d.nt + 1)])


def cpml_coefficients(d, npml, rcoef=1e-4, fc=3e3):
    """simulation.jl:222-223 -> cpmlcoeffs.jl:36-126.  Returns the eight
    vectors a_l, a_r, a_hl, a_hr, b_l, b_r, b_hl, b_hr (double)."""
    arr = {k: np.zeros(npml) for k in ("a_l", "a_r", "b_l", "b_r")}
    arr.update({k: np.zeros(npml + 1) for k in ("a_hl", "a_hr", "b_hl", "b_hr")})
    lib().ref_cpml_coefficients(d.c0, d.dt, npml, rcoef, d.dx * npml, fc,
                                *[_ptr(arr[k]) for k in ("a_l", "a_r", "a_hl", "a_hr", "b_l", "b_r", "b_hl", "b_hr")])
    return arr


def trilinear(d, point):
    """simulation.jl:254-341.  Returns (idx (4,3) int64 1-based, w (4,8))."""
    idx = np.zeros((4, 3), np.int64)
    w = np.zeros((4, 8))
    pt = _f64(point)
    rc = lib().ref_trilinear(C.byref(d), _ptr(pt), _ptr(idx), _ptr(w))
    if rc != 0:
        raise ValueError("point lies below every grid node (the reference throws here)")
    return idx, w


SHAPES = {"none": 0, "sphere": 1, "cube": 2, "cross": 3, "ellipsoid": 4}


def build_update_mask(d, shape, params):
    """geometry.jl:250-280 (None for :none, as the reference)."""
    if shape == "none":
        return None
    prm = _f64(np.atleast_1d(params).tolist() + [0.0] * 3)
    m = np.zeros((4, d.nz, d.ny, d.nx), np.uint8)
    lib().ref_build_update_mask(C.byref(d), SHAPES[shape], _ptr(prm), _ptr(m))
    return m


def fibonacci_sphere(n, radius=1.0):
    """illumination.jl:19-46 -> (points (n,3), normals (n,3))."""
    pts = np.zeros((3, n))
    nrm = np.zeros((3, n))
    lib().ref_fibonacci_sphere(n, radius, _ptr(pts), _ptr(nrm))
    return pts.T.copy(), nrm.T.copy()


def impulsive_wavelet(d, f_3db_hz):
    tf = np.zeros(d.nt)
    ts = lib().ref_impulsive_wavelet(C.byref(d), f_3db_hz, _ptr(tf))
    return tf, ts


def ricker_wavelet(d, fc=18000.0, injection=True):
    tf = np.zeros(d.nt)
    lib().ref_ricker_wavelet(C.byref(d), fc, 1 if injection else 0, _ptr(tf))
    return tf


def initial_field(d, fc, distance):
    """simulation.jl:468-518 in double (the reference stores it as Float32)."""
    f = np.zeros((4, d.nz, d.ny, d.nx))
    lib().ref_initial_field(C.byref(d), fc, distance, _ptr(f))
    return f


def c_for_n(N, r_inner, d):
    out = np.zeros(3)
    lib().ref_c_for_n(N, r_inner, d.dt, d.r0, d.dx, d.z0, _ptr(out))
    return {"C_f": out[0], "C_q": out[1], "dS_in": out[2]}


def transceivers_to_grid(d, txs, dtype=np.float64):
    """simulation.jl:387-431.  txs: list of (point, direction(4), tf(nt)).

    Returns a dict with idx, w, src, rec, dir, vn_src, vn_rec in the layouts
    listed in the module docstring, cast to ``dtype`` as Data.Array does.
    """
    ntx = len(txs)
    idx = np.zeros((4, ntx, 3), np.int64)
    w = np.zeros((4, ntx, 8))
    src = np.zeros((4, ntx, d.nt))
    dirs = np.zeros((ntx, 4))
    for s, (pt, dr, tf) in enumerate(txs):
        i4, w4 = trilinear(d, pt)
        idx[:, s, :] = i4
        w[:, s, :] = w4
        dr = np.asarray(dr, np.float64)
        dirs[s] = dr
        for ch in range(4):
            src[ch, s] = np.asarray(tf, np.float64) * dr[ch]
    return {
        "idx": idx,
        "w": w.astype(dtype),
        "src": src.astype(dtype),
        "rec": np.zeros((4, ntx, d.nt), dtype),
        "dir": dirs.astype(dtype),
        "vn_src": np.zeros((ntx, d.nt), dtype),
        "vn_rec": np.zeros((ntx, d.nt), dtype),
    }


# ------------------------------------------------------------- analytic KH

KERNELS = {"p_p": 0, "p_v": 1, "v_p": 2, "v_v": 3}


def eval_one_kernel(which, outer_pts, outer_nrm, inner_pts, inner_nrm, t_grid, sigma, c0, rho0):
    """kh_kernels.jl:106-157 -> K (nt, n_outer, n_inner) as numpy (n_inner, n_outer, nt)."""
    op, on = _f64(np.asarray(outer_pts).T), _f64(np.asarray(outer_nrm).T)
    ip, inn = _f64(np.asarray(inner_pts).T), _f64(np.asarray(inner_nrm).T)
    no, ni, nt = op.shape[1], ip.shape[1], len(t_grid)
    K = np.zeros((ni, no, nt))
    lib().ref_eval_one_kernel(KERNELS[which], no, _ptr(op), _ptr(on), ni, _ptr(ip), _ptr(inn), nt,
                              _ptr(_f64(t_grid)), sigma, c0, rho0, _ptr(K))
    return K


def eval_kernel_pedestal(which, outer_pts, outer_nrm, inner_pts, inner_nrm, rho0):
    op, on = _f64(np.asarray(outer_pts).T), _f64(np.asarray(outer_nrm).T)
    ip, inn = _f64(np.asarray(inner_pts).T), _f64(np.asarray(inner_nrm).T)
    no, ni = op.shape[1], ip.shape[1]
    Kdc = np.zeros((ni, no))
    lib().ref_eval_kernel_pedestal(KERNELS[which], no, _ptr(op), _ptr(on), ni, _ptr(ip), _ptr(inn), rho0, _ptr(Kdc))
    return Kdc


def direct_kh_accumulate(out, K, inp, alpha=1.0, K_dc=None, cumsum_in=None):
    """kh_kernels.jl:256-296.  out (n_inner, nt), K (n_inner, n_outer, ntK),
    inp (n_outer, nt); numpy layouts of the Julia (nt, ·) matrices."""
    assert out.dtype == np.float64 and out.flags.c_contiguous
    K = _f64(K)
    inp = _f64(inp)
    ni, no, ntK = K.shape
    nt = out.shape[1]
    rc = lib().ref_direct_kh_accumulate(_ptr(out), nt, ni, _ptr(K), ntK, no, _ptr(inp), alpha,
                                        _ptr(None if K_dc is None else _f64(K_dc)),
                                        _ptr(None if cumsum_in is None else _f64(cumsum_in)))
    if rc:
        raise ValueError("K longer than output")
    return out


# ---------------------------------------------------------------- solver


class RefSolver:
    """State of one reference run: field, CPML memory and coefficients,
    mask and packed transceivers, stepped by the C restatement of
    forward_onestep! (forward.jl:13-139).

    mode "f64": Float64 arrays (the reference algorithm with number type
    Float64); mode "r32": the reference's own Float32 arrays with Float64
    domain scalars.
    """

    def __init__(self, d, npml=0, rcoef=1e-4, fc=3e3, mask=None, txs=None, mode="f64", field=None):
        self.d = d
        self.mode = mode
        self.T = np.float64 if mode == "f64" else np.float32
        T = self.T
        nx, ny, nz, p = d.nx, d.ny, d.nz, npml
        self.p = p
        self.field = np.zeros((4, nz, ny, nx), T) if field is None else np.ascontiguousarray(field, T)
        self.mv = [np.zeros((2, nz, ny, p), T), np.zeros((2, nz, p, nx), T), np.zeros((2, p, ny, nx), T)]
        q = p + 1
        self.mp = [np.zeros((2, nz, ny, q), T), np.zeros((2, nz, q, nx), T), np.zeros((2, q, ny, nx), T)]
        c = cpml_coefficients(d, p, rcoef, fc)
        self.coef = c
        self.a = np.ascontiguousarray(np.stack([c["a_l"], c["a_r"]]), T)
        self.b = np.ascontiguousarray(np.stack([c["b_l"], c["b_r"]]), T)
        self.ah = np.ascontiguousarray(np.stack([c["a_hl"], c["a_hr"]]), T)
        self.bh = np.ascontiguousarray(np.stack([c["b_hl"], c["b_hr"]]), T)
        self.mask = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        if txs is None:
            txs = transceivers_to_grid(d, [], T)
        self.tx = {k: np.ascontiguousarray(v, np.int64 if k == "idx" else T) for k, v in txs.items()}
        self.ntx = self.tx["dir"].shape[0]
        self._gf_keep = None
        self.gf = None

    def set_gf_map(self, kind, tensors, iRec, iSrc, writeback=True):
        """kind: 'pv_from_pin' (p_pin, v_pin: numpy (nrec, nsrc, nt_gf)),
        'pv_from_pv' (pp, pv, vp, vv: numpy (nout, nm, nt_gf)) or
        'p_from_pin' (pos, neg: numpy (nsrc, nrec, nt_gf)).  iRec/iSrc 1-based."""
        kinds = {"pv_from_pin": 1, "pv_from_pv": 2, "p_from_pin": 3}
        g = RefGFMap()
        g.kind = kinds[kind]
        g.writeback = 1 if writeback else 0
        ts = [_f64(t) for t in tensors]
        g.nt_gf = ts[0].shape[2]
        g.nsrc = ts[0].shape[1]
        g.nrec = ts[0].shape[0]
        for i, t in enumerate(ts):
            g.g[i] = t.ctypes.data
        ir = np.ascontiguousarray(iRec, np.int64)
        isr = np.ascontiguousarray(iSrc, np.int64)
        g.iRec = ir.ctypes.data
        g.iSrc = isr.ctypes.data
        g.n_irec = len(ir)
        g.n_isrc = len(isr)
        self._gf_keep = (ts, ir, isr)
        self.gf = g

    def run(self, start_step=1, end_step=None):
        d = self.d
        if end_step is None:
            end_step = d.nt
        f = getattr(lib(), "ref_run_fdtd_" + ("f64" if self.mode == "f64" else "r32"))
        t = self.tx
        rc = f(C.byref(d), _ptr(self.field), self.p,
               *[_ptr(x) for x in self.mv], *[_ptr(x) for x in self.mp],
               _ptr(self.a), _ptr(self.b), _ptr(self.ah), _ptr(self.bh), _ptr(self.mask),
               self.ntx, _ptr(t["idx"]), _ptr(t["w"]), _ptr(t["src"]), _ptr(t["rec"]), _ptr(t["dir"]),
               _ptr(t["vn_src"]), _ptr(t["vn_rec"]), start_step, end_step,
               None if self.gf is None else C.byref(self.gf))
        if rc:
            raise RuntimeError(f"oracle run failed: {rc}")
        return self

    @property
    def p_rec(self):
        return self.tx["rec"][0]

    @property
    def vn_rec(self):
        return self.tx["vn_rec"]


def set_threads(n: int) -> None:
    """OpenMP threads used by the stencil loops of the oracle."""
    os.environ["OMP_NUM_THREADS"] = str(n)
    try:
        gomp = C.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(C.c_int(n))
    except OSError:
        pass


# --------------------------------------------- impulsive_gfs post-processing
#
# Plain numpy (float64) restatements, checked against scipy's natural cubic
# spline and closed forms in tests/test_oracle_pins.py.

def gf_denoise(A):
    """src/greens.jl:176-177: per trace (last axis = time here), zero every
    sample with |A| < 1e-4 * max_t A, where the maximum is the SIGNED
    maximum over time (``maximum(A, dims = 1)``)."""
    A = np.array(A, dtype=np.float64, copy=True)
    mx = A.max(axis=-1, keepdims=True)
    A[np.abs(A) < 1e-4 * mx] = 0.0
    return A


def natural_bspline_coefficients(y):
    """Coefficients c_0..c_{n+1} of the cubic B-spline interpolant of y at
    the nodes 1..n with ``Line(OnGrid())`` boundaries (Interpolations.jl's
    ``BSpline(Cubic(Line(OnGrid())))``, used by src/resample.jl:35-40): the
    rows (c_{i-1} + 4 c_i + c_{i+1}) / 6 = y_i plus a zero second difference
    c_0 - 2 c_1 + c_2 = 0 = c_{n-1} - 2 c_n + c_{n+1} at both ends (zero
    second derivative at the end nodes: the natural spline).  The end rows
    give c_1 = y_1 and c_n = y_n; the inner system is tridiagonal (Thomas
    algorithm along the last axis)."""
    y = np.asarray(y, dtype=np.float64)
    n = y.shape[-1]
    c = np.zeros(y.shape[:-1] + (n + 2,))
    if n == 1:
        c[..., :] = y[..., :1]
        return c
    c[..., 1] = y[..., 0]
    c[..., n] = y[..., n - 1]
    m = n - 2  # unknowns c_2..c_{n-1}
    if m > 0:
        rhs = 6.0 * y[..., 1:n - 1].copy()
        rhs[..., 0] -= c[..., 1]
        rhs[..., m - 1] -= c[..., n]
        cp = np.zeros(m)
        dp = np.zeros(y.shape[:-1] + (m,))
        cp[0] = 1.0 / 4.0
        dp[..., 0] = rhs[..., 0] / 4.0
        for i in range(1, m):
            den = 4.0 - cp[i - 1]
            cp[i] = 1.0 / den
            dp[..., i] = (rhs[..., i] - dp[..., i - 1]) / den
        x = np.zeros_like(dp)
        x[..., m - 1] = dp[..., m - 1]
        for i in range(m - 2, -1, -1):
            x[..., i] = dp[..., i] - cp[i] * x[..., i + 1]
        c[..., 2:n] = x
    c[..., 0] = 2.0 * c[..., 1] - c[..., 2]
    c[..., n + 1] = 2.0 * c[..., n] - c[..., n - 1]
    return c


def bspline_eval(c, xq):
    """Evaluate the cubic B-spline with coefficients c (last axis, nodes
    1..n at c[1..n]) at index coordinates xq, clamped to [1, n] (``Flat``
    extrapolation)."""
    c = np.asarray(c, dtype=np.float64)
    n = c.shape[-1] - 2
    xq = np.clip(np.asarray(xq, dtype=np.float64), 1.0, float(n))
    i = np.minimum(np.floor(xq).astype(np.int64), max(n - 1, 1))
    d = xq - i
    e = 1.0 - d
    w0 = e * e * e / 6.0
    w1 = 2.0 / 3.0 - d * d + d * d * d / 2.0
    w2 = 2.0 / 3.0 - e * e + e * e * e / 2.0
    w3 = d * d * d / 6.0
    g = lambda k: np.take(c, np.clip(i + k, 0, n + 1), axis=-1)
    return w0 * g(-1) + w1 * g(0) + w2 * g(1) + w3 * g(2)


def resample_time(A, t_in0, dt_in, t_out0, dt_out, n_out, scale_amplitude=True):
    """src/resample.jl:35-77 (kind = :cubic) along the last axis: the
    natural cubic B-spline through A on t_in = t_in0 + (k-1) dt_in, Flat
    outside, evaluated at t_out = t_out0 + (k-1) dt_out, times
    dt_out / dt_in."""
    c = natural_bspline_coefficients(A)
    xq = 1.0 + (t_out0 + np.arange(n_out) * dt_out - t_in0) / dt_in
    out = bspline_eval(c, xq)
    return out * (dt_out / dt_in) if scale_amplitude else out
