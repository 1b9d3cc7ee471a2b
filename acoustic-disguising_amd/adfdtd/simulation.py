"""Host side of the FDTD core, mirroring src/simulation.jl of the reference.

Same names, argument meanings and error behaviour as the Julia API
(`Domain`, `coords`, `Transceiver`, `Cpml`, `resetCPMLmemory!`,
`transceiversToGrid`, `initialField`, `run_fdtd!`); the time loop itself
runs in libadfdtd.so on the GPU.

Numpy layouts are the Julia column-major arrays with the axes reversed:
field (nx,ny,nz,4) -> (4,nz,ny,nx); a (nt,ntx) trace -> (ntx,nt); CPML
memory p[0] (p+1,ny,nz,2) -> (2,nz,ny,p+1), etc.
"""
from __future__ import annotations

import functools
import math
from dataclasses import dataclass, field as dc_field
from fractions import Fraction
from typing import Callable, List, Optional, Sequence

import numpy as np

# ----------------------------------------------------------------- Domain


@dataclass(frozen=True)
class Domain:
    """src/simulation.jl:25-44 (same fields, same order)."""

    tmax: float
    xmax: float
    ymax: float
    zmax: float
    nx: int
    ny: int
    nz: int
    c0: float
    r0: float
    cf: float
    dx: float
    dy: float
    dz: float
    fact_m0: float
    fact_m1: float
    dt: float
    nt: int
    z0: float


def make_domain(tmax=10e-4, xmax=1.0, ymax=1.0, zmax=1.0, nx=201, ny=201, nz=201,
                c0=1500.0, r0=1000.0, cf=0.25) -> Domain:
    """`Domain(; tmax, xmax, ..., cf)` (src/simulation.jl:46-79)."""
    tmax, xmax, ymax, zmax = float(tmax), float(xmax), float(ymax), float(zmax)
    nx, ny, nz = int(nx), int(ny), int(nz)
    c0, r0, cf = float(c0), float(r0), float(cf)
    dx = 2 * xmax / (nx - 1)
    dy = 2 * ymax / (ny - 1)
    dz = 2 * zmax / (nz - 1)
    dt = cf * min(dx, dy, dz) / c0 / math.sqrt(3)
    nt = int(math.ceil(tmax / dt))
    fact_m0 = dt * r0 * (c0 * c0)
    fact_m1 = dt / r0
    z0 = r0 * c0
    return Domain(tmax, xmax, ymax, zmax, nx, ny, nz, c0, r0, cf, dx, dy, dz, fact_m0, fact_m1, dt, nt, z0)


# ---------------------------------------------------------- coordinates
# range(-xmax, xmax, length=n)[i] is evaluated as the correctly rounded value
# of the exact rational xmax*(2i-1-n)/(n-1) (Julia's TwicePrecision ranges
# agree to the last ulp); the staggered grids subtract dx/2 exactly before
# the single rounding, like `x(d) .- dx/2` on a TwicePrecision range.


@functools.lru_cache(maxsize=64)
def _axis_cached(xmax: float, n: int, dx: float, staggered: bool) -> np.ndarray:
    fx = Fraction(xmax)
    h = Fraction(dx) / 2
    out = np.empty(n)
    for i in range(1, n + 1):
        v = fx * (2 * i - 1 - n) / (n - 1)
        if staggered:
            v -= h
        out[i - 1] = float(v)
    out.flags.writeable = False
    return out


def _axis(xmax: float, n: int, dx: float, staggered: bool) -> np.ndarray:
    # the exact-rational evaluation costs milliseconds per axis and every
    # transceivers_to_grid call of a domain needs the same six axes
    return _axis_cached(float(xmax), int(n), float(dx), bool(staggered)).copy()


def x(d: Domain) -> np.ndarray:
    """Cell-centred x (src/simulation.jl:89)."""
    return _axis(d.xmax, d.nx, d.dx, False)


def y(d: Domain) -> np.ndarray:
    return _axis(d.ymax, d.ny, d.dy, False)


def z(d: Domain) -> np.ndarray:
    return _axis(d.zmax, d.nz, d.dz, False)


def xv(d: Domain) -> np.ndarray:
    """Staggered x, shifted by -dx/2 (src/simulation.jl:96)."""
    return _axis(d.xmax, d.nx, d.dx, True)


def yv(d: Domain) -> np.ndarray:
    return _axis(d.ymax, d.ny, d.dy, True)


def zv(d: Domain) -> np.ndarray:
    return _axis(d.zmax, d.nz, d.dz, True)


def t(d: Domain) -> np.ndarray:
    """`range(0; stop=tmax, length=nt)` (src/simulation.jl:86).  Note its step
    is tmax/(nt-1), not dt."""
    ft = Fraction(d.tmax)
    if d.nt == 1:
        return np.zeros(1)
    return np.array([float(ft * (i - 1) / (d.nt - 1)) for i in range(1, d.nt + 1)])


def coords(d: Domain) -> dict:
    """src/simulation.jl:109-115."""
    X, Y, Z, XV, YV, ZV = x(d), y(d), z(d), xv(d), yv(d), zv(d)
    return {
        "x": {"p": X, "vx": XV, "vy": X, "vz": X},
        "y": {"p": Y, "vx": Y, "vy": YV, "vz": Y},
        "z": {"p": Z, "vx": Z, "vy": Z, "vz": ZV},
    }


# ---------------------------------------------------------- transceivers

CHANNELS = ("p", "vx", "vy", "vz")


def trilinear_interpolation(d: Domain, points: np.ndarray, cs: Optional[dict] = None):
    """src/simulation.jl:254-341 for many points at once.

    points (N,3) -> (idx int64 (4,N,3) 1-based, w (4,N,8) column-major 2x2x2).
    Raises ValueError where the reference throws (a point below every node).
    """
    pts = np.atleast_2d(np.asarray(points, np.float64))
    if cs is None:
        cs = coords(d)
    xp, yp, zp = pts[:, 0], pts[:, 1], pts[:, 2]
    N = pts.shape[0]
    idx = np.zeros((4, N, 3), np.int64)
    w = np.zeros((4, N, 8))
    ds = d.dx * d.dy
    for c, ch in enumerate(CHANNELS):
        xy = "p" if ch == "vz" else ch  # vz reuses p's x-y stencil (:281)
        xk, yk, zk = cs["x"][xy], cs["y"][xy], cs["z"][ch]
        ix = np.searchsorted(xk, xp, side="left")  # findlast(xk .< xp)
        iy = np.searchsorted(yk, yp, side="left")
        iz = np.searchsorted(zk, zp, side="left")
        if (ix == 0).any() or (iy == 0).any() or (iz == 0).any():
            raise ValueError("trilinearInterpolation: point below every grid node")
        okxy = (ix < d.nx) & (iy < d.ny)
        ixs, iys = np.minimum(ix, d.nx - 1), np.minimum(iy, d.ny - 1)
        x1, x2 = xk[ixs - 1], xk[ixs]
        y1, y2 = yk[iys - 1], yk[iys]
        dx0, dx1 = xp - x1, x2 - xp
        dy0, dy1 = yp - y1, y2 - yp
        wxy = np.stack([dx1 * dy1 / ds, dx0 * dy1 / ds, dx1 * dy0 / ds, dx0 * dy0 / ds], axis=1)
        wxy[~okxy] = 0.0
        okz = iz < d.nz
        izs = np.minimum(iz, d.nz - 1)
        z1, z2 = zk[izs - 1], zk[izs]
        lo = wxy * (z2 - zp)[:, None] / d.dz
        hi = wxy * (zp - z1)[:, None] / d.dz
        lo[~okz] = 0.0
        hi[~okz] = 0.0
        w[c, :, 0:4] = lo
        w[c, :, 4:8] = hi
        idx[c, :, 0], idx[c, :, 1], idx[c, :, 2] = ix, iy, iz
    return idx, w


@dataclass
class Transceiver:
    """src/simulation.jl:168-187: point, time function and direction
    [q, fx, fy, fz]."""

    point: np.ndarray
    tf: np.ndarray
    direction: np.ndarray = dc_field(default_factory=lambda: np.array([1.0, 0.0, 0.0, 0.0]))


def make_transceiver(d: Domain, point, direction, tf) -> Transceiver:
    return Transceiver(np.array(point, np.float64).ravel().copy(), np.asarray(tf, np.float64),
                       np.array(direction, np.float64).ravel().copy())


@dataclass
class TransceiverGrid:
    """src/simulation.jl:354-378 packed transceivers (numpy layouts):
    idx int64 (4,ntx,3) 1-based, w (4,ntx,8), src (4,ntx,nt), rec (4,ntx,nt),
    direction (ntx,4), vn_src (ntx,nt), vn_rec (ntx,nt)."""

    idx: np.ndarray
    w: np.ndarray
    src: np.ndarray
    rec: np.ndarray
    direction: np.ndarray
    vn_src: np.ndarray
    vn_rec: np.ndarray

    @property
    def ntx(self) -> int:
        return self.direction.shape[0]


def concat_grids(a: TransceiverGrid, b: TransceiverGrid) -> TransceiverGrid:
    """The packing of a's transceivers followed by b's: transceivers_to_grid
    packs each transceiver on its own (simulation.jl:387-431), so this equals
    packing the concatenated list."""
    return TransceiverGrid(idx=np.concatenate([a.idx, b.idx], axis=1), w=np.concatenate([a.w, b.w], axis=1),
                           src=np.concatenate([a.src, b.src], axis=1), rec=np.concatenate([a.rec, b.rec], axis=1),
                           direction=np.concatenate([a.direction, b.direction], axis=0),
                           vn_src=np.concatenate([a.vn_src, b.vn_src], axis=0),
                           vn_rec=np.concatenate([a.vn_rec, b.vn_rec], axis=0))


def p_src(t: TransceiverGrid) -> np.ndarray:
    """src/simulation.jl:449-452 accessors; rows are transceivers."""
    return t.src[0]


def p_rec(t: TransceiverGrid) -> np.ndarray:
    return t.rec[0]


def vn_src(t: TransceiverGrid) -> np.ndarray:
    return t.vn_src


def vn_rec(t: TransceiverGrid) -> np.ndarray:
    return t.vn_rec


def transceivers_to_grid(d: Domain, txs: Sequence[Transceiver], dtype=np.float32) -> TransceiverGrid:
    """`transceiversToGrid` (src/simulation.jl:387-431); `dtype` plays the
    role of Data.Array's number type."""
    ntx = len(txs)
    if ntx:
        pts = np.stack([t_.point for t_ in txs])
        idx, w = trilinear_interpolation(d, pts)
        dirs = np.stack([t_.direction for t_ in txs])
        tfs = np.stack([t_.tf for t_ in txs])
        if tfs.shape[1] != d.nt:
            raise ValueError("transceiver time functions must have length nt")
        src = tfs[None, :, :] * dirs.T[:, :, None]  # src[c,is,:] = tf * direction[c]
    else:
        idx = np.zeros((4, 0, 3), np.int64)
        w = np.zeros((4, 0, 8))
        dirs = np.zeros((0, 4))
        src = np.zeros((4, 0, d.nt))
    return TransceiverGrid(
        idx=idx,
        w=np.ascontiguousarray(w, dtype),
        src=np.ascontiguousarray(src, dtype),
        rec=np.zeros((4, ntx, d.nt), dtype),
        direction=np.ascontiguousarray(dirs, dtype),
        vn_src=np.zeros((ntx, d.nt), dtype),
        vn_rec=np.zeros((ntx, d.nt), dtype),
    )


# ----------------------------------------------------------------- CPML


def _calc_kab(halo: int, dt: float, npower: float, d0: float, alpha_max: float, halfgrid: bool):
    """src/kernels/cpmlcoeffs.jl:65-126 (K_max_pml = nothing)."""
    n = halo + (1 if halfgrid else 0)
    shift = 0.5 if halfgrid else 0.0
    if n <= 0:
        return [], [], [], []
    start = 0 - shift
    stop = (n - shift) - 1
    lendiv = max(n - 1, 1)
    dist = []
    for i in range(n):  # LinRange: (1-t)*start + t*stop, t = (i-1)/lendiv
        tt = i / lendiv
        dist.append((1 - tt) * start + tt * stop)
    if halfgrid:
        dist[0] = 0.0
    a_l, a_r, b_l, b_r = [], [], [], []
    for i in range(n):
        ndl, ndr = dist[n - 1 - i], dist[i]
        if halo != 0:
            ndl, ndr = ndl / halo, ndr / halo
        dl, dr = d0 * (ndl * ndl), d0 * (ndr * ndr)
        al, ar = alpha_max * (1.0 - ndl), alpha_max * (1.0 - ndr)
        bl = math.exp((-(dl / 1.0 + al)) * dt)
        br = math.exp((-(dr / 1.0 + ar)) * dt)
        b_l.append(bl)
        b_r.append(br)
        a_l.append(dl * (bl - 1.0) / (1.0 * (dl + 1.0 * al)) if (dl + al) != 0 else float("nan"))
        a_r.append(dr * (br - 1.0) / (1.0 * (dr + 1.0 * ar)) if (dr + ar) != 0 else float("nan"))
    return a_l, a_r, b_l, b_r


def compute_cpml_coefficients(vel_max: float, dt: float, halo: int, rcoef: float, thickness: float,
                              f0: float) -> dict:
    """src/kernels/cpmlcoeffs.jl:36-63 -> the eight coefficient vectors."""
    alpha_max = math.pi * f0
    npower = 2.0
    d0 = -(npower + 1) * vel_max * math.log(rcoef) / (2.0 * thickness) if halo != 0 else 0.0
    a_l, a_r, b_l, b_r = _calc_kab(halo, dt, npower, d0, alpha_max, False)
    a_hl, a_hr, b_hl, b_hr = _calc_kab(halo, dt, npower, d0, alpha_max, True)
    f = lambda v: np.array(v, np.float64)
    return {"a_l": f(a_l), "a_r": f(a_r), "a_hl": f(a_hl), "a_hr": f(a_hr),
            "b_l": f(b_l), "b_r": f(b_r), "b_hl": f(b_hl), "b_hr": f(b_hr)}


@dataclass
class Cpml:
    """src/simulation.jl:200-227: memory per axis (p: extent n+1, v: extent n)
    and the coefficients.  Memory in numpy layout (2, ., ., .)."""

    n: int
    p: List[np.ndarray]
    v: List[np.ndarray]
    c: dict
    rcoef: float = 1e-4
    fc: float = 3e3


def make_cpml(d: Domain, npml: int, rcoef: float = 1e-4, fc: float = 3e3, dtype=np.float32) -> Cpml:
    """`Cpml(d; npml, rcoef, fc)` (src/simulation.jl:206-227)."""
    nx, ny, nz, q = d.nx, d.ny, d.nz, npml + 1
    mp = [np.zeros((2, nz, ny, q), dtype), np.zeros((2, nz, q, nx), dtype), np.zeros((2, q, ny, nx), dtype)]
    mv = [np.zeros((2, nz, ny, npml), dtype), np.zeros((2, nz, npml, nx), dtype),
          np.zeros((2, npml, ny, nx), dtype)]
    c = compute_cpml_coefficients(d.c0, d.dt, npml, rcoef, d.dx * npml, fc)
    return Cpml(npml, mp, mv, c, rcoef, fc)


def reset_cpml_memory(cpml: Cpml) -> None:
    """`resetCPMLmemory!` (src/simulation.jl:235-241)."""
    for a in cpml.p + cpml.v:
        a[...] = 0


# --------------------------------------------------------- initial field


def initial_field(d: Domain, fc: float, distance: float) -> np.ndarray:
    """`initialField` (src/simulation.jl:468-518), plane-wave branch.
    Returns Float32 (4,nz,ny,nx) like the reference."""
    cs = coords(d)
    k = fc / d.c0
    f = np.zeros((4, d.nz, d.ny, d.nx), np.float32)
    cx = cs["x"]["p"]
    cxv = cs["x"]["vx"]
    w1 = math.pi * k * (cx + distance)
    w2 = math.pi * k * ((cxv + distance) - d.c0 * d.dt / 2)
    pw = (1 - 2 * (w1 * w1)) * np.exp(-(w1 * w1))
    vw = (1 - 2 * (w2 * w2)) * np.exp(-(w2 * w2)) / d.z0
    f[0] = pw.astype(np.float32)[None, None, :]
    f[1] = vw.astype(np.float32)[None, None, :]
    return f


# ------------------------------------------------------------- run_fdtd!


def run_fdtd(dom: Domain, field: np.ndarray, txs: TransceiverGrid, cpml: Cpml, va,
             start_step: int = 1, on_step: Optional[Callable] = None, gf_map=None,
             precision: str = "fast", solver=None, path: str = "auto", tile_y: int = 0, block_z: int = 0,
             waves: int = 0, pipeline: int = 1, launch: int = 0, store_nt: int = -1, tx_launch: int = 2,
             cells_per_lane: int = 0):
    """`run_fdtd!(dom, field, txs, cpml, va; start_step, on_step, gf_map)`
    (src/simulation.jl:542-557) on the GPU.

    Mutates `field`, `txs` (rec, vn_rec and the vn2v!-mutated src buffers)
    and `cpml` memory in place, as the reference does.  `va` is the update
    mask (uint8 (4,nz,ny,nx)) or None.  float64 arrays run the Float64
    algorithm; float32 arrays run the fast float path (precision="fast") or
    the reference's own Float32 arithmetic (precision="ref").
    """
    from .solver import Solver, dtype_code

    code = dtype_code(field.dtype, precision)
    s = solver if solver is not None else Solver.cached(dom, code, cpml)
    s.set_path(path, tile_y, block_z, waves, pipeline=pipeline, launch=launch, store_nt=store_nt,
               tx_launch=tx_launch, cells_per_lane=cells_per_lane)
    s.set_field(field)
    s.set_cpml_memory(cpml)
    s.set_mask(va)
    s.set_transceivers(txs)
    s.set_gf_map(gf_map)
    if on_step is None:
        s.run(start_step, dom.nt)
    else:
        for it in range(start_step, dom.nt + 1):
            s.run(it, it)
            s.get_field(field)
            on_step(it, field)
    s.get_field(field)
    s.get_cpml_memory(cpml)
    s.get_transceivers(txs)
    return field
