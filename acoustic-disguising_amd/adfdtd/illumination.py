"""Source/receiver placement and wavelets, mirroring src/illumination.jl."""
from __future__ import annotations

import math

import numpy as np

from .simulation import Domain, t


def fibonacci_sphere(n: int, radius: float = 1.0):
    """src/illumination.jl:19-46 -> (points (n,3), normals (n,3))."""
    if n <= 0:
        raise ValueError("fibonacci_sphere requires n > 0")
    pts = np.empty((n, 3))
    nrm = np.empty((n, 3))
    golden = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(1, n + 1):
        zz = 1.0 - (2.0 * (i - 1) + 1.0) / n
        r_xy = math.sqrt(max(0.0, 1.0 - zz * zz))
        th = golden * (i - 1)
        nx, ny = r_xy * math.cos(th), r_xy * math.sin(th)
        nrm[i - 1] = (nx, ny, zz)
        pts[i - 1] = (radius * nx, radius * ny, radius * zz)
    return pts, nrm


def r2_sphere(n: int, radius: float = 1.0):
    """src/illumination.jl:69-97 -> (points (n,3), normals (n,3)): Roberts'
    R2 low-discrepancy sequence, u = frac(i/phi), v = frac(i/phi^2) with
    phi the plastic number (the real root of x^3 = x + 1), mapped to the
    sphere by the Lambert equal-area map.  Point i depends only on i."""
    if n <= 0:
        raise ValueError("r2_sphere requires n > 0")
    phi = (9 + math.sqrt(69)) ** (1 / 3) / np.cbrt(18.0) + (9 - math.sqrt(69)) ** (1 / 3) / np.cbrt(18.0)
    a1, a2 = 1.0 / phi, 1.0 / phi ** 2
    pts = np.empty((n, 3))
    nrm = np.empty((n, 3))
    for i in range(1, n + 1):
        u = math.fmod(i * a1, 1.0)
        v = math.fmod(i * a2, 1.0)
        ph = 2 * math.pi * u
        zz = 1.0 - 2.0 * v
        r_xy = math.sqrt(max(0.0, 1.0 - zz * zz))
        nx, ny = r_xy * math.cos(ph), r_xy * math.sin(ph)
        nrm[i - 1] = (nx, ny, zz)
        pts[i - 1] = (radius * nx, radius * ny, radius * zz)
    return pts, nrm


def illumination_points(cfg: dict):
    """src/illumination.jl:119-130."""
    s = cfg["surfaces"]
    inner, inner_normals = fibonacci_sphere(int(s["nPoints_inner"]), float(s["radius_inner"]))
    outer, outer_normals = fibonacci_sphere(int(s["nPoints_outer"]), float(s["radius_outer"]))
    return {"inner": inner, "outer": outer, "inner_normals": inner_normals, "outer_normals": outer_normals}


def impulsive_wavelet(d: Domain, f_3db_hz: float):
    """src/illumination.jl:167-184 -> (tf (nt,), timeshift)."""
    sig = math.sqrt(math.log(2)) / (2 * math.pi * float(f_3db_hz))
    pl = int(round(10 * sig / d.dt))  # Python round: ties to even, as round(Int, x)
    pl = min(pl, d.nt)
    center = pl / 2
    width = pl / 10
    tf = np.zeros(d.nt)
    for k in range(1, pl + 1):
        u = k - center
        tf[k - 1] = math.exp(-(u * u) / (2 * (width * width)))
    s = 0.0
    for v in tf:
        s += v
    den = s * d.dt
    tf = tf / den
    return tf, center * d.dt


def ricker_wavelet(d: Domain, fc: float = 18000.0, injection: bool = True) -> np.ndarray:
    """src/illumination.jl:199-207 on t(d) (whose step is not dt)."""
    tt = t(d)
    tf = np.empty(d.nt)
    for i, ti in enumerate(tt):
        wn = math.pi * fc * (ti - 1.0 / fc)
        w2 = wn * wn
        tf[i] = (1.0 - 2.0 * w2) * math.exp(-w2)
    if injection:
        sc = d.dt / d.dx
        acc = 0.0
        for i in range(d.nt):
            acc += tf[i]
            tf[i] = acc * sc
    return tf


def c_for_n(N: int, r_inner: float, d: Domain) -> dict:
    """`C_for_N` (diagnostics/_check_inject_kh/injection_coefficients.jl:467-472):
    dS = 4*pi*R^2/N, C_f = dS*dt/(rho*dx^3), C_q = C_f*z0^2."""
    dS = 4 * math.pi * (r_inner * r_inner) / N
    C_f = dS * d.dt / (d.r0 * (d.dx * d.dx * d.dx))
    return {"C_f": C_f, "C_q": C_f * (d.z0 * d.z0), "dS_in": dS}
