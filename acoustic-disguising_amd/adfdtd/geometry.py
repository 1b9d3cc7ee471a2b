"""Scatterer masks, mirroring src/geometry.jl of the reference.

`build_update_mask` returns the reference's Bool (nx,ny,nz,4) update mask as
uint8 (4,nz,ny,nx), or None for :none (so run_fdtd takes the unmasked path,
src/geometry.jl:251).  The same mask can be built on the device with
`Solver.build_mask`; both evaluate the predicates at the pressure-grid
coordinates with the same floating-point operations.
"""
from __future__ import annotations

import numpy as np

from .simulation import Domain, coords

SCATTERER_DEFAULTS = {"sphere_radius": 0.15, "cube_length": 0.20, "cross_length": 0.30}


def outside_predicate(shape: str, params):
    """src/geometry.jl:86-143 (`_outside_predicate`), vectorised over numpy
    arrays.  'ellipsoid' (a, b, c) is an addition: outside iff
    x*x/(a*a) + y*y/(b*b) + z*z/(c*c) > 1."""
    prm = np.atleast_1d(np.asarray(params, np.float64)) if params is not None else np.zeros(1)
    if shape == "none":
        return lambda X, Y, Z: np.ones(np.broadcast(X, Y, Z).shape, bool)
    if shape == "sphere":
        r = prm[0]
        return lambda X, Y, Z: X * X + Y * Y + Z * Z > r * r
    if shape == "cube":
        r = prm[0]
        return lambda X, Y, Z: (np.abs(X) > r / 2) | (np.abs(Y) > r / 2) | (np.abs(Z) > r / 2)
    if shape == "cross":
        l = prm[0]
        s = l / 2
        sl = 2.0 * l / 3

        def inc(X, Y, Z):
            return (np.abs(X) < s) & (np.abs(Y) < s) & (np.abs(Z) < s)

        def f(X, Y, Z):
            arm = np.zeros(np.broadcast(X, Y, Z).shape, bool)
            for a, b in ((+1, +1), (-1, -1), (+1, -1), (-1, +1)):
                arm |= inc(X + a * sl, Y + b * sl, Z)
                arm |= inc(X + a * sl, Y, Z + b * sl)
                arm |= inc(X, Y + a * sl, Z + b * sl)
            return ~inc(X, Y, Z) | arm

        return f
    if shape == "ellipsoid":
        a, b, c = prm[0], prm[1], prm[2]
        return lambda X, Y, Z: X * X / (a * a) + Y * Y / (b * b) + Z * Z / (c * c) > 1.0
    raise ValueError(f"Unknown scatterer: {shape}. Expected one of none, sphere, cube, cross, ellipsoid.")


def build_scatterer(d: Domain, shape: str, params) -> np.ndarray:
    """src/geometry.jl:159-176: Bool (nz,ny,nx), True inside the scatterer."""
    cs = coords(d)
    if shape == "none":
        return np.zeros((d.nz, d.ny, d.nx), bool)
    X = cs["x"]["p"][None, None, :]
    Y = cs["y"]["p"][None, :, None]
    Z = cs["z"]["p"][:, None, None]
    return ~outside_predicate(shape, params)(X, Y, Z)


def build_update_mask(d: Domain, shape: str, params):
    """src/geometry.jl:250-280 -> uint8 (4,nz,ny,nx) or None for 'none'."""
    if shape == "none":
        return None
    s = build_scatterer(d, shape, params)
    f = ~s
    u = np.zeros((4, d.nz, d.ny, d.nx), np.uint8)
    u[0] = f
    u[1, :, :, 1:] = f[:, :, :-1] & f[:, :, 1:]
    u[2, :, 1:, :] = f[:, :-1, :] & f[:, 1:, :]
    u[3, 1:, :, :] = f[:-1, :, :] & f[1:, :, :]
    return u
