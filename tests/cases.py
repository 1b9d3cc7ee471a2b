"""Shared builders for the parity tests: one seeded case is turned into the
host arrays of the GPU library (built by adfdtd, the product) and of the
oracle (built by the oracle's own setup functions), and the two are
asserted equal bit for bit before anything runs."""
from __future__ import annotations

import numpy as np

import adfdtd as A
import oracle as O


def domain_cells(n, h, nt, c0=343.0, r0=1.21, cf=0.5):
    """Domain with cell size h, or (hx, hy, hz) per axis (xmax = (n-1)h/2
    per axis), and exactly nt steps (tmax = (nt-1/2) dt)."""
    nx, ny, nz = n if isinstance(n, (tuple, list)) else (n, n, n)
    hx, hy, hz = h if isinstance(h, (tuple, list)) else (h, h, h)
    xm, ym, zm = (nx - 1) * hx / 2, (ny - 1) * hy / 2, (nz - 1) * hz / 2
    d0 = A.make_domain(1.0, xm, ym, zm, nx, ny, nz, c0, r0, cf)
    return A.make_domain((nt - 0.5) * d0.dt, xm, ym, zm, nx, ny, nz, c0, r0, cf)


def build(n=32, h=0.00686, nt=80, npml=6, shape="none", shape_params=None, n_rec=20, rec_radius_cells=8,
          src_cells=(0.0, 0.0, 0.0), src_dir=(1.0, 0.0, 0.0, 0.0), fc=2000.0, extra=(), seed=0):
    """Point source + Fibonacci receivers.  `extra` appends (point, dir, tf)."""
    d = domain_cells(n, h, nt)
    tf = A.ricker_wavelet(d, fc, True)
    assert np.array_equal(tf, O.ricker_wavelet(oracle_domain(d), fc, True))
    pts, nrm = A.fibonacci_sphere(n_rec, rec_radius_cells * d.dx) if n_rec else (np.zeros((0, 3)), None)
    if n_rec:
        opts, onrm = O.fibonacci_sphere(n_rec, rec_radius_cells * d.dx)
        assert np.array_equal(pts, opts) and np.array_equal(nrm, onrm)
    txs = [A.make_transceiver(d, np.array(src_cells) * d.dx, src_dir, tf)]
    for i in range(n_rec):
        txs.append(A.make_transceiver(d, pts[i], np.r_[0.0, nrm[i]], np.zeros(d.nt)))
    for (pt, dr, tfx) in extra:
        txs.append(A.make_transceiver(d, pt, dr, tfx))
    cpml = A.make_cpml(d, npml, 1e-4, fc)
    mask = A.build_update_mask(d, shape, shape_params)
    omask = O.build_update_mask(oracle_domain(d), shape, shape_params)
    assert (mask is None and omask is None) or np.array_equal(mask, omask)
    return d, txs, cpml, mask


def oracle_domain(d):
    """The oracle's Domain from the same arguments; every field must equal
    the product's."""
    rd = O.domain(d.tmax, d.xmax, d.ymax, d.zmax, d.nx, d.ny, d.nz, d.c0, d.r0, d.cf)
    for k, _ in O.RefDomain._fields_:
        assert getattr(rd, k) == getattr(d, k), k
    return rd


def oracle_transceivers(rd, d, txs, T):
    """The oracle's packing of the transceivers, asserted equal to the
    product's (indices, weights, sources, directions)."""
    tx = O.transceivers_to_grid(rd, [(t.point, t.direction, t.tf) for t in txs], T)
    g = A.transceivers_to_grid(d, txs, T)
    assert np.array_equal(g.idx, tx["idx"])
    for k, r in (("w", "w"), ("src", "src"), ("direction", "dir")):
        assert np.array_equal(getattr(g, k), tx[r]), k
    return tx


def oracle_solver(O, d, txs, cpml, mask, mode, field=None):
    """Oracle state from the same case, its inputs built by the oracle."""
    rd = oracle_domain(d)
    T = np.float64 if mode == "f64" else np.float32
    tx = oracle_transceivers(rd, d, txs, T)
    if field is not None:
        field = np.array(field, T, copy=True)  # the oracle steps its field in place
    s = O.RefSolver(rd, npml=cpml.n, rcoef=cpml.rcoef, fc=cpml.fc, mask=mask, txs=tx, mode=mode, field=field)
    return s


def gpu_run(d, txs, cpml, mask, dtype, precision, field=None, gf=None, path="auto", tile_y=0, block_z=0, waves=0,
            pipeline=1, launch=0, store_nt=-1, tx_launch=2, cells_per_lane=0):
    g = A.transceivers_to_grid(d, txs, dtype)
    f = np.zeros((4, d.nz, d.ny, d.nx), dtype) if field is None else np.array(field, dtype)
    cp = A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, dtype)
    A.run_fdtd(d, f, g, cp, mask, gf_map=gf, precision=precision, path=path, tile_y=tile_y, block_z=block_z,
               waves=waves, pipeline=pipeline, launch=launch, store_nt=store_nt, tx_launch=tx_launch,
               cells_per_lane=cells_per_lane)
    return f, g, cp


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
