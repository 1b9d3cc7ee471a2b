"""The product's own host-side construction (adfdtd) against the oracle's,
bit for bit: domain scalars, coordinates, CPML tables, Fibonacci points,
wavelets, C_for_N, the plane-wave initial field, masks of every shape and
the packed transceivers (indices, weights, sources, directions) of the
benchmark configurations, including the 64^3 case's monopole exactly at
the centre, which sits on a tie between staggered nodes."""
import math

import numpy as np
import pytest

import adfdtd as A
import oracle as O

C_AIR, RHO_AIR, FC = 343.0, 1.21, 2000.0
DX = C_AIR / (2.5 * FC) / 10


def both_domains(n, nt, h=(DX, DX, DX), cf=0.5, c0=C_AIR, r0=RHO_AIR):
    nx, ny, nz = n if isinstance(n, tuple) else (n, n, n)
    xm, ym, zm = (nx - 1) * h[0] / 2, (ny - 1) * h[1] / 2, (nz - 1) * h[2] / 2
    a0 = A.make_domain(1.0, xm, ym, zm, nx, ny, nz, c0, r0, cf)
    tmax = (nt - 0.5) * a0.dt
    return A.make_domain(tmax, xm, ym, zm, nx, ny, nz, c0, r0, cf), O.domain(tmax, xm, ym, zm, nx, ny, nz, c0, r0, cf)


DOMAINS = [(64, 600), (256, 30), (192, 30), (128, 30), ((37, 29, 23), 40)]


@pytest.mark.parametrize("n,nt", DOMAINS)
def test_domain_and_coordinates(n, nt):
    a, o = both_domains(n, nt) if not isinstance(n, tuple) else both_domains(n, nt, h=(DX, 1.3 * DX, 0.8 * DX))
    for k, _ in O.RefDomain._fields_:
        assert getattr(a, k) == getattr(o, k), k
    cs = A.coords(a)
    for ax, (name, nn, dd) in enumerate((("x", a.nx, a.dx), ("y", a.ny, a.dy), ("z", a.nz, a.dz))):
        xm = getattr(a, name + "max")
        assert np.array_equal(cs[name]["p"], O.axis_coords(xm, nn, dd))
        stag = {"x": "vx", "y": "vy", "z": "vz"}[name]
        assert np.array_equal(cs[name][stag], O.axis_coords(xm, nn, dd, staggered=True))
    if nt <= 600:
        assert np.array_equal(A.t(a), O.time_axis(o))


@pytest.mark.parametrize("p", [0, 1, 2, 10, 12, 16])
def test_cpml_tables(p):
    a, o = both_domains(64, 10)
    cp = A.make_cpml(a, p, 1e-4, FC, np.float64)
    ref = O.cpml_coefficients(o, p, 1e-4, FC)
    if p == 0:
        return
    for k, v in ref.items():
        assert np.array_equal(cp.c[k], v), k


@pytest.mark.parametrize("n", [1, 2, 3, 16, 100, 500, 2000, 4000])
@pytest.mark.parametrize("radius", [1.0, 20 * DX, 40 * DX, 80 * DX])
def test_fibonacci_sphere(n, radius):
    pa, na = A.fibonacci_sphere(n, radius)
    po, no = O.fibonacci_sphere(n, radius)
    assert np.array_equal(pa, po) and np.array_equal(na, no)


@pytest.mark.parametrize("f", [2000.0, 7000.0])
def test_wavelets(f):
    a, o = both_domains(64, 600)
    assert np.array_equal(A.ricker_wavelet(a, f, True), O.ricker_wavelet(o, f, True))
    assert np.array_equal(A.ricker_wavelet(a, f, False), O.ricker_wavelet(o, f, False))
    ta, sa = A.impulsive_wavelet(a, f)
    to, so = O.impulsive_wavelet(o, f)
    assert np.array_equal(ta, to) and sa == so


def test_c_for_n():
    a, o = both_domains(192, 1500)
    for N, R in ((2000, 40 * DX), (400, 40 * DX), (300, 0.1)):
        ca, co = A.c_for_n(N, R, a), O.c_for_n(N, R, o)
        assert ca["C_f"] == co["C_f"] and ca["C_q"] == co["C_q"] and ca["dS_in"] == co["dS_in"]


def test_initial_field():
    a, o = both_domains((40, 12, 10), 30)
    for dist in (0.05, 0.1):
        fa = A.initial_field(a, FC, dist)
        fo = O.initial_field(o, FC, dist).astype(np.float32)  # the reference stores it as Float32
        assert fa.dtype == np.float32 and np.array_equal(fa, fo)


SHAPES = [("sphere", [24 * DX]), ("sphere", [16 * DX]), ("cube", [30 * DX]), ("cross", [40 * DX]),
          ("ellipsoid", [20 * DX, 12 * DX, 16 * DX]), ("ellipsoid", [10 * DX, 10 * DX, 10 * DX])]


@pytest.mark.parametrize("shape,prm", SHAPES, ids=[f"{s}-{i}" for i, (s, _) in enumerate(SHAPES)])
@pytest.mark.parametrize("n", [64, 96, 128, (75, 61, 57)])
def test_masks(shape, prm, n):
    a, o = both_domains(n, 10) if not isinstance(n, tuple) else both_domains(n, 10, h=(DX, 1.1 * DX, 0.9 * DX))
    ma = A.build_update_mask(a, shape, prm)
    mo = O.build_update_mask(o, shape, prm)
    assert ma.dtype == mo.dtype and np.array_equal(ma, mo)
    assert (ma[0] == 0).any()


def _oracle_tx(o, txs, T):
    return O.transceivers_to_grid(o, [(t.point, t.direction, t.tf) for t in txs], T)


def assert_grids_equal(g, ref):
    assert np.array_equal(g.idx, ref["idx"])
    for k, r in (("w", "w"), ("src", "src"), ("direction", "dir"), ("vn_src", "vn_src"), ("rec", "rec")):
        x, y = getattr(g, k), ref[r]
        assert x.dtype == y.dtype and np.array_equal(x, y), k


def cfg1_set(a):
    tf = A.ricker_wavelet(a, FC, True)
    pts, nrm = A.fibonacci_sphere(100, 20 * a.dx)
    z = np.zeros(a.nt)
    return [A.make_transceiver(a, [0.0, 0.0, 0.0], [1.0, 0, 0, 0], tf)] + \
        [A.make_transceiver(a, pts[i], np.r_[0.0, nrm[i]], z) for i in range(100)]


def test_transceivers_64_centre_monopole_and_100_receivers():
    a, o = both_domains(64, 600)
    txs = cfg1_set(a)
    for T in (np.float64, np.float32):
        g = A.transceivers_to_grid(a, txs, T)
        assert_grids_equal(g, _oracle_tx(o, txs, T))
    # the centre.  xv(d) = x(d) .- dx/2 (simulation.jl:95): node 33 of the
    # staggered axis is the exact xmax/63 - dx/2 rounded once, which is half
    # the rounding error of dx = 2 xmax/63: 1.8e-19 m, above zero, so
    # findlast(xv .< 0) (simulation.jl:266) takes node 32 and the monopole's
    # x weight falls almost wholly on node 33
    from fractions import Fraction
    cs = A.coords(a)
    node33 = float(Fraction(a.xmax) / 63 - Fraction(a.dx) / 2)
    assert cs["x"]["vx"][32] == node33 and 0.0 < node33 < 1e-18
    assert cs["x"]["p"][31] < 0.0 < cs["x"]["p"][32]
    g = A.transceivers_to_grid(a, txs[:1], np.float64)
    assert g.idx[0, 0].tolist() == [32, 32, 32] and g.idx[1, 0].tolist() == [32, 32, 32]
    np.testing.assert_allclose(g.w[0, 0], 0.125, rtol=1e-12)
    lo_x = g.w[1, 0][[0, 2, 4, 6]].sum()  # weight on node 32 of the Vx axis
    assert 0.0 < lo_x < 1e-15 and abs(g.w[1, 0].sum() - 1.0) < 1e-15


def test_transceivers_256_and_192_sets():
    for n, npts in ((256, 2000), (192, 2000)):
        a, o = both_domains(n, 20)
        kh, khn = A.fibonacci_sphere(npts, 40 * a.dx)
        ff, ffn = A.fibonacci_sphere(500, 80 * a.dx)
        rng = np.random.default_rng(n)
        txs = [A.make_transceiver(a, [60 * a.dx, 0.0, 0.0], [1.0, 0, 0, 0], A.ricker_wavelet(a, FC, True))]
        txs += [A.make_transceiver(a, kh[i], np.r_[0.0, khn[i]], rng.standard_normal(a.nt)) for i in range(npts)]
        txs += [A.make_transceiver(a, ff[i], np.r_[0.0, ffn[i]], np.zeros(a.nt)) for i in range(500)]
        for T in (np.float32, np.float64):
            assert_grids_equal(A.transceivers_to_grid(a, txs, T), _oracle_tx(o, txs, T))


def test_transceivers_anisotropic_and_outside_points():
    a, o = both_domains((37, 29, 23), 20, h=(DX, 1.3 * DX, 0.8 * DX))
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1, 1, (300, 3)) * np.array([a.xmax, a.ymax, a.zmax]) * 1.05
    pts = pts[(pts > -np.array([a.xmax, a.ymax, a.zmax]) + 0.6 * a.dx).all(1)]  # the reference throws below node 1
    txs = [A.make_transceiver(a, p, [0.5, 0.2, -0.3, 0.7], rng.standard_normal(a.nt)) for p in pts]
    assert_grids_equal(A.transceivers_to_grid(a, txs, np.float64), _oracle_tx(o, txs, np.float64))
    g = A.transceivers_to_grid(a, txs, np.float64)
    assert (g.idx[0, :, 0] >= a.nx).any()  # some points lie past the last node: zero weights
