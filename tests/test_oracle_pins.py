"""Pins for the oracle (the CPU restatement of the reference) that do not
retype its formulas: closed forms, invariants the mathematics fixes, brute
force on tiny inputs, and the reference's own test designs
(test/test_reciprocity.jl, test/test_extrapolation_writecount.jl).
Each check is chosen so a plausible slip (dropped term, sign, index,
transposed operand) in oracle/fdtd_ref.c fails it."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

C_AIR, RHO_AIR = 343.0, 1.21


def cfg1_domain(n=64, nt=600):
    fc = 2000.0
    dx = C_AIR / (2.5 * fc) / 10
    xmax = (n - 1) * dx / 2
    d0 = O.domain(1.0, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
    return O.domain((nt - 0.5) * d0.dt, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)


# ------------------------------------------------------------------ Domain

def test_domain_cfg1_numbers():
    # values quoted for the 64^3 air case: dx = lambda_min/10 = 6.86e-3 m,
    # xmax = 0.21609 m, dt = 0.5*dx/(c*sqrt(3)) = 5.7735e-6 s, nt = 600
    d = cfg1_domain()
    assert abs(d.dx - 6.86e-3) < 1e-15 and abs(d.xmax - 0.21609) < 1e-15
    assert abs(d.dt - 5.7735e-6) < 1e-10 and d.nt == 600
    assert d.dt == pytest.approx(0.5 * 6.86e-3 / 343.0 / math.sqrt(3), rel=1e-15)
    assert d.z0 == pytest.approx(RHO_AIR * C_AIR, rel=1e-15)
    # P update factor rho c^2 dt and V factor dt/rho: their product is (c dt)^2
    assert d.fact_m0 * d.fact_m1 == pytest.approx((C_AIR * d.dt) ** 2, rel=1e-14)


def test_coordinates_exact_and_symmetric():
    d = cfg1_domain(n=37, nt=10)
    x = O.axis_coords(d.xmax, d.nx, d.dx)
    xv = O.axis_coords(d.xmax, d.nx, d.dx, staggered=True)
    assert x[0] == -d.xmax and x[-1] == d.xmax and x[18] == 0.0
    assert np.array_equal(x, -x[::-1])  # symmetric bit for bit
    for i in (0, 5, 17, 36):  # correctly rounded exact values
        ex = Fraction(d.xmax) * (2 * (i + 1) - 1 - d.nx) / (d.nx - 1)
        assert x[i] == float(ex)
        assert xv[i] == float(ex - Fraction(d.dx) / 2)
    assert np.all(np.diff(x) > 0) and np.all(x - xv > 0)
    t = O.time_axis(d)
    assert t[0] == 0.0 and t[-1] == d.tmax and len(t) == d.nt


# --------------------------------------------------------------------- CPML

@pytest.mark.parametrize("p", [1, 5, 16])
def test_cpml_profile_is_the_quadratic_damping_law(p):
    # invert b = exp(-(d+alpha) dt), a = d (b-1)/(d+alpha) for d and alpha and
    # check d = d0 * (dist/p)^2 with d0 = -3 c ln(R) / (2 p dx) and
    # alpha = pi f0 (1 - dist/p), on-grid dist 0..p-1, half-grid 0,0.5..p-0.5
    d = cfg1_domain()
    R, f0 = 1e-4, 2000.0
    c = O.cpml_coefficients(d, p, R, f0)
    d0 = -3 * d.c0 * math.log(R) / (2 * d.dx * p)
    on = np.arange(p) / p
    half = np.r_[0.0, np.arange(1, p + 1) - 0.5] / p
    for side, dist in (("r", on), ("hr", half)):
        a, b = c["a_" + side], c["b_" + side]
        s = -np.log(b) / d.dt                      # d + alpha
        dd = np.where(np.abs(b - 1) > 0, a * s / (b - 1), 0.0)
        np.testing.assert_allclose(dd, d0 * dist ** 2, rtol=1e-9, atol=1e-6 * d0)
        np.testing.assert_allclose(s - dd, math.pi * f0 * (1 - dist), rtol=1e-9, atol=1e-6)
    # the left tables are the right ones reversed (cpmlcoeffs.jl:329-330)
    for k in ("a", "b"):
        np.testing.assert_array_equal(c[k + "_l"], c[k + "_r"][::-1])
        np.testing.assert_array_equal(c[k + "_hl"], c[k + "_hr"][::-1])
    assert np.all(c["a_r"] <= 0) and np.all((c["b_r"] > 0) & (c["b_r"] <= 1))
    assert c["a_r"][0] == 0 and c["a_hr"][0] == 0  # no damping at the inner edge


# -------------------------------------------------------------- trilinear

def test_trilinear_reproduces_linear_fields_and_brackets_the_point():
    d = cfg1_domain(n=20, nt=10)
    rng = np.random.default_rng(0)
    xs = {k: O.axis_coords(getattr(d, k + "max"), d.nx, d.dx) for k in "xyz"}
    xv = {k: O.axis_coords(getattr(d, k + "max"), d.nx, d.dx, True) for k in "xyz"}
    grids = [("x", "y", "z"), ("xv", "y", "z"), ("x", "yv", "z"), ("x", "y", "zv")]
    co = lambda name: xv[name[0]] if name.endswith("v") else xs[name]
    g = np.array([0.7, -1.3, 2.1])
    for _ in range(50):
        pt = rng.uniform(-0.8, 0.8, 3) * d.xmax
        idx, w = O.trilinear(d, pt)
        for ch in range(4):
            gx, gy, gz = (co(n) for n in grids[ch])
            I, J, K = idx[ch] - 1
            assert gx[I] < pt[0] <= gx[I + 1] and gy[J] < pt[1] <= gy[J + 1] and gz[K] < pt[2] <= gz[K + 1]
            val = 0.0
            for q in range(8):
                a, b, cc = q & 1, (q >> 1) & 1, q >> 2
                node = np.array([gx[I + a], gy[J + b], gz[K + cc]])
                val += w[ch, q] * (g @ node + 0.3)
            assert val == pytest.approx(g @ pt + 0.3, rel=1e-12, abs=1e-14)
            assert w[ch].sum() == pytest.approx(1.0, rel=1e-13)


def test_trilinear_zero_weights_beyond_last_node_and_error_below_first():
    d = cfg1_domain(n=12, nt=10)
    idx, w = O.trilinear(d, np.array([d.xmax * 1.2, 0.01, 0.0]))
    assert idx[0, 0] == d.nx and not w[0].any()
    with pytest.raises(ValueError):
        O.trilinear(d, np.array([-d.xmax * 1.5, 0.0, 0.0]))


# ------------------------------------------------------------ geometry

@pytest.mark.parametrize("shape,prm", [("sphere", [0.05]), ("cube", [0.07]), ("cross", [0.12]),
                                       ("ellipsoid", [0.06, 0.03, 0.045])])
def test_mask_against_brute_force_point_in_shape(shape, prm):
    d = cfg1_domain(n=30, nt=10)
    m = O.build_update_mask(d, shape, prm)
    x = O.axis_coords(d.xmax, d.nx, d.dx)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    if shape == "sphere":
        solid = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) <= prm[0]
    elif shape == "cube":
        solid = (np.abs(X) <= prm[0] / 2) & (np.abs(Y) <= prm[0] / 2) & (np.abs(Z) <= prm[0] / 2)
    elif shape == "ellipsoid":
        solid = (X / prm[0]) ** 2 + (Y / prm[1]) ** 2 + (Z / prm[2]) ** 2 <= 1
    else:  # Greek cross: centre cube of side l/3 plus six arms reaching l/2
        l = prm[0]
        a = np.abs(np.stack([X, Y, Z]))
        inner = a < l / 6 + 1e-15
        solid = (a < l / 2).all(0) & (inner.sum(0) >= 2)
    # exact ties at the boundary are excluded from the comparison
    far = np.ones_like(solid)
    if shape in ("sphere", "ellipsoid"):
        r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) if shape == "sphere" else None
        if r is not None:
            far = np.abs(r - prm[0]) > 1e-12
    assert np.array_equal((m[0] == 0)[far], solid[far])
    s = m[0] == 0
    np.testing.assert_array_equal(m[1][:, :, 1:], ~s[:, :, :-1] & ~s[:, :, 1:])
    np.testing.assert_array_equal(m[2][:, 1:, :], ~s[:, :-1, :] & ~s[:, 1:, :])
    np.testing.assert_array_equal(m[3][1:, :, :], ~s[:-1, :, :] & ~s[1:, :, :])
    assert not m[1][:, :, 0].any() and not m[2][:, 0, :].any() and not m[3][0].any()


def test_ellipsoid_with_equal_axes_is_the_sphere():
    d = cfg1_domain(n=33, nt=10)
    r = 9.3 * d.dx
    assert np.array_equal(O.build_update_mask(d, "ellipsoid", [r, r, r]), O.build_update_mask(d, "sphere", [r]))


# ------------------------------------------------------------ illumination

def test_fibonacci_sphere_points():
    n, R = 257, 0.3
    p, nrm = O.fibonacci_sphere(n, R)
    np.testing.assert_allclose(np.linalg.norm(p, axis=1), R, rtol=1e-14)
    np.testing.assert_allclose(p, R * nrm, rtol=1e-15)
    np.testing.assert_allclose(nrm[:, 2], 1 - (2 * np.arange(n) + 1) / n, rtol=0, atol=1e-15)
    assert np.abs(nrm.mean(0)).max() < 2.0 / n  # near-uniform
    ang = np.arctan2(nrm[1:, 1], nrm[1:, 0]) - np.arctan2(nrm[:-1, 1], nrm[:-1, 0])
    np.testing.assert_allclose(np.mod(ang, 2 * np.pi), np.mod(math.pi * (3 - math.sqrt(5)), 2 * np.pi), atol=1e-12)


def test_wavelets():
    d = cfg1_domain()
    tf, ts = O.impulsive_wavelet(d, 7000.0)
    assert tf.sum() * d.dt == pytest.approx(1.0, rel=1e-14)
    sig = math.sqrt(math.log(2)) / (2 * math.pi * 7000.0)
    pl = round(10 * sig / d.dt)
    assert ts == pytest.approx(pl / 2 * d.dt) and not tf[pl:].any()
    assert np.argmax(tf) + 1 in (pl // 2, pl // 2 + 1)
    raw = O.ricker_wavelet(d, 2000.0, injection=False)
    t = O.time_axis(d)
    k = np.argmax(raw)
    assert abs(t[k] - 1 / 2000.0) <= (t[1] - t[0]) and raw.max() <= 1.0
    inj = O.ricker_wavelet(d, 2000.0, injection=True)
    np.testing.assert_allclose(inj, np.cumsum(raw) * (d.dt / d.dx), rtol=1e-12, atol=1e-18)


def test_c_for_n_closed_form():
    d = cfg1_domain()
    c = O.c_for_n(300, 0.1, d)
    dS = 4 * math.pi * 0.01 / 300
    assert c["dS_in"] == pytest.approx(dS, rel=1e-15)
    assert c["C_f"] == pytest.approx(dS * d.dt / (d.r0 * d.dx ** 3), rel=1e-14)
    assert c["C_q"] == pytest.approx(c["C_f"] * d.z0 ** 2, rel=1e-14)


# -------------------------------------------------------------- the step

def _solver(d, npml=0, mask=None, txs=None, field=None, mode="f64"):
    return O.RefSolver(d, npml=npml, rcoef=1e-4, fc=2000.0, mask=mask, txs=txs, mode=mode, field=field)


def test_plane_wave_stays_one_dimensional_bitwise():
    # an x-travelling field uniform in y,z: dP/dy = 0 exactly, so Vy, Vz stay 0
    # and every (j,k) column evolves identically (simulation.jl:508-514) --
    # except near the (j=ny, k=nz) edge where P is never updated
    # (forward.jl:67-68,76), whose influence spreads at most two cells per
    # step in j and in k.
    s_steps = 8
    d = cfg1_domain(n=40, nt=s_steps)
    f0 = O.initial_field(d, 2000.0, 0.05)
    s = _solver(d, npml=0, field=f0).run()
    P = s.field[0]
    ref_row = P[0, 0]
    kk, jj = np.meshgrid(np.arange(d.nz), np.arange(d.ny), indexing="ij")
    clean = (d.ny - 1 - jj) + (d.nz - 1 - kk) > 4 * s_steps + 4
    assert clean.sum() > 0.5 * clean.size
    assert np.array_equal(P[clean], np.broadcast_to(ref_row, P[clean].shape))
    assert not s.field[2][clean].any() and not s.field[3][clean].any()
    assert not np.array_equal(P[-1, -1], ref_row)  # the edge does break the symmetry


def test_x_mirror_symmetry_before_boundary_contact():
    # source at x=0 on an odd grid: P(i) == P(n+1-i) and Vx(i) == -Vx(n+2-i)
    # (1-based) bit for bit while the numerical domain of dependence (two
    # cells per step) has not reached the x walls, where P(nx) and Vx(1) are
    # frozen (forward.jl:67-106)
    d = cfg1_domain(n=51, nt=10)
    tf = O.ricker_wavelet(d, 2000.0, True)
    tx = O.transceivers_to_grid(d, [(np.array([0.0, 0.3 * d.dx, -0.1 * d.dx]), [1, 0, 0, 0], tf)])
    s = _solver(d, txs=tx).run()
    P, Vx = s.field[0], s.field[1]
    assert np.abs(P).max() > 0
    np.testing.assert_array_equal(P, P[:, :, ::-1])
    np.testing.assert_array_equal(Vx[:, :, 1:], -Vx[:, :, :0:-1])
    assert np.abs(Vx).max() > 0


def _energy(s, d):
    P, Vx, Vy, Vz = s.field
    return 0.5 * (P[:-1, :-1, :-1] ** 2).sum() / (d.r0 * d.c0 ** 2) + 0.5 * d.r0 * (
        (Vx[:, :, 1:] ** 2).sum() + (Vy[:, 1:, :] ** 2).sum() + (Vz[1:] ** 2).sum())


def test_rigid_box_keeps_energy_and_cpml_absorbs_it():
    # lossless leapfrog in a closed box: energy bounded and non-decaying;
    # with CPML the pulse leaves the box and the energy collapses
    d = cfg1_domain(n=28, nt=400)
    tf = O.ricker_wavelet(d, 6000.0, injection=False)  # zero-mean: no net volume injected
    src = [(np.array([0.1, -0.2, 0.3]) * d.dx, [1, 0, 0, 0], tf)]
    e = {}
    for p in (0, 6):
        s = _solver(d, npml=p, txs=O.transceivers_to_grid(d, src))
        s.run(1, 150)
        e1 = _energy(s, d)
        s.run(151, 400)
        e[p] = (e1, _energy(s, d))
    assert 0.8 < e[0][1] / e[0][0] < 1.25
    assert e[6][1] / e[6][0] < 0.05


def test_plane_wave_pulse_travels_at_c():
    # the initialField plane wave (P and Vx staggered by c dt/2,
    # simulation.jl:508-511) moves in +x at c: after N steps its peak sits at
    # -distance + c N dt, and its shape is kept (25 points per wavelength)
    d = cfg1_domain(n=64, nt=60)
    dist = 0.1
    f0 = O.initial_field(d, 2000.0, dist)
    s = _solver(d, field=f0).run()
    x = O.axis_coords(d.xmax, d.nx, d.dx)
    prof = s.field[0, 0, 0]
    k = int(np.argmax(prof))
    y0, y1, y2 = prof[k - 1], prof[k], prof[k + 1]
    xpk = x[k] + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2) * d.dx
    # the staggered leapfrog's phase speed is ~0.2-0.6 % below c over this band
    assert xpk + dist == pytest.approx(d.c0 * d.nt * d.dt, rel=1e-2)
    assert prof.max() == pytest.approx(f0[0].max(), rel=2e-2)


def test_mask_freezes_solid_cells_and_closed_faces():
    d = cfg1_domain(n=26, nt=80)
    m = O.build_update_mask(d, "cube", [8 * d.dx])
    rng = np.random.default_rng(1)
    f0 = rng.standard_normal((4, d.nz, d.ny, d.nx)) * 1e-3
    s = _solver(d, npml=4, mask=m, field=f0.copy()).run()
    core = m[0] == 0
    core[-1, :, :] = core[:, -1, :] = core[:, :, -1] = False  # rigid planes ignore the mask
    assert np.array_equal(s.field[0][core], f0[0][core])
    for c in (1, 2, 3):
        closed = m[c] == 0
        closed[:, :, 0] = closed[:, 0, :] = closed[0] = False  # never-updated / mirrored planes
        assert np.array_equal(s.field[c][closed], f0[c][closed])


def test_free_space_monopole_matches_closed_form():
    # P-channel injection g[it]: p(r,t) = dx^3/(c^2 dt) * g'(t - r/c) / (4 pi r)
    # (the C_p mapping, greens_io.jl:163,182-185).  FDTD vs analytic is a few
    # percent (the reference reports 1-6 %, greens_io.jl:145-147).
    d = cfg1_domain()
    tf = O.ricker_wavelet(d, 2000.0, True)
    pts, nrm = O.fibonacci_sphere(100, 20 * d.dx)
    txs = [(np.zeros(3), [1, 0, 0, 0], tf)] + [(pts[i], np.r_[0, nrm[i]], np.zeros(d.nt)) for i in range(100)]
    s = _solver(d, npml=10, txs=O.transceivers_to_grid(d, txs)).run()
    g = s.tx["src"][0, 0]
    gd = np.gradient(g, d.dt)
    tt = np.arange(1, d.nt + 1) * d.dt
    errs = []
    for i in range(100):
        r = np.linalg.norm(pts[i])
        an = d.dx ** 3 / (d.c0 ** 2 * d.dt) * np.interp(tt - r / d.c0 + 0.5 * d.dt, tt, gd, left=0) / (4 * np.pi * r)
        errs.append(np.linalg.norm(s.p_rec[i + 1] - an) / np.linalg.norm(an))
    assert np.median(errs) < 0.08 and max(errs) < 0.15


def test_reciprocity_reference_setup():
    # test/test_reciprocity.jl:18-69: 51^3, tmax=4e-4, xmax=0.3, cf=0.5, npml=0,
    # water, Gaussian f3db=7 kHz, A=[0.10,-0.05,0.02], B=[-0.04,0.07,0.06];
    # rms(AB-BA)/peak < 1 %, in the reference's own Float32 arithmetic
    d = O.domain(4.0e-4, 0.3, 0.3, 0.3, 51, 51, 51, 1500.0, 1000.0, 0.5)
    tf, _ = O.impulsive_wavelet(d, 7000.0)
    A, B = np.array([0.10, -0.05, 0.02]), np.array([-0.04, 0.07, 0.06])

    def rec(a, b):
        tx = O.transceivers_to_grid(d, [(a, [1, 0, 0, 0], tf), (b, [1, 0, 0, 0], np.zeros(d.nt))], np.float32)
        s = _solver(d, txs=tx, mode="r32").run()
        return s.p_rec[1].astype(np.float64)

    ab, ba = rec(A, B), rec(B, A)
    rms = np.sqrt(np.mean((ab - ba) ** 2))
    assert np.abs(ab).max() > 0 and rms / np.abs(ab).max() < 0.01


def test_float32_reference_arithmetic_tracks_float64():
    d = cfg1_domain(n=32, nt=300)
    tf = O.ricker_wavelet(d, 2000.0, True)
    pts, nrm = O.fibonacci_sphere(20, 9 * d.dx)
    txs = [(np.array([0.2, 0.1, -0.3]) * d.dx, [1, 0, 0, 0], tf)] + [(pts[i], np.r_[0, nrm[i]], np.zeros(d.nt))
                                                                  for i in range(20)]
    s64 = _solver(d, npml=6, txs=O.transceivers_to_grid(d, txs, np.float64)).run()
    s32 = _solver(d, npml=6, txs=O.transceivers_to_grid(d, txs, np.float32), mode="r32").run()
    e = np.linalg.norm(s32.p_rec.astype(np.float64) - s64.p_rec) / np.linalg.norm(s64.p_rec)
    assert e < 1e-5


# ---------------------------------------------------------- extrapolation

def test_gf_gather_equals_scatter_form():
    # test_extrapolation_writecount.jl:354-370: the production gather and a
    # scatter (each input sample pushed into its K future outputs) agree.
    # The oracle's in-loop gather writes the sources of step it+1 from the
    # records up to it; an independent numpy scatter over the same run's
    # records must reproduce every written source sample.
    d = cfg1_domain(n=20, nt=40)
    rng = np.random.default_rng(7)
    pts, nrm = O.fibonacci_sphere(10, 5 * d.dx)
    tf = O.ricker_wavelet(d, 2000.0, True)
    txs = [(np.array([0.3, 0.2, 0.1]) * d.dx, [1, 0, 0, 0], tf)] + [(pts[i], np.r_[0, nrm[i]], np.zeros(d.nt))
                                                                  for i in range(10)]
    iRec, iSrc = np.array([2, 3, 4, 5, 6]), np.array([7, 8, 9])
    K = 9
    pp = rng.standard_normal((len(iRec), len(iSrc), K)) * 1e-3
    vp = rng.standard_normal((len(iRec), len(iSrc), K)) * 1e-3
    s = _solver(d, npml=3, txs=O.transceivers_to_grid(d, txs))
    s.set_gf_map("pv_from_pin", (pp, vp), iRec, iSrc, writeback=True)
    s.run()
    pin = (s.p_rec[iRec - 1] - s.vn_rec[iRec - 1] * d.z0) / 2  # (nrec, nt)
    out_p = np.zeros((len(iSrc), d.nt + K + 1))
    out_v = np.zeros((len(iSrc), d.nt + K + 1))
    for r in range(len(iRec)):
        for m in range(d.nt):
            for k in range(1, K + 1):
                out_p[:, m + k] += pp[r, :, k - 1] * pin[r, m]
                out_v[:, m + k] += vp[r, :, k - 1] * pin[r, m]
    # p_src <- vp-derived, vn_src <- pp-derived (extrapolation.jl:165-167)
    np.testing.assert_allclose(s.tx["src"][0][iSrc - 1, 1:], out_v[:, 1:d.nt], rtol=1e-9, atol=1e-25)
    np.testing.assert_allclose(s.tx["vn_src"][iSrc - 1, 1:], out_p[:, 1:d.nt], rtol=1e-9, atol=1e-25)
    assert np.abs(out_v[:, 1:d.nt]).max() > 0


def test_gf_delta_kernel_is_a_shift():
    # a GF with a single tap at lag L copies pin from L-1 steps earlier
    d = cfg1_domain(n=18, nt=30)
    pts, nrm = O.fibonacci_sphere(6, 4 * d.dx)
    tf = O.ricker_wavelet(d, 2000.0, True)
    txs = [(np.array([0.3, 0.2, 0.1]) * d.dx, [1, 0, 0, 0], tf)] + [(pts[i], np.r_[0, nrm[i]], np.zeros(d.nt))
                                                                  for i in range(6)]
    L = 4
    g = np.zeros((1, 1, 8))
    g[0, 0, L - 1] = 1.0
    s = _solver(d, npml=2, txs=O.transceivers_to_grid(d, txs))
    s.set_gf_map("pv_from_pin", (np.zeros_like(g), g), [2], [7], writeback=True)
    s.run()
    pin = (s.p_rec[1] - s.vn_rec[1] * d.z0) / 2
    src = s.tx["src"][0, 6]
    # extrapolate! at step it writes src[it+1] = sum_k g[k] pin[it+1-k]  -> lag L
    np.testing.assert_allclose(src[L:], pin[: d.nt - L], rtol=1e-12, atol=1e-300)


# ------------------------------------------------------------- analytic KH

def test_kh_kernels_closed_forms_and_pedestal_equivalence():
    rng = np.random.default_rng(5)
    op, on = O.fibonacci_sphere(7, 0.3)
    ip, inn = O.fibonacci_sphere(5, 0.1)
    sig = math.sqrt(math.log(2)) / (2 * math.pi * 3000.0)
    nt, dt = 900, 4e-6  # 3.6 ms: past the longest retardation (0.4 m / c) plus 10 sigma
    t = np.arange(nt) * dt
    c0, rho = 343.0, 1.21
    Kpp = O.eval_one_kernel("p_p", op, on, ip, inn, t, sig, c0, rho)
    Kvv = O.eval_one_kernel("v_v", op, on, ip, inn, t, sig, c0, rho)
    Kdc = O.eval_kernel_pedestal("v_v", op, on, ip, inn, rho)
    # p_p = f'(t - r/c)/(4 pi r): compare with a numerical derivative of the Gaussian
    r = np.linalg.norm(ip[2] - op[3])
    f = lambda tau: np.exp(-tau ** 2 / (2 * sig ** 2)) / (math.sqrt(2 * math.pi) * sig)
    h = 1e-9
    fd = (f(t - r / c0 + h) - f(t - r / c0 - h)) / (2 * h) / (4 * math.pi * r)
    np.testing.assert_allclose(Kpp[2, 3], fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
    # v_v tends to its pedestal at late times
    np.testing.assert_allclose(Kvv[:, :, -1], Kdc, rtol=1e-6)
    # truncated kernel + pedestal == full kernel (kh_kernels.jl:237-296)
    inp = rng.standard_normal((7, nt))
    full = O.direct_kh_accumulate(np.zeros((5, nt)), Kvv, inp, alpha=0.7)
    ntK = 500
    trunc = Kvv[:, :, :ntK] - Kdc[:, :, None]
    ped = O.direct_kh_accumulate(np.zeros((5, nt)), trunc, inp, alpha=0.7, K_dc=Kdc, cumsum_in=np.cumsum(inp, 1))
    np.testing.assert_allclose(ped, full, rtol=1e-9, atol=1e-9 * np.abs(full).max())


def test_direct_kh_delta_kernel():
    K = np.zeros((2, 3, 10))
    K[1, 2, 0] = 1.0
    inp = np.arange(3 * 12, dtype=float).reshape(3, 12)
    out = O.direct_kh_accumulate(np.zeros((2, 12)), K, inp, alpha=2.0)
    np.testing.assert_array_equal(out[1], 2 * inp[2])
    assert not out[0].any()


# ------------------------------------------- impulsive_gfs post-processing


def test_resample_is_the_natural_cubic_spline():
    """resample_time(kind = :cubic) uses BSpline(Cubic(Line(OnGrid())))
    (src/resample.jl:35-40): the C2 cubic with zero second derivative at the
    end nodes.  scipy's CubicSpline(bc_type="natural") is an independent
    implementation of that interpolant."""
    from scipy.interpolate import CubicSpline

    rng = np.random.default_rng(11)
    y = rng.standard_normal((3, 57))
    c = O.natural_bspline_coefficients(y)
    xq = np.sort(rng.uniform(1.0, 57.0, 400))
    ref = CubicSpline(np.arange(1, 58), y, axis=-1, bc_type="natural")(xq)
    assert np.max(np.abs(O.bspline_eval(c, xq) - ref)) < 1e-12
    # a dropped boundary row or a wrong tap weight would fail this: the
    # interpolant passes through the data and its second derivative vanishes
    # at the ends
    assert np.max(np.abs(O.bspline_eval(c, np.arange(1.0, 58.0)) - y)) < 1e-12
    assert np.allclose(c[:, 0] - 2 * c[:, 1] + c[:, 2], 0.0, atol=1e-12)
    assert np.allclose(c[:, -3] - 2 * c[:, -2] + c[:, -1], 0.0, atol=1e-12)


def test_resample_time_lines_flat_ends_and_amplitude_scale():
    """Linear data are reproduced exactly (Line ends), t outside t_in takes
    the end value (extrapolate(..., Flat())) and the output is scaled by
    step(t_out)/step(t_in) (src/resample.jl:77)."""
    t0, dt, n = -0.3, 0.01, 80
    tin = t0 + dt * np.arange(n)
    lin = 2.0 - 5.0 * tin
    out = O.resample_time(lin[None, :], t0, dt, 0.0, 0.0137, 40, scale_amplitude=False)[0]
    tout = 0.0137 * np.arange(40)
    expect = np.where(tout > tin[-1], lin[-1], 2.0 - 5.0 * tout)
    assert np.max(np.abs(out - expect)) < 1e-12
    before = O.resample_time(lin[None, :], t0, dt, -1.0, 0.1, 3, scale_amplitude=False)[0]
    assert np.max(np.abs(before - lin[0])) < 1e-13
    y = np.sin(tin * 7.0)[None, :]
    a = O.resample_time(y, t0, dt, 0.0, 0.02, 20, scale_amplitude=True)
    b = O.resample_time(y, t0, dt, 0.0, 0.02, 20, scale_amplitude=False)
    assert np.allclose(a, b * 2.0, rtol=1e-15, atol=0)


def test_gf_denoise_uses_the_signed_maximum():
    """greens.jl:176-177: A .= ifelse.(abs.(A) .< 1e-4 .* maximum(A, dims=1),
    0, A) -- the threshold is the signed maximum, so an all-negative trace
    keeps every sample, and the comparison is strict."""
    a = np.array([[1.0, 5e-5, -5e-5, 1e-4, -2.0, 3e-5]])
    d = O.gf_denoise(a)
    assert d.tolist() == [[1.0, 0.0, 0.0, 1e-4, -2.0, 0.0]]
    neg = np.array([[-1.0, -1e-9, -3.0]])
    assert np.array_equal(O.gf_denoise(neg), neg)


def test_concat_grids_equals_packing_the_list():
    # impulsive_gfs packs the shared receivers once and puts each run's source
    # in front: the same arrays as packing [source] + receivers
    import adfdtd as A
    d = A.make_domain(20 * 5.7735e-6, 0.21609, 0.21609, 0.21609, 64, 64, 64, 343.0, 1.21, 0.5)
    pts, nrm = A.fibonacci_sphere(30, 20 * d.dx)
    rng = np.random.default_rng(5)
    rec = [A.make_transceiver(d, pts[i], np.r_[0.0, nrm[i]], np.zeros(d.nt)) for i in range(30)]
    src = A.make_transceiver(d, np.array([1.3, -0.4, 2.2]) * d.dx, [1.0, 0.0, 0.0, 0.0], rng.standard_normal(d.nt))
    for T in (np.float32, np.float64):
        a = A.transceivers_to_grid(d, [src] + rec, T)
        b = A.concat_grids(A.transceivers_to_grid(d, [src], T), A.transceivers_to_grid(d, rec, T))
        for k in ("idx", "w", "src", "rec", "direction", "vn_src", "vn_rec"):
            x, y = getattr(a, k), getattr(b, k)
            assert x.dtype == y.dtype and np.array_equal(x, y), k
