"""Pins for the oracle's in-loop GF gather of every kind, its analytic KH
kernels p_v and v_p, and its CPML inside the time step.  None of them
retypes the oracle's formula:

- gather kinds: a numpy scatter form (each recorded sample pushed into the
  outputs it reaches) over the run's own records, the design of the
  reference test test/test_extrapolation_writecount.jl:354-370, and
  single-tap kernels that must reduce to delayed copies, with distinct
  tensors per component so a swapped operand shows;
- p_v and v_p: finite differences of the monopole kernel f(t - r/c)/(4 pi r)
  in the source position along n_s and in the receiver position along n_r;
- CPML: a run with a 10-cell CPML against the same run on a grid large
  enough that nothing returns from its walls in the window.
"""
import math

import numpy as np
import pytest

import oracle as O

C_AIR, RHO_AIR = 343.0, 1.21


def cfg1_domain(n=20, nt=40):
    fc = 2000.0
    dx = C_AIR / (2.5 * fc) / 10
    xmax = (n - 1) * dx / 2
    d0 = O.domain(1.0, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
    return O.domain((nt - 0.5) * d0.dt, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)


def _gather_case(n=20, nt=40, n_pts=10):
    d = cfg1_domain(n, nt)
    pts, nrm = O.fibonacci_sphere(n_pts, 5 * d.dx)
    tf = O.ricker_wavelet(d, 2000.0, True)
    txs = [(np.array([0.3, 0.2, 0.1]) * d.dx, [1, 0, 0, 0], tf)]
    txs += [(pts[i], np.r_[0, nrm[i]], np.zeros(d.nt)) for i in range(n_pts)]
    return d, O.transceivers_to_grid(d, txs)


def _scatter(G, x, nt):
    """out[o, t + k] += G[o, i, k - 1] * x[i, t] for every recorded sample
    t (0-based) and lag k = 1..K: the scatter form of the causal gather.
    G is (n_out, n_in, K), x (n_in, nt)."""
    n_out, n_in, K = G.shape
    out = np.zeros((n_out, nt + K + 1))
    for i in range(n_in):
        for t in range(nt):
            for k in range(1, K + 1):
                out[:, t + k] += G[:, i, k - 1] * x[i, t]
    return out[:, :nt]


def test_gf_gather_pv_from_pv_equals_scatter_form():
    """extrapolation.jl:101-137,202-218: psum = sum pv p + pp vn goes to
    vn_src, vsum = sum vv p + vp vn goes to p_src, at step it + 1; the
    tensors are (nt_gf, #iRec, #iSrc) in Julia, numpy (n_out, n_in, K)."""
    d, tx = _gather_case()
    rng = np.random.default_rng(8)
    iRec, iSrc = np.array([2, 3, 4, 5, 6]), np.array([7, 8, 9])
    K = 9
    pp, pv, vp, vv = (rng.standard_normal((len(iSrc), len(iRec), K)) * 1e-3 for _ in range(4))
    s = O.RefSolver(d, npml=3, rcoef=1e-4, fc=2000.0, txs=tx)
    s.set_gf_map("pv_from_pv", (pp, pv, vp, vv), iRec, iSrc, writeback=True)
    s.run()
    p = s.p_rec[iRec - 1]
    vn = s.vn_rec[iRec - 1]
    to_vn = _scatter(pv, p, d.nt) + _scatter(pp, vn, d.nt)
    to_p = _scatter(vv, p, d.nt) + _scatter(vp, vn, d.nt)
    np.testing.assert_allclose(s.tx["vn_src"][iSrc - 1, 1:], to_vn[:, 1:], rtol=1e-9, atol=1e-25)
    np.testing.assert_allclose(s.tx["src"][0][iSrc - 1, 1:], to_p[:, 1:], rtol=1e-9, atol=1e-25)
    assert np.abs(to_vn[:, 1:]).max() > 0 and np.abs(to_p[:, 1:]).max() > 0
    # every component contributes: a swapped pair gives another answer
    alt = _scatter(pp, p, d.nt) + _scatter(pv, vn, d.nt)
    assert np.abs(alt - to_vn).max() > 1e-3 * np.abs(to_vn).max()


def test_gf_gather_p_from_pin_equals_scatter_form():
    """extrapolation.jl:172-192,238-249: pos and neg are (nt_gf, #iRec,
    #iSrc) in Julia (numpy (n_out, n_in, K)); p_src = pos - neg and
    vn_src = (pos + neg) / z0 at step it + 1, pin = (p - z0 vn) / 2."""
    d, tx = _gather_case()
    rng = np.random.default_rng(9)
    iRec, iSrc = np.array([2, 3, 4, 5, 6]), np.array([8, 10, 11])
    K = 7
    pos = rng.standard_normal((len(iSrc), len(iRec), K)) * 1e-3
    neg = rng.standard_normal((len(iSrc), len(iRec), K)) * 1e-3
    s = O.RefSolver(d, npml=3, rcoef=1e-4, fc=2000.0, txs=tx)
    s.set_gf_map("p_from_pin", (pos, neg), iRec, iSrc, writeback=True)
    s.run()
    pin = (s.p_rec[iRec - 1] - s.vn_rec[iRec - 1] * d.z0) / 2
    a, b = _scatter(pos, pin, d.nt), _scatter(neg, pin, d.nt)
    np.testing.assert_allclose(s.tx["src"][0][iSrc - 1, 1:], (a - b)[:, 1:], rtol=1e-9, atol=1e-25)
    np.testing.assert_allclose(s.tx["vn_src"][iSrc - 1, 1:], ((a + b) / d.z0)[:, 1:], rtol=1e-9, atol=1e-28)
    assert np.abs(a - b).max() > 0


@pytest.mark.parametrize("component", ["pp", "pv", "vp", "vv"])
def test_gf_pv_from_pv_single_tap_is_a_delayed_copy(component):
    """One tap at lag L of one component copies one record channel to one
    source channel L - 1 steps later: pv maps p to vn_src, pp maps vn to
    vn_src, vv maps p to p_src, vp maps vn to p_src."""
    d, tx = _gather_case(n=18, nt=30, n_pts=6)
    L = 4
    G = {c: np.zeros((1, 1, 8)) for c in ("pp", "pv", "vp", "vv")}
    G[component][0, 0, L - 1] = 1.0
    s = O.RefSolver(d, npml=2, rcoef=1e-4, fc=2000.0, txs=tx)
    s.set_gf_map("pv_from_pv", (G["pp"], G["pv"], G["vp"], G["vv"]), [3], [6], writeback=True)
    s.run()
    rec = {"p": s.p_rec[2], "vn": s.vn_rec[2]}[{"pp": "vn", "pv": "p", "vp": "vn", "vv": "p"}[component]]
    dst = s.tx["vn_src"][5] if component in ("pp", "pv") else s.tx["src"][0, 5]
    other = s.tx["src"][0, 5] if component in ("pp", "pv") else s.tx["vn_src"][5]
    assert np.abs(rec).max() > 0
    np.testing.assert_allclose(dst[L:], rec[: d.nt - L], rtol=1e-12, atol=1e-300)
    assert not other.any()


@pytest.mark.parametrize("component", ["pos", "neg"])
def test_gf_p_from_pin_single_tap_is_a_delayed_copy(component):
    d, tx = _gather_case(n=18, nt=30, n_pts=6)
    L = 3
    g = np.zeros((1, 1, 6))
    g[0, 0, L - 1] = 1.0
    z = np.zeros_like(g)
    s = O.RefSolver(d, npml=2, rcoef=1e-4, fc=2000.0, txs=tx)
    s.set_gf_map("p_from_pin", (g, z) if component == "pos" else (z, g), [2], [7], writeback=True)
    s.run()
    pin = (s.p_rec[1] - s.vn_rec[1] * d.z0) / 2
    sign = 1.0 if component == "pos" else -1.0
    assert np.abs(pin).max() > 0
    np.testing.assert_allclose(s.tx["src"][0, 6, L:], sign * pin[: d.nt - L], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(s.tx["vn_src"][6, L:], pin[: d.nt - L] / d.z0, rtol=1e-12, atol=1e-300)


# ------------------------------------------------------------- analytic KH

def _monopole(sig, c0):
    f = lambda tau: np.exp(-tau ** 2 / (2 * sig ** 2)) / (math.sqrt(2 * math.pi) * sig)
    return lambda xr, xs, t: f(t - np.linalg.norm(xr - xs) / c0) / (4 * math.pi * np.linalg.norm(xr - xs))


def _kh_case():
    op, on = O.fibonacci_sphere(6, 0.3)      # outer = source side (s_pt, n_s)
    ip, inn = O.fibonacci_sphere(4, 0.12)    # inner = receiver side (r_pt, n_r)
    sig = math.sqrt(math.log(2)) / (2 * math.pi * 3000.0)
    t = np.arange(700) * 4e-6
    return op, on, ip, inn, sig, t


def test_kh_p_v_is_the_source_normal_derivative_of_the_monopole():
    """kh_kernels.jl:140-143: p_v = cos(theta_s) (f/(4 pi r^2) + f'/(4 pi r c))
    is n_s . grad_s of g = f(t - r/c)/(4 pi r), with r = |x_r - x_s|."""
    op, on, ip, inn, sig, t = _kh_case()
    c0 = C_AIR
    K = O.eval_one_kernel("p_v", op, on, ip, inn, t, sig, c0, RHO_AIR)
    g = _monopole(sig, c0)
    h = 1e-6
    for ir in range(len(ip)):
        for js in range(len(op)):
            fd = (g(ip[ir], op[js] + h * on[js], t) - g(ip[ir], op[js] - h * on[js], t)) / (2 * h)
            np.testing.assert_allclose(K[ir, js], fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
            assert np.abs(fd).max() > 0


def test_kh_v_p_is_minus_the_receiver_normal_derivative_over_rho():
    """kh_kernels.jl:145-150: v_p = (cos(theta_r)/rho)(f/(4 pi r^2) +
    f'/(4 pi r c)) = -(1/rho) n_r . grad_r of the same monopole kernel."""
    op, on, ip, inn, sig, t = _kh_case()
    c0 = C_AIR
    K = O.eval_one_kernel("v_p", op, on, ip, inn, t, sig, c0, RHO_AIR)
    g = _monopole(sig, c0)
    h = 1e-6
    for ir in range(len(ip)):
        for js in range(len(op)):
            fd = -(g(ip[ir] + h * inn[ir], op[js], t) - g(ip[ir] - h * inn[ir], op[js], t)) / (2 * h) / RHO_AIR
            np.testing.assert_allclose(K[ir, js], fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())


# ------------------------------------------------------------ CPML in the step

def _cpml_run(n, npml, nt=220, plant=None):
    fc = 2000.0
    dx = C_AIR / (2.5 * fc) / 10
    xm = (n - 1) * dx / 2
    d0 = O.domain(1.0, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)
    d = O.domain((nt - 0.5) * d0.dt, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)
    tf = O.ricker_wavelet(d, 4000.0, injection=False)
    # receivers 3 cells inside each face of the 41^3 grid's CPML, and two off-axis
    recs = [np.array(v) * dx for v in ([7, 0, 0], [-7, 1, 0], [0, 7.5, 0], [0, -7, 0.5], [0, 0, 7], [0.5, 0, -7],
                                       [5, 5, 5], [-5, 4, -5])]
    txs = [(np.array([0.3, -0.2, 0.1]) * dx, [1, 0, 0, 0], tf)] + [(r, [0, 1, 0, 0], np.zeros(d.nt)) for r in recs]
    s = O.RefSolver(d, npml=npml, rcoef=1e-4, fc=fc, txs=O.transceivers_to_grid(d, txs))
    if plant == "profile_reversed":
        s.a, s.b = np.ascontiguousarray(s.a[:, ::-1]), np.ascontiguousarray(s.b[:, ::-1])
        s.ah, s.bh = np.ascontiguousarray(s.ah[:, ::-1]), np.ascontiguousarray(s.bh[:, ::-1])
    elif plant == "slab_index_shifted":
        s.a, s.b = np.ascontiguousarray(np.roll(s.a, 1, axis=1)), np.ascontiguousarray(np.roll(s.b, 1, axis=1))
    elif plant == "no_a_term":
        s.a, s.ah = np.zeros_like(s.a), np.zeros_like(s.ah)
    s.run()
    return s.p_rec[1:].copy()


@pytest.fixture(scope="module")
def unbounded():
    # 121^3 without CPML: from the source, the nearest wall is 60 cells away,
    # more than half the 220-step window's travel (220 c dt = 63.5 dx) plus
    # the receivers' 7 cells, so nothing reflected reaches them
    return _cpml_run(121, 0)


def test_cpml_matches_an_unbounded_domain(unbounded):
    small = _cpml_run(41, 10)
    e = np.linalg.norm(small - unbounded) / np.linalg.norm(unbounded)
    per = np.linalg.norm(small - unbounded, axis=1) / np.linalg.norm(unbounded, axis=1)
    assert e < 0.03 and per.max() < 0.06, (e, per)
    rigid = _cpml_run(41, 0)
    assert np.linalg.norm(rigid - unbounded) / np.linalg.norm(unbounded) > 0.3  # what the CPML removes


@pytest.mark.parametrize("plant", ["profile_reversed", "slab_index_shifted", "no_a_term"])
def test_cpml_mistakes_planted_in_the_oracle_fail_the_bar(unbounded, plant):
    small = _cpml_run(41, 10, plant=plant)
    assert np.linalg.norm(small - unbounded) / np.linalg.norm(unbounded) > 0.1
