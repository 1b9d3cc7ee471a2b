"""The GF loaders' scaling chain (src/greens_io.jl:149-224,224-477,536-617)
pinned by its closed forms: C_p and C_v (C_v/C_p = -rho c^2), the
factorisation of the production kernels (docs/extrapolation_conventions.md
:91-97,144-146)
    gf[:v][:pin] = (dS_i dS_o e C_p) z0   (v_p + c v_v)
    gf[:p][:pin] = (dS_i dS_o e C_p) / z0 (p_p + c p_v),
the two-way form's relation to it, the split form's definition, the MDD
loader's linear resample, and the loaders' argument checks.  The cubic
resample is the oracle's restatement here (the device's is checked against
it in test_gpu_greens_io.py)."""
import math

import numpy as np
import pytest

import adfdtd as A
import oracle as O
from cases import domain_cells

H = 0.00686
CFG = {"surfaces": {"nPoints_outer": 6, "radius_outer": 0.12, "nPoints_inner": 6, "radius_inner": 0.05}}


def np_cubic(a, t_in0, dt_in, t_out0, dt_out, n_out, scale):
    flat = np.asarray(a, np.float64).reshape(-1, a.shape[-1])
    out = O.resample_time(flat, t_in0, dt_in, t_out0, dt_out, n_out, scale_amplitude=False) * scale
    return out.reshape(a.shape[:-1] + (n_out,))


def gf_file(dom, nt_save=50, dt_save_steps=3, seed=0, n=6, units="physical"):
    """A reference GF file in physical units: random kernels (n_src, n_rec,
    nt) sampled every dt_save_steps FDTD steps of `dom`."""
    rng = np.random.default_rng(seed)
    dt_save = dt_save_steps * dom.dt
    f = {c: rng.standard_normal((n, n, nt_save)).astype(np.float32) for c in ("p_p", "p_v", "v_p", "v_v")}
    f.update(t=np.arange(nt_save) * dt_save, dom_dt=dom.dt, dom_dx=dom.dx, backend="HIP", units_convention=units)
    return f


def test_area_and_unit_factors_closed_forms():
    d = domain_cells(30, H, 100)
    assert A.fib_sphere_area(0.05, 200) == pytest.approx(4 * math.pi * 0.0025 / 200, rel=1e-15)
    dt_save = 3 * d.dt
    C_p, C_v = A.gf_unit_factors(d, dt_save)
    assert C_p == pytest.approx(d.dx ** 3 * dt_save / (d.dt ** 2 * d.c0 ** 2), rel=1e-15)
    assert C_v / C_p == pytest.approx(-d.r0 * d.c0 ** 2, rel=1e-14)  # continuity coupling of a momentum source
    # the loader's factors reduce to these for GFs generated on the same domain
    c_p, c_v, dx = A.scaling_factors(d, d.dx, d.dt, dt_save)
    assert (c_p, c_v, dx) == (C_p, C_v, d.dx)
    # a file without dom_dx takes dx from cfg["greens"]["grid"]
    c_p2, _, dx2 = A.scaling_factors(d, float("nan"), d.dt, dt_save, {"greens": {"grid": {"n": 61, "xmax": 0.3}}})
    assert dx2 == 2 * 0.3 / 60 and c_p2 == pytest.approx(dx2 ** 3 * dt_save / (d.dt ** 2 * d.c0 ** 2), rel=1e-15)
    with pytest.raises(ValueError, match="dom_dx"):
        A.scaling_factors(d, float("nan"), d.dt, dt_save, {})


def test_physical_and_raw_units_are_inverse():
    d = domain_cells(30, H, 100)
    C_p, C_v = A.gf_unit_factors(d, 3 * d.dt)
    f = gf_file(d)
    back = A.gf_to_raw_fdtd(A.gf_to_physical(f, C_p, C_v), C_p, C_v)
    for c in ("p_p", "p_v", "v_p", "v_v"):
        np.testing.assert_allclose(back[c], f[c], rtol=3e-7, atol=0)
    assert back["t"] is f["t"]


def _chain(d, f, kind, dtype=np.float64, tmax=0.0):
    return A.load_gf_ref(f, d, CFG, extrap_type=kind, dtype=dtype, resample=np_cubic, tmax=tmax)


def _resampled_physical(d, f, tmax=0.0):
    """R(kernel): the physical kernels on 0:dt:t_end, with the amplitude
    factor dt / dt_save."""
    t = f["t"]
    n = t.shape[0]
    t_end = tmax if tmax > 0 else t[-1]
    n_out = int(np.floor(t_end / d.dt + 1e-9)) + 1
    step_in = (t[-1] - t[0]) / (n - 1)
    return {c: np_cubic(f[c].astype(np.float64), t[0], step_in, 0.0, d.dt, n_out, d.dt / step_in)
            for c in ("p_p", "p_v", "v_p", "v_v")}


def test_pin_kernels_factor_as_documented():
    d = domain_cells(30, H, 200)
    f = gf_file(d)
    g = _chain(d, f, "pv_from_pin")
    R = _resampled_physical(d, f)
    s = CFG["surfaces"]
    dS = 4 * math.pi * s["radius_inner"] ** 2 / s["nPoints_inner"] * 4 * math.pi * s["radius_outer"] ** 2 / \
        s["nPoints_outer"]
    # e C_p = c^2 dt_gf^2 dt / dx^3: C_p cancels out of the product
    eCp = d.c0 ** 2 * f["dom_dt"] ** 2 * d.dt / d.dx ** 3
    c = d.c0
    v_expect = dS * eCp * d.z0 * (R["v_p"] + c * R["v_v"])
    p_expect = dS * eCp / d.z0 * (R["p_p"] + c * R["p_v"])
    assert g.v_pin.shape == v_expect.shape == (6, 6, R["p_p"].shape[-1])
    np.testing.assert_allclose(g.v_pin, v_expect, rtol=2e-6, atol=2e-6 * np.abs(v_expect).max())
    np.testing.assert_allclose(g.p_pin, p_expect, rtol=2e-6, atol=2e-6 * np.abs(p_expect).max())
    assert g.p_pin.dtype == np.float64  # widened by z0 (greens_io.jl:466-468)


def test_two_way_form_and_split_form_against_the_pin_form():
    d = domain_cells(30, H, 200)
    f = gf_file(d, seed=3)
    pin = _chain(d, f, "pv_from_pin")
    pv = _chain(d, f, "pv_from_pv")
    z0 = d.z0
    # pv_from_pv scales each kernel by -(z0 ratio) dS e: the pin kernels are
    # -pp/z0 + pv and -vp/z0 + vv of the two-way kernels
    np.testing.assert_allclose(pin.p_pin, -pv.pp / z0 + pv.pv, rtol=1e-12,
                               atol=1e-12 * np.abs(pin.p_pin).max())
    np.testing.assert_allclose(pin.v_pin, -pv.vp / z0 + pv.vv, rtol=1e-12,
                               atol=1e-12 * np.abs(pin.v_pin).max())
    # the split form skips the impedance conversion (greens_io.jl:424-445
    # convert only for pv_from_pv and pv_from_pin): its p_pin is z0 times
    # the pin form's, its v_pin the pin form's over z0
    sp = _chain(d, f, "p_from_pin")
    np.testing.assert_allclose(sp.pos + sp.neg, pin.p_pin * z0, rtol=1e-12,
                               atol=1e-12 * np.abs(pin.p_pin * z0).max())
    np.testing.assert_allclose(sp.pos - sp.neg, pin.v_pin, rtol=1e-12, atol=1e-12 * np.abs(pin.v_pin).max())


def test_scattered_subtracts_the_homogeneous_companion_and_tmax_cuts():
    d = domain_cells(30, H, 200)
    het, hom = gf_file(d, seed=1), gf_file(d, seed=2)
    diff = dict(het, **{c: het[c] - hom[c] for c in ("p_p", "p_v", "v_p", "v_v")})
    a = A.load_gf_ref(het, d, CFG, gf_of="scattered", homogeneous=hom, dtype=np.float64, resample=np_cubic)
    b = A.load_gf_ref(diff, d, CFG, dtype=np.float64, resample=np_cubic)
    np.testing.assert_array_equal(a.p_pin, b.p_pin)
    t_end = 40 * d.dt
    c = A.load_gf_ref(het, d, CFG, dtype=np.float64, resample=np_cubic, tmax=t_end)
    assert c.p_pin.shape[-1] == 41


def test_loader_argument_checks():
    d = domain_cells(30, H, 200)
    f = gf_file(d)
    with pytest.raises(ValueError, match="extrap_type"):
        A.load_gf_ref(f, d, CFG, extrap_type="pv")
    with pytest.raises(ValueError, match="homogeneous"):
        A.load_gf_ref(f, d, CFG, gf_of="scattered")
    with pytest.raises(ValueError, match="units_convention"):
        A.load_gf_ref(gf_file(d, units="fdtd"), d, CFG)
    with pytest.raises(ValueError, match="empirical_scale"):
        A.load_gf_mdd({"p_p": f["p_p"], "v_p": f["v_p"], "t": f["t"], "units_convention": "physical"}, d, CFG)


def test_mdd_loader_linear_resample_and_scales():
    d = domain_cells(30, H, 200)
    f = gf_file(d, seed=5)
    mdd = {"p_p": f["p_p"], "v_p": f["v_p"], "t": f["t"], "units_convention": "physical", "empirical_scale": 1.25}
    g = A.load_gf_mdd(mdd, d, CFG, dtype=np.float64)
    t = f["t"]
    tin = np.linspace(0.0, t[-1], t.shape[0])  # range(0, t[end]; length = nt)
    n_out = int(np.floor(t[-1] / d.dt + 1e-9)) + 1
    tout = np.arange(n_out) * d.dt
    scale = d.dt / (t[-1] / (t.shape[0] - 1))
    s = CFG["surfaces"]
    dS = A.fib_sphere_area(s["radius_inner"], s["nPoints_inner"]) * A.fib_sphere_area(s["radius_outer"],
                                                                                        s["nPoints_outer"])
    es = np.float32(1.25)
    for got, key, zf in ((g.p_pin, "p_p", 1 / d.z0), (g.v_pin, "v_p", d.z0)):
        src = (f[key] * es).astype(np.float64)
        lin = np.stack([np.stack([np.interp(tout, tin, src[a, b]) for b in range(6)]) for a in range(6)])
        np.testing.assert_allclose(got, lin * scale * dS * zf, rtol=1e-12, atol=1e-12 * np.abs(got).max())


def test_npz_container_round_trip(tmp_path):
    d = domain_cells(30, H, 200)
    f = gf_file(d)
    p = str(tmp_path / "gf.npz")
    A.write_gf_file(p, f)
    r = A.read_gf_file(p)
    assert r["units_convention"] == "physical" and r["dom_dt"] == d.dt and r["backend"] == "HIP"
    for c in ("p_p", "p_v", "v_p", "v_v", "t"):
        assert np.array_equal(r[c], f[c])
    with pytest.raises(RuntimeError, match="h5py"):
        A.read_gf_file(str(tmp_path / "gf.h5"))
