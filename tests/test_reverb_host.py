"""Host side of the reverb stage (src/reverb.jl): the R2 illumination
sphere (src/illumination.jl:69-97), Julia's Float64 step ranges (the output
time axis `range(0.0, tmax; step = 1/fs_out)`), and the sweep's argument
checks and stop file, which run without a GPU."""
import math
import os

import numpy as np
import pytest

import adfdtd as A
from cases import domain_cells

PLASTIC = 1.3247179572447460259609088544780973407344040569017333645340150503


def test_r2_sphere_closed_form_and_prefix_property():
    # the plastic number is the real root of x^3 = x + 1
    phi = (9 + math.sqrt(69)) ** (1 / 3) / np.cbrt(18.0) + (9 - math.sqrt(69)) ** (1 / 3) / np.cbrt(18.0)
    assert abs(phi - PLASTIC) < 4e-16 and abs(phi ** 3 - phi - 1) < 1e-15
    pts, nrm = A.r2_sphere(50, 0.45)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-15)
    assert np.array_equal(pts, 0.45 * nrm)
    # point i uses u = frac(i/phi), v = frac(i/phi^2) on the Lambert map
    for i in (1, 2, 7, 50):
        u, v = math.fmod(i / PLASTIC, 1.0), math.fmod(i / PLASTIC ** 2, 1.0)
        z = 1 - 2 * v
        r = math.sqrt(1 - z * z)
        assert np.allclose(nrm[i - 1], [r * math.cos(2 * math.pi * u), r * math.sin(2 * math.pi * u), z], atol=1e-12)
    # position i depends only on i: growing n appends points
    p7, _ = A.r2_sphere(7, 0.45)
    assert np.array_equal(p7, pts[:7])
    # low discrepancy: the z (Lambert) coordinate covers [-1, 1] evenly
    h, _ = np.histogram(A.r2_sphere(1000)[1][:, 2], bins=10, range=(-1, 1))
    assert h.min() >= 95 and h.max() <= 105
    with pytest.raises(ValueError):
        A.r2_sphere(0)


def test_julia_step_ranges():
    # Julia's documented behaviour: 0.0:0.1:0.3 has four elements ending at
    # exactly 0.3, and 0.1:0.1:0.3 == [0.1, 0.2, 0.3] (not 0.1 + 2*0.1)
    r = A.julia_step_range(0.0, 0.1, 0.3)
    assert r.tolist() == [0.0, 0.1, 0.2, 0.3]
    assert A.julia_step_range(0.1, 0.1, 0.3).tolist() == [0.1, 0.2, 0.3]
    # a stop on the grid of the step counts, in exact rationals
    for dur, fs in ((5e-3, 60e3), (300e-4, 60e3), (4e-4, 25e3), (1e-3, 1e5)):
        r = A.julia_step_range(0.0, 1.0 / fs, dur)
        n = round(dur * fs)
        assert len(r) == n + 1 and r[-1] == dur, (dur, fs, len(r))
        # elements k/fs, correctly rounded (the exact rationals k * step)
        assert np.allclose(r, np.arange(n + 1) / fs, rtol=2.3e-16, atol=0)
    # off-grid stops: floor((stop - start) / step) + 1 elements
    for dur, fs in ((5.0001234e-3, 60e3), (2.9e-3, 12345.0), (4e-4, 24e3)):
        r = A.julia_step_range(0.0, 1.0 / fs, dur)
        assert len(r) == math.floor(dur * fs) + 1 and r[-1] <= dur
    assert len(A.julia_step_range(0.0, 1.0, -1.0)) == 0
    assert len(A.julia_step_range(2.0, 1.0, 2.0)) == 1


def _cfg(**rev):
    h = 0.00686
    return {"reverb": {"radius_ill": 10 * h, "fc_source": 2000.0, "bandwidth_factor": 2.5, **rev},
            "pml": {"n": 0, "rcoef": 1e-4, "fc": 2000.0},
            "surfaces": {"nPoints_outer": 4, "radius_outer": 7 * h, "nPoints_inner": 5, "radius_inner": 4 * h,
                         "sphere_radius": 2 * h}}


def test_build_reverb_data_argument_checks_and_stop_file(tmp_path):
    d = domain_cells(28, 0.00686, 60)
    cfg = _cfg()
    bad = _cfg()
    del bad["reverb"]["bandwidth_factor"]
    with pytest.raises(ValueError, match="bandwidth_factor"):
        A.build_reverb_data(d, bad, scatterer="sphere", n_ill=2)
    withpml = _cfg()
    withpml["pml"]["n"] = 4
    with pytest.raises(ValueError, match="resume"):
        A.build_reverb_data(d, withpml, scatterer="sphere", n_ill=2, n_saved=10)
    with pytest.raises(ValueError, match="exceeds"):
        A.build_reverb_data(d, cfg, scatterer="sphere", n_ill=2, n_saved=d.nt + 1)
    # a stop file that exists before the sweep: nothing is computed
    stop = tmp_path / "STOP"
    stop.write_text("")
    seen = []
    out = A.build_reverb_data(d, cfg, scatterer="sphere", n_ill=3, stopfile=str(stop),
                              on_source_complete=lambda i, s: seen.append(i))
    fs_out = 2 * 2.5 * 2000.0
    nt_ill = len(A.julia_step_range(0.0, 1 / fs_out, d.tmax))
    assert seen == [] and out["outer_p"].shape == (3, 4, nt_ill) and out["inner_vnz"].shape == (3, 5, nt_ill)
    assert not np.any(out["outer_p"]) and out["t"].dtype == np.float32
    assert np.array_equal(out["ill_positions"], A.r2_sphere(3, 10 * 0.00686)[0].astype(np.float32))
    # prefill shapes are checked
    with pytest.raises(ValueError, match="prefill"):
        A.build_reverb_data(d, cfg, scatterer="sphere", n_ill=3, stopfile=str(stop),
                            prefill={k: np.zeros((2, 4, nt_ill + 1), np.float32)
                                     for k in ("outer_p", "outer_vnz", "inner_p", "inner_vnz")})
