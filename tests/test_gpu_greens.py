"""impulsive_gfs on the device (src/greens.jl:48-234): the per-run
post-processing (adf_solver_gf_resample: self-record dropped, denoise,
cubic resample) against the oracle's restatement on the same traces, and
the whole source sweep against the oracle pipeline (oracle FDTD per source
and type, oracle denoise + resample).

Bars: the device post-processing equals the oracle's float64 computation
rounded to float32 (same operations in the same order); end to end, the
Float64 run matches to float32 rounding of the float64 result (its FDTD is
bit-exact, see test_gpu_parity.py) and the fast float32 run to 1e-5
relative L2 of the Float64 oracle per GF tensor."""
import numpy as np
import pytest

import adfdtd as A
import oracle as O
from adfdtd import _lib as L
from cases import build, domain_cells, oracle_solver, rel_l2

pytestmark = pytest.mark.gpu


def _oracle_post(rec_p, rec_vn, d, timeshift, nt_gf):
    """Oracle tail of impulsive_gfs for one run: traces (ntx, nt), column 0
    (the source's self-record) dropped."""
    tmax = d.tmax
    dt_in, dt_out = tmax / (d.nt - 1), tmax / (nt_gf - 1)
    out = []
    for tr in (rec_p, rec_vn):
        a = O.gf_denoise(np.asarray(tr[1:], np.float64))
        out.append(O.resample_time(a, -timeshift, dt_in, 0.0, dt_out, nt_gf).astype(np.float32))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gf_resample_equals_oracle_on_the_same_traces(dtype):
    d, txs, cpml, mask = build(n=32, nt=120, npml=5, n_rec=12, rec_radius_cells=7, src_cells=(1.3, -0.4, 0.6))
    g = A.transceivers_to_grid(d, txs, dtype)
    s = A.Solver(d, A.dtype_code(dtype, "fast"), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, dtype))
    s.set_transceivers(g)
    s.run(1, d.nt)
    _, timeshift = A.impulsive_wavelet(d, 2000.0)
    for nt_gf in (47, 120, 233):
        gp, gv = s.gf_resample(1, 12, timeshift, nt_gf)
        op, ov = _oracle_post(s.get_trace(L.ADF_REC_P), s.get_trace(L.ADF_VN_REC), d, timeshift, nt_gf)
        assert np.array_equal(gp, op), np.max(np.abs(gp - op))
        assert np.array_equal(gv, ov), np.max(np.abs(gv - ov))
    # a subset of the transceivers, starting at the source's own record
    gp, gv = s.gf_resample(0, 5, 0.0, 30)
    z = np.zeros((1, d.nt))
    op, ov = _oracle_post(np.vstack([z, s.get_trace(L.ADF_REC_P)[:5]]), np.vstack([z, s.get_trace(L.ADF_VN_REC)[:5]]),
                          d, 0.0, 30)
    assert np.array_equal(gp, op) and np.array_equal(gv, ov)
    with pytest.raises(L.AdfError):
        s.gf_resample(1, 13, 0.0, 10)  # past the last transceiver


def _cfg(n_outer, n_inner, dx, nt_save):
    return {"greens": {"f_3db_hz": 2000.0, "nt_save": nt_save},
            "pml": {"n": 5, "rcoef": 1e-4, "fc": 2000.0},
            "surfaces": {"nPoints_outer": n_outer, "radius_outer": 7 * dx, "nPoints_inner": n_inner,
                         "radius_inner": 4 * dx, "sphere_radius": 2 * dx}}


def _oracle_gfs(d, cfg, scatterer):
    """src/greens.jl:143-191 on the oracle (Float64)."""
    params = A.scatterer_params(scatterer, cfg)
    mask = A.build_update_mask(d, scatterer, params)
    srctf, timeshift = A.impulsive_wavelet(d, cfg["greens"]["f_3db_hz"])
    pts = A.illumination_points(cfg)
    rec = [A.make_transceiver(d, pts["inner"][i], np.r_[0.0, pts["inner_normals"][i]], np.zeros(d.nt))
           for i in range(len(pts["inner"]))]
    cpml = A.make_cpml(d, cfg["pml"]["n"], cfg["pml"]["rcoef"], cfg["pml"]["fc"])
    nt_gf = cfg["greens"]["nt_save"]
    n_src, n_rec = len(pts["outer"]), len(rec)
    out = {k: np.zeros((n_src, n_rec, nt_gf), np.float32) for k in ("p_p", "p_v", "v_p", "v_v")}
    for i in range(n_src):
        for kind in ("p", "v"):
            dr = [1.0, 0.0, 0.0, 0.0] if kind == "p" else np.r_[0.0, pts["outer_normals"][i]]
            txs = [A.make_transceiver(d, pts["outer"][i], dr, srctf)] + rec
            r = oracle_solver(O, d, txs, cpml, mask, "f64").run()
            gp, gv = _oracle_post(r.tx["rec"][0], r.tx["vn_rec"], d, timeshift, nt_gf)
            out["p_" + kind][i] = gp
            out["v_" + kind][i] = gv
    return out


def test_impulsive_gfs_matches_oracle_pipeline():
    h = 0.00686
    d = domain_cells(28, h, 140)
    cfg = _cfg(3, 6, d.dx, 70)
    ref = _oracle_gfs(d, cfg, "sphere")
    g64 = A.impulsive_gfs(d, cfg, scatterer="sphere", dtype=np.float64, sources_per_batch=2)
    g32 = A.impulsive_gfs(d, cfg, scatterer="sphere", dtype=np.float32)
    assert g64["p_p"].shape == (3, 6, 70)
    for k in ("p_p", "p_v", "v_p", "v_v"):
        assert np.array_equal(g64[k], ref[k]), (k, np.max(np.abs(g64[k] - ref[k])))
        assert rel_l2(g32[k], ref[k]) <= 1e-5, (k, rel_l2(g32[k], ref[k]))
        assert np.abs(ref[k]).max() > 0
    assert np.allclose(g64["t"], np.linspace(0, d.tmax, 70).astype(np.float32))
    # init: sources already present are skipped, the rest are computed
    part = {k: g64[k].copy() for k in ("p_p", "p_v", "v_p", "v_v")}
    for k in part:
        part[k][1] = 0
    seen = []
    g2 = A.impulsive_gfs(d, cfg, scatterer="sphere", dtype=np.float64, init=part,
                         on_source_complete=lambda i, st: seen.append(i))
    assert seen == [2]
    for k in part:
        assert np.array_equal(g2[k], g64[k])
