"""The reverb stage on the device (src/reverb.jl:72-452): the device
resample (adf_resample_f32) against the oracle's restatement on the same
float32 traces, and the whole sweep against the oracle pipeline (oracle
FDTD per illumination source, the reference's Float32 cast with vn times
z0, oracle cubic resample without amplitude scaling), including resume
from saved fields, prefill and field snapshots.

Bars: the device resample equals the oracle's float64 computation rounded
to float32; end to end, the Float64 sweep equals the oracle pipeline bit
for bit (its FDTD is bit-exact, see test_gpu_parity.py) and the fast
float32 sweep is within 1e-5 relative L2 of it per output tensor."""
import numpy as np
import pytest

import adfdtd as A
import oracle as O
from adfdtd import _lib as L
from cases import domain_cells, oracle_solver, rel_l2

pytestmark = pytest.mark.gpu

H = 0.00686
KEYS = ("outer_p", "outer_vnz", "inner_p", "inner_vnz")


def _cfg():
    return {"reverb": {"radius_ill": 10 * H, "fc_source": 2000.0, "bandwidth_factor": 2.5},
            "pml": {"n": 0, "rcoef": 1e-4, "fc": 2000.0},
            "surfaces": {"nPoints_outer": 4, "radius_outer": 7 * H, "nPoints_inner": 5, "radius_inner": 4 * H,
                         "sphere_radius": 2 * H}}


def test_resample_traces_equals_oracle():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((7, 133)).astype(np.float32)
    a[2] *= 1e-3
    a[3, ::3] = 1e-9
    cases = [(0.0, 1e-5, 0.0, 3.3e-5, 41, 1.0, False), (-2e-4, 1e-5, 0.0, 0.7e-5, 200, 0.7, True),
             (0.0, 1e-5, 0.0, 1e-5, 133, 1.0, False), (1e-4, 2e-6, 0.0, 1e-6, 20, 1.0, False)]
    for t_in0, dt_in, t_out0, dt_out, n_out, amp, den in cases:
        got = A.resample_traces(a, t_in0, dt_in, t_out0, dt_out, n_out, amp_scale=amp, denoise=den)
        src = O.gf_denoise(a) if den else a.astype(np.float64)
        exp = O.resample_time(src, t_in0, dt_in, t_out0, dt_out, n_out, scale_amplitude=False)
        exp = (exp * amp if amp != 1.0 else exp).astype(np.float32)
        assert got.shape == (7, n_out) and np.array_equal(got, exp), np.max(np.abs(got - exp))
    one = A.resample_traces(a[:, :1], 0.0, 1.0, 0.0, 0.5, 4)
    assert np.array_equal(one, np.repeat(a[:, :1], 4, axis=1))
    with pytest.raises(L.AdfError):
        A.resample_traces(a, 0.0, 0.0, 0.0, 1e-5, 10)


def _source(d, pt, tf):
    return A.make_transceiver(d, pt, [d.z0, 0.0, 0.0, 0.0], tf)


def _oracle_source(d, cfg, i, field=None, start=1, saved=None, n_saved=0, snap_steps=()):
    """One illumination source through the oracle: FDTD, cast, splice of
    saved rows; -> (full traces dict (n_rec, nt) float32, snapshots)."""
    pts = A.illumination_points(cfg)
    fc = cfg["reverb"]["fc_source"]
    ill, _ = A.r2_sphere(i, cfg["reverb"]["radius_ill"])
    z = np.zeros(d.nt)
    rec = [A.make_transceiver(d, pts["outer"][k], np.r_[0.0, pts["outer_normals"][k]], z) for k in range(4)]
    rec += [A.make_transceiver(d, pts["inner"][k], np.r_[0.0, pts["inner_normals"][k]], z) for k in range(5)]
    tf = np.zeros(d.nt) if saved is not None else A.ricker_wavelet(d, fc, True)
    txs = rec + [_source(d, ill[i - 1], tf)]
    mask = A.build_update_mask(d, "sphere", [2 * H])
    cpml = A.make_cpml(d, 0, 1e-4, fc)
    r = oracle_solver(O, d, txs, cpml, mask, "f64", field=field)
    snaps = {}
    cur = start
    for st in snap_steps:
        r.run(cur, st)
        snaps[st] = r.field.astype(np.float32).copy()
        cur = st + 1
    r.run(cur, d.nt)
    p = np.asarray(r.tx["rec"][0][:9]).astype(np.float32)
    vn = (np.asarray(r.tx["vn_rec"][:9]).astype(np.float32).astype(np.float64) * d.z0).astype(np.float32)
    full = {"outer_p": p[:4].copy(), "inner_p": p[4:].copy(), "outer_vnz": vn[:4].copy(), "inner_vnz": vn[4:].copy()}
    if saved is not None:
        for k in KEYS:
            full[k][:, :n_saved] = saved[k][:, :n_saved]
    return full, snaps, r.field


def _oracle_resample(d, full, fs_out):
    t_ill = A.julia_step_range(0.0, 1.0 / fs_out, d.tmax).astype(np.float32)
    n_out = len(t_ill)
    dt_out = float(t_ill[-1]) / (n_out - 1)
    return {k: O.resample_time(full[k].astype(np.float64), 0.0, d.tmax / (d.nt - 1), 0.0, dt_out, n_out,
                               scale_amplitude=False).astype(np.float32) for k in KEYS}


def test_build_reverb_data_matches_oracle_pipeline():
    d = domain_cells(28, H, 150)
    cfg = _cfg()
    fs_out = 2 * 2.5 * 2000.0
    ref, snaps_ref = {k: [] for k in KEYS}, None
    for i in (1, 2, 3):
        full, snaps, _ = _oracle_source(d, cfg, i, snap_steps=(40, 77) if i == 2 else ())
        if i == 2:
            snaps_ref = snaps
        rs = _oracle_resample(d, full, fs_out)
        for k in KEYS:
            ref[k].append(rs[k])
    ref = {k: np.stack(v) for k, v in ref.items()}
    seen, snaps = [], {}

    def done(i, st):
        seen.append(i)
        if st["snapshots"]:
            snaps[i] = st["snapshots"]

    g64 = A.build_reverb_data(d, cfg, scatterer="sphere", n_ill=3, dtype=np.float64, sources_per_batch=2,
                              on_source_complete=done, snapshot_steps=[40, 77], snapshot_isrc={2})
    g32 = A.build_reverb_data(d, cfg, scatterer="sphere", n_ill=3)
    assert seen == [1, 2, 3] and list(snaps) == [2]
    for st in (40, 77):
        assert np.array_equal(snaps[2][st], snaps_ref[st])
    assert g64["outer_p"].shape == (3, 4, len(A.julia_step_range(0.0, 1 / fs_out, d.tmax)))
    for k in KEYS:
        assert np.abs(ref[k]).max() > 0
        assert np.array_equal(g64[k], ref[k]), (k, np.max(np.abs(g64[k] - ref[k])))
        assert rel_l2(g32[k], ref[k]) <= 1e-5, (k, rel_l2(g32[k], ref[k]))
    assert np.array_equal(g64["t"], A.julia_step_range(0.0, 1 / fs_out, d.tmax).astype(np.float32))
    # prefill: populated slots are copied and skipped
    seen.clear()
    g2 = A.build_reverb_data(d, cfg, scatterer="sphere", n_ill=3, dtype=np.float64,
                             prefill={k: g64[k][:2] for k in KEYS}, on_source_complete=lambda i, st: seen.append(i))
    assert seen == [3]
    for k in KEYS:
        assert np.array_equal(g2[k], g64[k])


def test_resume_extends_saved_sources():
    """Sources saved after a shorter run continue from their Float32 field
    with a zero wavelet (reverb.jl:268-283), their saved rows spliced in."""
    cfg = _cfg()
    ds, dl = domain_cells(28, H, 90), domain_cells(28, H, 150)
    assert ds.dt == dl.dt
    saved = {}
    A.build_reverb_data(ds, cfg, scatterer="sphere", n_ill=2, dtype=np.float64,
                        save_isrc_done=lambda i, st: saved.__setitem__(i, st))
    assert sorted(saved) == [1, 2] and saved[1]["field"].dtype == np.float32
    out = A.build_reverb_data(dl, cfg, scatterer="sphere", n_ill=3, dtype=np.float64, n_saved=ds.nt,
                              load_isrc_saved=lambda i: saved.get(i))
    fs_out = 2 * 2.5 * 2000.0
    for i in (1, 2, 3):
        if i in saved:
            full, _, _ = _oracle_source(dl, cfg, i, field=saved[i]["field"].astype(np.float64), start=ds.nt + 1,
                                        saved=saved[i], n_saved=ds.nt)
        else:
            full, _, _ = _oracle_source(dl, cfg, i)
        rs = _oracle_resample(dl, full, fs_out)
        for k in KEYS:
            assert np.array_equal(out[k][i - 1], rs[k]), (i, k)
    # a fully saved sweep (n_saved == nt) is only resampled
    seen = []
    again = A.build_reverb_data(ds, cfg, scatterer="sphere", n_ill=2, dtype=np.float64, n_saved=ds.nt,
                                load_isrc_saved=lambda i: saved.get(i), on_source_complete=lambda i, st: seen.append(i))
    first = A.build_reverb_data(ds, cfg, scatterer="sphere", n_ill=2, dtype=np.float64)
    assert seen == []
    for k in KEYS:
        assert np.array_equal(again[k], first[k])
