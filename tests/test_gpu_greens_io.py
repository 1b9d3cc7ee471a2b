"""The GF loaders on the device: load_gf_ref with the device's cubic
resample equals the host restatement within float32 rounding, and the
paper's chain runs end to end on the device: impulsive_gfs on the GPU,
its GFs to physical units (gf_unit_factors, gf_to_physical, as the
impulsive script saves them), an .npz container, load_gf_ref, and
build_hologram(RefMode) with that map, whose traces equal the same map run
on the oracle within 1e-5."""
import numpy as np
import pytest

import adfdtd as A
import oracle as O
from cases import domain_cells, oracle_domain, oracle_transceivers, rel_l2
from test_greens_io import CFG as CFG_SMALL, gf_file, np_cubic

pytestmark = pytest.mark.gpu

H = 0.00686


@pytest.mark.parametrize("kind", ["pv_from_pin", "pv_from_pv", "p_from_pin"])
def test_device_resample_in_the_loader(kind):
    d = domain_cells(30, H, 300)
    f = gf_file(d, nt_save=90, seed=4)
    dev = A.load_gf_ref(f, d, CFG_SMALL, extrap_type=kind, dtype=np.float64)
    host = A.load_gf_ref(f, d, CFG_SMALL, extrap_type=kind, dtype=np.float64, resample=np_cubic)
    _, td = dev.kind_and_tensors()
    _, th = host.kind_and_tensors()
    for a, b in zip(td, th):
        assert a.shape == b.shape and rel_l2(a, b) < 1e-6


def test_impulsive_gfs_to_hologram_chain(tmp_path):
    n, nt = 36, 120
    d = domain_cells(n, H, nt)
    cfg = {"greens": {"f_3db_hz": 2000.0, "nt_save": 60},
           "pml": {"n": 4, "rcoef": 1e-4, "fc": 2000.0},
           "surfaces": {"nPoints_outer": 8, "radius_outer": 9 * H, "nPoints_inner": 8, "radius_inner": 5 * H,
                        "cube_length": 3 * H}}
    gfs = A.impulsive_gfs(d, cfg, scatterer="cube", dtype=np.float32)
    dt_save = d.tmax / (cfg["greens"]["nt_save"] - 1)
    C_p, C_v = A.gf_unit_factors(d, dt_save)
    phys = A.gf_to_physical(gfs, C_p, C_v)
    path = str(tmp_path / "gf_ref.npz")
    A.write_gf_file(path, {**{k: phys[k] for k in ("p_p", "p_v", "v_p", "v_v", "t")}, "dom_dt": d.dt,
                           "dom_dx": d.dx, "backend": "HIP", "units_convention": "physical"})
    gf = A.load_gf_ref(A.read_gf_file(path), d, cfg, extrap_type="pv_from_pin", dtype=np.float32)
    assert gf.p_pin.shape == (8, 8, len(A.julia_step_range(0.0, d.dt, float(phys["t"][-1]))))
    va = A.build_update_mask(d, "cube", [3 * H])
    out = A.build_hologram(d, cfg, gf, scatterer="cube", implementation=A.RefMode(), fc=2000.0,
                           initial_distance=0.08, va=va, dtype=np.float32)
    g = out["txs_on_grid"]
    assert np.isfinite(out["field_final"]).all() and np.abs(g.src[0][8:16]).max() > 0
    # the same map on the oracle
    rd = oracle_domain(d)
    rectf = np.zeros(d.nt)
    ill = A.illumination_points(cfg)
    txs = [A.make_transceiver(d, ill["outer"][i], np.r_[0.0, ill["outer_normals"][i]], rectf) for i in range(8)]
    txs += [A.make_transceiver(d, ill["inner"][i], np.r_[0.0, ill["inner_normals"][i]], rectf) for i in range(8)]
    tx = oracle_transceivers(rd, d, txs, np.float64)
    field = O.initial_field(rd, 2000.0, 0.08).astype(np.float32).astype(np.float64) * 1.5
    s = O.RefSolver(rd, npml=4, rcoef=1e-4, fc=2000.0, mask=va, txs=tx, mode="f64", field=field)
    s.set_gf_map("pv_from_pin", (gf.p_pin, gf.v_pin), list(range(1, 9)), list(range(9, 17)), writeback=True)
    s.run()
    assert rel_l2(g.rec[0], s.tx["rec"][0]) < 1e-5 and rel_l2(g.src[0], s.tx["src"][0]) < 1e-5
