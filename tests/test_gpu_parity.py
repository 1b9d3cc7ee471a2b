"""GPU (libadfdtd.so) vs the oracle (CPU restatement of the reference) on
identical seeded inputs, through the C ABI.

Bars:
  ADF_F64      vs oracle Float64:  rel-L2 <= 1e-12 (observed bit-exact)
  ADF_F32_REF  vs oracle r32 (the reference's Float32 arithmetic): <= 1e-6
  ADF_F32      vs oracle Float64:  rel-L2 <= 1e-5
on every transceiver trace, the field and the CPML memory.
"""
import numpy as np
import pytest

import adfdtd as A
import oracle as O
from cases import build, gpu_run, oracle_solver, rel_l2

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "ref": 1e-6, "fast": 1e-5}

# Both kernel paths (two-pass P/V kernels and the fused z-marching sweep,
# with several tile shapes, waves per workgroup, the pipelined and the
# plane-at-a-time routines, and the three launch forms) must give the same
# results, with ordinary and non-temporal stores, and with the transceivers
# of a step in one launch per level, in one launch, or in the sweep launch
# (records one step later).
# (path, tile_y, block_z, waves, pipeline, launch, store_nt, tx_launch)
PATHS = [("two-pass", 0, 0, 0, 0, 0, 0, 1), ("fused", 0, 3, 1, 0, 0, 0, 0), ("fused", 0, 0, 2, 1, 1, 0, 1),
         ("fused", 0, 5, 4, 1, 0, 1, 1), ("fused", 0, 2, 1, 1, 2, 0, 0), ("fused", 0, 4, 1, 1, 0, 1, 2),
         ("fused", 0, 3, 1, 0, 0, 0, 2)]


@pytest.fixture(params=PATHS, ids=lambda p: f"{p[0]}-ty{p[1]}-kb{p[2]}-w{p[3]}-pf{p[4]}-l{p[5]}-nt{p[6]}-tx{p[7]}")
def path(request):
    return request.param


def _compare(d, txs, cpml, mask, kind, field=None, gf_pair=None, path=("auto", 0, 0, 0)):
    """kind: 'f64' (GPU F64 vs oracle f64), 'ref' (GPU F32_REF vs oracle r32)
    or 'fast' (GPU F32 vs oracle f64)."""
    mode = "r32" if kind == "ref" else "f64"
    ref = oracle_solver(O, d, txs, cpml, mask, mode, field=field)
    if gf_pair is not None:
        ref.set_gf_map(*gf_pair[0])
    ref.run()
    dtype = np.float64 if kind == "f64" else np.float32
    prec = "ref" if kind == "ref" else "fast"
    f, g, cp = gpu_run(d, txs, cpml, mask, dtype, prec, field=field, gf=None if gf_pair is None else gf_pair[1],
                       path=path[0], tile_y=path[1], block_z=path[2], waves=path[3], pipeline=path[4],
                       launch=path[5], store_nt=path[6] if len(path) > 6 else 0,
                       tx_launch=path[7] if len(path) > 7 else 1)
    tol = TOL[kind]
    errs = {
        "p_rec": rel_l2(g.rec[0], ref.tx["rec"][0]),
        "vx_rec": rel_l2(g.rec[1], ref.tx["rec"][1]),
        "vy_rec": rel_l2(g.rec[2], ref.tx["rec"][2]),
        "vz_rec": rel_l2(g.rec[3], ref.tx["rec"][3]),
        "vn_rec": rel_l2(g.vn_rec, ref.tx["vn_rec"]),
        "src": rel_l2(g.src, ref.tx["src"]),
        "P": rel_l2(f[0], ref.field[0]),
        "V": rel_l2(f[1:], ref.field[1:]),
    }
    if cpml.n:
        errs["psi"] = max(rel_l2(a, b) for a, b in zip(cp.p + cp.v, ref.mp + ref.mv))
    print(kind, {k: float(f"{v:.3g}") for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, bad
    return errs


@pytest.mark.parametrize("kind", ["f64", "ref", "fast"])
def test_point_source_cpml(kind, path):
    d, txs, cpml, mask = build(n=32, nt=90, npml=6, n_rec=24, rec_radius_cells=8)
    errs = _compare(d, txs, cpml, mask, kind, path=path)
    if kind in ("f64", "ref"):
        assert errs["p_rec"] == 0.0 and errs["P"] == 0.0  # same operations, same order


@pytest.mark.parametrize("kind", ["f64", "ref", "fast"])
def test_sphere_mask_velocity_source(kind, path):
    d, txs, cpml, mask = build(n=(36, 33, 30), nt=80, npml=5, shape="sphere", shape_params=[5 * 0.00686],
                               n_rec=30, rec_radius_cells=9, src_cells=(8.3, -1.2, 0.7),
                               src_dir=(0.0, 0.6, 0.8, 0.0))
    _compare(d, txs, cpml, mask, kind, path=path)


@pytest.mark.parametrize("npml,shape", [(0, "none"), (0, "cube"), (1, "none"), (1, "cube"), (2, "cross")])
def test_cpml_guards(npml, shape, path):
    # update_p_cpml! skips CPML for npml == 1 while update_p_cpml_mask! does not
    # (updates.jl:46-47 vs :333-337); npml == 0 selects update_p!/update_p_mask!.
    d, txs, cpml, mask = build(n=24, nt=60, npml=npml, shape=shape, shape_params=[8 * 0.00686], n_rec=10,
                               rec_radius_cells=7, src_cells=(2.5, 1.5, -0.5))
    _compare(d, txs, cpml, mask, "f64", path=path)


@pytest.mark.parametrize("kind", ["f64", "ref", "fast"])
@pytest.mark.parametrize("shape,size", [("sphere", 5), ("cube", 20)])
def test_random_field_every_cell_active(kind, shape, size, path):
    # a random initial field makes every cell, every CPML memory entry and
    # every boundary row/plane active from the first step; the cube of edge
    # 20 cells reaches into the CPML slabs (masked cells inside the slabs)
    d, txs, cpml, mask = build(n=(40, 29, 27), nt=12, npml=5, shape=shape, shape_params=[size * 0.00686],
                               n_rec=10, rec_radius_cells=6)
    f0 = np.random.default_rng(7).standard_normal((4, d.nz, d.ny, d.nx))
    errs = _compare(d, txs, cpml, mask, kind, field=f0 if kind == "f64" else f0.astype(np.float32), path=path)
    if kind == "f64":
        assert all(v == 0.0 for v in errs.values()), errs


@pytest.mark.parametrize("kind", ["f64", "fast"])
def test_mask_closed_where_no_routine_reads_it(kind, path):
    """A mask closed only where no update reads it (P at i = nx-1, Vx at
    i = 0, Vy on row 0, Vz on plane 0) leaves every row open for the
    tile split, so those tiles run the unmasked routines; a few rows also
    have cells closed for real (closed pressure cells and their faces).
    Float64 equals the oracle with the same mask bit for bit, from a random
    field."""
    d, txs, cpml, _ = build(n=(40, 29, 27), nt=10, npml=5, n_rec=10, rec_radius_cells=6)
    m = np.ones((4, d.nz, d.ny, d.nx), np.uint8)
    m[0, :, :, d.nx - 1] = 0
    m[1, :, :, 0] = 0
    m[2, :, 0, :] = 0
    m[3, 0, :, :] = 0
    rng = np.random.default_rng(11)
    for _ in range(6):  # closed cells for real, away from the CPML: the tiles around them stay masked
        i, j, k = rng.integers(8, d.nx - 8), rng.integers(8, d.ny - 8), rng.integers(8, d.nz - 8)
        m[:, k, j, i] = 0
        m[1, k, j, i + 1] = m[2, k, j + 1, i] = m[3, k + 1, j, i] = 0
    f0 = np.random.default_rng(5).standard_normal((4, d.nz, d.ny, d.nx))
    errs = _compare(d, txs, cpml, m, kind, field=f0 if kind == "f64" else f0.astype(np.float32), path=path)
    if kind == "f64":
        assert all(v == 0.0 for v in errs.values()), errs


@pytest.mark.parametrize("kind", ["f64", "ref", "fast"])
def test_anisotropic_spacing(kind, path):
    # dx != dy != dz (different xmax, ymax, zmax per axis) and nx != ny != nz:
    # a mix-up of axes, spacings or CPML slabs shows here; random field so
    # every cell and memory entry is active, an ellipsoid scatterer, and a
    # source injecting into every channel
    d, txs, cpml, mask = build(n=(34, 30, 27), h=(0.00686, 0.0081, 0.0059), nt=40, npml=5, shape="ellipsoid",
                               shape_params=[0.03, 0.022, 0.026], n_rec=20, rec_radius_cells=6,
                               src_cells=(1.3, -0.7, 0.4), src_dir=(1.0, 0.3, -0.4, 0.5))
    assert len({d.dx, d.dy, d.dz}) == 3
    f0 = np.random.default_rng(5).standard_normal((4, d.nz, d.ny, d.nx)) * 1e-3
    errs = _compare(d, txs, cpml, mask, kind, field=f0 if kind == "f64" else f0.astype(np.float32), path=path)
    if kind == "f64":
        assert all(v == 0.0 for v in errs.values()), errs


def test_overlapping_injectors_keep_sequential_order(path):
    # several transceivers inject into shared cells: the result depends on the
    # order of the transceiver loop (derivatives.jl:30-60); levels keep it.
    d, txs, cpml, mask = build(n=28, nt=60, npml=4, n_rec=0)
    rng = np.random.default_rng(3)
    extra = []
    for k in range(12):
        pt = (np.array([0.3, -0.2, 0.1]) + rng.uniform(-0.6, 0.6, 3)) * d.dx
        extra.append((pt, [1.0, 0.3, -0.2, 0.5], rng.standard_normal(d.nt) * 1e-3))
    d, txs, cpml, mask = build(n=28, nt=60, npml=4, n_rec=0, extra=extra)
    for kind in ("f64", "ref"):
        errs = _compare(d, txs, cpml, mask, kind, path=path)
        assert errs["p_rec"] == 0.0


def test_velocity_stencils_on_mirror_planes(path):
    # a velocity transceiver touching Vy j<=2 / Vz k<=2 forces the separate
    # mirror pass (forward.jl:134-137 after the transceivers)
    d, txs, cpml, mask = build(n=20, nt=50, npml=0, n_rec=4, rec_radius_cells=4)
    lo = np.array([-d.xmax, -d.ymax, -d.zmax]) + np.array([5.2, 0.7, 0.4]) * d.dx
    extra = [(lo, [0.0, 0.2, 0.7, 0.7], np.sin(np.arange(d.nt) * 0.3) * 1e-4)]
    d, txs, cpml, mask = build(n=20, nt=50, npml=0, n_rec=4, rec_radius_cells=4, extra=extra)
    _compare(d, txs, cpml, mask, "f64", path=path)


def test_plane_wave_initial_field_hologram_like(path):
    # initialField plane wave (simulation.jl:468-518) with surfaces as receivers
    d, txs, cpml, mask = build(n=(40, 24, 24), nt=70, npml=5, shape="cube", shape_params=[6 * 0.00686],
                               n_rec=16, rec_radius_cells=7, src_dir=(0.0, 0.0, 0.0, 0.0))
    f0 = A.initial_field(d, 2000.0, 0.12).astype(np.float64) * np.float32(1.5)
    _compare(d, txs, cpml, mask, "f64", field=f0, path=path)
    _compare(d, txs, cpml, mask, "ref", field=f0.astype(np.float32), path=path)


def test_device_mask_bit_exact():
    d, _, _, _ = build(n=(45, 38, 33), nt=5, npml=3)
    for shape, prm in (("sphere", [9 * d.dx]), ("cube", [11 * d.dx]), ("cross", [18 * d.dx]),
                       ("ellipsoid", [12 * d.dx, 7 * d.dx, 9 * d.dx])):
        s = A.Solver(d, 1, None)
        s.build_mask(shape, prm)
        m_dev = s.get_mask()
        rd = O.domain(d.tmax, d.xmax, d.ymax, d.zmax, d.nx, d.ny, d.nz, d.c0, d.r0, d.cf)
        m_ref = O.build_update_mask(rd, shape, prm)
        assert np.array_equal(m_dev, m_ref), shape
        # and the uploaded reference-format mask round-trips
        s.set_mask(m_ref)
        assert np.array_equal(s.get_mask(), m_ref)


@pytest.mark.parametrize("writeback", [True, False])
def test_gf_map_pv_from_pin(writeback, path):
    # outer surface (receivers) drives inner sources through a random causal GF
    rng = np.random.default_rng(11)
    d, txs, cpml, mask = build(n=26, nt=60, npml=4, n_rec=14, rec_radius_cells=7)
    n_out, n_in = 8, 6
    iRec = list(range(2, 2 + n_out))            # 1-based transceiver columns
    iSrc = list(range(2 + n_out, 2 + n_out + n_in))
    ng = 17
    pp = rng.standard_normal((n_out, n_in, ng)) * 1e-2  # numpy of Julia (ng, nsrc, nrec)
    vp = rng.standard_normal((n_out, n_in, ng)) * 1e-2
    ref_args = ("pv_from_pin", (pp, vp), iRec, iSrc, writeback)
    gf = A.PVFromPin(pp, vp, iRec, iSrc, writeback)
    errs = _compare(d, txs, cpml, mask, "f64", gf_pair=(ref_args, gf), path=path)
    assert errs["src"] <= 1e-12


def test_solver_reset_and_rerun_identical():
    d, txs, cpml, mask = build(n=24, nt=40, npml=3, n_rec=6, rec_radius_cells=6)
    g = A.transceivers_to_grid(d, txs, np.float64)
    cp = A.make_cpml(d, 3, 1e-4, 2000.0, np.float64)
    s = A.Solver(d, 2, cp)
    s.set_transceivers(g)
    s.run()
    a = s.get_trace(0)
    s.reset()
    s.set_transceivers(g)
    s.run()
    assert np.array_equal(a, s.get_trace(0))


@pytest.mark.parametrize("kind", ["f64", "ref", "fast"])
@pytest.mark.parametrize("n", [(68, 14, 12), (132, 10, 13), (100, 9, 11), (256, 6, 7)])
def test_wide_rows_vector_lanes(kind, n, path):
    # nx > 64 puts 2 or 4 cells on each lane of the fused kernel (one wave per
    # x-row); CPML on every axis, a sphere mask and transceivers near the walls
    d, txs, cpml, mask = build(n=n, nt=40, npml=2, shape="sphere", shape_params=[3.1 * 0.00686], n_rec=12,
                               rec_radius_cells=2.2, src_cells=(1.3, 0.4, -0.2), src_dir=(1.0, 0.2, -0.3, 0.1))
    _compare(d, txs, cpml, mask, kind, path=path)


@pytest.mark.parametrize("kind", ["f64", "ref"])
def test_inject_inner_sources_after_setup(kind, path):
    # inject_inner_sources! (fdtd_runners.jl:39-45) writes vn_src = C_f*p_drive
    # and p_src = C_q*vn_drive into an already packed transceiver grid; the
    # library sees them through adf_solver_set_trace and must treat those
    # transceivers as injectors (levels, fixups) from then on.
    rng = np.random.default_rng(2)
    d, txs, cpml, mask = build(n=30, nt=50, npml=4, n_rec=20, rec_radius_cells=6, src_dir=(0.0, 0.0, 0.0, 0.0))
    inner = np.arange(1, 21)  # numpy rows of the 20 surface transceivers
    cf = A.c_for_n(20, 6 * d.dx, d)
    p_drive = rng.standard_normal((20, d.nt)) * 1e-3
    vn_drive = rng.standard_normal((20, d.nt)) * 1e-6
    T = np.float64 if kind == "f64" else np.float32
    g = A.transceivers_to_grid(d, txs, T)
    p_src = g.src[0].copy()
    vn = g.vn_src.copy()
    p_src[inner] = (cf["C_q"] * vn_drive).astype(T)
    vn[inner] = (cf["C_f"] * p_drive).astype(T)
    # oracle with the sources in place from the start
    ref = oracle_solver(O, d, txs, cpml, mask, "f64" if kind == "f64" else "r32")
    ref.tx["src"][0][...] = p_src
    ref.tx["vn_src"][...] = vn
    ref.run()
    cp = A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, T)
    s = A.Solver(d, A.dtype_code(T, "ref" if kind == "ref" else "fast"), cp)
    s.set_path(path[0], path[1], path[2], path[3], pipeline=path[4], launch=path[5], store_nt=path[6],
               tx_launch=path[7])
    s.set_transceivers(g)  # zero sources at packing time
    s.set_trace(A._lib.ADF_SRC_P, p_src)
    s.set_trace(A._lib.ADF_VN_SRC, vn)
    s.run()
    for which, r in ((0, ref.tx["rec"][0]), (4, ref.tx["vn_rec"]), (6, ref.tx["src"][1]), (8, ref.tx["src"][3])):
        e = rel_l2(s.get_trace(which), r)
        assert e <= TOL[kind], (which, e)


def test_injection_surface_groups_one_launch_equals_levels():
    # A KH-style surface: a pressure (q) and a velocity (f) injector at each
    # of 60 points (~4.6 cells apart) and 30 recorders further out.  The
    # injectors fall into small groups that share written cells, packed
    # several per chain workgroup of the one-launch transceiver step; the
    # result must be the per-level launches' bit for bit and the oracle's.
    from cases import domain_cells
    d = domain_cells(36, 0.00686, 50)
    rng = np.random.default_rng(11)
    pts, nrm = A.fibonacci_sphere(60, 10 * d.dx)
    rp, rn = A.fibonacci_sphere(30, 14 * d.dx)
    txs = [A.make_transceiver(d, pts[i], [1.0, 0.0, 0.0, 0.0], rng.standard_normal(d.nt) * 1e-3) for i in range(60)]
    txs += [A.make_transceiver(d, pts[i], np.r_[0.0, nrm[i]], rng.standard_normal(d.nt) * 1e-6) for i in range(60)]
    txs += [A.make_transceiver(d, rp[i], np.r_[0.0, rn[i]], np.zeros(d.nt)) for i in range(30)]
    cpml = A.make_cpml(d, 5, 1e-4, 2000.0)
    mask = A.build_update_mask(d, "sphere", [4 * d.dx])
    ref = oracle_solver(O, d, txs, cpml, mask, "f64")
    ref.run()
    runs = [gpu_run(d, txs, cpml, mask, np.float64, "fast", path="fused", tx_launch=t) for t in (1, 0, 2)]
    (f1, g1, _), (f0, g0, _), (f2, g2, _) = runs
    for f, g in ((f0, g0), (f2, g2)):
        assert np.array_equal(f1, f)
        assert np.array_equal(g1.rec, g.rec) and np.array_equal(g1.vn_rec, g.vn_rec)
        assert np.array_equal(g1.src, g.src)
    errs = [rel_l2(g1.rec[c], ref.tx["rec"][c]) for c in range(4)]
    errs += [rel_l2(g1.vn_rec, ref.tx["vn_rec"]), rel_l2(g1.src, ref.tx["src"]), rel_l2(f1, ref.field)]
    print("f64 vs oracle", [float(f"{e:.3g}") for e in errs], flush=True)
    assert max(errs) <= TOL["f64"], errs


def test_tx_in_sweep_split_runs_and_trace_reads():
    # Transceivers carried by the sweep launch take each step's records one
    # step later (in the next sweep launch, or in a closing launch at the
    # end of a run).  Runs split at any step, a trace read between runs and
    # a change of the transceiver launch form between runs must leave every
    # trace and the field equal, bit for bit, to one run with the one-launch
    # transceiver step.
    from cases import domain_cells
    d = domain_cells(36, 0.00686, 40)
    rng = np.random.default_rng(12)
    pts, nrm = A.fibonacci_sphere(40, 10 * d.dx)
    rp, rn = A.fibonacci_sphere(30, 14 * d.dx)
    txs = [A.make_transceiver(d, pts[i], [1.0, 0.0, 0.0, 0.0], rng.standard_normal(d.nt) * 1e-3) for i in range(40)]
    txs += [A.make_transceiver(d, pts[i], np.r_[0.0, nrm[i]], rng.standard_normal(d.nt) * 1e-6) for i in range(40)]
    txs += [A.make_transceiver(d, rp[i], np.r_[0.0, rn[i]], np.zeros(d.nt)) for i in range(30)]
    mask = A.build_update_mask(d, "sphere", [4 * d.dx])

    def run(parts):
        cp = A.make_cpml(d, 5, 1e-4, 2000.0, np.float64)
        s = A.Solver(d, A.dtype_code(np.float64, "fast"), cp)
        s.set_mask(mask)
        s.set_transceivers(A.transceivers_to_grid(d, txs, np.float64))
        seen = []
        for a, b, tx in parts:
            s.set_path("fused", tx_launch=tx)
            s.run(a, b)
            seen.append(s.get_trace(0)[:, :b].copy())
        return s.get_field(), [s.get_trace(w) for w in range(10)], seen

    f1, t1, _ = run([(1, d.nt, 1)])
    f2, t2, s2 = run([(1, 7, 2), (8, 8, 2), (9, 25, 1), (26, d.nt, 2)])
    f3, t3, _ = run([(1, d.nt, 2)])
    for f, t in ((f2, t2), (f3, t3)):
        assert np.array_equal(f1, f)
        for a, b in zip(t1, t):
            assert np.array_equal(a, b)
    for part, b in zip(s2, (7, 8, 25, d.nt)):  # every record of a run is there when it returns
        assert np.array_equal(part, t1[0][:, :b])
