"""The fused sweep on rows wider than one wave and on rows whose length is
not a multiple of the cells per lane: the widths of the reference's
production grids (configs/paper.toml:21,71,165: 501^3, 301^3, 545^3).

A row of more than 64 x 4 cells is cut into segments of at most 62 x 4
cells, one wave each, whose edge lanes recompute the neighbouring
segments' edge pressure; an odd row ends in a lane that stores cell by
cell.  Against the oracle at small y/z extents (every cell active through a
random field, a sphere and a source across a segment boundary, CPML in
every direction), batch and server members of those widths against their
solo runs, and the fused sweep against the two-pass kernels at the paper
grids themselves.
"""
import numpy as np
import pytest

import adfdtd as A
import oracle as O
from adfdtd import _lib as L
from cases import build, gpu_run, oracle_solver, rel_l2

pytestmark = pytest.mark.gpu

H = 0.00686
TOL = {"f64": 0.0, "ref": 1e-6, "fast": 1e-5}

# (path, tile_y, block_z, waves, pipeline, launch, store_nt, cells per
# lane): the pipelined routines (auto tiling), the plane-at-a-time routines,
# the pipelined ones with 2 waves per workgroup, two launches and
# non-temporal stores, and 2 and 1 cells per lane (more segments per row)
PATHS = [("fused", 0, 0, 0, 1, 0, 0, 0), ("fused", 0, 3, 1, 0, 0, 0, 0), ("fused", 0, 4, 2, 1, 1, 1, 0),
         ("fused", 0, 0, 0, 1, 0, 0, 2), ("fused", 0, 3, 1, 1, 0, 0, 1)]

# 97: 2 cells per lane, ragged; 130: 4 per lane, ragged, one segment;
# 257, 301: two segments, ragged; 520: three segments of 176 cells
WIDTHS = [97, 130, 257, 301, 520]


def first_boundary(nx):
    """Cell where the second row segment starts (the row's middle for
    one-segment rows): fewest segments of at most 62 cpl cells, the cells
    shared out evenly in multiples of cpl."""
    cpl = 4 if nx > 128 else (2 if nx > 64 else 1)
    if nx <= 64 * cpl:
        return nx // 2
    n = -(-nx // (62 * cpl))
    return -(-(-(-nx // n)) // cpl) * cpl


def _case(nx, nt=14):
    """A source and eight velocity receivers straddling the first segment
    boundary, a sphere of radius 5 cells at the centre, 5-cell CPML."""
    off = first_boundary(nx) - (nx - 1) / 2  # cells from the centre
    rng = np.random.default_rng(nx)
    pts = [np.array([off + rng.uniform(-3, 3), rng.uniform(-6, 6), rng.uniform(-5, 5)]) * H for _ in range(8)]
    extra = [(p, [0.0, 1.0, 0.0, 0.0], np.zeros(nt)) for p in pts]
    return build(n=(nx, 21, 19), nt=nt, npml=5, shape="sphere", shape_params=[5 * H], n_rec=0,
                 src_cells=(off - 0.4, 0.3, -0.2), src_dir=(1.0, 0.4, -0.3, 0.2), extra=extra)


def _arrays(f, g, cp):
    return [f, g.src, g.vn_rec] + list(g.rec) + list(cp.p) + list(cp.v)


@pytest.mark.parametrize("nx", WIDTHS)
@pytest.mark.parametrize("kind", ["f64", "ref", "fast"])
def test_wide_and_ragged_rows_match_oracle(nx, kind):
    d, txs, cpml, mask = _case(nx)
    assert d.nx == nx
    f0 = np.random.default_rng(3).standard_normal((4, d.nz, d.ny, d.nx)) * 1e-3
    mode = "r32" if kind == "ref" else "f64"
    field = f0 if kind == "f64" else f0.astype(np.float32)
    ref = oracle_solver(O, d, txs, cpml, mask, mode, field=field).run()
    dtype = np.float64 if kind == "f64" else np.float32
    prec = "ref" if kind == "ref" else "fast"
    two = _arrays(*gpu_run(d, txs, cpml, mask, dtype, prec, field=field, path="two-pass"))
    for path in PATHS:
        f, g, cp = gpu_run(d, txs, cpml, mask, dtype, prec, field=field, path=path[0], tile_y=path[1],
                           block_z=path[2], waves=path[3], pipeline=path[4], launch=path[5], store_nt=path[6],
                           cells_per_lane=path[7])
        errs = {"P": rel_l2(f[0], ref.field[0]), "V": rel_l2(f[1:], ref.field[1:]),
                "rec": max(rel_l2(g.rec[c], ref.tx["rec"][c]) for c in range(4)),
                "vn_rec": rel_l2(g.vn_rec, ref.tx["vn_rec"]), "src": rel_l2(g.src, ref.tx["src"]),
                "psi": max(rel_l2(a, b) for a, b in zip(cp.p + cp.v, ref.mp + ref.mv))}
        print(nx, kind, path, {k: float(f"{v:.3g}") for k, v in errs.items()}, flush=True)
        bad = {k: v for k, v in errs.items() if not v <= TOL[kind]}
        assert not bad, (path, bad)
        if kind != "fast":  # the same arithmetic as the two-pass kernels, value for value
            for q, (x, y) in enumerate(zip(_arrays(f, g, cp), two)):
                assert np.array_equal(x, y), (path, q)


def _solver(case, dtype, prec="fast", field=None):
    d, txs, cpml, mask = case
    s = A.Solver(d, A.dtype_code(dtype, prec), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, dtype))
    s.set_mask(mask)
    if field is not None:
        s.set_field(field)
    s.set_transceivers(A.transceivers_to_grid(d, txs, dtype))
    return s


def _state(s):
    cp = A.make_cpml(s.dom, s.npml, 1e-4, 2000.0, s.T)
    s.get_cpml_memory(cp)
    return [s.get_field()] + [s.get_trace(w) for w in range(10)] + list(cp.p) + list(cp.v)


@pytest.mark.parametrize("dtype,prec", [(np.float64, "fast"), (np.float32, "ref"), (np.float32, "fast")])
def test_batch_members_of_every_width_equal_solo(dtype, prec):
    cs = [_case(nx, nt=30) for nx in (301, 520, 130)]
    cs.append(build(n=(64, 16, 14), nt=30, npml=3, n_rec=6, rec_radius_cells=4))
    solo = [_solver(c, dtype, prec) for c in cs]
    for s in solo:
        s.run(1, 30)
    b = A.Batch()
    mem = [_solver(c, dtype, prec) for c in cs]
    for s in mem:
        b.add(s)
    b.run(11)
    b.run(19)
    for i, (a, c) in enumerate(zip(mem, solo)):
        for q, (x, y) in enumerate(zip(_state(a), _state(c))):
            assert np.array_equal(x, y), f"member {i} array {q}"


def test_server_takes_a_paper_width_request():
    d, txs, cpml, _ = _case(301, nt=20)
    s = _solver((d, txs, cpml, None), np.float32)
    s.build_mask("sphere", [5 * H])
    s.run(1, 20)
    want = [s.get_trace(w) for w in range(10)], s.get_field()
    with A.Server() as sv:
        rid = sv.submit(d, L.ADF_F32, A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, np.float32),
                        A.transceivers_to_grid(d, txs, np.float32), mask=("sphere", [5 * H]), start_step=1,
                        end_step=20, want=0x3FF, keep_field=True)
        assert sv.wait(rid, 120.0) == "done"
        for w in range(10):
            assert np.array_equal(sv.trace(rid, w), want[0][w]), w
        assert np.array_equal(sv.field(rid), want[1])


@pytest.mark.parametrize("n,dtype,prec,cpl", [(301, np.float64, "fast", 0), (301, np.float32, "ref", 2),
                                              (501, np.float32, "ref", 0), (545, np.float32, "ref", 0)])
def test_paper_grids_fused_equals_two_pass(n, dtype, prec, cpl):
    """Full paper grids, a few steps from a random field with CPML and a
    sphere: the fused sweep (several segments, ragged rows) equals the
    two-pass kernels bit for bit (both are bit-exact to the oracle on the
    small cases)."""
    d = A.make_domain(1e-3, (n - 1) * H / 2, (n - 1) * H / 2, (n - 1) * H / 2, n, n, n, 343.0, 1.21, 0.5)
    d = A.make_domain(5.5 * d.dt, d.xmax, d.ymax, d.zmax, n, n, n, 343.0, 1.21, 0.5)
    assert d.nt == 6
    rng = np.random.default_rng(n)
    f0 = rng.standard_normal((4, n, n, n), dtype=np.float32).astype(dtype) * dtype(1e-3)
    tf = A.ricker_wavelet(d, 2000.0, True)
    txs = [A.make_transceiver(d, np.array([0.3, -0.2, 0.1]) * H, [1.0, 0.2, 0.0, -0.1], tf)]
    txs += [A.make_transceiver(d, np.array([x, 0.0, 0.0]) * H, [0.0, 1.0, 0.0, 0.0], np.zeros(d.nt))
            for x in (-40.5, 2.2, 60.7)]
    out = []
    for path in ("fused", "two-pass"):
        s = A.Solver(d, A.dtype_code(dtype, prec), A.make_cpml(d, 10, 1e-4, 2000.0, dtype))
        s.build_mask("sphere", [20 * H])
        s.set_field(f0)
        s.set_transceivers(A.transceivers_to_grid(d, txs, dtype))
        s.set_path(path, cells_per_lane=cpl)
        s.run(1, d.nt)
        out.append(_state(s))
        del s
    for q, (x, y) in enumerate(zip(*out)):
        assert np.array_equal(x, y), f"array {q}: max diff {np.max(np.abs(x - y))}"


@pytest.mark.parametrize("n,kind", [(301, "f64"), (501, "ref")])
def test_paper_grid_matches_oracle(n, kind):
    """Four steps of a full paper grid against the oracle: 10-cell CPML, a
    sphere across the middle segment boundary, a random field (every cell
    and CPML entry active), a source and receivers spread over the row
    segments.  Float64 bit for bit; the reference's Float32 arithmetic
    within 1e-6."""
    d, txs, cpml, mask = build(n=n, nt=4, npml=10, shape="sphere", shape_params=[20 * H], n_rec=24,
                               rec_radius_cells=n // 3, src_cells=(first_boundary(n) - (n - 1) / 2 - 0.4, 0.3, -0.2),
                               src_dir=(1.0, 0.4, -0.3, 0.2))
    dtype = np.float64 if kind == "f64" else np.float32
    f0 = np.random.default_rng(n).standard_normal((4, n, n, n), dtype=np.float32).astype(dtype) * dtype(1e-3)
    ref = oracle_solver(O, d, txs, cpml, mask, "f64" if kind == "f64" else "r32", field=f0).run()
    f, g, cp = gpu_run(d, txs, cpml, mask, dtype, "fast" if kind == "f64" else "ref", field=f0, path="fused")
    got = [f, g.src, g.vn_rec] + list(g.rec) + list(cp.p) + list(cp.v)
    want = [ref.field, ref.tx["src"], ref.tx["vn_rec"]] + list(ref.tx["rec"]) + list(ref.mp) + list(ref.mv)
    for q, (a, b) in enumerate(zip(got, want)):
        if kind == "f64":
            assert np.array_equal(a, b), q
        else:
            assert rel_l2(a, b) <= 1e-6, (q, rel_l2(a, b))


@pytest.mark.parametrize("nx", [3, 5, 7, 9, 13, 31, 33, 63, 65, 127, 129, 131])
def test_narrow_and_odd_rows_at_every_cells_per_lane(nx):
    """Rows narrower than a lane's cells (nx = 3 at 4 cells per lane; the
    library's smallest grid is 3^3),
    odd and one-past-a-boundary widths, at 1, 2 and 4 cells per lane and
    the automatic choice: Float64 bit for bit against the oracle, random
    field, CPML where the row has room for it."""
    p = min(3, (nx - 1) // 3)
    d, txs, cpml, mask = build(n=(nx, 14, 13), nt=6, npml=p, n_rec=0, src_cells=(0.2 - (nx % 2) * 0.5, 0.3, -0.4),
                               src_dir=(1.0, 0.3, -0.2, 0.1))
    f0 = np.random.default_rng(nx).standard_normal((4, d.nz, d.ny, d.nx))
    ref = oracle_solver(O, d, txs, cpml, mask, "f64", field=f0).run()
    want = [ref.field, ref.tx["src"], ref.tx["vn_rec"]] + list(ref.tx["rec"]) + list(ref.mp) + list(ref.mv)
    for cpl in (0, 1, 2, 4):
        c = cpl or (4 if nx > 128 else (2 if nx > 64 else 1))
        if not x_slabs_apart(nx, p, c):
            # a lane would hold cells of both x slabs (nx = 5 at 4 cells per
            # lane): the fused path refuses, the automatic one runs two-pass
            with pytest.raises(L.AdfError, match="fused path"):
                gpu_run(d, txs, cpml, mask, np.float64, "fast", field=f0, path="fused", cells_per_lane=cpl)
            f, g, cp = gpu_run(d, txs, cpml, mask, np.float64, "fast", field=f0, cells_per_lane=cpl)
            got = [f, g.src, g.vn_rec] + list(g.rec) + list(cp.p) + list(cp.v)
            for q, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), (cpl, "auto", q)
            continue
        for pf in (1, 0):
            f, g, cp = gpu_run(d, txs, cpml, mask, np.float64, "fast", field=f0, path="fused", block_z=2,
                               pipeline=pf, cells_per_lane=cpl)
            got = [f, g.src, g.vn_rec] + list(g.rec) + list(cp.p) + list(cp.v)
            for q, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), (cpl, pf, q)


def x_slabs_apart(nx, p, cpl):
    """No lane of cpl cells (starting at multiples of cpl) holds updated
    cells of both x CPML slabs of a phase (pressure: cells [0, nx-1), slab
    width p; velocity: cells [1, nx), width p+1)."""
    for w, lo, hi in ((p, 0, nx - 1), (p + 1, 1, nx)):
        if w <= 0:
            continue
        for i0 in range(0, nx, cpl):
            cells = [i for i in range(i0, min(i0 + cpl, nx)) if lo <= i < hi]
            if any(i < w for i in cells) and any(w <= i and i >= nx - w for i in cells):
                return False
    return True


@pytest.mark.parametrize("p", [63, 64])
def test_x_slabs_that_meet_take_the_two_pass_kernels(p):
    """At nx = 130 (4 cells per lane) a 64-cell CPML puts the velocity
    phase's two x slabs side by side (the library allows npml up to
    (min(nx, ny, nz) - 1) / 2), and the lane holding cells 64-67 has
    cells of both (64 on the left, 65-67 on the right, where the memory
    coefficients are not zero).  The fused routines address a lane's x-slab
    memory from one side, so that grid runs on the two-pass kernels (slab
    test per cell) and refuses the fused path; with a 63-cell CPML the slabs
    stay apart and the grid runs fused.  Float64 equals the oracle bit for
    bit either way, ten steps from a random field (the memory of those
    cells turns nonzero from the first step)."""
    nx = 130
    assert x_slabs_apart(nx, p, 4) == (p == 63)
    d, txs, cpml, mask = build(n=(nx, 129, 129), nt=10, npml=p, n_rec=0, src_cells=(0.3, 0.2, -0.1),
                               src_dir=(1.0, 0.3, -0.2, 0.1))
    rng = np.random.default_rng(p)
    f0 = rng.standard_normal((4, d.nz, d.ny, d.nx))
    ref = oracle_solver(O, d, txs, cpml, mask, "f64", field=f0).run()
    want = [ref.field, ref.tx["src"], ref.tx["vn_rec"]] + list(ref.tx["rec"]) + list(ref.mp) + list(ref.mv)
    runs = [dict()]
    if p == 63:
        runs.append(dict(path="fused", pipeline=1, block_z=3))
    else:
        with pytest.raises(L.AdfError, match="fused path"):
            gpu_run(d, txs, cpml, mask, np.float64, "fast", field=f0, path="fused")
    for kw in runs:
        f, g, cp = gpu_run(d, txs, cpml, mask, np.float64, "fast", field=f0, **kw)
        got = [f, g.src, g.vn_rec] + list(g.rec) + list(cp.p) + list(cp.v)
        for q, (a, b) in enumerate(zip(got, want)):
            assert np.array_equal(a, b), (kw, q)
