"""Batched engine (adf_batch_*): each member of a batch must end in the state
`adf_solver_run` gives for the same steps, bit for bit, whatever else is in
the batch (grid sizes of every row-width class, masks, CPML widths,
injectors, mirrored-plane stencils, members joining mid-run), with and
without graph replay and the persistent tile queue; the Float64 members also
match the oracle exactly."""
import numpy as np
import pytest

import adfdtd as A
import oracle as O
from adfdtd import _lib as L
from cases import build, oracle_solver, rel_l2

pytestmark = pytest.mark.gpu

TRACES = range(10)


def make_solver(case, dtype, prec="fast", tile_y=0, block_z=0, waves=0):
    d, txs, cpml, mask = case
    s = A.Solver(d, A.dtype_code(dtype, prec), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, dtype))
    s.set_mask(mask)
    s.set_transceivers(A.transceivers_to_grid(d, txs, dtype))
    if tile_y or block_z or waves:
        s.set_path("fused", tile_y, block_z, waves)
    return s


def state(s):
    cp = A.make_cpml(s.dom, s.npml, 1e-4, 2000.0, s.T)
    s.get_cpml_memory(cp)
    return [s.get_field()] + [s.get_trace(w) for w in TRACES] + list(cp.p) + list(cp.v)


def assert_same(a, b, what):
    for q, (x, y) in enumerate(zip(state(a), state(b))):
        assert np.array_equal(x, y), f"{what}: array {q} differs (max {np.max(np.abs(x - y))})"


def cases():
    """Members of every row-width class (1, 2 and 4 cells per lane) and the
    features that take separate code in a step."""
    h = 0.00686
    out = []
    # point source, sphere mask, 6-cell CPML, 48 wide (1 cell per lane)
    out.append(build(n=(48, 20, 22), nt=70, npml=6, shape="sphere", shape_params=[4 * h], n_rec=12,
                     rec_radius_cells=7, src_cells=(2.3, -1.1, 0.4)))
    # velocity source, no mask, 96 wide (2 cells per lane), thin CPML
    out.append(build(n=(96, 18, 16), nt=70, npml=3, n_rec=10, rec_radius_cells=6, src_cells=(-3.2, 0.7, 1.9),
                     src_dir=(0.0, 0.6, 0.0, 0.8)))
    # 160 wide (4 cells per lane), cube mask, receivers on the mirrored
    # planes (separate mirror pass) and two overlapping p injectors
    d = build(n=(160, 14, 12), nt=70, npml=4, n_rec=0)[0]
    tf = A.ricker_wavelet(d, 2000.0, True)
    lo = np.array([-d.xmax, -d.ymax, -d.zmax])
    extra = [(lo + np.array([20.3, 0.4, 5.1]) * h, [0.0, 0.0, 1.0, 0.0], tf),
             (lo + np.array([30.5, 6.2, 0.6]) * h, [0.0, 0.0, 0.0, 1.0], tf),
             (np.array([1.2, 0.3, -0.4]) * h, [1.0, 0.0, 0.0, 0.0], tf),
             (np.array([1.7, 0.6, -0.2]) * h, [0.7, 0.0, 0.0, 0.0], 0.5 * tf)]
    out.append(build(n=(160, 14, 12), nt=70, npml=4, shape="cube", shape_params=[5 * h], n_rec=8,
                     rec_radius_cells=4, extra=extra))
    # 128 wide, no CPML
    out.append(build(n=(128, 12, 13), nt=70, npml=0, shape="cross", shape_params=[6 * h], n_rec=6,
                     rec_radius_cells=4, src_cells=(1.5, 0.5, -0.5)))
    return out


@pytest.mark.parametrize("dtype,prec", [(np.float64, "fast"), (np.float32, "fast"), (np.float32, "ref")])
@pytest.mark.parametrize("persistent,graph", [(1, 16), (0, 0), (1, 5)])
def test_batch_equals_solo(dtype, prec, persistent, graph):
    cs = cases()
    solo = [make_solver(c, dtype, prec) for c in cs]
    for s in solo:
        s.run(1, s.dom.nt)
    mem = [make_solver(c, dtype, prec) for c in cs]
    b = A.Batch(persistent=bool(persistent), graph_steps=graph)
    for s in mem:
        b.add(s)
    b.run(33)
    b.run(cs[0][0].nt - 33)
    for i, (a, c) in enumerate(zip(mem, solo)):
        assert a.next_step() == c.dom.nt + 1
        assert_same(a, c, f"member {i}")


@pytest.mark.parametrize("tx", [0, 1])
def test_batch_transceiver_launch_forms(tx):
    """The members' transceivers as the batch's per-level launches (the
    members' ADF_OPT_TX_LAUNCH 0) or as one launch for every member, which
    then also moves the members to the next step (no member here needs the
    separate mirror pass), across graph replays: each member equals its solo
    run bit for bit."""
    cs = [c for i, c in enumerate(cases()) if i != 2]
    solo = [make_solver(c, np.float64) for c in cs]
    for s in solo:
        s.run(1, s.dom.nt)
    mem = [make_solver(c, np.float64) for c in cs]
    b = A.Batch(graph_steps=16)
    for s in mem:
        s.set_option(L.ADF_OPT_TX_LAUNCH, tx)
        b.add(s)
    b.run(33)
    b.run(cs[0][0].nt - 33)
    for i, (a, c) in enumerate(zip(mem, solo)):
        assert a.next_step() == c.dom.nt + 1
        assert_same(a, c, f"member {i}")


def test_batch_f64_matches_oracle():
    cs = cases()
    b = A.Batch()
    mem = [make_solver(c, np.float64) for c in cs]
    for s in mem:
        b.add(s)
    b.run(cs[0][0].nt)
    for s, (d, txs, cpml, mask) in zip(mem, cs):
        ref = oracle_solver(O, d, txs, cpml, mask, "f64").run()
        for c in range(4):
            assert rel_l2(s.get_trace(L.ADF_REC_P + c), ref.tx["rec"][c]) == 0.0
        assert rel_l2(s.get_trace(L.ADF_VN_REC), ref.tx["vn_rec"]) == 0.0
        assert rel_l2(s.get_field(), ref.field) == 0.0


def test_members_join_and_leave_mid_run():
    """Continuous batching: members at different steps advance together."""
    cs = cases()
    dtype = np.float32
    solo = [make_solver(c, dtype) for c in cs]
    for s in solo:
        s.run(1, s.dom.nt)
    mem = [make_solver(c, dtype) for c in cs]
    b = A.Batch(graph_steps=8)
    b.add(mem[0])
    b.add(mem[1])
    b.run(20)
    b.add(mem[2])            # joins at its step 1 while the others are at 21
    b.run(25)
    b.remove(mem[0])         # leaves at step 46; finishes alone
    mem[0].run(46, mem[0].dom.nt)
    b.add(mem[3])
    b.run(25)                # member 1 reaches nt; members 2, 3 are behind
    assert mem[1].next_step() == mem[1].dom.nt + 1
    b.remove(mem[1])
    b.run(20)
    assert mem[2].next_step() == 71 and mem[3].next_step() == 46
    b.remove(mem[2])
    b.run(25)
    for i, (a, c) in enumerate(zip(mem, solo)):
        assert a.next_step() == c.dom.nt + 1
        assert_same(a, c, f"member {i}")


def test_tilings_do_not_change_results():
    cs = cases()[:3]
    ref = [make_solver(c, np.float32) for c in cs]
    for s in ref:
        s.run(1, s.dom.nt)
    b = A.Batch()
    mem = [make_solver(c, np.float32, tile_y=ty, block_z=kb, waves=w)
           for c, (ty, kb, w) in zip(cs, [(0, 3, 2), (0, 0, 1), (0, 7, 4)])]
    for s in mem:
        b.add(s)
    b.run(cs[0][0].nt)
    for i, (a, c) in enumerate(zip(mem, ref)):
        assert_same(a, c, f"member {i}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_store_hint_does_not_change_results(dtype):
    """Members with the non-temporal store hint (ADF_OPT_STORE_NT) sweep in
    a launch of their own beside the others; every member still ends as its
    solo run."""
    cs = cases()
    ref = [make_solver(c, dtype) for c in cs]
    for s in ref:
        s.run(1, s.dom.nt)
    b = A.Batch(graph_steps=8)
    mem = [make_solver(c, dtype) for c in cs]
    for i, s in enumerate(mem):
        s.set_path("auto", store_nt=i % 2)
        b.add(s)
    b.run(cs[0][0].nt)
    for i, (a, c) in enumerate(zip(mem, ref)):
        assert_same(a, c, f"member {i}")


def test_batch_errors():
    cs = cases()
    b = A.Batch()
    s32 = make_solver(cs[0], np.float32)
    s64 = make_solver(cs[1], np.float64)
    b.add(s32)
    with pytest.raises(L.AdfError, match="dtype"):
        b.add(s64)
    with pytest.raises(L.AdfError, match="already"):
        b.add(s32)
    with pytest.raises(L.AdfError, match="past its nt"):
        b.run(s32.dom.nt + 1)
    assert s32.next_step() == 1
    # rows of more than 63 segments of 62 x 4 cells
    d = A.make_domain(1e-3, 10.0, 0.02, 0.02, 15625, 4, 4, 343.0, 1.21, 0.5)
    wide = A.Solver(d, L.ADF_F32, None)
    with pytest.raises(L.AdfError, match="row segments"):
        b.add(wide)


def test_mask_and_transceivers_changed_mid_run():
    """A member whose mask and transceivers change between batch runs takes
    the new ones in its next step; a solo fused solver does the same.  In
    Float64 both equal the two-pass kernels, which read the mask cell by
    cell, bit for bit."""
    h = 0.00686
    kw = dict(n=(48, 20, 22), nt=60, npml=6, n_rec=8, rec_radius_cells=6)
    d, txs, cpml, m1 = build(shape="sphere", shape_params=[3 * h], **kw)
    m2 = build(shape="cube", shape_params=[7 * h], **kw)[3]
    txs2 = build(src_cells=(-2.6, 1.3, 0.8), **kw)[1]
    other = cases()[1]
    runs = []
    for mode in ("two-pass", "fused", "batch"):
        s = make_solver((d, txs, cpml, m1), np.float64)
        if mode != "batch":
            s.set_path(mode)
            runs.append((s, None))
            continue
        b = A.Batch(graph_steps=8)
        o = make_solver(other, np.float64)
        b.add(s)
        b.add(o)
        runs.append((s, (b, o)))
    for s, bo in runs:
        step = (lambda n: bo[0].run(n)) if bo else (lambda n, s=s: s.run(s.next_step(), s.next_step() + n - 1))
        step(20)
        s.set_mask(m2)
        step(20)
        s.set_transceivers(A.transceivers_to_grid(d, txs2, np.float64))
        step(d.nt - 40)
        assert s.next_step() == d.nt + 1
    assert_same(runs[1][0], runs[0][0], "fused solo")
    assert_same(runs[2][0], runs[0][0], "batch member")
    solo = make_solver(other, np.float64)
    solo.run(1, d.nt)
    assert_same(runs[2][1][1], solo, "unchanged member")
