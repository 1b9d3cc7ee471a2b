"""The in-loop KH gather (extrapolation.jl:101-249) in its blocked form:
all three map kinds against the oracle, with maps shorter and longer than a
block of steps, longer than one lag tile, and inputs spanning several
chunks; and runs split into pieces, stepped one at a time or reset and run
again, which must give the same bits as one run.

Bars: src traces within 1e-12 relative L2 of the oracle's (Float64; the
gather sums in another order), and bit-identical across run splittings."""
import numpy as np
import pytest

import adfdtd as A
import oracle as O
from adfdtd import _lib as L
from cases import build, oracle_solver, rel_l2

pytestmark = pytest.mark.gpu

KINDS = ("pv_from_pin", "pv_from_pv", "p_from_pin")


def _map(kind, n_in, n_out, ng, iRec, iSrc, seed, writeback=True):
    rng = np.random.default_rng(seed)
    r = lambda *sh: rng.standard_normal(sh) * 1e-2
    if kind == "pv_from_pin":   # numpy of Julia (ng, nsrc, nrec)
        t = (r(n_in, n_out, ng), r(n_in, n_out, ng))
        return (kind, t, iRec, iSrc, writeback), A.PVFromPin(*t, iRec, iSrc, writeback)
    if kind == "pv_from_pv":    # numpy of Julia (ng, nm, nout)
        t = (r(n_out, n_in, ng), r(n_out, n_in, ng), r(n_out, n_in, ng), r(n_out, n_in, ng))
        return (kind, t, iRec, iSrc, writeback), A.PVFromPV(*t, iRec, iSrc, writeback)
    t = (r(n_out, n_in, ng), r(n_out, n_in, ng))  # numpy of Julia (ng, nrec, nsrc)
    return (kind, t, iRec, iSrc, writeback), A.PFromPin(*t, iRec, iSrc, writeback)


def _case(n, nt, n_in, n_out):
    d, txs, cpml, mask = build(n=n, nt=nt, npml=3, n_rec=n_in + n_out, rec_radius_cells=n / 4)
    iRec = list(range(2, 2 + n_in))
    iSrc = list(range(2 + n_in, 2 + n_in + n_out))
    return d, txs, cpml, mask, iRec, iSrc


def _gpu(d, txs, cpml, gf, runs):
    s = A.Solver(d, A.dtype_code(np.float64), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, np.float64))
    s.set_transceivers(A.transceivers_to_grid(d, txs, np.float64))
    s.set_gf_map(gf)
    for a, b in runs:
        if a == "reset":
            s.reset()
            s.set_transceivers(A.transceivers_to_grid(d, txs, np.float64))
            continue
        s.run(a, b)
    return [s.get_trace(w) for w in range(10)]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("ng", [5, 40])
def test_blocked_gather_matches_oracle(kind, ng):
    d, txs, cpml, mask, iRec, iSrc = _case(22, 70, 40, 7)
    ref_args, gf = _map(kind, 40, 7, ng, iRec, iSrc, seed=ng)
    ref = oracle_solver(O, d, txs, cpml, mask, "f64")
    ref.set_gf_map(*ref_args)
    ref.run()
    tr = _gpu(d, txs, cpml, gf, [(1, d.nt)])
    src = np.stack([tr[L.ADF_SRC_P], tr[L.ADF_VN_SRC]])
    rsrc = np.stack([ref.tx["src"][0], ref.tx["vn_src"]])
    assert np.abs(rsrc[:, 8:]).max() > 0
    assert rel_l2(src, rsrc) <= 1e-12, rel_l2(src, rsrc)
    assert rel_l2(tr[L.ADF_REC_P], ref.tx["rec"][0]) <= 1e-12


def test_blocked_gather_long_map():
    # lags beyond one 1024-lag tile of the far pass
    d, txs, cpml, mask, iRec, iSrc = _case(14, 1150, 18, 3)
    ref_args, gf = _map("pv_from_pin", 18, 3, 1100, iRec, iSrc, seed=5)
    ref = oracle_solver(O, d, txs, cpml, mask, "f64")
    ref.set_gf_map(*ref_args)
    ref.run()
    tr = _gpu(d, txs, cpml, gf, [(1, d.nt)])
    src = np.stack([tr[L.ADF_SRC_P], tr[L.ADF_VN_SRC]])
    rsrc = np.stack([ref.tx["src"][0], ref.tx["vn_src"]])
    assert rel_l2(src, rsrc) <= 1e-12, rel_l2(src, rsrc)


@pytest.mark.parametrize("kind", ["pv_from_pin", "pv_from_pv"])
def test_blocked_gather_run_splittings_are_bit_identical(kind):
    d, txs, cpml, mask, iRec, iSrc = _case(18, 60, 20, 5)
    _, gf = _map(kind, 20, 5, 23, iRec, iSrc, seed=3)
    one = _gpu(d, txs, cpml, gf, [(1, d.nt)])
    pieces = _gpu(d, txs, cpml, gf, [(1, 17), (18, 40), (41, d.nt)])
    steps = _gpu(d, txs, cpml, gf, [(i, i) for i in range(1, d.nt + 1)])
    again = _gpu(d, txs, cpml, gf, [(1, 25), ("reset", None), (1, d.nt)])
    for other in (pieces, steps, again):
        assert all(np.array_equal(a, b) for a, b in zip(one, other))


def test_set_transceivers_keeps_the_map_consistent():
    d, txs, cpml, mask, iRec, iSrc = _case(16, 30, 6, 3)
    _, gf = _map("pv_from_pin", 6, 3, 9, iRec, iSrc, seed=2)
    s = A.Solver(d, A.dtype_code(np.float64), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, np.float64))
    s.set_transceivers(A.transceivers_to_grid(d, txs, np.float64))
    s.set_gf_map(gf)
    # fewer transceivers than the map's iRec/iSrc reach: refused, nothing changed
    with pytest.raises(L.AdfError, match="GF map"):
        s.set_transceivers(A.transceivers_to_grid(d, txs[:4], np.float64))
    s.run()
    a = s.get_trace(L.ADF_SRC_P)
    # the same transceivers again, after the map: the map's sources stay injectors
    s.reset()
    s.set_transceivers(A.transceivers_to_grid(d, txs, np.float64))
    s.run()
    assert np.array_equal(a, s.get_trace(L.ADF_SRC_P))


def _solver(d, cpml, g):
    s = A.Solver(d, A.dtype_code(np.float64), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, np.float64))
    s.set_transceivers(g)
    return s


def test_replacing_a_map_leaves_no_stale_outputs():
    """A transceiver with a vn drive of its own is an output of a first map
    and not of a second.  After the second map replaces the first (or the
    first is removed), the run equals one that only ever had the second map
    (or none), bit for bit: that transceiver's vn2v! runs again with the
    velocity transceivers instead of after the extrapolation."""
    d, txs, cpml, mask, iRec, iSrc = _case(18, 50, 8, 6)
    _, gf1 = _map("pv_from_pin", 8, 6, 11, iRec, iSrc, seed=8)
    _, gf2 = _map("pv_from_pin", 8, 5, 11, iRec, iSrc[:5], seed=9)
    g = A.transceivers_to_grid(d, txs, np.float64)
    g.vn_src[iSrc[5] - 1] = A.ricker_wavelet(d, 2000.0, True) * 1e-3

    def run(maps):
        s = _solver(d, cpml, g)
        for m in maps:
            s.set_gf_map(m)
        s.run()
        return [s.get_trace(w) for w in range(10)] + [s.get_field()]

    for got, want in ((run([gf1, gf2]), run([gf2])), (run([gf1, None]), run([]))):
        assert all(np.array_equal(a, b) for a, b in zip(got, want))
    assert np.abs(run([gf2])[L.ADF_SRC_VX][iSrc[5] - 1]).max() > 0


def test_duplicate_outputs_are_refused_with_write_back():
    d, txs, cpml, mask, iRec, iSrc = _case(16, 30, 6, 3)
    dup = [iSrc[0], iSrc[1], iSrc[0]]
    s = _solver(d, cpml, A.transceivers_to_grid(d, txs, np.float64))
    with pytest.raises(L.AdfError, match="twice"):
        s.set_gf_map(_map("pv_from_pin", 6, 3, 9, iRec, dup, seed=2)[1])
    # without write-back nothing is written, so the map is accepted
    s.set_gf_map(_map("pv_from_pin", 6, 3, 9, iRec, dup, seed=2, writeback=False)[1])
    s.run()


@pytest.mark.parametrize("kind", KINDS)
def test_blocked_gather_fp32_within_1e5_of_the_float64_oracle(kind):
    """The FP32 solve carries the maps as float32 words (sums still in
    double): extrapolated sources and the traces they drive stay within
    1e-5 relative L2 of the Float64 oracle, over several 16-step blocks."""
    d, txs, cpml, mask, iRec, iSrc = _case(22, 90, 40, 7)
    ref_args, gf = _map(kind, 40, 7, 40, iRec, iSrc, seed=21)
    ref = oracle_solver(O, d, txs, cpml, mask, "f64")
    ref.set_gf_map(*ref_args)
    ref.run()
    s = A.Solver(d, A.dtype_code(np.float32, "fast"), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, np.float32))
    s.set_transceivers(A.transceivers_to_grid(d, txs, np.float32))
    s.set_gf_map(gf)
    s.run()
    src = np.stack([s.get_trace(L.ADF_SRC_P), s.get_trace(L.ADF_VN_SRC)])
    rsrc = np.stack([ref.tx["src"][0], ref.tx["vn_src"]])
    errs = (rel_l2(src, rsrc), rel_l2(s.get_trace(L.ADF_REC_P), ref.tx["rec"][0]))
    print(kind, errs, flush=True)
    assert np.abs(rsrc[:, 8:]).max() > 0
    assert max(errs) <= 1e-5, errs


@pytest.mark.parametrize("kind", ["pv_from_pin", "pv_from_pv"])
def test_resume_in_a_fresh_solver_from_saved_state(kind):
    """A run resumed at step 41 in a fresh solver, from the field, CPML
    memory and the ten trace buffers saved after step 40, gives the same
    bits as one run: the extrapolation takes the history of steps 1-40 from
    the restored record traces, as the reference reads txs.p.rec and
    txs.vn_rec (extrapolation.jl:202-249), and its 16-step blocks do not
    depend on where a run starts."""
    d, txs, cpml, mask, iRec, iSrc = _case(18, 60, 20, 5)
    _, gf = _map(kind, 20, 5, 23, iRec, iSrc, seed=3)

    def solver():
        s = A.Solver(d, A.dtype_code(np.float64), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, np.float64))
        s.set_transceivers(A.transceivers_to_grid(d, txs, np.float64))
        s.set_gf_map(gf)
        return s

    one = solver()
    one.run(1, d.nt)
    a = solver()
    a.run(1, 40)
    cp = A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, np.float64)
    a.get_cpml_memory(cp)
    b = solver()
    b.set_field(a.get_field())
    b.set_cpml_memory(cp)
    for w in range(10):
        b.set_trace(w, a.get_trace(w))
    b.run(41, d.nt)
    assert np.abs(one.get_trace(L.ADF_SRC_P)[iSrc[0] - 1]).max() > 0
    assert all(np.array_equal(one.get_trace(w), b.get_trace(w)) for w in range(10))
    assert np.array_equal(one.get_field(), b.get_field())
