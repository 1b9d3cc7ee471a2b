"""Served form (adf_server_*): every request's results equal those of the
same solve run alone through adf_solver_run, value for value, while other
requests of other sizes, dtypes and step ranges are admitted, stepped and
retired around it; admission respects the device-memory budget."""
import time

import numpy as np
import pytest

import adfdtd as A
from adfdtd import _lib as L
from cases import build

pytestmark = pytest.mark.gpu

H = 0.00686


def requests():
    """(case, dtype code, mask spec, field, start, end) of a mixed stream."""
    out = []
    c = build(n=(48, 20, 22), nt=70, npml=6, n_rec=12, rec_radius_cells=7, src_cells=(2.3, -1.1, 0.4))
    out.append((c, L.ADF_F32, ("sphere", [4 * H]), None, 1, 70))
    c = build(n=(96, 18, 16), nt=60, npml=3, n_rec=10, rec_radius_cells=6, src_cells=(-3.2, 0.7, 1.9),
              src_dir=(0.0, 0.6, 0.0, 0.8))
    out.append((c, L.ADF_F64, None, None, 1, 60))
    c = build(n=(160, 14, 12), nt=50, npml=4, shape="cube", shape_params=[5 * H], n_rec=8, rec_radius_cells=4)
    out.append((c, L.ADF_F32, "array", None, 1, 50))
    c = build(n=(128, 12, 13), nt=80, npml=2, n_rec=6, rec_radius_cells=4, src_cells=(1.5, 0.5, -0.5))
    out.append((c, L.ADF_F32, ("ellipsoid", [4 * H, 3 * H, 2.5 * H]), None, 1, 80))
    # a hologram-like request: initial plane-wave field, receivers only, resumed at step 11
    c = build(n=(64, 16, 16), nt=40, npml=3, n_rec=10, rec_radius_cells=5, src_dir=(0.0, 0.0, 0.0, 0.0))
    f0 = A.initial_field(c[0], 2000.0, 0.05).astype(np.float32)
    out.append((c, L.ADF_F32_REF, None, f0, 11, 40))
    c = build(n=(32, 24, 20), nt=90, npml=5, n_rec=9, rec_radius_cells=5, src_cells=(0.2, 0.1, -0.3))
    out.append((c, L.ADF_F64, ("cross", [9 * H]), None, 1, 90))
    return out


def solo(req):
    (d, txs, cpml, mask), code, mspec, field, start, end = req
    T = np.float64 if code == L.ADF_F64 else np.float32
    s = A.Solver(d, code, A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, T))
    if isinstance(mspec, tuple):
        s.build_mask(*mspec)
    elif mspec == "array":
        s.set_mask(mask)
    if field is not None:
        s.set_field(field)
    s.set_transceivers(A.transceivers_to_grid(d, txs, T))
    s.run(start, end)
    return [s.get_trace(w) for w in range(10)], s.get_field()


def submit(sv, req, want=0x3FF, keep_field=True):
    (d, txs, cpml, mask), code, mspec, field, start, end = req
    T = np.float64 if code == L.ADF_F64 else np.float32
    m = mspec if isinstance(mspec, tuple) else (mask if mspec == "array" else None)
    return sv.submit(d, code, A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, T), A.transceivers_to_grid(d, txs, T),
                     mask=m, field=field, start_step=start, end_step=end, want=want, keep_field=keep_field)


def test_each_result_is_the_same_alone_and_in_the_mix():
    reqs = requests()
    ref = [solo(r) for r in reqs]
    # a small budget (about two of these solves) forces queueing, admission
    # as memory frees up, and members at different steps in one batch
    with A.Server(budget_bytes=5 << 20, chunk=7) as sv:
        ids = []
        for i, r in enumerate(reqs):
            ids.append(submit(sv, r))
            if i % 2:
                time.sleep(0.02)  # arrivals while others run
        for rid in ids:
            assert sv.wait(rid, 120.0) == "done"
        for i, (rid, (tr, fld)) in enumerate(zip(ids, ref)):
            for w in range(10):
                assert np.array_equal(sv.trace(rid, w), tr[w]), (i, w)
            assert np.array_equal(sv.field(rid), fld), i
            sv.release(rid)
        st = sv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["peak_bytes"] <= 5 << 20
        assert st["p50_ms"] > 0 and st["p99_ms"] >= st["p50_ms"]


def test_admission_by_budget_and_oversized_request_runs_alone():
    reqs = requests()[:3]
    with A.Server(budget_bytes=1, chunk=16) as sv:  # every request exceeds the budget
        ids = [submit(sv, r, want=0x1F, keep_field=False) for r in reqs]
        for rid in ids:
            assert sv.wait(rid, 120.0) == "done"
        st = sv.stats()
        assert st["completed"] == 3
        # one at a time: the peak is the largest single footprint
        assert st["running"] == 0 and st["queued"] == 0


def test_bad_requests_fail_cleanly():
    (d, txs, cpml, mask), code, *_ = requests()[0]
    with A.Server() as sv:
        with pytest.raises(L.AdfError, match="steps"):
            sv.submit(d, code, A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, np.float32),
                      A.transceivers_to_grid(d, txs, np.float32), start_step=5, end_step=d.nt + 1)
        # rows of more than 63 segments of 62 x 4 cells
        wide = A.make_domain(1e-3, 10.0, 0.02, 0.02, 15625, 4, 4, 343.0, 1.21, 0.5)
        with pytest.raises(L.AdfError, match="nx <= 15624"):
            sv.submit(wide, L.ADF_F32, None, None)
        rid = submit(sv, requests()[0], want=0x1, keep_field=False)
        assert sv.wait(rid, 60.0) == "done"
        with pytest.raises(L.AdfError, match="not requested"):
            sv.trace(rid, 4)
        with pytest.raises(L.AdfError, match="keep_field"):
            sv.field(rid)
