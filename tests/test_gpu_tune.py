"""Autotuner: the chosen tiling is one of the candidates, is cached in the
JSON file, and leaves results unchanged (tiling changes speed only)."""
import json

import numpy as np
import pytest

import adfdtd as A
from adfdtd import autotune
from cases import build

pytestmark = pytest.mark.gpu


def _solver(case, dtype=np.float32):
    d, txs, cpml, mask = case
    s = A.Solver(d, A.dtype_code(dtype, "fast"), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, dtype))
    s.set_mask(mask)
    return s, A.transceivers_to_grid(d, txs, dtype)


def test_tune_solver_cached_and_result_invariant(tmp_path):
    path = str(tmp_path / "tune.json")
    case = build(n=(96, 40, 36), nt=50, npml=5, shape="sphere", shape_params=[6 * 0.00686], n_rec=10,
                 rec_radius_cells=8)
    ref, g = _solver(case)
    ref.set_transceivers(g)
    ref.run()
    s, g = _solver(case)
    shape = autotune.tune_solver(s, steps=6, path=path)
    keys = ("tile_y", "block_z", "block_z_gen", "waves", "pipeline", "launch")
    assert {k: shape[k] for k in keys} in autotune.SOLVER_CANDIDATES
    assert shape.get("store_nt", 0) in (0, 1)
    data = json.load(open(path))
    assert len(data) == 1 and next(iter(data)).startswith("solver:")
    assert autotune.tune_solver(s, path=path) == shape  # cache hit
    s.set_transceivers(g)
    s.run()
    for w in range(5):
        assert np.array_equal(s.get_trace(w), ref.get_trace(w))
    assert np.array_equal(s.get_field(), ref.get_field())


def test_tune_batch_cached_and_result_invariant(tmp_path):
    path = str(tmp_path / "tune.json")
    cases = [build(n=(64, 24, 20), nt=40, npml=4, n_rec=6, rec_radius_cells=5),
             build(n=(128, 16, 18), nt=40, npml=3, shape="cube", shape_params=[4 * 0.00686], n_rec=6,
                   rec_radius_cells=4)]
    refs = []
    for c in cases:
        s, g = _solver(c)
        s.set_transceivers(g)
        s.run()
        refs.append(s)
    mem = []
    for c in cases:
        s, g = _solver(c)
        s.set_transceivers(g)
        mem.append(s)
    b = A.Batch()
    cands = autotune.BATCH_CANDIDATES[::7]
    shape = autotune.tune_batch(mem, steps=6, candidates=cands, path=path, batch=b)
    assert {k: shape[k] for k in ("tile_y", "block_z", "block_z_gen", "waves", "persistent", "pipeline")} in cands
    assert next(iter(json.load(open(path)))).startswith("batch:")
    for s in mem:
        b.add(s)
    b.run(cases[0][0].nt)
    for s, r in zip(mem, refs):
        for w in range(5):
            assert np.array_equal(s.get_trace(w), r.get_trace(w))
