"""Tile-shape autotuner for the fused sweep, with a JSON cache.

The sweep's tiling (planes per interior tile, planes per side tile, waves
per workgroup, the pipelined routines, the launch form and, for
batches, the persistent tile queue) changes only speed, never results.  The tuner times the candidates on
scratch copies of the solves (same grids, CPML, masks and dtype, zero
fields, no transceivers), so the caller's state is untouched, and stores
the best shape per workload signature in a JSON file
($ADFDTD_TUNE_CACHE, default ~/.cache/adfdtd/tune.json).
"""
from __future__ import annotations

import itertools
import json
import os
import threading
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from . import _lib as L
from .solver import Solver

_lock = threading.Lock()

# block_z: planes per interior tile; block_z_gen: planes per side tile (CPML
# slabs and boundary); pipeline: the prefetching tile routines; launch: one
# launch over all tiles, or side and interior tiles as two launches on two
# streams at the same time (single solves; ADF_SWEEP_*); store_nt: the
# non-temporal store hint, tried on the best tiling of a single solve or a
# batch; cpl:
# cells per lane (ADF_OPT_CELLS_PER_LANE), 2 tried beside the default for
# rows wider than 256 cells
SOLVER_CANDIDATES = [dict(tile_y=0, block_z=kb, block_z_gen=kbg, waves=1, pipeline=pf, launch=ln)
                     for kb, kbg, pf, ln in itertools.product((4, 8, 16, 32), (2, 4), (1, 0), (0, 1))]
BATCH_CANDIDATES = [dict(tile_y=0, block_z=kb, block_z_gen=kbg, waves=1, persistent=pers, pipeline=pf)
                    for kb, kbg, pers, pf in itertools.product((4, 8, 16), (2, 4, 8), (0, 1), (1, 0))]


def cache_path() -> str:
    return os.environ.get("ADFDTD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "adfdtd",
                                                             "tune.json"))


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, key: str, value: dict) -> None:
    with _lock:
        data = _load(path)
        data[key] = value
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def _signature(s: Solver) -> str:
    d = s.dom
    masked = int(getattr(s, "_masked", False))
    return f"{s.code}:{d.nx}x{d.ny}x{d.nz}:p{s.npml}:m{masked}"


def _scratch(s: Solver) -> Solver:
    """A solver with the same grid, dtype, CPML and mask as s."""
    x = Solver(s.dom, s.code, s._cpml, s.device)
    if getattr(s, "_masked", False):
        x.set_mask(s.get_mask())
    return x


def _apply(s: Solver, shape: dict) -> None:
    """Set a shape's options on s; a shape without a store hint leaves the
    automatic one (ADF_OPT_STORE_NT = -1), so candidates are timed with the
    stores they would run with."""
    s.set_path("auto", shape["tile_y"], shape["block_z"], shape["waves"], shape["block_z_gen"],
               pipeline=shape.get("pipeline", 1), launch=shape.get("launch", L.ADF_SWEEP_ONE_LAUNCH),
               store_nt=shape.get("store_nt", -1), cells_per_lane=shape.get("cpl", 0))


def tune_solver(s: Solver, steps: int = 20, candidates: Optional[Sequence[dict]] = None,
                path: Optional[str] = None, use_cache: bool = True) -> dict:
    """Pick the fused-sweep tiling of solver s (by the mean device time of
    the field-update launches) and apply it; returns the chosen shape."""
    path = path or cache_path()
    key = "solver:" + _signature(s)
    if use_cache:
        hit = _load(path).get(key)
        if hit:
            _apply(s, hit)
            return hit
    x = _scratch(s)
    x.set_option(L.ADF_OPT_TIME_SWEEP, 1)
    n = max(4, min(steps, s.dom.nt // 2))
    best, best_ms = None, float("inf")
    if candidates is None:
        candidates = SOLVER_CANDIDATES
        if s.dom.nx > 256:  # rows of several waves: 2 cells per lane fill more lanes per segment
            candidates = candidates + [dict(c, cpl=2) for c in SOLVER_CANDIDATES]
        tuned = True
    else:
        tuned = False
    for cand in candidates:
        try:
            _apply(x, cand)
            x.reset()
            x.run(1, 2)                   # builds the tile lists
            x.run(3, 2 + n)
        except L.AdfError:
            continue                      # shape not valid for this grid
        ms, _ = x.sweep_stats()
        if ms < best_ms:
            best, best_ms = dict(cand), ms
    if best is not None and best.get("pipeline", 1) and tuned:
        # the store hint on the best tiling, both timed again over more steps
        timed = {}
        m = max(1, min(2 * n, s.dom.nt - 2))
        for nt_ in (0, 1):
            _apply(x, dict(best, store_nt=nt_))
            x.reset()
            x.run(1, 2)
            x.run(3, 2 + m)
            timed[nt_], _ = x.sweep_stats()
        best["store_nt"] = int(timed[1] < timed[0])
        best_ms = timed[best["store_nt"]]
    del x
    if best is None:
        best = dict(SOLVER_CANDIDATES[0])
    best["ms"] = best_ms
    _store(path, key, best)
    _apply(s, best)
    return best


def tune_batch(members: Sequence[Solver], steps: int = 32, candidates: Optional[Sequence[dict]] = None,
               path: Optional[str] = None, use_cache: bool = True, batch=None) -> dict:
    """Pick the tiling of a batch of these solvers (by the device time per
    batched step on scratch copies) and apply it to every member (and to
    `batch`'s persistent-queue option when given); returns the shape."""
    from .batch import Batch

    path = path or cache_path()
    sigs = sorted(_signature(s) for s in members)
    counts = {}
    for g in sigs:
        counts[g] = counts.get(g, 0) + 1
    key = "batch:" + ",".join(f"{g}*{c}" for g, c in sorted(counts.items()))

    def apply_all(shape):
        for s in members:
            _apply(s, shape)
        if batch is not None:
            batch.set_option(L.ADF_BATCH_PERSISTENT, shape["persistent"])

    if use_cache:
        hit = _load(path).get(key)
        if hit:
            apply_all(hit)
            return hit
    scratch = [_scratch(s) for s in members]
    nmin = min(s.dom.nt for s in members)
    n = max(4, min(steps, nmin // 2 - 1))
    best, best_ms = None, float("inf")
    for cand in candidates or BATCH_CANDIDATES:
        b = Batch(persistent=bool(cand["persistent"]), graph_steps=min(16, n))
        try:
            for x in scratch:
                x.reset()
                _apply(x, cand)
                b.add(x)
            b.run(min(16, n))             # builds tables and the graph
            b.run(n)
            ms = b.last_step_ms()
        except L.AdfError:
            ms = float("inf")
        finally:
            del b
        if ms < best_ms:
            best, best_ms = dict(cand), ms
    if best is not None and best.get("pipeline", 1) and candidates is None:
        # the store hint on the best tiling, both timed again
        timed = {}
        for nt_ in (0, 1):
            shape = dict(best, store_nt=nt_)
            b = Batch(persistent=bool(shape["persistent"]), graph_steps=min(16, n))
            try:
                for x in scratch:
                    x.reset()
                    _apply(x, shape)
                    b.add(x)
                b.run(min(16, n))
                b.run(n)
                timed[nt_] = b.last_step_ms()
            except L.AdfError:
                timed[nt_] = float("inf")
            finally:
                del b
        best["store_nt"] = int(timed[1] < timed[0])
        best_ms = timed[best["store_nt"]]
    del scratch
    best = best or dict(BATCH_CANDIDATES[0])
    best["ms"] = best_ms
    _store(path, key, best)
    apply_all(best)
    return best
