"""Impulsive-source Green's functions, mirroring src/greens.jl.

`impulsive_gfs(dom, cfg; scatterer)` (src/greens.jl:48-234) runs two FDTD
solves per outer Fibonacci point (pressure monopole [1,0,0,0], then velocity
monopole [0, n]) recording p and vn on the inner sphere.  Here the solves of
many sources advance together in one batch (adf_batch_*), and each finished
solve's traces are post-processed on the device (adf_solver_gf_resample:
self-record dropped, denoise, cubic resample onto the GF time grid).

Layouts follow the rest of the package: Julia column-major arrays with the
axes reversed, so the reference's p_p (nt_gf, n_rec, n_src) is a numpy
(n_src, n_rec, nt_gf) float32 array.
"""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np

from . import _lib as L
from .geometry import SCATTERER_DEFAULTS
from .illumination import illumination_points, impulsive_wavelet
from .simulation import Domain, concat_grids, make_cpml, make_transceiver, transceivers_to_grid
from .solver import Solver, dtype_code

_SHAPE_PARAM = {"sphere": "sphere_radius", "cube": "cube_length", "cross": "cross_length"}


def scatterer_params(scatterer: str, cfg: dict):
    """src/geometry.jl:16-20 (`_scatterer_param`): the size of `scatterer`
    from cfg["surfaces"], or the package default."""
    if scatterer == "none":
        return None
    key = _SHAPE_PARAM[scatterer]
    s = cfg.get("surfaces", {})
    return [float(s.get(key, SCATTERER_DEFAULTS[key]))]


def impulsive_gfs(dom: Domain, cfg: dict, scatterer: str = "cross", precision: str = "fast",
                  dtype=np.float32, on_source_complete: Optional[Callable] = None, init: Optional[dict] = None,
                  sources_per_batch: int = 32) -> dict:
    """src/greens.jl:48-234.  cfg keys as in the reference: greens.f_3db_hz,
    greens.nt_save, pml.{n, rcoef, fc}, surfaces.{nPoints_inner,
    radius_inner, nPoints_outer, radius_outer} and the scatterer size.
    Returns {p_p, p_v, v_p, v_v (n_src, n_rec, nt_gf) float32, t (nt_gf,),
    src_positions, rec_positions} (float32).  Sources whose four slices in
    `init` hold any non-zero value are skipped (greens.jl:128-133);
    on_source_complete(iSrc, state) is called after each source, iSrc
    1-based, in source order."""
    from .batch import Batch

    params = scatterer_params(scatterer, cfg)
    f_3db = float(cfg["greens"]["f_3db_hz"])
    srctf, timeshift = impulsive_wavelet(dom, f_3db)
    rectf = np.zeros(dom.nt)
    pts = illumination_points(cfg)
    src_pos, src_nrm = pts["outer"], pts["outer_normals"]
    rec_pos, rec_nrm = pts["inner"], pts["inner_normals"]
    n_src, n_rec = src_pos.shape[0], rec_pos.shape[0]
    receivers = [make_transceiver(dom, rec_pos[i], np.r_[0.0, rec_nrm[i]], rectf) for i in range(n_rec)]
    pml = cfg["pml"]
    npml, rcoef, fc = int(pml["n"]), float(pml["rcoef"]), float(pml["fc"])
    nt_gf = int(cfg["greens"]["nt_save"])
    t_gf = np.linspace(0.0, dom.tmax, nt_gf)

    shape = (n_src, n_rec, nt_gf)
    if init is None:
        out = {k: np.zeros(shape, np.float32) for k in ("p_p", "p_v", "v_p", "v_v")}
    else:
        out = {}
        for k in ("p_p", "p_v", "v_p", "v_v"):
            a = np.asarray(init[k], np.float32)
            if a.shape != shape:
                raise ValueError(f"impulsive_gfs: init.{k} shape {a.shape} does not match expected {shape} "
                                 "(n_src, n_rec, nt_gf). Was it built with a different config?")
            out[k] = a.copy()

    def done(i):
        return any(np.any(out[k][i] != 0) for k in ("p_p", "p_v", "v_p", "v_v"))

    todo = [i for i in range(n_src) if not done(i)]
    code = dtype_code(dtype, precision)

    # every run packs the same receivers behind its source: pack them once
    rec_grid = transceivers_to_grid(dom, receivers, dtype)

    def make_solver(i, kind):
        direction = [1.0, 0.0, 0.0, 0.0] if kind == "p" else np.r_[0.0, src_nrm[i]]
        src = make_transceiver(dom, src_pos[i], direction, srctf)
        s = Solver(dom, code, make_cpml(dom, npml, rcoef, fc, dtype))
        if params is not None:
            s.build_mask(scatterer, params)
        s.set_transceivers(concat_grids(transceivers_to_grid(dom, [src], dtype), rec_grid))
        return s

    for c0 in range(0, len(todo), max(1, int(sources_per_batch))):
        chunk = todo[c0:c0 + max(1, int(sources_per_batch))]
        sol = {(i, k): make_solver(i, k) for i in chunk for k in ("p", "v")}
        b = Batch()
        try:
            for s in sol.values():
                b.add(s)
            batched = True
        except L.AdfError:  # grids the batched sweep does not take: one solve at a time
            batched = False
        if batched:
            b.run(dom.nt)
        del b
        if not batched:
            for s in sol.values():
                s.run(1, dom.nt)
        for i in chunk:
            gp, gv = sol[(i, "p")].gf_resample(1, n_rec, timeshift, nt_gf)
            out["p_p"][i], out["v_p"][i] = gp, gv
            gp, gv = sol[(i, "v")].gf_resample(1, n_rec, timeshift, nt_gf)
            out["p_v"][i], out["v_v"][i] = gp, gv
            if on_source_complete is not None:
                on_source_complete(i + 1, {**out, "t": t_gf, "src_positions": src_pos, "rec_positions": rec_pos,
                                           "n_src": n_src})
        del sol
    return {**out, "t": t_gf.astype(np.float32), "src_positions": src_pos.astype(np.float32),
            "rec_positions": rec_pos.astype(np.float32)}
