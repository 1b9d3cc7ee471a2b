"""Reverberating-source data for MDD, mirroring src/reverb.jl.

`build_reverb_data(dom, cfg; scatterer, ...)` (src/reverb.jl:72-417) fires a
pressure monopole from each point of an illumination sphere (Roberts' R2
sequence, src/illumination.jl:69-97), usually inside a closed box
(`[reverb.pml].n = 0`), and records p and the normal velocity on the outer
and inner control surfaces.  Each source's traces are cast to Float32 (vn
times z0) and resampled (natural cubic spline, no amplitude scaling) onto
`0:1/fs_out:tmax` (src/reverb.jl:424-452).

Here the fresh sources of a sweep advance together in batches (adf_batch_*),
and each finished source's resample runs on the device (adf_resample_f32).
Sources resumed from a saved field (`n_saved`) or asked to record field
snapshots run one at a time.

Layouts follow the rest of the package: Julia column-major arrays with the
axes reversed, so the reference's outer_p (nt_ill, n_outer, n_ill) is a
numpy (n_ill, n_outer, nt_ill) float32 array, and a per-source trace block
(dom.nt, n_outer) is (n_outer, nt).
"""
from __future__ import annotations

import math
import os
from fractions import Fraction
from typing import Callable, Optional

import numpy as np

from . import _lib as L
from .greens import scatterer_params
from .illumination import illumination_points, r2_sphere, ricker_wavelet
from .simulation import Domain, make_cpml, make_transceiver, transceivers_to_grid
from .solver import Solver, context, dtype_code


# ------------------------------------------------------------ Julia ranges


def _jl_rat(x: float):
    """Julia's Base.rat (base/twiceprecision.jl): the continued-fraction
    approximation a/b of x with |a|, |b| <= maxintfloat(Float32) that stops
    at the first convergent whose Float64 quotient equals x."""
    y = float(x)
    a = d = 1
    b = c = 0
    m = 16777216
    while abs(y) <= m:
        f = math.trunc(y)
        y -= f
        a, c = f * a + c, a
        b, d = f * b + d, b
        if max(abs(a), abs(b)) > m:
            return c, d
        if float(a) / float(b) == x:
            break
        y = math.inf if y == 0.0 else 1.0 / y
    return a, b


def _between(a, x, b):
    return a <= x <= b or b <= x <= a


def _tdiv(a: int, b: int) -> int:
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def julia_step_range(start: float, step: float, stop: float) -> np.ndarray:
    """The Float64 range `start:step:stop` as Julia builds it (base/
    twiceprecision.jl `(:)(start::T, step::T, stop::T)`): when start, step
    and stop have exact small rational forms, the length and the elements
    come from those rationals (so `0.0:1/60000:0.005` has 301 elements);
    otherwise length round((stop-start)/step)+1, less one on overshoot, and
    elements start + k step.  Elements are the correctly rounded values."""
    start, step, stop = float(start), float(step), float(stop)
    if step == 0:
        raise ValueError("range step cannot be zero")
    step_n, step_d = _jl_rat(step)
    if step_d != 0 and step_n / step_d == step:
        start_n, start_d = _jl_rat(start)
        stop_n, stop_d = _jl_rat(stop)
        if start_d != 0 and stop_d != 0 and start_n / start_d == start and stop_n / stop_d == stop:
            den = math.lcm(start_d, step_d)
            m = 2.0 ** 53
            if den != 0 and abs(start * den) <= m and abs(step * den) <= m and den % start_d == 0 \
                    and den % step_d == 0:
                sn = round(start * den)
                tn = round(step * den)
                n = max(0, _tdiv(den * stop_n - stop_d * sn + tn * stop_d, tn * stop_d))
                if _between(start, start + (n - 1) * step, stop + step / 2) and \
                        not _between(start, start + n * step, stop):
                    return np.array([float(Fraction(sn + k * tn, den)) for k in range(n)])
    lf = (stop - start) / step
    if lf < 0:
        n = 0
    elif lf == 0:
        n = 1
    else:
        n = round(lf) + 1
        stp = start + (n - 1) * step
        n -= int(start < stop < stp) + int(start > stop > stp)
    fs, ft = Fraction(start), Fraction(step)
    return np.array([float(fs + k * ft) for k in range(n)])


# ---------------------------------------------------------- device resample


def resample_traces(traces, t_in0: float, dt_in: float, t_out0: float, dt_out: float, n_out: int,
                    amp_scale: float = 1.0, denoise: bool = False, device: int = 0) -> np.ndarray:
    """src/resample.jl:35-77 (kind = :cubic) on the device along the last
    axis of float32 traces (..., n_in) -> (..., n_out) float32: natural cubic
    B-spline on t_in0 + k dt_in, Flat outside, sampled at t_out0 + k dt_out,
    times amp_scale (adf_resample_f32)."""
    a = np.ascontiguousarray(traces, dtype=np.float32)
    n_in = a.shape[-1]
    ncol = int(np.prod(a.shape[:-1], dtype=np.int64))
    out = np.empty(a.shape[:-1] + (int(n_out),), np.float32)
    lib = L.load()
    L.check(lib.adf_resample_f32(context(device), L.ptr(a), n_in, ncol, float(t_in0), float(dt_in), float(t_out0),
                                 float(dt_out), int(n_out), float(amp_scale), int(bool(denoise)), L.ptr(out)),
            "adf_resample_f32")
    return out


def _resample_isrc(out: dict, isrc: int, full: dict, tin0: float, tin_last: float, tout: np.ndarray) -> None:
    """src/reverb.jl:424-452: the raw traces (dom.dt rate) of source isrc
    (0-based) onto tout (float32), without amplitude scaling, into slot
    isrc of the outputs.  tin = range(tin0, tin_last; length = n_in), tout
    range(tout[1], tout[end]; length = length(tout))."""
    n_in = full["outer_p"].shape[-1]
    keys = ("outer_p", "inner_p", "outer_vnz", "inner_vnz")
    A = np.concatenate([full[k] for k in keys], axis=0)
    dt_in = (tin_last - tin0) / (n_in - 1) if n_in > 1 else 1.0
    n_out = tout.shape[0]
    t0, t1 = float(tout[0]), float(tout[-1])
    dt_out = (t1 - t0) / (n_out - 1) if n_out > 1 else 1.0
    R = resample_traces(A, tin0, dt_in, t0, dt_out, n_out)
    o = 0
    for k in keys:
        n = full[k].shape[0]
        out[k][isrc] = R[o:o + n]
        o += n


# ------------------------------------------------------- build_reverb_data


def _stop_requested(path: Optional[str]) -> bool:
    return path is not None and os.path.exists(path)


def build_reverb_data(dom: Domain, cfg: dict, scatterer: str = "cross", n_ill: Optional[int] = None,
                      on_source_complete: Optional[Callable] = None, n_saved: int = 0,
                      load_isrc_saved: Optional[Callable] = None, save_isrc_done: Optional[Callable] = None,
                      save_field: bool = True, snapshot_steps=(), snapshot_isrc=None, prefill: Optional[dict] = None,
                      stopfile: Optional[str] = None, precision: str = "fast", dtype=np.float32,
                      sources_per_batch: int = 32) -> dict:
    """src/reverb.jl:72-417.  cfg keys as in the reference: reverb.{radius_ill,
    fc_source, bandwidth_factor (required), nPoints_ill}, pml.{n, rcoef},
    surfaces.{nPoints_inner, radius_inner, nPoints_outer, radius_outer} and
    the scatterer size.  Sources are 1-based in the callbacks, as in the
    reference:

    - load_isrc_saved(iSrc) -> None or {field (4, nz, ny, nx), outer_p,
      outer_vnz, inner_p, inner_vnz (n_rec, n_saved) float32}: a saved
      source is resampled as is when n_saved == dom.nt, else continued from
      step n_saved + 1 from its field with a zero source wavelet;
    - prefill: {outer_p, outer_vnz, inner_p, inner_vnz} of shape (n_prefill,
      n_rec, nt_ill); populated slots are copied and skipped;
    - save_isrc_done(iSrc, {field, outer_p, outer_vnz, inner_p, inner_vnz})
      with the full-length float32 traces and (save_field) the final field;
    - on_source_complete(iSrc, state) after each source;
    - snapshot_steps / snapshot_isrc: fields at those steps for those
      sources (all when snapshot_isrc is None), in state["snapshots"];
    - stopfile: the sweep stops before the next batch of sources (before
      the next source, for sources run one at a time) once the file exists.

    Returns {inner_p, inner_vnz, outer_p, outer_vnz (n_ill, n_rec, nt_ill),
    t (nt_ill,), ill_positions, inner_positions, outer_positions} float32."""
    from .batch import Batch

    rev = cfg.get("reverb", {})
    r_ill = float(rev.get("radius_ill", 0.6))
    fc = float(rev.get("fc_source", 18000.0))
    bw = rev.get("bandwidth_factor")
    if bw is None:
        raise ValueError("[reverb].bandwidth_factor is required -- it sets the recorded Ricker band: "
                         "fs_out = 2*bandwidth_factor*fc_source.")
    fs_out = 2.0 * float(bw) * fc
    n_ill_ = int(n_ill) if n_ill is not None else int(rev.get("nPoints_ill", cfg["surfaces"]["nPoints_outer"]))
    pml = cfg["pml"]
    npml = int(pml["n"])
    if n_saved > 0 and npml > 0:
        raise ValueError("build_reverb_data: resume / extend is only supported with [pml].n = 0 "
                         f"(CPML memory is not checkpointed). Got pml.n = {npml}.")
    if n_saved > dom.nt:
        raise ValueError(f"build_reverb_data: n_saved={n_saved} exceeds dom.nt={dom.nt}. Refusing to shrink tmax.")

    params = scatterer_params(scatterer, cfg)
    srctf_full = ricker_wavelet(dom, fc, True)
    srctf_zero = np.zeros(dom.nt)
    rectf = np.zeros(dom.nt)

    ill_points, _ = r2_sphere(n_ill_, r_ill)
    pts = illumination_points(cfg)
    inner_pos, inner_nrm = pts["inner"], pts["inner_normals"]
    outer_pos, outer_nrm = pts["outer"], pts["outer_normals"]
    n_inner, n_outer = inner_pos.shape[0], outer_pos.shape[0]
    outer_txs = [make_transceiver(dom, outer_pos[i], np.r_[0.0, outer_nrm[i]], rectf) for i in range(n_outer)]
    inner_txs = [make_transceiver(dom, inner_pos[i], np.r_[0.0, inner_nrm[i]], rectf) for i in range(n_inner)]
    n_rec = n_outer + n_inner
    rcoef = float(pml["rcoef"])

    t_ill = julia_step_range(0.0, 1.0 / fs_out, dom.tmax)
    nt_ill = t_ill.shape[0]
    t_ill_arr = t_ill.astype(np.float32)
    shapes = {"outer_p": n_outer, "outer_vnz": n_outer, "inner_p": n_inner, "inner_vnz": n_inner}
    out = {k: np.zeros((n_ill_, n, nt_ill), np.float32) for k, n in shapes.items()}

    n_prefill = 0
    if prefill is not None:
        n_prefill = np.shape(prefill["outer_p"])[0]
        if n_prefill > n_ill_:
            raise ValueError(f"build_reverb_data: prefill has {n_prefill} iSrc but n_ill={n_ill_}")
        for k, n in shapes.items():
            a = np.asarray(prefill[k], np.float32)
            if a.shape != (n_prefill, n, nt_ill):
                raise ValueError(f"build_reverb_data: prefill.{k} shape {a.shape} does not match "
                                 f"(n_prefill={n_prefill}, n_rec={n}, nt_ill={nt_ill})")
            out[k][:n_prefill] = a

    z0 = dom.z0
    tmax = dom.tmax
    dt_t = tmax / (dom.nt - 1) if dom.nt > 1 else 0.0  # step of t(dom)
    code = dtype_code(dtype, precision)
    snap_set = set(int(s) for s in snapshot_steps)

    def cast(p_tr, vn_tr):
        """Float32(p) and Float32(vn) * z0 stored as Float32 (reverb.jl:335-352)."""
        return (np.asarray(p_tr).astype(np.float32),
                (np.asarray(vn_tr).astype(np.float32).astype(np.float64) * z0).astype(np.float32))

    # first pass: saved and prefilled sources (reverb.jl:236-257)
    fully_done = [False] * n_ill_
    has_saved = [False] * n_ill_
    for i in range(n_ill_):
        saved = load_isrc_saved(i + 1) if load_isrc_saved is not None else None
        if saved is not None:
            has_saved[i] = True
            fully_done[i] = n_saved >= dom.nt
            full = {k: np.asarray(saved[k], np.float32) for k in shapes}
            ns = full["outer_p"].shape[-1]
            _resample_isrc(out, i, full, 0.0, float(Fraction(tmax) * (ns - 1) / (dom.nt - 1)) if ns > 1 else 0.0,
                           t_ill_arr)
        elif i < n_prefill and np.any(out["outer_p"][i] != 0):
            fully_done[i] = True

    def make_solver(i, resumed):
        tf = srctf_zero if resumed else srctf_full
        src = make_transceiver(dom, ill_points[i], [z0, 0.0, 0.0, 0.0], tf)
        s = Solver(dom, code, make_cpml(dom, npml, rcoef, fc, dtype))
        if params is not None:
            s.build_mask(scatterer, params)
        s.set_transceivers(transceivers_to_grid(dom, outer_txs + inner_txs + [src], dtype))
        return s

    def finish(i, s, saved, start_step, snapshots):
        p_tr, vn_tr = cast(s.get_trace(L.ADF_REC_P)[:n_rec], s.get_trace(L.ADF_VN_REC)[:n_rec])
        full = {"outer_p": p_tr[:n_outer].copy(), "inner_p": p_tr[n_outer:].copy(),
                "outer_vnz": vn_tr[:n_outer].copy(), "inner_vnz": vn_tr[n_outer:].copy()}
        if saved is not None:
            for k in shapes:
                full[k][:, :n_saved] = np.asarray(saved[k], np.float32)[:, :n_saved]
        if not (np.all(np.isfinite(full["outer_p"])) and np.all(np.isfinite(full["inner_p"]))):
            import warnings
            warnings.warn(f"build_reverb_data: non-finite trace values (iSrc {i + 1})")
        final_field = s.get_field().astype(np.float32) if save_field else None
        if save_isrc_done is not None:
            save_isrc_done(i + 1, {"field": final_field, **full})
        _resample_isrc(out, i, full, 0.0, tmax, t_ill_arr)
        if on_source_complete is not None:
            on_source_complete(i + 1, {**out, "t": t_ill_arr, "ill_positions": ill_points.astype(np.float32),
                                       "inner_positions": inner_pos.astype(np.float32),
                                       "outer_positions": outer_pos.astype(np.float32), "n_ill": n_ill_,
                                       "snapshots": snapshots, "dom": dom})

    def run_one(i):
        saved = load_isrc_saved(i + 1) if has_saved[i] else None
        resumed = saved is not None
        s = make_solver(i, resumed)
        start = n_saved + 1 if resumed else 1
        if resumed:
            s.set_field(np.asarray(saved["field"], dtype))
        snaps = {}
        capture = snapshot_isrc is None or (i + 1) in snapshot_isrc
        cur = start
        for st in sorted(snap_set) if capture else ():
            if cur <= st <= dom.nt:
                s.run(cur, st)
                cur = st + 1
                snaps[st] = s.get_field().astype(np.float32)
        if cur <= dom.nt:
            s.run(cur, dom.nt)
        finish(i, s, saved, start, snaps)
        del s

    todo = [i for i in range(n_ill_) if not fully_done[i]]
    solo = set(i for i in todo if has_saved[i] or (snap_set and (snapshot_isrc is None or (i + 1) in snapshot_isrc)))
    # sources in the reference's order: solo sources run when their turn comes
    step = max(1, int(sources_per_batch))
    k = 0
    while k < len(todo):
        if _stop_requested(stopfile):
            break
        i = todo[k]
        if i in solo:
            run_one(i)
            k += 1
            continue
        chunk = []
        while k < len(todo) and todo[k] not in solo and len(chunk) < step:
            chunk.append(todo[k])
            k += 1
        sol = {i: make_solver(i, False) for i in chunk}
        b = Batch()
        try:
            for s in sol.values():
                b.add(s)
            ok = True
        except L.AdfError:  # grids the batched sweep does not take: one solve at a time
            ok = False
        if ok:
            b.run(dom.nt)
        del b
        for i in chunk:
            if not ok:
                sol[i].run(1, dom.nt)
            finish(i, sol[i], None, 1, {})
        del sol
    return {**out, "t": t_ill_arr, "ill_positions": ill_points.astype(np.float32),
            "inner_positions": inner_pos.astype(np.float32), "outer_positions": outer_pos.astype(np.float32)}
