#!/usr/bin/env python3
"""Benchmark of the 3-D staggered-grid acoustic FDTD step on one MI355X.

Workload (one step = one forward_onestep! of the reference):
  256^3 grid, air (c = 343 m/s, rho = 1.21 kg/m^3), dx = lambda_min/10 for a
  2 kHz Ricker (lambda_min = c / (2.5 fc)), cf = 0.5, 16-cell CPML,
  rigid sphere mask of radius 24 cells at the centre, Ricker pressure source
  60 cells off-centre, p and vn recorded every step on a 2000-point
  Fibonacci sphere of radius 40 cells plus 500 far-field receivers at
  80 cells.  Synthetic inputs, FP32 storage and arithmetic (the fast path).

Prints one JSON line: GLUPS of the whole job, the roofline of the dominant
kernel (the fused field-update sweep, timed with HIP events on its stream),
the FP64 (bit-reproducible) path's GLUPS, and the oracle (plain-C
restatement of the reference) timed on the host's cores as the CPU baseline.

Modes for profiling (see profiles/):
  --profile-steps N   run N steps only (no CPU leg), for rocprofv3
  --profile-batch N   run the 64 x 128^3 batched leg for N steps only
  --profile-gf        run the in-loop KH gather leg only
  --profile-hologram  run the 192^3 hologram synthesis leg only
  --profile-kh        run the post-hoc KH leg only (p_v, 2000 -> 500 points, 2000 steps)
  --profile-grid N    run the production-grid leg at N^3 only (301, 501 or 545;
                      --profile-steps K steps, default 50)
  --calibrate-copy    HBM ceilings of the library's streaming probe kernels
                      (copy/read/write/z-march; --probe-kind K for one kind, the
                      run the PMC scale factors are read from)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "acoustic-disguising_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
T_START = time.perf_counter()


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.perf_counter() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)
C_AIR, RHO_AIR, FC = 343.0, 1.21, 2000.0


SWEEP_SAMPLE = 16


def build_workload(n=256, npml=16, nt=2050):
    import adfdtd as A

    dx = C_AIR / (2.5 * FC) / 10
    xmax = (n - 1) * dx / 2
    d0 = A.make_domain(1.0, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
    d = A.make_domain((nt - 0.5) * d0.dt, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
    tf = A.ricker_wavelet(d, FC, True)
    kh, khn = A.fibonacci_sphere(2000, 40 * d.dx)
    ff, ffn = A.fibonacci_sphere(500, 80 * d.dx)
    txs = [A.make_transceiver(d, [60 * d.dx, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0], tf)]
    z = np.zeros(d.nt)
    txs += [A.make_transceiver(d, kh[i], np.r_[0.0, khn[i]], z) for i in range(2000)]
    txs += [A.make_transceiver(d, ff[i], np.r_[0.0, ffn[i]], z) for i in range(500)]
    return d, txs, npml


def batch_specs(n_solves, sizes, nt, seed=0):
    """Impulsive-GF solves of the batched workload: solves are split evenly
    over `sizes` (cells per axis); grid size 128 has a 10-cell CPML, a rigid
    sphere of radius 16 cells, its source on a Fibonacci sphere of radius 40
    cells (monopole, compact-Gaussian wavelet, f_3db = 2 kHz) and 100
    receivers (p and vn) at radius 20 cells; other sizes scale the radii by
    n/128."""
    import adfdtd as A

    dx = C_AIR / (2.5 * FC) / 10
    specs = []
    per = n_solves // len(sizes)
    for n in sizes:
        xmax = (n - 1) * dx / 2
        d0 = A.make_domain(1.0, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
        d = A.make_domain((nt - 0.5) * d0.dt, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
        sc = n / 128
        src, _ = A.fibonacci_sphere(per, 40 * sc * d.dx)
        rec, rn = A.fibonacci_sphere(100, 20 * sc * d.dx)
        tf, _ = A.impulsive_wavelet(d, FC)
        for i in range(per):
            specs.append({"d": d, "npml": 10, "sphere": 16 * sc * d.dx, "src": src[i], "tf": tf, "rec": rec,
                          "rec_n": rn})
    return specs


def make_batch_solver(sp, dtype=np.float32):
    import adfdtd as A

    d = sp["d"]
    s = A.Solver(d, A.dtype_code(dtype, "fast"), A.make_cpml(d, sp["npml"], 1e-4, FC, dtype))
    s.build_mask("sphere", [sp["sphere"]])
    z = np.zeros(d.nt)
    txs = [A.make_transceiver(d, sp["src"], [1.0, 0.0, 0.0, 0.0], sp["tf"])]
    txs += [A.make_transceiver(d, sp["rec"][i], np.r_[0.0, sp["rec_n"][i]], z) for i in range(len(sp["rec"]))]
    s.set_transceivers(A.transceivers_to_grid(d, txs, dtype))
    return s


def algorithmic_bytes_per_step(n, p, es, masked=True):
    """8 n^3 field words read+written, the CPML memory touched (read+write),
    and a 1-bit scatterer bitmap."""
    n_psi = 3 * ((2 * p - 1) * (n - 1) ** 2 + (2 * p + 1) * n ** 2) if p > 0 else 0
    return es * (8 * n ** 3 + 2 * n_psi) + (n ** 3 / 8 if masked else 0)


def read_traffic(name):
    """HBM bytes per launch of the sweep kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    fn = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(fn):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(fn).read(), re.M)
    return float(m.group(1)) if m else None


def sweep_sample_period(steps):
    """Every how many steps of a timed region the sweep is bracketed by HIP
    events: every 16th step on long regions, and at least 5 samples (at most
    every step) on short ones."""
    return max(1, min(SWEEP_SAMPLE, steps // 5))


def plan_steps(args):
    """Step ranges of the 256^3 legs: the domain's step count and, per leg,
    (warm-up first, warm-up last, timed first, timed last), 1-based and
    inclusive like adf_solver_run.  Every range lies inside 1..nt."""
    legs = {}
    for name, steps in (("f32", args.steps), ("f64", args.fp64_steps)):
        if steps > 0:
            legs[name] = (1, args.warmup, args.warmup + 1, args.warmup + steps)
    nt = max([r[3] for r in legs.values()] + [1]) + 1
    return nt, legs


def gpu_leg(args):
    import adfdtd as A
    from adfdtd import _lib as L

    nt, legs = plan_steps(args)
    d, txs, p = build_workload(nt=nt)
    out = {}
    for name, dtype, prec in (("f32", np.float32, "fast"), ("f64", np.float64, "fast")):
        if name not in legs:
            continue
        w0, w1, t_first, t_last = legs[name]
        steps = t_last - t_first + 1
        cp = A.make_cpml(d, p, 1e-4, FC, dtype)
        s = A.Solver(d, A.dtype_code(dtype, prec), cp)
        s.build_mask("sphere", [24 * d.dx])
        shape = A.tune_solver(s)
        s.set_transceivers(A.transceivers_to_grid(d, txs, dtype))
        if w1 >= w0:
            s.run(w0, w1)
        # events around the sweep of sampled steps of the timed region
        s.set_option(L.ADF_OPT_TIME_SWEEP, sweep_sample_period(steps))
        t0 = time.perf_counter()
        s.run(t_first, t_last)
        wall = time.perf_counter() - t0
        sweep_ms, launches = s.sweep_stats()
        peak = float(np.abs(s.get_trace(L.ADF_REC_P)[1:]).max())
        # host-buffer boundary: what a run_fdtd! call adds around the steps
        # (field, CPML memory and transceivers in; field, CPML memory and all
        # ten trace buffers out)
        g = A.transceivers_to_grid(d, txs, dtype)
        f = np.zeros((4, d.nz, d.ny, d.nx), dtype)
        t1 = time.perf_counter()
        s.set_field(f)
        s.set_cpml_memory(cp)
        s.set_transceivers(g)
        t2 = time.perf_counter()
        s.get_field(f)
        s.get_cpml_memory(cp)
        s.get_transceivers(g)
        t3 = time.perf_counter()
        out[name + "_pcie"] = {"upload_s": t2 - t1, "download_s": t3 - t2}
        out[name] = {"wall_s": wall, "steps": steps, "ms_per_step": wall / steps * 1e3,
                     "dev_ms_per_step": s.last_step_ms(), "sweep_ms": sweep_ms, "sweep_launches": launches,
                     "glups": d.nx * d.ny * d.nz * steps / wall / 1e9, "tiling": shape,
                     "p_rec_peak": peak}
        del s
    return d, p, out


PAPER_GRIDS = (301, 501, 545)


def paper_grid_leg(n, steps=50, warmup=10):
    """The reference's production grid sizes (configs/paper.toml:21,71,165:
    [grid] n = 501 hologram box, [greens.grid] n = 301, [scattering.grid]
    n = 545), rows wider than one wave and not a multiple of its cells per
    lane: n^3 cells over a 2.2 m box with rigid walls (paper.toml's pml
    n = 0), a rigid sphere of radius 0.1 m, a Ricker point source at the
    centre (fc 10 kHz) and p+vn recorded on the 400-point inner (0.2 m) and
    outer (0.3 m) Fibonacci surfaces.  FP32, fused sweep, autotuned tiling."""
    import adfdtd as A
    from adfdtd import _lib as L

    nt = warmup + steps + 1
    d0 = A.make_domain(1.0, 1.1, 1.1, 1.1, n, n, n, C_AIR, RHO_AIR, 0.5)
    d = A.make_domain((nt - 0.5) * d0.dt, 1.1, 1.1, 1.1, n, n, n, C_AIR, RHO_AIR, 0.5)
    tf = A.ricker_wavelet(d, 10e3, True)
    z = np.zeros(d.nt)
    txs = [A.make_transceiver(d, [0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0], tf)]
    for r in (0.2, 0.3):
        pts, nrm = A.fibonacci_sphere(400, r)
        txs += [A.make_transceiver(d, pts[i], np.r_[0.0, nrm[i]], z) for i in range(400)]
    s = A.Solver(d, A.dtype_code(np.float32, "fast"), None)
    s.build_mask("sphere", [0.1])
    shape = A.tune_solver(s)
    s.set_transceivers(A.transceivers_to_grid(d, txs, np.float32))
    s.run(1, warmup)
    s.set_option(L.ADF_OPT_TIME_SWEEP, sweep_sample_period(steps))
    t0 = time.perf_counter()
    s.run(warmup + 1, warmup + steps)
    wall = time.perf_counter() - t0
    sweep_ms, launches = s.sweep_stats()
    b = algorithmic_bytes_per_step(n, 0, 4)
    del s
    return {"grid": [n, n, n], "steps": steps, "ms_per_step": round(wall / steps * 1e3, 4),
            "glups": round(n ** 3 * steps / wall / 1e9, 2), "sweep_ms": round(sweep_ms, 4),
            "sweep_frac": round(b / (sweep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": b, "launches_timed": launches, "tiling": shape}


def host_cpu():
    """The host's CPU model, the logical CPUs this process may run on, the
    physical cores among them (distinct (package, core) pairs), and the
    thread share the environment grants (OMP_NUM_THREADS, when set)."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            cores.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": len(cores), "thread_share": share or None}


def cpu_threads(info):
    """Every physical core the process may use, capped by the environment's
    thread share when it sets one."""
    n = info["physical_cores"]
    return min(n, info["thread_share"]) if info["thread_share"] else n


def cpu_sample_steps(nt, one_step_s, max_seconds):
    """Steps of the CPU sample after two warm-up steps: about max_seconds of
    work, at least 2, and within the domain's nt (steps 3..2+k)."""
    return int(max(2, min(nt - 2, max_seconds / max(one_step_s, 1e-3))))


def _cpu_steps(d, txs, p, sphere, mode, max_seconds):
    """The oracle (OpenMP over k on every physical core) stepping the given
    solve: (steps timed, seconds, threads, host info), on a bounded sample of
    consecutive steps after two warm-up steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker / CPU baseline only
    import adfdtd as A

    info = host_cpu()
    threads = cpu_threads(info)
    O.set_threads(threads)
    T = np.float64 if mode == "f64" else np.float32
    rd = O.domain(d.tmax, d.xmax, d.ymax, d.zmax, d.nx, d.ny, d.nz, d.c0, d.r0, d.cf)
    g = A.transceivers_to_grid(d, txs, T)
    tx = {"idx": g.idx, "w": g.w, "src": g.src, "rec": g.rec, "dir": g.direction, "vn_src": g.vn_src,
          "vn_rec": g.vn_rec}
    mask = A.build_update_mask(d, "sphere", [sphere])
    s = O.RefSolver(rd, npml=p, rcoef=1e-4, fc=FC, mask=mask, txs=tx, mode=mode)
    s.run(1, 1)
    t0 = time.perf_counter()
    s.run(2, 2)
    one = time.perf_counter() - t0
    k = cpu_sample_steps(d.nt, one, max_seconds)
    t0 = time.perf_counter()
    s.run(3, 2 + k)
    return k, time.perf_counter() - t0, threads, info


def cpu_leg(d, txs, p, mode="r32", max_seconds=15.0):
    """The oracle in the reference's own Float32 arithmetic (or Float64),
    OpenMP over k, on a bounded sample: consecutive steps of the same 256^3
    workload."""
    k, el, threads, info = _cpu_steps(d, txs, p, 24 * d.dx, mode, max_seconds)
    arith = "Float64" if mode == "f64" else "the reference's Float32"
    return {"value": d.nx * d.ny * d.nz * k / el / 1e9, "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"{k} consecutive steps (3..{2 + k}) of the same 256^3 workload, oracle in {arith} "
                      f"arithmetic, OpenMP over k on {threads} threads, {el:.1f} s",
            "host": info}


def cpu_batch_leg(nt, max_seconds=10.0):
    """The batched workload on the host: its solves one after another, each
    on every physical core (oracle, the reference's Float32 arithmetic),
    from a bounded sample of one 128^3 solve's steps."""
    sp = batch_specs(64, [128], nt)[0]
    d = sp["d"]
    z = np.zeros(d.nt)
    import adfdtd as A

    txs = [A.make_transceiver(d, sp["src"], [1.0, 0.0, 0.0, 0.0], sp["tf"])]
    txs += [A.make_transceiver(d, sp["rec"][i], np.r_[0.0, sp["rec_n"][i]], z) for i in range(len(sp["rec"]))]
    k, el, threads, info = _cpu_steps(d, txs, sp["npml"], sp["sphere"], "r32", max_seconds)
    per_solve = el / k * nt
    return {"value": 1.0 / per_solve, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{k} consecutive steps of one 128^3 solve of the batched workload ({nt} steps per solve), "
                      f"oracle in the reference's Float32 arithmetic on {threads} threads, {el:.1f} s; solves "
                      f"run one after another"}


def calibrate_copy(kinds=None, sizes=(1 << 28, 1 << 30), reps=20):
    """HBM ceilings with the library's own streaming kernels
    (adf_bandwidth_probe): copy with the sweep's 128-bit buffer loads and
    stores, copy with global loads and stores, read only, write only, copy
    with non-temporal stores; buffers of 256 MiB (the size of the GPU's
    last-level cache) and 1 GiB.  The z-march probes (the sweep's access
    pattern without its arithmetic, ordinary and non-temporal stores) also
    run at 64 MiB per buffer, where both buffers fit the last-level cache."""
    import ctypes as C

    from adfdtd import _lib as L
    from adfdtd.solver import context

    names = {L.ADF_BW_COPY_BUF: "copy_buffer", L.ADF_BW_COPY_GLOBAL: "copy_global", L.ADF_BW_READ: "read",
             L.ADF_BW_WRITE: "write", L.ADF_BW_COPY_NT: "copy_nontemporal", L.ADF_BW_ZMARCH: "zmarch",
             L.ADF_BW_ZMARCH_NT: "zmarch_nontemporal"}
    zmarch = (L.ADF_BW_ZMARCH, L.ADF_BW_ZMARCH_NT)
    lib = L.load()
    ctx = context(0)
    out = {}
    for kind in kinds if kinds is not None else sorted(names):
        for nb in ((1 << 26,) + tuple(sizes) if kind in zmarch and kinds is None else sizes):
            ms = C.c_double()
            L.check(lib.adf_bandwidth_probe(ctx, kind, nb, reps, C.byref(ms)), "adf_bandwidth_probe")
            moved = nb * (2 if kind in (L.ADF_BW_COPY_BUF, L.ADF_BW_COPY_GLOBAL, L.ADF_BW_COPY_NT) + zmarch else 1)
            out[f"{names[kind]}_{nb >> 20}MiB"] = {"ms": round(ms.value, 4), "bytes_moved": moved,
                                                   "GBps": round(moved / (ms.value * 1e-3) / 1e9, 1),
                                                   "frac_of_peak": round(moved / (ms.value * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    print(json.dumps(out))


def batch_bytes(specs):
    return sum(algorithmic_bytes_per_step(sp["d"].nx, sp["npml"], 4) for sp in specs)


def batched_leg(n_solves, sizes, nt, tune=True):
    """One batch of impulsive-GF solves run to completion (nt steps each) in
    one adf_batch_run; solves/s from the device time of the whole run."""
    import adfdtd as A

    specs = batch_specs(n_solves, sizes, nt)
    sol = [make_batch_solver(sp) for sp in specs]
    b = A.Batch(persistent=False, graph_steps=16)
    shape = A.tune_batch(sol, batch=b) if tune else None
    for x in sol:
        b.add(x)
    t0 = time.perf_counter()
    b.run(nt)
    wall = time.perf_counter() - t0
    ms = b.last_step_ms()
    run_s = ms * nt / 1e3
    byts = batch_bytes(specs)
    cells = sum(sp["d"].nx * sp["d"].ny * sp["d"].nz for sp in specs)
    peak = float(np.abs(sol[0].get_trace(0)).max())
    del b, sol
    traffic = read_traffic("sweep_batch_128_f32.md") if sizes == [128] and n_solves == 64 else None
    return {"solves": len(specs), "grids": sorted({sp["d"].nx for sp in specs}), "steps": nt,
            "run_s": round(run_s, 4), "wall_s": round(wall, 4), "solves_per_s": round(len(specs) / run_s, 3),
            "ms_per_batched_step": round(ms, 4), "glups": round(cells / (ms * 1e-3) / 1e9, 2),
            "algorithmic_GBps": round(byts / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "tiling": shape,
            "sweep_traffic_bytes": traffic, "algorithmic_bytes_per_step": byts,
            "p_rec_peak_solve0": peak}


def plain_solves_leg(n_solves, nt):
    """The plain single-solve GPU path on the same 128^3 solves: one solver,
    per solve reset + transceivers + nt steps + traces back, one after
    another (impulsive_gfs' loop, src/greens.jl:143-171)."""
    import adfdtd as A
    from adfdtd import _lib as L

    specs = batch_specs(n_solves, [128], nt)
    d = specs[0]["d"]
    s = A.Solver(d, A.dtype_code(np.float32, "fast"), A.make_cpml(d, specs[0]["npml"], 1e-4, FC, np.float32))
    s.build_mask("sphere", [specs[0]["sphere"]])
    A.tune_solver(s)
    z = np.zeros(d.nt)
    grids = []
    for sp in specs:
        txs = [A.make_transceiver(d, sp["src"], [1.0, 0.0, 0.0, 0.0], sp["tf"])]
        txs += [A.make_transceiver(d, sp["rec"][i], np.r_[0.0, sp["rec_n"][i]], z) for i in range(len(sp["rec"]))]
        grids.append(A.transceivers_to_grid(d, txs, np.float32))
    s.reset()
    s.set_transceivers(grids[0])
    s.run(1, 20)  # warm-up
    t0 = time.perf_counter()
    dev = 0.0
    for g in grids:
        s.reset()
        s.set_transceivers(g)
        s.run(1, nt)
        dev += s.last_step_ms() * nt / 1e3
        s.get_trace(L.ADF_REC_P)
        s.get_trace(L.ADF_VN_REC)
    wall = time.perf_counter() - t0
    return {"solves": n_solves, "steps": nt, "wall_s": round(wall, 4), "device_s": round(dev, 4),
            "solves_per_s": round(n_solves / wall, 3), "solves_per_s_device": round(n_solves / dev, 3)}


def gfs_leg(n_src=32, n=128, nt=1200, n_rec=100, nt_save=600):
    """impulsive_gfs end to end (src/greens.jl:48-234): 2 n_src solves (a
    pressure and a velocity monopole per outer point) on the 128^3 batched
    workload's grid, scatterer and surfaces, advanced as one batch, then the
    device post-processing of every run (self-record dropped, denoise, cubic
    resample to nt_save samples).  Wall time of the whole call, set-up and
    host transfers included."""
    import adfdtd as A

    dx = C_AIR / (2.5 * FC) / 10
    xmax = (n - 1) * dx / 2
    d0 = A.make_domain(1.0, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
    d = A.make_domain((nt - 0.5) * d0.dt, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
    cfg = {"greens": {"f_3db_hz": FC, "nt_save": nt_save}, "pml": {"n": 10, "rcoef": 1e-4, "fc": FC},
           "surfaces": {"nPoints_outer": n_src, "radius_outer": 40 * d.dx, "nPoints_inner": n_rec,
                        "radius_inner": 20 * d.dx, "sphere_radius": 16 * d.dx}}
    t0 = time.perf_counter()
    g = A.impulsive_gfs(d, cfg, scatterer="sphere", sources_per_batch=n_src)
    wall = time.perf_counter() - t0
    return {"solves": 2 * n_src, "grid": n, "steps": nt, "receivers": n_rec, "nt_save": nt_save,
            "wall_s": round(wall, 3), "solves_per_s": round(2 * n_src / wall, 3),
            "p_p_peak": float(np.abs(g["p_p"]).max())}


def reverb_leg(n_ill=32, n=128, nt=1200, n_rec=100):
    """build_reverb_data end to end (src/reverb.jl:72-417): n_ill pressure
    monopoles on the R2 illumination sphere (r = 56 cells) in a closed
    128^3 box (npml = 0) around the sphere scatterer (r = 16 cells), p and
    vn recorded on outer (r = 40 cells) and inner (r = 20 cells) Fibonacci
    surfaces of n_rec points each, resampled to fs_out = 2 * 2.5 * fc.
    Fresh sources advance as one batch.  Wall time of the whole call, set-up
    and host transfers included."""
    import adfdtd as A

    dx = C_AIR / (2.5 * FC) / 10
    xmax = (n - 1) * dx / 2
    d0 = A.make_domain(1.0, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
    d = A.make_domain((nt - 0.5) * d0.dt, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
    cfg = {"reverb": {"radius_ill": 56 * d.dx, "fc_source": FC, "bandwidth_factor": 2.5, "nPoints_ill": n_ill},
           "pml": {"n": 0, "rcoef": 1e-4, "fc": FC},
           "surfaces": {"nPoints_outer": n_rec, "radius_outer": 40 * d.dx, "nPoints_inner": n_rec,
                        "radius_inner": 20 * d.dx, "sphere_radius": 16 * d.dx}}
    t0 = time.perf_counter()
    g = A.build_reverb_data(d, cfg, scatterer="sphere", save_field=False, sources_per_batch=n_ill)
    wall = time.perf_counter() - t0
    byts = n_ill * algorithmic_bytes_per_step(n, 0, 4)
    return {"sources": n_ill, "grid": n, "steps": nt, "receivers": 2 * n_rec, "nt_ill": int(g["t"].shape[0]),
            "wall_s": round(wall, 3), "sources_per_s": round(n_ill / wall, 3),
            "glups_wall": round(n_ill * n ** 3 * nt / wall / 1e9, 2),
            "hbm_frac_wall": round(byts * nt / wall / 1e9 / HBM_PEAK_GBS, 4),
            "inner_p_peak": float(np.abs(g["inner_p"]).max())}


def hologram_leg(n=192, n_kh=2000, n_ff=500, steps=1500):
    """Hologram synthesis (build_hologram's FDTD, src/kernels/forward.jl
    with C_for_N injection, illumination.jl): 192^3 FP32, 12-cell CPML, cube
    of edge 30 cells; every point of a 2000-point KH sphere (r = 40 cells)
    injects q and f through C_for_N with seeded Ricker-filtered Gaussian
    noise drives, and 500 far-field receivers (r = 80 cells) record.  Whole
    steps by wall clock, the sweep by HIP events every 16th step."""
    import adfdtd as A
    from adfdtd import _lib as L

    p = 12
    dx = C_AIR / (2.5 * FC) / 10
    xm = (n - 1) * dx / 2
    d0 = A.make_domain(1.0, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)
    d = A.make_domain((steps + 1 - 0.5) * d0.dt, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)
    kh, khn = A.fibonacci_sphere(n_kh, 40 * d.dx)
    ff, ffn = A.fibonacci_sphere(n_ff, 80 * d.dx)
    rng = np.random.default_rng(2)
    rick = A.ricker_wavelet(d, FC, False)
    drive = lambda: np.convolve(rng.standard_normal(d.nt), rick)[: d.nt]
    cf = A.c_for_n(n_kh, 40 * d.dx, d)
    txs = [A.make_transceiver(d, kh[i], [cf["C_q"], 0.0, 0.0, 0.0], drive()) for i in range(n_kh)]
    txs += [A.make_transceiver(d, kh[i], np.r_[0.0, cf["C_f"] * khn[i]], drive()) for i in range(n_kh)]
    z = np.zeros(d.nt)
    txs += [A.make_transceiver(d, ff[i], np.r_[0.0, ffn[i]], z) for i in range(n_ff)]
    s = A.Solver(d, A.dtype_code(np.float32, "fast"), A.make_cpml(d, p, 1e-4, FC, np.float32))
    s.build_mask("cube", [30 * d.dx])
    shape = A.tune_solver(s)
    s.set_transceivers(A.transceivers_to_grid(d, txs, np.float32))
    s.set_option(L.ADF_OPT_TIME_SWEEP, SWEEP_SAMPLE)
    t0 = time.perf_counter()
    s.run(1, steps)
    wall = time.perf_counter() - t0
    sweep_ms, launches = s.sweep_stats()
    peak = float(np.abs(s.get_trace(L.ADF_REC_P)[2 * n_kh:]).max())
    b = algorithmic_bytes_per_step(n, p, 4)
    return {"grid": n, "steps": steps, "injectors": 2 * n_kh, "receivers": n_ff, "wall_s": round(wall, 4),
            "ms_per_step": round(wall / steps * 1e3, 5), "glups": round(n ** 3 * steps / wall / 1e9, 2),
            "sweep_ms": round(sweep_ms, 5), "hbm_frac_sweep": round(b / (sweep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "hbm_frac_step": round(b / (wall / steps) / 1e9 / HBM_PEAK_GBS, 4), "tiling": shape,
            "far_field_p_peak": peak}


def gf_map_leg(n=192, nio=400, ng=827, steps=100):
    """In-loop KH gather (extrapolation.jl:101-249) at paper scale: a 192^3
    FP32 solve (12-cell CPML, cube r=30 cells) whose 400 outer receivers
    (r = 80 cells) drive 400 inner sources (r = 40 cells) through a
    PVFromPin map of nt_gf = 827 lags (random tensors).  Device time per
    step with and without the map, after nt_gf steps so every step uses
    the whole map; the gather reads 2 * 400 * 400 * 827 float32 per step's
    worth of MACs."""
    import adfdtd as A

    dx = C_AIR / (2.5 * FC) / 10
    xm = (n - 1) * dx / 2
    nt = ng + steps + 60
    d0 = A.make_domain(1.0, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)
    d = A.make_domain((nt - 0.5) * d0.dt, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)
    ro, rn = A.fibonacci_sphere(nio, 80 * d.dx)
    io, inn = A.fibonacci_sphere(nio, 40 * d.dx)
    z = np.zeros(d.nt)
    txs = [A.make_transceiver(d, [60 * d.dx, 0, 0], [1, 0, 0, 0], A.ricker_wavelet(d, FC, True))]
    txs += [A.make_transceiver(d, ro[i], np.r_[0.0, rn[i]], z) for i in range(nio)]
    txs += [A.make_transceiver(d, io[i], [1.0, inn[i][0], inn[i][1], inn[i][2]], z) for i in range(nio)]
    rng = np.random.default_rng(1)
    pp = (rng.standard_normal((nio, nio, ng)) * 1e-3).astype(np.float32)
    vp = (rng.standard_normal((nio, nio, ng)) * 1e-3).astype(np.float32)
    out = {}
    for use in (0, 1):
        s = A.Solver(d, A.dtype_code(np.float32, "fast"), A.make_cpml(d, 12, 1e-4, FC, np.float32))
        s.build_mask("cube", [30 * d.dx])
        s.set_transceivers(A.transceivers_to_grid(d, txs, np.float32))
        if use:
            s.set_gf_map(A.PVFromPin(pp, vp, list(range(2, 2 + nio)), list(range(2 + nio, 2 + 2 * nio)), True))
        s.run(1, ng + 40)
        s.run(ng + 41, ng + 40 + steps)
        out[use] = s.last_step_ms()
        del s
    g = out[1] - out[0]
    macs = 2 * nio * nio * ng
    return {"grid": n, "inputs": nio, "outputs": nio, "nt_gf": ng, "steps": steps,
            "step_ms_without_map": round(out[0], 4), "step_ms_with_map": round(out[1], 4),
            "gather_ms_per_step": round(g, 4), "gather_gmacs": round(macs / (g * 1e-3) / 1e9, 1),
            "map_bytes": macs * 4}


def kh_leg(which="p_v"):
    """Post-hoc KH of the 256^3 workload's surfaces: the 2000-point KH sphere
    (r = 40 cells) extrapolated to the 500 far-field points (r = 80 cells)
    over 2000 steps, analytic kernel evaluated on the device (no dense K:
    it would be 8 GB).  The work is the arrival-window MACs; VALU-bound."""
    import adfdtd as A

    d, _, _ = build_workload(nt=2000)
    op, on = A.fibonacci_sphere(2000, 40 * d.dx)
    ip, inn = A.fibonacci_sphere(500, 80 * d.dx)
    sigma = math.sqrt(math.log(2)) / (2 * math.pi * FC)
    nt = 2000
    tg = np.arange(nt) * d.dt
    inp = np.random.default_rng(7).standard_normal((2000, nt))
    out = np.zeros((500, nt))
    A.kh_analytic_accumulate(out, which, op, on, ip, inn, tg, sigma, C_AIR, RHO_AIR, inp[:, :nt], dtype=np.float32)
    t0 = time.perf_counter()
    A.kh_analytic_accumulate(out, which, op, on, ip, inn, tg, sigma, C_AIR, RHO_AIR, inp, dtype=np.float32)
    wall = time.perf_counter() - t0
    ms = A.kh.kh_last_ms()
    # window MACs: per (r, s) pair, sum over t of the window samples <= t
    r = np.linalg.norm(ip[None, :, :] - op[:, None, :], axis=2) / C_AIR
    lo = np.clip(np.floor((r - 10 * sigma) / d.dt), 0, nt)
    hi = np.clip(np.floor((r + 10 * sigma) / d.dt) + 1, 0, nt)
    macs = float(np.sum((hi - lo) * (nt - (lo + hi) / 2)))
    dense = 500 * 2000 * nt * (nt + 1) / 2
    tflops = 2 * macs / (ms * 1e-3) / 1e12
    return {"kernel": which, "n_outer": 2000, "n_inner": 500, "nt": nt, "device_ms": round(ms, 3),
            "wall_s_incl_transfers": round(wall, 4), "window_macs": macs, "dense_equivalent_macs": dense,
            "tflops_window": round(tflops, 2), "frac_fp32_valu_peak": round(tflops / 157.3, 4),
            "frac_fp32_fma_unpacked_peak": round(tflops / 78.6, 4),
            "dense_equivalent_tmacs_per_s": round(dense / (ms * 1e-3) / 1e12, 1),
            "note": "157.3 TFLOP/s counts packed FP32 FMAs (two per lane per cycle); the library is compiled without "
                    "them, which caps FP32 FMAs at 78.6 TFLOP/s"}


def served_specs(n_req, seed=5, sizes=(96, 128, 192, 256), steps=(800, 3000), rate=20.0):
    """Poisson arrivals (rate per s, seeded) of impulsive-GF requests: grid
    uniform over `sizes`, steps uniform in `steps`, scatterer uniform over
    sphere/cube/ellipsoid with radius 10-30 % of the grid half-extent, a
    monopole source on a Fibonacci sphere of radius 40 % of the half-extent,
    100 receivers (p, vn) at 20 %."""
    import adfdtd as A

    rng = np.random.default_rng(seed)
    dx = C_AIR / (2.5 * FC) / 10
    t = 0.0
    out = []
    for i in range(n_req):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        n = int(rng.choice(sizes))
        nt = int(rng.integers(steps[0], steps[1] + 1))
        xmax = (n - 1) * dx / 2
        d0 = A.make_domain(1.0, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
        d = A.make_domain((nt - 0.5) * d0.dt, xmax, xmax, xmax, n, n, n, C_AIR, RHO_AIR, 0.5)
        shape = str(rng.choice(["sphere", "cube", "ellipsoid"]))
        r = rng.uniform(0.1, 0.3, 3) * xmax
        params = {"sphere": [r[0]], "cube": [2 * r[0]], "ellipsoid": list(r)}[shape]
        srcs, _ = A.fibonacci_sphere(16, 0.4 * xmax)
        rec, rn = A.fibonacci_sphere(100, 0.2 * xmax)
        tf, _ = A.impulsive_wavelet(d, FC)
        out.append({"t": t, "d": d, "npml": 10, "shape": shape, "params": params, "src": srcs[i % 16],
                    "tf": tf, "rec": rec, "rec_n": rn})
    return out


def served_leg(reqs, budget=200 << 30):
    """Requests submitted at their arrival times (key "t", seconds) to a
    running adf_server; latency is submission to retirement."""
    import adfdtd as A

    n_req = len(reqs)
    prepared = []
    for q in reqs:
        d = q["d"]
        z = np.zeros(d.nt)
        txs = [A.make_transceiver(d, q["src"], [1.0, 0.0, 0.0, 0.0], q["tf"])]
        txs += [A.make_transceiver(d, q["rec"][i], np.r_[0.0, q["rec_n"][i]], z) for i in range(len(q["rec"]))]
        prepared.append((q, A.make_cpml(d, q["npml"], 1e-4, FC, np.float32),
                         A.transceivers_to_grid(d, txs, np.float32)))
    code = A.dtype_code(np.float32, "fast")
    lups = sum(q["d"].nx ** 3 * q["d"].nt for q in reqs)
    with A.Server(budget_bytes=budget, chunk=64) as sv:
        t0 = time.perf_counter()
        ids = []
        for q, cp, g in prepared:
            dt = q["t"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(sv.submit(q["d"], code, cp, g, mask=(q["shape"], q["params"]), want=0x11))
        for rid in ids:
            sv.wait(rid, 600.0)
        wall = time.perf_counter() - t0
        st = sv.stats()
    return {"requests": n_req, "grids": sorted({q["d"].nx for q in reqs}),
            "steps": [min(q["d"].nt for q in reqs), max(q["d"].nt for q in reqs)],
            "budget_GB": budget / 2 ** 30, "wall_s": round(wall, 3), "solves_per_s": round(n_req / wall, 3),
            "p50_ms": round(st["p50_ms"], 1), "p99_ms": round(st["p99_ms"], 1), "mean_ms": round(st["mean_ms"], 1),
            "glups_busy": round(st["glups"], 2), "glups_wall": round(lups / wall / 1e9, 2),
            "peak_GB": round(st["peak_bytes"] / 2 ** 30, 3), "failed": st["failed"]}


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--fp64-steps", type=int, default=300)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=0)
    ap.add_argument("--profile-batch", type=int, default=0)
    ap.add_argument("--profile-gf", action="store_true")
    ap.add_argument("--profile-hologram", action="store_true")
    ap.add_argument("--profile-kh", action="store_true")
    ap.add_argument("--profile-grid", type=int, default=0)
    ap.add_argument("--skip-paper", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--probe-kind", type=int, default=-1, help="with --calibrate-copy: one probe kind only")
    ap.add_argument("--probe-bytes", type=int, default=1 << 30)
    ap.add_argument("--probe-reps", type=int, default=20)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-served", action="store_true")
    ap.add_argument("--batch-solves", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=1200)
    ap.add_argument("--served-requests", type=int, default=48)
    return ap.parse_args(argv)


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {args.gpus} requested); the library runs each "
              "solve and batch on one GPU, and independent solves would shard across GPUs without "
              "communication", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        if args.probe_kind >= 0:
            calibrate_copy([args.probe_kind], (args.probe_bytes,), args.probe_reps)
        else:
            calibrate_copy()
        return
    if args.profile_gf:
        print(json.dumps(gf_map_leg()))
        return
    if args.profile_hologram:
        print(json.dumps(hologram_leg()))
        return
    if args.profile_kh:
        print(json.dumps(kh_leg()))
        return
    if args.profile_grid:
        print(json.dumps(paper_grid_leg(args.profile_grid, steps=args.profile_steps or 50)))
        return
    if args.profile_batch:
        print(json.dumps(batched_leg(args.batch_solves, [128], args.profile_batch)))
        return
    if args.profile_steps:
        args.steps, args.warmup, args.fp64_steps = args.profile_steps, 2, 0
        args.skip_cpu = args.skip_batch = args.skip_served = args.skip_paper = True

    d, p, g = gpu_leg(args)
    log("256^3 plain-path legs done")
    n = d.nx
    f32 = g["f32"]
    b32 = algorithmic_bytes_per_step(n, p, 4)
    achieved = b32 / (f32["sweep_ms"] * 1e-3) / 1e9
    traffic = read_traffic("sweep_256_f32.md")
    res = {
        "metric": "GLUPS (256^3 FDTD, one step = one forward_onestep!)",
        "value": round(f32["glups"], 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "256^3 impulsive GF with rigid sphere scatterer r=24 cells, 16-cell CPML, Ricker "
                               "source at 60 cells, p+vn on 2000-point KH sphere r=40 cells + 500 receivers r=80 cells",
                   "grid": [n, n, n], "npml": p, "transceivers": 2501, "storage": "float32"},
        "hbm_frac_step": round(b32 / (f32["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_sweep (field-update sweep of one step: one launch over interior and side tiles, carrying the step's transceiver work)",
                     "algorithmic_bytes_per_launch": b32,
                     "avg_launch_ms": round(f32["sweep_ms"], 5), "launches": f32["sweep_launches"],
                     "timed_every": sweep_sample_period(args.steps),
                     "tiling": f32["tiling"]},
    }
    io = g["f32_pcie"]
    res["pcie_inclusive"] = {
        "glups": round(n ** 3 * args.steps / (f32["wall_s"] + io["upload_s"] + io["download_s"]) / 1e9, 3),
        "upload_s": round(io["upload_s"], 4), "download_s": round(io["download_s"], 4),
        "note": "value plus one host-to-device upload of field, CPML memory and transceivers and one download "
                "of field, CPML memory and the ten trace buffers (a run_fdtd! call through the host boundary)"}
    if "f64" in g:
        res["f64"] = {"glups": round(g["f64"]["glups"], 3), "ms_per_step": round(g["f64"]["ms_per_step"], 5),
                      "steps": g["f64"]["steps"], "sweep_ms": round(g["f64"]["sweep_ms"], 5),
                      "hbm_frac_sweep": round(algorithmic_bytes_per_step(n, p, 8) / (g["f64"]["sweep_ms"] * 1e-3)
                                              / 1e9 / HBM_PEAK_GBS, 4)}
    if not args.skip_batch:
        nb, ns = args.batch_solves, args.batch_steps
        bt = batched_leg(nb, [128], ns)
        log("batched 128^3 leg done")
        un = batched_leg(48, [96, 128, 160], ns)
        log("batched uneven leg done")
        pl = plain_solves_leg(4, ns)
        log("plain-path solves leg done")
        res["batched"] = {"workload": f"{nb} impulsive-GF solves, 128^3 FP32, 10-cell CPML, sphere r=16 cells, "
                                      f"sources on a 40-cell Fibonacci sphere, 100 receivers at 20 cells, {ns} steps",
                          **bt, "plain_path": pl,
                          "speedup_vs_plain": round(bt["solves_per_s"] / pl["solves_per_s"], 2)}
        res["batched_uneven"] = {"workload": f"48 solves, 16 each at 96^3/128^3/160^3 (radii scaled by n/128), "
                                             f"{ns} steps", **un}
    if not args.skip_batch:
        res["impulsive_gfs"] = {"workload": "impulsive_gfs end to end: 32 outer points x (p, v) sources = 64 solves, "
                                            "128^3 FP32, sphere r=16 cells, 100 inner receivers, 1200 steps, "
                                            "resampled to 600 samples", **gfs_leg()}
        log("impulsive_gfs leg done")
        res["reverb"] = {"workload": "build_reverb_data end to end: 32 R2 illumination sources, closed 128^3 box "
                                     "(npml 0), sphere r=16 cells, 100 outer + 100 inner receivers, 1200 steps, "
                                     "resampled to fs_out = 5 fc", **reverb_leg()}
        log("reverb leg done")
        res["hologram"] = {"workload": "hologram synthesis: 192^3 FP32, 12-cell CPML, cube edge 30 cells, 2000 KH points "
                                       "injecting q and f via C_for_N (seeded Ricker-filtered noise drives), 500 "
                                       "far-field receivers at 80 cells, 1500 steps", **hologram_leg()}
        log("hologram leg done")
        res["kh_in_loop"] = {"workload": "in-loop KH gather, 192^3 FP32 solve, 400 outer receivers -> 400 inner "
                                         "sources through a PVFromPin map of 827 lags (paper scale)", **gf_map_leg()}
        log("in-loop KH gather leg done")
        res["kh_post_hoc"] = {"workload": "analytic KH p_v kernel, 2000-point KH sphere -> 500 far-field points, "
                                          "2000 steps, FP32 kernel samples and inputs, double sums", **kh_leg()}
        log("post-hoc KH leg done")
    if not args.skip_paper:
        res["paper_grids"] = {"workload": "the reference's production grid sizes (paper.toml 301/501/545^3): 2.2 m "
                                          "box, rigid walls, sphere r=0.1 m, centre Ricker source (10 kHz), p+vn "
                                          "on 400+400 surface points, FP32, 50 steps",
                              "runs": [paper_grid_leg(n) for n in PAPER_GRIDS]}
        log("production-grid legs done")
    if not args.skip_served:
        res["served"] = {"workload": "seeded Poisson arrivals (20 req/s) of FP32 impulsive-GF requests: grid "
                                     "96/128/192/256^3, 800-3000 steps, sphere/cube/ellipsoid r=10-30 % of the "
                                     "half-extent, 100 receivers, 200 GB budget",
                         **served_leg(served_specs(args.served_requests))}
        log("served Poisson leg done")
        b128 = [dict(sp, t=0.0, shape="sphere", params=[sp["sphere"]])
                for sp in batch_specs(args.batch_solves, [128], args.batch_steps)]
        res["served_128"] = {"workload": f"the {args.batch_solves} solves of the 128^3 batched workload as requests, "
                                         "all submitted at t=0", **served_leg(b128)}
    if not args.skip_cpu:
        d2, txs, p2 = build_workload(nt=4000)
        res["cpu_baseline"] = cpu_leg(d2, txs, p2)
        res["cpu_baseline_f64"] = cpu_leg(d2, txs, p2, mode="f64", max_seconds=10.0)
        if "batched" in res:
            res["batched"]["cpu_baseline"] = cpu_batch_leg(args.batch_steps)
        log("CPU baseline legs done")
    print(json.dumps(res))


if __name__ == "__main__":
    main()
