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
  --calibrate-copy    time a plain 1 GiB device copy (PMC scale factor)
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
C_AIR, RHO_AIR, FC = 343.0, 1.21, 2000.0


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
    n_psi = 3 * ((2 * p - 1) * (n - 1) ** 2 + (2 * p + 1) * n ** 2)
    return es * (8 * n ** 3 + 2 * n_psi) + (n ** 3 / 8 if masked else 0)


def read_traffic(name):
    """HBM bytes per launch of the sweep kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    fn = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(fn):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(fn).read(), re.M)
    return float(m.group(1)) if m else None


def gpu_leg(args):
    import adfdtd as A
    from adfdtd import _lib as L

    nt = args.warmup + args.steps
    d, txs, p = build_workload(nt=nt + 1)
    out = {}
    for name, dtype, prec, steps in (("f32", np.float32, "fast", args.steps), ("f64", np.float64, "fast", args.fp64_steps)):
        if steps <= 0:
            continue
        cp = A.make_cpml(d, p, 1e-4, FC, dtype)
        s = A.Solver(d, A.dtype_code(dtype, prec), cp)
        s.build_mask("sphere", [24 * d.dx])
        s.set_transceivers(A.transceivers_to_grid(d, txs, dtype))
        s.run(1, args.warmup)
        s.set_option(L.ADF_OPT_TIME_SWEEP, 1)
        t0 = time.perf_counter()
        s.run(args.warmup + 1, args.warmup + steps)
        wall = time.perf_counter() - t0
        sweep_ms, launches = s.sweep_stats()
        out[name] = {"wall_s": wall, "steps": steps, "ms_per_step": wall / steps * 1e3,
                     "dev_ms_per_step": s.last_step_ms(), "sweep_ms": sweep_ms, "sweep_launches": launches,
                     "glups": d.nx * d.ny * d.nz * steps / wall / 1e9,
                     "p_rec_peak": float(np.abs(s.get_trace(L.ADF_REC_P)[1:]).max())}
        del s
    return d, p, out


def cpu_leg(d, txs, p, max_seconds=15.0):
    """The oracle in the reference's own Float32 arithmetic, OpenMP over k,
    on a bounded sample: consecutive steps of the same 256^3 workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker / CPU baseline only
    import adfdtd as A

    threads = min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    rd = O.domain(d.tmax, d.xmax, d.ymax, d.zmax, d.nx, d.ny, d.nz, d.c0, d.r0, d.cf)
    g = A.transceivers_to_grid(d, txs, np.float32)
    tx = {"idx": g.idx, "w": g.w, "src": g.src, "rec": g.rec, "dir": g.direction, "vn_src": g.vn_src,
          "vn_rec": g.vn_rec}
    mask = A.build_update_mask(d, "sphere", [24 * d.dx])
    s = O.RefSolver(rd, npml=p, rcoef=1e-4, fc=FC, mask=mask, txs=tx, mode="r32")
    s.run(1, 1)
    t0 = time.perf_counter()
    s.run(2, 2)
    one = time.perf_counter() - t0
    k = int(max(2, min(d.nt - 3, max_seconds / max(one, 1e-3))))
    t0 = time.perf_counter()
    s.run(3, 2 + k)
    el = time.perf_counter() - t0
    return {"value": d.nx * d.ny * d.nz * k / el / 1e9, "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"{k} consecutive steps (3..{2 + k}) of the same 256^3 workload, oracle in the reference's "
                      f"Float32 arithmetic, OpenMP over k, {el:.1f} s"}


def calibrate_copy():
    import torch

    n = 1 << 28  # 1 GiB of float32
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / 10
    print(json.dumps({"copy_bytes_each_way": n * 4, "ms": el * 1e3, "GBps_read_plus_write": 2 * n * 4 / el / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--fp64-steps", type=int, default=300)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=0)
    ap.add_argument("--calibrate-copy", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "independent solves shard across GPUs without communication, see DESIGN.md", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.profile_steps:
        args.steps, args.warmup, args.fp64_steps = args.profile_steps, 2, 0
        args.skip_cpu = True

    d, p, g = gpu_leg(args)
    n = d.nx
    f32 = g["f32"]
    b32 = algorithmic_bytes_per_step(n, p, 4)
    achieved = b32 / (f32["sweep_ms"] * 1e-3) / 1e9
    traffic = read_traffic("fused_sweep_256_f32_pmc.txt")
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
                     "kernel": "k_fused (field-update sweep)", "algorithmic_bytes_per_launch": b32,
                     "avg_launch_ms": round(f32["sweep_ms"], 5), "launches": f32["sweep_launches"]},
    }
    if "f64" in g:
        res["f64"] = {"glups": round(g["f64"]["glups"], 3), "ms_per_step": round(g["f64"]["ms_per_step"], 5),
                      "steps": g["f64"]["steps"], "sweep_ms": round(g["f64"]["sweep_ms"], 5),
                      "hbm_frac_sweep": round(algorithmic_bytes_per_step(n, p, 8) / (g["f64"]["sweep_ms"] * 1e-3)
                                              / 1e9 / HBM_PEAK_GBS, 4)}
    if not args.skip_cpu:
        d2, txs, p2 = build_workload(nt=4000)
        res["cpu_baseline"] = cpu_leg(d2, txs, p2)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
