"""The benchmark's step ranges, checked on the CPU from the exact command
lines it is run with: every leg must stay inside its domain's step count."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

ARGVS = [
    ["--gpus", "1", "--steps", "20", "--warmup", "5"],
    [],
    ["--steps", "1", "--warmup", "0"],
    ["--steps", "2000", "--warmup", "50", "--fp64-steps", "0"],
    ["--profile-steps", "10"],
]


@pytest.mark.parametrize("argv", ARGVS, ids=[" ".join(a) or "defaults" for a in ARGVS])
def test_every_256_leg_fits_its_domain(argv):
    args = bench.parse_args(argv)
    if args.profile_steps:
        args.steps, args.warmup, args.fp64_steps = args.profile_steps, 2, 0
    nt, legs = bench.plan_steps(args)
    assert "f32" in legs
    for name, (w0, w1, t0, t1) in legs.items():
        assert w0 == 1 and w1 == args.warmup and t0 == w1 + 1
        assert t1 - t0 + 1 == (args.steps if name == "f32" else args.fp64_steps)
        assert 1 <= t0 <= t1 <= nt - 1, (name, legs[name], nt)
    # the solver refuses start > end + 1 and end > nt (adf_solver_run)
    assert nt == max(r[3] for r in legs.values()) + 1


def test_driver_command_runs_the_fp64_leg_in_range():
    args = bench.parse_args(["--gpus", "1", "--steps", "20", "--warmup", "5"])
    nt, legs = bench.plan_steps(args)
    assert legs["f64"] == (1, 5, 6, 5 + args.fp64_steps) and legs["f64"][3] < nt


@pytest.mark.parametrize("steps", [1, 2, 5, 20, 79, 80, 2000])
def test_sweep_samples_cover_the_timed_region(steps):
    per = bench.sweep_sample_period(steps)
    samples = len(range(0, steps, per))
    assert samples >= min(5, steps)
    assert per <= bench.SWEEP_SAMPLE


@pytest.mark.parametrize("one", [1e-6, 0.01, 0.5, 100.0])
def test_cpu_sample_stays_in_the_domain(one):
    for nt in (5, 4000):
        k = bench.cpu_sample_steps(nt, one, 15.0)
        assert 2 <= k and 2 + k <= nt


def test_cpu_threads_use_physical_cores_within_the_share():
    info = bench.host_cpu()
    assert info["physical_cores"] >= 1 and info["logical_cpus"] >= info["physical_cores"]
    assert bench.cpu_threads(dict(info, thread_share=None)) == info["physical_cores"]
    assert bench.cpu_threads(dict(info, physical_cores=128, thread_share=16)) == 16
    assert bench.cpu_threads(dict(info, physical_cores=8, thread_share=16)) == 8


def test_gpus_other_than_one_exit_with_a_message(capsys):
    sys_argv = sys.argv
    sys.argv = ["bench.py", "--gpus", "2"]
    try:
        with pytest.raises(SystemExit) as e:
            bench.main()
    finally:
        sys.argv = sys_argv
    assert e.value.code == 2
    assert "one GPU" in capsys.readouterr().err
