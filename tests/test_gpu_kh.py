"""Post-hoc KH on the device (adf_kh_analytic_accumulate) against the
oracle's eval_one_kernel / eval_kernel_pedestal / direct_kh_accumulate
(kh_kernels.jl:106-157,174-203,256-296): Float64 within 1e-12 relative L2,
Float32 within 1e-5 of the Float64 oracle, for every kernel, truncated and
untruncated kernel lengths, with and without the pedestal form."""
import numpy as np
import pytest

import adfdtd as A
import oracle as O

pytestmark = pytest.mark.gpu

C0, RHO0, F3DB = 343.0, 1.21, 2000.0


def setup(n_outer=60, n_inner=7, nt=320, seed=1):
    rng = np.random.default_rng(seed)
    op, on = A.fibonacci_sphere(n_outer, 0.08)
    ip, inn = A.fibonacci_sphere(n_inner, 0.25)
    sigma = np.sqrt(np.log(2)) / (2 * np.pi * F3DB)
    dt = sigma / 4
    inp = rng.standard_normal((n_outer, nt))
    return op, on, ip, inn, sigma, dt, inp


def reference(which, op, on, ip, inn, tg, sigma, inp, nt, alpha, pedestal):
    K = O.eval_one_kernel(which, op, on, ip, inn, tg, sigma, C0, RHO0)
    out = np.zeros((ip.shape[0], nt))
    if pedestal:
        Kdc = O.eval_kernel_pedestal(which, op, on, ip, inn, RHO0)
        K = K - Kdc[:, :, None]
        O.direct_kh_accumulate(out, K, inp, alpha, K_dc=Kdc, cumsum_in=np.cumsum(inp, axis=1))
    else:
        O.direct_kh_accumulate(out, K, inp, alpha)
    return out


@pytest.mark.parametrize("which", ["p_p", "p_v", "v_p", "v_v"])
@pytest.mark.parametrize("ntK", [320, 200])
def test_analytic_kh_matches_oracle(which, ntK):
    op, on, ip, inn, sigma, dt, inp = setup()
    nt = inp.shape[1]
    tg = np.arange(ntK) * dt
    peds = [False, True] if which == "v_v" else [False]
    for ped in peds:
        ref = reference(which, op, on, ip, inn, tg, sigma, inp, nt, -0.7, ped)
        base = np.full_like(ref, 0.25)  # accumulates into existing values
        g64 = A.kh_analytic_accumulate(base.copy(), which, op, on, ip, inn, tg, sigma, C0, RHO0, inp, alpha=-0.7,
                                       pedestal=ped, dtype=np.float64) - 0.25
        e64 = np.linalg.norm(g64 - ref) / np.linalg.norm(ref)
        g32 = A.kh_analytic_accumulate(base.copy(), which, op, on, ip, inn, tg, sigma, C0, RHO0, inp, alpha=-0.7,
                                       pedestal=ped, dtype=np.float32) - 0.25
        e32 = np.linalg.norm(g32 - ref) / np.linalg.norm(ref)
        print(which, ntK, ped, f"f64 {e64:.2e} f32 {e32:.2e}", flush=True)
        assert e64 <= 1e-12 and e32 <= 1e-5, (which, ntK, ped, e64, e32)


def test_non_uniform_time_grid_and_wide_window():
    # window wider than one LDS chunk (sigma / dt large) and a stretched grid
    op, on, ip, inn, sigma, dt, inp = setup(n_outer=20, n_inner=3, nt=1500, seed=4)
    tg = np.cumsum(np.full(1500, sigma / 60.0)) * (1 + 0.1 * np.linspace(0, 1, 1500))
    ref = reference("v_v", op, on, ip, inn, tg, sigma, inp, 1500, 1.0, False)
    g = A.kh_analytic_accumulate(np.zeros_like(ref), "v_v", op, on, ip, inn, tg, sigma, C0, RHO0, inp)
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) <= 1e-12
