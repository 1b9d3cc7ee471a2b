"""Parity at the benchmark workloads' full sizes (256^3 single solve, 192^3
hologram synthesis, 64 x 128^3 batch), a few steps each from a seeded random
field, so that every cell, every CPML slab cell, the scatterer faces and all
transceivers are active from the first step.

Bars: the GPU's Float64 path equals the oracle's Float64 bit for bit (field,
CPML memory, every trace); the fast float32 path is within 1e-5 relative L2
of it; batch members equal their solo runs bit for bit."""
import numpy as np
import pytest

import adfdtd as A
import oracle as O
from adfdtd import _lib as L
from cases import oracle_domain, oracle_solver, rel_l2

pytestmark = pytest.mark.gpu

C_AIR, RHO_AIR, FC = 343.0, 1.21, 2000.0
DX = C_AIR / (2.5 * FC) / 10


def _domain(n, nt):
    xm = (n - 1) * DX / 2
    d0 = A.make_domain(1.0, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)
    return A.make_domain((nt - 0.5) * d0.dt, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)


def _state(s, cp_dtype, d, npml):
    cp = A.make_cpml(d, npml, 1e-4, FC, cp_dtype)
    s.get_cpml_memory(cp)
    return s.get_field(), [s.get_trace(w) for w in range(10)], list(cp.p) + list(cp.v)


def _compare_with_oracle(d, txs, npml, mask_shape, mask_params, field, steps, store_nt=False):
    """store_nt: also run the float32 solve with non-temporal stores, which
    must give the same bits."""
    mask = O.build_update_mask(oracle_domain(d), mask_shape, mask_params)
    assert np.array_equal(mask, A.build_update_mask(d, mask_shape, mask_params))
    cpml = A.make_cpml(d, npml, 1e-4, FC)
    ref = oracle_solver(O, d, txs, cpml, mask, "f64", field=field).run(1, steps)
    ref_psi = ref.mp + ref.mv
    out = {}
    runs = [(np.float64, 0), (np.float32, 0)] + ([(np.float32, 1)] if store_nt else [])
    for dtype, nt in runs:
        s = A.Solver(d, A.dtype_code(dtype, "fast"), A.make_cpml(d, npml, 1e-4, FC, dtype))
        s.build_mask(mask_shape, mask_params)
        s.set_field(np.asarray(field, dtype))
        s.set_transceivers(A.transceivers_to_grid(d, txs, dtype))
        s.set_path("auto", store_nt=nt)
        s.run(1, steps)
        f, tr, psi = _state(s, dtype, d, npml)
        if nt:
            a = out["f32_state"]
            assert np.array_equal(f, a[0]) and all(np.array_equal(x, y) for x, y in zip(tr + psi, a[1] + a[2]))
            del s
            continue
        if dtype == np.float32:
            out["f32_state"] = (f, tr, psi)
        errs = {"field": rel_l2(f, ref.field),
                "psi": max(rel_l2(a, b) for a, b in zip(psi, ref_psi)),
                "p_rec": rel_l2(tr[0], ref.tx["rec"][0]), "vn_rec": rel_l2(tr[4], ref.tx["vn_rec"]),
                "src": rel_l2(np.stack(tr[5:9]), ref.tx["src"])}
        out[dtype] = errs
        del s
    out.pop("f32_state", None)
    print(out, flush=True)
    assert all(v == 0.0 for v in out[np.float64].values()), out[np.float64]
    assert all(v <= 1e-5 for v in out[np.float32].values()), out[np.float32]


def test_256_sphere_cpml16_2501_transceivers_full_size():
    """The 256^3 benchmark workload: rigid sphere r = 24 cells, 16-cell
    CPML, Ricker source 60 cells off-centre, 2000-point KH sphere (r = 40
    cells) and 500 far-field receivers (r = 80 cells)."""
    steps = 5
    d = _domain(256, steps + 1)
    tf = A.ricker_wavelet(d, FC, True)
    kh, khn = A.fibonacci_sphere(2000, 40 * d.dx)
    ff, ffn = A.fibonacci_sphere(500, 80 * d.dx)
    z = np.zeros(d.nt)
    txs = [A.make_transceiver(d, [60 * d.dx, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0], tf)]
    txs += [A.make_transceiver(d, kh[i], np.r_[0.0, khn[i]], z) for i in range(2000)]
    txs += [A.make_transceiver(d, ff[i], np.r_[0.0, ffn[i]], z) for i in range(500)]
    field = np.random.default_rng(1).standard_normal((4, 256, 256, 256))
    _compare_with_oracle(d, txs, 16, "sphere", [24 * d.dx], field, steps, store_nt=True)


def test_192_hologram_cube_kh_injection_full_size():
    """The 192^3 hologram workload: cube of edge 30 cells, 12-cell CPML,
    2000 KH points driven through C_for_N (seeded Ricker-filtered noise as
    p and vn drives), 500 far-field receivers at 80 cells."""
    steps = 5
    d = _domain(192, steps + 1)
    kh, khn = A.fibonacci_sphere(2000, 40 * d.dx)
    ff, ffn = A.fibonacci_sphere(500, 80 * d.dx)
    rng = np.random.default_rng(2)
    rick = A.ricker_wavelet(d, FC, False)
    drive = lambda: np.convolve(rng.standard_normal(d.nt), rick)[: d.nt]
    cf = A.c_for_n(2000, 40 * d.dx, d)
    txs = []
    for i in range(2000):
        txs.append(A.make_transceiver(d, kh[i], [cf["C_q"], 0.0, 0.0, 0.0], drive()))
    for i in range(2000):
        txs.append(A.make_transceiver(d, kh[i], np.r_[0.0, cf["C_f"] * khn[i]], drive()))
    z = np.zeros(d.nt)
    txs += [A.make_transceiver(d, ff[i], np.r_[0.0, ffn[i]], z) for i in range(500)]
    field = np.random.default_rng(3).standard_normal((4, 192, 192, 192))
    _compare_with_oracle(d, txs, 12, "cube", [30 * d.dx], field, steps)


def test_batch_64_members_128_full_size():
    """The batched workload: 64 solves at 128^3 (sphere r = 16 cells, 10-cell
    CPML, sources on a 40-cell Fibonacci sphere, 100 receivers at 20 cells),
    random initial fields; each member equals its solo run bit for bit."""
    steps = 6
    d = _domain(128, steps + 1)
    src, _ = A.fibonacci_sphere(64, 40 * d.dx)
    rec, rn = A.fibonacci_sphere(100, 20 * d.dx)
    tf, _ = A.impulsive_wavelet(d, FC)
    z = np.zeros(d.nt)
    rng = np.random.default_rng(4)

    def member(i):
        s = A.Solver(d, A.dtype_code(np.float32, "fast"), A.make_cpml(d, 10, 1e-4, FC, np.float32))
        s.build_mask("sphere", [16 * d.dx])
        txs = [A.make_transceiver(d, src[i], [1.0, 0.0, 0.0, 0.0], tf)]
        txs += [A.make_transceiver(d, rec[k], np.r_[0.0, rn[k]], z) for k in range(100)]
        s.set_transceivers(A.transceivers_to_grid(d, txs, np.float32))
        return s

    fields = [rng.standard_normal((4, 128, 128, 128)).astype(np.float32) for _ in range(64)]
    b = A.Batch()
    mem = []
    for i in range(64):
        s = member(i)
        s.set_field(fields[i])
        b.add(s)
        mem.append(s)
    b.run(steps)
    for i in (0, 17, 40, 63):
        solo = member(i)
        solo.set_field(fields[i])
        solo.run(1, steps)
        a, c = _state(mem[i], np.float32, d, 10), _state(solo, np.float32, d, 10)
        assert np.array_equal(a[0], c[0]), i
        assert all(np.array_equal(x, y) for x, y in zip(a[1], c[1])), i
        assert all(np.array_equal(x, y) for x, y in zip(a[2], c[2])), i
        del solo
