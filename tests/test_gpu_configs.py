"""The benchmark configurations not covered at their own sizes elsewhere:

- the 64^3 Float64 impulsive GF (air, dx = lambda_min/10 at 2 kHz, 10-cell
  CPML, 600 steps, Ricker monopole exactly at the centre, 100 receivers on a
  Fibonacci sphere of radius 20 cells), whole run: GPU Float64 equals the
  oracle bit for bit, float32 within 1e-5, and the traces match the
  analytic free-space monopole within the reference's reported 1-6 %
  (greens_io.jl:145-147);
- the uneven batch: 48 solves, 16 each at 96^3, 128^3 and 160^3 in one
  batch, each member equal to its solo run bit for bit;
- the served stream: requests of 256^3, 192^3, 128^3 and 96^3 with sphere,
  cube and ellipsoid scatterers arriving while others run, under a budget
  that holds the 256^3 request but not it and the 192^3 one together; each
  result equals its solo run bit for bit.
"""
import time

import numpy as np
import pytest

import adfdtd as A
import oracle as O
from adfdtd import _lib as L
from cases import oracle_solver, rel_l2

pytestmark = pytest.mark.gpu

C_AIR, RHO_AIR, FC = 343.0, 1.21, 2000.0
DX = C_AIR / (2.5 * FC) / 10


def _domain(n, nt):
    xm = (n - 1) * DX / 2
    d0 = A.make_domain(1.0, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)
    return A.make_domain((nt - 0.5) * d0.dt, xm, xm, xm, n, n, n, C_AIR, RHO_AIR, 0.5)


def _traces(s):
    return [s.get_trace(w) for w in range(10)]


def test_64_fp64_600_steps_centre_monopole_100_receivers():
    d = _domain(64, 600)
    assert abs(d.dx - 6.86e-3) < 1e-15 and abs(d.dt - 5.7735e-6) < 1e-10 and d.nt == 600
    tf = A.ricker_wavelet(d, FC, True)
    pts, nrm = A.fibonacci_sphere(100, 20 * d.dx)
    z = np.zeros(d.nt)
    txs = [A.make_transceiver(d, [0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0], tf)]
    txs += [A.make_transceiver(d, pts[i], np.r_[0.0, nrm[i]], z) for i in range(100)]
    cpml = A.make_cpml(d, 10, 1e-4, FC)
    ref = oracle_solver(O, d, txs, cpml, None, "f64").run()
    res = {}
    for dtype in (np.float64, np.float32):
        s = A.Solver(d, A.dtype_code(dtype, "fast"), A.make_cpml(d, 10, 1e-4, FC, dtype))
        s.set_transceivers(A.transceivers_to_grid(d, txs, dtype))
        s.run(1, d.nt)
        cp = A.make_cpml(d, 10, 1e-4, FC, dtype)
        s.get_cpml_memory(cp)
        res[dtype] = (s.get_field(), _traces(s), cp.p + cp.v)
        del s
    f, tr, psi = res[np.float64]
    assert np.array_equal(f, ref.field)
    assert np.array_equal(tr[0], ref.tx["rec"][0]) and np.array_equal(tr[4], ref.tx["vn_rec"])
    for a, b in zip(psi, ref.mp + ref.mv):
        assert np.array_equal(a, b)
    f32, tr32, _ = res[np.float32]
    assert rel_l2(tr32[0], ref.tx["rec"][0]) < 1e-5 and rel_l2(tr32[4], ref.tx["vn_rec"]) < 1e-5
    # analytic free-space monopole: a P-channel injection g records
    # p(r, t) = dx^3/(c^2 dt) g'(t - r/c)/(4 pi r) (C_p, greens_io.jl:163,182-185)
    g = tr[5][0]  # p source trace of the monopole
    gd = np.gradient(g, d.dt)
    tt = np.arange(1, d.nt + 1) * d.dt
    errs = []
    for i in range(100):
        r = np.linalg.norm(pts[i])
        an = d.dx ** 3 / (d.c0 ** 2 * d.dt) * np.interp(tt - r / d.c0 + 0.5 * d.dt, tt, gd, left=0) / (4 * np.pi * r)
        errs.append(np.linalg.norm(tr[0][i + 1] - an) / np.linalg.norm(an))
    print("analytic monopole rel-L2: median", np.median(errs), "max", max(errs), flush=True)
    assert np.median(errs) < 0.06 and max(errs) < 0.12


def _batch_member(n, i, nt, per=16):
    """Member i of the uneven batch: the 128^3 spec (sphere r = 16 cells,
    10-cell CPML, source on a 40-cell Fibonacci sphere, 100 receivers at 20
    cells) with every radius scaled by n/128."""
    d = _domain(n, nt)
    sc = n / 128
    src, _ = A.fibonacci_sphere(per, 40 * sc * d.dx)
    rec, rn = A.fibonacci_sphere(100, 20 * sc * d.dx)
    tf, _ = A.impulsive_wavelet(d, FC)
    z = np.zeros(d.nt)
    s = A.Solver(d, A.dtype_code(np.float32, "fast"), A.make_cpml(d, 10, 1e-4, FC, np.float32))
    s.build_mask("sphere", [16 * sc * d.dx])
    txs = [A.make_transceiver(d, src[i], [1.0, 0.0, 0.0, 0.0], tf)]
    txs += [A.make_transceiver(d, rec[k], np.r_[0.0, rn[k]], z) for k in range(100)]
    s.set_transceivers(A.transceivers_to_grid(d, txs, np.float32))
    return s


def test_uneven_batch_48_solves_96_128_160_equal_solo():
    steps = 6
    sizes = (96, 128, 160)
    rng = np.random.default_rng(11)
    fields = {(n, i): rng.standard_normal((4, n, n, n), dtype=np.float32) for n in sizes for i in (0, 7, 15)}
    b = A.Batch()
    mem = {}
    for n in sizes:
        for i in range(16):
            s = _batch_member(n, i, steps + 1)
            if (n, i) in fields:
                s.set_field(fields[(n, i)])
            b.add(s)
            mem[(n, i)] = s
    assert len(b) == 48
    b.run(steps)
    for key, fld in fields.items():
        solo = _batch_member(key[0], key[1], steps + 1)
        solo.set_field(fld)
        solo.run(1, steps)
        assert np.array_equal(mem[key].get_field(), solo.get_field()), key
        assert all(np.array_equal(x, y) for x, y in zip(_traces(mem[key]), _traces(solo))), key
        del solo
    # members started from rest carry their impulsive source into the traces
    assert np.abs(mem[(128, 3)].get_trace(L.ADF_REC_P)).max() > 0


def _served_requests():
    """(domain, mask spec, transceivers, field) of a served mix: one request
    per grid size, the scatterers sphere, cube and ellipsoid."""
    spec = [(256, 70, "sphere", lambda x: [0.2 * x]), (192, 60, "cube", lambda x: [0.5 * x]),
            (128, 80, "ellipsoid", lambda x: [0.25 * x, 0.15 * x, 0.2 * x]), (96, 50, "sphere", lambda x: [0.12 * x]),
            (192, 45, "ellipsoid", lambda x: [0.1 * x, 0.3 * x, 0.2 * x])]
    out = []
    for k, (n, nt, shape, prm) in enumerate(spec):
        d = _domain(n, nt)
        srcs, _ = A.fibonacci_sphere(16, 0.4 * d.xmax)
        rec, rn = A.fibonacci_sphere(100, 0.2 * d.xmax)
        tf, _ = A.impulsive_wavelet(d, FC)
        z = np.zeros(d.nt)
        txs = [A.make_transceiver(d, srcs[k], [1.0, 0.0, 0.0, 0.0], tf)]
        txs += [A.make_transceiver(d, rec[i], np.r_[0.0, rn[i]], z) for i in range(100)]
        out.append((d, (shape, prm(d.xmax)), txs))
    return out


def _served_solo(req):
    d, mspec, txs = req
    s = A.Solver(d, L.ADF_F32, A.make_cpml(d, 10, 1e-4, FC, np.float32))
    s.build_mask(*mspec)
    s.set_transceivers(A.transceivers_to_grid(d, txs, np.float32))
    s.run(1, d.nt)
    out = _traces(s), s.get_field()
    del s
    return out


def test_served_mix_with_256_and_192_requests_equals_solo():
    reqs = _served_requests()
    ref = [_served_solo(r) for r in reqs]
    # the 256^3 request holds about 612 MB of device memory (two field
    # buffers, CPML memory, mask, traces) and a 192^3 one about 270 MB: a
    # 700 MB budget admits the first alone, so the requests behind it queue
    # and are admitted as memory frees up
    budget = 700 << 20
    with A.Server(budget_bytes=budget, chunk=16) as sv:
        ids = []
        for k, (d, mspec, txs) in enumerate(reqs):
            ids.append(sv.submit(d, L.ADF_F32, A.make_cpml(d, 10, 1e-4, FC, np.float32),
                                 A.transceivers_to_grid(d, txs, np.float32), mask=mspec, want=0x3FF,
                                 keep_field=True))
            time.sleep(0.01 * (k % 2))  # some arrivals while others run
        for rid in ids:
            assert sv.wait(rid, 300.0) == "done"
        for k, (rid, (tr, fld)) in enumerate(zip(ids, ref)):
            for w in range(10):
                assert np.array_equal(sv.trace(rid, w), tr[w]), (k, w)
            assert np.array_equal(sv.field(rid), fld), k
            sv.release(rid)
        st = sv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["peak_bytes"] <= budget
        assert st["p99_ms"] >= st["p50_ms"] > 0
