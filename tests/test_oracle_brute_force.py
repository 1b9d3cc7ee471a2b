"""Brute force on tiny grids: the oracle's C step against a second,
independently written numpy restatement of the reference step
(tests/naive_step.py), bit for bit, in both of the oracle's arithmetics.

The grids are anisotropic (dx, dy, dz all different, nx, ny, nz all
different) so an axis mix-up shows, start from a random field and random
CPML memory so every cell and memory entry takes part from the first step,
and carry injectors on every channel plus a velocity source (vn2v!)."""
import numpy as np
import pytest

import oracle as O
from naive_step import NaiveState


def aniso_domain(n=(15, 12, 10), h=(0.0071, 0.0093, 0.0057), nt=14, c0=343.0, r0=1.21, cf=0.5):
    nx, ny, nz = n
    xm, ym, zm = (nx - 1) * h[0] / 2, (ny - 1) * h[1] / 2, (nz - 1) * h[2] / 2
    d0 = O.domain(1.0, xm, ym, zm, nx, ny, nz, c0, r0, cf)
    return O.domain((nt - 0.5) * d0.dt, xm, ym, zm, nx, ny, nz, c0, r0, cf)


def make_case(mode, npml, shape, seed=3):
    d = aniso_domain()
    assert len({d.dx, d.dy, d.dz}) == 3
    T = np.float64 if mode == "f64" else np.float32
    rng = np.random.default_rng(seed)
    pts = [np.array([0.21, -0.33, 0.12]) * d.xmax, np.array([-0.4, 0.25, -0.3]) * d.xmax,
           np.array([0.05, 0.41, 0.2]) * d.xmax, np.array([0.33, 0.1, -0.45]) * d.xmax,
           np.array([0.33, 0.1, -0.45]) * d.xmax + 0.3 * d.dx]  # two overlapping stencils
    dirs = [[1.0, 0, 0, 0], [0, 0.6, -0.8, 0.0], [0.3, 0.2, 0.5, -0.7], [0, 0, 0, 1.0], [1.0, 0.5, 0.5, 0.5]]
    txs = [(pts[i], dirs[i], rng.standard_normal(d.nt)) for i in range(len(pts))]
    tx = O.transceivers_to_grid(d, txs, T)
    tx["vn_src"][...] = rng.standard_normal(tx["vn_src"].shape).astype(T)
    mask = O.build_update_mask(d, shape, [4.2 * d.dx]) if shape != "none" else None
    field = (rng.standard_normal((4, d.nz, d.ny, d.nx)) * 1e-2).astype(T)
    s = O.RefSolver(d, npml=npml, rcoef=1e-4, fc=2000.0, mask=mask, txs=tx, mode=mode, field=field.copy())
    for arr in s.mv + s.mp:
        arr[...] = (rng.standard_normal(arr.shape) * 1e-3).astype(T)
    naive = NaiveState(d, field.copy(), [m.copy() for m in s.mv], [m.copy() for m in s.mp], s.a, s.b, s.ah, s.bh,
                       mask, {k: v.copy() for k, v in s.tx.items()})
    return d, s, naive


@pytest.mark.parametrize("mode", ["f64", "r32"])
@pytest.mark.parametrize("npml,shape", [(3, "cube"), (3, "none"), (1, "none"), (1, "cube"), (2, "sphere"),
                                        (0, "cube"), (0, "none")])
def test_oracle_step_equals_independent_restatement(mode, npml, shape):
    d, s, naive = make_case(mode, npml, shape)
    for it in range(1, d.nt + 1):
        s.run(it, it)
        naive.step(it)
        np.testing.assert_array_equal(s.field, naive.f, err_msg=f"field after step {it}")
    for a in range(3):
        np.testing.assert_array_equal(s.mv[a], naive.mv[a], err_msg=f"P-phase memory axis {a}")
        np.testing.assert_array_equal(s.mp[a], naive.mp[a], err_msg=f"V-phase memory axis {a}")
    for k in ("rec", "src", "vn_rec"):
        np.testing.assert_array_equal(s.tx[k], naive.tx[k], err_msg=k)
    # the case exercised every part it claims to
    assert np.abs(s.tx["rec"]).min(axis=2).max() > 0
    if npml > 1:
        assert all(np.abs(m).max() > 0 for m in s.mv)


@pytest.mark.parametrize("plant", ["half_grid_tables_on_grid", "left_right_swapped", "dy_for_dx"])
def test_planted_oracle_mistakes_are_caught(plant):
    """Mistakes planted in the oracle's inputs (on the CPU) make the two
    restatements disagree: using the on-grid profile for the velocity
    memory, swapping the left and right slabs, or reading the domain with
    dy in place of dx."""
    d, s, naive = make_case("f64", 3, "cube")
    if plant == "half_grid_tables_on_grid":
        s.ah = np.ascontiguousarray(np.c_[s.a, s.a[:, -1:]])
        s.bh = np.ascontiguousarray(np.c_[s.b, s.b[:, -1:]])
    elif plant == "left_right_swapped":
        s.a, s.b = np.ascontiguousarray(s.a[::-1]), np.ascontiguousarray(s.b[::-1])
    else:
        s.d = type(d).from_buffer_copy(d)  # the oracle's own copy of the domain
        s.d.dx = d.dy
    s.run(1, 4)
    for it in range(1, 5):
        naive.step(it)
    assert not np.array_equal(s.field, naive.f)
