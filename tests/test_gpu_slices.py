"""adf_solver_get_pressure_slice: the on_step snapshots of build_hologram
(src/hologram.jl:238-247).  Indices are 1-based, as in Julia; axis 0 is
the plane i = index (a (nz, ny) array in the reversed-axes layout), axis 1
j = index, axis 2 k = index.  Every axis, the boundary and interior planes,
both storage types, after a set_field and after steps, against the field
the solver holds; out-of-range axes and indices are refused."""
import numpy as np
import pytest

import adfdtd as A
from adfdtd import _lib as L
from cases import build

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_slices_equal_the_held_field(dtype):
    d, txs, cpml, mask = build(n=(37, 30, 26), h=(0.00686, 0.0075, 0.006), nt=12, npml=3, n_rec=0)
    s = A.Solver(d, A.dtype_code(dtype, "fast"), A.make_cpml(d, cpml.n, cpml.rcoef, cpml.fc, dtype))
    f0 = np.random.default_rng(4).standard_normal((4, d.nz, d.ny, d.nx)).astype(dtype)
    s.set_field(f0)
    s.set_transceivers(A.transceivers_to_grid(d, txs, dtype))
    for when in ("after set_field", "after steps"):
        if when == "after steps":
            s.run(1, 7)
        P = s.get_field()[0]
        for i in (1, 2, d.nx // 2, d.nx - 1, d.nx):
            assert np.array_equal(s.pressure_slice(0, i), P[:, :, i - 1]), (when, "x", i)
        for j in (1, d.ny // 3, d.ny):
            assert np.array_equal(s.pressure_slice(1, j), P[:, j - 1, :]), (when, "y", j)
        for k in (1, d.nz // 2 + 1, d.nz):
            assert np.array_equal(s.pressure_slice(2, k), P[k - 1, :, :]), (when, "z", k)
    for axis, index in ((0, 0), (0, d.nx + 1), (1, d.ny + 1), (2, 0), (2, d.nz + 1)):
        with pytest.raises(L.AdfError, match="slice index out of range"):
            s.pressure_slice(axis, index)
    out = np.empty((d.nz, d.ny), dtype)
    for axis in (-1, 3):
        with pytest.raises(L.AdfError, match="slice index out of range"):
            L.check(s.lib.adf_solver_get_pressure_slice(s.h, axis, 1, L.ptr(out)), "pressure_slice")
