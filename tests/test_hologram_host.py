"""build_hologram's host logic (src/hologram.jl:18-68,154-160,225-230):
mode parsing, the GF map's index vectors, the argument checks and the
argmin slice planes."""
import numpy as np
import pytest

import adfdtd as A
from cases import domain_cells


def test_hologram_mode_parsing():
    assert A.hologram_mode("real") == A.RealMode() and A.hologram_mode(":ref") == A.RefMode()
    assert A.hologram_mode("mdd") == A.MDDMode() and str(A.MDDMode()) == "mdd"
    with pytest.raises(ValueError, match="unknown :focus"):
        A.hologram_mode("focus")


def test_gf_map_indices_are_outer_then_inner():
    from adfdtd.hologram import _prepare_gf_map

    gf = A.PVFromPin(np.zeros((3, 2, 4)), np.zeros((3, 2, 4)))
    out = _prepare_gf_map(A.RefMode(), gf, range(1, 4), range(4, 6))
    assert out is gf and gf.iRec == [1, 2, 3] and gf.iSrc == [4, 5]
    assert _prepare_gf_map(A.RealMode(), gf, range(1, 4), range(4, 6)) is None
    with pytest.raises(ValueError, match="requires `gf`"):
        _prepare_gf_map(A.MDDMode(), None, range(1, 4), range(4, 6))


def test_real_mode_with_a_scatterer_needs_the_mask():
    d = domain_cells(20, 0.00686, 10)
    cfg = {"surfaces": {"nPoints_outer": 4, "radius_outer": 0.03, "nPoints_inner": 3, "radius_inner": 0.02},
           "pml": {"n": 2, "rcoef": 1e-4, "fc": 2000.0}}
    with pytest.raises(ValueError, match="requires `va`"):
        A.build_hologram(d, cfg, None, scatterer="cube", implementation=A.RealMode())


@pytest.mark.parametrize("offsets", [[-0.3, 0.0], [0.0137, -0.02], [1.0, -1.0]])
def test_slice_planes_are_the_nearest_pressure_nodes(offsets):
    d = domain_cells((41, 30, 27), (0.00686, 0.0071, 0.0066), 10)
    cs = A.coords(d)
    for o, got in zip(offsets, A.slice_indices(d, offsets)):
        for a in "xyz":
            dist = np.abs(cs[a]["p"] - o)
            first = min(i for i in range(len(dist)) if dist[i] == dist.min())  # Julia's argmin: first minimum
            assert got[a] == first + 1
