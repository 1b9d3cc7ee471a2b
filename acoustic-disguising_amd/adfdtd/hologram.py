"""Hologram synthesis, mirroring src/hologram.jl.

`build_hologram(dom, cfg, gf; ...)` (src/hologram.jl:141-278) drives one
FDTD solve from the initial plane wave (scaled by `initial_amplitude`) with
the outer Fibonacci surface, the inner surface and any extra transceivers;
in RefMode / MDDMode the GF map's iRec is the outer surface and its iSrc
the inner one (`_prepare_gf_map`, :62-68), so the in-loop extrapolation
drives the inner surface from the outer records.  The solve runs on the
device (adf_solver_*); with `snapshot_every` the pressure on three
orthogonal planes per offset is copied out every N steps
(adf_solver_get_pressure_slice), the reference's `on_step` closure
(:225-247).

Layouts follow the package: Julia arrays with their axes reversed, so a
reference x-slice (ny, nz) is a numpy (nz, ny) array.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from .illumination import illumination_points
from .simulation import Domain, coords, initial_field, make_cpml, make_transceiver, t, transceivers_to_grid
from .solver import Solver, dtype_code


class HologramMode:
    """src/hologram.jl:18-37: which extrapolation path build_hologram takes."""

    name = ""

    def __repr__(self):
        return f"{type(self).__name__}()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))

    def __str__(self):
        return self.name


class RealMode(HologramMode):
    """No GF extrapolation: the surfaces only record."""

    name = "real"


class RefMode(HologramMode):
    """Extrapolation with load_gf_ref's map (analytical / impulsive GFs)."""

    name = "ref"


class MDDMode(HologramMode):
    """Extrapolation with load_gf_mdd's map."""

    name = "mdd"


def hologram_mode(s) -> HologramMode:
    """src/hologram.jl:46-52."""
    s = str(s).lstrip(":")
    if s == "real":
        return RealMode()
    if s == "ref":
        return RefMode()
    if s == "mdd":
        return MDDMode()
    raise ValueError(f"hologram_mode: unknown :{s} (expected :real, :ref, or :mdd)")


def _prepare_gf_map(mode: HologramMode, gf, outer_range, inner_range):
    """src/hologram.jl:62-68: the outer surface's columns feed the map, the
    inner surface's columns receive its output."""
    if isinstance(mode, RealMode):
        return None
    if gf is None:
        raise ValueError("build_hologram: this implementation requires `gf` (use load_gf_ref/load_gf_mdd).")
    gf.iRec = list(outer_range)
    gf.iSrc = list(inner_range)
    return gf


def slice_indices(dom: Domain, offsets_m: Sequence[float]):
    """1-based indices of the pressure-grid planes nearest each offset:
    argmin(abs.(cs[:a][:p] .- o)) per axis (src/hologram.jl:225-230; the
    first index wins a tie, as Julia's argmin)."""
    cs = coords(dom)
    return [{a: int(np.argmin(np.abs(cs[a]["p"] - float(o)))) + 1 for a in ("x", "y", "z")} for o in offsets_m]


def build_hologram(dom: Domain, cfg: dict, gf, scatterer: str = "cross",
                   implementation: HologramMode = MDDMode(), gf_of: str = "heterogeneous", fc: float = 4000.0,
                   initial_distance: float = 0.65, initial_amplitude: float = 1.5, va=None,
                   txs_extra: Sequence = (), snapshot_every: Optional[int] = None,
                   slice_offsets_m: Sequence[float] = (-0.3, 0.0), dtype=np.float32, precision: str = "fast",
                   solver: Optional[Solver] = None) -> dict:
    """src/hologram.jl:141-278.  Returns field_final (4, nz, ny, nx),
    xSlices / ySlices / zSlices (per offset, a list of per-snapshot pressure
    planes: (nz, ny), (nz, nx), (ny, nx)), snapshot_steps, slice_offsets_m,
    slice_indices (1-based), times, txs_on_grid (the transceivers with
    their traces after the solve) and metadata.  `dtype` plays the role of
    Data.Array's number type: float32 runs the fast float path
    (precision="fast") or the reference's Float32 arithmetic
    (precision="ref"); float64 the Float64 algorithm.  `solver`: an
    existing Solver of this domain, dtype and CPML to reuse."""
    if isinstance(implementation, RealMode) and scatterer != "none" and va is None:
        raise ValueError(f"build_hologram: implementation=RealMode(), scatterer=:{scatterer} requires `va` "
                         "(build with AcousticDisguising.build_update_mask).")
    T = np.dtype(dtype).type
    rectf = np.zeros(dom.nt)
    ill = illumination_points(cfg)
    outer = [make_transceiver(dom, ill["outer"][i], np.r_[0.0, ill["outer_normals"][i]], rectf)
             for i in range(ill["outer"].shape[0])]
    inner = [make_transceiver(dom, ill["inner"][i], np.r_[0.0, ill["inner_normals"][i]], rectf)
             for i in range(ill["inner"].shape[0])]
    txs = outer + inner + list(txs_extra)
    n_outer = int(cfg["surfaces"]["nPoints_outer"])
    n_inner = int(cfg["surfaces"]["nPoints_inner"])
    outer_range = range(1, n_outer + 1)
    inner_range = range(n_outer + 1, n_outer + n_inner + 1)

    pml = cfg["pml"]
    cpml = make_cpml(dom, int(pml["n"]), float(pml["rcoef"]), float(pml["fc"]), T)
    txs_on_grid = transceivers_to_grid(dom, txs, T)
    gf_map = _prepare_gf_map(implementation, gf, outer_range, inner_range)

    # Data.Array(field_host) * Float32(initial_amplitude): the Float32 field
    # in the storage type times the amplitude rounded to Float32
    field = initial_field(dom, float(fc), float(initial_distance)).astype(T) * T(np.float32(initial_amplitude))

    idx = slice_indices(dom, slice_offsets_m)
    n_off = len(idx)
    xs, ys, zs = [[] for _ in range(n_off)], [[] for _ in range(n_off)], [[] for _ in range(n_off)]
    steps = []

    s = solver if solver is not None else Solver(dom, dtype_code(T, precision), cpml)
    s.set_field(field)
    s.set_cpml_memory(cpml)
    s.set_mask(va)
    s.set_transceivers(txs_on_grid)
    s.set_gf_map(gf_map)
    if snapshot_every is None:
        s.run(1, dom.nt)
    else:
        N = int(snapshot_every)
        if N < 1:
            raise ValueError("build_hologram: snapshot_every must be >= 1")
        it = 1
        while it <= dom.nt:
            end = min(dom.nt, ((it + N - 1) // N) * N)  # the next multiple of N, or nt
            s.run(it, end)
            if end % N == 0:
                for k in range(n_off):
                    xs[k].append(s.pressure_slice(0, idx[k]["x"]))
                    ys[k].append(s.pressure_slice(1, idx[k]["y"]))
                    zs[k].append(s.pressure_slice(2, idx[k]["z"]))
                steps.append(end)
            it = end + 1
    s.get_field(field)
    s.get_transceivers(txs_on_grid)

    metadata = {"scatterer": scatterer, "implementation": implementation, "gf_of": gf_of, "fc": float(fc),
                "initial_distance": float(initial_distance), "initial_amplitude": float(initial_amplitude),
                "n_outer": n_outer, "n_inner": n_inner, "outer_range": outer_range, "inner_range": inner_range}
    return {"field_final": field, "xSlices": xs, "ySlices": ys, "zSlices": zs, "snapshot_steps": steps,
            "slice_offsets_m": [float(o) for o in slice_offsets_m], "slice_indices": idx, "times": t(dom),
            "txs_on_grid": txs_on_grid, "metadata": metadata}
