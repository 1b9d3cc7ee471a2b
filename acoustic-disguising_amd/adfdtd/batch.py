"""Batched engine: many independent solves advanced together (adf_batch_*).

The reference runs its solves one after another: `impulsive_gfs` calls
`run_fdtd!` once per source and source type (src/greens.jl:143-164) and
`build_reverb_data` once per illumination point (src/reverb.jl:265-312).
A `Batch` holds any number of `Solver`s (each with its own grid, CPML, mask,
transceivers and step counter) and advances all of them by the same number
of steps per call; each member's result is the one `Solver.run` gives for
the same steps, whatever else is in the batch.
"""
from __future__ import annotations

import ctypes as C
from typing import List

from . import _lib as L
from .solver import Solver, context


class Batch:
    """Wraps adf_batch_* (include/adfdtd.h)."""

    def __init__(self, device: int = 0, persistent: bool = False, graph_steps: int = 16):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.adf_batch_create(context(device), C.byref(h)), "adf_batch_create")
        self.h = h
        self.members: List[Solver] = []
        self.set_option(L.ADF_BATCH_PERSISTENT, 1 if persistent else 0)
        self.set_option(L.ADF_BATCH_GRAPH_STEPS, graph_steps)

    def __del__(self):
        try:
            if getattr(self, "h", None) and self.h.value:
                self.lib.adf_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def __len__(self) -> int:
        return len(self.members)

    def add(self, s: Solver) -> None:
        L.check(self.lib.adf_batch_add(self.h, s.h), "batch_add")
        self.members.append(s)  # keeps the solver alive while it is a member

    def remove(self, s: Solver) -> None:
        L.check(self.lib.adf_batch_remove(self.h, s.h), "batch_remove")
        self.members.remove(s)

    def run(self, nsteps: int) -> None:
        """Advance every member by nsteps from its own next step."""
        L.check(self.lib.adf_batch_run(self.h, int(nsteps)), "batch_run")

    def set_option(self, option: int, value: int) -> None:
        L.check(self.lib.adf_batch_set_option(self.h, option, int(value)), "batch_set_option")

    def last_step_ms(self) -> float:
        v = C.c_double()
        L.check(self.lib.adf_batch_last_step_ms(self.h, C.byref(v)), "batch_last_step_ms")
        return v.value
