"""GF maps for the in-loop extrapolation, mirroring src/kernels/extrapolation.jl.

Tensors are numpy arrays in the reversed layout of the Julia arrays:
Julia (nt_gf, a, b) -> numpy (b, a, nt_gf).  iRec / iSrc are 1-based
transceiver columns, populated by the caller as in src/hologram.jl:62-68.
`writeback=True` adds the extrapolated sources into the source slots at
it+1, as documented (docs/impulsive_gf_hologram_path.md:154-171);
`writeback=False` computes and discards them, which is what the reference's
copying views do (extrapolation.jl:215-216,226-232,246-247).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np


@dataclass
class PVFromPV:
    """Two-way form (extrapolation.jl:45-54): pp, pv, vp, vv (nt_gf, nm, nout)."""

    pp: np.ndarray
    pv: np.ndarray
    vp: np.ndarray
    vv: np.ndarray
    iRec: List[int] = field(default_factory=list)
    iSrc: List[int] = field(default_factory=list)
    writeback: bool = True

    def kind_and_tensors(self):
        return 2, (self.pp, self.pv, self.vp, self.vv)


@dataclass
class PVFromPin:
    """One-way incoming-pressure form (extrapolation.jl:64-71): p_pin, v_pin
    (nt_gf, nsrc, nrec)."""

    p_pin: np.ndarray
    v_pin: np.ndarray
    iRec: List[int] = field(default_factory=list)
    iSrc: List[int] = field(default_factory=list)
    writeback: bool = True

    def kind_and_tensors(self):
        return 1, (self.p_pin, self.v_pin)


@dataclass
class PFromPin:
    """Split one-way form (extrapolation.jl:80-87): pos, neg (nt_gf, nrec, nsrc)."""

    pos: np.ndarray
    neg: np.ndarray
    iRec: List[int] = field(default_factory=list)
    iSrc: List[int] = field(default_factory=list)
    writeback: bool = True

    def kind_and_tensors(self):
        return 3, (self.pos, self.neg)
