"""A second, independently written restatement of one reference time step,
in whole-plane numpy operations, for brute-force checks of the oracle on
tiny grids.  It follows the Julia source directly:

  forward_onestep!          src/kernels/forward.jl:13-139
  update_p! / *_cpml! / *_mask! / *_cpml_mask!
                            src/kernels/updates.jl:21-142, 288-378
  update_v{x,y,z}! / *_mask!
                            src/kernels/updates.jl:192-275, 380-501
  getBoundaryIndex          src/kernels/derivatives.jl:66-79
  interp_trilinear!         src/kernels/derivatives.jl:30-60
  v2vn! / vn2v!             src/kernels/derivatives.jl:85-97

Arrays use the oracle's numpy layouts (the Julia column-major arrays with
their axes reversed): field (4, nz, ny, nx); P-phase memory mv[a] and
V-phase memory mp[a] as (2, ., ., .) with the CPML extent on the axis of a;
coefficient tables (2, p) and (2, p+1); transceiver idx (4, ntx, 3)
1-based, w (4, ntx, 8), src/rec (4, ntx, nt), dir (ntx, 4), vn_* (ntx, nt).

Arithmetic follows Julia's promotion rules for the array type T: with
Float32 arrays, differences of array values stay Float32 and are promoted
to Float64 by the Float64 domain scalars (dx, fq, fv); memory updates mix
Float32 table products with Float64 derivatives and are stored back as
Float32.  No step here fuses a multiply and an add.
"""
from __future__ import annotations

import numpy as np


def boundary_index(i, ndim, npml):
    """getBoundaryIndex (derivatives.jl:66-79), 1-based: (l, side) with
    side 0 outside the slabs, 1 left, 2 right."""
    if i <= npml:
        return i, 1
    ii = i - (ndim - npml)
    return (ii, 2) if ii > 0 else (ii, 0)


class NaiveState:
    def __init__(self, d, field, mv, mp, a, b, ah, bh, mask, tx):
        self.d = d
        self.T = field.dtype.type
        self.f = field
        self.mv, self.mp = mv, mp
        self.a, self.b, self.ah, self.bh = a, b, ah, bh
        self.mask = None if mask is None else mask.astype(self.T)  # Data.Array cast of the Bool mask
        self.tx = tx

    # ----------------------------------------------------------- pressure
    def _p_rigid_planes(self):
        """update_p! over (1:nx-1, ny, 1:nz-1) then (1:nx-1, 1:ny-1, nz)
        (forward.jl:66-70): no memory, no mask; -2 V / d at the rigid side."""
        d, F = self.d, self.f
        P, Vx, Vy, Vz = F
        dx, dy, dz, fq = (np.float64(v) for v in (d.dx, d.dy, d.dz, d.fact_m0))
        nx, ny, nz = d.nx, d.ny, d.nz
        # plane j = ny: i in 1..nx-1, k in 1..nz-1
        sl = (slice(0, nz - 1), ny - 1, slice(0, nx - 1))
        ddx = (Vx[0:nz - 1, ny - 1, 1:nx] - Vx[0:nz - 1, ny - 1, 0:nx - 1]) / dx
        ddy = np.float64(-2) * Vy[sl] / dy
        ddz = (Vz[1:nz, ny - 1, 0:nx - 1] - Vz[0:nz - 1, ny - 1, 0:nx - 1]) / dz
        P[sl] = P[sl] - fq * (ddx + ddy + ddz)
        # plane k = nz: i in 1..nx-1, j in 1..ny-1
        sl = (nz - 1, slice(0, ny - 1), slice(0, nx - 1))
        ddx = (Vx[nz - 1, 0:ny - 1, 1:nx] - Vx[nz - 1, 0:ny - 1, 0:nx - 1]) / dx
        ddy = (Vy[nz - 1, 1:ny, 0:nx - 1] - Vy[nz - 1, 0:ny - 1, 0:nx - 1]) / dy
        ddz = np.float64(-2) * Vz[sl] / dz
        P[sl] = P[sl] - fq * (ddx + ddy + ddz)

    def _memory(self, dv, psi, a, b, axis, n_axis, npml, lo):
        """Memory update of one phase along one numpy axis (0 = z, 1 = y,
        2 = x) over a region whose first index on that axis is the 1-based
        `lo`; dv is the derivative over the region (Float64), updated in
        place, psi the (2, ...) memory array of the full grid."""
        T = self.T
        for pos in range(dv.shape[axis]):
            i = lo + pos
            l, side = boundary_index(i, n_axis, npml)
            if side == 0:
                continue
            dsl = [slice(None)] * 3
            dsl[axis] = pos
            dsl = tuple(dsl)
            # the memory spans the whole grid on the other two axes; every
            # update range starts at index 1 on them
            msl = tuple([side - 1] + [(l - 1) if ax == axis else slice(0, dv.shape[ax]) for ax in range(3)])
            m_new = (b[side - 1, l - 1] * psi[msl]) + (a[side - 1, l - 1] * dv[dsl])
            psi[msl] = m_new.astype(T)
            dv[dsl] = dv[dsl] + psi[msl]

    def _p_interior(self):
        d, F = self.d, self.f
        P, Vx, Vy, Vz = F
        dx, dy, dz, fq = (np.float64(v) for v in (d.dx, d.dy, d.dz, d.fact_m0))
        nx, ny, nz = d.nx, d.ny, d.nz
        p = self.a.shape[1]
        R = (slice(0, nz - 1), slice(0, ny - 1), slice(0, nx - 1))
        ddx = (Vx[0:nz - 1, 0:ny - 1, 1:nx] - Vx[R]) / dx
        ddy = (Vy[0:nz - 1, 1:ny, 0:nx - 1] - Vy[R]) / dy
        ddz = (Vz[1:nz, 0:ny - 1, 0:nx - 1] - Vz[R]) / dz
        if p > 0:
            masked = self.mask is not None
            # update_p_cpml! skips an axis whose memory extent is 1 (the
            # `npml > 1` guard); update_p_cpml_mask! has no guard
            if masked or p > 1:
                self._memory(ddx, self.mv[0], self.a, self.b, 2, nx, p, 1)
                self._memory(ddy, self.mv[1], self.a, self.b, 1, ny, p, 1)
                self._memory(ddz, self.mv[2], self.a, self.b, 0, nz, p, 1)
        if self.mask is None:
            P[R] = P[R] - fq * (ddx + ddy + ddz)
        else:
            coeff = fq * self.mask[0][R]
            P[R] = P[R] - coeff * (ddx + ddy + ddz)

    # ----------------------------------------------------------- velocity
    def _v(self, c):
        d, F = self.d, self.f
        P = F[0]
        V = F[c]
        nx, ny, nz = d.nx, d.ny, d.nz
        fv = np.float64(d.fact_m1)
        h = np.float64((d.dx, d.dy, d.dz)[c - 1])
        q = self.ah.shape[1]
        axis = 3 - c  # numpy axis of x, y, z
        n_axis = (nx, ny, nz)[c - 1]
        R = [slice(None)] * 3
        R[axis] = slice(1, n_axis)
        R = tuple(R)
        lo = [slice(None)] * 3
        lo[axis] = slice(0, n_axis - 1)
        dp = (P[R] - P[tuple(lo)]) / h
        if q > 1:  # the `npml > 1` guard on the p+1 extent
            self._memory(dp, self.mp[c - 1], self.ah, self.bh, axis, n_axis, q, 2)
        if self.mask is None:
            V[R] = V[R] - fv * dp
        else:
            coeff = fv * self.mask[c][R]
            V[R] = V[R] - coeff * dp

    # -------------------------------------------------------- transceivers
    def _interp(self, ch, it):
        """interp_trilinear! for every transceiver in order, corners with i
        fastest: record, then inject."""
        d, tx = self.d, self.tx
        Fc = self.f[ch]
        for s in range(tx["idx"].shape[1]):
            I, J, K = (int(v) for v in tx["idx"][ch, s])
            if not (1 <= I < d.nx and 1 <= J < d.ny and 1 <= K < d.nz):
                continue
            for qd in range(8):
                a, b, c = qd & 1, (qd >> 1) & 1, qd >> 2
                w = tx["w"][ch, s, qd]
                cell = (K - 1 + c, J - 1 + b, I - 1 + a)
                tx["rec"][ch, s, it - 1] = tx["rec"][ch, s, it - 1] + w * Fc[cell]
                Fc[cell] = Fc[cell] + w * tx["src"][ch, s, it - 1]

    def step(self, it):
        d, tx = self.d, self.tx
        self._p_rigid_planes()
        self._p_interior()
        self._interp(0, it)
        for c in (1, 2, 3):
            self._v(c)
        for c in (1, 2, 3):
            self._interp(c, it)
        n = tx["dir"]
        r = tx["rec"]
        tx["vn_rec"][:, it - 1] = r[1, :, it - 1] * n[:, 1] + r[2, :, it - 1] * n[:, 2] + r[3, :, it - 1] * n[:, 3]
        if it < d.nt:
            for c in (1, 2, 3):
                tx["src"][c, :, it] = tx["src"][c, :, it] + tx["vn_src"][:, it] * n[:, c]
        F = self.f
        F[2][:, 0, :] = -F[2][:, 1, :]
        F[3][0, :, :] = -F[3][1, :, :]
