// Interior tiles of the fused sweep, written for occupancy: no y/z CPML, no
// rigid or mirrored planes and no clamped rows can occur in them (the host's
// tile split guarantees it), so the only special cases left are along x and
// they are handled without branches:
//   - the x-slab CPML coefficients of a lane's cells are lane constants that
//     are zero outside the slab, so the memory update runs on every cell and
//     reduces to dv + 0 outside (the same value the reference computes);
//     the memory is loaded as one vector per row at the lane's slab offset
//     and stored back only for the cells inside the slab;
//   - P at i = nx-1 and Vx at i = 0 (never updated by the reference) are
//     kept with a select.
// Every row access is a buffer load/store with the row offset in a scalar
// register.  Per plane: all loads, then P1 on rows -1..TY-1, then the
// velocities of the owned rows and the stores (updates.jl:21-142,192-275,
// 288-501; forward.jl:66-107).
#pragma once
#include "adf_fused.hpp"

namespace adf {

// Lane constants of the x direction for one phase: per-cell coefficients
// (zero outside the slab), the lane's element offset into the phase's
// x-slab CPML array (side and in-row position of its first cell) and the
// bitmask of its cells inside the slab.
template <typename T, int CPL>
struct XSlab {
    T a[CPL], b[CPL];
    int off;
    uint32_t in;
};

// Phase with slab width w (p for P, p+1 for V) over cells [lo, hi) of the
// row; `base` is the array's element offset in psi0/psi1.
template <typename T, int CPL>
__device__ __forceinline__ XSlab<T, CPL> xslab(const Fused<T>& F, int i0, int w, int lo, int hi, bool on,
                                               int coef_a, int coef_b, int base) {
    XSlab<T, CPL> X;
    X.in = 0;
    X.off = base;
    int side = 0, l0 = 0;
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        const int i = i0 + q;
        int l = 0, sd = 0;
        if (on && i >= lo && i < hi) sd = slab(i, F.nx, w, l);
        X.a[q] = sd ? F.coef[coef_a + l + w * (sd - 1)] : (T)0;
        X.b[q] = sd ? F.coef[coef_b + l + w * (sd - 1)] : (T)0;
        if (sd) {
            X.in |= 1u << q;
            if (!side) { side = sd; l0 = l - q; }
        }
    }
    // element offset of cell i0 (may point before/after the slab row for
    // cells outside it; those lanes only load, never store)
    if (side) X.off = base + l0 + w * F.ny * F.nz * (side - 1);
    return X;
}

// Loaded rows of one plane of an interior tile (rows -1..TY-1, Vy to TY).
template <typename T, int CPL, int TY>
struct IPlane {
    T p0[TY + 1][CPL], vx0[TY + 1][CPL], vy0[TY + 2][CPL], vzn[TY + 1][CPL];
    T mPx[TY + 1][CPL], mVx[TY + 1][CPL];
    uint32_t mp_[TY + 1], mvx_[TY + 1], mvy_[TY + 1], mvz_[TY + 1];
};

template <typename T, bool EXACT, int CPL, int TY, bool PF = false>
__device__ __forceinline__ void sweep_int_tile(const Fused<T>& F, int tyi, int k0, int k1, bool masked) {
    typedef FA<T, EXACT> M_;
    typedef typename M_::A A;
    constexpr int ES = (int)sizeof(T);
    const int nx = F.nx, ny = F.ny, nz = F.nz, p = F.p, qv = p + 1;
    const int sxy = nx * ny;
    const int lane = threadIdx.x & 63;
    const int i0 = lane * CPL;
    const bool lane_ok = i0 < nx;
    const int i0c = lane_ok ? i0 : nx - CPL;
    const int j0 = tyi * TY;
    const bool cpP = F.cpml_p != 0, cpV = qv > 1;

    const uint32_t chb = (uint32_t)F.n3 * (uint32_t)ES;
    const Rsrc rF0 = rsrc_of(F.f0, 4u * chb), rF1 = rsrc_of(F.f1, 4u * chb);
    const Rsrc rS0 = rsrc_of(F.psi0, (uint32_t)F.psi_n * ES), rS1 = rsrc_of(F.psi1, (uint32_t)F.psi_n * ES);
    const uint32_t* Mk = masked ? F.mask : nullptr;
    const Rsrc rM = rsrc_of(Mk, Mk ? (uint32_t)(4 * F.nwx * ny * nz) * 4u : 0u);
    const int vl = i0c * ES;
    const int vm = (i0c >> 5) * 4;
    const int sh = i0c & 31;
    auto rowb = [&](int j, int k) { return (int)((uint32_t)(nx * j + sxy * k) * (uint32_t)ES); };
    auto chrow = [&](int c, int rb) { return (int)((uint32_t)c * chb + (uint32_t)rb); };
    auto maskb = [&](int plane, int j, int k) { return (((plane * nz + k) * ny + j) * F.nwx) * 4; };

    // x-slab constants: P phase over i < nx-1, V phase over 1 <= i < nx
    const XSlab<T, CPL> XP = xslab<T, CPL>(F, i0, p, 0, nx - 1, cpP, 0, 2 * p, F.mv_off(0));
    const XSlab<T, CPL> XV = xslab<T, CPL>(F, i0, qv, 1, nx, cpV, 4 * p, 4 * p + 2 * qv, F.mp_off(0));
    const int vP = XP.off * ES, vV = XV.off * ES;
    const bool anyP = cpP && XP.in, anyV = cpV && XV.in;

    T vzc[TY + 1][CPL];     // Vz0 at the current plane, rows -1..TY-1
    T p1prev[TY + 1][CPL];  // P1 at the previous plane
#pragma unroll
    for (int r = 0; r <= TY; ++r) bld<T, CPL>(rF0, vl, chrow(3, rowb(j0 + r - 1, k0 - 1)), vzc[r]);

    // Every load of plane k into L (mask words kept raw: no arithmetic on a
    // loaded value here, so the loads of the next plane stay in flight while
    // the current one is computed).
    auto load = [&](IPlane<T, CPL, TY>& L, int k) {
        const bool owned = k >= k0;
#pragma unroll
        for (int r = 0; r <= TY + 1; ++r) {
            const int j = j0 + r - 1;
            const int rb = rowb(j, k);
            bld<T, CPL>(rF0, vl, chrow(2, rb), L.vy0[r]);
            if (r > TY) continue;
            bld<T, CPL>(rF0, vl, rb, L.p0[r]);
            bld<T, CPL>(rF0, vl, chrow(1, rb), L.vx0[r]);
            bld<T, CPL>(rF0, vl, chrow(3, rowb(j, k + 1)), L.vzn[r]);
            L.mp_[r] = L.mvx_[r] = L.mvy_[r] = L.mvz_[r] = 0xffffffffu;
            if (Mk) {
                L.mp_[r] = bld_u32(rM, vm, maskb(0, j, k));
                if (owned && r > 0) {
                    L.mvx_[r] = bld_u32(rM, vm, maskb(1, j, k));
                    L.mvy_[r] = bld_u32(rM, vm, maskb(2, j, k));
                    L.mvz_[r] = bld_u32(rM, vm, maskb(3, j, k));
                }
            }
            if (cpP) bld<T, CPL>(rS0, vP, p * (j + ny * k) * ES, L.mPx[r]);
            if (cpV && owned && r > 0) bld<T, CPL>(rS0, vV, qv * (j + ny * k) * ES, L.mVx[r]);
        }
    };

    // P1 on rows -1..TY-1 of plane k, then the velocities and stores of the
    // owned rows (shuffles in uniform control flow; only lanes inside the
    // row store)
    auto compute = [&](IPlane<T, CPL, TY>& L, int k) {
        const bool owned = k >= k0;
        T p1[TY + 1][CPL];
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
            const T vxl = shfl_down1<T>(L.vx0[r][0]);
            const uint32_t mp = L.mp_[r] >> sh;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const T vxn = q + 1 < CPL ? L.vx0[r][q + 1] : vxl;
                A dvx = M_::diff(vxn, L.vx0[r][q], F.dx, F.rdx);
                if (cpP) dvx = M_::psi(L.mPx[r][q], XP.b[q], XP.a[q], dvx);
                const A dvy = M_::diff(L.vy0[r + 1][q], L.vy0[r][q], F.dy, F.rdy);
                const A dvz = M_::diff(L.vzn[r][q], vzc[r][q], F.dz, F.rdz);
                const T out = M_::upd(L.p0[r][q], F.fq, F.fqf, (mp >> q) & 1u, (dvx + dvy) + dvz);
                p1[r][q] = (i0 + q < nx - 1) ? out : L.p0[r][q];
            }
        }
        if (owned) {
#pragma unroll
            for (int r = 1; r <= TY; ++r) {
                const int j = j0 + r - 1;
                const T pml = shfl_up1<T>(p1[r][CPL - 1]);
                const int rb = rowb(j, k);
                const uint32_t mvx = L.mvx_[r] >> sh, mvy = L.mvy_[r] >> sh, mvz = L.mvz_[r] >> sh;
                T vx1[CPL], vy1[CPL], vz1[CPL];
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const T pm = q > 0 ? p1[r][q - 1] : pml;
                    A dpx = M_::diff(p1[r][q], pm, F.dx, F.rdx);
                    if (cpV) dpx = M_::psi(L.mVx[r][q], XV.b[q], XV.a[q], dpx);
                    const T vx = M_::upd(L.vx0[r][q], F.fv, F.fvf, (mvx >> q) & 1u, dpx);
                    vx1[q] = (i0 + q >= 1) ? vx : L.vx0[r][q];
                    vy1[q] = M_::upd(L.vy0[r][q], F.fv, F.fvf, (mvy >> q) & 1u,
                                     M_::diff(p1[r][q], p1[r - 1][q], F.dy, F.rdy));
                    vz1[q] = M_::upd(vzc[r][q], F.fv, F.fvf, (mvz >> q) & 1u,
                                     M_::diff(p1[r][q], p1prev[r][q], F.dz, F.rdz));
                }
                if (!lane_ok) continue;
                const int vs = i0 * ES;
                bst<T, CPL>(rF1, vs, rb, p1[r]);
                bst<T, CPL>(rF1, vs, chrow(1, rb), vx1);
                bst<T, CPL>(rF1, vs, chrow(2, rb), vy1);
                bst<T, CPL>(rF1, vs, chrow(3, rb), vz1);
                // CPML memory of the slab cells
                if (anyP) {
                    const int so = p * (j + ny * k) * ES;
                    if (XP.in == (1u << CPL) - 1) {
                        bst<T, CPL>(rS1, vP, so, L.mPx[r]);
                    } else {
#pragma unroll
                        for (int q = 0; q < CPL; ++q)
                            if ((XP.in >> q) & 1u) {
                                const T one[1] = {L.mPx[r][q]};
                                bst<T, 1>(rS1, vP + q * ES, so, one);
                            }
                    }
                }
                if (anyV) {
                    const int so = qv * (j + ny * k) * ES;
                    if (XV.in == (1u << CPL) - 1) {
                        bst<T, CPL>(rS1, vV, so, L.mVx[r]);
                    } else {
#pragma unroll
                        for (int q = 0; q < CPL; ++q)
                            if ((XV.in >> q) & 1u) {
                                const T one[1] = {L.mVx[r][q]};
                                bst<T, 1>(rS1, vV + q * ES, so, one);
                            }
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r <= TY; ++r)
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                p1prev[r][q] = p1[r][q];
                vzc[r][q] = L.vzn[r][q];
            }
    };

    if constexpr (PF) {
        // software pipeline: the loads of plane k+1 are issued before plane
        // k is computed (two plane buffers, loop unrolled by two)
        IPlane<T, CPL, TY> B0, B1;
        load(B0, k0 - 1);
        int k = k0 - 1;
        for (;;) {
            if (k + 1 < k1) load(B1, k + 1);
            compute(B0, k);
            if (++k >= k1) break;
            if (k + 1 < k1) load(B0, k + 1);
            compute(B1, k);
            if (++k >= k1) break;
        }
    } else {
        for (int k = k0 - 1; k < k1; ++k) {
            IPlane<T, CPL, TY> L;
            load(L, k);
            compute(L, k);
        }
    }
}

// One solve's tiles (entries as xcd_tile describes), GEN: general tiles
// (CPML slabs in y or z, boundary planes, clamped rows), else interior
// tiles; W waves per workgroup take W consecutive tiles.  PF: interior
// tiles load plane k+1 before computing plane k.
template <typename T, bool EXACT, int CPL, int TY, bool GEN, bool PF = false>
__global__ __launch_bounds__(256) void k_sweep(Fused<T> F, const uint2* __restrict__ tiles, int ntiles) {
    const int W = (int)(blockDim.x >> 6);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int t = xcd_tile((int)blockIdx.x, (int)gridDim.x) * W + w;
    if (t >= ntiles) return;
    const uint2 e = tiles[t];
    const int tyi = (int)(e.x & 0xffffu), k0 = (int)(e.y >> 16), k1 = k0 + (int)(e.y & 0xffffu);
    const bool masked = (e.x >> 31) != 0;
    if constexpr (GEN) fused_tile<T, EXACT, CPL, TY, true>(F, tyi, k0, k1, masked);
    else sweep_int_tile<T, EXACT, CPL, TY, PF>(F, tyi, k0, k1, masked);
}

}  // namespace adf
