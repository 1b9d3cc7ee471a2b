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
// x-slab CPML array (side and in-row position of its first cell), the
// bitmask of its cells inside the slab and the bitmask of its cells whose
// position lies inside the slab row's extent [0, w) (a lane at the edge of
// the slab may reach into the neighbouring row's memory).
template <typename T, int CPL>
struct XSlab {
    T a[CPL], b[CPL];
    int off;
    uint32_t in, ext;
};

// Phase with slab width w (p for P, p+1 for V) over cells [lo, hi) of the
// row; `base` is the array's element offset in psi0/psi1.  carry: see below.
template <typename T, int CPL>
__device__ __forceinline__ XSlab<T, CPL> xslab(const Fused<T>& F, int i0, int w, int lo, int hi, bool on,
                                               int coef_a, int coef_b, int base, bool carry = false) {
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
    // carry: a cell inside the extent but outside [lo, hi) (P at i = nx-1,
    // Vx at i = 0: never updated) gets b = 1, a = 0, so its memory passes
    // through the update unchanged (b m + 0 dv; dv there is a finite field
    // difference, and the cell keeps its field value, so dv + psi is not
    // used) and a vector store writes it back as loaded, without a select.
    // The exact routines keep it with a select instead (m + 0 would turn a
    // memory of -0.0 into +0.0).
    X.ext = 0;
#pragma unroll
    for (int q = 0; q < CPL; ++q)
        if (side && l0 + q >= 0 && l0 + q < w) {
            X.ext |= 1u << q;
            if (carry && !((X.in >> q) & 1u)) X.b[q] = (T)1;
        }
    return X;
}

// Loaded rows of one plane of an interior tile (rows -1..TY-1, Vy to TY).
template <typename T, int CPL, int TY>
struct IPlane {
    T p0[TY + 1][CPL], vx0[TY + 1][CPL], vy0[TY + 2][CPL], vzn[TY + 1][CPL];
    T mPx[TY + 1][CPL], mVx[TY + 1][CPL];
    uint32_t mp_[TY + 1], mvx_[TY + 1], mvy_[TY + 1], mvz_[TY + 1];
};

template <typename T, bool EXACT, int CPL, int TY, bool RAG = false>
__device__ __forceinline__ void sweep_int_tile(const Fused<T>& F, int j0, int k0, int k1, bool masked, int seg) {
    typedef FA<T, EXACT> M_;
    typedef typename M_::A A;
    constexpr int ES = (int)sizeof(T);
    const int nx = F.nx, ny = F.ny, nz = F.nz, p = F.p, qv = p + 1;
    const int sxy = nx * ny;
    const XLane XL = xlane<CPL>(nx, F.xseg, seg);
    const int i0 = XL.i0, i0c = XL.i0c;
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
                if (!XL.own) continue;
                const int vs = i0 * ES;
                if (RAG && XL.part) {
#pragma unroll
                    for (int q = 0; q < CPL; ++q)
                        if (i0 + q < nx) {
                            const T o[4][1] = {{p1[r][q]}, {vx1[q]}, {vy1[q]}, {vz1[q]}};
#pragma unroll
                            for (int c = 0; c < 4; ++c) bst<T, 1>(rF1, vs + q * ES, chrow(c, rb), o[c]);
                        }
                } else {
                    bst<T, CPL>(rF1, vs, rb, p1[r]);
                    bst<T, CPL>(rF1, vs, chrow(1, rb), vx1);
                    bst<T, CPL>(rF1, vs, chrow(2, rb), vy1);
                    bst<T, CPL>(rF1, vs, chrow(3, rb), vz1);
                }
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

    for (int k = k0 - 1; k < k1; ++k) {
        IPlane<T, CPL, TY> L;
        load(L, k);
        compute(L, k);
    }
}

// Byte offset that every buffer access treats as out of range: loads at it
// return zero and stores to it are dropped (the range check covers the
// vector and instruction offsets, not the scalar one).
constexpr uint32_t OOR = 0xfffff000u;

// Loaded rows of one plane of a pipelined tile: the interior rows of
// IPlane plus, for tiles in the y (SY) or z (SZ) CPML slabs, the slab
// memory rows of both phases.
template <typename T, int CPL, int TY, bool SY, bool SZ>
struct SPlane : IPlane<T, CPL, TY> {
    T mPy[SY ? TY + 1 : 1][CPL], mVy[SY ? TY + 1 : 1][CPL];
    T mPz[SZ ? TY + 1 : 1][CPL], mVz[SZ ? TY + 1 : 1][CPL];
    // z slab of the plane (uniform): psi row offsets of row j0 (bytes) and
    // the coefficients, zero outside the slab
    int zoP, zoV;
    bool zinP, zinV;
    T azP, bzP, azV, bzV;
};

// The regular-tile routine as a software pipeline.  The loads of plane k+1
// are issued before plane k is computed and stored, and every memory
// instruction of the loop body is issued unconditionally (disabled accesses
// go to a null resource or to the out-of-range offset), so each iteration
// issues the same sequence and the wait for plane k's data never waits for
// stores or for the loads of plane k+1.
//
// A regular tile keeps away from the mirrored rows/planes (owned rows and
// planes >= 2), the rigid ones (rows and planes <= n-2) and the grid edge
// (halo row and plane >= 1, Vy of row j0+TY and Vz of plane k1 inside the
// grid); the host's tile split guarantees it.  SY / SZ: some row / plane of
// the tile lies in a y / z CPML slab of either phase; the slab memory update
// then runs on every row / plane of the tile with zero coefficients (and a
// zero-reading null resource) outside the slab, which leaves the derivative
// unchanged, as the reference's slab test does (updates.jl:21-103,192-275,
// 320-378,380-501).  The x slab is handled the same way per lane.
// NOMASK: the instantiation for tiles that meet no masked cell (no mask
// loads or registers).
template <typename T, bool EXACT, int CPL, int TY, bool SY, bool SZ, int POL = 0, bool NOMASK = false,
          bool RAG = false>
__device__ __forceinline__ void sweep_tile_pf(const Fused<T>& F, int j0, int k0, int k1, bool masked, int seg) {
    typedef FA<T, EXACT> M_;
    typedef typename M_::A A;
    typedef SPlane<T, CPL, TY, SY, SZ> Plane;
    constexpr int ES = (int)sizeof(T);
    const int nx = F.nx, ny = F.ny, nz = F.nz, p = F.p, qv = p + 1;
    const int sxy = nx * ny;
    const XLane XL = xlane<CPL>(nx, F.xseg, seg);
    const int i0 = XL.i0, i0c = XL.i0c;
    // lanes storing one vector per row; with RAG the partial lane stores
    // cell by cell at offsets vq (out of range for every other lane)
    const bool lane_ok = XL.own && !(RAG && XL.part);
    int vq[RAG ? CPL - 1 : 1];
#pragma unroll
    for (int q = 0; q < (RAG ? CPL - 1 : 1); ++q)
        vq[q] = (RAG && XL.part && i0 + q < nx) ? (i0 + q) * ES : (int)OOR;
    const bool cpP = F.cpml_p != 0, cpV = qv > 1;

    const uint32_t chb = (uint32_t)F.n3 * (uint32_t)ES;
    const Rsrc rF0 = rsrc_of(F.f0, 4u * chb), rF1 = rsrc_of(F.f1, 4u * chb);
    const uint32_t psib = (uint32_t)F.psi_n * ES;
    const Rsrc rS1 = rsrc_of(F.psi1, psib);
    // null resources (every access out of range) for the disabled streams
    const Rsrc rSP0 = rsrc_of(F.psi0, cpP ? psib : 0u), rSV0 = rsrc_of(F.psi0, cpV ? psib : 0u);
    const Rsrc rM = rsrc_of(F.mask, masked ? (uint32_t)(4 * F.nwx * ny * nz) * 4u : 0u);
    const uint32_t mfill = (masked && !NOMASK) ? 0u : 0xffffffffu;  // all open without a mask
    const int vl = i0c * ES;
    const int vs = lane_ok ? i0 * ES : (int)OOR;
    // one lane offset for the four field channels and one for the mask
    // planes: the channel's start goes into the scalar offset with the
    // row's (the range check that drops the out-of-range stores covers the
    // vector offset only)
    const int vm = (i0c >> 5) * 4;
    const int sh = i0c & 31;
    auto rowb = [&](int j, int k) { return (int)((uint32_t)(nx * j + sxy * k) * (uint32_t)ES); };
    auto chs = [&](int c, int rb) { return (int)((uint32_t)rb + (uint32_t)c * chb); };
    auto maskb = [&](int plane, int j, int k) { return (((plane * nz + k) * ny + j) * F.nwx) * 4; };

    const XSlab<T, CPL> XP = xslab<T, CPL>(F, i0, p, 0, nx - 1, cpP, 0, 2 * p, F.mv_off(0), !EXACT);
    const XSlab<T, CPL> XV = xslab<T, CPL>(F, i0, qv, 1, nx, cpV, 4 * p, 4 * p + 2 * qv, F.mp_off(0), !EXACT);
    const int vP = XP.off * ES, vV = XV.off * ES;
    constexpr uint32_t FULL = (1u << CPL) - 1u;
    // CPML memory stores of the x-slab cells.  A lane with cells in the slab
    // whose positions all lie in the slab row's extent stores one vector
    // (its cells outside the update range carry their memory unchanged, see
    // xslab).  A lane reaching past the row's extent (into the neighbouring
    // row's memory) stores its updated cells one by one.  The other lanes'
    // offsets are out of range (those stores are dropped).
    const bool stP = XL.own && cpP && XP.in, stV = XL.own && cpV && XV.in;
    const int sPv = (stP && XP.ext == FULL) ? vP : (int)OOR;
    const int sVv = (stV && XV.ext == FULL) ? vV : (int)OOR;
    int sPq[CPL], sVq[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        sPq[q] = (stP && XP.ext != FULL && ((XP.in >> q) & 1u)) ? vP + q * ES : (int)OOR;
        sVq[q] = (stV && XV.ext != FULL && ((XV.in >> q) & 1u)) ? vV + q * ES : (int)OOR;
    }

    // y slab of each row (fixed for the tile): psi row offset at plane 0
    // (bytes; plane k adds k times the plane stride) and the coefficients
    int yoP[SY ? TY + 1 : 1], yoV[SY ? TY + 1 : 1];
    bool yinP[SY ? TY + 1 : 1], yinV[SY ? TY + 1 : 1];
    T ayP[SY ? TY + 1 : 1], byP[SY ? TY + 1 : 1], ayV[SY ? TY + 1 : 1], byV[SY ? TY + 1 : 1];
    if constexpr (SY) {
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
            const int j = j0 + r - 1;
            int l = 0;
            const int sd = cpP ? slab(j, ny, p, l) : 0;
            yinP[r] = sd != 0;
            yoP[r] = sd ? (F.mv_off(1) + nx * (l + p * nz * (sd - 1))) * ES : 0;
            ayP[r] = sd ? F.ca(l + p * (sd - 1)) : (T)0;
            byP[r] = sd ? F.cb(l + p * (sd - 1)) : (T)0;
            int lv = 0;
            const int sv = (cpV && r > 0) ? slab(j, ny, qv, lv) : 0;
            yinV[r] = sv != 0;
            yoV[r] = sv ? (F.mp_off(1) + nx * (lv + qv * nz * (sv - 1))) * ES : 0;
            ayV[r] = sv ? F.cah(lv + qv * (sv - 1)) : (T)0;
            byV[r] = sv ? F.cbh(lv + qv * (sv - 1)) : (T)0;
        }
    }
    const int ypP = nx * p * ES, ypV = nx * qv * ES;  // y-slab psi plane strides

    T vzc[TY + 1][CPL];
    T p1prev[TY + 1][CPL];
#pragma unroll
    for (int r = 0; r <= TY; ++r) bld<T, CPL>(rF0, vl, chs(3, rowb(j0 + r - 1, k0 - 1)), vzc[r]);

    auto load = [&](Plane& L, int k) {
        if constexpr (SZ) {
            int l = 0;
            const int sd = cpP ? slab(k, nz, p, l) : 0;
            L.zinP = sd != 0;
            L.zoP = sd ? (F.mv_off(2) + nx * (j0 - 1 + ny * (l + p * (sd - 1)))) * ES : 0;
            L.azP = sd ? F.ca(l + p * (sd - 1)) : (T)0;
            L.bzP = sd ? F.cb(l + p * (sd - 1)) : (T)0;
            int lv = 0;
            const int sv = cpV ? slab(k, nz, qv, lv) : 0;
            L.zinV = sv != 0;
            L.zoV = sv ? (F.mp_off(2) + nx * (j0 - 1 + ny * (lv + qv * (sv - 1)))) * ES : 0;
            L.azV = sv ? F.cah(lv + qv * (sv - 1)) : (T)0;
            L.bzV = sv ? F.cbh(lv + qv * (sv - 1)) : (T)0;
        }
#pragma unroll
        for (int r = 0; r <= TY + 1; ++r) {
            const int j = j0 + r - 1;
            const int rb = rowb(j, k);
            bld<T, CPL>(rF0, vl, chs(2, rb), L.vy0[r]);
            if (r > TY) continue;
            bld<T, CPL>(rF0, vl, rb, L.p0[r]);
            bld<T, CPL>(rF0, vl, chs(1, rb), L.vx0[r]);
            bld<T, CPL>(rF0, vl, chs(3, rowb(j, k + 1)), L.vzn[r]);
            bld<T, CPL>(rSP0, vP, p * (j + ny * k) * ES, L.mPx[r]);
            if constexpr (SY)
                bld<T, CPL>(rsrc_of(F.psi0, yinP[r] ? psib : 0u), vl, yoP[r] + k * ypP, L.mPy[r]);
            if constexpr (SZ) bld<T, CPL>(rsrc_of(F.psi0, L.zinP ? psib : 0u), vl, L.zoP + r * nx * ES, L.mPz[r]);
            if (r > 0) {
                bld<T, CPL>(rSV0, vV, qv * (j + ny * k) * ES, L.mVx[r]);
                if constexpr (SY)
                    bld<T, CPL>(rsrc_of(F.psi0, yinV[r] ? psib : 0u), vl, yoV[r] + k * ypV, L.mVy[r]);
                if constexpr (SZ)
                    bld<T, CPL>(rsrc_of(F.psi0, L.zinV ? psib : 0u), vl, L.zoV + r * nx * ES, L.mVz[r]);
            }
        }
        // update masks (a tile instantiated for unmasked tiles reads none:
        // mfill opens every cell)
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
            const int j = j0 + r - 1;
            if constexpr (NOMASK) {
                L.mp_[r] = L.mvx_[r] = L.mvy_[r] = L.mvz_[r] = 0u;
                continue;
            }
            L.mp_[r] = bld_u32(rM, vm, maskb(0, j, k));
            if (r > 0) {
                L.mvx_[r] = bld_u32(rM, vm, maskb(1, j, k));
                L.mvy_[r] = bld_u32(rM, vm, maskb(2, j, k));
                L.mvz_[r] = bld_u32(rM, vm, maskb(3, j, k));
            }
        }
    };

    auto pressure = [&](Plane& L, T (&p1)[TY + 1][CPL]) {
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
            const T vxl = shfl_down1<T>(L.vx0[r][0]);
            const uint32_t mp = (L.mp_[r] | mfill) >> sh;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const T vxn = q + 1 < CPL ? L.vx0[r][q + 1] : vxl;
                A dvx = M_::diff(vxn, L.vx0[r][q], F.dx, F.rdx);
                // P at i = nx-1 is never updated (forward.jl:67-68): that
                // cell keeps P and its slab memory (see xslab for the x slab)
                const bool upd = i0 + q < nx - 1;
                if (cpP) {
                    if constexpr (EXACT) dvx = M_::psi_keep(L.mPx[r][q], XP.b[q], XP.a[q], dvx, !upd);
                    else dvx = M_::psi(L.mPx[r][q], XP.b[q], XP.a[q], dvx);
                }
                A dvy = M_::diff(L.vy0[r + 1][q], L.vy0[r][q], F.dy, F.rdy);
                if constexpr (SY) {
                    const T m0 = L.mPy[r][q];
                    dvy = M_::psi(L.mPy[r][q], byP[r], ayP[r], dvy);
                    if (!upd) L.mPy[r][q] = m0;
                }
                A dvz = M_::diff(L.vzn[r][q], vzc[r][q], F.dz, F.rdz);
                if constexpr (SZ) {
                    const T m0 = L.mPz[r][q];
                    dvz = M_::psi(L.mPz[r][q], L.bzP, L.azP, dvz);
                    if (!upd) L.mPz[r][q] = m0;
                }
                const T out = M_::upd(L.p0[r][q], F.fq, F.fqf, (mp >> q) & 1u, (dvx + dvy) + dvz);
                p1[r][q] = upd ? out : L.p0[r][q];
            }
        }
    };
    auto carry = [&](Plane& L, const T (&p1)[TY + 1][CPL]) {
#pragma unroll
        for (int r = 0; r <= TY; ++r)
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                p1prev[r][q] = p1[r][q];
                vzc[r][q] = L.vzn[r][q];
            }
    };
    // an owned plane: pressure, velocities of the owned rows, all stores
    auto step = [&](Plane& L, int k) {
        T p1[TY + 1][CPL];
        pressure(L, p1);
#pragma unroll
        for (int r = 1; r <= TY; ++r) {
            const int j = j0 + r - 1;
            const T pml = shfl_up1<T>(p1[r][CPL - 1]);
            const int rb = rowb(j, k);
            const uint32_t mvx = (L.mvx_[r] | mfill) >> sh, mvy = (L.mvy_[r] | mfill) >> sh,
                           mvz = (L.mvz_[r] | mfill) >> sh;
            T vx1[CPL], vy1[CPL], vz1[CPL];
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const T pm = q > 0 ? p1[r][q - 1] : pml;
                A dpx = M_::diff(p1[r][q], pm, F.dx, F.rdx);
                if (cpV) {
                    if constexpr (EXACT) dpx = M_::psi_keep(L.mVx[r][q], XV.b[q], XV.a[q], dpx, i0 + q < 1);
                    else dpx = M_::psi(L.mVx[r][q], XV.b[q], XV.a[q], dpx);
                }
                const T vx = M_::upd(L.vx0[r][q], F.fv, F.fvf, (mvx >> q) & 1u, dpx);
                vx1[q] = (i0 + q >= 1) ? vx : L.vx0[r][q];
                A dpy = M_::diff(p1[r][q], p1[r - 1][q], F.dy, F.rdy);
                if constexpr (SY) dpy = M_::psi(L.mVy[r][q], byV[r], ayV[r], dpy);
                vy1[q] = M_::upd(L.vy0[r][q], F.fv, F.fvf, (mvy >> q) & 1u, dpy);
                A dpz = M_::diff(p1[r][q], p1prev[r][q], F.dz, F.rdz);
                if constexpr (SZ) dpz = M_::psi(L.mVz[r][q], L.bzV, L.azV, dpz);
                vz1[q] = M_::upd(vzc[r][q], F.fv, F.fvf, (mvz >> q) & 1u, dpz);
            }
            bst<T, CPL, POL>(rF1, vs, rb, p1[r]);
            bst<T, CPL, POL>(rF1, vs, chs(1, rb), vx1);
            bst<T, CPL, POL>(rF1, vs, chs(2, rb), vy1);
            bst<T, CPL, POL>(rF1, vs, chs(3, rb), vz1);
            if constexpr (RAG) {
#pragma unroll
                for (int q = 0; q < CPL - 1; ++q) {
                    const T o[4][1] = {{p1[r][q]}, {vx1[q]}, {vy1[q]}, {vz1[q]}};
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        bst<T, 1, POL>(rF1, vq[q], (int)((uint32_t)rb + (uint32_t)c * chb), o[c]);
                }
            }
            // x-slab memory: 64-byte pieces of a row, written with ordinary
            // stores whatever POL is (with the non-temporal hint they cost
            // 5 % of the 256^3 sweep: each piece then goes to HBM alone)
            const int soP = p * (j + ny * k) * ES, soV = qv * (j + ny * k) * ES;
            bst<T, CPL>(rS1, sPv, soP, L.mPx[r]);
            bst<T, CPL>(rS1, sVv, soV, L.mVx[r]);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const T a1[1] = {L.mPx[r][q]}, b1[1] = {L.mVx[r][q]};
                bst<T, 1>(rS1, sPq[q], soP, a1);
                bst<T, 1>(rS1, sVq[q], soV, b1);
            }
            // y/z slab memory (row vectors of the slab arrays)
            auto sslab = [&](Rsrc rs, int soff, const T (&o)[CPL]) {
                bst<T, CPL, POL>(rs, vs, soff, o);
                if constexpr (RAG) {
#pragma unroll
                    for (int q = 0; q < CPL - 1; ++q) {
                        const T one[1] = {o[q]};
                        bst<T, 1, POL>(rs, vq[q], soff, one);
                    }
                }
            };
            if constexpr (SY) {
                sslab(rsrc_of(F.psi1, yinP[r] ? psib : 0u), yoP[r] + k * ypP, L.mPy[r]);
                sslab(rsrc_of(F.psi1, yinV[r] ? psib : 0u), yoV[r] + k * ypV, L.mVy[r]);
            }
            if constexpr (SZ) {
                sslab(rsrc_of(F.psi1, L.zinP ? psib : 0u), L.zoP + r * nx * ES, L.mPz[r]);
                sslab(rsrc_of(F.psi1, L.zinV ? psib : 0u), L.zoV + r * nx * ES, L.mVz[r]);
            }
        }
        carry(L, p1);
    };

    Plane B0, B1;
    load(B0, k0 - 1);  // halo plane: pressure only
    load(B1, k0);
    {
        T p1[TY + 1][CPL];
        pressure(B0, p1);
        carry(B0, p1);
    }
    // the scheduling barriers keep each prefetch ahead of the computation
    // it overlaps (the scheduler would otherwise sink the loads to shorten
    // register lifetimes)
    // No prefetch after the last plane, except where registers are at
    // the limit (the branch costs a few): there the last prefetch repeats
    // the last plane.
    constexpr bool SKIP_TAIL = ES == 4 || CPL < 4 || NOMASK;
    int k = k0;
    for (;;) {
        if (!SKIP_TAIL || k + 1 < k1) load(B0, min(k + 1, k1 - 1));
        __builtin_amdgcn_sched_barrier(0);
        step(B1, k);
        if (++k >= k1) break;
        if (!SKIP_TAIL || k + 1 < k1) load(B1, min(k + 1, k1 - 1));
        __builtin_amdgcn_sched_barrier(0);
        step(B0, k);
        if (++k >= k1) break;
    }
}

// Tile kinds: interior tiles (regular, outside every y/z slab; sweep_ty_int
// rows), regular tiles in the y and/or z CPML slabs, and general tiles
// (boundary rows and planes, clamped rows); the last three are side tiles
// (sweep_ty_side(CPL) rows).  Entries carry their kind in bits 28-30 of the x
// word.  Lists: interior tiles, side tiles, or all tiles (side tiles first
// within each XCD's chunk).
enum { TILE_INT = 0, TILE_SY = 1, TILE_SZ = 2, TILE_SYZ = 3, TILE_GEN = 4 };
enum { LIST_INT = 0, LIST_SIDE = 1, LIST_ALL = 2, TILE_LISTS = 3 };

// One tile of a solve.  PF: the pipelined routine for every regular tile;
// otherwise the plane-at-a-time routines (fused_tile for all side tiles).
// General tiles always take fused_tile.
// RAG: the solve's rows are not a multiple of CPL long (fused_tile).
// INTONLY: the tile is an unmasked interior tile (the interior list holds
// only those): a kernel instantiated with INTONLY compiles the one routine
// they take, so its register count, and with it the waves per SIMD, is
// that routine's alone.
template <typename T, bool EXACT, int CPL, bool PF, int POL = 0, bool RAG = false, bool INTONLY = false>
__device__ __forceinline__ void sweep_tile(const Fused<T>& F, uint2 e) {
    const int j0 = (int)(e.x & 0xffffu), k0 = (int)(e.y >> 16), k1 = k0 + (int)(e.y & 0x3ffu);
    const int seg = (int)((e.y >> 10) & 0x3fu);
    const bool masked = (e.x >> 31) != 0;
    const int kind = (int)((e.x >> 28) & 7u);
    constexpr int TI = sweep_ty_int((int)sizeof(T), CPL), TS = sweep_ty_side(CPL);
    if constexpr (INTONLY) {
        sweep_tile_pf<T, EXACT, CPL, TI, false, false, POL, true, RAG>(F, j0, k0, k1, false, seg);
        return;
    }
    if (kind == TILE_INT) {
        if constexpr (PF) {
            if (masked) sweep_tile_pf<T, EXACT, CPL, TI, false, false, POL, false, RAG>(F, j0, k0, k1, true, seg);
            else sweep_tile_pf<T, EXACT, CPL, TI, false, false, POL, true, RAG>(F, j0, k0, k1, false, seg);
        } else {
            sweep_int_tile<T, EXACT, CPL, TI, RAG>(F, j0, k0, k1, masked, seg);
        }
    } else if (PF && kind == TILE_SY) {
        sweep_tile_pf<T, EXACT, CPL, TS, true, false, POL, false, RAG>(F, j0, k0, k1, masked, seg);
    } else if (PF && kind == TILE_SZ) {
        sweep_tile_pf<T, EXACT, CPL, TS, false, true, POL, false, RAG>(F, j0, k0, k1, masked, seg);
    } else if (PF && kind == TILE_SYZ) {
        sweep_tile_pf<T, EXACT, CPL, TS, true, true, POL, false, RAG>(F, j0, k0, k1, masked, seg);
    } else {
        fused_tile<T, EXACT, CPL, TS, true, POL, RAG>(F, j0, k0, k1, masked, seg);
    }
}

// One solve's tiles of one list (entries as xcd_tile describes); W waves
// per workgroup take W consecutive tiles.  POL: cache policy of the
// stores (0 or ST_NT).  With S.nbt > 0 (one wave per workgroup) the step's
// transceiver work rides in the launch (TxSweep): S.nrb record blocks
// first, then S.nbt blocks of one tile each, then the chain blocks.
template <typename T, bool EXACT, int CPL, bool PF, int POL, bool RAG, bool INTONLY = false>
__global__ __launch_bounds__(256) void k_sweep(Fused<T> F, const uint2* __restrict__ tiles, int ntiles,
                                               TxSweep<T> S) {
    if (S.nbt > 0) {
        const int b = (int)blockIdx.x - S.nrb;
        if (b < 0) {
            tx_sweep_records<T>(F, S, b + S.nrb);
            return;
        }
        if (b >= S.nbt) {
            tx_sweep_chain<T, EXACT>(F, S, b - S.nbt);
            return;
        }
        const int t = xcd_tile(b, S.nbt);
        const uint2 e = tiles[t];
        sweep_tile<T, EXACT, CPL, PF, POL, RAG, INTONLY>(F, e);
        if (e.x & TILE_SIGNAL) tx_tile_signal<T>(S, t);
        return;
    }
    const int W = (int)(blockDim.x >> 6);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int t = xcd_tile((int)blockIdx.x, (int)gridDim.x) * W + w;
    if (t >= ntiles) return;
    sweep_tile<T, EXACT, CPL, PF, POL, RAG, INTONLY>(F, tiles[t]);
}

}  // namespace adf
