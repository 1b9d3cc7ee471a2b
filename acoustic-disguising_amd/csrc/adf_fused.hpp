// Fused single-pass step for gfx950: the pressure update and the three
// velocity updates of one time step in one z-marching sweep.
//
// Buffers are ping-ponged: the sweep reads the old field / CPML memory (0)
// and writes the new ones (1), so a wave may recompute the pressure of a
// neighbouring tile's edge cells (its halo) from untouched old values.
//
// One wave64 owns an x-row segment (lane l holds CPL consecutive cells as
// one vector load/store; see xlane), TY rows in y held in registers, and
// marches KB planes in z.  Per plane it loads every row of the
// tile plus the halo row below (j0-1), computes P1 on all of them, takes the
// x-neighbours Vx0(i+1) and P1(i-1) from the adjacent lane with wave
// shuffles, and writes P1, Vx1, Vy1, Vz1 of its own rows.  P1 of the previous
// plane stays in registers for Vz; a tile's first plane gets it from a
// recomputed halo plane (k0-1).  No LDS, no barriers: waves are independent.
//
// Cells the reference never updates (P at i = nx, the (ny,nz) edge, Vx at
// i = 1, Vy at j = 1 / Vz at k = 1 except through the mirrors;
// forward.jl:67-106,134-137) are copied or mirrored so the new buffer is
// complete.  Velocity cells next to injected pressure cells are recomputed
// afterwards by k_fixup_v from the old buffers and the injected P1, which
// keeps the order of the reference step (P update, P injection, V update).
#pragma once
#include <type_traits>

#include "adf_kernels.hpp"

namespace adf {

// Rows per wave of the fused sweep's tiles: interior tiles (SWEEP_TY, also
// the row granularity of the tile split) and side tiles (CPML-slab and
// boundary tiles), whose routines need more registers per row: one row
// at 4 cells per lane, where two rows would not fit the register budget of
// the interior routine, else two.
constexpr int SWEEP_TY = 2;
__host__ __device__ constexpr int sweep_ty_side(int cpl) { return cpl >= 4 ? 1 : 2; }
// Interior rows per tile for es-byte words at cpl cells per lane: SWEEP_TY,
// except Float64 at 4 cells per lane, where a 2-row tile of the pipelined
// routine needs more than a wave's 512 registers and spilled to scratch.
__host__ __device__ constexpr int sweep_ty_int(int es, int cpl) { return es == 8 && cpl >= 4 ? 1 : SWEEP_TY; }

// Kernel arguments of one solve, kept small so the fused kernel does not
// run out of scalar registers: one base pointer per buffer, offsets derived
// on the device.  f0/f1: old/new field (P,Vx,Vy,Vz at c*n3); psi0/psi1:
// old/new CPML memory, the six arrays mv_x,mv_y,mv_z (P phase) then
// mp_x,mp_y,mp_z (V phase) at 256-byte aligned offsets; coef: a,b (2p each)
// then ah,bh (2(p+1) each).
template <typename T>
struct Fused {
    const T* f0;
    T* f1;
    const T* psi0;
    T* psi1;
    const T* coef;
    const uint32_t* mask;
    int nx, ny, nz, p, nwx, n3;
    int psi_n;  // elements in psi0 / psi1
    int cpml_p, fuse_mirror;
    int ty;                        // rows per wave
    int xseg;                      // owned cells per row segment (xlane)
    int off_mv[3], off_mp[3];      // element offsets of the CPML arrays in psi0/psi1
    double dx, dy, dz, fq, fv;
    float rdx, rdy, rdz, fqf, fvf;

    __device__ __forceinline__ const T* P0() const { return f0; }
    __device__ __forceinline__ const T* V0(int a) const { return f0 + (size_t)(a + 1) * n3; }
    __device__ __forceinline__ T* P1() const { return f1; }
    __device__ __forceinline__ T* V1(int a) const { return f1 + (size_t)(a + 1) * n3; }
    // CPML memory offsets (each array in the reference layout, 256-byte aligned)
    __device__ __forceinline__ int mv_off(int a) const { return off_mv[a]; }
    __device__ __forceinline__ int mp_off(int a) const { return off_mp[a]; }
    __device__ __forceinline__ T ca(int i) const { return coef[i]; }
    __device__ __forceinline__ T cb(int i) const { return coef[2 * p + i]; }
    __device__ __forceinline__ T cah(int i) const { return coef[4 * p + i]; }
    __device__ __forceinline__ T cbh(int i) const { return coef[4 * p + 2 * (p + 1) + i]; }
};

template <typename T, int N>
__device__ __forceinline__ void ldv(const T* p, T (&o)[N]) {
    if constexpr (N == 1) {
        o[0] = *p;
    } else {
        typedef T V __attribute__((ext_vector_type(N)));
        const V v = *reinterpret_cast<const V*>(p);
#pragma unroll
        for (int q = 0; q < N; ++q) o[q] = v[q];
    }
}

template <typename T, int N>
__device__ __forceinline__ void stv(T* p, const T (&o)[N]) {
    if constexpr (N == 1) {
        *p = o[0];
    } else {
        typedef T V __attribute__((ext_vector_type(N)));
        V v;
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] = o[q];
        *reinterpret_cast<V*>(p) = v;
    }
}

// Buffer-resource access for the row loads and stores of the sweep: the
// descriptor (base, size) sits in scalar registers, the lane's byte offset
// in one vector register and the row's byte offset in a scalar register, so
// a row access costs no 64-bit vector address.  Word 3 is the gfx9 default
// (32-bit data format, no swizzle); accesses past `bytes` read zero.
typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ Rsrc rsrc_of(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}

template <typename T, int N>
__device__ __forceinline__ void bld(Rsrc r, int voff, int soff, T (&o)[N]) {
    constexpr int B = (int)sizeof(T) * N;
    if constexpr (B == 4) {
        const uint32_t u = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
        __builtin_memcpy(&o[0], &u, 4);
    } else if constexpr (B == 8) {
        const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        __builtin_memcpy(&o[0], &u, 8);
    } else if constexpr (B == 16) {
        const u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
        __builtin_memcpy(&o[0], &u, 16);
    } else {
        static_assert(B == 32, "row chunk of 4, 8, 16 or 32 bytes");
        const u32x4 u0 = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
        const u32x4 u1 = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16, soff, 0);
        __builtin_memcpy(&o[0], &u0, 16);
        __builtin_memcpy(&o[N / 2], &u1, 16);
    }
}

// A 128-bit buffer store followed by eight wait states during which its
// data registers are held (the asm reads them, so nothing can overwrite
// them before it).  Without this, 256^3 float32 sweeps stored, in some
// lanes, a value that a later VALU or LDS instruction wrote into the first
// data register of such a store (results varying from run to run); the
// compiler inserts no wait state after stores with an SGPR offset.  The
// same hazard in compiler-written spill code is why no kernel here may use
// scratch memory (the Makefile checks it).
template <int POL>
__device__ __forceinline__ void store_b128(u32x4 u, Rsrc r, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b128(u, r, voff, soff, POL);
    asm volatile("s_nop 7" ::"v"(u));
}

// Cache policy of a store: 0 ordinary, ST_NT the non-temporal hint (the
// instruction's nt bit).
constexpr int ST_NT = 2;

template <typename T, int N, int POL = 0>
__device__ __forceinline__ void bst(Rsrc r, int voff, int soff, const T (&o)[N]) {
    constexpr int B = (int)sizeof(T) * N;
    if constexpr (B == 4) {
        uint32_t u;
        __builtin_memcpy(&u, &o[0], 4);
        __builtin_amdgcn_raw_buffer_store_b32(u, r, voff, soff, POL);
    } else if constexpr (B == 8) {
        u32x2 u;
        __builtin_memcpy(&u, &o[0], 8);
        __builtin_amdgcn_raw_buffer_store_b64(u, r, voff, soff, POL);
    } else if constexpr (B == 16) {
        u32x4 u;
        __builtin_memcpy(&u, &o[0], 16);
        store_b128<POL>(u, r, voff, soff);
    } else {
        static_assert(B == 32, "row chunk of 4, 8, 16 or 32 bytes");
        u32x4 u0, u1;
        __builtin_memcpy(&u0, &o[0], 16);
        __builtin_memcpy(&u1, &o[N / 2], 16);
        store_b128<POL>(u0, r, voff, soff);
        store_b128<POL>(u1, r, voff + 16, soff);
    }
}

// x placement of a lane.  A row of up to 64 CPL cells is one segment,
// lane l holding cells l CPL..; a wider row is cut into segments of xseg =
// 62 CPL owned cells, one wave each, whose lane 0 (after the first
// segment) and lane 63 (before the last) hold the neighbouring segments'
// edge cells: the wave computes their new pressure from the old buffers
// like the owner does, so the x-neighbours Vx0(i+1) and P1(i-1) of its own
// edge cells come from adjacent lanes as inside the row.  With nx not a
// multiple of CPL the last lane of the row runs past its end (`part`): its
// loads read cells of the next row (never used) and it stores cell by cell.
struct XLane {
    int i0;     // first cell of the lane
    int i0c;    // first cell of its loads: i0 clamped into the row (a multiple of CPL)
    bool own;   // the lane's cells belong to this segment (it stores them)
    bool part;  // an owned lane whose cells run past the row end
};
template <int CPL>
__device__ __forceinline__ XLane xlane(int nx, int xseg, int seg) {
    const int lane = threadIdx.x & 63;
    const int xs = seg * xseg;
    XLane X;
    X.i0 = (seg > 0 ? xs - CPL : 0) + lane * CPL;
    X.own = X.i0 >= xs && X.i0 < min(xs + xseg, nx);
    X.part = X.own && X.i0 + CPL > nx;
    X.i0c = min(X.i0, (nx - 1) / CPL * CPL);
    return X;
}

__device__ __forceinline__ uint32_t bld_u32(Rsrc r, int voff, int soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}

template <typename T>
__device__ __forceinline__ T shfl_down1(T v) { return __shfl_down(v, 1, 64); }
template <typename T>
__device__ __forceinline__ T shfl_up1(T v) { return __shfl_up(v, 1, 64); }

// New pressure at (i,j,k) from old values (updates.jl:21-142,288-378).
// vxn = Vx0(i+1), vyn = Vy0(j+1), vzn = Vz0(k+1) (ignored where unused).
template <typename T, bool EXACT>
__device__ __forceinline__ T p_new(const Fused<T>& F, int i, int j, int k, T p0, T vx0, T vxn, T vy0, T vyn,
                                   T vz0, T vzn, bool fluid, bool own) {
    const int nx = F.nx, ny = F.ny, nz = F.nz;
    if (i >= nx - 1) return p0;
    const bool rj = (j == ny - 1), rk = (k == nz - 1);
    if (rj && rk) return p0;
    if (rj || rk) {
        if constexpr (EXACT) {
            const double dvx = (double)(T)(vxn - vx0) / F.dx;
            const double dvy = rj ? (double)((T)(-2) * vy0) / F.dy : (double)(T)(vyn - vy0) / F.dy;
            const double dvz = rk ? (double)((T)(-2) * vz0) / F.dz : (double)(T)(vzn - vz0) / F.dz;
            return (T)((double)p0 - F.fq * ((dvx + dvy) + dvz));
        } else {
            const float dvx = (float)(vxn - vx0) * F.rdx;
            const float dvy = rj ? (-2.0f * (float)vy0) * F.rdy : (float)(vyn - vy0) * F.rdy;
            const float dvz = rk ? (-2.0f * (float)vz0) * F.rdz : (float)(vzn - vz0) * F.rdz;
            return (T)fmaf(-F.fqf, (dvx + dvy) + dvz, (float)p0);
        }
    }
    const int p = F.p;
    if constexpr (EXACT) {
        double dvx = (double)(T)(vxn - vx0) / F.dx;
        double dvy = (double)(T)(vyn - vy0) / F.dy;
        double dvz = (double)(T)(vzn - vz0) / F.dz;
        if (F.cpml_p) {
            int l, sd;
            if ((sd = slab(i, nx, p, l))) {
                const int mi = F.mv_off(0) + l + p * (j + ny * (k + nz * (sd - 1)));
                T m = F.psi0[mi];
                dvx = psi_exact(&m, F.cb(l + p * (sd - 1)), F.ca(l + p * (sd - 1)), dvx);
                if (own) F.psi1[mi] = m;
            }
            if ((sd = slab(j, ny, p, l))) {
                const int mi = F.mv_off(1) + i + nx * (l + p * (k + nz * (sd - 1)));
                T m = F.psi0[mi];
                dvy = psi_exact(&m, F.cb(l + p * (sd - 1)), F.ca(l + p * (sd - 1)), dvy);
                if (own) F.psi1[mi] = m;
            }
            if ((sd = slab(k, nz, p, l))) {
                const int mi = F.mv_off(2) + i + nx * (j + ny * (l + p * (sd - 1)));
                T m = F.psi0[mi];
                dvz = psi_exact(&m, F.cb(l + p * (sd - 1)), F.ca(l + p * (sd - 1)), dvz);
                if (own) F.psi1[mi] = m;
            }
        }
        const double coeff = fluid ? F.fq : 0.0;
        return (T)((double)p0 - coeff * ((dvx + dvy) + dvz));
    } else {
        float dvx = (float)(vxn - vx0) * F.rdx;
        float dvy = (float)(vyn - vy0) * F.rdy;
        float dvz = (float)(vzn - vz0) * F.rdz;
        if (F.cpml_p) {
            int l, sd;
            if ((sd = slab(i, nx, p, l))) {
                const int mi = F.mv_off(0) + l + p * (j + ny * (k + nz * (sd - 1)));
                float m = (float)F.psi0[mi];
                dvx = psi_fast(&m, F.cb(l + p * (sd - 1)), F.ca(l + p * (sd - 1)), dvx);
                if (own) F.psi1[mi] = (T)m;
            }
            if ((sd = slab(j, ny, p, l))) {
                const int mi = F.mv_off(1) + i + nx * (l + p * (k + nz * (sd - 1)));
                float m = (float)F.psi0[mi];
                dvy = psi_fast(&m, F.cb(l + p * (sd - 1)), F.ca(l + p * (sd - 1)), dvy);
                if (own) F.psi1[mi] = (T)m;
            }
            if ((sd = slab(k, nz, p, l))) {
                const int mi = F.mv_off(2) + i + nx * (j + ny * (l + p * (sd - 1)));
                float m = (float)F.psi0[mi];
                dvz = psi_fast(&m, F.cb(l + p * (sd - 1)), F.ca(l + p * (sd - 1)), dvz);
                if (own) F.psi1[mi] = (T)m;
            }
        }
        return fluid ? (T)fmaf(-F.fqf, (dvx + dvy) + dvz, (float)p0) : p0;
    }
}

// New velocity component AX at (i,j,k) from V0, the new pressure at the cell
// (pc) and at its minus neighbour along AX (pm) (updates.jl:192-275,380-501).
template <typename T, bool EXACT, int AX>
__device__ __forceinline__ T v_new(const Fused<T>& F, int i, int j, int k, T v0, T pc, T pm, bool open) {
    const int nx = F.nx, ny = F.ny, nz = F.nz;
    const int q = F.p + 1;
    const T* mp0 = F.psi0 + F.mp_off(AX);
    T* mp1 = F.psi1 + F.mp_off(AX);
    const int idx = AX == 0 ? i : (AX == 1 ? j : k);
    const int n = AX == 0 ? nx : (AX == 1 ? ny : nz);
    if constexpr (EXACT) {
        const double h = AX == 0 ? F.dx : (AX == 1 ? F.dy : F.dz);
        double dp = (double)(T)(pc - pm) / h;
        if (q > 1) {
            int l, sd;
            if ((sd = slab(idx, n, q, l))) {
                const int mi = AX == 0 ? l + q * (j + ny * (k + nz * (sd - 1)))
                             : (AX == 1 ? i + nx * (l + q * (k + nz * (sd - 1)))
                                        : i + nx * (j + ny * (l + q * (sd - 1))));
                T m = mp0[mi];
                dp = psi_exact(&m, F.cbh(l + q * (sd - 1)), F.cah(l + q * (sd - 1)), dp);
                mp1[mi] = m;
            }
        }
        const double coeff = open ? F.fv : 0.0;
        return (T)((double)v0 - coeff * dp);
    } else {
        const float rh = AX == 0 ? F.rdx : (AX == 1 ? F.rdy : F.rdz);
        float dp = (float)(pc - pm) * rh;
        if (q > 1) {
            int l, sd;
            if ((sd = slab(idx, n, q, l))) {
                const int mi = AX == 0 ? l + q * (j + ny * (k + nz * (sd - 1)))
                             : (AX == 1 ? i + nx * (l + q * (k + nz * (sd - 1)))
                                        : i + nx * (j + ny * (l + q * (sd - 1))));
                float m = (float)mp0[mi];
                dp = psi_fast(&m, F.cbh(l + q * (sd - 1)), F.cah(l + q * (sd - 1)), dp);
                mp1[mi] = (T)m;
            }
        }
        return open ? (T)fmaf(-F.fvf, dp, (float)v0) : v0;
    }
}

__device__ __forceinline__ uint32_t mask_word(const uint32_t* m, int plane, int nwx, int ny, int nz, int i,
                                              int j, int k) {
    return m[((size_t)plane * nz * ny + (size_t)k * ny + j) * nwx + (i >> 5)];
}

// Arithmetic of the fused sweep.  EXACT: the reference's (T differences,
// double scalars, true divisions); otherwise float with reciprocals.
template <typename T, bool EXACT>
struct FA {
    typedef typename std::conditional<EXACT, double, float>::type A;
    static __device__ __forceinline__ A diff(T hi, T lo, double h, float rh) {
        if constexpr (EXACT) return (double)(T)(hi - lo) / h;
        else return (float)(hi - lo) * rh;
    }
    static __device__ __forceinline__ A rigid(T v, double h, float rh) {
        if constexpr (EXACT) return (double)((T)(-2) * v) / h;
        else return (-2.0f * (float)v) * rh;
    }
    // psi = b*psi + a*dv (updated in place); returns dv + psi
    static __device__ __forceinline__ A psi(T& m, T bc, T ac, A dv) {
        if constexpr (EXACT) {
            const T bm = bc * m;
            m = (T)((double)bm + (double)ac * dv);
            return dv + (double)m;
        } else {
            m = (T)fmaf((float)bc, (float)m, (float)ac * dv);
            return dv + (float)m;
        }
    }
    // The same where a cell with `keep` set keeps its memory (a select, so
    // the memory comes back bit for bit, -0.0 included)
    static __device__ __forceinline__ A psi_keep(T& m, T bc, T ac, A dv, bool keep) {
        const T m0 = m;
        const A d = psi(m, bc, ac, dv);
        if (keep) m = m0;
        return d;
    }
    // x1 = x0 - f*s (f = 0 where the mask closes the cell)
    static __device__ __forceinline__ T upd(T x0, double f, float ff, bool open, A s) {
        if constexpr (EXACT) return (T)((double)x0 - (open ? f : 0.0) * s);
        else return open ? (T)fmaf(-ff, s, (float)x0) : x0;
    }
};

// One (y-tile, z-block) of one solve, executed by one wave64.  Every load of
// a plane (field rows, mask words, CPML memory of the slabs the cells lie in)
// is issued before any arithmetic, and CPML memory is written back after it,
// so a plane costs one memory round trip.
//
// GEN = false is the interior variant: the caller guarantees that every row
// j0-1..j0+TY-1 and every plane k0-1..k1-1 of the tile lies outside the y/z
// CPML slabs of both phases, away from the mirrored (1,2) and rigid (n)
// planes, so only the x direction keeps special cases.  Loads use clamped
// indices (always in bounds) instead of guards; values of rows, lanes and
// planes outside the grid are computed but never stored.
// RAG: rows whose length is not a multiple of CPL (the partial lane stores
// cell by cell); seg: the row segment (xlane).
template <typename T, bool EXACT, int CPL, int TY, bool GEN, int POL = 0, bool RAG = false>
__device__ __forceinline__ void fused_tile(const Fused<T>& F, int j0, int k0, int k1, bool masked, int seg) {
    typedef FA<T, EXACT> M_;
    typedef typename M_::A A;
    const int nx = F.nx, ny = F.ny, nz = F.nz, p = F.p, qv = p + 1;
    const int sxy = nx * ny;
    const XLane XL = xlane<CPL>(nx, F.xseg, seg);
    const int i0 = XL.i0;
    const bool lane_in = i0 < nx;  // the lane has cells in the row (loads of CPML memory)
    const int i0c = XL.i0c;        // clamped lane start (field and mask loads)
    const int ks = k0 > 0 ? k0 - 1 : k0;  // first plane (halo plane when k0 > 0)
    const uint32_t* Mk = masked ? F.mask : nullptr;  // tiles without closed cells skip the mask
    const int sh = i0c & 31;
    const bool cpP = F.cpml_p != 0, cpV = qv > 1;
    const T* __restrict__ psi0 = F.psi0;
    T* __restrict__ psi1 = F.psi1;
    auto crow = [&](int j) { return j < 0 ? 0 : (j >= ny ? ny - 1 : j); };
    constexpr int ES = (int)sizeof(T);
    // one descriptor per field buffer (the host keeps 4 n3 ES < 4 GiB for
    // this kernel); channel c starts at byte c * chb
    const uint32_t chb = (uint32_t)F.n3 * (uint32_t)ES;
    const Rsrc rF0 = rsrc_of(F.f0, 4u * chb), rF1 = rsrc_of(F.f1, 4u * chb);
    const Rsrc rS0 = rsrc_of(psi0, (uint32_t)F.psi_n * ES), rS1 = rsrc_of(psi1, (uint32_t)F.psi_n * ES);
    const Rsrc rM = rsrc_of(Mk, Mk ? (uint32_t)(4 * F.nwx * ny * nz) * 4u : 0u);
    const int vl = i0c * ES;         // lane byte offset of loads (clamped lane)
    const int vs = i0 * ES;          // lane byte offset of stores (owned lanes only)
    const int vm = (i0c >> 5) * 4;   // lane byte offset of its mask word
    // byte offset of row j (already clamped) of plane k in a field channel
    // (unsigned: a Float64 channel may exceed 2 GiB)
    auto rowb = [&](int j, int k) { return (int)((uint32_t)(nx * j + sxy * k) * (uint32_t)ES); };
    auto chrow = [&](int c, int rb) { return (int)((uint32_t)c * chb + (uint32_t)rb); };
    auto maskb = [&](int plane, int j, int k) { return (((plane * nz + k) * ny + j) * F.nwx) * 4; };
    // a row store of the lane's cells (cell by cell for the partial lane)
    auto srow = [&](Rsrc rs, int soff, const T (&o)[CPL]) {
        if (RAG && XL.part) {
#pragma unroll
            for (int q = 0; q < CPL; ++q)
                if (i0 + q < nx) {
                    const T one[1] = {o[q]};
                    bst<T, 1, POL>(rs, vs + q * ES, soff, one);
                }
        } else {
            bst<T, CPL, POL>(rs, vs, soff, o);
        }
    };

    // per-lane x-slab data (fixed for the whole sweep)
    int sxP[CPL], lxP[CPL], sxV[CPL], lxV[CPL];
    T axP[CPL], bxP[CPL], axV[CPL], bxV[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        const int i = i0 + q;
        lxP[q] = lxV[q] = 0;
        sxP[q] = (cpP && i < nx - 1) ? slab(i, nx, p, lxP[q]) : 0;
        sxV[q] = (cpV && i >= 1 && i < nx) ? slab(i, nx, qv, lxV[q]) : 0;
        axP[q] = sxP[q] ? F.ca(lxP[q] + p * (sxP[q] - 1)) : (T)0;
        bxP[q] = sxP[q] ? F.cb(lxP[q] + p * (sxP[q] - 1)) : (T)0;
        axV[q] = sxV[q] ? F.cah(lxV[q] + qv * (sxV[q] - 1)) : (T)0;
        bxV[q] = sxV[q] ? F.cbh(lxV[q] + qv * (sxV[q] - 1)) : (T)0;
    }

    // rows r = -1..TY-1 are stored at index r+1
    T vzc[TY + 1][CPL];    // Vz0 at the current plane
    T p1prev[TY + 1][CPL]; // P1 at the previous plane
#pragma unroll
    for (int r = 0; r <= TY; ++r) {
        bld<T, CPL>(rF0, vl, chrow(3, rowb(crow(j0 + r - 1), ks)), vzc[r]);
#pragma unroll
        for (int q = 0; q < CPL; ++q) p1prev[r][q] = (T)0;
    }

    for (int k = ks; k < k1; ++k) {
        const bool owned = k >= k0;
        const int kc = sxy * k;
        const bool rk = GEN && (k == nz - 1);
        // z-slab of this plane (uniform; never for interior tiles)
        int lzP = 0, lzV = 0, szP = 0, szV = 0;
        T azP = (T)0, bzP = (T)0, azV = (T)0, bzV = (T)0;
        if constexpr (GEN) {
            szP = (cpP && !rk) ? slab(k, nz, p, lzP) : 0;
            szV = (cpV && k >= 1 && owned) ? slab(k, nz, qv, lzV) : 0;
            if (szP) { azP = F.ca(lzP + p * (szP - 1)); bzP = F.cb(lzP + p * (szP - 1)); }
            if (szV) { azV = F.cah(lzV + qv * (szV - 1)); bzV = F.cbh(lzV + qv * (szV - 1)); }
        }

        T p0[TY + 1][CPL], vx0[TY + 1][CPL], vy0[TY + 2][CPL], vzn[TY + 1][CPL], p1[TY + 1][CPL];
        T mPx[TY + 1][CPL], mVx[TY + 1][CPL];
        T mPy[GEN ? TY + 1 : 1][CPL], mPz[GEN ? TY + 1 : 1][CPL];
        T mVy[GEN ? TY + 1 : 1][CPL], mVz[GEN ? TY + 1 : 1][CPL];
        uint32_t mp_[TY + 1], mvx_[TY + 1], mvy_[TY + 1], mvz_[TY + 1];
        int syP[TY + 1], lyP[TY + 1], syV[TY + 1], lyV[TY + 1];
        const int kn = (k + 1 < nz) ? k + 1 : k;

        // ---- all loads of this plane
#pragma unroll
        for (int r = 0; r <= TY + 1; ++r) {
            const int j = j0 + r - 1;
            const int jc = crow(j);
            const int rb = rowb(jc, k);
            bld<T, CPL>(rF0, vl, chrow(2, rb), vy0[r]);
            if (r > TY) continue;
            bld<T, CPL>(rF0, vl, rb, p0[r]);
            bld<T, CPL>(rF0, vl, chrow(1, rb), vx0[r]);
            bld<T, CPL>(rF0, vl, chrow(3, rowb(jc, kn)), vzn[r]);
            mp_[r] = mvx_[r] = mvy_[r] = mvz_[r] = 0xffffffffu;
            if (Mk) {
                mp_[r] = bld_u32(rM, vm, maskb(0, jc, k)) >> sh;
                if (owned && r > 0) {
                    mvx_[r] = bld_u32(rM, vm, maskb(1, jc, k)) >> sh;
                    mvy_[r] = bld_u32(rM, vm, maskb(2, jc, k)) >> sh;
                    mvz_[r] = bld_u32(rM, vm, maskb(3, jc, k)) >> sh;
                }
            }
            const bool rowv = j >= 0 && j < ny;
            const bool rj = GEN && (j == ny - 1);
            const bool ok = lane_in && rowv;
            lyP[r] = lyV[r] = 0;
            syP[r] = syV[r] = 0;
            if constexpr (GEN) {
                syP[r] = (cpP && rowv && !rj && !rk) ? slab(j, ny, p, lyP[r]) : 0;
                syV[r] = (cpV && rowv && j >= 1 && owned && r > 0) ? slab(j, ny, qv, lyV[r]) : 0;
            }
            if (cpP && ok && !rj && !rk) {
#pragma unroll
                for (int q = 0; q < CPL; ++q)
                    if (sxP[q]) mPx[r][q] = psi0[F.mv_off(0) + lxP[q] + p * (j + ny * (k + nz * (sxP[q] - 1)))];
            }
            if (cpV && ok && owned && r > 0) {
#pragma unroll
                for (int q = 0; q < CPL; ++q)
                    if (sxV[q]) mVx[r][q] = psi0[F.mp_off(0) + lxV[q] + qv * (j + ny * (k + nz * (sxV[q] - 1)))];
            }
            if constexpr (GEN) {
                if (ok && syP[r]) bld<T, CPL>(rS0, vs, (F.mv_off(1) + nx * (lyP[r] + p * (k + nz * (syP[r] - 1)))) * ES, mPy[r]);
                if (ok && szP && !rj) bld<T, CPL>(rS0, vs, (F.mv_off(2) + nx * (j + ny * (lzP + p * (szP - 1)))) * ES, mPz[r]);
                if (ok && syV[r]) bld<T, CPL>(rS0, vs, (F.mp_off(1) + nx * (lyV[r] + qv * (k + nz * (syV[r] - 1)))) * ES, mVy[r]);
                if (ok && owned && r > 0 && szV) bld<T, CPL>(rS0, vs, (F.mp_off(2) + nx * (j + ny * (lzV + qv * (szV - 1)))) * ES, mVz[r]);
            }
        }

        // ---- new pressure on rows -1..TY-1
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
            const int j = j0 + r - 1;
            const T vxl = shfl_down1<T>(vx0[r][0]);  // Vx0 of the next lane's first cell
            const bool rj = GEN && (j == ny - 1);
            T ayP = (T)0, byP = (T)0;
            if constexpr (GEN) {
                if (syP[r]) { ayP = F.ca(lyP[r] + p * (syP[r] - 1)); byP = F.cb(lyP[r] + p * (syP[r] - 1)); }
            }
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int i = i0 + q;
                const T vxn = q + 1 < CPL ? vx0[r][q + 1] : vxl;
                T out = p0[r][q];
                if (i < nx - 1 && !(rj && rk)) {
                    if (GEN && (rj || rk)) {
                        // rigid planes: -2V/d at j==ny / k==nz, no CPML, no mask
                        const A dvx = M_::diff(vxn, vx0[r][q], F.dx, F.rdx);
                        const A dvy = rj ? M_::rigid(vy0[r][q], F.dy, F.rdy) : M_::diff(vy0[r + 1][q], vy0[r][q], F.dy, F.rdy);
                        const A dvz = rk ? M_::rigid(vzc[r][q], F.dz, F.rdz) : M_::diff(vzn[r][q], vzc[r][q], F.dz, F.rdz);
                        out = M_::upd(p0[r][q], F.fq, F.fqf, true, (dvx + dvy) + dvz);
                    } else {
                        A dvx = M_::diff(vxn, vx0[r][q], F.dx, F.rdx);
                        A dvy = M_::diff(vy0[r + 1][q], vy0[r][q], F.dy, F.rdy);
                        A dvz = M_::diff(vzn[r][q], vzc[r][q], F.dz, F.rdz);
                        if (sxP[q]) dvx = M_::psi(mPx[r][q], bxP[q], axP[q], dvx);
                        if constexpr (GEN) {
                            if (syP[r]) dvy = M_::psi(mPy[r][q], byP, ayP, dvy);
                            if (szP) dvz = M_::psi(mPz[r][q], bzP, azP, dvz);
                        }
                        out = M_::upd(p0[r][q], F.fq, F.fqf, (mp_[r] >> q) & 1u, (dvx + dvy) + dvz);
                    }
                }
                p1[r][q] = out;
            }
        }

        if (owned) {
            // ---- velocities, stores and CPML memory of the owned rows
#pragma unroll
            for (int r = 1; r <= TY; ++r) {
                const int j = j0 + r - 1;
                const T pml = shfl_up1<T>(p1[r][CPL - 1]);  // P1 of the previous lane's last cell
                if (!XL.own || j >= ny) continue;
                const int rb = rowb(j, k);
                const bool rj = GEN && (j == ny - 1);
                if (cpP && !rj && !rk) {
#pragma unroll
                    for (int q = 0; q < CPL; ++q)
                        if (sxP[q]) psi1[F.mv_off(0) + lxP[q] + p * (j + ny * (k + nz * (sxP[q] - 1)))] = mPx[r][q];
                }
                T ayV = (T)0, byV = (T)0;
                if constexpr (GEN) {
                    if (syP[r] && !rj) srow(rS1, (F.mv_off(1) + nx * (lyP[r] + p * (k + nz * (syP[r] - 1)))) * ES, mPy[r]);
                    if (szP && !rj) srow(rS1, (F.mv_off(2) + nx * (j + ny * (lzP + p * (szP - 1)))) * ES, mPz[r]);
                    if (syV[r]) { ayV = F.cah(lyV[r] + qv * (syV[r] - 1)); byV = F.cbh(lyV[r] + qv * (syV[r] - 1)); }
                }
                T vx1[CPL], vy1[CPL], vz1[CPL];
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int i = i0 + q;
                    const T pm = q > 0 ? p1[r][q - 1] : pml;
                    vx1[q] = vx0[r][q];
                    if (i >= 1) {
                        A dp = M_::diff(p1[r][q], pm, F.dx, F.rdx);
                        if (sxV[q]) dp = M_::psi(mVx[r][q], bxV[q], axV[q], dp);
                        vx1[q] = M_::upd(vx0[r][q], F.fv, F.fvf, (mvx_[r] >> q) & 1u, dp);
                    }
                    vy1[q] = vy0[r][q];
                    if (!GEN || j >= 1) {
                        A dp = M_::diff(p1[r][q], p1[r - 1][q], F.dy, F.rdy);
                        if constexpr (GEN) {
                            if (syV[r]) dp = M_::psi(mVy[r][q], byV, ayV, dp);
                        }
                        vy1[q] = M_::upd(vy0[r][q], F.fv, F.fvf, (mvy_[r] >> q) & 1u, dp);
                    }
                    vz1[q] = vzc[r][q];
                    if (!GEN || k >= 1) {
                        A dp = M_::diff(p1[r][q], p1prev[r][q], F.dz, F.rdz);
                        if constexpr (GEN) {
                            if (szV) dp = M_::psi(mVz[r][q], bzV, azV, dp);
                        }
                        vz1[q] = M_::upd(vzc[r][q], F.fv, F.fvf, (mvz_[r] >> q) & 1u, dp);
                    }
                }
                srow(rF1, rb, p1[r]);
                srow(rF1, chrow(1, rb), vx1);
                if (!GEN || j >= 1 || !F.fuse_mirror) srow(rF1, chrow(2, rb), vy1);
                if (!GEN || k >= 1 || !F.fuse_mirror) srow(rF1, chrow(3, rb), vz1);
                if constexpr (GEN) {
                    if (F.fuse_mirror && (j == 1 || k == 1)) {
                        T neg[CPL];
                        if (j == 1) {  // forward.jl:135
#pragma unroll
                            for (int q = 0; q < CPL; ++q) neg[q] = -vy1[q];
                            srow(rF1, chrow(2, (int)((uint32_t)rb - (uint32_t)(nx * ES))), neg);
                        }
                        if (k == 1) {  // forward.jl:136
#pragma unroll
                            for (int q = 0; q < CPL; ++q) neg[q] = -vz1[q];
                            srow(rF1, chrow(3, (int)((uint32_t)rb - (uint32_t)(sxy * ES))), neg);
                        }
                    }
                }
                if (cpV) {
#pragma unroll
                    for (int q = 0; q < CPL; ++q)
                        if (sxV[q]) psi1[F.mp_off(0) + lxV[q] + qv * (j + ny * (k + nz * (sxV[q] - 1)))] = mVx[r][q];
                }
                if constexpr (GEN) {
                    if (syV[r]) srow(rS1, (F.mp_off(1) + nx * (lyV[r] + qv * (k + nz * (syV[r] - 1)))) * ES, mVy[r]);
                    if (szV) srow(rS1, (F.mp_off(2) + nx * (j + ny * (lzV + qv * (szV - 1)))) * ES, mVz[r]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r <= TY; ++r)
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                p1prev[r][q] = p1[r][q];
                vzc[r][q] = vzn[r][q];
            }
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8), each with its own L2.  Block b takes the tiles of chunk b % 8
// of the list (8 contiguous chunks, chunk x spanning [x q + min(x, r),
// +q + (x < r)) for n = 8 q + r), so each XCD sweeps a contiguous run of
// neighbouring tiles, which re-read each other's halo rows from its L2.
// Tile entries: x = masked << 31 | kind << 28 | member << 16 | first row,
// y = first plane << 16 | planes; `masked` is set when a cell of the tile or
// its halo is closed by the scatterer mask (other tiles skip the mask loads).
__device__ __forceinline__ int xcd_chunk_start(int x, int n) { return x * (n >> 3) + min(x, n & 7); }
__device__ __forceinline__ int xcd_chunk_size(int x, int n) { return (n >> 3) + (x < (n & 7) ? 1 : 0); }
__device__ __forceinline__ int xcd_tile(int b, int n) { return xcd_chunk_start(b & 7, n) + (b >> 3); }

// One velocity cell whose update reads an injected pressure cell,
// recomputed from V0 / psi0 and the injected P1 (after the pressure
// transceivers).  e: packed (axis << 30 | cell) entry.
template <typename T, bool EXACT>
__device__ __forceinline__ void fixup_cell(const Fused<T>& F, uint32_t e) {
    const int ax = (int)(e >> 30);
    const int c = (int)(e & 0x3fffffffu);
    const int nx = F.nx, ny = F.ny, nz = F.nz;
    const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    const bool open = F.mask ? ((mask_word(F.mask, 1 + ax, F.nwx, ny, nz, i, j, k) >> (i & 31)) & 1u) : true;
    const T* P1 = F.P1();
    if (ax == 0) {
        F.V1(0)[c] = v_new<T, EXACT, 0>(F, i, j, k, F.V0(0)[c], P1[c], P1[c - 1], open);
    } else if (ax == 1) {
        const T v = v_new<T, EXACT, 1>(F, i, j, k, F.V0(1)[c], P1[c], P1[c - nx], open);
        F.V1(1)[c] = v;
        if (j == 1 && F.fuse_mirror) F.V1(1)[c - nx] = -v;
    } else {
        const T v = v_new<T, EXACT, 2>(F, i, j, k, F.V0(2)[c], P1[c], P1[c - nx * ny], open);
        F.V1(2)[c] = v;
        if (k == 1 && F.fuse_mirror) F.V1(2)[c - nx * ny] = -v;
    }
}

template <typename T, bool EXACT>
__global__ void k_fixup_v(Fused<T> F, const uint32_t* cells, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) fixup_cell<T, EXACT>(F, cells[t]);
}

// The transceiver work of one fused step in one launch (forward.jl:92-132;
// k_tx_p levels, k_fixup_v and k_tx_v levels otherwise).  The chain is the
// transceivers that inject, or whose stencils touch a cell written by an
// injection or by the velocity fix-up.  It falls apart into groups that
// share no written cell; each chain block takes some whole groups with
// their fix-up cells, in their order_p / order_v level order.  All other
// transceivers only record.
constexpr int TXC_MAXLEV = 8;
// per chain block: offsets into `list` of the pressure levels
// [0, TXC_MAXLEV], of the velocity levels [TXC_MAXLEV + 1, 2 TXC_MAXLEV + 1],
// then the first and end offsets of the fix-up cells
constexpr int TXC_HDR = 2 * (TXC_MAXLEV + 1) + 2;
struct TxChain {
    const uint32_t* list;  // per chain block: order_p entries, order_v entries, fix-up cells
    const int* hdr;        // TXC_HDR offsets per chain block
    const uint32_t* rec;   // indices of the record-only transceivers
    int nchain, n_rec;
    int nlp, nlv;          // levels of the two orders
};

// Blocks 0..nchain-1 run their groups in the reference order: pressure
// levels, fix-up, velocity levels, with a workgroup barrier between phases
// that hold work (every wave of a block is on one CU, so the barrier also
// orders the block's global stores and loads).  The other blocks take 8
// record-only transceivers each (32 lanes per transceiver): their records
// read cells nothing in this launch writes, so they run beside the chain.
template <typename T, bool EXACT>
__device__ __forceinline__ void tx_step_block(const Fused<T>& F, const Txs<T>& X, const TxChain& C, int it0,
                                              const GfHist& H, int blk) {
    const int tid = (int)threadIdx.x, lw = tid & 63;
    T* P = F.P1();
    T* Vx = F.V1(0);
    T* Vy = F.V1(1);
    T* Vz = F.V1(2);
    if (blk < C.nchain) {
        // the block's offsets: one vector load, then lane reads (a scalar
        // load per offset would wait a memory round trip each)
        const int hv = lw < TXC_HDR ? C.hdr[(size_t)blk * TXC_HDR + lw] : 0;
        auto h = [&](int i) { return __builtin_amdgcn_readlane(hv, i); };
        // Pressure entries fill the block from thread 0, velocity entries
        // from thread 64 and fix-up cells from thread 128 (rotated), so with
        // few entries the first velocity pass's loads (entry, index, weight,
        // record, source) and the fix-up entries are issued at the start by
        // waves the pressure phase leaves idle, beside its loads.  None of
        // them reads a cell this launch writes.
        const int vt = (tid + 192) & 255, ft = (tid + 128) & 255;
        int lv0 = 0;
        while (lv0 < C.nlv && h(TXC_MAXLEV + 1 + lv0) == h(TXC_MAXLEV + 2 + lv0)) ++lv0;
        const int vlo = lv0 < C.nlv ? h(TXC_MAXLEV + 1 + lv0) : 0;
        const int vhi = lv0 < C.nlv ? h(TXC_MAXLEV + 2 + lv0) : 0;
        const int flo = h(2 * (TXC_MAXLEV + 1)), fhi = h(2 * (TXC_MAXLEV + 1) + 1);
        const bool vany0 = vlo + (vt >> 5) < vhi;
        const TxV<T> v0 = tx_v_load<T>(F.nx, F.ny, F.nz, X, vany0 ? C.list[vlo + (vt >> 5)] : 0u, vany0, lw, it0);
        const uint32_t fe0 = flo + ft < fhi ? C.list[flo + ft] : 0u;
        bool pend = false;  // a phase with work ran since the last barrier
        for (int L = 0; L < C.nlp; ++L) {
            const int lo = h(L), hi = h(L + 1);
            if (lo == hi) continue;
            if (pend) __syncthreads();
            for (int b = lo; b < hi; b += 32) {
                const int g = b + (tid >> 3);
                const bool live = g < hi;
                tx_p_entry<T>(P, F.nx, F.ny, F.nz, X, live ? C.list[g] : 0u, live, lw, it0);
            }
            pend = true;
        }
        if (flo < fhi) {
            if (pend) __syncthreads();
            for (int b = flo; b < fhi; b += 256)
                if (b + ft < fhi) fixup_cell<T, EXACT>(F, b == flo ? fe0 : C.list[b + ft]);
            pend = true;
        }
        for (int L = 0; L < C.nlv; ++L) {
            const int lo = h(TXC_MAXLEV + 1 + L), hi = h(TXC_MAXLEV + 2 + L);
            if (lo == hi) continue;
            if (pend) __syncthreads();
            for (int b = lo; b < hi; b += 8) {
                const int g = b + (vt >> 5);
                const bool any = g < hi;
                if (L == lv0 && b == lo)
                    tx_v_apply<T>(Vx, Vy, Vz, X, v0, any, lw, it0, H);
                else
                    tx_v_entry<T>(Vx, Vy, Vz, F.nx, F.ny, F.nz, X, any ? C.list[g] : 0u, any, lw, it0, H);
            }
            pend = true;
        }
        return;
    }
    const int g = (blk - C.nchain) * 8 + (tid >> 5);
    const bool any = g < C.n_rec;
    tx_record_entry<T>(P, Vx, Vy, Vz, F.nx, F.ny, F.nz, X, any ? (int)C.rec[g] : 0, any, lw, it0, H);
}

template <typename T, bool EXACT>
__global__ __launch_bounds__(256) void k_tx_step(Fused<T> F, Txs<T> X, TxChain C, int it0, GfHist H) {
    tx_step_block<T, EXACT>(F, X, C, it0, H, (int)blockIdx.x);
}

// Transceiver work carried by a single solve's sweep launch
// (ADF_OPT_TX_LAUNCH = 2): one-wave blocks before and after the nbt tile
// blocks.
//   - Record blocks (first in the grid) take the record-only transceivers
//     of the previous step, TXS_REC per block, from the sweep's old
//     buffers: the state that step left, which the sweep only reads.  At
//     the front they also stagger the tiles' start (the launch ran ~2 us
//     shorter than with them after the tiles, and ~1 us shorter than
//     without them, at 256^3).
//   - Chain blocks (last in the grid) run k_tx_step's chain groups on the
//     new buffers.  Each first waits until the tiles owning the cells its
//     groups read or write have stored them: such a tile (TILE_SIGNAL in
//     its entry) publishes the launch's epoch in tflag after an agent-scope
//     release, and the block acquires it.  No other tile reads the new
//     buffers or writes those cells.
// Blocks are dispatched in index order, so every tile a chain block waits
// for is running or done when it starts and waits for nothing itself; the
// wait is bounded all the same (a wait that runs out sets *err and the
// block skips its work, so the results are then invalid and the host
// reports it).
template <typename T>
struct TxSweep {
    Txs<T> X;
    TxChain C;
    const int* wait_off;  // chain block b waits for tiles wait[wait_off[b] .. wait_off[b + 1])
    const int* wait;
    int* tflag;           // per tile: epoch of the launch that last stored it
    int* err;
    int epoch;
    int nbt;              // tile blocks (0: the launch carries no transceiver work)
    int nrb;              // record blocks
    int it_rec;           // 0-based step of the records
    int it0;              // 0-based step of the chain
};
constexpr uint32_t TILE_SIGNAL = 1u << 27;  // tile entry bit (free in a single solve's entries)
constexpr int TXS_SPIN = 1 << 17;           // polls of s_sleep 4 (~0.1 us each)

template <typename T>
__device__ __forceinline__ void tx_tile_signal(const TxSweep<T>& S, int t) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(S.tflag + t, S.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Record blocks take TXS_REC record-only transceivers each, 32 lanes per
// transceiver (tx_record_entry's lanes) in TXS_REC / 2 rounds: the loads
// that do not read the field for every round first, then the corners, then
// the sums in the reference's corner order, as tx_record_entry does.
constexpr int TXS_REC = 8;

template <typename T>
__device__ __forceinline__ void tx_sweep_records(const Fused<T>& F, const TxSweep<T>& S, int rb) {
    constexpr int NR = TXS_REC / 2;
    const int lw = (int)(threadIdx.x & 63);
    const int ch = (lw >> 3) & 3, q = lw & 7, g32 = lw & ~31, g0 = lw & ~7;
    const int c = ch < 3 ? ch + 1 : 0;  // buffer channel: P, Vx, Vy, Vz
    const Txs<T>& X = S.X;
    const int ntx = X.ntx, nt = X.nt;
    const size_t cs = (size_t)ntx * nt;
    const T* Fc = F.f0 + (size_t)c * F.n3;
    int is[NR];
    bool any[NR];
    TxLane<T> L[NR];
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int g = rb * TXS_REC + 2 * u + (lw >> 5);
        any[u] = g < S.C.n_rec;
        is[u] = any[u] ? (int)S.C.rec[g] : 0;
    }
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const size_t o = (size_t)is[u] * nt + S.it_rec + cs * c;
        L[u] = tx_lane_load<T>(X.idx + 3 * (is[u] + ntx * c), F.nx, F.ny, F.nz, X.w + 8 * (is[u] + ntx * c),
                               X.rec + o, X.src + o, false, any[u], q);
    }
    T f[NR];
#pragma unroll
    for (int u = 0; u < NR; ++u) f[u] = L[u].ok ? Fc[L[u].c] : (T)0;
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        // tx_lane_apply with the corner already loaded (records only)
        const T prod = L[u].w * f[u];
        T r = L[u].r0;
#pragma unroll
        for (int qq = 0; qq < 8; ++qq) r = r + __shfl(prod, g0 + qq, 64);
        if (!L[u].ok) r = L[u].r0;
        const size_t o = (size_t)is[u] * nt + S.it_rec;
        if (L[u].ok && q == 0) X.rec[o + cs * c] = r;
        const T rx = __shfl(r, g32, 64), ry = __shfl(r, g32 + 8, 64), rz = __shfl(r, g32 + 16, 64);
        if (any[u] && lw == g32) {
            const T* d = X.dir + 4 * is[u];
            X.vn_rec[o] = rx * d[1] + ry * d[2] + rz * d[3];
        }
    }
}

template <typename T, bool EXACT>
__device__ __forceinline__ void tx_sweep_chain(const Fused<T>& F, const TxSweep<T>& S, int cb) {
    const int lw = (int)(threadIdx.x & 63);
    const TxChain& C = S.C;
    const int hv = lw < TXC_HDR ? C.hdr[(size_t)cb * TXC_HDR + lw] : 0;
    auto h = [&](int i) { return __builtin_amdgcn_readlane(hv, i); };
    const int w0 = S.wait_off[cb], w1 = S.wait_off[cb + 1];
    bool ok = true;
    for (int base = w0; base < w1 && ok; base += 64) {
        const int t = base + lw < w1 ? S.wait[base + lw] : -1;
        for (int n = 0;; ++n) {
            const bool done =
                t < 0 || __hip_atomic_load(S.tflag + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == S.epoch;
            if (__all(done)) break;
            if (n >= TXS_SPIN) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (!ok) {
        if (lw == 0) atomicOr(S.err, 1);
        return;
    }
    T* P = F.P1();
    T* Vx = F.V1(0);
    T* Vy = F.V1(1);
    T* Vz = F.V1(2);
    const GfHist H{};
    bool pend = false;  // a phase with work ran since the last barrier
    for (int L = 0; L < C.nlp; ++L) {
        const int lo = h(L), hi = h(L + 1);
        if (lo == hi) continue;
        if (pend) __syncthreads();
        for (int b = lo; b < hi; b += 8) {
            const int g = b + (lw >> 3);
            const bool live = g < hi;
            tx_p_entry<T>(P, F.nx, F.ny, F.nz, S.X, live ? C.list[g] : 0u, live, lw, S.it0);
        }
        pend = true;
    }
    const int flo = h(2 * (TXC_MAXLEV + 1)), fhi = h(2 * (TXC_MAXLEV + 1) + 1);
    if (flo < fhi) {
        if (pend) __syncthreads();
        for (int b = flo; b < fhi; b += 64)
            if (b + lw < fhi) fixup_cell<T, EXACT>(F, C.list[b + lw]);
        pend = true;
    }
    for (int L = 0; L < C.nlv; ++L) {
        const int lo = h(TXC_MAXLEV + 1 + L), hi = h(TXC_MAXLEV + 2 + L);
        if (lo == hi) continue;
        if (pend) __syncthreads();
        for (int b = lo; b < hi; b += 2) {
            const int g = b + (lw >> 5);
            const bool any = g < hi;
            tx_v_entry<T>(Vx, Vy, Vz, F.nx, F.ny, F.nz, S.X, any ? C.list[g] : 0u, any, lw, S.it0, H);
        }
        pend = true;
    }
}

// The records of the last step of a run (the sweep launches take each
// step's records one step later): from the current buffers, which F holds
// as its old ones.
template <typename T>
__global__ __launch_bounds__(64) void k_tx_records(Fused<T> F, TxSweep<T> S) {
    tx_sweep_records<T>(F, S, (int)blockIdx.x);
}

}  // namespace adf
