// adfdtd device kernels (gfx950, wave64).
//
// Arithmetic modes (template flag EXACT):
//   EXACT = true   the reference's arithmetic: neighbour differences in the
//                  storage type T, everything that touches a Float64 domain
//                  scalar in double, true divisions by dx/dy/dz, no
//                  contraction (the file is built with -ffp-contract=off),
//                  stores rounded to T.  T = double gives the Float64
//                  algorithm, T = float the reference's own Float32 runs
//                  (src/kernels/updates.jl with Float32 arrays and Float64
//                  `dx`, `fact_m0`, `fact_m1`, src/simulation.jl:25-44).
//   EXACT = false  float storage, float arithmetic, reciprocal multiplies
//                  and explicit FMAs (fast path).
//
// All indices are 0-based here; the reference's 1-based ranges are noted.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace adf {

// Scalars and pointers of one solve.  Arrays follow the reference's Julia
// layouts: field channel c at F + c*n3 with x fastest; CPML memory
// mv_x (p,ny,nz,2), mv_y (nx,p,nz,2), mv_z (nx,ny,p,2), mp_* with p+1;
// coefficients a,b (p,2), ah,bh (p+1,2).
template <typename T>
struct Solve {
    T* P; T* Vx; T* Vy; T* Vz;
    T* mvx; T* mvy; T* mvz;
    T* mpx; T* mpy; T* mpz;
    const T* a; const T* b; const T* ah; const T* bh;
    const uint32_t* mask;  // 4 bit-planes of nz*ny*nwx words, or null
    int nx, ny, nz, p, nwx;
    int cpml_p;            // CPML in the P update (p>0 && (mask || p>1))
    int fuse_mirror;       // write the rigid mirrors inside the V update
    double dx, dy, dz, fq, fv;
    float rdx, rdy, rdz, fqf, fvf;
};

__device__ __forceinline__ bool mask_bit(const uint32_t* m, int plane, int nwx, int ny, int nz,
                                         int i, int j, int k) {
    const uint32_t w = m[((size_t)plane * nz * ny + (size_t)k * ny + j) * nwx + (i >> 5)];
    return (w >> (i & 31)) & 1u;
}

// derivatives.jl:66-79 (getBoundaryIndex), 0-based: returns side 0 (none),
// 1 (left, l = i) or 2 (right, l = i - (n - npml)).
__device__ __forceinline__ int slab(int i, int n, int npml, int& l) {
    if (i < npml) { l = i; return 1; }
    l = i - (n - npml);
    return l >= 0 ? 2 : 0;
}

// ---------------------------------------------------------------- exact
// One CPML memory update psi = b*psi + a*dv ; dv += psi (updates.jl:58-59).
template <typename T>
__device__ __forceinline__ double psi_exact(T* m, T bc, T ac, double dv) {
    const T bm = bc * (*m);
    const T nm = (T)((double)bm + (double)ac * dv);
    *m = nm;
    return dv + (double)nm;
}
__device__ __forceinline__ float psi_fast(float* m, float bc, float ac, float dv) {
    const float nm = fmaf(bc, *m, ac * dv);
    *m = nm;
    return dv + nm;
}

// Pressure at cell (i,j,k), i in [0,nx-2], (j,k) != (ny-1,nz-1).
// forward.jl:66-89; updates.jl:21-142,288-378.
template <typename T, bool EXACT>
__device__ __forceinline__ void p_cell(const Solve<T>& S, int i, int j, int k) {
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int sxy = nx * ny;
    const int c = i + nx * (j + ny * k);
    const bool rj = (j == ny - 1), rk = (k == nz - 1);
    if (rj || rk) {
        // rigid planes: update_p! with -2V/d at j==ny / k==nz, no CPML, no mask
        if constexpr (EXACT) {
            const double dvx = (double)(T)(S.Vx[c + 1] - S.Vx[c]) / S.dx;
            const double dvy = rj ? (double)((T)(-2) * S.Vy[c]) / S.dy
                                  : (double)(T)(S.Vy[c + nx] - S.Vy[c]) / S.dy;
            const double dvz = rk ? (double)((T)(-2) * S.Vz[c]) / S.dz
                                  : (double)(T)(S.Vz[c + sxy] - S.Vz[c]) / S.dz;
            S.P[c] = (T)((double)S.P[c] - S.fq * ((dvx + dvy) + dvz));
        } else {
            const float dvx = (float)(S.Vx[c + 1] - S.Vx[c]) * S.rdx;
            const float dvy = rj ? (-2.0f * (float)S.Vy[c]) * S.rdy : (float)(S.Vy[c + nx] - S.Vy[c]) * S.rdy;
            const float dvz = rk ? (-2.0f * (float)S.Vz[c]) * S.rdz : (float)(S.Vz[c + sxy] - S.Vz[c]) * S.rdz;
            S.P[c] = (T)fmaf(-S.fqf, (dvx + dvy) + dvz, (float)S.P[c]);
        }
        return;
    }
    const int p = S.p;
    if constexpr (EXACT) {
        double dvx = (double)(T)(S.Vx[c + 1] - S.Vx[c]) / S.dx;
        double dvy = (double)(T)(S.Vy[c + nx] - S.Vy[c]) / S.dy;
        double dvz = (double)(T)(S.Vz[c + sxy] - S.Vz[c]) / S.dz;
        if (S.cpml_p) {
            int l, sd;
            if ((sd = slab(i, nx, p, l)))
                dvx = psi_exact(&S.mvx[l + p * (j + ny * (k + nz * (sd - 1)))], S.b[l + p * (sd - 1)],
                                S.a[l + p * (sd - 1)], dvx);
            if ((sd = slab(j, ny, p, l)))
                dvy = psi_exact(&S.mvy[i + nx * (l + p * (k + nz * (sd - 1)))], S.b[l + p * (sd - 1)],
                                S.a[l + p * (sd - 1)], dvy);
            if ((sd = slab(k, nz, p, l)))
                dvz = psi_exact(&S.mvz[i + nx * (j + ny * (l + p * (sd - 1)))], S.b[l + p * (sd - 1)],
                                S.a[l + p * (sd - 1)], dvz);
        }
        double coeff = S.fq;
        if (S.mask && !mask_bit(S.mask, 0, S.nwx, ny, nz, i, j, k)) coeff = 0.0;
        S.P[c] = (T)((double)S.P[c] - coeff * ((dvx + dvy) + dvz));
    } else {
        float dvx = (float)(S.Vx[c + 1] - S.Vx[c]) * S.rdx;
        float dvy = (float)(S.Vy[c + nx] - S.Vy[c]) * S.rdy;
        float dvz = (float)(S.Vz[c + sxy] - S.Vz[c]) * S.rdz;
        if (S.cpml_p) {
            int l, sd;
            if ((sd = slab(i, nx, p, l)))
                dvx = psi_fast((float*)&S.mvx[l + p * (j + ny * (k + nz * (sd - 1)))], S.b[l + p * (sd - 1)],
                               S.a[l + p * (sd - 1)], dvx);
            if ((sd = slab(j, ny, p, l)))
                dvy = psi_fast((float*)&S.mvy[i + nx * (l + p * (k + nz * (sd - 1)))], S.b[l + p * (sd - 1)],
                               S.a[l + p * (sd - 1)], dvy);
            if ((sd = slab(k, nz, p, l)))
                dvz = psi_fast((float*)&S.mvz[i + nx * (j + ny * (l + p * (sd - 1)))], S.b[l + p * (sd - 1)],
                               S.a[l + p * (sd - 1)], dvz);
        }
        if (!S.mask || mask_bit(S.mask, 0, S.nwx, ny, nz, i, j, k))
            S.P[c] = (T)fmaf(-S.fqf, (dvx + dvy) + dvz, (float)S.P[c]);
    }
}

// One velocity component at cell (i,j,k); axis 0 needs i>=1, 1 needs j>=1,
// 2 needs k>=1 (the reference ranges 2:nx, 2:ny, 2:nz).  Returns the new
// value.  updates.jl:192-275,380-501.
template <typename T, bool EXACT, int AX>
__device__ __forceinline__ T v_cell(const Solve<T>& S, int i, int j, int k, int c, T Pc) {
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int q = S.p + 1;
    T* V = AX == 0 ? S.Vx : (AX == 1 ? S.Vy : S.Vz);
    T* mp = AX == 0 ? S.mpx : (AX == 1 ? S.mpy : S.mpz);
    const int st = AX == 0 ? 1 : (AX == 1 ? nx : nx * ny);
    const int idx = AX == 0 ? i : (AX == 1 ? j : k);
    const int n = AX == 0 ? nx : (AX == 1 ? ny : nz);
    T out;
    if constexpr (EXACT) {
        const double h = AX == 0 ? S.dx : (AX == 1 ? S.dy : S.dz);
        double dp = (double)(T)(Pc - S.P[c - st]) / h;
        if (q > 1) {
            int l, sd;
            if ((sd = slab(idx, n, q, l))) {
                const int mi = AX == 0 ? l + q * (j + ny * (k + nz * (sd - 1)))
                             : (AX == 1 ? i + nx * (l + q * (k + nz * (sd - 1)))
                                        : i + nx * (j + ny * (l + q * (sd - 1))));
                dp = psi_exact(&mp[mi], S.bh[l + q * (sd - 1)], S.ah[l + q * (sd - 1)], dp);
            }
        }
        double coeff = S.fv;
        if (S.mask && !mask_bit(S.mask, 1 + AX, S.nwx, ny, nz, i, j, k)) coeff = 0.0;
        out = (T)((double)V[c] - coeff * dp);
    } else {
        const float rh = AX == 0 ? S.rdx : (AX == 1 ? S.rdy : S.rdz);
        float dp = (float)(Pc - S.P[c - st]) * rh;
        if (q > 1) {
            int l, sd;
            if ((sd = slab(idx, n, q, l))) {
                const int mi = AX == 0 ? l + q * (j + ny * (k + nz * (sd - 1)))
                             : (AX == 1 ? i + nx * (l + q * (k + nz * (sd - 1)))
                                        : i + nx * (j + ny * (l + q * (sd - 1))));
                dp = psi_fast((float*)&mp[mi], S.bh[l + q * (sd - 1)], S.ah[l + q * (sd - 1)], dp);
            }
        }
        out = V[c];
        if (!S.mask || mask_bit(S.mask, 1 + AX, S.nwx, ny, nz, i, j, k))
            out = (T)fmaf(-S.fvf, dp, (float)V[c]);
    }
    V[c] = out;
    return out;
}

// ------------------------------------------------------------ v1 kernels
// Block (64,4); each thread walks KZ planes of one (i,j) column.

template <typename T, bool EXACT>
__global__ __launch_bounds__(256) void k_update_p(Solve<T> S, int kz) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    if (i >= S.nx - 1 || j >= S.ny) return;
    const int k0 = blockIdx.z * kz;
    const int k1 = min(k0 + kz, S.nz);
    for (int k = k0; k < k1; ++k) {
        if (j == S.ny - 1 && k == S.nz - 1) continue;
        p_cell<T, EXACT>(S, i, j, k);
    }
}

template <typename T, bool EXACT>
__global__ __launch_bounds__(256) void k_update_v(Solve<T> S, int kz) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int nx = S.nx, ny = S.ny;
    if (i >= nx || j >= ny) return;
    const int k0 = blockIdx.z * kz;
    const int k1 = min(k0 + kz, S.nz);
    for (int k = k0; k < k1; ++k) {
        const int c = i + nx * (j + ny * k);
        const T Pc = S.P[c];
        if (i >= 1) v_cell<T, EXACT, 0>(S, i, j, k, c, Pc);
        if (j >= 1) {
            const T v = v_cell<T, EXACT, 1>(S, i, j, k, c, Pc);
            if (S.fuse_mirror && j == 1) S.Vy[c - nx] = -v;  // forward.jl:135
        }
        if (k >= 1) {
            const T v = v_cell<T, EXACT, 2>(S, i, j, k, c, Pc);
            if (S.fuse_mirror && k == 1) S.Vz[c - nx * ny] = -v;  // forward.jl:136
        }
    }
}

// forward.jl:134-137 as a separate pass (used when a velocity transceiver
// touches the mirrored planes).
template <typename T>
__global__ void k_mirror(Solve<T> S) {
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nx * nz) {
        const int i = t % nx, k = t / nx;
        S.Vy[i + nx * ny * k] = -S.Vy[i + nx * (1 + ny * k)];
    }
    if (t < nx * ny) S.Vz[t] = -S.Vz[t + nx * ny];
}

// ------------------------------------------------------------ transceivers
// Packed transceivers of one solve (device copy of transceiversToGrid,
// simulation.jl:387-431): idx int32 [4][ntx][3] 1-based, w [4][ntx][8],
// src/rec [4][ntx][nt], dir [ntx][4], vn_src/vn_rec [ntx][nt].
template <typename T>
struct Txs {
    const int32_t* idx;
    const T* w;
    T* src;
    T* rec;
    const T* dir;
    T* vn_src;
    T* vn_rec;
    int ntx, nt;
};

// One transceiver on one channel over 8 lanes of a wave (q = corner, i
// fastest; derivatives.jl:30-60): each lane loads its corner and weight,
// the record is summed in the reference's corner order (every lane runs the
// same sequence of shuffles, so every lane holds the sum), then each lane
// injects into its corner.  The eight corners are distinct cells, so this
// equals the reference's record-then-inject loop value for value.  `live`
// lanes belong to a real entry; all lanes must call this (shuffles).
// Returns the new record (the old one when the stencil is skipped).
// tx_lane_load does the loads that do not read the field (index, weight,
// old record, source sample); tx_lane_apply reads the corner and records /
// injects, so a caller may issue the first part ahead.
template <typename T>
struct TxLane {
    int c;    // the lane's corner cell
    bool ok;  // the stencil lies inside the grid
    T w, r0, sv;
};

template <typename T>
__device__ __forceinline__ TxLane<T> tx_lane_load(const int32_t* id3, int nx, int ny, int nz, const T* w8,
                                                  const T* rec, const T* src, bool inj, bool live, int q) {
    TxLane<T> L;
    int I = 0, J = 0, K = 0;
    if (live) { I = id3[0]; J = id3[1]; K = id3[2]; }
    L.ok = live && I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz;
    L.c = (I - 1) + nx * ((J - 1) + ny * (K - 1)) + (q & 1) + nx * ((q >> 1) & 1) + nx * ny * (q >> 2);
    L.w = L.ok ? w8[q] : (T)0;
    L.r0 = live ? *rec : (T)0;
    L.sv = (L.ok && inj) ? *src : (T)0;
    return L;
}

template <typename T>
__device__ __forceinline__ T tx_lane_apply(T* F, const TxLane<T>& L, T* rec, bool inj, int q, int g0) {
    const T f = L.ok ? F[L.c] : (T)0;
    const T prod = L.w * f;
    T r = L.r0;
#pragma unroll
    for (int qq = 0; qq < 8; ++qq) r = r + __shfl(prod, g0 + qq, 64);
    if (!L.ok) r = L.r0;
    if (L.ok && q == 0) *rec = r;
    if (L.ok && inj) F[L.c] = f + L.w * L.sv;
    return r;
}

template <typename T>
__device__ __forceinline__ T tx_corners(T* F, const int32_t* id3, int nx, int ny, int nz, const T* w8, T* rec,
                                        const T* src, bool inj, bool live, int q, int g0) {
    return tx_lane_apply<T>(F, tx_lane_load<T>(id3, nx, ny, nz, w8, rec, src, inj, live, q), rec, inj, q, g0);
}

// Pressure transceivers of one conflict-free level (forward.jl:92), 8 lanes
// per transceiver.  order entries carry the injector flag in bit 31:
// transceivers whose p source is zero for the whole run only record.
// One order_p entry e over 8 lanes (lw: lane in the wave); all 8 lanes call.
template <typename T>
__device__ __forceinline__ void tx_p_entry(T* P, int nx, int ny, int nz, const Txs<T>& X, uint32_t e, bool live,
                                           int lw, int it0) {
    const int is = (int)(e & 0x7fffffffu);
    const size_t o = (size_t)is * X.nt + it0;
    tx_corners<T>(P, X.idx + 3 * is, nx, ny, nz, X.w + 8 * is, X.rec + o, X.src + o, (e >> 31) != 0, live, lw & 7,
                  lw & ~7);
}

template <typename T>
__global__ __launch_bounds__(256) void k_tx_p(T* P, int nx, int ny, int nz, Txs<T> X, const uint32_t* order, int n,
                                              int it0) {
    const int g = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 3);
    const bool live = g < n;
    tx_p_entry<T>(P, nx, ny, nz, X, live ? order[g] : 0u, live, (int)(threadIdx.x & 63), it0);
}

// The history of a GF map's inputs, [n_irec][nt] doubles, which k_tx_v
// writes once both records of a step exist:
// PVFromPin / PFromPin: pin = (p_rec - vn_rec*z0)/2 (extrapolation.jl:224,241)
// PVFromPV: h0 = p_rec, h1 = vn_rec (extrapolation.jl:207-208).
// first[is]: the first input slot reading transceiver is (-1: none), next[r]:
// the next slot reading the same transceiver; first is null without a map
// with write-back (nothing reads the history then).
struct GfHist {
    const int32_t* first;
    const int32_t* next;
    double* h0;
    double* h1;
    int kind;
    double z0;
};

// order_p / order_v entries: transceiver index | TX_INJ (the transceiver
// injects on that phase) | TX_GF_OUT (order_v: an output of a GF map with
// write-back, whose vn2v! runs after the extrapolation, in k_gf_near).
constexpr uint32_t TX_INJ = 1u << 31, TX_GF_OUT = 1u << 30, TX_INDEX = TX_GF_OUT - 1;

// The GF history of step it0 for transceiver is from its records p and vn.
__device__ __forceinline__ void gf_hist_write(const GfHist& H, int is, int nt, int it0, double p, double vn) {
    for (int r = H.first[is]; r >= 0; r = H.next[r]) {
        const size_t hb = (size_t)r * nt + it0;
        if (H.kind == 2) {
            H.h0[hb] = p;
            H.h1[hb] = vn;
        } else {
            H.h0[hb] = (p - vn * H.z0) / 2.0;
        }
    }
}

// One order_v entry e over the 32 lanes of a wave half (lw: lane in the
// wave; channel x, y, z on lanes 0-7, 8-15, 16-23 of the half; `any`: the
// entry is real).  All 32 lanes must call this (shuffles).  tx_v_load /
// tx_v_apply split it as tx_lane_load / tx_lane_apply do.
template <typename T>
struct TxV {
    uint32_t e;
    TxLane<T> L;
};

template <typename T>
__device__ __forceinline__ TxV<T> tx_v_load(int nx, int ny, int nz, const Txs<T>& X, uint32_t e, bool any, int lw,
                                            int it0) {
    const int ch = (lw >> 3) & 3, q = lw & 7;
    const int is = (int)(e & TX_INDEX);
    const size_t o = (size_t)is * X.nt + it0;
    const size_t cs = (size_t)X.ntx * X.nt;  // channel stride of src/rec
    const int c1 = (ch < 3 ? ch : 2) + 1;
    TxV<T> v;
    v.e = e;
    v.L = tx_lane_load<T>(X.idx + 3 * (is + X.ntx * c1), nx, ny, nz, X.w + 8 * (is + X.ntx * c1),
                          X.rec + o + cs * c1, X.src + o + cs * c1, (e & TX_INJ) != 0, any && ch < 3, q);
    return v;
}

template <typename T>
__device__ __forceinline__ void tx_v_apply(T* Vx, T* Vy, T* Vz, const Txs<T>& X, const TxV<T>& v, bool any, int lw,
                                           int it0, const GfHist& H) {
    const int ch = (lw >> 3) & 3, q = lw & 7, g32 = lw & ~31, g0 = lw & ~7;
    const bool live = any && ch < 3;
    const uint32_t e = v.e;
    const int is = (int)(e & TX_INDEX);
    const bool inj = (e & TX_INJ) != 0, gf_out = (e & TX_GF_OUT) != 0;
    const int ntx = X.ntx, nt = X.nt;
    const size_t o = (size_t)is * nt + it0;
    const size_t cs = (size_t)ntx * nt;
    const int c1 = (ch < 3 ? ch : 2) + 1;
    T* F = c1 == 1 ? Vx : (c1 == 2 ? Vy : Vz);
    const T r = tx_lane_apply<T>(F, v.L, X.rec + o + cs * c1, inj, q, g0);
    const T rx = __shfl(r, g32, 64), ry = __shfl(r, g32 + 8, 64), rz = __shfl(r, g32 + 16, 64);
    if (!any) return;
    const T* d = X.dir + 4 * is;
    if (lw == g32) {
        const T vnr = rx * d[1] + ry * d[2] + rz * d[3];
        X.vn_rec[o] = vnr;
        if (H.first) gf_hist_write(H, is, nt, it0, (double)X.rec[o], (double)vnr);
    }
    if (inj && !gf_out && it0 + 1 < nt && live && q == 0) {
        const T vn = X.vn_src[o + 1];
        X.src[o + 1 + cs * c1] = X.src[o + 1 + cs * c1] + vn * d[c1];
    }
}

template <typename T>
__device__ __forceinline__ void tx_v_entry(T* Vx, T* Vy, T* Vz, int nx, int ny, int nz, const Txs<T>& X, uint32_t e,
                                           bool any, int lw, int it0, const GfHist& H) {
    tx_v_apply<T>(Vx, Vy, Vz, X, tx_v_load<T>(nx, ny, nz, X, e, any, lw, it0), any, lw, it0, H);
}

// Velocity transceivers of one level (forward.jl:111-115), then v2vn!
// (forward.jl:120) and vn2v! into it+1 (forward.jl:130-132) for the same
// transceiver, and the GF history of the step.  Entries with bit 30 set
// are outputs of a GF map with write-back: their vn2v! runs after the
// extrapolation writes their sources (k_gf_near).  32 lanes per
// transceiver: channel x, y, z on lanes 0-7, 8-15, 16-23.
template <typename T>
__global__ __launch_bounds__(256) void k_tx_v(T* Vx, T* Vy, T* Vz, int nx, int ny, int nz, Txs<T> X,
                                              const uint32_t* order, int n, int it0, GfHist H) {
    const int g = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 5);
    const bool any = g < n;
    tx_v_entry<T>(Vx, Vy, Vz, nx, ny, nz, X, any ? order[g] : 0u, any, (int)(threadIdx.x & 63), it0, H);
}

// Pressure and velocity records of one transceiver that injects nothing
// and whose stencils touch no cell written by the step's injections or
// velocity fix-up (so its records do not depend on the transceiver order),
// over the 32 lanes of a wave half: channel x, y, z on lanes 0-7, 8-15,
// 16-23 and pressure on lanes 24-31, then v2vn! and the GF history.  The
// same operations as k_tx_p + k_tx_v for such an entry.
template <typename T>
__device__ __forceinline__ void tx_record_entry(T* P, T* Vx, T* Vy, T* Vz, int nx, int ny, int nz, const Txs<T>& X,
                                                int is, bool any, int lw, int it0, const GfHist& H) {
    const int ch = (lw >> 3) & 3, q = lw & 7, g32 = lw & ~31, g0 = lw & ~7;
    const int ntx = X.ntx, nt = X.nt;
    const size_t o = (size_t)is * nt + it0;
    const size_t cs = (size_t)ntx * nt;
    const int c = ch < 3 ? ch + 1 : 0;  // buffer channel: P, Vx, Vy, Vz
    T* F = c == 0 ? P : (c == 1 ? Vx : (c == 2 ? Vy : Vz));
    const T r = tx_corners<T>(F, X.idx + 3 * (is + ntx * c), nx, ny, nz, X.w + 8 * (is + ntx * c),
                              X.rec + o + cs * c, X.src + o + cs * c, false, any, q, g0);
    const T rx = __shfl(r, g32, 64), ry = __shfl(r, g32 + 8, 64), rz = __shfl(r, g32 + 16, 64);
    const T rp = __shfl(r, g32 + 24, 64);
    if (!any || lw != g32) return;
    const T* d = X.dir + 4 * is;
    const T vnr = rx * d[1] + ry * d[2] + rz * d[3];
    X.vn_rec[o] = vnr;
    if (H.first) gf_hist_write(H, is, nt, it0, (double)rp, (double)vnr);
}

// ------------------------------------------------------ GF extrapolation
// The GF history of input r at times [t0, t1) from the record traces, as
// k_tx_v writes it at each step.
template <typename T>
__global__ void k_gf_hist_fill(Txs<T> X, const int32_t* iRec, int nin, int kind, double z0, double* h0, double* h1,
                               int t0, int t1) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int len = t1 - t0;
    if (i >= (int64_t)nin * len) return;
    const int r = (int)(i / len), tt = t0 + (int)(i % len);
    const size_t o = (size_t)iRec[r] * X.nt + tt, hb = (size_t)r * X.nt + tt;
    const double p = (double)X.rec[o], vn = (double)X.vn_rec[o];
    if (kind == 2) {
        h0[hb] = p;
        h1[hb] = vn;
    } else {
        h0[hb] = (p - vn * z0) / 2.0;
    }
}

// The gather of extrapolation.jl:101-192: for each output s at step t
// (0-based; written into the source slots at t+1)
//   out(t)[s] = sum_r sum_{k=0}^{min(t, ng-1)} G[k, s, r] * hist[r][t-k]
// (two outputs, and for PVFromPV two histories and four tensors).  Every
// element of G is used once per step, so a step-by-step gather reads the
// whole map every step.  It is split in time instead: a far pass, once per
// block of GF_M steps starting at step B, sums for every step t = B+m of
// the block the terms whose history lies before the block (lags
// k >= m+1), reading each element of G once for all GF_M steps; a near
// pass at each step adds the terms on the block's own history (lags
// k <= m) and the far sums, in a fixed order.
//
// The far pass is a product of a GF_M x K Toeplitz matrix of history
// values, A[m][(r,k)] = hist[r][B+m-k] for B+m-k in [0, B) and 0 elsewhere,
// with the K x n_out map, K = (input, lag): it runs on the FP64 matrix
// cores (v_mfma_f64_16x16x4f64: M = the 16 steps of the block, N = 16
// outputs, 4 lags per instruction).
constexpr int GF_M = 16;         // steps per far pass = the MFMA's M
constexpr int GF_FAR_WAVES = 8192;  // far-pass waves aimed for (inputs per wave are chosen to reach it)

// Inputs per far-pass wave for a map of nin inputs and nout outputs.
inline int gf_far_rc(int64_t nin, int64_t nout) {
    const int64_t n_og = (nout + 15) / 16;
    return (int)std::max<int64_t>(1, (nin * n_og + GF_FAR_WAVES - 1) / GF_FAR_WAVES);
}

// Element offset of G[0, s, r] in the map as given: PVFromPin (ng, n_isrc,
// n_irec); PVFromPV and PFromPin (ng, n_irec, n_isrc).
inline size_t gf_base(int kind, int64_t s, int64_t r, int64_t ng, int64_t d2) {
    return kind == 1 ? (size_t)ng * ((size_t)s + (size_t)d2 * r) : (size_t)ng * ((size_t)r + (size_t)d2 * s);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

typedef double gf_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ gf_d4 mfma_f64(double a, double b, gf_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// The far pass reads the map tile-major: for input r, output group og
// (outputs 16*og..16*og+15) and 16-lag block kb, the 256 words
//   tile[((r * n_og + og) * nkb + kb) * 256 + kk * 64 + n * 4 + j]
//     = G[lag 16*kb + 4*kk + j, output 16*og + n, input r]
// (zero past nt_gf and past the last output), so lane (kk, n) finds its four
// B operands of a 16-lag step in one aligned 16-byte run and a wave reads
// 256 consecutive words.
__host__ __device__ constexpr size_t gf_tile_index(int64_t r, int64_t og, int64_t kb, int n_og, int nkb, int kk,
                                                   int n, int j) {
    return (((size_t)r * n_og + og) * nkb + kb) * 256 + kk * 64 + n * 4 + j;
}

// Far pass of the block [B, B+Mb).  One wave per (16 outputs, chunk of rc
// inputs): wave w takes output group w % n_og and chunk w / n_og.  Lane l
// holds A[m = l%16][kk = l/16] and B[kk = l/16][n = l%16]; instruction j of
// a 16-lag step puts lag 16*kb + 4*kk + j at K index kk (A and B agree, so
// the sum is over all 16 lags).  The accumulator holds D[4v + l/16][l%16] in
// element v.  Writes part[((c * GF_M + m) * 2 + ch) * nout + s] for m < Mb.
// No LDS and no barriers: each wave runs on its own.
template <int KIND, typename G>
__global__ __launch_bounds__(256) void k_gf_far(const G* __restrict__ g0, const G* __restrict__ g1,
                                                const G* __restrict__ g2, const G* __restrict__ g3, int ng,
                                                int nin, int nout, const double* __restrict__ h0,
                                                const double* __restrict__ h1, int nt, int B, int Mb, int rc,
                                                int nchunk, double* __restrict__ part) {
    typedef G G4 __attribute__((ext_vector_type(4)));
    const int n_og = (nout + 15) >> 4, nkb = (ng + 15) >> 4;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n_og * nchunk) return;
    const int og = w % n_og, c = w / n_og;
    const int l = threadIdx.x & 63, q = l >> 4, n = l & 15;
    const int s = og * 16 + n;
    const int r0 = c * rc, r1 = min(r0 + rc, nin);
    const int kmax = min(ng - 1, B + Mb - 1);
    const int nkbu = kmax / 16 + 1;  // blocks holding lags 0..kmax
    // U steps of 16 lags per iteration: all their loads are issued before
    // their MFMAs, so one memory round trip covers 16*U lags.  Loads are
    // unconditional (a guarded load becomes a branch with its own wait):
    // block and time are clamped into the buffers and the terms outside the
    // sum are zeroed after.
    constexpr int U = KIND == 2 ? 2 : 4;
    gf_d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    for (int r = r0; r < r1; ++r) {
        const double* hp = h0 + (size_t)r * nt;
        const double* hv = KIND == 2 ? h1 + (size_t)r * nt : hp;
        const size_t tb = gf_tile_index(r, og, 0, n_og, nkb, q, n, 0);
        for (int kb0 = 0; kb0 < nkbu; kb0 += U) {
            double ap[U][4], av[U][4];
            G4 b0[U], b1[U], b2[U], b3[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kb = kb0 + u;
                const size_t o = tb + (size_t)min(kb, nkbu - 1) * 256;
                const bool kok = kb < nkbu;
                const G4 z = {G(0), G(0), G(0), G(0)};
                const G4 y0 = *(const G4*)(g0 + o), y1 = *(const G4*)(g1 + o);
                b0[u] = kok ? y0 : z;
                b1[u] = kok ? y1 : z;
                if constexpr (KIND == 2) {
                    const G4 y2 = *(const G4*)(g2 + o), y3 = *(const G4*)(g3 + o);
                    b2[u] = kok ? y2 : z;
                    b3[u] = kok ? y3 : z;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int tt = B + n - (16 * kb + 4 * q + j);  // history time of A[m = n][kk = q]
                    const bool in = kok && tt >= 0 && tt < B;
                    const int tc = min(max(tt, 0), nt - 1);
                    const double xp = hp[tc];
                    ap[u][j] = in ? xp : 0.0;
                    if constexpr (KIND == 2) {
                        const double xv = hv[tc];
                        av[u][j] = in ? xv : 0.0;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if constexpr (KIND == 2) {
                        acc0 = mfma_f64(ap[u][j], (double)b1[u][j], acc0);  // pv*p
                        acc0 = mfma_f64(av[u][j], (double)b0[u][j], acc0);  // + pp*vn
                        acc1 = mfma_f64(ap[u][j], (double)b3[u][j], acc1);  // vv*p
                        acc1 = mfma_f64(av[u][j], (double)b2[u][j], acc1);  // + vp*vn
                    } else {
                        acc0 = mfma_f64(ap[u][j], (double)b0[u][j], acc0);
                        acc1 = mfma_f64(ap[u][j], (double)b1[u][j], acc1);
                    }
                }
            }
        }
    }
    if (s >= nout) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int m = 4 * v + q;
        if (m < Mb) {
            part[((size_t)(c * GF_M + m) * 2) * nout + s] = acc0[v];
            part[((size_t)(c * GF_M + m) * 2 + 1) * nout + s] = acc1[v];
        }
    }
}

// Sum of the far pass's chunk partials in chunk order, into chunk 0's slots:
// part[(m * 2 + ch) * nout + s], m < Mb.
static __global__ __launch_bounds__(256) void k_gf_far_sum(double* __restrict__ part, int nout, int Mb, int nchunk) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Mb * 2 * nout) return;
    const size_t stride = (size_t)GF_M * 2 * nout;
    double a = part[i];
    int c = 1;
    for (; c + 4 <= nchunk; c += 4) {
        const double x0 = part[i + c * stride], x1 = part[i + (c + 1) * stride];
        const double x2 = part[i + (c + 2) * stride], x3 = part[i + (c + 3) * stride];
        a += x0;
        a += x1;
        a += x2;
        a += x3;
    }
    for (; c < nchunk; ++c) a += part[i + c * stride];
    part[i] = a;
}

// Near pass at step t = B + m: lags 0..min(m, t, ng-1) of every input, plus
// the far sums, written into the source slots at t+1, then vn2v! for those
// transceivers (forward.jl:130-132; the others' runs in k_tx_v).  1024
// threads per output; lane l of each 16-lane group takes lag l%16 of one
// input.  g0..g3 are the maps' first GF_M lags laid out [s][r][GF_M], so a
// workgroup reads one contiguous block of them.
static_assert(GF_M <= 16, "the near pass has 16 lanes per input, one per lag of the block");
template <int KIND, typename T, typename G>
__global__ __launch_bounds__(1024) void k_gf_near(Txs<T> X, const G* __restrict__ g0, const G* __restrict__ g1,
                                                  const G* __restrict__ g2, const G* __restrict__ g3, int ng,
                                                  const int32_t* __restrict__ iSrc, int nout, int nin,
                                                  const double* __restrict__ h0, const double* __restrict__ h1,
                                                  double z0, int t, int m,
                                                  const double* __restrict__ part) {
    const int s = blockIdx.x, tid = threadIdx.x;
    const int kn = min(m, min(t, ng - 1));
    const int k = tid & 15;
    double acc0 = 0.0, acc1 = 0.0;
    if (k <= kn) {
#pragma unroll 4
        for (int r = tid >> 4; r < nin; r += 64) {
            const size_t gb = ((size_t)s * nin + r) * GF_M + k;
            const size_t hb = (size_t)r * X.nt + (t - k);
            const double x0 = h0[hb], x1 = KIND == 2 ? h1[hb] : 0.0;
            if constexpr (KIND == 2) {
                acc0 += (double)g1[gb] * x0 + (double)g0[gb] * x1;  // pv*p + pp*vn
                acc1 += (double)g3[gb] * x0 + (double)g2[gb] * x1;  // vv*p + vp*vn
            } else {
                acc0 += (double)g0[gb] * x0;
                acc1 += (double)g1[gb] * x0;
            }
        }
    }
    __shared__ double red[2][16];
    const double s0 = wave_sum(acc0), s1 = wave_sum(acc1);
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = s0;
        red[1][tid >> 6] = s1;
    }
    __syncthreads();
    if (tid == 0 && t + 1 < X.nt) {
        double a0 = 0.0, a1 = 0.0;
        for (int w = 0; w < 16; ++w) {
            a0 += red[0][w];
            a1 += red[1][w];
        }
        a0 += part[(size_t)(m * 2) * nout + s];
        a1 += part[(size_t)(m * 2 + 1) * nout + s];
        const int is = iSrc[s];
        const size_t o = (size_t)is * X.nt + t + 1;
        T vn;
        if constexpr (KIND == 1 || KIND == 2) {
            // p_src <- vp-derived, vn_src <- pp-derived (extrapolation.jl:165-167,215-216)
            X.src[o] = (T)a1;
            vn = (T)a0;
        } else {
            // PFromPin: p_src = pos - neg, vn_src = (pos + neg)/z0 (extrapolation.jl:246-247)
            X.src[o] = (T)(a0 - a1);
            vn = (T)((a0 + a1) / z0);
        }
        X.vn_src[o] = vn;
        const size_t cs = (size_t)X.ntx * X.nt;
        const T* d = X.dir + 4 * is;
#pragma unroll
        for (int ch = 1; ch <= 3; ++ch) X.src[o + cs * ch] = X.src[o + cs * ch] + vn * d[ch];
    }
}

// ------------------------------------------------------------------ mask
// geometry.jl:86-143 predicates at pressure-grid coordinates, packed into
// the scatterer bit-plane s (1 = solid); then build_update_mask's 4 planes
// (geometry.jl:250-280).  shape: 1 sphere, 2 cube, 3 cross, 4 ellipsoid.
__device__ __forceinline__ bool inside_shape(int shape, const double* prm, double x, double y, double z) {
    if (shape == 1) { const double r = prm[0]; return !(x * x + y * y + z * z > r * r); }
    if (shape == 2) {
        const double r = prm[0];
        return !(fabs(x) > r / 2 || fabs(y) > r / 2 || fabs(z) > r / 2);
    }
    if (shape == 3) {
        const double l = prm[0], s = l / 2, sl = 2.0 * l / 3;
        auto inc = [&](double X, double Y, double Z) { return fabs(X) < s && fabs(Y) < s && fabs(Z) < s; };
        if (!inc(x, y, z)) return false;
        if (inc(x + sl, y + sl, z) || inc(x - sl, y - sl, z) || inc(x + sl, y - sl, z) || inc(x - sl, y + sl, z)) return false;
        if (inc(x + sl, y, z + sl) || inc(x - sl, y, z - sl) || inc(x + sl, y, z - sl) || inc(x - sl, y, z + sl)) return false;
        if (inc(x, y + sl, z + sl) || inc(x, y - sl, z - sl) || inc(x, y + sl, z - sl) || inc(x, y - sl, z + sl)) return false;
        return true;
    }
    if (shape == 4) {
        const double a = prm[0], b = prm[1], c = prm[2];
        return !(x * x / (a * a) + y * y / (b * b) + z * z / (c * c) > 1.0);
    }
    return false;
}

// One thread per 32-cell word of the solid plane.
static __global__ void k_mask_solid(uint32_t* solid, int shape, const double* prm, const double* xc,
                             const double* yc, const double* zc, int nx, int ny, int nz, int nwx) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)nwx * ny * nz;
    if (t >= total) return;
    const int wx = (int)(t % nwx);
    const int j = (int)((t / nwx) % ny);
    const int k = (int)(t / ((size_t)nwx * ny));
    uint32_t bits = 0;
    for (int b = 0; b < 32; ++b) {
        const int i = wx * 32 + b;
        if (i < nx && inside_shape(shape, prm, xc[i], yc[j], zc[k])) bits |= 1u << b;
    }
    solid[t] = bits;
}

// Four update planes from the solid plane (geometry.jl:257-272).
static __global__ void k_mask_planes(uint32_t* planes, const uint32_t* solid, int nx, int ny, int nz, int nwx) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t pl = (size_t)nwx * ny * nz;
    if (t >= pl) return;
    const int wx = (int)(t % nwx);
    const int j = (int)((t / nwx) % ny);
    const int k = (int)(t / ((size_t)nwx * ny));
    const uint32_t valid = (wx * 32 + 32 <= nx) ? 0xffffffffu : ((1u << (nx - wx * 32)) - 1u);
    const uint32_t s = solid[t];
    const uint32_t fl = ~s & valid;
    // fluid bit of cell i-1 for each i of this word
    const uint32_t prevw = wx > 0 ? (~solid[t - 1]) : 0u;
    const uint32_t flm = (fl << 1) | (wx > 0 ? (prevw >> 31) & 1u : 0u);
    uint32_t vx = fl & flm;
    if (wx == 0) vx &= ~1u;  // i == 1 stays closed
    const uint32_t vy = j > 0 ? (fl & ~solid[t - nwx]) : 0u;
    const uint32_t vz = k > 0 ? (fl & ~solid[t - (size_t)nwx * ny]) : 0u;
    planes[t] = fl;
    planes[pl + t] = vx & valid;
    planes[2 * pl + t] = vy & valid;
    planes[3 * pl + t] = vz & valid;
}

}  // namespace adf
