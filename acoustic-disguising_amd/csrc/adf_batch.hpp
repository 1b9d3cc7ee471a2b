// Batched engine: many independent solves advanced together, one step at a
// time, with every per-solve quantity that changes between steps (step
// index, ping-pong parity) kept on the device so a step can be captured in a
// hipGraph and replayed.
//
// The field sweep of all member solves runs as one persistent tile-queue
// kernel per cells-per-lane class: a fixed grid of waves takes (solve, tile)
// entries from an atomic counter, general (boundary) tiles first.  The
// transceiver levels, the velocity fixups and the rigid mirrors of all
// members run as one launch each, then k_advance moves every solve to the
// next step.  When every member's transceivers fit the one-launch step,
// they all run in one launch instead (k_tx_step_batch), which also
// advances the members unless some need the separate mirror launch.
#pragma once
#include "adf_sweep.hpp"

namespace adf {

// Transceivers and field buffers of one member solve.
template <typename T>
struct TxDev {
    Txs<T> X;
    T* fb[2];  // the two field buffers (P,Vx,Vy,Vz at c*n3)
    int nx, ny, nz, n3;
    int fuse_mirror;
};

// Tile entries as xcd_tile describes, in (solve, z-block, y-tile) order;
// GEN selects the general or the interior tile routine.  RAG: the group's
// rows are not a multiple of CPL long.
// Non-persistent: W waves per workgroup, block b takes tiles
// xcd_tile(b) * W + w.  Persistent: a fixed grid of waves; a wave on XCD x
// takes tiles from chunk x's queue (ctr[x]) and, when it is empty, from the
// other chunks' queues in turn; the last wave to finish resets the queue
// heads for the next launch.  POL: cache policy of the pipelined routines'
// field stores (0 or ST_NT).
template <typename T, bool EXACT, int CPL, bool PF, bool RAG, int POL = 0, bool INTONLY = false>
__global__ __launch_bounds__(256) void k_sweep_batch(const Fused<T>* __restrict__ fs, const int* __restrict__ cur,
                                                     const uint2* __restrict__ tiles, int ntiles, unsigned* ctr,
                                                     int persistent) {
    // One call site of the tile routine for both modes: two would be
    // compiled as two copies (or as a call with a stack frame).
    const int home = (int)(blockIdx.x & 7);
    int q = 0;  // persistent: queues (home + q) & 7 drained so far
    for (bool first = true;; first = false) {
        int t;
        if (!persistent) {
            if (!first) break;
            const int W = (int)(blockDim.x >> 6);
            t = xcd_tile((int)blockIdx.x, (int)gridDim.x) * W + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
            if (t >= ntiles) break;
        } else {
            t = -1;
            while (q < 8) {
                const int x = (home + q) & 7;
                unsigned v = 0;
                if ((threadIdx.x & 63) == 0) v = atomicAdd(&ctr[x], 1u);
                // wave-uniform: keep the tile index and the per-solve loads scalar
                const int u = __builtin_amdgcn_readfirstlane((int)v);
                if (u < xcd_chunk_size(x, ntiles)) {
                    t = xcd_chunk_start(x, ntiles) + u;
                    break;
                }
                ++q;
            }
            if (t < 0) break;
        }
        const uint2 e = tiles[t];
        const int s = (int)((e.x >> 16) & 0xfffu);
        const Fused<T> F = fs[2 * s + cur[s]];
        sweep_tile<T, EXACT, CPL, PF, POL, RAG, INTONLY>(F, e);
    }
    if (!persistent) return;
    if ((threadIdx.x & 63) == 0) {
        __threadfence();
        const unsigned d = atomicAdd(&ctr[8], 1u);
        if (d == gridDim.x - 1) {
            for (int x = 0; x < 9; ++x) atomicExch(&ctr[x], 0u);
        }
    }
}

// entry = solve << 20 | injector << 19 | transceiver; 8 lanes per entry
// (tx_corners)
template <typename T>
__global__ __launch_bounds__(256) void k_tx_p_batch(const TxDev<T>* __restrict__ td, const int* __restrict__ it0s,
                                                    const int* __restrict__ cur, const uint32_t* __restrict__ ent,
                                                    int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = t >> 3, q = t & 7, g0 = (int)(threadIdx.x & 63) & ~7;
    const bool live = g < n;
    const uint32_t e = live ? ent[g] : 0u;
    const int s = (int)(e >> 20);
    const int is = (int)(e & 0x7ffffu);
    const TxDev<T> D = td[s];
    const Txs<T>& X = D.X;
    T* P = (cur[s] ^ 1) ? D.fb[1] : D.fb[0];  // a select: a runtime index would put D on the stack
    const size_t o = (size_t)is * X.nt + it0s[s];
    tx_corners<T>(P, X.idx + 3 * is, D.nx, D.ny, D.nz, X.w + 8 * is, X.rec + o, X.src + o, ((e >> 19) & 1u) != 0,
                  live, q, g0);
}

// 32 lanes per entry: channels x, y, z on lanes 0-7, 8-15, 16-23 (k_tx_v).
template <typename T>
__global__ __launch_bounds__(256) void k_tx_v_batch(const TxDev<T>* __restrict__ td, const int* __restrict__ it0s,
                                                    const int* __restrict__ cur, const uint32_t* __restrict__ ent,
                                                    int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = t >> 5, ch = (t >> 3) & 3, q = t & 7;
    const int lw = (int)(threadIdx.x & 63), g32 = lw & ~31, g0 = lw & ~7;
    const bool any = g < n;
    const bool live = any && ch < 3;
    const uint32_t e = any ? ent[g] : 0u;
    const int s = (int)(e >> 20);
    const int is = (int)(e & 0x7ffffu);
    const bool inj = ((e >> 19) & 1u) != 0;
    const TxDev<T> D = td[s];
    const Txs<T>& X = D.X;
    const int ntx = X.ntx, nt = X.nt, it0 = it0s[s];
    const size_t o = (size_t)is * nt + it0;
    const size_t cs = (size_t)ntx * nt;
    const int c1 = (ch < 3 ? ch : 2) + 1;
    T* F = ((cur[s] ^ 1) ? D.fb[1] : D.fb[0]) + (size_t)c1 * D.n3;
    const T r = tx_corners<T>(F, X.idx + 3 * (is + ntx * c1), D.nx, D.ny, D.nz, X.w + 8 * (is + ntx * c1),
                              X.rec + o + cs * c1, X.src + o + cs * c1, inj, live, q, g0);
    const T rx = __shfl(r, g32, 64), ry = __shfl(r, g32 + 8, 64), rz = __shfl(r, g32 + 16, 64);
    if (!any) return;
    const T* d = X.dir + 4 * is;
    if (lw == g32) X.vn_rec[o] = rx * d[1] + ry * d[2] + rz * d[3];
    if (inj && it0 + 1 < nt && live && q == 0) {  // vn2v! into it+1 (forward.jl:130-132)
        const T vn = X.vn_src[o + 1];
        X.src[o + 1 + cs * c1] = X.src[o + 1 + cs * c1] + vn * d[c1];
    }
}

// entry.x = solve, entry.y = axis << 30 | cell
template <typename T, bool EXACT>
__global__ void k_fixup_batch(const Fused<T>* __restrict__ fs, const int* __restrict__ cur,
                              const uint2* __restrict__ ent, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint2 e = ent[t];
    const int s = (int)e.x;
    const Fused<T> F = fs[2 * s + cur[s]];
    const int ax = (int)(e.y >> 30);
    const int c = (int)(e.y & 0x3fffffffu);
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

// forward.jl:134-137 for the members whose transceivers touch the mirrored
// planes; blockIdx.y = index into `solves`.
template <typename T>
__global__ void k_mirror_batch(const TxDev<T>* __restrict__ td, const int* __restrict__ cur,
                               const int* __restrict__ solves) {
    const int s = solves[blockIdx.y];
    const TxDev<T> D = td[s];
    T* F = (cur[s] ^ 1) ? D.fb[1] : D.fb[0];
    T* Vy = F + 2 * (size_t)D.n3;
    T* Vz = F + 3 * (size_t)D.n3;
    const int nx = D.nx, ny = D.ny, nz = D.nz;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < max(nx * nz, nx * ny); t += gridDim.x * blockDim.x) {
        if (t < nx * nz) {
            const int i = t % nx, k = t / nx;
            Vy[i + nx * ny * k] = -Vy[i + nx * (1 + ny * k)];
        }
        if (t < nx * ny) Vz[t] = -Vz[t + nx * ny];
    }
}

// Every member's transceiver work of one batched step in one launch: block b
// runs block map[b].y of member map[b].x's one-launch transceiver step
// (k_tx_step: its chain blocks, then its record blocks) on the member's
// current buffers and step.  With adv, the last block to finish moves every
// member to the next step as k_advance does: each block has read its
// member's step and buffer parity (and used them) before it counts itself
// finished, and the next kernels see the new values at the kernel
// boundary, so the count needs no fence (an agent-scope fence per block
// would write back its XCD's L2 each time).
template <typename T, bool EXACT>
__global__ __launch_bounds__(256) void k_tx_step_batch(const Fused<T>* __restrict__ fs,
                                                       const TxDev<T>* __restrict__ td,
                                                       const TxChain* __restrict__ chains,
                                                       const int2* __restrict__ map, int* it0s, int* cur,
                                                       unsigned* done, int nmem, int adv) {
    const int2 m = map[blockIdx.x];
    const int s = m.x;
    const GfHist H{};
    tx_step_block<T, EXACT>(fs[2 * s + cur[s]], td[s].X, chains[s], it0s[s], H, m.y);
    if (!adv) return;
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(done, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    // the members in parallel (one thread walking them would wait a memory
    // round trip per member: the two arrays may alias)
    for (int t = (int)threadIdx.x; t < nmem; t += (int)blockDim.x) {
        it0s[t] += 1;
        cur[t] ^= 1;
    }
    if (threadIdx.x == 0) atomicExch(done, 0u);
}

static __global__ void k_advance(int* it0s, int* cur, int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    it0s[s] += 1;
    cur[s] ^= 1;
}

}  // namespace adf
