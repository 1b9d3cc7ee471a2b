// Definitions of the fused-sweep launches declared in adf_launch.hpp;
// included by one translation unit per (storage type, arithmetic).
#pragma once
#include "adf_launch.hpp"
#include "adf_sweep.hpp"
#include "adf_batch.hpp"

namespace adf {

// No kernel of this library may spill to scratch memory (the Makefile
// fails the build if one does): the compiler's spill code stores 128-bit
// registers to scratch and reloads the same registers at once, which has
// corrupted spilled values (the store's data is read after the next
// instruction has overwritten it).  The pipelined routines fit in Float64
// at 4 cells per lane only with 1-row interior tiles (sweep_ty_int).

// The tiles of list LIST on stream st.  RAG: rows not a multiple of CPL
// long.  X (LIST_ALL only): the step's transceiver work rides in the launch
// (TxSweep), over the solve's signal-marked tile order.
template <typename T, bool EXACT, int CPL, bool RAG, int LIST>
static void launch_list(adf_solver* s, const Fused<T>& F, hipStream_t st, const TxSweep<T>* X = nullptr) {
    const int n = s->ntl[LIST];
    if (!n) return;
    const int W = s->waves ? s->waves : 1;
    dim3 g((unsigned)((n + W - 1) / W));
    const uint2* tl = (const uint2*)s->tiles[LIST];
    TxSweep<T> S{};
    if (X) {
        S = *X;
        tl = s->txs_tiles;
        g = dim3((unsigned)(S.nbt + S.nrb + S.C.nchain));
    }
    // the interior list holds only unmasked interior tiles: its launch
    // takes the kernel that compiles their routine alone
    constexpr bool IO = LIST == LIST_INT;
    if (s->pipeline && s->store_nt > 0) {  // automatic: ordinary stores for a single solve
        hipLaunchKernelGGL((k_sweep<T, EXACT, CPL, true, ST_NT, RAG, IO>), g, dim3(64 * W), 0, st, F, tl, n, S);
        return;
    }
    if (s->pipeline) {
        hipLaunchKernelGGL((k_sweep<T, EXACT, CPL, true, 0, RAG, IO>), g, dim3(64 * W), 0, st, F, tl, n, S);
        return;
    }
    hipLaunchKernelGGL((k_sweep<T, EXACT, CPL, false, 0, RAG>), g, dim3(64 * W), 0, st, F, tl, n, S);
}

// One launch over all tiles, or side and interior tiles as two launches
// (on two streams of equal priority at the same time through fork/join
// events, or one after the other).
template <typename T, bool EXACT, int CPL, bool RAG>
static void launch_fused_cpl(adf_solver* s, const Fused<T>& F, hipStream_t st, const TxSweep<T>* X) {
    if (s->launch == ADF_SWEEP_ONE_LAUNCH) {
        launch_list<T, EXACT, CPL, RAG, LIST_ALL>(s, F, st, X);
        return;
    }
    hipStream_t side = s->ctx->side;
    const bool fork = s->ntl[LIST_SIDE] && s->ntl[LIST_INT] && s->launch == ADF_SWEEP_TWO_STREAMS;
    if (fork) {
        (void)hipEventRecord(s->ev_fork, st);
        (void)hipStreamWaitEvent(side, s->ev_fork, 0);
    }
    launch_list<T, EXACT, CPL, RAG, LIST_SIDE>(s, F, fork ? side : st);
    launch_list<T, EXACT, CPL, RAG, LIST_INT>(s, F, st);
    if (fork) {
        (void)hipEventRecord(s->ev_join, side);
        (void)hipStreamWaitEvent(st, s->ev_join, 0);
    }
}

// Rows of one cell per lane are never ragged.
template <typename T, bool EXACT>
void launch_fused(adf_solver* s, const Fused<T>& F, int cpl, hipStream_t st, const TxSweep<T>* X) {
    const bool rag = F.nx % cpl != 0;
    if (cpl == 4 && rag) launch_fused_cpl<T, EXACT, 4, true>(s, F, st, X);
    else if (cpl == 4) launch_fused_cpl<T, EXACT, 4, false>(s, F, st, X);
    else if (cpl == 2 && rag) launch_fused_cpl<T, EXACT, 2, true>(s, F, st, X);
    else if (cpl == 2) launch_fused_cpl<T, EXACT, 2, false>(s, F, st, X);
    else launch_fused_cpl<T, EXACT, 1, false>(s, F, st, X);
}


template <typename T, bool EXACT, int CPL, bool RAG, int LIST>
static void batch_launch_list(adf_batch* b, const adf_batch::Group& g, hipStream_t st) {
    if (!g.n[LIST]) return;
    const Fused<T>* fs = (const Fused<T>*)b->fs;
    const int pers = b->persistent ? 1 : 0;
    const dim3 blk(64 * g.waves);
    constexpr bool IO = LIST == LIST_INT;  // unmasked interior tiles only
    if (g.pf && g.store_nt) {
        hipLaunchKernelGGL((k_sweep_batch<T, EXACT, CPL, true, RAG, ST_NT, IO>), dim3(g.grid[LIST]), blk, 0, st, fs,
                           b->cur, g.tiles[LIST], g.n[LIST], g.ctr[LIST], pers);
        return;
    }
    if (g.pf) {
        hipLaunchKernelGGL((k_sweep_batch<T, EXACT, CPL, true, RAG, 0, IO>), dim3(g.grid[LIST]), blk, 0, st, fs,
                           b->cur, g.tiles[LIST], g.n[LIST], g.ctr[LIST], pers);
        return;
    }
    hipLaunchKernelGGL((k_sweep_batch<T, EXACT, CPL, false, RAG>), dim3(g.grid[LIST]), blk, 0, st, fs, b->cur,
                       g.tiles[LIST], g.n[LIST], g.ctr[LIST], pers);
}

template <typename T, bool EXACT, int CPL, bool RAG>
static void batch_launch_cpl(adf_batch* b, const adf_batch::Group& g, hipStream_t st) {
    if (g.merged) {
        batch_launch_list<T, EXACT, CPL, RAG, LIST_ALL>(b, g, st);
    } else {
        batch_launch_list<T, EXACT, CPL, RAG, LIST_SIDE>(b, g, st);
        batch_launch_list<T, EXACT, CPL, RAG, LIST_INT>(b, g, st);
    }
}

template <typename T, bool EXACT>
void launch_fused_batch(adf_batch* b, hipStream_t st) {
    for (const auto& g : b->groups) {
        if (g.cpl == 4 && g.rag) batch_launch_cpl<T, EXACT, 4, true>(b, g, st);
        else if (g.cpl == 4) batch_launch_cpl<T, EXACT, 4, false>(b, g, st);
        else if (g.cpl == 2 && g.rag) batch_launch_cpl<T, EXACT, 2, true>(b, g, st);
        else if (g.cpl == 2) batch_launch_cpl<T, EXACT, 2, false>(b, g, st);
        else batch_launch_cpl<T, EXACT, 1, false>(b, g, st);
    }
}

}  // namespace adf
