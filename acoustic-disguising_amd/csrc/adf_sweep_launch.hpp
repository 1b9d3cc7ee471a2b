// Definitions of the fused-sweep launches declared in adf_launch.hpp;
// included by one translation unit per (storage type, arithmetic).
#pragma once
#include "adf_launch.hpp"
#include "adf_sweep.hpp"
#include "adf_batch.hpp"

namespace adf {

// Interior tiles on the main stream, the side list on the side stream, at
// the same time (fork/join through events; both
// streams default priority).
template <typename T, bool EXACT, int CPL, int TY, int KIND>
static void launch_kind(adf_solver* s, const Fused<T>& F, hipStream_t st) {
    const int n = s->ntl[KIND];
    if (!n) return;
    const int W = s->waves ? s->waves : 1;
    const dim3 g((unsigned)((n + W - 1) / W));
    if (s->pipeline)
        hipLaunchKernelGGL((k_sweep<T, EXACT, CPL, TY, KIND, true>), g, dim3(64 * W), 0, st, F,
                           (const uint2*)s->tiles[KIND], n);
    else
        hipLaunchKernelGGL((k_sweep<T, EXACT, CPL, TY, KIND, false>), g, dim3(64 * W), 0, st, F,
                           (const uint2*)s->tiles[KIND], n);
}

template <typename T, bool EXACT, int CPL, int TY>
static void launch_fused_ty(adf_solver* s, const Fused<T>& F, hipStream_t st) {
    hipStream_t side = s->ctx->side;
    const bool fork = s->ntl[LIST_SIDE] && s->ntl[LIST_INT] && s->concurrent;
    if (fork) {
        (void)hipEventRecord(s->ev_fork, st);
        (void)hipStreamWaitEvent(side, s->ev_fork, 0);
    }
    launch_kind<T, EXACT, CPL, SWEEP_TY_SIDE, LIST_SIDE>(s, F, fork ? side : st);
    launch_kind<T, EXACT, CPL, TY, LIST_INT>(s, F, st);
    if (fork) {
        (void)hipEventRecord(s->ev_join, side);
        (void)hipStreamWaitEvent(st, s->ev_join, 0);
    }
}

template <typename T, bool EXACT>
void launch_fused(adf_solver* s, const Fused<T>& F, int cpl, hipStream_t st) {
    if (cpl == 4) launch_fused_ty<T, EXACT, 4, SWEEP_TY>(s, F, st);
    else if (cpl == 2) launch_fused_ty<T, EXACT, 2, SWEEP_TY>(s, F, st);
    else launch_fused_ty<T, EXACT, 1, SWEEP_TY>(s, F, st);
}


template <typename T, bool EXACT, int CPL, int TY, int KIND>
static void batch_launch_kind(adf_batch* b, const adf_batch::Group& g, hipStream_t st) {
    if (!g.n[KIND]) return;
    const Fused<T>* fs = (const Fused<T>*)b->fs;
    const int pers = b->persistent ? 1 : 0;
    const dim3 blk(64 * g.waves);
    if (g.pf)
        hipLaunchKernelGGL((k_sweep_batch<T, EXACT, CPL, TY, KIND, true>), dim3(g.grid[KIND]), blk, 0, st, fs, b->cur,
                           g.tiles[KIND], g.n[KIND], g.ctr[KIND], pers);
    else
        hipLaunchKernelGGL((k_sweep_batch<T, EXACT, CPL, TY, KIND, false>), dim3(g.grid[KIND]), blk, 0, st, fs, b->cur,
                           g.tiles[KIND], g.n[KIND], g.ctr[KIND], pers);
}

template <typename T, bool EXACT, int CPL, int TY>
static void batch_launch_ty(adf_batch* b, const adf_batch::Group& g, hipStream_t st) {
    batch_launch_kind<T, EXACT, CPL, SWEEP_TY_SIDE, LIST_SIDE>(b, g, st);
    batch_launch_kind<T, EXACT, CPL, TY, LIST_INT>(b, g, st);
}

template <typename T, bool EXACT>
void launch_fused_batch(adf_batch* b, hipStream_t st) {
    for (const auto& g : b->groups) {
        if (g.cpl == 4) batch_launch_ty<T, EXACT, 4, SWEEP_TY>(b, g, st);
        else if (g.cpl == 2) batch_launch_ty<T, EXACT, 2, SWEEP_TY>(b, g, st);
        else batch_launch_ty<T, EXACT, 1, SWEEP_TY>(b, g, st);
    }
}

}  // namespace adf
