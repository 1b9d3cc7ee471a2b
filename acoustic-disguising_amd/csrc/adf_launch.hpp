// Launches of the fused field-update sweep.  The kernels are instantiated
// per (storage type, arithmetic) in their own translation units
// (adf_sweep_f32.hip, adf_sweep_r32.hip, adf_sweep_f64.hip) so they compile
// in parallel.
#pragma once
#include "adf_internal.hpp"
#include "adf_fused.hpp"

namespace adf {

// Rows per wave of the fused sweep's tiles: interior tiles (SWEEP_TY, the
// row granularity of the tile split) and side tiles (slab and general
// tiles, whose routines need more registers per row).
constexpr int SWEEP_TY = 2;
constexpr int SWEEP_TY_SIDE = 1;

// One sweep of solve s (old buffers -> new buffers as F describes), cpl
// cells per lane, on stream st (general and slab tiles on the context's
// side stream when s->concurrent).
template <typename T, bool EXACT>
void launch_fused(adf_solver* s, const Fused<T>& F, int cpl, hipStream_t st);

// One sweep of every member of batch b on stream st: the side tiles, then
// the interior tiles.  One stream: a batched step is captured in a HIP
// graph, where the two-stream fork/join of the single-solve sweep gave
// intermittently different results on replay.
template <typename T, bool EXACT>
void launch_fused_batch(adf_batch* b, hipStream_t st);

}  // namespace adf
