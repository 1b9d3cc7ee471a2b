// Launches of the fused field-update sweep.  The kernels are instantiated
// per (storage type, arithmetic) in their own translation units
// (adf_sweep_f32.hip, adf_sweep_r32.hip, adf_sweep_f64.hip) so they compile
// in parallel.
#pragma once
#include "adf_internal.hpp"
#include "adf_fused.hpp"

namespace adf {

// One sweep of solve s (old buffers -> new buffers as F describes), cpl
// cells per lane, on stream st (side tiles on the context's side stream
// when s->launch is ADF_SWEEP_TWO_STREAMS).  X: the step's transceiver
// work carried by the launch (one launch, one wave per workgroup), or
// nullptr.
template <typename T, bool EXACT>
void launch_fused(adf_solver* s, const Fused<T>& F, int cpl, hipStream_t st, const TxSweep<T>* X);

// One sweep of every member of batch b on stream st: the side tiles, then
// the interior tiles.  One stream: a batched step is captured in a HIP
// graph, where the two-stream fork/join of the single-solve sweep gave
// intermittently different results on replay.
template <typename T, bool EXACT>
void launch_fused_batch(adf_batch* b, hipStream_t st);

}  // namespace adf
