// Fused-sweep kernels and launches for float storage, float arithmetic (ADF_F32).
#include "adf_sweep_launch.hpp"

namespace adf {
template void launch_fused<float, false>(adf_solver*, const Fused<float>&, int, hipStream_t, const TxSweep<float>*);
template void launch_fused_batch<float, false>(adf_batch*, hipStream_t);
}  // namespace adf
