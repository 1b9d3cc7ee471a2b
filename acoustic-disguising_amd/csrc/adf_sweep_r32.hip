// Fused-sweep kernels and launches for float storage, the reference's Float32 arithmetic (ADF_F32_REF).
#include "adf_sweep_launch.hpp"

namespace adf {
template void launch_fused<float, true>(adf_solver*, const Fused<float>&, int, hipStream_t, const TxSweep<float>*);
template void launch_fused_batch<float, true>(adf_batch*, hipStream_t);
}  // namespace adf
