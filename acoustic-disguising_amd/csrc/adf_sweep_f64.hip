// Fused-sweep kernels and launches for double storage and arithmetic (ADF_F64).
#include "adf_sweep_launch.hpp"

namespace adf {
template void launch_fused<double, true>(adf_solver*, const Fused<double>&, int, hipStream_t, const TxSweep<double>*);
template void launch_fused_batch<double, true>(adf_batch*, hipStream_t);
}  // namespace adf
