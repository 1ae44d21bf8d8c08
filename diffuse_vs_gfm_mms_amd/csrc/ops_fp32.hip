// ops_fp32.hip — the FP32-mixed policy's kernels.
#include "ops_impl.hpp"

DIMHIP_INSTANTIATE_POLICY(dimhip::Fp32Mixed)
