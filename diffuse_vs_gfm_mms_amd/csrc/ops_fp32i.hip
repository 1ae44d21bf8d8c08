// ops_fp32i.hip — the FP32-mixed policy's kernels for contexts whose
// materials are all ideal gases under the Kapila lambda model.
#include "ops_impl.hpp"

DIMHIP_INSTANTIATE_POLICY(dimhip::Fp32Ideal)
