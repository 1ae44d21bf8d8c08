// ops_fp64.hip — the FP64 policy's kernels (ideal and stiffened gases) and
// the device-function entry points with the same parameter block.
#include "ops_impl.hpp"

DIMHIP_INSTANTIATE_POLICY(dimhip::Fp64)
DIMHIP_INSTANTIATE_KAT(dimhip::EosParams)
