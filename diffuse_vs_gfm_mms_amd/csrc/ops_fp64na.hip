// ops_fp64na.hip — the FP64 policy's kernels compiled with the noble-Abel
// law (contexts with a noble-Abel material), the device-function entry points
// with that parameter block, and the driver's input/output row kernels.
#include "ops_impl.hpp"

DIMHIP_INSTANTIATE_POLICY(dimhip::Fp64NA)
DIMHIP_INSTANTIATE_KAT(dimhip::EosParamsNA)
template struct dimhip::IoOps<1, 1>;
template struct dimhip::IoOps<2, 1>;
template struct dimhip::IoOps<3, 1>;
template struct dimhip::IoOps<1, 2>;
template struct dimhip::IoOps<2, 2>;
template struct dimhip::IoOps<3, 2>;
template struct dimhip::IoOps<1, 3>;
template struct dimhip::IoOps<2, 3>;
template struct dimhip::IoOps<3, 3>;
