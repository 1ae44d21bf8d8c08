// kernels_fp32.hpp — mixed-precision (FP32 state, FP64 reductions) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "dim_params.hpp"

namespace dimhip {

struct CflArgs;

template<int D, int M, int DIR>
void launch_sweep32(hipStream_t, const Grid&, const EosParamsF&, const EosParams&, StepCtl*, const CflArgs&, int,
                    bool, const float*, float*)
{
}

template<int D, int M>
__global__ void bc32_kernel(float*, Grid, EosParamsF, EosParams, BcArgs, StepCtl*, int)
{
}

template<int D, int M>
__global__ void cfl32_kernel(const float*, Grid, EosParams, StepCtl*, CflArgs)
{
}

template<int D, int M>
__global__ void cons_totals32_kernel(const float*, Grid, double, long long, double*)
{
}

__global__ void aos_to_soa32_kernel(const double*, float*, long long, long long, int, long long) {}
__global__ void soa32_to_aos_kernel(const float*, double*, long long, long long, int, long long) {}

}  // namespace dimhip
