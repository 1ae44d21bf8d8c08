// kernels_fp64.hpp — step-control kernels and the FP64 known-answer kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "sweep.hpp"

namespace dimhip {

template<int D, int M>
struct Layout {
    static constexpr int NV = 2 * M + D;
    static constexpr int NP = 2 * M + D + 1;
    static constexpr int NF = M + D + 1;
};

// Start of a step (64 threads): the CFL minimum is the min over the current
// parity's dt slots (min is exact in any order); dt = min(dt_max, minimum)
// exactly as the reference's running std::min from dt = dt_max
// (cfl.hpp:43-63); the next parity's slots are cleared.
static __global__ void __launch_bounds__(64) begin_step_kernel(StepCtl* ctl, double dt_max_host)
{
    const int s = ctl->slot;
    const int k = threadIdx.x;
    double m = bits_to_double(dt_slots(ctl, s)[k * kDtStride]);
    dt_slots(ctl, s ^ 1)[k * kDtStride] = kInfBits;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = smin(m, __shfl_xor(m, off, 64));
    if (k != 0) return;
    ctl->dtmin[s] = double_to_bits(m);
    double dt_max = dt_max_host;
    int active = 1;
    if (ctl->use_t_end) {
        active = ctl->time < ctl->t_end - 1e-14;
        dt_max = ctl->t_end - ctl->time;
    }
    // a step that raised an error ends the run, as the reference's exception
    // does: later steps pass the state through
    if (ctl->err) active = 0;
    const double dt = smin(dt_max, m);
    if (active && !(dt > 0.0)) {
        ctl->err |= ERR_NONPOSITIVE_DT;
        active = 0;
    }
    ctl->dt = dt;
    ctl->dt_max = dt_max;
    ctl->active = active;
    ctl->cfl_parity = s ^ 1;
    ctl->dtmin[s ^ 1] = kInfBits;
}

// 64 threads: clear the dt slots of parity slot ^ which_rel
static __global__ void __launch_bounds__(64) reset_dtmin_kernel(StepCtl* ctl, int which_rel)
{
    const int s = ctl->slot ^ which_rel;
    dt_slots(ctl, s)[threadIdx.x * kDtStride] = kInfBits;
    if (threadIdx.x == 0) ctl->dtmin[s] = kInfBits;
}

// 64 threads: StepCtl::dtmin[slot] = min over the current parity's dt slots
static __global__ void __launch_bounds__(64) reduce_dtmin_kernel(StepCtl* ctl)
{
    const int s = ctl->slot;
    double m = bits_to_double(dt_slots(ctl, s)[threadIdx.x * kDtStride]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = smin(m, __shfl_xor(m, off, 64));
    if (threadIdx.x == 0) ctl->dtmin[s] = double_to_bits(m);
}

static __global__ void set_active_kernel(StepCtl* ctl, double dt)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->dt = dt;
        ctl->active = 1;
    }
}

// ---- known-answer kernels (one item per thread) ---------------------------
template<int D, int M>
__device__ __forceinline__ void load_packed(const double* p, State<D, M>& U)
{
    int o = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) U.m[k] = p[o++];
#pragma unroll
    for (int d = 0; d < D; ++d) U.mom[d] = p[o++];
    U.E = p[o++];
#pragma unroll
    for (int k = 0; k < M - 1; ++k) U.a[k] = p[o++];
}

template<int D, int M>
__device__ __forceinline__ void store_packed(double* p, const State<D, M>& U)
{
    int o = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) p[o++] = U.m[k];
#pragma unroll
    for (int d = 0; d < D; ++d) p[o++] = U.mom[d];
    p[o++] = U.E;
#pragma unroll
    for (int k = 0; k < M - 1; ++k) p[o++] = U.a[k];
}

template<int D, int M>
__device__ __forceinline__ void load_packed_prim(const double* p, Prim<D, M>& Q)
{
    int o = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) Q.rho[k] = p[o++];
#pragma unroll
    for (int k = 0; k < M; ++k) Q.alpha[k] = p[o++];
#pragma unroll
    for (int d = 0; d < D; ++d) Q.vel[d] = p[o++];
    Q.p = p[o++];
}

template<int D, int M>
__device__ __forceinline__ void store_packed_prim(double* p, const Prim<D, M>& Q)
{
    int o = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) p[o++] = Q.rho[k];
#pragma unroll
    for (int k = 0; k < M; ++k) p[o++] = Q.alpha[k];
#pragma unroll
    for (int d = 0; d < D; ++d) p[o++] = Q.vel[d];
    p[o++] = Q.p;
}

// EP: EosParams, or EosParamsNA when a material is noble-Abel (the sweeps'
// own choice, so each parameter block's device functions are the ones tested)
template<int D, int M, int DIR, typename EP>
__global__ void kat64_kernel(EP P, int which, int count, const double* __restrict__ in0,
                             const double* __restrict__ in1, double* __restrict__ out0, double* __restrict__ out1,
                             int* __restrict__ err)
{
    using LY = Layout<D, M>;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    if (which == 0) {
        State<D, M> U;
        Prim<D, M> Q;
        load_packed<D, M>(in0 + (long long)t * LY::NV, U);
        cons_to_prim<D, M>(P, U, Q);
        store_packed_prim<D, M>(out0 + (long long)t * LY::NP, Q);
    } else if (which == 1) {
        State<D, M> U;
        Prim<D, M> Q;
        load_packed_prim<D, M>(in0 + (long long)t * LY::NP, Q);
        prim_to_cons<D, M>(P, Q, U);
        store_packed<D, M>(out0 + (long long)t * LY::NV, U);
    } else if (which == 2) {
        State<D, M> U;
        load_packed<D, M>(in0 + (long long)t * LY::NV, U);
        repair_state<D, M>(P, U);
        store_packed<D, M>(out0 + (long long)t * LY::NV, U);
    } else if (which == 3) {
        State<D, M> UL, UR;
        Prim<D, M> PL, PR;
        load_packed<D, M>(in0 + (long long)t * LY::NV, UL);
        load_packed<D, M>(in1 + (long long)t * LY::NV, UR);
        cons_to_prim<D, M>(P, UL, PL);
        cons_to_prim<D, M>(P, UR, PR);
        Flux<D, M> F;
        double fv = 0.0;
        if (const int e = hllc_flux<D, M, DIR>(P, UL, UR, PL, PR, F, fv)) {
            atomicOr(err, e);
            // err[1..2]: the smallest (item, check) key, as a 64-bit value
            atomicMin((unsigned long long*)(err + 2), hllc_err_key(0, 0, t, e));
        }
        double* o = out0 + (long long)t * LY::NF;
        int j = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) o[j++] = F.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) o[j++] = F.mom[d];
        o[j++] = F.E;
        out1[t] = fv;
    } else if (which == 6) {
        // [rho_total, e, alpha[M], rho[M]] -> pressure (out0 and out1)
        const double* r = in0 + (long long)t * (2 + 2 * M);
        out0[t] = pressure_from_internal_energy<M>(P, r[0], r[1], r + 2, r + 2 + M);
        out1[t] = out0[t];
    } else if (which == 4) {
        Prim<D, M> Q;
        load_packed_prim<D, M>(in0 + (long long)t * LY::NP, Q);
        double rhs[M];
        alpha_rhs_coefficients<D, M>(P, Q, rhs);
#pragma unroll
        for (int k = 0; k < M; ++k) out0[(long long)t * M + k] = rhs[k];
    }
}

static __global__ void kat_ieee_kernel(int count, const double* __restrict__ in, double* __restrict__ div_out,
                                double* __restrict__ sqrt_out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const double a = in[2 * t], b = in[2 * t + 1];
    div_out[t] = a / b;
    sqrt_out[t] = __builtin_sqrt(a);
}

}  // namespace dimhip
