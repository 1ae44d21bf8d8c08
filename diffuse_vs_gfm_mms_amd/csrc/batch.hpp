// batch.hpp — many independent 2D two-material instances advanced together:
// one launch per stage over a work table of (instance, tile) entries, walked
// by persistent blocks.
#pragma once

#include "policy_fp32.hpp"
#include "sweep.hpp"

namespace dimhip {

template<typename Pol>
struct BatchInst {
    Grid g;
    typename Pol::Eos P;
    BcArgs bc;
    CflArgs cfl;
    void* q[2];
    StepCtl* ctl;
};

// work entry: x = instance, y/z = tile coordinates (sweeps) or ring /
// cell range [y, z) (boundary and CFL passes); edge: a sweep tile next to
// the instance's ring (outputs at index 0, 1, N-2 or N-1 on some axis)
struct Work {
    int inst;
    int a;
    int b;
    int edge;
};

// schedule: 0 = cyclic (block k takes entries k, k+G, ...), 1 = contiguous
// chunks (block k takes [k*chunk, (k+1)*chunk)).  k is the block's logical
// index in XCD order (xcd_block): the blocks that run on one XCD take
// consecutive k, so the tiles one XCD works on at a time are neighbours in
// the work table, whose halo rows and shared cache lines then stay in that
// XCD's L2.
__device__ __forceinline__ void work_range(int nwork, int schedule, int& beg, int& end, int& stride)
{
    int k, unused_y, unused_z;
    xcd_block(k, unused_y, unused_z);
    if (schedule == 0) {
        beg = k;
        end = nwork;
        stride = gridDim.x;
    } else {
        const int chunk = (nwork + gridDim.x - 1) / gridDim.x;
        beg = k * chunk;
        end = min(nwork, beg + chunk);
        stride = 1;
    }
}

// the single-grid sweep's register cap (8 waves per SIMD for the FP32-mixed
// two-material tiles), which the persistent loop's few extra registers fit
// RING: copy-kind boundaries on every instance, whose ring treatment runs
// inside the sweep (sweep_tile)
template<typename Pol, int DIR, int L, int W, bool FUSE, bool RING>
__global__ void __launch_bounds__(L* W) __attribute__((amdgpu_waves_per_eu(Pol::template min_waves_per_eu<2, 2>())))
batch_sweep_kernel(const BatchInst<Pol>* __restrict__ insts, const Work* __restrict__ work, int nwork, int schedule,
                   int cur)
{
    int beg, end, stride;
    work_range(nwork, schedule, beg, end, stride);
    for (int w = beg; w < end; w += stride) {
        const Work it = work[w];
        const BatchInst<Pol>& I = insts[it.inst];
        sweep_tile<Pol, 2, 2, DIR, L, W, 0, FUSE, true, RING>(I.q[cur], I.q[cur ^ 1], I.g, I.P, I.ctl, I.cfl,
                                                               nullptr, nullptr, it.a, it.b, 0, &I.bc, it.edge);
    }
}

// schedule 2: one work entry per block (no persistent loop), entries in
// XCD order as in the single-grid sweeps
template<typename Pol, int DIR, int L, int W, bool FUSE, bool RING>
__global__ void __launch_bounds__(L* W) __attribute__((amdgpu_waves_per_eu(Pol::template min_waves_per_eu<2, 2>())))
batch_tile_kernel(const BatchInst<Pol>* __restrict__ insts, const Work* __restrict__ work, int cur)
{
    int w, unused_y, unused_z;
    xcd_block(w, unused_y, unused_z);
    const Work it = work[w];
    const BatchInst<Pol>& I = insts[it.inst];
    sweep_tile<Pol, 2, 2, DIR, L, W, 0, FUSE, false, RING>(I.q[cur], I.q[cur ^ 1], I.g, I.P, I.ctl, I.cfl, nullptr,
                                                            nullptr, it.a, it.b, 0, &I.bc, it.edge);
}

template<typename Pol>
__global__ void __launch_bounds__(256)
batch_bc_kernel(const BatchInst<Pol>* __restrict__ insts, const Work* __restrict__ work, int nwork, int cur,
                int end_step)
{
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const Work it = work[w];
        const BatchInst<Pol>& I = insts[it.inst];
        const bool active = I.ctl->active;
        for (long long t = it.a + (long long)threadIdx.x; active && t < it.b; t += blockDim.x)
            bc_cell<Pol, 2, 2>(I.q[cur], I.g, I.P, I.bc, t);
        // the entry that starts the instance's ring closes its step
        if (end_step && it.a == 0 && threadIdx.x == 0 && active) end_step_ctl(I.ctl);
    }
}

template<typename Pol>
__global__ void __launch_bounds__(256)
batch_cfl_kernel(const BatchInst<Pol>* __restrict__ insts, const Work* __restrict__ work, int nwork, int cur)
{
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const Work it = work[w];
        const BatchInst<Pol>& I = insts[it.inst];
        double cand = cfl_identity<Pol>();
        for (long long c = it.a + (long long)threadIdx.x; c < it.b; c += blockDim.x) {
            typename Pol::template Cell<2, 2> U;
            Pol::template load<2, 2>(I.q[cur], I.g.pitch, c, U);
            double pc;
            cand = cfl_combine<Pol>(cand, Pol::template cfl_value<2, 2>(I.P, U, I.cfl, pc));
        }
        cfl_block_atomic<Pol, 256>(cand, I.ctl, I.ctl->slot, w);
    }
}

// one 64-thread block per instance: reduce the instance's dt slots
template<typename Pol>
__global__ void __launch_bounds__(64)
batch_begin_step_kernel(const BatchInst<Pol>* __restrict__ insts, int n, int reset_current)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    StepCtl* ctl = insts[i].ctl;
    const int s = ctl->slot;
    const int k = threadIdx.x;
    if (reset_current) {
        dt_slots(ctl, s)[k * kDtStride] = kInfBits;
        if (k == 0) ctl->dtmin[s] = kInfBits;
        return;
    }
    double m = bits_to_double(dt_slots(ctl, s)[k * kDtStride]);
    dt_slots(ctl, s ^ 1)[k * kDtStride] = kInfBits;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = smin(m, __shfl_xor(m, off, 64));
    if (k != 0) return;
    ctl->dtmin[s] = double_to_bits(m);
    int active = ctl->time < ctl->t_end - 1e-14 && !ctl->err;
    const double dt_max = ctl->t_end - ctl->time;
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

}  // namespace dimhip
