// sweep.hpp — the DIM kernels, generic over an arithmetic policy
// (Fp64: the reference's arithmetic bit for bit; Fp32Mixed: FP32 bulk with
// FP64 masses, reductions and interface-cell pressure recovery).
#pragma once

#include <hip/hip_runtime.h>

#include "policy_fp64.hpp"

namespace dimhip {

// LDS record helpers: a per-thread record stored as SoA planes sh[f * S + slot]
template<typename R, int D, int M, template<int, int> class PrT>
__device__ __forceinline__ void put_prim(R* sh, int S, int slot, int off, const PrT<D, M>& P)
{
    int f = off;
#pragma unroll
    for (int k = 0; k < M; ++k) sh[(f++) * S + slot] = P.rho[k];
#pragma unroll
    for (int k = 0; k < M; ++k) sh[(f++) * S + slot] = P.alpha[k];
#pragma unroll
    for (int d = 0; d < D; ++d) sh[(f++) * S + slot] = P.vel[d];
    sh[(f++) * S + slot] = P.p;
}

template<typename R, int D, int M, template<int, int> class PrT>
__device__ __forceinline__ void get_prim(const R* sh, int S, int slot, int off, PrT<D, M>& P)
{
    int f = off;
#pragma unroll
    for (int k = 0; k < M; ++k) P.rho[k] = sh[(f++) * S + slot];
#pragma unroll
    for (int k = 0; k < M; ++k) P.alpha[k] = sh[(f++) * S + slot];
#pragma unroll
    for (int d = 0; d < D; ++d) P.vel[d] = sh[(f++) * S + slot];
    P.p = sh[(f++) * S + slot];
}

template<typename R, int D, int M, template<int, int> class StT>
__device__ __forceinline__ void put_state(R* sh, int S, int slot, int off, const StT<D, M>& U)
{
    int f = off;
#pragma unroll
    for (int k = 0; k < M; ++k) sh[(f++) * S + slot] = U.m[k];
#pragma unroll
    for (int d = 0; d < D; ++d) sh[(f++) * S + slot] = U.mom[d];
    sh[(f++) * S + slot] = U.E;
#pragma unroll
    for (int k = 0; k < M - 1; ++k) sh[(f++) * S + slot] = U.a[k];
}

template<typename R, int D, int M, template<int, int> class StT>
__device__ __forceinline__ void get_state(const R* sh, int S, int slot, int off, StT<D, M>& U)
{
    int f = off;
#pragma unroll
    for (int k = 0; k < M; ++k) U.m[k] = sh[(f++) * S + slot];
#pragma unroll
    for (int d = 0; d < D; ++d) U.mom[d] = sh[(f++) * S + slot];
    U.E = sh[(f++) * S + slot];
#pragma unroll
    for (int k = 0; k < M - 1; ++k) U.a[k] = sh[(f++) * S + slot];
}

// Block-wide min of positive doubles (or +inf) and one atomic per block on
// the bit pattern (positive doubles order like unsigned integers).
// CFL reduction of one block into *target (dt slot, bit-ordered atomicMin of
// a positive double).  A policy contributes either dt candidates combined by
// min (the FP64 form: min_cell min_d cfl dx_d/(|u_d|+c), cfl.hpp:43-63) or
// wave rates combined by max, inverted once per block (the FP32-mixed form).
template<typename Pol>
__device__ __forceinline__ double cfl_identity()
{
    return Pol::kCflRate ? 0.0 : bits_to_double(kInfBits);
}

template<typename Pol>
__device__ __forceinline__ double cfl_combine(double a, double b)
{
    return Pol::kCflRate ? smax(a, b) : smin(a, b);
}

// Non-negative doubles (and +inf) order like their bit patterns as unsigned
// integers, so the wave reduction runs on bits: min (candidates) or max
// (rates); a NaN pattern is above +inf and never wins a min, as in the
// reference's running std::min.
template<bool MAX>
__device__ __forceinline__ unsigned long long u64_combine(unsigned long long a, unsigned long long b)
{
    return MAX ? (a > b ? a : b) : (a < b ? a : b);
}

// v from lane (per DPP control) of this wave; lanes outside row_mask get `idle`
template<int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v, unsigned long long idle)
{
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)idle, (int)(unsigned)v, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(idle >> 32), (int)(unsigned)(v >> 32), CTRL, ROW_MASK,
                                               0xF, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// Reduction over a full wave with DPP moves (quad swaps, half-row and row
// mirrors, row broadcasts 15 and 31): lane 63 ends with the wave's result.
template<bool MAX>
__device__ __forceinline__ unsigned long long wave_reduce_u64(unsigned long long v)
{
    constexpr unsigned long long idle = MAX ? 0ull : ~0ull;
    v = u64_combine<MAX>(v, dpp_u64<0xB1, 0xF>(v, idle));   // quad_perm [1,0,3,2]
    v = u64_combine<MAX>(v, dpp_u64<0x4E, 0xF>(v, idle));   // quad_perm [2,3,0,1]
    v = u64_combine<MAX>(v, dpp_u64<0x141, 0xF>(v, idle));  // row_half_mirror
    v = u64_combine<MAX>(v, dpp_u64<0x140, 0xF>(v, idle));  // row_mirror
    v = u64_combine<MAX>(v, dpp_u64<0x142, 0xA>(v, idle));  // row_bcast:15 into rows 1, 3
    v = u64_combine<MAX>(v, dpp_u64<0x143, 0xC>(v, idle));  // row_bcast:31 into rows 2, 3
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)v, 63);
    const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// Max over a full wave of 32-bit unsigned values: a DPP move and an
// integer max per step (quad swaps, half-row and row mirrors, row broadcasts
// 15 and 31); lane 63 ends with the result, which is returned to every lane.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
    const auto step = [&](unsigned o) { v = v > o ? v : o; };
    step((unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    step((unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    step((unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    step((unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));  // row_mirror
    step((unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));  // row_bcast:15 into rows 1, 3
    step((unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));  // row_bcast:31 into rows 2, 3
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Max over a full wave of non-negative, non-NaN doubles (the wave rates of
// the FP32-mixed CFL: 0 or positive, cfl_value never yields NaN), as their
// bit patterns, which order like unsigned integers: the wave's largest high
// word first, then the largest low word among the lanes that hold it.  The
// maximum is exact, so this equals an FP64 max reduction.
__device__ __forceinline__ unsigned long long wave_max_bits_nonneg(double v)
{
    const unsigned long long b = double_to_bits(v);
    const unsigned hi = wave_max_u32((unsigned)(b >> 32));
    const unsigned lo = wave_max_u32(((unsigned)(b >> 32) == hi) ? (unsigned)b : 0u);
    return ((unsigned long long)hi << 32) | lo;
}

// the wave's combined CFL value as a bit pattern: rates by max, candidates
// by the bit-ordered min
template<typename Pol>
__device__ __forceinline__ unsigned long long wave_cfl(double v)
{
    if constexpr (Pol::kCflRate) return wave_max_bits_nonneg(v);
    else return wave_reduce_u64<false>(double_to_bits(v));
}

template<typename Pol>
__device__ __forceinline__ unsigned long long cfl_combine_bits(unsigned long long a, unsigned long long b)
{
    return u64_combine<Pol::kCflRate>(a, b);
}

// CFL reduction of one block into its dt slots (bit-ordered atomicMin of a
// positive double).  A policy contributes either dt candidates combined by
// min (the FP64 form: min_cell min_d cfl dx_d/(|u_d|+c), cfl.hpp:43-63) or
// wave rates combined by max, inverted once per block (the FP32-mixed form).
// Every thread of the block calls it.
// PER_WAVE: each wave inverts its own result and issues its own atomic, with
// no barrier or LDS; the min over waves of RN(1/r_w) is RN(1/max_w r_w), as
// the reciprocal is monotone, so the slots end with the same minimum.  Used
// by the persistent batched sweeps, whose blocks run many tiles: there the
// two barriers of the block reduction cost more than the extra atomics.
template<typename Pol, int NT, bool PER_WAVE = false>
__device__ __forceinline__ void cfl_block_atomic(double v, StepCtl* ctl, int parity, int blk)
{
    const int tid = threadIdx.x + threadIdx.y * blockDim.x + threadIdx.z * blockDim.x * blockDim.y;
    if constexpr (PER_WAVE) {
        constexpr int NW = NT / 64 > 0 ? NT / 64 : 1;
        const unsigned long long w = wave_cfl<Pol>(v);
        if ((tid & 63) == 0) {
            unsigned long long* target = dt_slots(ctl, parity) + ((blk * NW + (tid >> 6)) % kDtSlots) * kDtStride;
            const double r = bits_to_double(w);
            const double dt = Pol::kCflRate ? (r > 0.0 ? 1.0 / r : bits_to_double(kInfBits)) : r;
            if (dt < bits_to_double(kInfBits)) atomicMin(target, double_to_bits(dt));
        }
        return;
    }
    unsigned long long* target = dt_slots(ctl, parity) + (blk % kDtSlots) * kDtStride;
    __shared__ unsigned long long red[NT / 64 > 0 ? NT / 64 : 1];
    __syncthreads();  // the block may reuse red[] for its next tile
    const unsigned long long w = wave_cfl<Pol>(v);
    if ((tid & 63) == 0) red[tid >> 6] = w;
    __syncthreads();
    if (tid == 0) {
        unsigned long long m = red[0];
#pragma unroll
        for (int k = 1; k < NT / 64; ++k) m = cfl_combine_bits<Pol>(m, red[k]);
        const double r = bits_to_double(m);
        const double dt = Pol::kCflRate ? (r > 0.0 ? 1.0 / r : bits_to_double(kInfBits)) : r;
        if (dt < bits_to_double(kInfBits)) atomicMin(target, double_to_bits(dt));
    }
}

// MI355X dispatches workgroups round-robin over its 8 XCDs, each with its own
// L2.  Renumber blocks so that consecutive logical blocks (neighbouring tiles,
// which share cache lines and halo cells) run on the same XCD:
// physical block p is the (p/8)-th block of XCD p%8 and gets logical index
// start(p%8) + p/8, a bijection for any grid size.
__device__ __forceinline__ void xcd_block(int& bx, int& by, int& bz)
{
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const unsigned total = gx * gy * gz;
    const unsigned p = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned x = p & 7u, q = total >> 3, r = total & 7u;
    const unsigned l = x * q + (x < r ? x : r) + (p >> 3);
    bx = (int)(l % gx);
    by = (int)((l / gx) % gy);
    bz = (int)(l / (gx * gy));
}

// The same per-XCD logical index, walked in column strips: strip s holds
// block columns [s*band, (s+1)*band) (the last strip what remains), each strip
// row by row.  With band = ceil(gx/8) each XCD's share of a 2D grid is about
// one strip, so a y tile's halo rows, which the tiles above and below it read
// too, stay in that XCD's L2.  A bijection for any grid and band >= 1.
__device__ __forceinline__ void xcd_block_banded(int& bx, int& by, int& bz, unsigned band)
{
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const unsigned total = gx * gy * gz;
    const unsigned p = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned x = p & 7u, q = total >> 3, r = total & 7u;
    const unsigned l = x * q + (x < r ? x : r) + (p >> 3);
    const unsigned plane = gx * gy;
    const unsigned z = l / plane;
    const unsigned m = l - z * plane;
    const unsigned full = gx / band;
    const unsigned fc = full * band * gy;
    unsigned ox, oy;
    if (m < fc) {
        const unsigned t = m / band;
        const unsigned b = t / gy;
        oy = t - b * gy;
        ox = b * band + (m - t * band);
    } else {
        const unsigned w = gx - full * band, rr = m - fc;
        oy = rr / w;
        ox = full * band + (rr - oy * w);
    }
    bx = (int)ox;
    by = (int)oy;
    bz = (int)z;
}

// ---------------------------------------------------------------------------
// Ring boundary treatment fused into a sweep (boundary.hpp:165-226 for the
// copy kinds: transmissive, reflection, reflective).  Every ring cell (index
// 0 or N-1 on an axis longer than 1) takes the state of its source cell, the
// ring cell's index clamped to [1, N-2] on every such axis, with the normal
// momentum of each reflective side it lies on negated.  With extent >= 3 on
// every axis longer than 1 the source is not a ring cell, and each ring cell
// is reached from exactly one source by moving some of the source's axes
// from index 1 to 0 or from N-2 to N-1.  So the thread that computes a source
// cell also stores the ring cells it feeds (one per non-empty set of such
// moves), and ring cells are not stored by the sweep itself: the result is
// the separate boundary pass's, without its launch or its re-read.
// pout: the recovered-pressure plane, whose ring entries copy the source's.
// ---------------------------------------------------------------------------
struct BcArgs {
    int bc_lo[3];
    int bc_hi[3];
    long long face_start[7];  // prefix sums of ring cells per face (d, side) = 2d+side
};

template<int D>
__device__ __forceinline__ bool ring_cell(const Grid& g, const int* idx)
{
    bool r = false;
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (g.N[d] > 1 && (idx[d] == 0 || idx[d] == g.N[d] - 1)) r = true;
    return r;
}

// axes on which the cell at idx is the source of the ring cell at index 0
// (lo) or N-1 (hi)
template<int D>
__device__ __forceinline__ void ring_sources(const Grid& g, const int* idx, int& lo, int& hi)
{
    lo = 0;
    hi = 0;
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (g.N[d] > 1) {
            if (idx[d] == 1) lo |= 1 << d;
            if (idx[d] == g.N[d] - 2) hi |= 1 << d;
        }
}

// The ring cells a source feeds: per axis d the source's index is kept or
// moved to the ring (to 0 where lo has bit d, to N-1 where hi has it; both
// only for an extent of 3), so with r_d = 1 + lo_d + hi_d choices per axis
// the source feeds prod_d r_d - 1 ring cells (usually 1; 3 at a 2D corner, 7
// at a 3D corner).  Number c in [1, prod r_d) names one: digit d of c in
// radix r_d, 0 keeping the index, 1 the first move, 2 the move to N-1.
// Returns its linear index and the axes whose momentum the copy negates.
template<int D>
__device__ __forceinline__ int ring_count(int lo, int hi)
{
    int n = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) n *= 1 + ((lo >> d) & 1) + ((hi >> d) & 1);
    return n;
}

template<int D>
__device__ __forceinline__ long long ring_target(const Grid& g, const BcArgs& bc, const int* idx, int lo, int hi,
                                                 int c, int& flip)
{
    long long lin = 0;
    flip = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int l = (lo >> d) & 1, h = (hi >> d) & 1;
        int digit = 0;
        if (l + h == 1) {
            digit = c & 1;
            c >>= 1;
        } else if (l + h == 2) {
            digit = c % 3;
            c /= 3;
        }
        int id = idx[d];
        if (digit != 0) {
            const bool to_lo = digit == 1 && l;
            id = to_lo ? 0 : g.N[d] - 1;
            if ((to_lo ? bc.bc_lo[d] : bc.bc_hi[d]) == BC_REFLECTIVE) flip |= 1 << d;
        }
        lin += (long long)id * g.stride[d];
    }
    return lin;
}

// the ring cells a source feeds get its state
// an x-face ring cell whose other indices are interior: its source is its
// x-neighbour, which the sweep's tile holds in the next lane of the wave
template<int D>
__device__ __forceinline__ bool xring_cell(const Grid& g, const int* idx)
{
    bool r = g.N[0] > 1 && (idx[0] == 0 || idx[0] == g.N[0] - 1);
#pragma unroll
    for (int d = 1; d < D; ++d)
        if (g.N[d] > 1 && (idx[d] == 0 || idx[d] == g.N[d] - 1)) r = false;
    return r;
}

// a tile holding the x-face outputs x = 0 or x = N0-1 (block-uniform)
template<int DIR, int T, int W>
__device__ __forceinline__ bool xring_tile(const Grid& g, int bx)
{
    const int span = DIR == 0 ? T : W;
    return bx == 0 || (bx + 1) * span >= g.N[0];
}

template<typename Pol, int D, int M>
__device__ __forceinline__ void ring_copies(void* __restrict__ q, const Grid& g, const BcArgs& bc, const int* idx,
                                            const typename Pol::template Cell<D, M>& r)
{
    using Cell = typename Pol::template Cell<D, M>;
    int lo, hi;
    ring_sources<D>(g, idx, lo, hi);
    const int n = ring_count<D>(lo, hi);
    for (int c = 1; c < n; ++c) {
        int flip;
        const long long lin = ring_target<D>(g, bc, idx, lo, hi, c, flip);
        Cell st = r;
#pragma unroll
        for (int d = 0; d < D; ++d)
            if ((flip >> d) & 1) st.mom[d] *= -1;
        Pol::template store<D, M>(q, g.pitch, lin, st);
    }
}

// ... and its recovered pressure
template<int D>
__device__ __forceinline__ void ring_pressures(double* __restrict__ pout, const Grid& g, const BcArgs& bc,
                                               const int* idx, double pc)
{
    int lo, hi;
    ring_sources<D>(g, idx, lo, hi);
    const int n = ring_count<D>(lo, hi);
    for (int c = 1; c < n; ++c) {
        int flip;
        pout[ring_target<D>(g, bc, idx, lo, hi, c, flip)] = pc;
    }
}

__device__ __forceinline__ void end_step_ctl(StepCtl* ctl)
{
    ctl->time += ctl->dt;
    ctl->step += 1;
    ctl->slot ^= 1;
}

// ---------------------------------------------------------------------------
// Directional sweep: sweep_direction_dispatch + solve_line
// (sweep_core.hpp:16-202, muscl.hpp:197-329, hllc.hpp:62-195).
//
// A block owns W lines x L positions along the sweep axis; positions
// [3, L-3) are written (T = L-6 cells per tile); the 3-cell halo on each side
// is the reference's stencil radius.  Phases, each followed by an LDS
// exchange with the neighbouring positions:
//   1  P^n = cons_to_prim(U^n); P_line = enforce_positive(P^n)
//   2  face primitives from P_line[i-1..i+1], Hancock half step -> U_half, P_half
//   3  face primitives from P_half -> time-centred face states and their
//      primitives (the HLLC inputs)
//   4  HLLC flux on face i+1/2 and the upwind alpha flux
//   5  conservative update, repair, non-conservative alpha update, repair
// Cells 0 and n-1 of a line are first order and unpredicted and the end
// faces copy their neighbours, as in the reference.
//
// MODE 0: sweep qin -> qout.  MODE 1: unsplit accumulation
// qout += (sweep(qin) - qin), then repair (advance_step.hpp:198-222).
// FUSE_CFL: reduce next step's CFL candidate over interior cells too (ring
// cells become copies of interior cells with |u| preserved in the boundary
// pass, so the interior minimum is the next step's dt).
// pin (Pol::kStoredPressure): the recovered pressure of every input cell,
// stored by the CFL pass that read the same state; phase 1 takes it instead
// of repeating the bisection.  pout (with FUSE_CFL): the recovered pressure
// of every interior output cell, for the next sweep.
// PERSISTENT: the block runs many tiles (the batched launches).  Its CFL
// epilogue then reduces per wave (cfl_block_atomic's PER_WAVE).
// RING (MODE 0, copy-kind boundaries bc only; the batched launches): the
// ring boundary treatment is fused (ring_copies) in the tiles the caller
// marks with edge_tile; a fused last sweep
// (FUSE_CFL) also closes the step (time, step count, dt-slot parity), which
// no sweep block reads: the CFL epilogue takes its parity from
// ctl->cfl_parity, set when the step began.
// ---------------------------------------------------------------------------
template<typename Pol, int D, int M, int DIR, int L, int W, int MODE, bool FUSE_CFL, bool PERSISTENT = false,
         bool RING = false>
__device__ __forceinline__ void sweep_tile(const void* __restrict__ qin, void* __restrict__ qout, const Grid& g,
                                           const typename Pol::Eos& P, StepCtl* __restrict__ ctl,
                                           const CflArgs& cflargs, const double* __restrict__ pin,
                                           double* __restrict__ pout, int bx, int by, int bz,
                                           const BcArgs* __restrict__ bc = nullptr, int edge_tile = 0)
{
    static_assert(!RING || MODE == 0, "the fused ring treatment belongs to split sweeps");
    using R = typename Pol::R;
    using St = typename Pol::template St<D, M>;
    using Pr = typename Pol::template Pr<D, M>;
    using Fx = typename Pol::template Fx<D, M>;
    using Cell = typename Pol::template Cell<D, M>;
    constexpr int S = L * W;
    constexpr int NV = 2 * M + D;
    constexpr int NPR = 2 * M + D + 1;  // primitive record: rho[M], alpha[M], vel[D], p
    constexpr int T = L - 6;
    // LDS records, SoA rows of S slots.  Pol::ping_pong: consecutive
    // exchanges alternate between two regions, A (P_line, then the face
    // records) and B (the half-step primitives, then the fluxes), so a phase
    // writes its output as soon as it is computed -- the region it overwrites
    // was last read before the previous barrier -- and each exchange takes one
    // barrier.  The face record is then the face primitive with the face
    // state's momentum and energy (its partial masses and fraction code are
    // rebuilt from the primitive), or with Pol::prim_faces the face primitive
    // alone (the state rebuilt from it by prim_to_cons).  Otherwise one region
    // holds every exchange, the full face state and primitive, with a barrier
    // on each side of every write.
    constexpr bool PP = Pol::template ping_pong<D, M, DIR>();
    constexpr bool PRIM_FACES = PP && Pol::template prim_faces<D, M, DIR>();
    constexpr int NFACE = PP ? (PRIM_FACES ? NPR : NPR + D + 1) : NV + NPR;
    constexpr int NFLX = 2 * M + D + 2;  // F.m[M], F.mom[D], F.E, face velocity, alpha flux[M]
    constexpr int OFF_B = PP ? (NPR > NFACE ? NPR : NFACE) : 0;
    constexpr int NREC = PP ? OFF_B + (NPR > NFLX ? NPR : NFLX) : NV + NPR;
    // Pol::kParkAlpha: P^n's fractions and rhs_n wait in NPARK extra LDS rows
    // of the thread's own slot from phase 1 to phase 5 (for nmat <= 2, whose
    // tiles still fit the occupancy the register cap allows; not with the
    // two regions, whose records already fill that budget)
    constexpr bool PARK = Pol::template park<D, M, DIR>();
    constexpr int NPARK = PARK ? 2 * M : 0;
    __shared__ R sh[(NREC + NPARK) * S];

    if (PERSISTENT) __syncthreads();  // the block's previous tile may still read the LDS
    const typename Pol::Factors fac = Pol::factors(ctl->dt, g.dx[DIR], cflargs.inv_dx[DIR]);

    int pos, lane;
    if (DIR == 0) {
        pos = threadIdx.x;
        lane = threadIdx.y;
    } else {
        lane = threadIdx.x;
        pos = threadIdx.y;
    }
    const int slot = (DIR == 0) ? lane * L + pos : pos * W + lane;
    const int slot_l = (DIR == 0) ? slot - 1 : slot - W;
    const int slot_r = (DIR == 0) ? slot + 1 : slot + W;

    const int n = g.N[DIR];
    int idx[3] = {0, 0, 0};
    bool line_ok;
    int i;
    if (DIR == 0) {
        i = bx * T - 3 + pos;
        idx[1] = by * W + lane;
        idx[2] = bz;
        line_ok = idx[1] < g.N[1];
    } else if (DIR == 1) {
        idx[0] = bx * W + lane;
        i = by * T - 3 + pos;
        idx[2] = bz;
        line_ok = idx[0] < g.N[0];
    } else {
        idx[0] = bx * W + lane;
        idx[1] = bz;
        i = by * T - 3 + pos;
        line_ok = idx[0] < g.N[0];
    }
    idx[DIR] = i;
    const bool in = line_ok && i >= 0 && i < n;
    const long long cell = (long long)idx[0] + (long long)idx[1] * g.stride[1] + (long long)idx[2] * g.stride[2];
    // x and y sweeps address the tile's cells as cbase (uniform across the
    // block) + crel (a 32-bit offset below 2 * stride1 + L or L * stride1),
    // so loads and stores take a scalar base and a 32-bit lane offset; z
    // sweeps, whose tiles span L planes, keep 64-bit cell indices
    constexpr bool REL = DIR < 2;
    long long cbase = 0;
    unsigned crel = 0;
    if (DIR == 0) {
        cbase = (long long)(bx * T - 3) + (long long)(by * W) * g.stride[1] + (long long)bz * g.stride[2];
        crel = (unsigned)pos + (unsigned)lane * (unsigned)g.stride[1];
    } else if (DIR == 1) {
        cbase = (long long)(bx * W) + (long long)(by * T - 3) * g.stride[1] + (long long)bz * g.stride[2];
        crel = (unsigned)lane + (unsigned)pos * (unsigned)g.stride[1];
    }
    const auto load_cell = [&](const void* q, Cell& U) {
        if constexpr (REL) Pol::template load_rel<D, M>(q, g.pitch, cbase, crel, U);
        else Pol::template load<D, M>(q, g.pitch, cell, U);
    };
    const auto store_cell = [&](void* q, const Cell& U) {
        if constexpr (REL) Pol::template store_rel<D, M>(q, g.pitch, cbase, crel, U);
        else Pol::template store<D, M>(q, g.pitch, cell, U);
    };
    const bool end_cell = (i == 0) || (i == n - 1);
    const bool own = in && pos >= 3 && pos < L - 3;
    // the fused ring treatment acts only in tiles whose outputs reach index
    // 0, 1, N-2 or N-1 on some axis (edge_tile, from the caller's work table)
    const bool edge = RING && edge_tile != 0;
    // x-face ring cells taken from the x-neighbour's lane (cflargs.xring;
    // 3D grids, whose x faces cost the ring pass a cache line per cell and
    // field)
    constexpr bool XRING = D == 3 && MODE == 0 && !RING;
    const bool xedge = XRING && cflargs.xring && xring_tile<DIR, T, W>(g, bx);
    // y-face ring rows written by the 2D x sweep (cflargs.yring), in the
    // tiles whose lines include y = 0, 1, N1-2 or N1-1
    constexpr bool YRING = D == 2 && DIR == 0 && MODE == 0 && !RING;
    const bool yedge = YRING && cflargs.yring && (by * W <= 1 || (by + 1) * W >= g.N[1] - 1);

    if (!ctl->active) {
        // a step past t_end (or after an error): the state passes through
        // unchanged.  The host flips the double buffer once per sweep, D
        // times per split step; with D even the flips return to the input
        // buffer, which no inactive sweep writes, so nothing is copied.  With
        // D odd the first sweep copies and the others, each reading or
        // writing a buffer that already holds the state, need not.
        if (MODE == 0 && D % 2 == 1 && DIR == 0 && own) {
            Cell U;
            load_cell(qin, U);
            store_cell(qout, U);
        }
        return;
    }
    if (RING && FUSE_CFL && bx == 0 && by == 0 && bz == 0 && threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0)
        end_step_ctl(ctl);

    // phase 1
    Cell Uc;
    St U;
    Pr Pn, Pl;
    R rhs_n[M];  // alpha_rhs_coefficients(P^n), used by the update
    if (in) {
        load_cell(qin, Uc);
        Pol::template to_state<D, M>(Uc, U);
        if constexpr (Pol::kStoredPressure) {
            if (pin) Pol::template cons_to_prim_stored<D, M>(P, U, pin[cell], Pn);
            else Pol::template cons_to_prim<D, M>(P, U, Pn);
        } else {
            Pol::template cons_to_prim<D, M>(P, U, Pn);
        }
        Pl = Pn;
        Pol::template enforce_c2p<D, M>(Pl);
        Pol::template alpha_rhs<D, M>(P, Pn, rhs_n);
        put_prim<R, D, M>(sh, S, slot, 0, Pl);
        if constexpr (PARK) {
#pragma unroll
            for (int k = 0; k < M; ++k) {
                sh[(NREC + k) * S + slot] = Pn.alpha[k];
                sh[(NREC + M + k) * S + slot] = rhs_n[k];
            }
        }
    }
    __syncthreads();

    // phase 2
    St Uh;
    Pr Ph;
    const bool ph2 = in && pos >= 1 && pos < L - 1;
    if (ph2) {
        if constexpr (Pol::kReloadInput) {
            // the cell's state and P_line are read back (global, an L2 hit;
            // the thread's own LDS slot) instead of held across the barrier
            Cell U2;
            load_cell(qin, U2);
            Pol::template to_state<D, M>(U2, U);
            get_prim<R, D, M>(sh, S, slot, 0, Pl);
        }
        if (end_cell) {
            Uh = U;
            Ph = Pl;
        } else {
            Pr PL, PR, lf, rf;
            get_prim<R, D, M>(sh, S, slot_l, 0, PL);
            get_prim<R, D, M>(sh, S, slot_r, 0, PR);
            Pol::template reconstruct<D, M>(PL, Pl, PR, lf, rf);
            St ul, ur;
            Pol::template face_cons<D, M>(P, lf, ul);
            Pol::template face_cons<D, M>(P, rf, ur);
            R rhs_l[M];
            Pol::template alpha_rhs_line<D, M>(P, Pl, rhs_n, rhs_l);
            Pol::template hancock<D, M, DIR>(P, U, Pl, ul, ur, lf, rf, rhs_l, fac, Uh);
            Pol::template cons_to_prim<D, M>(P, Uh, Ph);
            Pol::template enforce_c2p<D, M>(Ph);
        }
        if (PP) put_prim<R, D, M>(sh, S, slot, OFF_B, Ph);
    }
    if (!PP) {
        __syncthreads();
        if (ph2) put_prim<R, D, M>(sh, S, slot, 0, Ph);
    }
    __syncthreads();

    // phase 3
    St Ulh, Urh;
    Pr PUl, PUr;
    const bool ph3 = in && pos >= 2 && pos < L - 2;
    if (ph3) {
        if (end_cell) {
            // an end cell's faces are the cell itself (UL = UR = U^n and its
            // primitive); with Pol::kReloadInput rebuilt from memory rather
            // than held in registers since phase 1
            if constexpr (Pol::kReloadInput) {
                Cell Ue;
                load_cell(qin, Ue);
                Pol::template to_state<D, M>(Ue, Ulh);
                Pol::template cons_to_prim<D, M>(P, Ulh, PUl);
            } else {
                Ulh = U;
                PUl = Pn;
            }
            Urh = Ulh;
            PUr = PUl;
        } else {
            Pr PL, PR, lf, rf;
            get_prim<R, D, M>(sh, S, slot_l, OFF_B, PL);
            get_prim<R, D, M>(sh, S, slot_r, OFF_B, PR);
            Pol::template reconstruct<D, M>(PL, Ph, PR, lf, rf);
            Pol::template face_state<D, M>(P, lf, Ulh, PUl);
            Pol::template face_state<D, M>(P, rf, Urh, PUr);
        }
        if constexpr (PP) {
            put_prim<R, D, M>(sh, S, slot, 0, PUl);
            if constexpr (!PRIM_FACES) {
#pragma unroll
                for (int d = 0; d < D; ++d) sh[(NPR + d) * S + slot] = Ulh.mom[d];
                sh[(NPR + D) * S + slot] = Ulh.E;
            }
        }
    }
    if (!PP) {
        __syncthreads();
        if (ph3) {
            put_state<R, D, M>(sh, S, slot, 0, Ulh);
            put_prim<R, D, M>(sh, S, slot, NV, PUl);
        }
    }
    __syncthreads();

    // phase 4: face i+1/2
    Fx F;
    R fv = R(0);
    R af[M];
    const bool ph4 = in && pos >= 2 && pos < L - 3 && i <= n - 2;
    if (ph4) {
        St UR;
        Pr PR;
        if constexpr (PRIM_FACES) {
            get_prim<R, D, M>(sh, S, slot_r, 0, PR);
            Pol::template face_from_prim<D, M>(P, PR, UR);
        } else if constexpr (PP) {
            get_prim<R, D, M>(sh, S, slot_r, 0, PR);
            R mom[D];
#pragma unroll
            for (int d = 0; d < D; ++d) mom[d] = sh[(NPR + d) * S + slot_r];
            Pol::template face_from_record<D, M>(PR, mom, sh[(NPR + D) * S + slot_r], UR);
        } else {
            get_state<R, D, M>(sh, S, slot_r, 0, UR);
            get_prim<R, D, M>(sh, S, slot_r, NV, PR);
        }
        if (const int e = Pol::template hllc<D, M, DIR>(P, Urh, UR, PUr, PR, F, fv)) {
            // the reference's line index: the other axes ascending, last fastest
            long long line = 0;
#pragma unroll
            for (int d = 0; d < D; ++d)
                if (d != DIR) line = line * g.N[d] + idx[d];
            atomicOr(&ctl->err, e);
            atomicMin(&ctl->err_key, hllc_err_key(DIR, line, i, e));
        }
        const R* up = (fv >= R(0)) ? PUr.alpha : PR.alpha;
#pragma unroll
        for (int k = 0; k < M; ++k) af[k] = up[k] * fv;
    }
    const auto put_flux = [&]() {
        int f = OFF_B;
#pragma unroll
        for (int k = 0; k < M; ++k) sh[(f++) * S + slot] = F.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) sh[(f++) * S + slot] = F.mom[d];
        sh[(f++) * S + slot] = F.E;
        sh[(f++) * S + slot] = fv;
#pragma unroll
        for (int k = 0; k < M; ++k) sh[(f++) * S + slot] = af[k];
    };
    if (PP) {
        if (ph4) put_flux();
    } else {
        __syncthreads();
        if (ph4) put_flux();
    }
    __syncthreads();

    // phase 5: update cell i
    double cand = cfl_identity<Pol>();
    if (own) {
        if constexpr (PARK) {
#pragma unroll
            for (int k = 0; k < M; ++k) {
                Pn.alpha[k] = sh[(NREC + k) * S + slot];
                rhs_n[k] = sh[(NREC + M + k) * S + slot];
            }
        }
        // with Pol::kReloadInput the cell's input state is read again (an L2
        // hit) rather than held in registers across the four phases
        Cell Uin;
        if constexpr (Pol::kReloadInput) load_cell(qin, Uin);
        else Uin = Uc;
        Cell r = Uin;
        if (n > 1) {
            Fx Fl, Fr;
            R fvl, fvr, afl[M], afr[M];
            if (i <= n - 2) {
                Fr = F;
                fvr = fv;
#pragma unroll
                for (int k = 0; k < M; ++k) afr[k] = af[k];
            }
            if (i >= 1) {
                int f = OFF_B;
#pragma unroll
                for (int k = 0; k < M; ++k) Fl.m[k] = sh[(f++) * S + slot_l];
#pragma unroll
                for (int d = 0; d < D; ++d) Fl.mom[d] = sh[(f++) * S + slot_l];
                Fl.E = sh[(f++) * S + slot_l];
                fvl = sh[(f++) * S + slot_l];
#pragma unroll
                for (int k = 0; k < M; ++k) afl[k] = sh[(f++) * S + slot_l];
            }
            if (i == 0) {  // F[0] = F[1]; alpha_flux[0] = alpha_full[0] * fv[0]
                Fl = Fr;
                fvl = fvr;
#pragma unroll
                for (int k = 0; k < M; ++k) afl[k] = Pn.alpha[k] * fvl;
            }
            if (i == n - 1) {  // F[n] = F[n-1]; alpha_flux[n] = alpha_full[n-1] * fv[n]
                Fr = Fl;
                fvr = fvl;
#pragma unroll
                for (int k = 0; k < M; ++k) afr[k] = Pn.alpha[k] * fvr;
            }
            Pol::template update<D, M>(P, Uin, Pn, rhs_n, Fl, Fr, fvl, fvr, afl, afr, fac, r);
        }
        if (MODE == 1) {
            Cell acc;
            load_cell(qout, acc);
            Pol::template accumulate<D, M>(P, acc, r, Uin);
            r = acc;
        }
        // x-face ring cells: the source's state from the neighbouring lane
        // (every lane of the tile's own range is active here; the sources
        // are own cells of the same wave)
        bool xr = false;
        int xsrc = 0;
        if (XRING && xedge) {
            const int me = __lane_id();
            xr = xring_cell<D>(g, idx);
            xsrc = xr ? (idx[0] == 0 ? me + 1 : me - 1) : me;
            Cell s;
#pragma unroll
            for (int k = 0; k < M; ++k) s.m[k] = __shfl(r.m[k], xsrc);
#pragma unroll
            for (int d = 0; d < D; ++d) s.mom[d] = __shfl(r.mom[d], xsrc);
            s.E = __shfl(r.E, xsrc);
#pragma unroll
            for (int k = 0; k < M - 1; ++k) s.a[k] = __shfl(r.a[k], xsrc);
            if (xr) {
                if ((cflargs.xflip >> (idx[0] == 0 ? 0 : 1)) & 1) s.mom[0] *= -1;
                r = s;
            }
        }
        // y-face ring rows (2D x sweep): rows 1 and N1-2 store their copies
        // for interior x, and the ring rows store nothing themselves
        bool yskip = false;
        if (YRING && yedge) {
            const int y = idx[1];
            yskip = y == 0 || y == g.N[1] - 1;
            if (idx[0] >= 1 && idx[0] <= g.N[0] - 2) {
                if (y == 1) {
                    Cell s = r;
                    if (cflargs.yflip & 1) s.mom[1] *= -1;
                    Pol::template store<D, M>(qout, g.pitch, cell - g.stride[1], s);
                }
                if (y == g.N[1] - 2) {
                    Cell s = r;
                    if (cflargs.yflip & 2) s.mom[1] *= -1;
                    Pol::template store<D, M>(qout, g.pitch, cell + g.stride[1], s);
                }
            }
        }
        // with the fused ring treatment a ring cell's value comes from its
        // source's ring_copies; only tiles next to the ring hold either
        const bool ring_out = edge && ring_cell<D>(g, idx);
        if (!ring_out && !yskip) store_cell(qout, r);
        if (edge && !ring_out) ring_copies<Pol, D, M>(qout, g, *bc, idx, r);
        if (FUSE_CFL) {
            double pc = 0.0;
            if (!ring_cell<D>(g, idx)) {
                cand = Pol::template cfl_value<D, M>(P, r, cflargs, pc);
                if constexpr (Pol::kStoredPressure)
                    if (pout) {
                        pout[cell] = pc;
                        if (edge) ring_pressures<D>(pout, g, *bc, idx, pc);
                    }
            }
            // an x-face ring cell's recovered pressure is its source's
            if constexpr (Pol::kStoredPressure)
                if (XRING && xedge && pout) {
                    const double ps = __shfl(pc, xsrc);
                    if (xr) pout[cell] = ps;
                }
        }
    }
    if (FUSE_CFL)
        cfl_block_atomic<Pol, S, PERSISTENT>(cand, ctl, ctl->cfl_parity,
                                 (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)));
}

// one tile per block, blocks in XCD-aware order
// Pol::min_waves_per_eu caps the register budget: the FP64 sweep's divide and
// bisection chains are latency-bound, and 4 waves per SIMD (128 VGPRs, a few
// spills) beat 3 unconstrained waves by ~9%.
template<typename Pol, int D, int M, int DIR, int L, int W, int MODE, bool FUSE_CFL>
__global__ void __launch_bounds__(L* W) __attribute__((amdgpu_waves_per_eu(Pol::template min_waves_per_eu<D, M>())))
sweep_kernel(const void* __restrict__ qin, void* __restrict__ qout, Grid g, typename Pol::Eos P,
             StepCtl* __restrict__ ctl, CflArgs cflargs, const double* __restrict__ pin, double* __restrict__ pout)
{
    int bx, by, bz;
    if (DIR >= 1 && cflargs.band > 0) xcd_block_banded(bx, by, bz, (unsigned)cflargs.band);
    else xcd_block(bx, by, bz);
    sweep_tile<Pol, D, M, DIR, L, W, MODE, FUSE_CFL>(qin, qout, g, P, ctl, cflargs, pin, pout, bx, by, bz);
}

// ---------------------------------------------------------------------------
// Ring boundary pass (boundary.hpp:165-226).  Each ring cell is owned by the
// face of its lowest boundary axis; with extent >= 3 on every axis longer
// than 1 the source cells are interior, so ring cells are independent.
// ---------------------------------------------------------------------------
template<typename Pol, int D, int M>
__device__ __forceinline__ void nonreflective_state(const typename Pol::Eos& P, const void* q, const Grid& g,
                                                    const int* idx, const int* src,
                                                    typename Pol::template Cell<D, M>& out)
{
    // local_nonreflective_state (boundary.hpp:150-218)
    using Cell = typename Pol::template Cell<D, M>;
    int second[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d) {
        second[d] = src[d];
        if (idx[d] == 0 && g.N[d] > 2) second[d] = 2;
        else if (idx[d] == g.N[d] - 1 && g.N[d] > 2) second[d] = g.N[d] - 3;
    }
    long long sl = 0, tl = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        sl += (long long)src[d] * g.stride[d];
        tl += (long long)second[d] * g.stride[d];
    }
    Cell first, sec, ex;
    Pol::template load<D, M>(q, g.pitch, sl, first);
    Pol::template load<D, M>(q, g.pitch, tl, sec);
    Cell st = first;
#pragma unroll
    for (int k = 0; k < M; ++k) ex.m[k] = first.m[k] + (first.m[k] - sec.m[k]);
#pragma unroll
    for (int d = 0; d < D; ++d) ex.mom[d] = first.mom[d] + (first.mom[d] - sec.mom[d]);
    ex.E = first.E + (first.E - sec.E);
#pragma unroll
    for (int k = 0; k < M - 1; ++k) ex.a[k] = first.a[k] + (first.a[k] - sec.a[k]);
    int axis = 0;
#pragma unroll
    for (int d = D - 1; d >= 0; --d)
        if (idx[d] == 0 || idx[d] == g.N[d] - 1) axis = d;
    const double outward = (idx[axis] == 0) ? -1.0 : 1.0;
    double ex_rho = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) ex_rho += finite_or((double)ex.m[k]);
    ex_rho = smax(ex_rho, 1e-12);
    double exm = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (d == axis) exm = ex.mom[d];
    const double ex_un = outward * safe_div(exm, ex_rho);
    if (ex_un > 0.0) {
        double rho = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) rho += finite_or((double)st.m[k]);
        rho = smax(rho, 1e-12);
        double m2b = 0.0, m2a = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) m2b += (double)st.mom[d] * (double)st.mom[d];
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (d == axis) st.mom[d] = ex.mom[d];
#pragma unroll
        for (int d = 0; d < D; ++d) m2a += (double)st.mom[d] * (double)st.mom[d];
        st.E += 0.5 * (m2a - m2b) / rho;
    }
    Pol::template repair_cell<D, M>(P, st);
    out = st;
}

// Ring cell number t (0 <= t < bc.face_start[2D], below 2^32): faces in the
// order (axis, side) = 2 axis + side, each face's cells with the earlier
// axes' ring cells left out (a cell belongs to the face of its lowest
// boundary axis), the remaining axes ascending with the first fastest.  Its
// index, its source (the index clamped to [1, N-2] on every axis longer than
// 1) and both linear indices.
template<int D>
__device__ __forceinline__ void ring_index(const Grid& g, const BcArgs& bc, long long t, int* idx, int* src,
                                           long long& lin, long long& sl)
{
    // every array index below is a compile-time constant after unrolling,
    // so nothing is staged through private (scratch) memory
    int face = 0;
    long long base = 0;
#pragma unroll
    for (int f = 1; f < 2 * D; ++f)
        if (t >= bc.face_start[f]) {
            face = f;
            base = bc.face_start[f];
        }
    unsigned local = (unsigned)(t - base);
    const int fd = face >> 1, side = face & 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) idx[d] = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (d == fd) {
            idx[d] = side ? g.N[d] - 1 : 0;
            continue;
        }
        unsigned lo = 0, cnt = (unsigned)g.N[d];
        if (d < fd && g.N[d] > 1) {
            lo = 1;
            cnt = (unsigned)g.N[d] - 2;
        }
        idx[d] = (int)(lo + local % cnt);
        local /= cnt;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) src[d] = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        src[d] = idx[d];
        if (g.N[d] <= 1) continue;
        if (idx[d] == 0) src[d] = 1;
        else if (idx[d] == g.N[d] - 1) src[d] = g.N[d] - 2;
    }
    lin = 0;
    sl = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lin += (long long)idx[d] * g.stride[d];
        sl += (long long)src[d] * g.stride[d];
    }
}

// boundary treatment of ring cell number t (0 <= t < bc.face_start[2D])
template<typename Pol, int D, int M>
__device__ __forceinline__ void bc_cell(void* __restrict__ q, const Grid& g, const typename Pol::Eos& P,
                                        const BcArgs& bc, long long t, double* __restrict__ pstore = nullptr)
{
    using Cell = typename Pol::template Cell<D, M>;
    {
        int idx[3], src[3];
        long long lin, sl;
        ring_index<D>(g, bc, t, idx, src, lin, sl);
        // the first axis on which the cell touches the boundary names it
        int name = BC_TRANSMISSIVE;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            if (idx[d] == 0) name = bc.bc_lo[d];
            else if (idx[d] == g.N[d] - 1) name = bc.bc_hi[d];
        }
        Cell st;
        if (name == BC_NONREFLECTIVE) {
            nonreflective_state<Pol, D, M>(P, q, g, idx, src, st);
        } else {
            Pol::template load<D, M>(q, g.pitch, sl, st);
#pragma unroll
            for (int d = 0; d < D; ++d)
                if ((idx[d] == 0 && bc.bc_lo[d] == BC_REFLECTIVE) || (idx[d] == g.N[d] - 1 && bc.bc_hi[d] == BC_REFLECTIVE))
                    st.mom[d] *= -1;
            // a copy, with or without the normal momentum flipped, keeps the
            // source's recovered pressure bit for bit
            if (pstore) pstore[lin] = pstore[sl];
        }
        Pol::template store<D, M>(q, g.pitch, lin, st);
    }
}

template<typename Pol, int D, int M>
__global__ void __launch_bounds__(256)
bc_kernel(void* __restrict__ q, Grid g, typename Pol::Eos P, BcArgs bc, StepCtl* __restrict__ ctl, int end_step,
          double* __restrict__ pstore)
{
    const long long total = bc.face_start[2 * D];
    const bool active = ctl->active;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; active && t < total;
         t += (long long)gridDim.x * blockDim.x)
        bc_cell<Pol, D, M>(q, g, P, bc, t, pstore);
    if (end_step && blockIdx.x == 0 && threadIdx.x == 0 && active) end_step_ctl(ctl);
}

// The ring pass for copy-kind boundaries (no nonreflective side): the
// copies of bc_cell, field by field.  blockIdx.y is the stored field (the
// per-cell order m[M], mom[D], E, a[M-1]; with pstore one more for the
// recovered pressure) and the x blocks walk the ring cells, consecutive
// threads taking consecutive ring cells.  Each thread moves one value, so a
// wave has 64 independent loads in flight where bc_cell's thread chains a
// cell's fields.  Momentum d is negated on each reflective side of axis d
// the ring cell lies on, as in bc_cell.
template<typename Pol, int D, int M>
__global__ void __launch_bounds__(256)
bc_copy_kernel(void* __restrict__ q, Grid g, BcArgs bc, StepCtl* __restrict__ ctl, int end_step,
               double* __restrict__ pstore, int skip_xring)
{
    constexpr int NF = 2 * M + D;
    const long long total = bc.face_start[2 * D];
    const bool active = ctl->active;
    const int f = blockIdx.y;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; active && t < total;
         t += (long long)gridDim.x * blockDim.x) {
        int idx[3], src[3];
        long long lin, sl;
        ring_index<D>(g, bc, t, idx, src, lin, sl);
        // x-face cells with interior other indices, done by the sweeps
        if (skip_xring && xring_cell<D>(g, idx)) continue;
        if (f == NF) {
            pstore[lin] = pstore[sl];
            continue;
        }
        bool neg = false;
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (f == M + d && ((idx[d] == 0 && bc.bc_lo[d] == BC_REFLECTIVE) ||
                               (idx[d] == g.N[d] - 1 && bc.bc_hi[d] == BC_REFLECTIVE)))
                neg = true;
        Pol::template copy_field<D, M>(q, g.pitch, f, sl, lin, neg);
    }
    if (end_step && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && active) end_step_ctl(ctl);
}

// compute_dt_cfl over every cell into ctl->dtmin[ctl->slot]
template<typename Pol, int D, int M>
__global__ void __launch_bounds__(256)
cfl_kernel(const void* __restrict__ q, Grid g, typename Pol::Eos P, StepCtl* __restrict__ ctl, CflArgs a,
           double* __restrict__ pout)
{
    using Cell = typename Pol::template Cell<D, M>;
    double cand = cfl_identity<Pol>();
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncells;
         c += (long long)gridDim.x * blockDim.x) {
        Cell U;
        Pol::template load<D, M>(q, g.pitch, c, U);
        double pc = 0.0;
        cand = cfl_combine<Pol>(cand, Pol::template cfl_value<D, M>(P, U, a, pc));
        if constexpr (Pol::kStoredPressure)
            if (pout) pout[c] = pc;
    }
    cfl_block_atomic<Pol, 256>(cand, ctl, ctl->slot, (int)blockIdx.x);
}

// packed per-cell doubles <-> policy storage
template<typename Pol, int D, int M>
__global__ void pack_in_kernel(typename Pol::Eos P, const double* __restrict__ aos, void* __restrict__ q,
                               long long first, long long count, long long pitch)
{
    using Cell = typename Pol::template Cell<D, M>;
    constexpr int NV = 2 * M + D;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (long long)gridDim.x * blockDim.x) {
        Cell U;
        Pol::template from_packed<D, M>(P, aos + t * NV, U);
        Pol::template store<D, M>(q, pitch, first + t, U);
    }
}

template<typename Pol, int D, int M>
__global__ void pack_out_kernel(typename Pol::Eos P, const void* __restrict__ q, double* __restrict__ aos,
                                long long first, long long count, long long pitch)
{
    using Cell = typename Pol::template Cell<D, M>;
    constexpr int NV = 2 * M + D;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (long long)gridDim.x * blockDim.x) {
        Cell U;
        Pol::template load<D, M>(q, pitch, first + t, U);
        Pol::template to_packed<D, M>(P, U, aos + t * NV);
    }
}

// Conservation totals (conservation_report.hpp:188-227) in FP64, fixed order:
// block b sums cells [b*chunk, (b+1)*chunk) in a fixed tree; a second pass
// sums the block partials in order.
template<typename Pol, int D, int M>
__global__ void __launch_bounds__(256)
totals_kernel(const void* __restrict__ q, Grid g, typename Pol::Eos P, double volume, long long chunk,
              double* __restrict__ partial)
{
    using Cell = typename Pol::template Cell<D, M>;
    constexpr int NO = M + 1 + D + 1;
    __shared__ double red[NO][256];
    double acc[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) acc[o] = 0.0;
    const long long beg = (long long)blockIdx.x * chunk;
    long long end = beg + chunk;
    if (end > g.ncells) end = g.ncells;
    for (long long c = beg + threadIdx.x; c < end; c += 256) {
        Cell U;
        Pol::template load<D, M>(q, g.pitch, c, U);
        double cm = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double part = Pol::template mass<D, M>(U, k) * volume;
            acc[k] += part;
            cm += part;
        }
        acc[M] += cm;
#pragma unroll
        for (int d = 0; d < D; ++d) acc[M + 1 + d] += Pol::template mom<D, M>(U, d) * volume;
        acc[M + 1 + D] += Pol::template energy<D, M>(P, U) * volume;
    }
#pragma unroll
    for (int o = 0; o < NO; ++o) red[o][threadIdx.x] = acc[o];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int o = 0; o < NO; ++o) red[o][threadIdx.x] += red[o][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int o = 0; o < NO; ++o) partial[(long long)blockIdx.x * NO + o] = red[o][0];
}

static __global__ void sum_partials_kernel(const double* __restrict__ partial, int nblocks, int no, double* __restrict__ out)
{
    const int o = threadIdx.x;
    if (o >= no) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[(long long)b * no + o];
    out[o] = s;
}

}  // namespace dimhip
