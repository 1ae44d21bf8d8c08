// kernels_fp64.hpp — FP64 kernels: directional sweep, ring boundary pass,
// CFL reduction, step control, packing and conservation sums.
#pragma once

#include <hip/hip_runtime.h>

#include "dim_fp64.hpp"

namespace dimhip {

constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

template<int D, int M>
struct Layout {
    static constexpr int NV = 2 * M + D;
    static constexpr int NP = 2 * M + D + 1;
    static constexpr int NF = M + D + 1;
};

template<int D, int M>
__device__ __forceinline__ void load_state(const double* __restrict__ q, long long pitch, long long c,
                                           State<D, M>& U)
{
    int f = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) U.m[k] = q[(f++) * pitch + c];
#pragma unroll
    for (int d = 0; d < D; ++d) U.mom[d] = q[(f++) * pitch + c];
    U.E = q[(f++) * pitch + c];
#pragma unroll
    for (int k = 0; k < M - 1; ++k) U.a[k] = q[(f++) * pitch + c];
}

template<int D, int M>
__device__ __forceinline__ void store_state(double* __restrict__ q, long long pitch, long long c,
                                            const State<D, M>& U)
{
    int f = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) q[(f++) * pitch + c] = U.m[k];
#pragma unroll
    for (int d = 0; d < D; ++d) q[(f++) * pitch + c] = U.mom[d];
    q[(f++) * pitch + c] = U.E;
#pragma unroll
    for (int k = 0; k < M - 1; ++k) q[(f++) * pitch + c] = U.a[k];
}

// LDS exchange helpers: a per-thread record of NF doubles stored as SoA
// planes sh[f * S + slot].
template<int D, int M>
__device__ __forceinline__ void put_prim(double* sh, int S, int slot, int off, const Prim<D, M>& P)
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

template<int D, int M>
__device__ __forceinline__ void get_prim(const double* sh, int S, int slot, int off, Prim<D, M>& P)
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

template<int D, int M>
__device__ __forceinline__ void put_state(double* sh, int S, int slot, int off, const State<D, M>& U)
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

template<int D, int M>
__device__ __forceinline__ void get_state(const double* sh, int S, int slot, int off, State<D, M>& U)
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

__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long double_to_bits(double x) { return (unsigned long long)__double_as_longlong(x); }

// Block-wide min of a positive double (or +inf), then one atomic per block.
template<int NT>
__device__ __forceinline__ void block_min_atomic(double v, unsigned long long* target)
{
    __shared__ double red[NT / 64 > 0 ? NT / 64 : 1];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = smin(v, o);
    }
    const int tid = threadIdx.x + threadIdx.y * blockDim.x + threadIdx.z * blockDim.x * blockDim.y;
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double m = red[0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) m = smin(m, red[w]);
        if (m < bits_to_double(kInfBits)) atomicMin(target, double_to_bits(m));
    }
}

// CFL candidate min over d of cfl*dx[d]/(|u_d|+c) for one cell
// (cfl.hpp:46-63); unsplit: cfl / sum_d (|u_d|+c)/dx[d] (cfl.hpp:94-116).
template<int D, int M>
__device__ __forceinline__ double cfl_candidate(const EosParams& P, const State<D, M>& U, const double* cfl_dx,
                                                const double* dx, double cfl, bool unsplit)
{
    Prim<D, M> Q;
    cons_to_prim<D, M>(P, U, Q);
    const double c = mixture_sound_speed<M>(P, total_density<D, M>(U), Q.p, Q.alpha, Q.rho);
    double best = bits_to_double(kInfBits);
    if (!unsplit) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double ws = __builtin_fabs(Q.vel[d]) + c;
            if (ws > 1e-14) best = smin(best, cfl_dx[d] / ws);
        }
    } else {
        double spectral = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) spectral += (__builtin_fabs(Q.vel[d]) + c) / dx[d];
        if (spectral > 1e-14) best = smin(best, cfl / spectral);
    }
    return best;
}

struct CflArgs {
    double cfl_dx[3];
    double dx[3];
    double cfl;
    int unsplit;
};

// ---------------------------------------------------------------------------
// Directional sweep: sweep_direction_dispatch + solve_line
// (sweep_core.hpp:16-202, muscl.hpp:197-329, hllc.hpp:62-195).
//
// A block owns W lines x L positions along the sweep axis; positions
// [3, L-3) are written (T = L-6 cells per tile), the 3-cell halo on each side
// is the reference stencil radius.  Phases, each separated by an LDS
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
// MODE 0: plain sweep qin -> qout.  MODE 1: unsplit accumulation
// qout += (sweep(qin) - qin), then repair (advance_step.hpp:198-222).
// FUSE_CFL: also reduce next step's CFL candidate over interior cells.
// ---------------------------------------------------------------------------
template<int D, int M, int DIR, int L, int W, int MODE, bool FUSE_CFL>
__global__ void __launch_bounds__(L* W)
sweep64_kernel(const double* __restrict__ qin, double* __restrict__ qout, Grid g, EosParams P,
               StepCtl* __restrict__ ctl, CflArgs cflargs)
{
    constexpr int S = L * W;
    constexpr int T = L - 6;
    constexpr int NREC = 4 * M + 2 * D + 1;
    __shared__ double sh[NREC * S];

    const double dt = ctl->dt;
    const double dxs = g.dx[DIR];

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
        i = blockIdx.x * T - 3 + pos;
        idx[1] = blockIdx.y * W + lane;
        idx[2] = blockIdx.z;
        line_ok = idx[1] < g.N[1];
    } else if (DIR == 1) {
        idx[0] = blockIdx.x * W + lane;
        i = blockIdx.y * T - 3 + pos;
        idx[2] = blockIdx.z;
        line_ok = idx[0] < g.N[0];
    } else {
        idx[0] = blockIdx.x * W + lane;
        idx[1] = blockIdx.z;
        i = blockIdx.y * T - 3 + pos;
        line_ok = idx[0] < g.N[0];
    }
    idx[DIR] = i;
    const bool in = line_ok && i >= 0 && i < n;
    const long long cell = (long long)idx[0] + (long long)idx[1] * g.stride[1] + (long long)idx[2] * g.stride[2];
    const bool end_cell = (i == 0) || (i == n - 1);

    if (!ctl->active) {
        // a step past t_end: the state passes through unchanged, so the
        // host-side double-buffer flip stays consistent
        if (MODE == 0 && in && pos >= 3 && pos < L - 3) {
            State<D, M> U;
            load_state<D, M>(qin, g.pitch, cell, U);
            store_state<D, M>(qout, g.pitch, cell, U);
        }
        return;
    }

    // phase 1
    State<D, M> U;
    Prim<D, M> Pn, Pl;
    if (in) {
        load_state<D, M>(qin, g.pitch, cell, U);
        cons_to_prim<D, M>(P, U, Pn);
        Pl = Pn;
        enforce_positive_primitive<D, M>(Pl);
        put_prim<D, M>(sh, S, slot, 0, Pl);
    }
    __syncthreads();

    // phase 2
    State<D, M> Uh;
    Prim<D, M> Ph;
    const bool ph2 = in && pos >= 1 && pos < L - 1;
    if (ph2) {
        if (end_cell) {
            Uh = U;
            Ph = Pl;
        } else {
            Prim<D, M> PL, PR, lf, rf;
            get_prim<D, M>(sh, S, slot_l, 0, PL);
            get_prim<D, M>(sh, S, slot_r, 0, PR);
            reconstruct_cell_faces<D, M>(PL, Pl, PR, lf, rf);
            State<D, M> ul, ur;
            prim_to_cons<D, M>(P, lf, ul);
            prim_to_cons<D, M>(P, rf, ur);
            repair_state<D, M>(P, ul);
            repair_state<D, M>(P, ur);
            hancock_predict<D, M, DIR>(P, U, Pl, ul, ur, lf, rf, dt, dxs, Uh);
            cons_to_prim<D, M>(P, Uh, Ph);
            enforce_positive_primitive<D, M>(Ph);
        }
    }
    __syncthreads();
    if (ph2) put_prim<D, M>(sh, S, slot, 0, Ph);
    __syncthreads();

    // phase 3
    State<D, M> Ulh, Urh;
    Prim<D, M> PUl, PUr;
    const bool ph3 = in && pos >= 2 && pos < L - 2;
    if (ph3) {
        if (end_cell) {
            Ulh = U;
            Urh = U;
            PUl = Pn;
            PUr = Pn;
        } else {
            Prim<D, M> PL, PR, lf, rf;
            get_prim<D, M>(sh, S, slot_l, 0, PL);
            get_prim<D, M>(sh, S, slot_r, 0, PR);
            reconstruct_cell_faces<D, M>(PL, Ph, PR, lf, rf);
            prim_to_cons<D, M>(P, lf, Ulh);
            prim_to_cons<D, M>(P, rf, Urh);
            repair_state<D, M>(P, Ulh);
            repair_state<D, M>(P, Urh);
            cons_to_prim<D, M>(P, Ulh, PUl);
            cons_to_prim<D, M>(P, Urh, PUr);
        }
    }
    __syncthreads();
    if (ph3) {
        put_state<D, M>(sh, S, slot, 0, Ulh);
        put_prim<D, M>(sh, S, slot, Layout<D, M>::NV, PUl);
    }
    __syncthreads();

    // phase 4: face i+1/2
    Flux<D, M> F;
    double fv = 0.0;
    double af[M];
    const bool ph4 = in && pos >= 2 && pos < L - 3 && i <= n - 2;
    if (ph4) {
        State<D, M> UR;
        Prim<D, M> PR;
        get_state<D, M>(sh, S, slot_r, 0, UR);
        get_prim<D, M>(sh, S, slot_r, Layout<D, M>::NV, PR);
        if (!hllc_flux<D, M, DIR>(P, Urh, UR, PUr, PR, F, fv)) atomicOr(&ctl->err, ERR_NONFINITE_HLLC);
        const double* up = (fv >= 0.0) ? PUr.alpha : PR.alpha;
#pragma unroll
        for (int k = 0; k < M; ++k) af[k] = up[k] * fv;
    }
    __syncthreads();
    if (ph4) {
        int f = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) sh[(f++) * S + slot] = F.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) sh[(f++) * S + slot] = F.mom[d];
        sh[(f++) * S + slot] = F.E;
        sh[(f++) * S + slot] = fv;
#pragma unroll
        for (int k = 0; k < M; ++k) sh[(f++) * S + slot] = af[k];
    }
    __syncthreads();

    // phase 5: update cell i
    double cand = bits_to_double(kInfBits);
    if (in && pos >= 3 && pos < L - 3) {
        State<D, M> r = U;
        if (n > 1) {
            Flux<D, M> Fl, Fr;
            double fvl, fvr, afl[M], afr[M];
            if (i <= n - 2) {
                Fr = F;
                fvr = fv;
#pragma unroll
                for (int k = 0; k < M; ++k) afr[k] = af[k];
            }
            if (i >= 1) {
                int f = 0;
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
            const double lambda = dt / dxs;
#pragma unroll
            for (int k = 0; k < M; ++k) r.m[k] -= lambda * (Fr.m[k] - Fl.m[k]);
#pragma unroll
            for (int d = 0; d < D; ++d) r.mom[d] -= lambda * (Fr.mom[d] - Fl.mom[d]);
            r.E -= lambda * (Fr.E - Fl.E);
            repair_state<D, M>(P, r);

            double rhs[M], an[M];
            alpha_rhs_coefficients<D, M>(P, Pn, rhs);
            const double div_u = (fvr - fvl) / dxs;
#pragma unroll
            for (int k = 0; k < M; ++k) {
                an[k] = Pn.alpha[k];
                an[k] -= lambda * (afr[k] - afl[k]);
                an[k] += dt * rhs[k] * div_u;
            }
            sanitise_alpha<M>(an, 0.0);
#pragma unroll
            for (int k = 0; k < M - 1; ++k) r.a[k] = an[k];
            repair_state<D, M>(P, r);
        }
        if (MODE == 1) {
            // accumulate_delta (advance_step.hpp:198-222): acc += U_dir - U_base; repair
            State<D, M> acc;
            load_state<D, M>(qout, g.pitch, cell, acc);
#pragma unroll
            for (int k = 0; k < M; ++k) acc.m[k] += r.m[k] - U.m[k];
#pragma unroll
            for (int d = 0; d < D; ++d) acc.mom[d] += r.mom[d] - U.mom[d];
            acc.E += r.E - U.E;
#pragma unroll
            for (int k = 0; k < M - 1; ++k) acc.a[k] += r.a[k] - U.a[k];
            repair_state<D, M>(P, acc);
            r = acc;
        }
        store_state<D, M>(qout, g.pitch, cell, r);
        if (FUSE_CFL) {
            // ring cells become copies (|u| preserved) of interior cells in the
            // boundary pass, so the interior minimum is the next step's CFL dt
            bool interior = true;
#pragma unroll
            for (int d = 0; d < D; ++d)
                if (g.N[d] > 1 && (idx[d] == 0 || idx[d] == g.N[d] - 1)) interior = false;
            if (interior) cand = cfl_candidate<D, M>(P, r, cflargs.cfl_dx, cflargs.dx, cflargs.cfl, cflargs.unsplit);
        }
    }
    if (FUSE_CFL) block_min_atomic<S>(cand, &ctl->dtmin[ctl->slot ^ 1]);
}

// ---------------------------------------------------------------------------
// Ring boundary pass (boundary.hpp:165-226).  Each ring cell is owned by the
// face of its lowest boundary axis; source cells are interior (N >= 3 per
// axis with extent > 1), so cells are independent.
// ---------------------------------------------------------------------------
struct BcArgs {
    int bc_lo[3];
    int bc_hi[3];
    long long face_start[7];  // prefix sums of ring cells per face (d, side) = 2d+side
};

template<int D, int M>
__device__ __forceinline__ void nonreflective_state(const EosParams& P, const double* __restrict__ q, const Grid& g,
                                                    const int* idx, const int* src, State<D, M>& out)
{
    // local_nonreflective_state (boundary.hpp:150-218)
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
    State<D, M> first, sec, ex;
    load_state<D, M>(q, g.pitch, sl, first);
    load_state<D, M>(q, g.pitch, tl, sec);
    State<D, M> st = first;
#pragma unroll
    for (int k = 0; k < M; ++k) ex.m[k] = first.m[k] + (first.m[k] - sec.m[k]);
#pragma unroll
    for (int d = 0; d < D; ++d) ex.mom[d] = first.mom[d] + (first.mom[d] - sec.mom[d]);
    ex.E = first.E + (first.E - sec.E);
#pragma unroll
    for (int k = 0; k < M - 1; ++k) ex.a[k] = first.a[k] + (first.a[k] - sec.a[k]);
    int axis = 0;
    for (int d = D - 1; d >= 0; --d)
        if (idx[d] == 0 || idx[d] == g.N[d] - 1) axis = d;
    const double outward = (idx[axis] == 0) ? -1.0 : 1.0;
    const double ex_rho = smax(total_density<D, M>(ex), 1e-12);
    double exm = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (d == axis) exm = ex.mom[d];
    const double ex_un = outward * safe_div(exm, ex_rho);
    if (ex_un > 0.0) {
        const double rho = smax(total_density<D, M>(st), 1e-12);
        double m2b = 0.0, m2a = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) m2b += st.mom[d] * st.mom[d];
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (d == axis) st.mom[d] = ex.mom[d];
#pragma unroll
        for (int d = 0; d < D; ++d) m2a += st.mom[d] * st.mom[d];
        st.E += 0.5 * (m2a - m2b) / rho;
    }
    repair_state<D, M>(P, st);
    out = st;
}

template<int D, int M>
__global__ void __launch_bounds__(256)
bc64_kernel(double* __restrict__ q, Grid g, EosParams P, BcArgs bc, StepCtl* __restrict__ ctl, int end_step)
{
    const long long total = bc.face_start[2 * D];
    const bool active = ctl->active;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; active && t < total;
         t += (long long)gridDim.x * blockDim.x) {
        int face = 0;
        while (face < 2 * D - 1 && t >= bc.face_start[face + 1]) ++face;
        long long local = t - bc.face_start[face];
        const int fd = face >> 1, side = face & 1;
        int idx[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (d == fd) continue;
            int lo = 0, cnt = g.N[d];
            if (d < fd && g.N[d] > 1) {
                lo = 1;
                cnt = g.N[d] - 2;
            }
            idx[d] = lo + (int)(local % cnt);
            local /= cnt;
        }
        idx[fd] = side ? g.N[fd] - 1 : 0;
        int src[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; ++d) {
            src[d] = idx[d];
            if (g.N[d] <= 1) continue;
            if (idx[d] == 0) src[d] = 1;
            else if (idx[d] == g.N[d] - 1) src[d] = g.N[d] - 2;
        }
        int name = BC_TRANSMISSIVE;
        for (int d = 0; d < D; ++d) {
            if (idx[d] == 0) {
                name = bc.bc_lo[d];
                break;
            }
            if (idx[d] == g.N[d] - 1) {
                name = bc.bc_hi[d];
                break;
            }
        }
        long long lin = 0, sl = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            lin += (long long)idx[d] * g.stride[d];
            sl += (long long)src[d] * g.stride[d];
        }
        State<D, M> st;
        if (name == BC_NONREFLECTIVE) {
            nonreflective_state<D, M>(P, q, g, idx, src, st);
        } else {
            load_state<D, M>(q, g.pitch, sl, st);
#pragma unroll
            for (int d = 0; d < D; ++d)
                if ((idx[d] == 0 && bc.bc_lo[d] == BC_REFLECTIVE) || (idx[d] == g.N[d] - 1 && bc.bc_hi[d] == BC_REFLECTIVE))
                    st.mom[d] *= -1.0;
        }
        store_state<D, M>(q, g.pitch, lin, st);
    }
    if (end_step && blockIdx.x == 0 && threadIdx.x == 0 && active) {
        ctl->time += ctl->dt;
        ctl->step += 1;
        ctl->slot ^= 1;
    }
}

// compute_dt_cfl over every cell into ctl->dtmin[ctl->slot]
template<int D, int M>
__global__ void __launch_bounds__(256)
cfl64_kernel(const double* __restrict__ q, Grid g, EosParams P, StepCtl* __restrict__ ctl, CflArgs a)
{
    double cand = bits_to_double(kInfBits);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncells;
         c += (long long)gridDim.x * blockDim.x) {
        State<D, M> U;
        load_state<D, M>(q, g.pitch, c, U);
        cand = smin(cand, cfl_candidate<D, M>(P, U, a.cfl_dx, a.dx, a.cfl, a.unsplit));
    }
    block_min_atomic<256>(cand, &ctl->dtmin[ctl->slot]);
}

// Start of a step: dt = min(dt_max, CFL minimum) exactly as the reference's
// running std::min from dt = dt_max (cfl.hpp:43-63); clear the next slot.
__global__ void begin_step_kernel(StepCtl* ctl, double dt_max_host)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int s = ctl->slot;
    const double m = bits_to_double(ctl->dtmin[s]);
    double dt_max = dt_max_host;
    int active = 1;
    if (ctl->use_t_end) {
        active = ctl->time < ctl->t_end - 1e-14;
        dt_max = ctl->t_end - ctl->time;
    }
    const double dt = smin(dt_max, m);
    if (active && !(dt > 0.0)) {
        ctl->err |= ERR_NONPOSITIVE_DT;
        active = 0;
    }
    ctl->dt = dt;
    ctl->dt_max = dt_max;
    ctl->active = active;
    ctl->dtmin[s ^ 1] = kInfBits;
}

__global__ void reset_dtmin_kernel(StepCtl* ctl, int which_rel)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->dtmin[ctl->slot ^ which_rel] = kInfBits;
}

__global__ void set_active_kernel(StepCtl* ctl, double dt)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->dt = dt;
        ctl->active = 1;
    }
}

// packed AoS (cell-major, nv doubles per cell) <-> SoA planes
__global__ void aos_to_soa_kernel(const double* __restrict__ aos, double* __restrict__ soa, long long first,
                                  long long count, int nv, long long pitch)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count * nv;
         t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / nv;
        const int f = (int)(t - c * nv);
        soa[(long long)f * pitch + first + c] = aos[t];
    }
}

__global__ void soa_to_aos_kernel(const double* __restrict__ soa, double* __restrict__ aos, long long first,
                                  long long count, int nv, long long pitch)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count * nv;
         t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / nv;
        const int f = (int)(t - c * nv);
        aos[t] = soa[(long long)f * pitch + first + c];
    }
}

// Conservation totals, pass 1: block b sums cells [b*chunk, (b+1)*chunk) in
// a fixed tree order; pass 2 sums the block partials in a fixed order.
template<int D, int M>
__global__ void __launch_bounds__(256)
cons_totals_kernel(const double* __restrict__ q, Grid g, double volume, long long chunk, double* __restrict__ partial)
{
    constexpr int NO = M + 1 + D + 1;
    __shared__ double red[NO][256];
    double acc[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) acc[o] = 0.0;
    const long long beg = (long long)blockIdx.x * chunk;
    long long end = beg + chunk;
    if (end > g.ncells) end = g.ncells;
    for (long long c = beg + threadIdx.x; c < end; c += 256) {
        State<D, M> U;
        load_state<D, M>(q, g.pitch, c, U);
        double cm = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double part = U.m[k] * volume;
            acc[k] += part;
            cm += part;
        }
        acc[M] += cm;
#pragma unroll
        for (int d = 0; d < D; ++d) acc[M + 1 + d] += U.mom[d] * volume;
        acc[M + 1 + D] += U.E * volume;
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

__global__ void sum_partials_kernel(const double* __restrict__ partial, int nblocks, int no, double* __restrict__ out)
{
    const int o = threadIdx.x;
    if (o >= no) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[(long long)b * no + o];
    out[o] = s;
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

template<int D, int M, int DIR>
__global__ void kat64_kernel(EosParams P, int which, int count, const double* __restrict__ in0,
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
        if (!hllc_flux<D, M, DIR>(P, UL, UR, PL, PR, F, fv)) atomicOr(err, 1);
        double* o = out0 + (long long)t * LY::NF;
        int j = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) o[j++] = F.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) o[j++] = F.mom[d];
        o[j++] = F.E;
        out1[t] = fv;
    } else if (which == 4) {
        Prim<D, M> Q;
        load_packed_prim<D, M>(in0 + (long long)t * LY::NP, Q);
        double rhs[M];
        alpha_rhs_coefficients<D, M>(P, Q, rhs);
#pragma unroll
        for (int k = 0; k < M; ++k) out0[(long long)t * M + k] = rhs[k];
    }
}

__global__ void kat_ieee_kernel(int count, const double* __restrict__ in, double* __restrict__ div_out,
                                double* __restrict__ sqrt_out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const double a = in[2 * t], b = in[2 * t + 1];
    div_out[t] = a / b;
    sqrt_out[t] = __builtin_sqrt(a);
}

}  // namespace dimhip
