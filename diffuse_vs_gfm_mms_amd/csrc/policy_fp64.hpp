// policy_fp64.hpp — the FP64 arithmetic policy of the generic DIM kernels:
// every operation is the reference's (dim_fp64.hpp), state stored as FP64
// SoA planes in the reference's field order.
#pragma once

#include <type_traits>

#include "dim_fp64.hpp"

namespace dimhip {

struct CflArgs {
    double cfl_dx[3];
    double inv_cfl_dx[3];  // 1 / (cfl * dx): the FP32-mixed form's wave-rate reduction
    double dx[3];
    double inv_dx[3];
    double cfl;
    double inv_cfl;
    int unsplit;
    // single-grid 3D split steps with copy-kind boundaries: the sweeps give the
    // x-face ring cells (other indices interior) their x-neighbour's new
    // state, with the x momentum negated on a reflective side (xflip bit 0:
    // low side, bit 1: high side), and the ring pass leaves them out
    int xring;
    int xflip;
    // single-grid 2D split steps with copy-kind boundaries: the x sweep
    // writes the y-face ring rows (x interior) from rows 1 and N1-2, with
    // the y momentum negated on a reflective side (yflip bit 0: low, bit 1:
    // high), and no ring pass follows it (the y sweep reads no other ring
    // cell for an interior output; the pass after it restores the whole ring)
    int yring;
    int yflip;
    // y and z sweeps: block order in column strips of `band` tiles
    // (xcd_block_banded); 0 keeps the row-major order of xcd_block
    int band;
};

constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;
__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long double_to_bits(double x) { return (unsigned long long)__double_as_longlong(x); }

// NA selects the parameter block: EosParams (ideal and stiffened gases) or
// EosParamsNA (also noble-Abel); Fp64 and Fp64NA below.
template<bool NA>
struct Fp64T {
    // register cap of the sweep kernels: 4 waves per SIMD (128 VGPRs); the
    // three-material kernels (without noble-Abel) take 3 (168 VGPRs: the 2D
    // ones fit without spills, the 4-wave cap spilled about 700 B per lane)
    template<int D, int M>
    static constexpr int min_waves_per_eu() { return (M >= 3 && !NA) ? 3 : 4; }
    // the sweep re-reads a cell's input state instead of keeping it in
    // registers across its phases (register pressure, not bandwidth, binds here)
    static constexpr bool kReloadInput = true;
    // the update's inputs from the first phase (P^n's volume fractions and
    // the alpha source coefficients) wait in LDS, not registers, while the
    // reconstruction and flux phases run (fewer spills at the 4-wave cap)
    static constexpr bool kParkAlpha = true;
    // one LDS region for every exchange of the sweep tile: the faces' full
    // FP64 states travel through LDS (two regions would halve the occupancy)
    template<int D, int M, int DIR>
    static constexpr bool ping_pong() { return false; }
    template<int D, int M, int DIR>
    static constexpr bool prim_faces() { return false; }
    template<int D, int M, int DIR>
    static constexpr bool park() { return kParkAlpha && M <= 2; }
    // sweep tiles: x 128 positions x 2 lines; y/z 4 lines x 128 positions
    // (halo 6 of 128; 8 lines x 64 positions was 1.6 % slower)
    static constexpr int kSweepL0 = 128, kSweepW0 = 2, kSweepL1 = 128, kSweepW1 = 4;
    // 2D y sweeps keep the row-major block order: one column strip per XCD
    // (xcd_block_banded) measured 3 % slower for these VALU-bound tiles
    static constexpr bool kXcdStrips = false;
    // CFL reduction: each cell contributes its dt candidate, combined by min
    static constexpr bool kCflRate = false;
    // the CFL pass stores each cell's recovered pressure (cons_to_prim's p,
    // bisection included) and the next sweep's first phase reads it back
    // instead of repeating the bisection: same bits, one bisection less
    static constexpr bool kStoredPressure = true;
    using R = double;
    using Eos = std::conditional_t<NA, EosParamsNA, EosParams>;
    template<int D, int M> using St = State<D, M>;
    template<int D, int M> using Pr = Prim<D, M>;
    template<int D, int M> using Fx = Flux<D, M>;
    template<int D, int M> using Cell = State<D, M>;

    // storage: nv double planes, field f of cell c at q[f * pitch + c]
    template<int D, int M>
    static __device__ __forceinline__ void load(const void* base, long long pitch, long long c, Cell<D, M>& U)
    {
        const double* q = (const double*)base;
        int f = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) U.m[k] = q[(f++) * pitch + c];
#pragma unroll
        for (int d = 0; d < D; ++d) U.mom[d] = q[(f++) * pitch + c];
        U.E = q[(f++) * pitch + c];
#pragma unroll
        for (int k = 0; k < M - 1; ++k) U.a[k] = q[(f++) * pitch + c];
    }

    // load / store of cell cb + cr with cb uniform across the wave and cr a
    // 32-bit offset (cr * 8 < 2^32)
    template<int D, int M>
    static __device__ __forceinline__ void load_rel(const void* base, long long pitch, long long cb, unsigned cr,
                                                    Cell<D, M>& U)
    {
        const double* q = (const double*)base + cb;
        int f = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) U.m[k] = plane_at(q + (f++) * pitch, cr);
#pragma unroll
        for (int d = 0; d < D; ++d) U.mom[d] = plane_at(q + (f++) * pitch, cr);
        U.E = plane_at(q + (f++) * pitch, cr);
#pragma unroll
        for (int k = 0; k < M - 1; ++k) U.a[k] = plane_at(q + (f++) * pitch, cr);
    }

    template<int D, int M>
    static __device__ __forceinline__ void store_rel(void* base, long long pitch, long long cb, unsigned cr,
                                                     const Cell<D, M>& U)
    {
        double* q = (double*)base + cb;
        int f = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) plane_at(q + (f++) * pitch, cr) = U.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) plane_at(q + (f++) * pitch, cr) = U.mom[d];
        plane_at(q + (f++) * pitch, cr) = U.E;
#pragma unroll
        for (int k = 0; k < M - 1; ++k) plane_at(q + (f++) * pitch, cr) = U.a[k];
    }

    template<int D, int M>
    static __device__ __forceinline__ void store(void* base, long long pitch, long long c, const Cell<D, M>& U)
    {
        double* q = (double*)base;
        int f = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) q[(f++) * pitch + c] = U.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) q[(f++) * pitch + c] = U.mom[d];
        q[(f++) * pitch + c] = U.E;
#pragma unroll
        for (int k = 0; k < M - 1; ++k) q[(f++) * pitch + c] = U.a[k];
    }

    // stored field f (the per-cell order m[M], mom[D], E, a[M-1]) of cell src
    // copied to cell dst, negated when neg (a bit-exact sign flip: x * -1)
    template<int D, int M>
    static __device__ __forceinline__ void copy_field(void* base, long long pitch, int f, long long src,
                                                      long long dst, bool neg)
    {
        double* p = (double*)base + (long long)f * pitch;
        const double v = p[src];
        p[dst] = neg ? v * -1 : v;
    }

    // packed per-cell doubles <-> Cell
    template<int D, int M>
    static __device__ __forceinline__ void from_packed(const Eos&, const double* p, Cell<D, M>& U)
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
    static __device__ __forceinline__ void to_packed(const Eos&, const Cell<D, M>& U, double* p)
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
    static __device__ __forceinline__ void to_state(const Cell<D, M>& c, St<D, M>& U) { U = c; }

    // cons_to_prim with the pressure recovered earlier for the same state
    template<int D, int M>
    static __device__ __forceinline__ void cons_to_prim_stored(const Eos&, const St<D, M>& U, double p, Pr<D, M>& Q)
    {
        double ie;
        cons_to_prim_base<D, M>(U, Q, ie);
        Q.p = p;
    }

    template<int D, int M>
    static __device__ __forceinline__ void cons_to_prim(const Eos& P, const St<D, M>& U, Pr<D, M>& Q)
    {
        dimhip::cons_to_prim<D, M>(P, U, Q);
    }
    template<int D, int M>
    static __device__ __forceinline__ void enforce(Pr<D, M>& Q) { enforce_positive_primitive<D, M>(Q); }
    // enforce_positive_primitive after cons_to_prim (muscl.hpp:231-232, 294-295)
    template<int D, int M>
    static __device__ __forceinline__ void enforce_c2p(Pr<D, M>& Q) { enforce_positive_primitive<D, M>(Q); }

    // alpha_rhs_coefficients(P^n) (sweep_core.hpp:131-137)
    template<int D, int M>
    static __device__ __forceinline__ void alpha_rhs(const Eos& P, const Pr<D, M>& Pn, double* rhs)
    {
        alpha_rhs_coefficients<D, M>(P, Pn, rhs);
    }
    // the predictor's coefficients come from P_line (muscl.hpp:173-179)
    template<int D, int M>
    static __device__ __forceinline__ void alpha_rhs_line(const Eos& P, const Pr<D, M>& Pl, const double*,
                                                          double* rhs)
    {
        alpha_rhs_coefficients<D, M>(P, Pl, rhs);
    }
    template<int D, int M>
    static __device__ __forceinline__ void reconstruct(const Pr<D, M>& L, const Pr<D, M>& C, const Pr<D, M>& Rt,
                                                       Pr<D, M>& lf, Pr<D, M>& rf)
    {
        reconstruct_cell_faces<D, M>(L, C, Rt, lf, rf);
    }
    template<int D, int M>
    static __device__ __forceinline__ void prim_to_cons(const Eos& P, const Pr<D, M>& Q, St<D, M>& U)
    {
        dimhip::prim_to_cons<D, M>(P, Q, U);
    }
    template<int D, int M>
    static __device__ __forceinline__ void repair(const Eos& P, St<D, M>& U) { repair_state<D, M>(P, U); }

    // reconstructed face primitive -> face state, as the reference builds it
    // (muscl.hpp:263-266, 319-322): prim_to_cons then repair_state
    template<int D, int M>
    static __device__ __forceinline__ void face_cons(const Eos& P, const Pr<D, M>& Q, St<D, M>& U)
    {
        dimhip::prim_to_cons<D, M>(P, Q, U);
        repair_state<D, M>(P, U);
    }

    // ... and its primitive for the Riemann solver (sweep_core.hpp:78-81)
    template<int D, int M>
    static __device__ __forceinline__ void face_state(const Eos& P, const Pr<D, M>& Q, St<D, M>& U, Pr<D, M>& QU)
    {
        dimhip::prim_to_cons<D, M>(P, Q, U);
        repair_state<D, M>(P, U);
        dimhip::cons_to_prim<D, M>(P, U, QU);
    }
    // a sweep's step sizes: dt and dx, which the reference's expressions
    // combine themselves (dt/dx, 0.5*dt/dx) for their bits
    struct Factors {
        double dt, dx;
    };
    static __device__ __forceinline__ Factors factors(double dt, double dx, double /*inv_dx*/) { return {dt, dx}; }

    template<int D, int M, int DIR>
    static __device__ __forceinline__ void hancock(const Eos& P, const St<D, M>& Uc, const Pr<D, M>& Pc,
                                                   const St<D, M>& ul, const St<D, M>& ur, const Pr<D, M>& lf,
                                                   const Pr<D, M>& rf, const double* rhs, const Factors& f,
                                                   St<D, M>& out)
    {
        hancock_predict<D, M, DIR>(P, Uc, Pc, ul, ur, lf, rf, rhs, f.dt, f.dx, out);
    }
    template<int D, int M, int DIR>
    static __device__ __forceinline__ int hllc(const Eos& P, const St<D, M>& UL, const St<D, M>& UR,
                                                const Pr<D, M>& PL, const Pr<D, M>& PR, Fx<D, M>& F, double& fv)
    {
        return hllc_flux<D, M, DIR>(P, UL, UR, PL, PR, F, fv);
    }

    // solve_line's update of one cell (sweep_core.hpp:101-148)
    template<int D, int M>
    static __device__ __forceinline__ void update(const Eos& P, const Cell<D, M>& Uin, const Pr<D, M>& Pn,
                                                  const double* rhs, const Fx<D, M>& Fl, const Fx<D, M>& Fr,
                                                  double fvl, double fvr,
                                                  const double* afl, const double* afr, const Factors& f,
                                                  Cell<D, M>& r)
    {
        const double dt = f.dt, dx = f.dx;
        const double lambda = dt / dx;
        r = Uin;
#pragma unroll
        for (int k = 0; k < M; ++k) r.m[k] -= lambda * (Fr.m[k] - Fl.m[k]);
#pragma unroll
        for (int d = 0; d < D; ++d) r.mom[d] -= lambda * (Fr.mom[d] - Fl.mom[d]);
        r.E -= lambda * (Fr.E - Fl.E);
        repair_state<D, M>(P, r);
        double an[M];
        const double div_u = (fvr - fvl) / dx;
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

    // accumulate_delta (advance_step.hpp:198-222): acc += U_dir - U_base; repair
    template<int D, int M>
    static __device__ __forceinline__ void accumulate(const Eos& P, Cell<D, M>& acc, const Cell<D, M>& r,
                                                      const Cell<D, M>& U)
    {
#pragma unroll
        for (int k = 0; k < M; ++k) acc.m[k] += r.m[k] - U.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) acc.mom[d] += r.mom[d] - U.mom[d];
        acc.E += r.E - U.E;
#pragma unroll
        for (int k = 0; k < M - 1; ++k) acc.a[k] += r.a[k] - U.a[k];
        repair_state<D, M>(P, acc);
    }

    template<int D, int M>
    static __device__ __forceinline__ void repair_cell(const Eos& P, Cell<D, M>& U) { repair_state<D, M>(P, U); }

    // CFL candidate (cfl.hpp:46-63 / 94-116); p_out receives the cell's
    // recovered pressure, which the next sweep reuses (kStoredPressure)
    template<int D, int M>
    static __device__ __forceinline__ double cfl_value(const Eos& P, const Cell<D, M>& U, const CflArgs& a,
                                                       double& p_out)
    {
        Prim<D, M> Q;
        dimhip::cons_to_prim<D, M>(P, U, Q);
        p_out = Q.p;
        const double c = mixture_sound_speed<M>(P, total_density<D, M>(U), Q.p, Q.alpha, Q.rho);
        double best = bits_to_double(kInfBits);
        if (!a.unsplit) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double ws = __builtin_fabs(Q.vel[d]) + c;
                if (ws > 1e-14) best = smin(best, a.cfl_dx[d] / ws);
            }
        } else {
            double spectral = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) spectral += (__builtin_fabs(Q.vel[d]) + c) / a.dx[d];
            if (spectral > 1e-14) best = smin(best, a.cfl / spectral);
        }
        return best;
    }

    template<int D, int M>
    static __device__ __forceinline__ double mass(const Cell<D, M>& U, int k) { return U.m[k]; }
    template<int D, int M>
    static __device__ __forceinline__ double mom(const Cell<D, M>& U, int d) { return U.mom[d]; }
    template<int D, int M>
    static __device__ __forceinline__ double energy(const Eos&, const Cell<D, M>& U) { return U.E; }
};

using Fp64 = Fp64T<false>;
using Fp64NA = Fp64T<true>;

}  // namespace dimhip
