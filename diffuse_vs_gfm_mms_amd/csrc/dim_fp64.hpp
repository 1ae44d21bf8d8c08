// dim_fp64.hpp — FP64 device arithmetic of the DIM update, faithful to the
// reference's evaluation order so results match it bit for bit.
//
// Compiled with -ffp-contract=off: every + - * / sqrt below is one IEEE-754
// double operation rounded to nearest, as on the reference's x86-64 SSE2
// build.  std::max/std::min are restated as (a<b)?b:a and (b<a)?b:a because
// their NaN behaviour differs from fmax/fmin.
//
// Templates: D = spatial dimension (1..3), M = number of materials (1..4).
#pragma once

#include <hip/hip_runtime.h>

#include "dim_params.hpp"

namespace dimhip {

// ---- src/math/numerical_safety.hpp -------------------------------------
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ bool finite(double x) { return __builtin_isfinite(x); }
// numerical_safety.hpp:12-18
__device__ __forceinline__ double clamp_min(double x, double tol = 1e-10)
{
    return finite(x) ? smax(x, tol) : tol;
}
// numerical_safety.hpp:21-24
__device__ __forceinline__ double finite_or(double x, double fb = 0.0) { return finite(x) ? x : fb; }
// numerical_safety.hpp:27-39
__device__ __forceinline__ double safe_denom(double d, double tol = 1e-10)
{
    if (!finite(d)) return (d < 0.0) ? -tol : tol;
    if (__builtin_fabs(d) < tol) return (d < 0.0) ? -tol : tol;
    return d;
}
// numerical_safety.hpp:41-45
__device__ __forceinline__ double safe_div(double n, double d, double tol = 1e-10)
{
    return finite_or(n) / safe_denom(d, tol);
}

template<int D, int M>
struct State {
    double m[M];
    double mom[D];
    double E;
    double a[M > 1 ? M - 1 : 1];
};

template<int D, int M>
struct Prim {
    double rho[M];
    double alpha[M];
    double vel[D];
    double p;
};

template<int D, int M>
struct Flux {
    double m[M];
    double mom[D];
    double E;
};

// ---- per-material EoS (src/euler/eos.hpp) --------------------------------
// pressure<1> of a zero-momentum Conserved<1>: ideal and stiffened gas
// (eos.hpp:451-466, 555-571) and noble-Abel (eos.hpp:665-677, whose
// internal energy is internal_energy_from_conserved, eos.hpp:13-25), for a
// density r that is already clamp_min'ed (finite, >= 1e-10).  The
// reference's kinetic energy 0.5 * safe_div(0, r) is then +0, and E - (+0)
// is E bit for bit (signed zeros, infinities and NaN included), so the
// division and the subtraction are not formed.
template<typename EP>
__device__ __forceinline__ double mat_pressure_c1_clamped(const EP& P, int k, double r, double E)
{
    if (EP::kNA && P.kind[k] == EOS_NOBLE_ABEL) {
        const double e = smax(safe_div(E, r), 0.0);
        const double denom = safe_denom(1.0 - P.covol[k] * r, 1e-14);
        return clamp_min(P.gm1[k] * r * e / denom);
    }
    if (P.kind[k] == EOS_STIFFENED) return clamp_min(P.gm1[k] * E - P.gpinf[k]);
    return clamp_min(P.gm1[k] * E);
}

// internal_energy (eos.hpp:482-489, 587-595, 692-702); gm1 = gamma-1,
// gpinf = gamma*p_inf.  The noble-Abel form clamps its inputs and divides
// without a guard, as the reference does.
template<typename EP>
__device__ __forceinline__ double mat_internal_energy(const EP& P, int k, double rho, double p)
{
    if (EP::kNA && P.kind[k] == EOS_NOBLE_ABEL) {
        const double r = clamp_min(rho);
        const double pp = clamp_min(p);
        return pp * (1.0 - P.covol[k] * r) / (P.gm1[k] * r);
    }
    if (P.kind[k] == EOS_STIFFENED) return safe_div(p + P.gpinf[k], P.gm1[k] * rho);
    return safe_div(p, P.gm1[k] * rho);
}

// eos.hpp:16-27
template<typename EP>
__device__ __forceinline__ double material_pressure_from_density_energy(const EP& P, int k,
                                                                         double rho, double e)
{
    const double r = clamp_min(rho);
    return mat_pressure_c1_clamped(P, k, r, r * smax(e, 0.0));
}

// eos.hpp:29-40 with sound_speed<1> (eos.hpp:469-479, 574-583, 680-689)
template<typename EP>
__device__ __forceinline__ double material_sound_speed(const EP& P, int k, double rho, double p)
{
    // clamp_min is idempotent: the reference's second clamp of r (in
    // sound_speed<1> and pressure<1>) returns r
    const double r = clamp_min(rho);
    const double E = r * mat_internal_energy(P, k, r, p);
    const double pp = mat_pressure_c1_clamped(P, k, r, E);
    if (EP::kNA && P.kind[k] == EOS_NOBLE_ABEL) {
        const double denom = safe_denom(r * (1.0 - P.covol[k] * r), 1e-14);
        return __builtin_sqrt(clamp_min(P.gamma[k] * pp / denom));
    }
    if (P.kind[k] == EOS_STIFFENED) return __builtin_sqrt(P.gamma[k] * safe_div(pp + P.pinf[k], r));
    return __builtin_sqrt(P.gamma[k] * safe_div(pp, r));
}

// eos.hpp:42-57
template<typename EP>
__device__ __forceinline__ double material_energy_pressure_derivative(const EP& P, int k,
                                                                      double rho, double p)
{
    rho = clamp_min(rho);
    p = clamp_min(p);
    const double dp = smax(1e-6 * smax(__builtin_fabs(p), 1.0), 1e-8);
    const double p_lo = smax(p - dp, 1e-12);
    const double p_hi = p + dp;
    const double e_lo = mat_internal_energy(P, k, rho, p_lo);
    const double e_hi = mat_internal_energy(P, k, rho, p_hi);
    return smax(safe_div(e_hi - e_lo, p_hi - p_lo, 1e-14), 1e-14);
}

// ---- mixture EoS (src/dim/eos.hpp) ---------------------------------------
// Loop-invariant part of the residual of the pressure bisection:
// coef_k = max(finite_or(alpha_k),0) * clamp_min(rho_k), den_k = (gamma_k-1)*clamp_min(rho_k)
// (guarded by safe_denom except for noble-Abel, whose internal energy divides
// unguarded), omb_k = 1 - b_k*clamp_min(rho_k) for noble-Abel materials.
// NA (EP::kNA of the parameter block) selects the variant that carries omb;
// EosParams (no noble-Abel material) compiles the residual without it.
template<int M, bool NA>
struct MixCoef {
    double coef[M];
    double den[M];
    double rden[M];  // RN(1/den): the residual's divisions by den become div_rn
    double omb[NA ? M : 1];
};

// a/b correctly rounded from y = RN(1/b) (Markstein's theorem: q = a*y is
// within one ulp of a/b, r = a - b*q is exact with an FMA, and RN(q + r*y) is
// RN(a/b)).  Valid for finite a and normal b with a/b, a*y and r = a - b*q
// in the normal range, which the bisection's operands satisfy (|a| >= 1e-10
// or a = 0, 1e-10*(gamma-1) <= b, |a/b| < 1e300): three FMA-pipe operations in
// place of the ten-instruction IEEE division sequence, with the same bits.
__device__ __forceinline__ double div_rn(double a, double b, double y)
{
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
}

template<int M, bool NA, typename EP>
__device__ __forceinline__ MixCoef<M, NA> mix_coef(const EP& P, const double* alpha, const double* rho)
{
    MixCoef<M, NA> c;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double rk = clamp_min(rho[k]);
        c.coef[k] = smax(finite_or(alpha[k]), 0.0) * rk;
        if (NA && P.kind[k] == EOS_NOBLE_ABEL) {
            c.den[k] = P.gm1[k] * rk;
            c.omb[k] = 1.0 - P.covol[k] * rk;
        } else {
            c.den[k] = safe_denom(P.gm1[k] * rk);
        }
        c.rden[k] = 1.0 / c.den[k];
    }
    return c;
}

// den_k * e_k(rho_k, p) as the reference forms it before its division, at
// pk = clamp_min(p): ideal p (guarded), stiffened p + gamma p_inf (guarded),
// noble-Abel p (1 - b rho_k) (eos.hpp:482-489, 587-595, 692-702)
template<int M, bool NA, typename EP>
__device__ __forceinline__ double mix_num(const EP& P, const MixCoef<M, NA>& c, int k, double pk)
{
    if (NA && P.kind[k] == EOS_NOBLE_ABEL) return pk * c.omb[NA ? k : 0];
    return finite_or((P.kind[k] == EOS_STIFFENED) ? (pk + P.gpinf[k]) : pk);
}

// eos.hpp:59-75 (mixture_internal_energy_density) with the invariants hoisted
template<int M, bool NA, typename EP>
__device__ __forceinline__ double mix_energy(const EP& P, const MixCoef<M, NA>& c, double p)
{
    const double pk = clamp_min(p);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) s += c.coef[k] * div_rn(mix_num<M, NA>(P, c, k, pk), c.den[k], c.rden[k]);
    return finite_or(s);
}

// The certain halvings of the bisection, as one wave-uniform loop of at most
// 80 passes.  On entry `go` (a subset of the active lanes) holds the lanes
// that may fast-forward; pass p (p = 0, 1, ...) forms mid = 0.5 * (lo + hi),
// the reference's midpoint, on every lane still going, and a lane whose mid
// lies above p_up takes hi = mid, below p_dn lo = mid; a lane whose mid lies
// in [p_dn, p_up] stops with n = p (the halvings it took), and a lane still
// going after pass 79 ends with n = 80.  Lanes outside `go` keep lo, hi, n.
// The per-lane choice is made with the execution mask (64-bit moves under
// the lanes that go up, then those that go down), so a pass costs two FP64
// operations, two compares and two moves.  The loop runs on scalar registers
// only (counter, masks), touches no memory and restores the execution mask.
__device__ __forceinline__ void fast_forward_halvings(double& lo, double& hi, int& n, double p_up, double p_dn,
                                                      unsigned long long go)
{
    double mid;
    unsigned long long save, up, dn, t1, t2, t3;
    unsigned cnt;
    asm volatile(
        "s_mov_b64 %[save], exec\n\t"
        "s_and_b64 exec, exec, %[go]\n\t"
        "s_mov_b32 %[cnt], 0\n\t"
        "s_cbranch_execz .Lff_end_%=\n"
        ".Lff_loop_%=:\n\t"
        "v_add_f64 %[mid], %[lo], %[hi]\n\t"
        "v_mul_f64 %[mid], %[mid], 0.5\n\t"
        "v_cmp_gt_f64_e64 %[up], %[mid], %[pup]\n\t"
        "v_cmp_lt_f64_e64 %[dn], %[mid], %[pdn]\n\t"
        "s_or_b64 %[t1], %[up], %[dn]\n\t"
        "s_andn2_b64 %[t2], exec, %[t1]\n\t"
        "s_cmp_eq_u64 %[t2], 0\n\t"
        "s_cbranch_scc1 .Lff_keep_%=\n\t"
        "s_mov_b64 %[t3], exec\n\t"
        "s_mov_b64 exec, %[t2]\n\t"
        "v_mov_b32 %[n], %[cnt]\n\t"
        "s_mov_b64 exec, %[t3]\n"
        ".Lff_keep_%=:\n\t"
        "s_and_b64 exec, exec, %[t1]\n\t"
        "s_and_b64 %[t2], exec, %[up]\n\t"
        "s_andn2_b64 %[t3], exec, %[up]\n\t"
        "s_mov_b64 %[t1], exec\n\t"
        "s_mov_b64 exec, %[t2]\n\t"
        "v_mov_b64 %[hi], %[mid]\n\t"
        "s_mov_b64 exec, %[t3]\n\t"
        "v_mov_b64 %[lo], %[mid]\n\t"
        "s_mov_b64 exec, %[t1]\n\t"
        "s_add_u32 %[cnt], %[cnt], 1\n\t"
        "s_cmp_eq_u32 %[cnt], 80\n\t"
        "s_cbranch_scc1 .Lff_full_%=\n\t"
        "s_cbranch_execnz .Lff_loop_%=\n\t"
        "s_branch .Lff_end_%=\n"
        ".Lff_full_%=:\n\t"
        "v_mov_b32 %[n], 80\n"
        ".Lff_end_%=:\n\t"
        "s_mov_b64 exec, %[save]"
        : [lo] "+v"(lo), [hi] "+v"(hi), [n] "+v"(n), [mid] "=&v"(mid), [save] "=&s"(save), [up] "=&s"(up),
          [dn] "=&s"(dn), [t1] "=&s"(t1), [t2] "=&s"(t2), [t3] "=&s"(t3), [cnt] "=&s"(cnt)
        : [go] "s"(go), [pup] "v"(p_up), [pdn] "v"(p_dn)
        : "scc");
}

// eos.hpp:95-142: bracket doubling from [1e-12, max(1, fallback)], then
// bisection to |f| <= 1e-10 max(1, rho e)
template<int M, bool NA, typename EP>
__device__ __forceinline__ double bisect_pressure(const EP& P, double target, double fallback,
                                                  const double* alpha, const double* rho)
{
    const MixCoef<M, NA> c = mix_coef<M, NA>(P, alpha, rho);
    double lo = 1e-12;
    double hi = smax(1.0, fallback);
    const double flo = mix_energy<M, NA>(P, c, lo) - target;
    double fhi = mix_energy<M, NA>(P, c, hi) - target;
    for (int n = 0; n < 80 && finite(fhi) && fhi < 0.0; ++n) {
        hi *= 2.0;
        fhi = mix_energy<M, NA>(P, c, hi) - target;
    }
    if (!finite(flo) || !finite(fhi) || !(flo <= 0.0 && fhi >= 0.0)) return clamp_min(fallback);

    const double tol = 1e-10 * smax(1.0, target);
    int n = 0;
    // Fast-forward over the halvings whose outcome is certain.  For pk =
    // clamp_min(p) the residual is linear up to rounding: f(p) = S pk + C -
    // target with S = sum_k coef_k s_k / den_k and C = sum_k coef_k b_k / den_k
    // (numerator s_k pk + b_k: s = 1 or 1 - b rho_k, b = 0 or gamma p_inf).
    // The rounded residual fmid that the loop below would compute differs from
    // e = fma(S~, pk, C~ - target) by at most 15.4 u B, where B = Sabs pk +
    // Cabs + |target| bounds the magnitudes involved (Sabs, Cabs: the same
    // sums of absolute values; u = 2^-53; the three roundings of each term,
    // M - 1 additions and the subtraction of target in the residual, and those
    // of S~, C~ and the FMA in e).  So where |e| > tol + 32 u B the residual is
    // finite, |fmid| > tol (no return) and fmid has the sign of e: the halving
    // is taken from e without the M divisions.  The midpoints are the same
    // expressions in the same order, so the result is the same bits.  Near
    // the root (|e| <= tol + 32 u B) the exact evaluation takes over.
    {
        double S = 0.0, C = 0.0, Sabs = 0.0, Cabs = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double w = c.coef[k] * c.rden[k];
            const double sk = (NA && P.kind[k] == EOS_NOBLE_ABEL) ? c.omb[NA ? k : 0] : 1.0;
            const double bk = (P.kind[k] == EOS_STIFFENED) ? P.gpinf[k] : 0.0;
            S += w * sk;
            C += w * bk;
            Sabs += w * __builtin_fabs(sk);
            Cabs += w * __builtin_fabs(bk);
        }
        const double Cmt = C - target;
        const double B = Sabs * smax(hi, 1e-10) + Cabs + __builtin_fabs(target);  // pk <= max(hi, 1e-10) from here on
        const double thr = tol + 3.5527136788005009e-15 * B;                      // 32 u B
        // The halvings run as a wave-uniform loop (fast_forward_halvings):
        // each lane goes on until its first uncertain halving, and n counts the
        // halvings it took (its loop index on leaving the per-lane loop this
        // replaces).
        //
        // The test |e| > thr is taken on mid itself.  With S~ > 0 and L(p) =
        // S~ p + Cmt in exact arithmetic, e = RN(L(pk)), so L(pk) > T =
        // RN(thr (1 + 4u)) >= thr (1 + 2u) gives e >= L (1 - u) > thr, and
        // L(pk) < -T gives e < -thr.  L(pk) > T holds for pk > q_up = (T -
        // Cmt) / S~ and L(pk) < -T for pk < q_dn = (-T - Cmt) / S~.  Their
        // rounded values (one reciprocal of S~, a subtraction and a product)
        // carry a relative error below 3.01 u; moved outward by 2^-50 = 8u of
        // their magnitude (p_up up, p_dn down, either sign, the move's own
        // rounding included) they bound the exact ones.  lo and hi are finite
        // here (hi passed the finite(fhi) test) and lo >= 1e-12, so mid is
        // finite and positive, and pk = max(mid, 1e-10) >= mid: mid > p_up is
        // a certain upward halving; mid < p_dn is a certain downward one when
        // p_dn > 1e-10 (then pk < p_dn whichever of mid and 1e-10 it is), and
        // is never taken otherwise (p_dn = 0).  p_dn < p_up, and every mid with
        // |e| <= thr lies in [p_dn, p_up], so a lane stops at or before the
        // halving where the test on e would stop it.
        const double T = thr * 1.0000000000000004;  // 1 + 2^-51
        const double rS = 1.0 / S;
        const double q_up = (T - Cmt) * rS, q_dn = (-T - Cmt) * rS;
        const double p_up = q_up + __builtin_fabs(q_up) * 8.8817841970012523e-16;  // 2^-50
        double p_dn = q_dn - __builtin_fabs(q_dn) * 8.8817841970012523e-16;
        p_dn = (p_dn > 1e-10) ? p_dn : 0.0;
        const bool go = finite(S) && S > 0.0 && finite(Cmt) && B < 1e290 && finite(p_up) && finite(p_dn);
        fast_forward_halvings(lo, hi, n, p_up, p_dn, __builtin_amdgcn_ballot_w64(go));
    }
    for (; n < 80; ++n) {
        const double mid = 0.5 * (lo + hi);
        const double fmid = mix_energy<M, NA>(P, c, mid) - target;
        if (!finite(fmid)) return clamp_min(fallback);
        if (__builtin_fabs(fmid) <= tol) return clamp_min(mid);
        if (fmid > 0.0) hi = mid;
        else lo = mid;
    }
    return clamp_min(0.5 * (lo + hi));
}

// eos.hpp:77-143 — bracket doubling + bisection with the reference's
// midpoints, decisions and stopping test; loop-invariant subexpressions are
// hoisted (mix_coef) and the halvings far from the root are decided from the
// linear form of the residual (bisect_pressure), which leaves every rounded
// value unchanged.
template<int M, typename EP>
__device__ double pressure_from_internal_energy(const EP& P, double rho_total, double e,
                                                const double* alpha, const double* rho)
{
    rho_total = clamp_min(rho_total);
    e = finite_or(e);
    const double target = finite_or(rho_total * e);
    double fallback = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double ek = target / rho_total;
        fallback += smax(finite_or(alpha[k]), 0.0) *
                    material_pressure_from_density_energy(P, k, clamp_min(rho[k]), ek);
    }
    fallback = clamp_min(fallback);
    return bisect_pressure<M, EP::kNA>(P, target, fallback, alpha, rho);
}

// eos.hpp:59-75 at one pressure with plain divisions (same bits as div_rn)
template<int M, bool NA, typename EP>
__device__ __forceinline__ double mix_energy_once(const EP& P, double p, const double* alpha,
                                                  const double* rho)
{
    const MixCoef<M, NA> c = mix_coef<M, NA>(P, alpha, rho);
    const double pk = clamp_min(p);
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) sum += c.coef[k] * (mix_num<M, NA>(P, c, k, pk) / c.den[k]);
    return finite_or(sum);
}

template<int M, typename EP>
__device__ __forceinline__ double mixture_internal_energy_density(const EP& P, double p,
                                                                  const double* alpha, const double* rho)
{
    return mix_energy_once<M, EP::kNA>(P, p, alpha, rho);
}

// eos.hpp:163-181 via eos.hpp:59-75
template<int M, typename EP>
__device__ __forceinline__ double internal_energy_from_pressure(const EP& P, double rho_total,
                                                                double p, const double* alpha,
                                                                const double* rho)
{
    rho_total = clamp_min(rho_total);
    p = clamp_min(p);
    return mixture_internal_energy_density<M>(P, p, alpha, rho) / rho_total;
}

// eos.hpp:195-221 (Wood) and eos.hpp:223-275 (allaire / equal_velocity).
// Always inlined: left to the compiler, the three-material and noble-Abel
// sweeps called it, and the call saved their live registers to scratch
// (368-384 B per lane) at every sound speed.
template<int M, typename EP>
__device__ __forceinline__ double mixture_sound_speed(const EP& P, double rho_total, double p,
                                      const double* alpha, const double* rho)
{
    if (P.lambda_model == LAMBDA_KAPILA) {
        rho_total = clamp_min(rho_total);
        p = clamp_min(p);
        double beta_mix = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double rho_k = clamp_min(rho[k]);
            const double ck = material_sound_speed(P, k, rho_k, p);
            beta_mix += smax(finite_or(alpha[k]), 0.0) / safe_denom(rho_k * ck * ck, 1e-14);
        }
        return __builtin_sqrt(1.0 / (rho_total * safe_denom(beta_mix, 1e-14)));
    }
    double cs[M], xi[M];
    double xi_mix = 0.0, inverse_sound_sum = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double alpha_k = smax(finite_or(alpha[k]), 0.0);
        const double rho_k = clamp_min(rho[k]);
        cs[k] = material_sound_speed(P, k, rho_k, p);
        xi[k] = rho_k * material_energy_pressure_derivative(P, k, rho_k, p);
        xi_mix += alpha_k * xi[k];
        inverse_sound_sum += alpha_k / safe_denom(cs[k], 1e-14);
    }
    xi_mix = safe_denom(xi_mix, 1e-14);
    inverse_sound_sum = safe_denom(inverse_sound_sum, 1e-14);
    double rc2 = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double alpha_k = smax(finite_or(alpha[k]), 0.0);
        double lambda_k = 1.0;
        if (P.lambda_model == LAMBDA_EQUAL_VELOCITY)
            lambda_k = safe_div(1.0 / safe_denom(cs[k], 1e-14), inverse_sound_sum, 1e-14);
        rc2 += lambda_k * alpha_k * xi[k] * cs[k] * cs[k] / xi_mix;
    }
    return __builtin_sqrt(smax(rc2 / clamp_min(rho_total), 1e-14));
}

// ---- state algebra (src/dim/primitives.hpp) ---------------------------
// primitives.hpp:18-42
template<int D, int M>
__device__ __forceinline__ void full_alpha(const State<D, M>& U, double* alpha)
{
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < M - 1; ++k) {
        alpha[k] = U.a[k];
        sum += alpha[k];
    }
    alpha[M - 1] = 1.0 - sum;
}

// primitives.hpp:45-70
template<int M>
__device__ __forceinline__ void sanitise_alpha(double* alpha, double floor)
{
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        alpha[k] = finite_or(alpha[k], floor);
        alpha[k] = smax(alpha[k], floor);
        sum += alpha[k];
    }
    if (!finite(sum) || sum <= 0.0) {
#pragma unroll
        for (int k = 0; k < M; ++k) alpha[k] = (k == 0) ? 1.0 : 0.0;
        return;
    }
    // x / 1.0 is x for every finite x, and the fractions are finite here: the
    // M divisions run only where the sum differs from one (for two materials
    // stored as alpha_0 and 1 - alpha_0 the rounded sum is exactly one)
    if (sum != 1.0) {
#pragma unroll
        for (int k = 0; k < M; ++k) alpha[k] /= sum;
    }
}

// primitives.hpp:99-108
template<int D, int M>
__device__ __forceinline__ double total_density(const State<D, M>& U)
{
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) r += finite_or(U.m[k]);
    return r;
}

// cons_to_prim up to the pressure recovery: fractions, velocities and
// material densities into R; returns rho_total and sets the specific
// internal energy ie
template<int D, int M>
__device__ __forceinline__ double cons_to_prim_base(const State<D, M>& U, Prim<D, M>& R, double& ie)
{
    full_alpha<D, M>(U, R.alpha);
    sanitise_alpha<M>(R.alpha, 1e-12);
    double rho_total = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) rho_total += finite_or(U.m[k]);
    rho_total = smax(rho_total, 1e-12);
    double v2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        R.vel[d] = safe_div(U.mom[d], rho_total);
        v2 += R.vel[d] * R.vel[d];
    }
#pragma unroll
    for (int k = 0; k < M; ++k)
        R.rho[k] = smax(safe_div(finite_or(U.m[k]), smax(R.alpha[k], 1e-12)), 1e-12);
    const double kinetic = 0.5 * rho_total * v2;
    ie = safe_div(finite_or(U.E, kinetic) - kinetic, rho_total);
    return rho_total;
}

// primitives.hpp:111-162
template<int D, int M, typename EP>
__device__ __forceinline__ void cons_to_prim(const EP& P, const State<D, M>& U, Prim<D, M>& R)
{
    double ie;
    const double rho_total = cons_to_prim_base<D, M>(U, R, ie);
    R.p = smax(pressure_from_internal_energy<M>(P, rho_total, ie, R.alpha, R.rho), 1e-12);
}

// primitives.hpp:166-212
template<int D, int M, typename EP>
__device__ __forceinline__ void prim_to_cons(const EP& P, const Prim<D, M>& Pin, State<D, M>& U)
{
    double alpha[M], rho[M];
#pragma unroll
    for (int k = 0; k < M; ++k) alpha[k] = Pin.alpha[k];
    sanitise_alpha<M>(alpha, 1e-12);
    double rho_total = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        rho[k] = smax(Pin.rho[k], 1e-12);
        U.m[k] = alpha[k] * rho[k];
        rho_total += U.m[k];
    }
    rho_total = smax(rho_total, 1e-12);
    const double p = smax(Pin.p, 1e-12);
    double v2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        U.mom[d] = rho_total * Pin.vel[d];
        v2 += Pin.vel[d] * Pin.vel[d];
    }
    const double ie = internal_energy_from_pressure<M>(P, rho_total, p, alpha, rho);
    U.E = rho_total * ie + 0.5 * rho_total * v2;
#pragma unroll
    for (int k = 0; k < M - 1; ++k) U.a[k] = alpha[k];
}

// primitives.hpp:216-322
template<int D, int M, typename EP>
__device__ __forceinline__ void alpha_rhs_coefficients(const EP& P, const Prim<D, M>& Pr,
                                                       double* rhs)
{
    const double floor = 1e-12;
#pragma unroll
    for (int k = 0; k < M; ++k) rhs[k] = Pr.alpha[k];
    sanitise_alpha<M>(rhs, 0.0);
    if (P.lambda_model == LAMBDA_ALLAIRE) return;
    const double pp = smax(finite_or(Pr.p), floor);
    if (P.lambda_model == LAMBDA_EQUAL_VELOCITY) {
        double inv_sum = 0.0;
        double cs[M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            cs[k] = floor;
            const double alpha_k = smax(finite_or(rhs[k]), 0.0);
            if (alpha_k <= floor || alpha_k < smax(P.alpha_source_floor, 0.0)) continue;
            cs[k] = material_sound_speed(P, k, smax(finite_or(Pr.rho[k]), floor), pp);
            inv_sum += alpha_k / safe_denom(cs[k], 1e-14);
        }
        if (inv_sum <= floor || !finite(inv_sum)) return;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double alpha_k = smax(finite_or(rhs[k]), 0.0);
            if (alpha_k <= floor) {
                rhs[k] = 0.0;
                continue;
            }
            rhs[k] = alpha_k * safe_div(1.0 / safe_denom(cs[k], 1e-14), inv_sum, 1e-14);
        }
        return;
    }
    const double source_floor = smax(P.alpha_source_floor, 0.0);
    double beta_mix = 0.0;
    double imp[M];
    bool active[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        imp[k] = floor;
        active[k] = false;
        const double alpha_k = smax(finite_or(rhs[k]), 0.0);
        if (alpha_k <= floor || alpha_k < source_floor) continue;
        active[k] = true;
        const double rho_k = smax(finite_or(Pr.rho[k]), floor);
        const double c_k = material_sound_speed(P, k, rho_k, pp);
        imp[k] = smax(rho_k * c_k * c_k, floor);
        beta_mix += alpha_k / imp[k];
    }
    if (beta_mix <= floor || !finite(beta_mix)) return;
    const double q_mix = 1.0 / beta_mix;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double alpha_k = smax(finite_or(rhs[k]), 0.0);
        if (!active[k] || alpha_k <= floor) {
            rhs[k] = 0.0;
            continue;
        }
        rhs[k] = alpha_k + alpha_k * (q_mix / imp[k] - 1.0);
    }
}

// primitives.hpp:326-457.  The reference runs a pressure bisection
// (pressure_before) on every call but reads it only when inactive mass is
// removed; here it is evaluated on that branch only, from the same inputs.
template<int D, int M, typename EP>
__device__ __forceinline__ void repair_state(const EP& P, State<D, M>& U)
{
    const double mass_floor = 1e-12, p_floor = 1e-12, inactive = 1e-10;
#pragma unroll
    for (int k = 0; k < M; ++k) U.m[k] = smax(finite_or(U.m[k]), 0.0);
#pragma unroll
    for (int k = 0; k < M - 1; ++k) U.a[k] = finite_or(U.a[k]);
#pragma unroll
    for (int d = 0; d < D; ++d) U.mom[d] = finite_or(U.mom[d]);
    U.E = finite_or(U.E);

    // rho_before, vel_before and v2_before are read only where inactive mass
    // is removed, and are formed there from the same (unchanged) momenta
    const double rho_before = smax(total_density<D, M>(U), mass_floor);
    // alpha_before equals alpha_full below: both are sanitise(full_alpha(U.alpha), 0)
    // of the same (already finite_or'd) values
    double alpha_full[M];
    full_alpha<D, M>(U, alpha_full);
    sanitise_alpha<M>(alpha_full, 0.0);
#pragma unroll
    for (int k = 0; k < M - 1; ++k) U.a[k] = alpha_full[k];

    bool removed = false;
    double m_before[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        m_before[k] = U.m[k];
        if (alpha_full[k] <= inactive && U.m[k] > mass_floor) {
            U.m[k] = 0.0;
            removed = true;
        }
    }
    double rho_total = total_density<D, M>(U);
    if (!finite(rho_total) || rho_total <= 0.0) {
        int dom = 0;
#pragma unroll
        for (int k = 1; k < M; ++k)
            if (alpha_full[k] > alpha_full[dom]) dom = k;
#pragma unroll
        for (int k = 0; k < M; ++k)
            if (k == dom) U.m[k] = mass_floor;
        rho_total = mass_floor;
    }
    if (removed) {
        double vel_before[D];
        double v2_before = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            vel_before[d] = safe_div(U.mom[d], rho_before);
            v2_before += vel_before[d] * vel_before[d];
        }
        double rho_mb[M];
#pragma unroll
        for (int k = 0; k < M; ++k)
            rho_mb[k] = smax(safe_div(smax(m_before[k], 0.0), smax(alpha_full[k], inactive)), mass_floor);
        const double kinetic_before = 0.5 * rho_before * v2_before;
        // U.E is finite here, so finite_or(U.E, kinetic_before) == U.E
        const double internal_before = safe_div(U.E - kinetic_before, rho_before);
        const double pressure_before = smax(
            pressure_from_internal_energy<M>(P, rho_before, internal_before, alpha_full, rho_mb), p_floor);
#pragma unroll
        for (int d = 0; d < D; ++d) U.mom[d] = rho_total * vel_before[d];
        double rho_m[M];
#pragma unroll
        for (int k = 0; k < M; ++k)
            rho_m[k] = smax(safe_div(U.m[k], smax(alpha_full[k], inactive)), mass_floor);
        const double ie = internal_energy_from_pressure<M>(P, rho_total, pressure_before, alpha_full, rho_m);
        U.E = rho_total * ie + 0.5 * rho_total * v2_before;
    }
    // The energy floor U.E = max(U.E, kinetic + minimum_internal) leaves U.E
    // unchanged where U.E exceeds the floor's rounded value R, which is tested
    // first without its divisions.  With rho = rho_total in [1e-10, 1e300)
    // (safe_div's denominator is then rho itself), the reference's kinetic =
    // RN(RN(0.5 rho) v2), v2 the rounded sum of RN(RN(mom_d / rho)^2), is at
    // most 0.5 sum_d mom_d^2 / rho (1 + 8.01u) for D <= 3, and each term
    // RN(coef_k RN(num_k / den_k)) of minimum_internal at most alpha_k
    // emin_ub[k] (1 + 5.01u) (coef = alpha rho_k, den = (gamma-1) rho_k or its
    // 1e-10 guard, which only lowers the term); with the M <= 4 additions and
    // the final one R <= (0.5 sum mom^2 / rho + sum_k alpha_k emin_ub[k])
    // (1 + 14u).  Below, rhs >= (0.5 m2 + rho a)(1 + 2^-40)(1 - 7.01u) with
    // m2 = sum mom^2 and a = sum alpha emin_ub exact (at most max(D, M + 1) +
    // 2 roundings, each lowering it by at most u), and lhs <= rho U.E (1 + u),
    // so lhs > rhs gives U.E > (0.5 m2 / rho + a)(1 + 2^-40 - 8.02u) > R
    // (2^-40 = 8192u): the floor is skipped with the same result.  Otherwise
    // (U.E <= 0, or a non-finite term) the floor is evaluated.
    bool floor_needed = true;
    if (rho_total >= 1e-10 && rho_total < 1e300) {
        double m2 = 0.0, a = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) m2 += U.mom[d] * U.mom[d];
#pragma unroll
        for (int k = 0; k < M; ++k) a += alpha_full[k] * P.emin_ub[k];
        const double rhs = (0.5 * m2 + rho_total * a) * 1.0000000000009095;  // 1 + 2^-40
        const double lhs = U.E * rho_total;
        floor_needed = !(lhs > rhs && lhs < 1.7e308);
    }
    if (floor_needed) {
        double v2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double v = safe_div(U.mom[d], rho_total);
            v2 += v * v;
        }
        double rho_m[M];
#pragma unroll
        for (int k = 0; k < M; ++k) rho_m[k] = smax(safe_div(U.m[k], smax(alpha_full[k], inactive)), mass_floor);
        const double kinetic = 0.5 * rho_total * v2;
        const double minimum_internal = mixture_internal_energy_density<M>(P, p_floor, alpha_full, rho_m);
        U.E = smax(finite_or(U.E), kinetic + minimum_internal);
    }
}

// ---- reconstruction (src/core/fv/limiters.hpp, src/dim/reconstruction/muscl.hpp)
// limiters.hpp:9-15
__device__ __forceinline__ double minmod(double a, double b)
{
    if (a * b <= 0.0) return 0.0;
    return __builtin_fabs(a) < __builtin_fabs(b) ? a : b;
}

// muscl.hpp:26-40
template<int D, int M>
__device__ __forceinline__ void enforce_positive_primitive(Prim<D, M>& Q)
{
#pragma unroll
    for (int k = 0; k < M; ++k) Q.rho[k] = smax(Q.rho[k], 1e-12);
    sanitise_alpha<M>(Q.alpha, 0.0);
    Q.p = smax(Q.p, 1e-12);
}

// muscl.hpp:43-112
template<int D, int M>
__device__ __forceinline__ void reconstruct_cell_faces(const Prim<D, M>& L, const Prim<D, M>& C,
                                                       const Prim<D, M>& R, Prim<D, M>& lf,
                                                       Prim<D, M>& rf)
{
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double s = minmod(C.rho[k] - L.rho[k], R.rho[k] - C.rho[k]);
        lf.rho[k] = C.rho[k] - 0.5 * s;
        rf.rho[k] = C.rho[k] + 0.5 * s;
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double s = minmod(C.alpha[k] - L.alpha[k], R.alpha[k] - C.alpha[k]);
        lf.alpha[k] = C.alpha[k] - 0.5 * s;
        rf.alpha[k] = C.alpha[k] + 0.5 * s;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double s = minmod(C.vel[d] - L.vel[d], R.vel[d] - C.vel[d]);
        lf.vel[d] = C.vel[d] - 0.5 * s;
        rf.vel[d] = C.vel[d] + 0.5 * s;
    }
    const double s = minmod(C.p - L.p, R.p - C.p);
    lf.p = C.p - 0.5 * s;
    rf.p = C.p + 0.5 * s;
    enforce_positive_primitive<D, M>(lf);
    enforce_positive_primitive<D, M>(rf);
}

// dot(vel, e_dir) as the reference sums it (vector_ops.hpp:9-19): a running
// sum from +0.0 over every component, so -0.0 becomes +0.0
template<int D, int DIR>
__device__ __forceinline__ double normal_velocity(const double* vel)
{
    double un = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) un += vel[d] * (d == DIR ? 1.0 : 0.0);
    return un;
}

// flux.hpp:11-43 with normal = e_DIR
template<int D, int M, int DIR>
__device__ __forceinline__ void flux_normal(const State<D, M>& U, const Prim<D, M>& Q, Flux<D, M>& F)
{
    const double rho_total = total_density<D, M>(U);
    const double un = normal_velocity<D, DIR>(Q.vel);
#pragma unroll
    for (int k = 0; k < M; ++k) F.m[k] = U.m[k] * un;
#pragma unroll
    for (int d = 0; d < D; ++d) F.mom[d] = rho_total * Q.vel[d] * un + Q.p * (d == DIR ? 1.0 : 0.0);
    F.E = (U.E + Q.p) * un;
}

// muscl.hpp:143-191
template<int D, int M, int DIR, typename EP>
__device__ __forceinline__ void hancock_predict(const EP& P, const State<D, M>& Uc,
                                                const Prim<D, M>& Pc, const State<D, M>& ULf,
                                                const State<D, M>& URf, const Prim<D, M>& PLf,
                                                const Prim<D, M>& PRf, const double* rhs, double dt, double dx,
                                                State<D, M>& out)
{
    Flux<D, M> FL, FR;
    flux_normal<D, M, DIR>(ULf, PLf, FL);
    flux_normal<D, M, DIR>(URf, PRf, FR);
    const double factor = 0.5 * dt / dx;
    out = Uc;
#pragma unroll
    for (int k = 0; k < M; ++k) out.m[k] -= factor * (FR.m[k] - FL.m[k]);
#pragma unroll
    for (int d = 0; d < D; ++d) out.mom[d] -= factor * (FR.mom[d] - FL.mom[d]);
    out.E -= factor * (FR.E - FL.E);
    const double u_left = PLf.vel[DIR];
    const double u_right = PRf.vel[DIR];
    const double div_u = (u_right - u_left) / dx;
    // rhs = alpha_rhs_coefficients(Pc), computed by the caller
    double an[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        an[k] = Pc.alpha[k];
        an[k] -= factor * (PRf.alpha[k] * u_right - PLf.alpha[k] * u_left);
        an[k] += 0.5 * dt * rhs[k] * div_u;
    }
    sanitise_alpha<M>(an, 0.0);
#pragma unroll
    for (int k = 0; k < M - 1; ++k) out.a[k] = an[k];
    repair_state<D, M>(P, out);
}

// ---- HLLC (src/dim/riemann/hllc.hpp:15-195) ------------------------------
// Returns false on the reference's throwing branch (non-finite density or
// pressure in require_positive, numerical_safety.hpp:67-78).
template<int D, int M, int DIR, typename EP>
__device__ __forceinline__ int hllc_flux(const EP& P, const State<D, M>& UL,
                                          const State<D, M>& UR, const Prim<D, M>& PL,
                                          const Prim<D, M>& PR, Flux<D, M>& F, double& fv)
{
    double rhoL = total_density<D, M>(UL);
    double rhoR = total_density<D, M>(UR);
    double pL = PL.p, pR = PR.p;
    // require_positive on each input in the reference's order: a non-finite
    // one fails the face (status = its error bit), a non-positive one is raised
    if (!finite(rhoL)) return ERR_HLLC_LEFT_DENSITY;
    if (!finite(rhoR)) return ERR_HLLC_RIGHT_DENSITY;
    if (!finite(pL)) return ERR_HLLC_LEFT_PRESSURE;
    if (!finite(pR)) return ERR_HLLC_RIGHT_PRESSURE;
    if (rhoL <= 0.0) rhoL = 1e-10;
    if (rhoR <= 0.0) rhoR = 1e-10;
    if (pL <= 0.0) pL = 1e-10;
    if (pR <= 0.0) pR = 1e-10;
    const double unL = normal_velocity<D, DIR>(PL.vel);
    const double unR = normal_velocity<D, DIR>(PR.vel);
    const double cL = mixture_sound_speed<M>(P, rhoL, pL, PL.alpha, PL.rho);
    const double cR = mixture_sound_speed<M>(P, rhoR, pR, PR.alpha, PR.rho);
    const double sL = smin(unL - cL, unR - cR);
    const double sR = smax(unL + cL, unR + cR);
    const double numerator = (pR - pL) + rhoL * unL * (sL - unL) - rhoR * unR * (sR - unR);
    const double denom = rhoL * (sL - unL) - rhoR * (sR - unR);
    double s_star = safe_div(numerator, denom);
    if (__builtin_fabs(pL - pR) < 1e-10 && __builtin_fabs(unL - unR) < 1e-10) s_star = unL;

    Flux<D, M> FL, FR;
    flux_normal<D, M, DIR>(UL, PL, FL);
    flux_normal<D, M, DIR>(UR, PR, FR);
    const bool average = !finite(sL) || !finite(sR) || !finite(s_star) ||
                         (!(0.0 <= sL) && !(sR <= 0.0) && !(sL < sR));
    if (average) {
#pragma unroll
        for (int k = 0; k < M; ++k) F.m[k] = 0.5 * (FL.m[k] + FR.m[k]);
#pragma unroll
        for (int d = 0; d < D; ++d) F.mom[d] = 0.5 * (FL.mom[d] + FR.mom[d]);
        F.E = 0.5 * (FL.E + FR.E);
        fv = 0.5 * (unL + unR);
        return 0;
    }
    if (0.0 <= sL) {
        F = FL;
        fv = unL;
        return 0;
    }
    if (sR <= 0.0) {
        F = FR;
        fv = unR;
        return 0;
    }
    const bool left = (0.0 <= s_star);  // sL < 0 here
    const State<D, M>& UK = left ? UL : UR;
    const Prim<D, M>& PK = left ? PL : PR;
    const Flux<D, M>& FK = left ? FL : FR;
    const double rhoK = left ? rhoL : rhoR;
    const double pK = left ? pL : pR;
    const double sK = left ? sL : sR;
    const double unK = left ? unL : unR;
    const double rho_star = rhoK * safe_div(sK - unK, sK - s_star);
    const double E_star =
        rho_star * (safe_div(UK.E, rhoK) + (s_star - unK) * (s_star + safe_div(pK, rhoK * (sK - unK))));
    fv = s_star;
    if (rho_star <= 0.0 || !finite(rho_star) || !finite(E_star)) {
        // hll_flux (hllc.hpp:47-59): (sR FL - sL FR + sL sR (UR-UL)) * (1/(sR-sL))
        const double inv = 1.0 / (sR - sL);
        const double sLsR = sL * sR;
#pragma unroll
        for (int k = 0; k < M; ++k) F.m[k] = inv * ((sR * FL.m[k] - sL * FR.m[k]) + sLsR * (UR.m[k] - UL.m[k]));
#pragma unroll
        for (int d = 0; d < D; ++d)
            F.mom[d] = inv * ((sR * FL.mom[d] - sL * FR.mom[d]) + sLsR * (UR.mom[d] - UL.mom[d]));
        F.E = inv * ((sR * FL.E - sL * FR.E) + sLsR * (UR.E - UL.E));
        return 0;
    }
    const double mass_scale = safe_div(rho_star, rhoK);
    const double delta = s_star - unK;
#pragma unroll
    for (int k = 0; k < M; ++k) F.m[k] = FK.m[k] + sK * (UK.m[k] * mass_scale - UK.m[k]);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double u_star = PK.vel[d];
        u_star += delta * (d == DIR ? 1.0 : 0.0);
        F.mom[d] = FK.mom[d] + sK * (rho_star * u_star - UK.mom[d]);
    }
    F.E = FK.E + sK * (E_star - UK.E);
    return 0;
}

}  // namespace dimhip
