// policy_fp32.hpp — the mixed-precision policy of the generic DIM kernels.
//
// Same scheme as the reference (two-stage MUSCL-Hancock, HLLC, Kapila alpha
// source, repair, ring BCs), different arithmetic:
//  * reconstruction, fluxes, half step and the update of momentum, energy and
//    volume fraction run in FP32 and are stored in FP32;
//  * partial masses are stored in FP64 and updated in FP64 from the FP32
//    face fluxes, so mass telescopes to FP64 rounding;
//  * the mixture pressure is the closed-form root of
//    rho e = sum_k alpha_k (p + gamma_k p_inf_k)/(gamma_k - 1) (the equation the
//    reference bisects, src/dim/eos.hpp:77-143), evaluated in FP64 in
//    interface cells (1e-6 < alpha_k < 1-1e-6) where stiffened-gas
//    cancellation bites, in FP32 elsewhere;
//  * the CFL candidate cfl*dx/(|u|+c) and its reduction are FP64.
// Divisions use the hardware reciprocal (v_rcp_f32, 1 ulp) and square roots
// v_sqrt_f32 / v_rsq_f32.  Floors are v_max_f32 / v_min_f32, for which a NaN
// operand yields the other operand (the floor); states are finite by
// construction (every stored or half-step state has been through repair), so
// the reference's finite_or guards are kept only where a state is repaired
// and in the Riemann solver's input check.
#pragma once

#include "policy_fp64.hpp"

// Products and sums within one expression are fused into FMAs (llvm.fmuladd)
// in this policy's code only: it has no bit-parity contract, and a fused
// a*b+c is one instruction and one rounding.  The decision is made per
// expression at parse time, so every inlining context (single-grid and
// batched sweeps) computes the same bits.  The FP64 policy, compiled with
// -ffp-contract=off, is outside this region.
#pragma clang fp contract(on)

namespace dimhip {

struct EosMixed {
    int kind[kMaxMat];
    float gamma[kMaxMat];
    float pinf[kMaxMat];
    float gpinf[kMaxMat];     // gamma * p_inf
    float inv_gm1[kMaxMat];   // 1/(gamma-1)
    float b[kMaxMat];         // gamma p_inf/(gamma-1)
    float bmin[kMaxMat];      // 1e-10/(gamma-1): shifted energy floor per unit volume fraction
    double inv_gm1_d[kMaxMat];
    double b_d[kMaxMat];
    float gm1[kMaxMat];       // gamma - 1
    float coef_lo[kMaxMat];   // min(1, 1/(gamma-1)): lower bound of ecoef
    float b_ghost[kMaxMat];   // b for gamma < 2 (where a near-massless fraction is down-weighted), else 0
    int lambda_model;
    float alpha_source_floor;
    int stiff;                // any stiffened-gas material: B(alpha) is not identically zero
    int stiff_ghost;          // any b_ghost != 0
};

// The same parameter block for contexts whose materials are all ideal gases
// under the Kapila lambda model (the helium shock-bubble, the batched
// shock-interface instances): the kernels instantiated with it drop the
// stiffening and lambda-model branches at compile time.
struct EosMixedIdeal : EosMixed {
    EosMixedIdeal() = default;
    explicit EosMixedIdeal(const EosMixed& e) : EosMixed(e) {}
};
static_assert(sizeof(EosMixedIdeal) == sizeof(EosMixed), "layout-identical parameter blocks");

__device__ __forceinline__ bool law_stiff(const EosMixed& P) { return P.stiff != 0; }
__device__ __forceinline__ constexpr bool law_stiff(const EosMixedIdeal&) { return false; }
__device__ __forceinline__ bool law_stiff_ghost(const EosMixed& P) { return P.stiff_ghost != 0; }
__device__ __forceinline__ constexpr bool law_stiff_ghost(const EosMixedIdeal&) { return false; }
__device__ __forceinline__ int law_lambda(const EosMixed& P) { return P.lambda_model; }
__device__ __forceinline__ constexpr int law_lambda(const EosMixedIdeal&) { return LAMBDA_KAPILA; }

namespace f32 {

__device__ __forceinline__ float fmaxs(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float fmins(float a, float b) { return __builtin_fminf(a, b); }
template<typename T>
__device__ __forceinline__ bool ffinite(T x) { return __builtin_isfinite(x); }
__device__ __forceinline__ float ffinite_or(float x, float fb = 0.f) { return ffinite(x) ? x : fb; }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
// |d| raised to tol, keeping the sign of d (numerical_safety.hpp:38-63)
__device__ __forceinline__ float safe_denom_f(float d, float tol = 1e-10f)
{
    return __builtin_copysignf(__builtin_fmaxf(__builtin_fabsf(d), tol), d);
}
__device__ __forceinline__ float safe_div_f(float n, float d, float tol = 1e-10f)
{
    return n * rcp(safe_denom_f(d, tol));
}

// A material's weight in the mixture energy per unit volume fraction and
// pressure, w_k/(gamma_k - 1).  The reference evaluates each material law at
// clamp_min(rho_k) (>= 1e-10) and divides by safe_denom((gamma_k - 1) rho_k)
// (>= 1e-10) (src/dim/eos.hpp:59-75, src/euler/eos.hpp:607-614), so a
// fraction whose partial density is below 1e-10/(gamma_k - 1) -- one that
// carries (almost) no mass, as inactive-mass removal leaves behind
// (primitives.hpp:392-399) -- enters with w_k = (gamma_k - 1)
// max(rho_k, 1e-10)/1e-10 < 1 instead of 1:
//   w_k/(gamma_k - 1) = min(1/(gamma_k - 1), max(1e10 rho_k, 1))
//                     = med3(1e10 rho_k, min(1, 1/(gamma_k - 1)), 1/(gamma_k - 1)).
// Materials with gamma_k >= 2 always have w_k = 1.
template<typename EP>
__device__ __forceinline__ float ecoef(const EP& P, int k, float rho_k)
{
    return __builtin_amdgcn_fmed3f(rho_k * 1e10f, P.coef_lo[k], P.inv_gm1[k]);
}
// EosMixedIdeal contexts have gamma_k <= 2 for every material (the host
// selects these kernels only then), so min(1, 1/(gamma_k - 1)) is 1: an
// inline constant instead of a second parameter register
__device__ __forceinline__ float ecoef(const EosMixedIdeal& P, int k, float rho_k)
{
    return __builtin_amdgcn_fmed3f(rho_k * 1e10f, 1.0f, P.inv_gm1[k]);
}

// rho_k c_k^2 of the reference's material_sound_speed (src/dim/eos.hpp:29-40):
// the pressure recomputed from the energy built at the clamped density,
// p' = clamp_min(w_k (p + gamma p_inf) - gamma p_inf), and
// rho' c^2 = gamma (p' + p_inf); coef = ecoef(rho_k)
template<typename EP>
__device__ __forceinline__ float impedance(const EP& P, int k, float p, float coef)
{
    const float w = coef * P.gm1[k];
    if (law_stiff(P)) return P.gamma[k] * (fmaxs(w * (p + P.gpinf[k]) - P.gpinf[k], 1e-10f) + P.pinf[k]);
    return P.gamma[k] * fmaxs(w * p, 1e-10f);
}

// sum_k alpha_k (1 - w_k) b_k: the stiffening energy a down-weighted fraction
// does not carry (only for stiffened materials with gamma < 2)
template<int M, typename EP>
__device__ __forceinline__ float ghost_shift(const EP& P, const float* alpha, const float* coef)
{
    float s = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) s += fmaxs(alpha[k], 0.f) * P.b_ghost[k] * (1.f - coef[k] * P.gm1[k]);
    return s;
}

template<int D, int M>
struct St {
    float m[M];
    float mom[D];
    float E;
    float a[M > 1 ? M - 1 : 1];
};

template<int D, int M>
struct Pr {
    float rho[M];
    float alpha[M];
    float vel[D];
    float p;
};

template<int D, int M>
struct Fx {
    float m[M];
    float mom[D];
    float E;
};

template<int D, int M>
struct Cell {
    double m[M];
    float mom[D];
    float E;
    float a[M > 1 ? M - 1 : 1];
};

// sanitise_alpha (primitives.hpp:45-70).  Fractions that already sum to
// within 1e-4 of one are renormalised with 1 - (sum - 1), whose error
// (sum - 1)^2 <= 1e-8 is below FP32 resolution; other sums take the
// reciprocal.
template<int M>
__device__ __forceinline__ void sanitise(float* alpha, float floor)
{
    float sum = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        alpha[k] = fmaxs(alpha[k], floor);
        sum += alpha[k];
    }
    const float dev = sum - 1.f;
    float inv = 1.f - dev;
    if (!(__builtin_fabsf(dev) <= 1e-4f)) {
        if (!ffinite(sum) || sum <= 0.f) {
#pragma unroll
            for (int k = 0; k < M; ++k) alpha[k] = (k == 0) ? 1.f : 0.f;
            return;
        }
        inv = rcp(sum);
    }
#pragma unroll
    for (int k = 0; k < M; ++k) alpha[k] *= inv;
}

// Volume fractions are stored as the reference's nmat-1 independent values,
// except that for two materials the stored value is alpha_0 when
// alpha_0 <= 1/2 and -alpha_1 otherwise: the smaller of alpha and 1-alpha
// then keeps full FP32 relative precision (a cell that is 1-1e-9 air still
// knows it holds 1e-9 helium).
template<int M>
__device__ __forceinline__ void decode_alpha(const float* a, float* alpha)
{
    if constexpr (M == 2) {
        // the smaller fraction is |a| and the larger 1 - |a| (1 + a for a
        // negative a, the same rounding); the sign bit says which one is
        // alpha_0, so -0.0 encodes alpha_1 = 0
        const float small = __builtin_fabsf(a[0]);
        const float large = 1.f - small;
        const bool neg = __builtin_signbit(a[0]);
        alpha[0] = neg ? large : small;
        alpha[1] = neg ? small : large;
    } else {
        float sum = -0.f;
#pragma unroll
        for (int k = 0; k < M - 1; ++k) {
            alpha[k] = a[k];
            sum += alpha[k];
        }
        alpha[M - 1] = 1.f - sum;
    }
}

// alpha: a sanitised full vector (non-negative, summing to one)
template<int M>
__device__ __forceinline__ void encode_alpha(const float* alpha, float* a)
{
    if constexpr (M == 2) {
        a[0] = (alpha[0] <= 0.5f) ? alpha[0] : -alpha[1];
    } else {
#pragma unroll
        for (int k = 0; k < M - 1; ++k) a[k] = alpha[k];
    }
}

// Energy is carried shifted by the stiffening constant,
//   E_hat = E - B(alpha),  B(alpha) = sum_k alpha_k gamma_k p_inf_k / (gamma_k - 1),
// with alpha the decoded stored volume fractions, so that
// rho e - B = sum_k alpha_k p/(gamma_k - 1) carries no cancellation between
// terms of size gamma p_inf and the pressure (water: 7.8e8 J/m^3 vs 1e5 Pa).
template<int M, typename EP>
__device__ __forceinline__ float bshift(const EP& P, const float* from, const float* to)
{
    if (!law_stiff(P)) return -0.f;
    float s = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) s += P.b[k] * (from[k] - to[k]);
    return s;
}

template<int M, typename EP>
__device__ __forceinline__ float bdot(const EP& P, const float* alpha)
{
    if (!law_stiff(P)) return -0.f;
    float s = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) s += P.b[k] * alpha[k];
    return s;
}

// The 1e-12 floor of cons_to_prim / prim_to_cons on fractions that are
// already non-negative and sum to one: raising a fraction to 1e-12 moves the
// sum by at most 1e-12 * M, far below half an FP32 ulp of 1, so the
// renormalisation is the identity in FP32 and only the floor remains.  For
// three or more materials the general sanitise runs.
template<int M>
__device__ __forceinline__ void floor_fractions(float* alpha)
{
    if constexpr (M <= 2) {
#pragma unroll
        for (int k = 0; k < M; ++k) alpha[k] = fmaxs(alpha[k], 1e-12f);
    } else {
        sanitise<M>(alpha, 1e-12f);
    }
}

// a stored two-material code is valid when |a| <= 1/2 (NaN is not)
template<int M>
__device__ __forceinline__ bool valid_code(const float* a)
{
    if constexpr (M == 2) return __builtin_fabsf(a[0]) <= 0.5f;
    else return false;
}

template<int M>
__device__ __forceinline__ bool interface_cell(const float* alpha)
{
    // two materials: the fractions sum to one, so both lie in
    // (1e-6, 1-1e-6) exactly when the smaller one exceeds 1e-6
    if constexpr (M == 2) return fmins(alpha[0], alpha[1]) > 1e-6f;
    bool itf = false;
#pragma unroll
    for (int k = 0; k < M; ++k) itf |= (alpha[k] > 1e-6f && alpha[k] < 1.f - 1e-6f);
    return itf;
}

// p from the shifted internal energy density
// (rho e - B(alpha)) = sum_k alpha_k w_k p/(gamma_k-1) - sum_k alpha_k (1 - w_k) b_k,
// coef[k] = ecoef of the material's partial density.  The fractions are
// non-negative (floored by cons_to_prim, sanitised or decoded from a valid
// code by repair).
template<int M, typename EP>
__device__ __forceinline__ float pressure_shifted(const EP& P, float rho_e_hat, const float* alpha,
                                                  const float* coef)
{
    float S = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) S += alpha[k] * coef[k];
    if (law_stiff_ghost(P)) rho_e_hat += ghost_shift<M>(P, alpha, coef);
    return rho_e_hat * rcp(S);
}

// rho e - B at pressure p: sum_k alpha_k w_k max(p,1e-10)/(gamma_k-1) - sum_k alpha_k (1 - w_k) b_k,
// for p >= 1e-10 and non-negative fractions (its callers' faces lie between
// cons_to_prim outputs, whose pressure is floored at 1e-10; repair's
// rebuild floors its pressure itself)
template<int M, typename EP>
__device__ __forceinline__ float energy_shifted(const EP& P, float p, const float* alpha, const float* coef)
{
    float s = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) s += alpha[k] * (p * coef[k]);
    if (law_stiff_ghost(P)) s -= ghost_shift<M>(P, alpha, coef);
    return s;
}

// cons_to_prim (primitives.hpp:111-162) with the closed-form pressure, on a
// state with a valid fraction code (stored, half-step and face states)
template<int D, int M, typename EP>
__device__ __forceinline__ void cons_to_prim(const EP& P, const St<D, M>& U, Pr<D, M>& R)
{
    float ad[M];
    decode_alpha<M>(U.a, ad);
#pragma unroll
    for (int k = 0; k < M; ++k) R.alpha[k] = ad[k];
    floor_fractions<M>(R.alpha);
    float rho = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) rho += U.m[k];
    rho = fmaxs(rho, 1e-10f);
    float inv_rho, ia[M];
    if constexpr (M == 2) {
        // one reciprocal for 1/rho, 1/alpha_0' and 1/alpha_1' (alpha' =
        // max(alpha, 1e-10), so the product stays above 1e-30)
        const float a0 = fmaxs(R.alpha[0], 1e-10f), a1 = fmaxs(R.alpha[1], 1e-10f);
        const float a01 = a0 * a1;
        const float X = rcp(a01 * rho);
        const float rX = rho * X;
        inv_rho = a01 * X;
        ia[0] = a1 * rX;
        ia[1] = a0 * rX;
    } else {
        inv_rho = rcp(rho);
#pragma unroll
        for (int k = 0; k < M; ++k) ia[k] = rcp(fmaxs(R.alpha[k], 1e-10f));
    }
    float v2 = -0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        R.vel[d] = U.mom[d] * inv_rho;
        v2 += R.vel[d] * R.vel[d];
    }
    float coef[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        R.rho[k] = fmaxs(U.m[k] * ia[k], 1e-12f);
        coef[k] = ecoef(P, k, R.rho[k]);
    }
    const float kinetic = 0.5f * rho * v2;
    const float shift = law_stiff(P) ? bshift<M>(P, ad, R.alpha) : 0.f;
    // Stiffened materials: the shifted energy of a fast, low-pressure liquid
    // is mostly kinetic (water at 480 m/s and 3e4 Pa: 1 part in 1e4), so
    // there the kinetic energy is subtracted in FP64 too, whatever the
    // fractions; otherwise cells next to the interface-cell threshold would
    // alternate between the two recoveries with rounding noise.
    const bool fast_liquid = law_stiff(P) && kinetic > 0.5f * (U.E + shift);
    float p;
    if (interface_cell<M>(R.alpha) || fast_liquid) {
        // FP64 recovery: the energy minus kinetic energy difference, where the
        // cancellation sits, in FP64 (float products are exact in FP64); 1/rho
        // from the FP32 reciprocal and one Newton step (relative error ~1e-14)
        double m2 = -0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) m2 += (double)U.mom[d] * (double)U.mom[d];
        const double rd = (double)rho;
        double inv = (double)inv_rho;
        inv = inv * (2.0 - rd * inv);
        const double rho_e_hat = (double)U.E + (double)shift - 0.5 * m2 * inv;
        p = pressure_shifted<M>(P, (float)rho_e_hat, R.alpha, coef);
    } else {
        p = pressure_shifted<M>(P, U.E + shift - kinetic, R.alpha, coef);
    }
    R.p = fmaxs(p, 1e-10f);
}

// prim_to_cons (primitives.hpp:166-212) of a reconstructed face primitive:
// its values lie between those of cons_to_prim outputs (MinBee faces), so
// densities are >= 1e-12 and the pressure >= 1e-10, and for two materials
// the fractions are >= 1e-12 (three or more materials: non-negative and
// summing to one after enforce); alpha receives the floored fractions the
// state was built from.  The reference's floors at these values are
// therefore left out where they cannot act.
template<int D, int M, typename EP>
__device__ __forceinline__ void prim_to_cons(const EP& P, const Pr<D, M>& Q, St<D, M>& U, float* alpha)
{
#pragma unroll
    for (int k = 0; k < M; ++k) alpha[k] = Q.alpha[k];
    if constexpr (M > 2) floor_fractions<M>(alpha);
    float rho = -0.f, coef[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const float rk = Q.rho[k];
        U.m[k] = alpha[k] * rk;
        coef[k] = ecoef(P, k, rk);
        rho += U.m[k];
    }
    rho = fmaxs(rho, 1e-12f);
    float v2 = -0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        U.mom[d] = rho * Q.vel[d];
        v2 += Q.vel[d] * Q.vel[d];
    }
    encode_alpha<M>(alpha, U.a);
    float shift = -0.f;
    if (law_stiff(P)) {
        float ad[M];
        decode_alpha<M>(U.a, ad);
        shift = bshift<M>(P, alpha, ad);
    }
    U.E = energy_shifted<M>(P, Q.p, alpha, coef) + shift + 0.5f * rho * v2;
}

// a fraction at or below the inactive threshold (1e-10) that carries more
// than 1e-12 of mass (primitives.hpp:392-399)
template<int M, typename MT>
__device__ __forceinline__ bool inactive_mass(const float* alpha, const MT* m)
{
    bool r = false;
#pragma unroll
    for (int k = 0; k < M; ++k) r |= (alpha[k] <= 1e-10f && m[k] > MT(1e-12));
    return r;
}

// sum_k alpha_k bmin_k: the energy floor's internal-energy part at the
// floor pressure
template<int M, typename EP>
__device__ __forceinline__ float energy_floor(const EP& P, const float* alpha)
{
    float emin = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) emin += alpha[k] * P.bmin[k];
    return emin;
}

// repair_state (primitives.hpp:326-457) on a state whose partial masses are
// of type MT (float for face/half states, double for stored cells).  The
// common case (valid fraction code, no inactive mass) is the energy floor
// alone; fraction renormalisation and the removal of inactive mass with its
// pressure-preserving rebuild run only when needed.
template<int D, int M, typename S, typename EP>
__device__ __forceinline__ void repair(const EP& P, S& U)
{
    using MT = decltype(U.m[0] + 0.f);
    MT m_in[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        // negative or NaN mass -> 0
        U.m[k] = (U.m[k] > MT(0)) ? U.m[k] : MT(0);
        m_in[k] = U.m[k];
    }
    float ad0[M], alpha[M];
    decode_alpha<M>(U.a, ad0);
#pragma unroll
    for (int k = 0; k < M; ++k) alpha[k] = ad0[k];
    float fix = -0.f;           // B(alpha) - B(stored code): floors refer to the true energy
    float e_hat_before = U.E;  // rho e - B(alpha) + kinetic, before any change
    if (!valid_code<M>(U.a)) {
#pragma unroll
        for (int d = 0; d < D; ++d) U.mom[d] = ffinite_or(U.mom[d]);
        U.E = ffinite_or(U.E);
        sanitise<M>(alpha, 0.f);
        encode_alpha<M>(alpha, U.a);
        if (law_stiff(P)) {
            float ad1[M];
            decode_alpha<M>(U.a, ad1);
            e_hat_before = U.E + bshift<M>(P, ad0, alpha);
            U.E += bshift<M>(P, ad0, ad1);
            fix = bshift<M>(P, alpha, ad1);
        } else {
            e_hat_before = U.E;
        }
    }
    bool removed = false;
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (alpha[k] <= 1e-10f && U.m[k] > MT(1e-12)) {
            U.m[k] = MT(0);
            removed = true;
        }
    MT rt = MT(0);
#pragma unroll
    for (int k = 0; k < M; ++k) rt += U.m[k];
    if (!(rt > MT(0))) {
        int dom = 0;
#pragma unroll
        for (int k = 1; k < M; ++k)
            if (alpha[k] > alpha[dom]) dom = k;
#pragma unroll
        for (int k = 0; k < M; ++k)
            if (k == dom) U.m[k] = MT(1e-12);
        rt = MT(1e-12);
    }
    const float rho = (float)rt;
    if (removed) {
        // rebuild momentum and energy at the velocity and pressure before removal
        MT rb = MT(0);
#pragma unroll
        for (int k = 0; k < M; ++k) rb += m_in[k];
        const float rho_before = fmaxs((float)rb, 1e-12f);
        const float inv_rb = rcp(fmaxs(rho_before, 1e-10f));
        float v2b = -0.f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float v = U.mom[d] * inv_rb;
            U.mom[d] = rho * v;
            v2b += v * v;
        }
        // material densities before and after the removal, as the reference's
        // rho_material_before / rho_material (primitives.hpp:352-360, 404-410)
        float cb[M], ca[M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const float ia = rcp(fmaxs(alpha[k], 1e-10f));
            cb[k] = ecoef(P, k, fmaxs((float)m_in[k] * ia, 1e-12f));
            ca[k] = ecoef(P, k, fmaxs((float)U.m[k] * ia, 1e-12f));
        }
        const float p_before =
            fmaxs(pressure_shifted<M>(P, e_hat_before - 0.5f * rho_before * v2b, alpha, cb), 1e-10f);
        U.E = energy_shifted<M>(P, p_before, alpha, ca) + fix + 0.5f * rho * v2b;
    }
    const float inv = rcp(fmaxs(rho, 1e-10f));
    float v2 = -0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float v = U.mom[d] * inv;
        v2 += v * v;
    }
    U.E = fmaxs(U.E, 0.5f * rho * v2 + energy_floor<M>(P, alpha) + fix);
}

// muscl.hpp:26-40
template<int D, int M>
__device__ __forceinline__ void enforce(Pr<D, M>& Q)
{
#pragma unroll
    for (int k = 0; k < M; ++k) Q.rho[k] = fmaxs(Q.rho[k], 1e-12f);
    sanitise<M>(Q.alpha, 0.f);
    Q.p = fmaxs(Q.p, 1e-12f);
}

// MinBee (limiters.hpp:9-15): 0 for opposite signs, else the smaller
// magnitude, which is the median of (a, b, 0): one v_med3_f32
__device__ __forceinline__ float minmod(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.f); }

// muscl.hpp:43-112
template<int D, int M>
__device__ __forceinline__ void reconstruct(const Pr<D, M>& L, const Pr<D, M>& C, const Pr<D, M>& R, Pr<D, M>& lf,
                                            Pr<D, M>& rf)
{
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const float s = 0.5f * minmod(C.rho[k] - L.rho[k], R.rho[k] - C.rho[k]);
        lf.rho[k] = C.rho[k] - s;
        rf.rho[k] = C.rho[k] + s;
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const float s = 0.5f * minmod(C.alpha[k] - L.alpha[k], R.alpha[k] - C.alpha[k]);
        lf.alpha[k] = C.alpha[k] - s;
        rf.alpha[k] = C.alpha[k] + s;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float s = 0.5f * minmod(C.vel[d] - L.vel[d], R.vel[d] - C.vel[d]);
        lf.vel[d] = C.vel[d] - s;
        rf.vel[d] = C.vel[d] + s;
    }
    const float s = 0.5f * minmod(C.p - L.p, R.p - C.p);
    lf.p = C.p - s;
    rf.p = C.p + s;
    // enforce_positive_primitive (muscl.hpp:26-40).  A MinBee face
    // C -/+ s/2 with s = fl(C-L) or fl(R-C) (the smaller, same sign) or 0
    // lies between its neighbours' values even after rounding, and the
    // inputs are floored primitives (densities >= 1e-12, fractions >= 0,
    // p >= 1e-12), so the floors cannot act.  For two materials the faces'
    // fractions sum to one to FP32 rounding (the two slopes are each other's
    // negatives), and the renormalisation only moves them by that rounding;
    // both are left out.  Three or more materials keep the full enforce.
    if constexpr (M > 2) {
        enforce<D, M>(lf);
        enforce<D, M>(rf);
    }
}

// sum_k alpha_k / (rho_k c_k^2), the Wood compressibility (eos.hpp:195-221),
// with rho_k c_k^2 = gamma_k (p' + p_inf_k) at the clamped partial density
// (impedance; gamma_k (p + p_inf_k) unless the fraction is near-massless)
template<int M, typename EP>
__device__ __forceinline__ float compressibility(const EP& P, float p, const float* alpha, const float* rho_k)
{
    float beta = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) beta += alpha[k] * rcp(fmaxs(impedance(P, k, p, ecoef(P, k, rho_k[k])), 1e-12f));
    return beta;
}

// Wood sound speed (eos.hpp:195-221) and the allaire / equal_velocity
// variants (eos.hpp:223-275), closed form for ideal and stiffened gases.
// Its inputs are cons_to_prim outputs or reconstructed faces: p >= 1e-10
// and fractions >= 1e-12 (non-negative for three or more materials), so the
// reference's floors on them cannot act; the density floor stays.
template<int M, typename EP>
__device__ __forceinline__ float sound_speed(const EP& P, float rho, float p, const float* alpha, const float* rho_k)
{
    rho = fmaxs(rho, 1e-10f);
    if (law_lambda(P) == LAMBDA_KAPILA) {
        if constexpr (M == 2) {
            // c^2 = 1/(rho (a0/I0 + a1/I1)) = Y/(rho X), X = a0 I1 + a1 I0, Y = I0 I1,
            // so c = Y rsq(rho X Y): one transcendental (X >= 2e-24)
            const float I0 = fmaxs(impedance(P, 0, p, ecoef(P, 0, rho_k[0])), 1e-12f);
            const float I1 = fmaxs(impedance(P, 1, p, ecoef(P, 1, rho_k[1])), 1e-12f);
            const float a0 = alpha[0], a1 = alpha[1];
            const float X = a0 * I1 + a1 * I0;
            const float Y = I0 * I1;
            return Y * frsq(rho * X * Y);
        } else {
            const float beta = compressibility<M>(P, p, alpha, rho_k);
            return frsq(rho * fmaxs(beta, 1e-14f));
        }
    }
    // xi_k = rho_k de_k/dp = w_k/(gamma_k - 1) (eos.hpp:42-57 on the ideal and
    // stiffened laws), c_k^2 = rho' c^2 / rho'
    float cs[M], xi[M], xi_mix = -0.f, inv_sum = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const float a = alpha[k];
        xi[k] = ecoef(P, k, rho_k[k]);
        cs[k] = fsqrt(impedance(P, k, p, xi[k]) * rcp(fmaxs(rho_k[k], 1e-10f)));
        xi_mix += a * xi[k];
        inv_sum += a * rcp(fmaxs(cs[k], 1e-14f));
    }
    float rc2 = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const float a = alpha[k];
        const float lam = (law_lambda(P) == LAMBDA_EQUAL_VELOCITY) ? rcp(fmaxs(cs[k], 1e-14f) * inv_sum) : 1.f;
        rc2 += lam * a * xi[k] * cs[k] * cs[k];
    }
    return fsqrt(fmaxs(rc2 * rcp(xi_mix) * rcp(rho), 1e-14f));
}

// acoustic_impedance of alpha_rhs_coefficients (primitives.hpp:288-297):
// max(rho_k c_k^2, 1e-12) with rho_k the primitive's partial density (floored
// at 1e-12) and c_k from the clamped one (1e-10), rho_k/rho' = min(1, 1e10 rho_k)
template<typename EP>
__device__ __forceinline__ float kapila_impedance(const EP& P, int k, float p, float rho_k)
{
    return fmaxs(fmins(rho_k * 1e10f, 1.f) * impedance(P, k, p, ecoef(P, k, rho_k)), 1e-12f);
}

// alpha_rhs_coefficients (primitives.hpp:216-322)
// Q comes from cons_to_prim: its fractions are sanitised already, so the
// reference's second sanitise (floor 0) is the identity here.
template<int D, int M, typename EP>
__device__ __forceinline__ void alpha_rhs(const EP& P, const Pr<D, M>& Q, float* rhs)
{
#pragma unroll
    for (int k = 0; k < M; ++k) rhs[k] = Q.alpha[k];
    if (law_lambda(P) == LAMBDA_ALLAIRE) return;
    const float p = fmaxs(Q.p, 1e-12f);
    const float srcf = fmaxs(P.alpha_source_floor, 0.f);
    if (law_lambda(P) == LAMBDA_EQUAL_VELOCITY) {
        float cs[M], inv_sum = -0.f;
        bool act[M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const float a = rhs[k];
            act[k] = !(a <= 1e-12f || a < srcf);
            cs[k] = act[k] ? fsqrt(impedance(P, k, p, ecoef(P, k, Q.rho[k])) * rcp(fmaxs(Q.rho[k], 1e-10f))) : 1e-12f;
            if (act[k]) inv_sum += a * rcp(cs[k]);
        }
        if (inv_sum <= 1e-12f || !ffinite(inv_sum)) return;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const float a = rhs[k];
            rhs[k] = (a <= 1e-12f) ? 0.f : a * rcp(cs[k] * inv_sum);
        }
        return;
    }
    if constexpr (M == 2) {
        // both phases active: q/I0 = I1/X and q/I1 = I0/X with X = a0 I1 + a1 I0,
        // one reciprocal for the pair
        const float a0 = rhs[0], a1 = rhs[1];
        if (!(a0 <= 1e-12f || a0 < srcf) && !(a1 <= 1e-12f || a1 < srcf)) {
            const float I0 = kapila_impedance(P, 0, p, Q.rho[0]);
            const float I1 = kapila_impedance(P, 1, p, Q.rho[1]);
            const float invX = rcp(a0 * I1 + a1 * I0);
            rhs[0] = a0 + a0 * (I1 * invX - 1.f);
            rhs[1] = a1 + a1 * (I0 * invX - 1.f);
            return;
        }
    }
    float imp[M];
    bool act[M];
    float beta = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const float a = rhs[k];
        act[k] = !(a <= 1e-12f || a < srcf);
        imp[k] = kapila_impedance(P, k, p, Q.rho[k]);
        if (act[k]) beta += a * rcp(imp[k]);
    }
    if (beta <= 1e-12f || !ffinite(beta)) return;
    const float q = rcp(beta);
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const float a = rhs[k];
        rhs[k] = act[k] ? a + a * (q * rcp(imp[k]) - 1.f) : 0.f;
    }
}

// physical flux (flux.hpp:11-43) with the true energy E = E_hat + B(alpha).
// The face states here are built by prim_to_cons, whose momentum is rho u
// of the primitive, so rho u_d u_n is taken as mom_d u_n.
template<int D, int M, int DIR, typename EP>
__device__ __forceinline__ void flux_normal(const EP& P, const St<D, M>& U, const Pr<D, M>& Q, Fx<D, M>& F)
{
    const float un = Q.vel[DIR];
#pragma unroll
    for (int k = 0; k < M; ++k) F.m[k] = U.m[k] * un;
#pragma unroll
    for (int d = 0; d < D; ++d) F.mom[d] = U.mom[d] * un + (d == DIR ? Q.p : 0.f);
    float B = -0.f;
    if (law_stiff(P)) {
        float ad[M];
        decode_alpha<M>(U.a, ad);
        B = bdot<M>(P, ad);
    }
    F.E = (U.E + B + Q.p) * un;
}

// muscl.hpp:143-191
template<int D, int M, int DIR, typename EP>
__device__ __forceinline__ void hancock(const EP& P, const St<D, M>& Uc, const Pr<D, M>& Pc, const St<D, M>& ul,
                                        const St<D, M>& ur, const Pr<D, M>& lf, const Pr<D, M>& rf, const float* rhs,
                                        float dt, float inv_dx, St<D, M>& out)
{
    Fx<D, M> FL, FR;
    flux_normal<D, M, DIR>(P, ul, lf, FL);
    flux_normal<D, M, DIR>(P, ur, rf, FR);
    const float factor = 0.5f * dt * inv_dx;
    out = Uc;
#pragma unroll
    for (int k = 0; k < M; ++k) out.m[k] -= factor * (FR.m[k] - FL.m[k]);
#pragma unroll
    for (int d = 0; d < D; ++d) out.mom[d] -= factor * (FR.mom[d] - FL.mom[d]);
    out.E -= factor * (FR.E - FL.E);
    const float ul_ = lf.vel[DIR], ur_ = rf.vel[DIR];
    const float div_u = (ur_ - ul_) * inv_dx;
    float an[M];
#pragma unroll
    for (int k = 0; k < M; ++k)
        an[k] = Pc.alpha[k] - factor * (rf.alpha[k] * ur_ - lf.alpha[k] * ul_) + 0.5f * dt * rhs[k] * div_u;
    sanitise<M>(an, 0.f);
    if (law_stiff(P)) {
        float ad_old[M], ad_new[M];
        decode_alpha<M>(out.a, ad_old);
        encode_alpha<M>(an, out.a);
        decode_alpha<M>(out.a, ad_new);
        out.E += bshift<M>(P, ad_old, ad_new);
    } else {
        encode_alpha<M>(an, out.a);
    }
    repair<D, M>(P, out);
}

// hllc.hpp:62-195
template<int D, int M, int DIR, typename EP>
__device__ __forceinline__ int hllc(const EP& P, const St<D, M>& UL, const St<D, M>& UR, const Pr<D, M>& PL,
                                     const Pr<D, M>& PR, Fx<D, M>& F, float& fv)
{
    float rhoL = -0.f, rhoR = -0.f;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        rhoL += UL.m[k];
        rhoR += UR.m[k];
    }
    // require_positive (hllc.hpp:88-91) in the reference's order.  The four
    // values are non-negative or not finite, so their sum is finite exactly
    // when all four are (barring an overflow of the sum, which the
    // individual checks then clear).  Face pressures are >= 1e-10 (cons_to_prim
    // outputs and faces between them), so the pressure clamp to 1e-10 cannot
    // act; the density clamp stays (an end cell's FP32 masses can round to 0)
    if (!ffinite((rhoL + rhoR) + (PL.p + PR.p))) {
        if (!ffinite(rhoL)) return ERR_HLLC_LEFT_DENSITY;
        if (!ffinite(rhoR)) return ERR_HLLC_RIGHT_DENSITY;
        if (!ffinite(PL.p)) return ERR_HLLC_LEFT_PRESSURE;
        if (!ffinite(PR.p)) return ERR_HLLC_RIGHT_PRESSURE;
    }
    rhoL = rhoL <= 0.f ? 1e-10f : rhoL;
    rhoR = rhoR <= 0.f ? 1e-10f : rhoR;
    const float pL = PL.p;
    const float pR = PR.p;
    const float unL = PL.vel[DIR], unR = PR.vel[DIR];
    const float cL = sound_speed<M>(P, rhoL, pL, PL.alpha, PL.rho);
    const float cR = sound_speed<M>(P, rhoR, pR, PR.alpha, PR.rho);
    const float sL = fmins(unL - cL, unR - cR);
    const float sR = fmaxs(unL + cL, unR + cR);
    const float num = (pR - pL) + rhoL * unL * (sL - unL) - rhoR * unR * (sR - unR);
    const float den = rhoL * (sL - unL) - rhoR * (sR - unR);
    float s_star = safe_div_f(num, den);
    if (__builtin_fabsf(pL - pR) < 1e-10f && __builtin_fabsf(unL - unR) < 1e-10f) s_star = unL;
    Fx<D, M> FL, FR;
    flux_normal<D, M, DIR>(P, UL, PL, FL);
    flux_normal<D, M, DIR>(P, UR, PR, FR);
    const bool average = !ffinite(sL) || !ffinite(sR) || !ffinite(s_star) ||
                         (!(0.f <= sL) && !(sR <= 0.f) && !(sL < sR));
    if (average) {
#pragma unroll
        for (int k = 0; k < M; ++k) F.m[k] = 0.5f * (FL.m[k] + FR.m[k]);
#pragma unroll
        for (int d = 0; d < D; ++d) F.mom[d] = 0.5f * (FL.mom[d] + FR.mom[d]);
        F.E = 0.5f * (FL.E + FR.E);
        fv = 0.5f * (unL + unR);
        return 0;
    }
    if (0.f <= sL) {
        F = FL;
        fv = unL;
        return 0;
    }
    if (sR <= 0.f) {
        F = FR;
        fv = unR;
        return 0;
    }
    const bool left = (0.f <= s_star);
    const St<D, M>& UK = left ? UL : UR;
    const Pr<D, M>& PK = left ? PL : PR;
    const Fx<D, M>& FK = left ? FL : FR;
    const float rhoK = left ? rhoL : rhoR;
    const float pK = left ? pL : pR;
    const float sK = left ? sL : sR;
    const float unK = left ? unL : unR;
    // star state (hllc.hpp:15-41, 144-194) written through
    // delta = rho*/rho_K - 1 = (s* - u_K)/(s_K - s*) so that E* - E_K carries
    // no cancellation against the stiffened-gas energy
    float BK = -0.f;
    if (law_stiff(P)) {
        float adK[M];
        decode_alpha<M>(UK.a, adK);
        BK = bdot<M>(P, adK);
    }
    const float EK = UK.E + BK;
    const float inv_gap = rcp(safe_denom_f(sK - s_star));
    const float ratio = (sK - unK) * inv_gap;
    const float rho_star = rhoK * ratio;
    const float delta = (s_star - unK) * inv_gap;
    const float e_star_term = rho_star * (s_star - unK) * (s_star + safe_div_f(pK, rhoK * (sK - unK)));
    const float E_star = rho_star * safe_div_f(EK, rhoK) + e_star_term;
    fv = s_star;
    if (rho_star <= 0.f || !ffinite(rho_star) || !ffinite(E_star)) {
        float dB = -0.f;
        if (law_stiff(P)) {
            float adL[M], adR[M];
            decode_alpha<M>(UL.a, adL);
            decode_alpha<M>(UR.a, adR);
            dB = bshift<M>(P, adR, adL);
        }
        const float inv = rcp(sR - sL);
        const float sLsR = sL * sR;
#pragma unroll
        for (int k = 0; k < M; ++k) F.m[k] = inv * ((sR * FL.m[k] - sL * FR.m[k]) + sLsR * (UR.m[k] - UL.m[k]));
#pragma unroll
        for (int d = 0; d < D; ++d)
            F.mom[d] = inv * ((sR * FL.mom[d] - sL * FR.mom[d]) + sLsR * (UR.mom[d] - UL.mom[d]));
        F.E = inv * ((sR * FL.E - sL * FR.E) + sLsR * ((UR.E - UL.E) + dB));
        return 0;
    }
#pragma unroll
    for (int k = 0; k < M; ++k) F.m[k] = FK.m[k] + sK * (UK.m[k] * delta);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float u_star = (d == DIR) ? s_star : PK.vel[d];
        F.mom[d] = FK.mom[d] + sK * (rho_star * u_star - UK.mom[d]);
    }
    F.E = FK.E + sK * (EK * delta + e_star_term);
    return 0;
}

}  // namespace f32

template<typename EP>
struct Fp32MixedT {
    // register cap of the sweep kernels: 8 waves per SIMD (64 VGPRs) fits the
    // 1D/2D two-material sweeps without spills; the 3D and three-material
    // sweeps need more registers and run uncapped (a cap spills them)
    template<int D, int M>
    static constexpr int min_waves_per_eu() { return (D <= 2 && M <= 2) ? 8 : 1; }
    static constexpr bool kReloadInput = false;
    // the update's inputs from the first phase (P^n's volume fractions and
    // the alpha source coefficients) wait in LDS while the reconstruction and
    // flux phases run, where the LDS budget has room for them (3D and
    // three-material tiles)
    static constexpr bool kParkAlpha = true;
    // 1D/2D two-material tiles exchange through two alternating LDS
    // regions.  x tiles carry face records of primitive, momentum and energy
    // and hold the update's fraction terms in registers (18 floats per slot:
    // eight 256-thread tiles per CU); y tiles carry the face primitive alone
    // and park the fraction terms (19 floats per slot: four 512-thread tiles
    // per CU, where the other records spilled or cost a tile per CU).
    template<int D, int M, int DIR>
    static constexpr bool ping_pong() { return D <= 2 && M <= 2; }
    template<int D, int M, int DIR>
    static constexpr bool prim_faces() { return DIR != 0; }
    template<int D, int M, int DIR>
    static constexpr bool park() { return kParkAlpha && M <= 2 && !(ping_pong<D, M, DIR>() && DIR == 0); }
    // sweep tiles: x 128 positions x 2 lines; y/z 8 lines x 64 positions
    static constexpr int kSweepL0 = 128, kSweepW0 = 2, kSweepL1 = 64, kSweepW1 = 8;
    // 2D y sweeps: one column strip of tiles per XCD (xcd_block_banded), so
    // the halo rows a tile shares with the tiles above and below it are read
    // by the same L2 (4096x2048: 197 -> 181 us)
    static constexpr bool kXcdStrips = true;
    // CFL reduction: each cell contributes its wave rate max_d (|u_d|+c)/(cfl dx_d),
    // combined by max in FP64; the block's dt is 1/rate (reciprocal is monotone,
    // so this equals the minimum of the per-cell 1/rate)
    static constexpr bool kCflRate = true;
    // the recovered pressure is not carried between sweeps (closed form)
    static constexpr bool kStoredPressure = false;
    using R = float;
    using Eos = EP;
    template<int D, int M> using St = f32::St<D, M>;
    template<int D, int M> using Pr = f32::Pr<D, M>;
    template<int D, int M> using Fx = f32::Fx<D, M>;
    template<int D, int M> using Cell = f32::Cell<D, M>;

    // storage: M double planes (partial masses) then D + M float planes
    // (momentum, energy, independent volume fractions)
    template<int D, int M>
    static __device__ __forceinline__ void load(const void* base, long long pitch, long long c, Cell<D, M>& U)
    {
        const double* qm = (const double*)base;
        const float* qf = (const float*)(qm + M * pitch);
#pragma unroll
        for (int k = 0; k < M; ++k) U.m[k] = qm[k * pitch + c];
        int f = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) U.mom[d] = qf[(f++) * pitch + c];
        U.E = qf[(f++) * pitch + c];
#pragma unroll
        for (int k = 0; k < M - 1; ++k) U.a[k] = qf[(f++) * pitch + c];
    }

    // load / store of cell cb + cr with cb uniform across the wave and cr a
    // 32-bit offset (cr * 8 < 2^32)
    template<int D, int M>
    static __device__ __forceinline__ void load_rel(const void* base, long long pitch, long long cb, unsigned cr,
                                                    Cell<D, M>& U)
    {
        const double* qm = (const double*)base;
        const float* qf = (const float*)(qm + M * pitch) + cb;
        qm += cb;
#pragma unroll
        for (int k = 0; k < M; ++k) U.m[k] = plane_at(qm + k * pitch, cr);
        int f = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) U.mom[d] = plane_at(qf + (f++) * pitch, cr);
        U.E = plane_at(qf + (f++) * pitch, cr);
#pragma unroll
        for (int k = 0; k < M - 1; ++k) U.a[k] = plane_at(qf + (f++) * pitch, cr);
    }

    template<int D, int M>
    static __device__ __forceinline__ void store_rel(void* base, long long pitch, long long cb, unsigned cr,
                                                     const Cell<D, M>& U)
    {
        double* qm = (double*)base;
        float* qf = (float*)(qm + M * pitch) + cb;
        qm += cb;
#pragma unroll
        for (int k = 0; k < M; ++k) plane_at(qm + k * pitch, cr) = U.m[k];
        int f = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) plane_at(qf + (f++) * pitch, cr) = U.mom[d];
        plane_at(qf + (f++) * pitch, cr) = U.E;
#pragma unroll
        for (int k = 0; k < M - 1; ++k) plane_at(qf + (f++) * pitch, cr) = U.a[k];
    }

    template<int D, int M>
    static __device__ __forceinline__ void store(void* base, long long pitch, long long c, const Cell<D, M>& U)
    {
        double* qm = (double*)base;
        float* qf = (float*)(qm + M * pitch);
#pragma unroll
        for (int k = 0; k < M; ++k) qm[k * pitch + c] = U.m[k];
        int f = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) qf[(f++) * pitch + c] = U.mom[d];
        qf[(f++) * pitch + c] = U.E;
#pragma unroll
        for (int k = 0; k < M - 1; ++k) qf[(f++) * pitch + c] = U.a[k];
    }

    // stored field f (the per-cell order m[M], mom[D], E, a[M-1]) of cell src
    // copied to cell dst, negated when neg (a bit-exact sign flip: x * -1)
    template<int D, int M>
    static __device__ __forceinline__ void copy_field(void* base, long long pitch, int f, long long src,
                                                      long long dst, bool neg)
    {
        double* qm = (double*)base;
        if (f < M) {
            double* p = qm + (long long)f * pitch;
            p[dst] = p[src];
            return;
        }
        float* p = (float*)(qm + M * pitch) + (long long)(f - M) * pitch;
        const float v = p[src];
        p[dst] = neg ? v * -1 : v;
    }

    // packed FP64 cell -> storage: volume fractions clamped to [0,1] and
    // encoded, energy shifted by B(alpha) in FP64
    template<int D, int M>
    static __device__ __forceinline__ void from_packed(const Eos& P, const double* p, Cell<D, M>& U)
    {
        int o = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) U.m[k] = p[o++];
#pragma unroll
        for (int d = 0; d < D; ++d) U.mom[d] = (float)p[o++];
        const double E = p[o++];
        double al[M];
        double sum = -0.0;
#pragma unroll
        for (int k = 0; k < M - 1; ++k) {
            al[k] = p[o++];
            sum += al[k];
        }
        al[M - 1] = 1.0 - sum;
        if constexpr (M == 2) {
            const double a0 = al[0] < 0.0 ? 0.0 : (al[0] > 1.0 ? 1.0 : al[0]);
            U.a[0] = (a0 <= 0.5) ? (float)a0 : -(float)(1.0 - a0);
        } else {
#pragma unroll
            for (int k = 0; k < M - 1; ++k) U.a[k] = (float)al[k];
        }
        U.E = (float)(E - Bd<M>(P, U.a));
    }

    // B(decoded alpha) in FP64
    template<int M>
    static __device__ __forceinline__ double Bd(const Eos& P, const float* a)
    {
        double al[M];
        if constexpr (M == 2) {
            const bool pos = !__builtin_signbit(a[0]);
            al[0] = pos ? (double)a[0] : 1.0 + (double)a[0];
            al[1] = pos ? 1.0 - (double)a[0] : -(double)a[0];
        } else {
            double sum = -0.0;
#pragma unroll
            for (int k = 0; k < M - 1; ++k) {
                al[k] = a[k];
                sum += al[k];
            }
            al[M - 1] = 1.0 - sum;
        }
        double b = -0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) b += P.b_d[k] * al[k];
        return b;
    }

    template<int D, int M>
    static __device__ __forceinline__ void to_packed(const Eos& P, const Cell<D, M>& U, double* p)
    {
        int o = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) p[o++] = U.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) p[o++] = U.mom[d];
        p[o++] = (double)U.E + Bd<M>(P, U.a);
        if constexpr (M == 2) {
            p[o++] = !__builtin_signbit(U.a[0]) ? (double)U.a[0] : 1.0 + (double)U.a[0];
        } else {
#pragma unroll
            for (int k = 0; k < M - 1; ++k) p[o++] = U.a[k];
        }
    }

    template<int D, int M>
    static __device__ __forceinline__ void to_state(const Cell<D, M>& c, St<D, M>& U)
    {
#pragma unroll
        for (int k = 0; k < M; ++k) U.m[k] = (float)c.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) U.mom[d] = c.mom[d];
        U.E = c.E;
#pragma unroll
        for (int k = 0; k < M - 1; ++k) U.a[k] = c.a[k];
    }

    template<int D, int M>
    static __device__ __forceinline__ void cons_to_prim(const Eos& P, const St<D, M>& U, Pr<D, M>& Q)
    {
        f32::cons_to_prim<D, M>(P, U, Q);
    }
    template<int D, int M>
    static __device__ __forceinline__ void enforce(Pr<D, M>& Q) { f32::enforce<D, M>(Q); }
    // after cons_to_prim the densities are floored and the fractions
    // sanitised already; only the pressure floor of enforce_positive applies
    template<int D, int M>
    static __device__ __forceinline__ void enforce_c2p(Pr<D, M>& Q) { Q.p = f32::fmaxs(Q.p, 1e-12f); }

    template<int D, int M>
    static __device__ __forceinline__ void alpha_rhs(const Eos& P, const Pr<D, M>& Pn, float* rhs)
    {
        f32::alpha_rhs<D, M>(P, Pn, rhs);
    }
    // P_line differs from P^n only by a second renormalisation of the
    // fractions: the predictor reuses the coefficients of P^n
    template<int D, int M>
    static __device__ __forceinline__ void alpha_rhs_line(const Eos&, const Pr<D, M>&, const float* rhs_n, float* rhs)
    {
#pragma unroll
        for (int k = 0; k < M; ++k) rhs[k] = rhs_n[k];
    }
    template<int D, int M>
    static __device__ __forceinline__ void reconstruct(const Pr<D, M>& L, const Pr<D, M>& C, const Pr<D, M>& Rt,
                                                       Pr<D, M>& lf, Pr<D, M>& rf)
    {
        f32::reconstruct<D, M>(L, C, Rt, lf, rf);
    }
    template<int D, int M>
    static __device__ __forceinline__ void prim_to_cons(const Eos& P, const Pr<D, M>& Q, St<D, M>& U)
    {
        float alpha[M];
        f32::prim_to_cons<D, M>(P, Q, U, alpha);
    }
    template<int D, int M>
    static __device__ __forceinline__ void repair(const Eos& P, St<D, M>& U) { f32::repair<D, M>(P, U); }

    // repair_state(prim_to_cons(Q)) changes nothing unless a fraction at or
    // below the inactive threshold (1e-10) carries more than 1e-12 mass: masses
    // are non-negative, fractions sanitised and the energy above its floor by
    // construction.  Only then is the full repair run.
    template<int D, int M>
    static __device__ __forceinline__ bool needs_repair(const St<D, M>& U, const float* alpha)
    {
        bool r = false;
#pragma unroll
        for (int k = 0; k < M; ++k) r |= (alpha[k] <= 1e-10f && U.m[k] > 1e-12f);
        return r;
    }

    template<int D, int M>
    static __device__ __forceinline__ void face_cons(const Eos& P, const Pr<D, M>& Q, St<D, M>& U)
    {
        float alpha[M];
        f32::prim_to_cons<D, M>(P, Q, U, alpha);
        if (needs_repair<D, M>(U, alpha)) f32::repair<D, M>(P, U);
    }

    // the face state of a compact face record (its primitive Q, momentum and
    // energy): the partial masses and the fraction code are those
    // prim_to_cons builds from Q (alpha_k rho_k, the code of Q's fractions),
    // exactly so unless the face went through repair or is an end cell,
    // where they agree to rounding
    template<int D, int M>
    static __device__ __forceinline__ void face_from_record(const Pr<D, M>& Q, const float* mom, float E, St<D, M>& U)
    {
#pragma unroll
        for (int k = 0; k < M; ++k) U.m[k] = Q.alpha[k] * Q.rho[k];
#pragma unroll
        for (int d = 0; d < D; ++d) U.mom[d] = mom[d];
        U.E = E;
        f32::encode_alpha<M>(Q.alpha, U.a);
    }

    // the face state of a face primitive record: prim_to_cons of it, which
    // is the face state itself unless that went through repair or is an end
    // cell (then the two agree to rounding)
    template<int D, int M>
    static __device__ __forceinline__ void face_from_prim(const Eos& P, const Pr<D, M>& Q, St<D, M>& U)
    {
        float alpha[M];
        f32::prim_to_cons<D, M>(P, Q, U, alpha);
    }

    // face state and its primitive: when repair has nothing to do the
    // primitive of prim_to_cons(Q) is Q with the same floors, which the face
    // already satisfies (densities >= 1e-12, p >= 1e-10; the fractions are
    // the ones the state was built from), so the pressure round trip is
    // skipped
    template<int D, int M>
    static __device__ __forceinline__ void face_state(const Eos& P, const Pr<D, M>& Q, St<D, M>& U, Pr<D, M>& QU)
    {
        float alpha[M];
        f32::prim_to_cons<D, M>(P, Q, U, alpha);
        if (needs_repair<D, M>(U, alpha)) {
            f32::repair<D, M>(P, U);
            f32::cons_to_prim<D, M>(P, U, QU);
            return;
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
            QU.alpha[k] = alpha[k];
            QU.rho[k] = Q.rho[k];
        }
#pragma unroll
        for (int d = 0; d < D; ++d) QU.vel[d] = Q.vel[d];
        QU.p = Q.p;
    }
    // a sweep's step sizes: dt and the host's 1/dx (uniform values that stay
    // in scalar registers); the FP32 factors are formed where they are used
    struct Factors {
        double dt, inv_dx;
    };
    static __device__ __forceinline__ Factors factors(double dt, double /*dx*/, double inv_dx) { return {dt, inv_dx}; }

    template<int D, int M, int DIR>
    static __device__ __forceinline__ void hancock(const Eos& P, const St<D, M>& Uc, const Pr<D, M>& Pc,
                                                   const St<D, M>& ul, const St<D, M>& ur, const Pr<D, M>& lf,
                                                   const Pr<D, M>& rf, const float* rhs, const Factors& f,
                                                   St<D, M>& out)
    {
        f32::hancock<D, M, DIR>(P, Uc, Pc, ul, ur, lf, rf, rhs, (float)f.dt, (float)f.inv_dx, out);
    }
    template<int D, int M, int DIR>
    static __device__ __forceinline__ int hllc(const Eos& P, const St<D, M>& UL, const St<D, M>& UR,
                                                const Pr<D, M>& PL, const Pr<D, M>& PR, Fx<D, M>& F, float& fv)
    {
        return f32::hllc<D, M, DIR>(P, UL, UR, PL, PR, F, fv);
    }

    template<int D, int M>
    static __device__ __forceinline__ void update(const Eos& P, const Cell<D, M>& Uin, const Pr<D, M>& Pn,
                                                  const float* rhs, const Fx<D, M>& Fl, const Fx<D, M>& Fr,
                                                  float fvl, float fvr,
                                                  const float* afl, const float* afr, const Factors& fac,
                                                  Cell<D, M>& r)
    {
        const double lambda = fac.dt * fac.inv_dx;
        const float lf = (float)lambda;
        r = Uin;
#pragma unroll
        for (int k = 0; k < M; ++k) r.m[k] -= lambda * ((double)Fr.m[k] - (double)Fl.m[k]);
#pragma unroll
        for (int d = 0; d < D; ++d) r.mom[d] -= lf * (Fr.mom[d] - Fl.mom[d]);
        r.E -= lf * (Fr.E - Fl.E);
        float an[M];
        const float div_u = (fvr - fvl) * (float)fac.inv_dx;
        const float dtf = (float)fac.dt;
#pragma unroll
        for (int k = 0; k < M; ++k) an[k] = Pn.alpha[k] - lf * (afr[k] - afl[k]) + dtf * rhs[k] * div_u;
        f32::sanitise<M>(an, 0.f);
        const auto set_fractions = [&]() {
            if (law_stiff(P)) {
                float ad_old[M], ad_new[M];
                f32::decode_alpha<M>(r.a, ad_old);
                f32::encode_alpha<M>(an, r.a);
                f32::decode_alpha<M>(r.a, ad_new);
                r.E += f32::bshift<M>(P, ad_old, ad_new);
            } else {
                f32::encode_alpha<M>(an, r.a);
            }
        };
        // The reference repairs the state before and after the fraction
        // update.  When neither repair renormalises a fraction code or
        // removes inactive mass (and the density is positive), both are the
        // energy floor alone, with the same partial masses and momentum: the
        // density and kinetic term are formed once and each floor applied
        // with its own fractions -- the same operations as the two full
        // repairs, so the same bits.  Otherwise both run in full.
#pragma unroll
        for (int k = 0; k < M; ++k) r.m[k] = (r.m[k] > 0.0) ? r.m[k] : 0.0;
        float ad0[M];
        f32::decode_alpha<M>(r.a, ad0);
        double rt = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) rt += r.m[k];
        if (f32::valid_code<M>(r.a) && !f32::inactive_mass<M>(ad0, r.m) && rt > 0.0) {
            const float rho = (float)rt;
            const float inv = f32::rcp(f32::fmaxs(rho, 1e-10f));
            float v2 = -0.f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const float v = r.mom[d] * inv;
                v2 += v * v;
            }
            const float h = 0.5f * rho;
            r.E = f32::fmaxs(r.E, h * v2 + f32::energy_floor<M>(P, ad0));
            set_fractions();
            float ad1[M];
            f32::decode_alpha<M>(r.a, ad1);
            if (f32::valid_code<M>(r.a) && !f32::inactive_mass<M>(ad1, r.m))
                r.E = f32::fmaxs(r.E, h * v2 + f32::energy_floor<M>(P, ad1));
            else
                f32::repair<D, M>(P, r);
        } else {
            f32::repair<D, M>(P, r);
            set_fractions();
            f32::repair<D, M>(P, r);
        }
    }

    template<int D, int M>
    static __device__ __forceinline__ void accumulate(const Eos& P, Cell<D, M>& acc, const Cell<D, M>& r,
                                                      const Cell<D, M>& U)
    {
#pragma unroll
        for (int k = 0; k < M; ++k) acc.m[k] += r.m[k] - U.m[k];
#pragma unroll
        for (int d = 0; d < D; ++d) acc.mom[d] += r.mom[d] - U.mom[d];
        float aa[M], ar[M], au[M], sum[M], dn[M];
        f32::decode_alpha<M>(acc.a, aa);
        f32::decode_alpha<M>(r.a, ar);
        f32::decode_alpha<M>(U.a, au);
#pragma unroll
        for (int k = 0; k < M; ++k) sum[k] = aa[k] + (ar[k] - au[k]);
        if constexpr (M == 2) {
            float cl[2] = {sum[0], sum[1]};
            if (!(cl[0] >= 0.f)) {
                cl[0] = 0.f;
                cl[1] = 1.f;
            } else if (!(cl[1] >= 0.f)) {
                cl[0] = 1.f;
                cl[1] = 0.f;
            }
            f32::encode_alpha<M>(cl, acc.a);
        } else {
#pragma unroll
            for (int k = 0; k < M - 1; ++k) acc.a[k] = sum[k];
        }
        f32::decode_alpha<M>(acc.a, dn);
        acc.E += (r.E - U.E) + f32::bshift<M>(P, sum, dn);
        f32::repair<D, M>(P, acc);
    }

    template<int D, int M>
    static __device__ __forceinline__ void repair_cell(const Eos& P, Cell<D, M>& U) { f32::repair<D, M>(P, U); }

    // FP64 wave rate of the cell, 1/dt of cfl.hpp:43-63 (split: max_d
    // (|u_d|+c)/(cfl dx_d); unsplit: sum_d (|u_d|+c)/dx_d / cfl), from the
    // FP32 wave speed; 0 when no direction has |u_d|+c > 1e-14
    template<int D, int M>
    static __device__ __forceinline__ double cfl_value(const Eos& P, const Cell<D, M>& C, const CflArgs& a,
                                                       double& /*p_out: unused*/)
    {
        St<D, M> U;
        to_state<D, M>(C, U);
        Pr<D, M> Q;
        f32::cons_to_prim<D, M>(P, U, Q);
        float rho = -0.f;
#pragma unroll
        for (int k = 0; k < M; ++k) rho += U.m[k];
        const double c = (double)f32::sound_speed<M>(P, rho, Q.p, Q.alpha, Q.rho);
        double rate = 0.0;
        if (!a.unsplit) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double ws = (double)__builtin_fabsf(Q.vel[d]) + c;
                if (ws > 1e-14) rate = __builtin_fmax(rate, ws * a.inv_cfl_dx[d]);
            }
            return rate;
        }
        double spectral = -0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) spectral += ((double)__builtin_fabsf(Q.vel[d]) + c) * a.inv_dx[d];
        return spectral > 1e-14 ? spectral * a.inv_cfl : 0.0;
    }

    template<int D, int M>
    static __device__ __forceinline__ double mass(const Cell<D, M>& U, int k) { return U.m[k]; }
    template<int D, int M>
    static __device__ __forceinline__ double mom(const Cell<D, M>& U, int d) { return (double)U.mom[d]; }
    template<int D, int M>
    static __device__ __forceinline__ double energy(const Eos& P, const Cell<D, M>& U) { return (double)U.E + Bd<M>(P, U.a); }
};

using Fp32Mixed = Fp32MixedT<EosMixed>;
using Fp32Ideal = Fp32MixedT<EosMixedIdeal>;

}  // namespace dimhip

#pragma clang fp contract(off)
