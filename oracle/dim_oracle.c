/*
 * dim_oracle.c — CPU restatement of the reference DIM update (plain C11).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  It is imported
 * by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg,
 * never by the product path.
 *
 * Every function follows the reference's evaluation order operation by
 * operation (IEEE double, no FMA contraction: build with -ffp-contract=off),
 * because the reference's results are reproduced bit for bit.  Citations are
 * reference file:line.
 */
#include "dim_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[256];

static int fail(const char* msg)
{
    snprintf(g_err, sizeof g_err, "%s", msg);
    return 1;
}

const char* orc_last_error(void) { return g_err; }

/* ---- src/math/numerical_safety.hpp ------------------------------------ */
/* std::max / std::min semantics: max(a,b) = (a<b)?b:a, min(a,b) = (b<a)?b:a */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

/* numerical_safety.hpp:12-18 */
static inline double clamp_min_t(double x, double tol)
{
    if (!isfinite(x)) return tol;
    return smax(x, tol);
}
static inline double clamp_min(double x) { return clamp_min_t(x, 1e-10); }

/* numerical_safety.hpp:21-24 */
static inline double finite_or(double x, double fb) { return isfinite(x) ? x : fb; }

/* numerical_safety.hpp:27-39 */
static inline double safe_denom(double d, double tol)
{
    if (!isfinite(d)) return (d < 0.0) ? -tol : tol;
    if (fabs(d) < tol) return (d < 0.0) ? -tol : tol;
    return d;
}

/* numerical_safety.hpp:41-45 */
static inline double safe_div(double n, double d, double tol)
{
    return finite_or(n, 0.0) / safe_denom(d, tol);
}

/* numerical_safety.hpp:67-78 ; returns 0 on the throwing branch */
static inline int require_positive(double x, double* out)
{
    if (!isfinite(x)) return 0;
    *out = (x <= 0.0) ? 1e-10 : x;
    return 1;
}

/* ---- per-material EoS: src/euler/eos.hpp ------------------------------ */
typedef struct {
    double m[ORC_MAX_MAT];
    double mom[ORC_MAX_DIM];
    double E;
    double a[ORC_MAX_MAT]; /* nmat-1 independent volume fractions */
} st_t;

typedef struct {
    double rho[ORC_MAX_MAT];
    double alpha[ORC_MAX_MAT];
    double vel[ORC_MAX_DIM];
    double p;
} pr_t;

typedef struct {
    double m[ORC_MAX_MAT];
    double mom[ORC_MAX_DIM];
    double E;
} fx_t;

/* IdealGasEOS::pressure / StiffenedGasEOS::pressure / NobleAbelGasEOS::pressure
 * on a Conserved<1> with zero momentum (eos.hpp:451-466, 555-571, 665-677;
 * the noble-Abel internal energy is internal_energy_from_conserved,
 * eos.hpp:13-25). */
static double mat_pressure_c1(const orc_params* P, int k, double rho_in, double E)
{
    const double rho = clamp_min(rho_in);
    const double mom2 = 0.0 + 0.0 * 0.0;
    const double kinetic = 0.5 * safe_div(mom2, rho, 1e-10);
    if (P->kind[k] == ORC_EOS_NOBLE_ABEL) {
        const double e = smax(safe_div(E - kinetic, rho, 1e-10), 0.0);
        const double denom = safe_denom(1.0 - P->covolume[k] * rho, 1e-14);
        return clamp_min((P->gamma[k] - 1.0) * rho * e / denom);
    }
    if (P->kind[k] == ORC_EOS_STIFFENED) {
        const double p_raw = (P->gamma[k] - 1.0) * (E - kinetic) - P->gamma[k] * P->pinf[k];
        return clamp_min(p_raw);
    }
    const double p_raw = (P->gamma[k] - 1.0) * (E - kinetic);
    return clamp_min(p_raw);
}

/* internal_energy (eos.hpp:482-489, 587-595, 692-702); the noble-Abel form
 * clamps rho and p and divides without a guard */
static double mat_internal_energy(const orc_params* P, int k, double rho, double p)
{
    if (P->kind[k] == ORC_EOS_NOBLE_ABEL) {
        const double r = clamp_min(rho);
        const double pp = clamp_min(p);
        return pp * (1.0 - P->covolume[k] * r) / ((P->gamma[k] - 1.0) * r);
    }
    if (P->kind[k] == ORC_EOS_STIFFENED)
        return safe_div((p + P->gamma[k] * P->pinf[k]), ((P->gamma[k] - 1.0) * rho), 1e-10);
    return safe_div(p, ((P->gamma[k] - 1.0) * rho), 1e-10);
}

/* sound_speed on a Conserved<1> (eos.hpp:469-479, 574-583, 680-689) */
static double mat_sound_speed_c1(const orc_params* P, int k, double rho_in, double E)
{
    const double rho = clamp_min(rho_in);
    const double p = mat_pressure_c1(P, k, rho_in, E);
    if (P->kind[k] == ORC_EOS_NOBLE_ABEL) {
        const double denom = safe_denom(rho * (1.0 - P->covolume[k] * rho), 1e-14);
        return sqrt(clamp_min(P->gamma[k] * p / denom));
    }
    if (P->kind[k] == ORC_EOS_STIFFENED)
        return sqrt(P->gamma[k] * safe_div((p + P->pinf[k]), rho, 1e-10));
    return sqrt(P->gamma[k] * safe_div(p, rho, 1e-10));
}

/* ---- mixture EoS: src/dim/eos.hpp ------------------------------------- */
/* eos.hpp:16-27 */
static double material_pressure_from_density_energy(const orc_params* P, int k, double rho,
                                                    double e)
{
    const double r = clamp_min(rho);
    const double E = r * smax(e, 0.0);
    return mat_pressure_c1(P, k, r, E);
}

/* eos.hpp:29-40 */
static double material_sound_speed(const orc_params* P, int k, double rho, double p)
{
    const double r = clamp_min(rho);
    const double E = r * mat_internal_energy(P, k, r, p);
    return mat_sound_speed_c1(P, k, r, E);
}

/* eos.hpp:42-57 */
static double material_energy_pressure_derivative(const orc_params* P, int k, double rho,
                                                  double p)
{
    rho = clamp_min(rho);
    p = clamp_min(p);
    const double dp = smax(1e-6 * smax(fabs(p), 1.0), 1e-8);
    const double p_lo = smax(p - dp, 1e-12);
    const double p_hi = p + dp;
    const double e_lo = mat_internal_energy(P, k, rho, p_lo);
    const double e_hi = mat_internal_energy(P, k, rho, p_hi);
    return smax(safe_div(e_hi - e_lo, p_hi - p_lo, 1e-14), 1e-14);
}

/* eos.hpp:59-75 */
static double mixture_internal_energy_density(const orc_params* P, double p,
                                              const double* alpha, const double* rho)
{
    double energy_density = 0.0;
    for (int k = 0; k < P->nmat; ++k) {
        const double alpha_k = smax(finite_or(alpha[k], 0.0), 0.0);
        const double rho_k = clamp_min(rho[k]);
        const double p_k = clamp_min(p);
        energy_density += alpha_k * rho_k * mat_internal_energy(P, k, rho_k, p_k);
    }
    return finite_or(energy_density, 0.0);
}

/* eos.hpp:77-143: bracket-doubling + bisection for sum_k a_k r_k e_k(r_k,p) = rho e */
static double pressure_from_internal_energy(const orc_params* P, double rho_total, double e,
                                            const double* alpha, const double* rho)
{
    rho_total = clamp_min(rho_total);
    e = finite_or(e, 0.0);
    const double target = finite_or(rho_total * e, 0.0);
    double fallback = 0.0;
    for (int k = 0; k < P->nmat; ++k) {
        const double ek = target / rho_total;
        fallback += smax(finite_or(alpha[k], 0.0), 0.0) *
                    material_pressure_from_density_energy(P, k, clamp_min(rho[k]), ek);
    }
    fallback = clamp_min(fallback);

    double lo = 1e-12;
    double hi = smax(1.0, fallback);
    double flo = mixture_internal_energy_density(P, lo, alpha, rho) - target;
    double fhi = mixture_internal_energy_density(P, hi, alpha, rho) - target;
    for (int n = 0; n < 80 && isfinite(fhi) && fhi < 0.0; ++n) {
        hi *= 2.0;
        fhi = mixture_internal_energy_density(P, hi, alpha, rho) - target;
    }
    if (!isfinite(flo) || !isfinite(fhi) || !(flo <= 0.0 && fhi >= 0.0))
        return clamp_min(fallback);

    for (int n = 0; n < 80; ++n) {
        const double mid = 0.5 * (lo + hi);
        const double fmid = mixture_internal_energy_density(P, mid, alpha, rho) - target;
        if (!isfinite(fmid)) return clamp_min(fallback);
        if (fabs(fmid) <= 1e-10 * smax(1.0, target)) return clamp_min(mid);
        if (fmid > 0.0) hi = mid;
        else lo = mid;
    }
    return clamp_min(0.5 * (lo + hi));
}

/* eos.hpp:163-181 */
static double internal_energy_from_pressure(const orc_params* P, double rho_total, double p,
                                            const double* alpha, const double* rho)
{
    rho_total = clamp_min(rho_total);
    p = clamp_min(p);
    return mixture_internal_energy_density(P, p, alpha, rho) / rho_total;
}

/* eos.hpp:195-221 (Wood) */
static double mixture_sound_speed(const orc_params* P, double rho_total, double p,
                                  const double* alpha, const double* rho)
{
    rho_total = clamp_min(rho_total);
    p = clamp_min(p);
    double beta_mix = 0.0;
    for (int k = 0; k < P->nmat; ++k) {
        const double rho_k = clamp_min(rho[k]);
        const double c = material_sound_speed(P, k, rho_k, p);
        beta_mix += smax(finite_or(alpha[k], 0.0), 0.0) / safe_denom(rho_k * c * c, 1e-14);
    }
    return sqrt(1.0 / (rho_total * safe_denom(beta_mix, 1e-14)));
}

/* eos.hpp:223-275 */
static double generic_mixture_sound_speed(const orc_params* P, double rho_total, double p,
                                          const double* alpha, const double* rho)
{
    if (P->lambda_model == ORC_LAMBDA_KAPILA)
        return mixture_sound_speed(P, rho_total, p, alpha, rho);

    double c[ORC_MAX_MAT], xi[ORC_MAX_MAT];
    double xi_mix = 0.0, inverse_sound_sum = 0.0;
    for (int k = 0; k < P->nmat; ++k) {
        const double alpha_k = smax(finite_or(alpha[k], 0.0), 0.0);
        const double rho_k = clamp_min(rho[k]);
        c[k] = material_sound_speed(P, k, rho_k, p);
        xi[k] = rho_k * material_energy_pressure_derivative(P, k, rho_k, p);
        xi_mix += alpha_k * xi[k];
        inverse_sound_sum += alpha_k / safe_denom(c[k], 1e-14);
    }
    xi_mix = safe_denom(xi_mix, 1e-14);
    inverse_sound_sum = safe_denom(inverse_sound_sum, 1e-14);
    double rho_c_squared = 0.0;
    for (int k = 0; k < P->nmat; ++k) {
        const double alpha_k = smax(finite_or(alpha[k], 0.0), 0.0);
        double lambda_k = 1.0;
        if (P->lambda_model == ORC_LAMBDA_EQUAL_VELOCITY)
            lambda_k = safe_div(1.0 / safe_denom(c[k], 1e-14), inverse_sound_sum, 1e-14);
        rho_c_squared += lambda_k * alpha_k * xi[k] * c[k] * c[k] / xi_mix;
    }
    return sqrt(smax(rho_c_squared / clamp_min(rho_total), 1e-14));
}

/* ---- state algebra: src/dim/primitives.hpp ---------------------------- */
/* primitives.hpp:18-42 */
static void full_alpha(const orc_params* P, const double* indep, double* alpha)
{
    double sum = 0.0;
    for (int k = 0; k < P->nmat - 1; ++k) {
        alpha[k] = indep[k];
        sum += alpha[k];
    }
    alpha[P->nmat - 1] = 1.0 - sum;
}

/* primitives.hpp:45-70 */
static void sanitise_alpha(int n, double* alpha, double floor)
{
    double sum = 0.0;
    for (int k = 0; k < n; ++k) {
        alpha[k] = finite_or(alpha[k], floor);
        alpha[k] = smax(alpha[k], floor);
        sum += alpha[k];
    }
    if (!isfinite(sum) || sum <= 0.0) {
        for (int k = 0; k < n; ++k) alpha[k] = 0.0;
        alpha[0] = 1.0;
        return;
    }
    for (int k = 0; k < n; ++k) alpha[k] /= sum;
}

/* primitives.hpp:99-108 */
static double total_density(const orc_params* P, const st_t* U)
{
    double r = 0.0;
    for (int k = 0; k < P->nmat; ++k) r += finite_or(U->m[k], 0.0);
    return r;
}

/* primitives.hpp:111-162 */
static void cons_to_prim(const orc_params* P, const st_t* U, pr_t* out)
{
    const double rho_floor = 1e-12, p_floor = 1e-12, alpha_floor = 1e-12;
    pr_t R;
    memset(&R, 0, sizeof R);
    full_alpha(P, U->a, R.alpha);
    sanitise_alpha(P->nmat, R.alpha, alpha_floor);
    for (int k = 0; k < P->nmat; ++k) R.rho[k] = rho_floor;

    double rho_total = 0.0;
    for (int k = 0; k < P->nmat; ++k) rho_total += finite_or(U->m[k], 0.0);
    rho_total = smax(rho_total, rho_floor);

    double v2 = 0.0;
    for (int d = 0; d < P->dim; ++d) {
        R.vel[d] = safe_div(U->mom[d], rho_total, 1e-10);
        v2 += R.vel[d] * R.vel[d];
    }
    for (int k = 0; k < P->nmat; ++k) {
        const double alpha_k = smax(R.alpha[k], alpha_floor);
        R.rho[k] = smax(safe_div(finite_or(U->m[k], 0.0), alpha_k, 1e-10), rho_floor);
    }
    const double kinetic = 0.5 * rho_total * v2;
    const double ie = safe_div(finite_or(U->E, kinetic) - kinetic, rho_total, 1e-10);
    R.p = smax(pressure_from_internal_energy(P, rho_total, ie, R.alpha, R.rho), p_floor);
    *out = R;
}

/* primitives.hpp:166-212 */
static void prim_to_cons(const orc_params* P, const pr_t* Pin, st_t* out)
{
    const double rho_floor = 1e-12, p_floor = 1e-12, alpha_floor = 1e-12;
    pr_t Q = *Pin;
    sanitise_alpha(P->nmat, Q.alpha, alpha_floor);
    st_t U;
    memset(&U, 0, sizeof U);
    double rho_total = 0.0;
    for (int k = 0; k < P->nmat; ++k) {
        Q.rho[k] = smax(Q.rho[k], rho_floor);
        U.m[k] = Q.alpha[k] * Q.rho[k];
        rho_total += U.m[k];
    }
    rho_total = smax(rho_total, rho_floor);
    Q.p = smax(Q.p, p_floor);
    double v2 = 0.0;
    for (int d = 0; d < P->dim; ++d) {
        U.mom[d] = rho_total * Q.vel[d];
        v2 += Q.vel[d] * Q.vel[d];
    }
    const double ie = internal_energy_from_pressure(P, rho_total, Q.p, Q.alpha, Q.rho);
    U.E = rho_total * ie + 0.5 * rho_total * v2;
    for (int k = 0; k < P->nmat - 1; ++k) U.a[k] = Q.alpha[k];
    *out = U;
}

/* primitives.hpp:216-322 */
static void alpha_rhs_coefficients(const orc_params* P, const pr_t* Pr, double* rhs)
{
    const double floor = 1e-12;
    const double active_alpha_floor = P->alpha_source_floor;
    for (int k = 0; k < P->nmat; ++k) rhs[k] = Pr->alpha[k];
    sanitise_alpha(P->nmat, rhs, 0.0);
    if (P->lambda_model == ORC_LAMBDA_ALLAIRE) return;

    if (P->lambda_model == ORC_LAMBDA_EQUAL_VELOCITY) {
        double inverse_sound_sum = 0.0;
        double c[ORC_MAX_MAT];
        for (int k = 0; k < P->nmat; ++k) c[k] = floor;
        for (int k = 0; k < P->nmat; ++k) {
            const double alpha_k = smax(finite_or(rhs[k], 0.0), 0.0);
            if (alpha_k <= floor || alpha_k < smax(active_alpha_floor, 0.0)) continue;
            c[k] = material_sound_speed(P, k, smax(finite_or(Pr->rho[k], 0.0), floor),
                                        smax(finite_or(Pr->p, 0.0), floor));
            inverse_sound_sum += alpha_k / safe_denom(c[k], 1e-14);
        }
        if (inverse_sound_sum <= floor || !isfinite(inverse_sound_sum)) return;
        for (int k = 0; k < P->nmat; ++k) {
            const double alpha_k = smax(finite_or(rhs[k], 0.0), 0.0);
            if (alpha_k <= floor) {
                rhs[k] = 0.0;
                continue;
            }
            const double lambda_k =
                safe_div(1.0 / safe_denom(c[k], 1e-14), inverse_sound_sum, 1e-14);
            rhs[k] = alpha_k * lambda_k;
        }
        return;
    }

    const double source_alpha_floor = smax(active_alpha_floor, 0.0);
    double beta_mix = 0.0;
    double imp[ORC_MAX_MAT];
    int active[ORC_MAX_MAT];
    for (int k = 0; k < P->nmat; ++k) {
        imp[k] = floor;
        active[k] = 0;
    }
    for (int k = 0; k < P->nmat; ++k) {
        const double alpha_k = smax(finite_or(rhs[k], 0.0), 0.0);
        if (alpha_k <= floor || alpha_k < source_alpha_floor) continue;
        active[k] = 1;
        const double rho_k = smax(finite_or(Pr->rho[k], 0.0), floor);
        const double c_k = material_sound_speed(P, k, rho_k, smax(finite_or(Pr->p, 0.0), floor));
        imp[k] = smax(rho_k * c_k * c_k, floor);
        beta_mix += alpha_k / imp[k];
    }
    if (beta_mix <= floor || !isfinite(beta_mix)) return;
    const double q_mix = 1.0 / beta_mix;
    for (int k = 0; k < P->nmat; ++k) {
        const double alpha_k = smax(finite_or(rhs[k], 0.0), 0.0);
        if (!active[k] || alpha_k <= floor) {
            rhs[k] = 0.0;
            continue;
        }
        const double K_k = alpha_k * (q_mix / imp[k] - 1.0);
        rhs[k] = alpha_k + K_k;
    }
}

/* primitives.hpp:326-457 */
static void repair_state(const orc_params* P, st_t* U)
{
    const double mass_floor = 1e-12, p_floor = 1e-12, inactive_alpha_floor = 1e-10;
    const int nmat = P->nmat, dim = P->dim;
    for (int k = 0; k < nmat; ++k) U->m[k] = smax(finite_or(U->m[k], 0.0), 0.0);
    for (int k = 0; k < nmat - 1; ++k) U->a[k] = finite_or(U->a[k], 0.0);
    for (int d = 0; d < dim; ++d) U->mom[d] = finite_or(U->mom[d], 0.0);
    U->E = finite_or(U->E, 0.0);

    const double rho_before = smax(total_density(P, U), mass_floor);
    double vel_before[ORC_MAX_DIM];
    double v2_before = 0.0;
    for (int d = 0; d < dim; ++d) {
        vel_before[d] = safe_div(U->mom[d], rho_before, 1e-10);
        v2_before += vel_before[d] * vel_before[d];
    }
    double alpha_before[ORC_MAX_MAT];
    full_alpha(P, U->a, alpha_before);
    sanitise_alpha(nmat, alpha_before, 0.0);
    double rho_mat_before[ORC_MAX_MAT];
    for (int k = 0; k < nmat; ++k) {
        const double alpha_k = smax(alpha_before[k], inactive_alpha_floor);
        rho_mat_before[k] = smax(safe_div(smax(U->m[k], 0.0), alpha_k, 1e-10), mass_floor);
    }
    const double kinetic_before = 0.5 * rho_before * v2_before;
    const double internal_before =
        safe_div(finite_or(U->E, kinetic_before) - kinetic_before, rho_before, 1e-10);
    const double pressure_before = smax(
        pressure_from_internal_energy(P, rho_before, internal_before, alpha_before, rho_mat_before),
        p_floor);

    for (int k = 0; k < nmat; ++k) U->m[k] = smax(finite_or(U->m[k], 0.0), 0.0);
    double alpha_full[ORC_MAX_MAT];
    full_alpha(P, U->a, alpha_full);
    sanitise_alpha(nmat, alpha_full, 0.0);
    for (int k = 0; k < nmat - 1; ++k) U->a[k] = alpha_full[k];

    int removed = 0;
    for (int k = 0; k < nmat; ++k) {
        if (alpha_full[k] <= inactive_alpha_floor && U->m[k] > mass_floor) {
            U->m[k] = 0.0;
            removed = 1;
        }
    }
    double rho_total = total_density(P, U);
    if (!isfinite(rho_total) || rho_total <= 0.0) {
        int dom = 0;
        for (int k = 1; k < nmat; ++k)
            if (alpha_full[k] > alpha_full[dom]) dom = k;
        U->m[dom] = mass_floor;
        rho_total = mass_floor;
    }
    if (removed) {
        for (int d = 0; d < dim; ++d) U->mom[d] = rho_total * vel_before[d];
        double rho_mat[ORC_MAX_MAT];
        for (int k = 0; k < nmat; ++k) {
            const double alpha_k = smax(alpha_full[k], inactive_alpha_floor);
            rho_mat[k] = smax(safe_div(U->m[k], alpha_k, 1e-10), mass_floor);
        }
        const double ie =
            internal_energy_from_pressure(P, rho_total, pressure_before, alpha_full, rho_mat);
        U->E = rho_total * ie + 0.5 * rho_total * v2_before;
    }
    double v2 = 0.0;
    for (int d = 0; d < dim; ++d) {
        const double v = safe_div(U->mom[d], rho_total, 1e-10);
        v2 += v * v; /* std::pow(v, 2) */
    }
    double rho_mat[ORC_MAX_MAT];
    for (int k = 0; k < nmat; ++k) {
        const double alpha_k = smax(alpha_full[k], inactive_alpha_floor);
        rho_mat[k] = smax(safe_div(U->m[k], alpha_k, 1e-10), mass_floor);
    }
    const double kinetic = 0.5 * rho_total * v2;
    const double minimum_internal = mixture_internal_energy_density(P, p_floor, alpha_full, rho_mat);
    U->E = smax(finite_or(U->E, 0.0), kinetic + minimum_internal);
}

/* ---- reconstruction: src/core/fv/limiters.hpp, src/dim/reconstruction/muscl.hpp */
/* limiters.hpp:9-15 */
static inline double minmod(double a, double b)
{
    if (a * b <= 0.0) return 0.0;
    return fabs(a) < fabs(b) ? a : b;
}

/* muscl.hpp:26-40 */
static void enforce_positive_primitive(const orc_params* P, pr_t* Q)
{
    for (int k = 0; k < P->nmat; ++k) Q->rho[k] = smax(Q->rho[k], 1e-12);
    sanitise_alpha(P->nmat, Q->alpha, 0.0);
    Q->p = smax(Q->p, 1e-12);
}

/* muscl.hpp:43-112 */
static void reconstruct_cell_faces(const orc_params* P, const pr_t* L, const pr_t* C,
                                   const pr_t* R, pr_t* left_face, pr_t* right_face)
{
    double s_rho[ORC_MAX_MAT], s_alpha[ORC_MAX_MAT], s_vel[ORC_MAX_DIM], s_p;
    for (int k = 0; k < P->nmat; ++k) s_rho[k] = minmod(C->rho[k] - L->rho[k], R->rho[k] - C->rho[k]);
    for (int k = 0; k < P->nmat; ++k)
        s_alpha[k] = minmod(C->alpha[k] - L->alpha[k], R->alpha[k] - C->alpha[k]);
    for (int d = 0; d < P->dim; ++d) s_vel[d] = minmod(C->vel[d] - L->vel[d], R->vel[d] - C->vel[d]);
    s_p = minmod(C->p - L->p, R->p - C->p);

    *left_face = *C;
    *right_face = *C;
    for (int k = 0; k < P->nmat; ++k) {
        left_face->rho[k] = C->rho[k] - 0.5 * s_rho[k];
        right_face->rho[k] = C->rho[k] + 0.5 * s_rho[k];
    }
    for (int k = 0; k < P->nmat; ++k) {
        left_face->alpha[k] = C->alpha[k] - 0.5 * s_alpha[k];
        right_face->alpha[k] = C->alpha[k] + 0.5 * s_alpha[k];
    }
    for (int d = 0; d < P->dim; ++d) {
        left_face->vel[d] = C->vel[d] - 0.5 * s_vel[d];
        right_face->vel[d] = C->vel[d] + 0.5 * s_vel[d];
    }
    left_face->p = C->p - 0.5 * s_p;
    right_face->p = C->p + 0.5 * s_p;
    enforce_positive_primitive(P, left_face);
    enforce_positive_primitive(P, right_face);
}

/* flux.hpp:11-43 with normal = e_dir */
static void flux_normal(const orc_params* P, const st_t* U, const pr_t* Q, int dir, fx_t* F)
{
    double n[ORC_MAX_DIM] = {0.0, 0.0, 0.0};
    n[dir] = 1.0;
    const double rho_total = total_density(P, U);
    double un = 0.0;
    for (int d = 0; d < P->dim; ++d) un += Q->vel[d] * n[d];
    for (int k = 0; k < P->nmat; ++k) F->m[k] = U->m[k] * un;
    for (int d = 0; d < P->dim; ++d) F->mom[d] = rho_total * Q->vel[d] * un + Q->p * n[d];
    F->E = (U->E + Q->p) * un;
}

/* muscl.hpp:143-191 */
static void hancock_predict(const orc_params* P, const st_t* Uc, const pr_t* Pc, const st_t* ULf,
                            const st_t* URf, const pr_t* PLf, const pr_t* PRf, int dir,
                            double dt, double dx, st_t* out)
{
    fx_t FL, FR;
    flux_normal(P, ULf, PLf, dir, &FL);
    flux_normal(P, URf, PRf, dir, &FR);
    const double factor = 0.5 * dt / dx;
    st_t pred = *Uc;
    for (int k = 0; k < P->nmat; ++k) pred.m[k] -= factor * (FR.m[k] - FL.m[k]);
    for (int d = 0; d < P->dim; ++d) pred.mom[d] -= factor * (FR.mom[d] - FL.mom[d]);
    pred.E -= factor * (FR.E - FL.E);

    const double u_left = PLf->vel[dir];
    const double u_right = PRf->vel[dir];
    const double div_u = (u_right - u_left) / dx;
    double rhs[ORC_MAX_MAT];
    alpha_rhs_coefficients(P, Pc, rhs);
    double alpha_new[ORC_MAX_MAT];
    for (int k = 0; k < P->nmat; ++k) alpha_new[k] = Pc->alpha[k];
    for (int k = 0; k < P->nmat; ++k) {
        alpha_new[k] -= factor * (PRf->alpha[k] * u_right - PLf->alpha[k] * u_left);
        alpha_new[k] += 0.5 * dt * rhs[k] * div_u;
    }
    sanitise_alpha(P->nmat, alpha_new, 0.0);
    for (int k = 0; k < P->nmat - 1; ++k) pred.a[k] = alpha_new[k];
    repair_state(P, &pred);
    *out = pred;
}

/* muscl.hpp:197-329 */
static int reconstruct_line_interfaces(const orc_params* P, int dir, int n, const st_t* U,
                                       double dt, double dx, st_t* ULf, st_t* URf)
{
    if (n < 2) return fail("dim::reconstruct_line_interfaces: need at least two cells");
    if (n < 3) {
        for (int f = 0; f < n - 1; ++f) {
            ULf[f] = U[f];
            URf[f] = U[f + 1];
        }
        return 0;
    }
    pr_t* Pl = (pr_t*)malloc(sizeof(pr_t) * (size_t)n);
    pr_t* Ph = (pr_t*)malloc(sizeof(pr_t) * (size_t)n);
    st_t* Uh = (st_t*)malloc(sizeof(st_t) * (size_t)n);
    st_t* Ulh = (st_t*)malloc(sizeof(st_t) * (size_t)n);
    st_t* Urh = (st_t*)malloc(sizeof(st_t) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        cons_to_prim(P, &U[i], &Pl[i]);
        enforce_positive_primitive(P, &Pl[i]);
    }
    Uh[0] = U[0];
    Uh[n - 1] = U[n - 1];
    for (int i = 1; i < n - 1; ++i) {
        pr_t pl, pr;
        st_t ul, ur;
        reconstruct_cell_faces(P, &Pl[i - 1], &Pl[i], &Pl[i + 1], &pl, &pr);
        prim_to_cons(P, &pl, &ul);
        prim_to_cons(P, &pr, &ur);
        repair_state(P, &ul);
        repair_state(P, &ur);
        hancock_predict(P, &U[i], &Pl[i], &ul, &ur, &pl, &pr, dir, dt, dx, &Uh[i]);
    }
    for (int i = 0; i < n; ++i) {
        cons_to_prim(P, &Uh[i], &Ph[i]);
        enforce_positive_primitive(P, &Ph[i]);
    }
    Ulh[0] = Uh[0];
    Urh[0] = Uh[0];
    Ulh[n - 1] = Uh[n - 1];
    Urh[n - 1] = Uh[n - 1];
    for (int i = 1; i < n - 1; ++i) {
        pr_t pl, pr;
        reconstruct_cell_faces(P, &Ph[i - 1], &Ph[i], &Ph[i + 1], &pl, &pr);
        prim_to_cons(P, &pl, &Ulh[i]);
        prim_to_cons(P, &pr, &Urh[i]);
        repair_state(P, &Ulh[i]);
        repair_state(P, &Urh[i]);
    }
    for (int f = 0; f < n - 1; ++f) {
        ULf[f] = Urh[f];
        URf[f] = Ulh[f + 1];
    }
    free(Pl);
    free(Ph);
    free(Uh);
    free(Ulh);
    free(Urh);
    return 0;
}

/* ---- HLLC: src/dim/riemann/hllc.hpp ----------------------------------- */
/* hllc.hpp:15-24 */
static inline double star_density(double rho, double sK, double unK, double s_star)
{
    return rho * safe_div(sK - unK, sK - s_star, 1e-10);
}

/* hllc.hpp:27-41 */
static inline double star_energy(double rho_star, double E, double rho, double p, double sK,
                                 double unK, double s_star)
{
    return rho_star * (safe_div(E, rho, 1e-10) +
                       (s_star - unK) * (s_star + safe_div(p, rho * (sK - unK), 1e-10)));
}

static void flux_avg(const orc_params* P, const fx_t* A, const fx_t* B, fx_t* F)
{
    for (int k = 0; k < P->nmat; ++k) F->m[k] = 0.5 * (A->m[k] + B->m[k]);
    for (int d = 0; d < P->dim; ++d) F->mom[d] = 0.5 * (A->mom[d] + B->mom[d]);
    F->E = 0.5 * (A->E + B->E);
}

/* hllc.hpp:47-59: (sR FL - sL FR + sL sR (UR-UL)) / (sR - sL), the division
 * being a multiplication by 1/(sR-sL) (conservative.hpp:80-88). */
static void hll_flux(const orc_params* P, const st_t* UL, const st_t* UR, const fx_t* FL,
                     const fx_t* FR, double sL, double sR, fx_t* F)
{
    const double inv = 1.0 / (sR - sL);
    const double sLsR = sL * sR;
    for (int k = 0; k < P->nmat; ++k)
        F->m[k] = inv * ((sR * FL->m[k] - sL * FR->m[k]) + sLsR * (UR->m[k] - UL->m[k]));
    for (int d = 0; d < P->dim; ++d)
        F->mom[d] = inv * ((sR * FL->mom[d] - sL * FR->mom[d]) + sLsR * (UR->mom[d] - UL->mom[d]));
    F->E = inv * ((sR * FL->E - sL * FR->E) + sLsR * (UR->E - UL->E));
}

/* hllc.hpp:62-195.  PL/PR are cons_to_prim(UL/UR) (hllc.hpp:85-86). */
static int hllc_flux(const orc_params* P, const st_t* UL, const st_t* UR, const pr_t* PL,
                     const pr_t* PR, int dir, fx_t* F, double* fv)
{
    double n[ORC_MAX_DIM] = {0.0, 0.0, 0.0};
    n[dir] = 1.0;
    double rhoL, rhoR, pL, pR;
    if (!require_positive(total_density(P, UL), &rhoL))
        return fail("dim::hllc_flux_normal: invalid left density (non-finite)");
    if (!require_positive(total_density(P, UR), &rhoR))
        return fail("dim::hllc_flux_normal: invalid right density (non-finite)");
    if (!require_positive(PL->p, &pL))
        return fail("dim::hllc_flux_normal: invalid left pressure (non-finite)");
    if (!require_positive(PR->p, &pR))
        return fail("dim::hllc_flux_normal: invalid right pressure (non-finite)");

    double unL = 0.0, unR = 0.0;
    for (int d = 0; d < P->dim; ++d) unL += PL->vel[d] * n[d];
    for (int d = 0; d < P->dim; ++d) unR += PR->vel[d] * n[d];

    const double cL = generic_mixture_sound_speed(P, rhoL, pL, PL->alpha, PL->rho);
    const double cR = generic_mixture_sound_speed(P, rhoR, pR, PR->alpha, PR->rho);
    const double sL = smin(unL - cL, unR - cR);
    const double sR = smax(unL + cL, unR + cR);
    const double numerator = (pR - pL) + rhoL * unL * (sL - unL) - rhoR * unR * (sR - unR);
    const double denom = rhoL * (sL - unL) - rhoR * (sR - unR);
    double s_star = safe_div(numerator, denom, 1e-10);
    if (fabs(pL - pR) < 1e-10 && fabs(unL - unR) < 1e-10) s_star = unL;

    fx_t FL, FR;
    flux_normal(P, UL, PL, dir, &FL);
    flux_normal(P, UR, PR, dir, &FR);

    if (!isfinite(sL) || !isfinite(sR) || !isfinite(s_star)) {
        flux_avg(P, &FL, &FR, F);
        *fv = 0.5 * (unL + unR);
        return 0;
    }
    if (0.0 <= sL) {
        *F = FL;
        *fv = unL;
        return 0;
    }
    if (sR <= 0.0) {
        *F = FR;
        *fv = unR;
        return 0;
    }
    if (!(sL < sR)) {
        flux_avg(P, &FL, &FR, F);
        *fv = 0.5 * (unL + unR);
        return 0;
    }
    const int left = (sL <= 0.0 && 0.0 <= s_star);
    const st_t* UK = left ? UL : UR;
    const pr_t* PK = left ? PL : PR;
    const fx_t* FK = left ? &FL : &FR;
    const double rhoK = left ? rhoL : rhoR;
    const double pK = left ? pL : pR;
    const double sK = left ? sL : sR;
    const double unK = left ? unL : unR;

    const double rho_star = star_density(rhoK, sK, unK, s_star);
    const double E_star = star_energy(rho_star, UK->E, rhoK, pK, sK, unK, s_star);
    if (rho_star <= 0.0 || !isfinite(rho_star) || !isfinite(E_star)) {
        hll_flux(P, UL, UR, &FL, &FR, sL, sR, F);
        *fv = s_star;
        return 0;
    }
    const double mass_scale = safe_div(rho_star, rhoK, 1e-10);
    const double delta = s_star - unK;
    for (int k = 0; k < P->nmat; ++k) F->m[k] = FK->m[k] + sK * (UK->m[k] * mass_scale - UK->m[k]);
    for (int d = 0; d < P->dim; ++d) {
        double u_star = PK->vel[d];
        u_star += delta * n[d];
        F->mom[d] = FK->mom[d] + sK * (rho_star * u_star - UK->mom[d]);
    }
    F->E = FK->E + sK * (E_star - UK->E);
    *fv = s_star;
    return 0;
}

/* ---- line solve: src/dim/solver/advance/sweep_core.hpp:16-149 ---------- */
static int solve_line(const orc_params* P, int dir, int n, const st_t* Uin, st_t* Uout,
                      double dx, double dt)
{
    const int nmat = P->nmat;
    if (n == 0) return fail("dim::solve_line: empty line");
    if (n == 1) {
        Uout[0] = Uin[0];
        return 0;
    }
    const double lambda = dt / dx;
    fx_t* F = (fx_t*)calloc((size_t)n + 1, sizeof(fx_t));
    double* fv = (double*)calloc((size_t)n + 1, sizeof(double));
    pr_t* prim = (pr_t*)malloc(sizeof(pr_t) * (size_t)n);
    st_t* ULf = (st_t*)malloc(sizeof(st_t) * (size_t)n);
    st_t* URf = (st_t*)malloc(sizeof(st_t) * (size_t)n);
    pr_t* PLf = (pr_t*)malloc(sizeof(pr_t) * (size_t)n);
    pr_t* PRf = (pr_t*)malloc(sizeof(pr_t) * (size_t)n);
    double* aflux = (double*)calloc(((size_t)n + 1) * ORC_MAX_MAT, sizeof(double));
    int rc = 0;

    for (int i = 0; i < n; ++i) cons_to_prim(P, &Uin[i], &prim[i]);
    rc = reconstruct_line_interfaces(P, dir, n, Uin, dt, dx, ULf, URf);
    if (rc) goto done;
    for (int f = 0; f < n - 1; ++f) {
        cons_to_prim(P, &ULf[f], &PLf[f]);
        cons_to_prim(P, &URf[f], &PRf[f]);
    }
    for (int i = 0; i < n - 1; ++i) {
        rc = hllc_flux(P, &ULf[i], &URf[i], &PLf[i], &PRf[i], dir, &F[i + 1], &fv[i + 1]);
        if (rc) goto done;
    }
    F[0] = F[1];
    F[n] = F[n - 1];
    fv[0] = fv[1];
    fv[n] = fv[n - 1];

    for (int i = 0; i < n; ++i) {
        st_t r = Uin[i];
        for (int k = 0; k < nmat; ++k) r.m[k] -= lambda * (F[i + 1].m[k] - F[i].m[k]);
        for (int d = 0; d < P->dim; ++d) r.mom[d] -= lambda * (F[i + 1].mom[d] - F[i].mom[d]);
        r.E -= lambda * (F[i + 1].E - F[i].E);
        repair_state(P, &r);
        Uout[i] = r;
    }
    for (int k = 0; k < nmat; ++k) {
        aflux[0 * ORC_MAX_MAT + k] = prim[0].alpha[k] * fv[0];
        aflux[(size_t)n * ORC_MAX_MAT + k] = prim[n - 1].alpha[k] * fv[n];
    }
    for (int f = 1; f < n; ++f) {
        const double* up = (fv[f] >= 0.0) ? PLf[f - 1].alpha : PRf[f - 1].alpha;
        for (int k = 0; k < nmat; ++k) aflux[(size_t)f * ORC_MAX_MAT + k] = up[k] * fv[f];
    }
    for (int i = 0; i < n; ++i) {
        double alpha_new[ORC_MAX_MAT], rhs[ORC_MAX_MAT];
        for (int k = 0; k < nmat; ++k) alpha_new[k] = prim[i].alpha[k];
        alpha_rhs_coefficients(P, &prim[i], rhs);
        const double div_u = (fv[i + 1] - fv[i]) / dx;
        for (int k = 0; k < nmat; ++k) {
            alpha_new[k] -= lambda * (aflux[(size_t)(i + 1) * ORC_MAX_MAT + k] -
                                      aflux[(size_t)i * ORC_MAX_MAT + k]);
            alpha_new[k] += dt * rhs[k] * div_u;
        }
        sanitise_alpha(nmat, alpha_new, 0.0);
        for (int k = 0; k < nmat - 1; ++k) Uout[i].a[k] = alpha_new[k];
        repair_state(P, &Uout[i]);
    }
done:
    free(F);
    free(fv);
    free(prim);
    free(ULf);
    free(URf);
    free(PLf);
    free(PRf);
    free(aflux);
    return rc;
}

/* ---- packing ----------------------------------------------------------- */
int orc_nv(const orc_params* p) { return 2 * p->nmat + p->dim; }

static void unpack(const orc_params* P, const double* src, st_t* U)
{
    memset(U, 0, sizeof *U);
    int o = 0;
    for (int k = 0; k < P->nmat; ++k) U->m[k] = src[o++];
    for (int d = 0; d < P->dim; ++d) U->mom[d] = src[o++];
    U->E = src[o++];
    for (int k = 0; k < P->nmat - 1; ++k) U->a[k] = src[o++];
}

static void pack(const orc_params* P, const st_t* U, double* dst)
{
    int o = 0;
    for (int k = 0; k < P->nmat; ++k) dst[o++] = U->m[k];
    for (int d = 0; d < P->dim; ++d) dst[o++] = U->mom[d];
    dst[o++] = U->E;
    for (int k = 0; k < P->nmat - 1; ++k) dst[o++] = U->a[k];
}

/* primitive packing: rho[nmat], alpha[nmat], vel[dim], p */
static int prim_len(const orc_params* P) { return 2 * P->nmat + P->dim + 1; }

static void pack_prim(const orc_params* P, const pr_t* Q, double* dst)
{
    int o = 0;
    for (int k = 0; k < P->nmat; ++k) dst[o++] = Q->rho[k];
    for (int k = 0; k < P->nmat; ++k) dst[o++] = Q->alpha[k];
    for (int d = 0; d < P->dim; ++d) dst[o++] = Q->vel[d];
    dst[o++] = Q->p;
}

static void unpack_prim(const orc_params* P, const double* src, pr_t* Q)
{
    memset(Q, 0, sizeof *Q);
    int o = 0;
    for (int k = 0; k < P->nmat; ++k) Q->rho[k] = src[o++];
    for (int k = 0; k < P->nmat; ++k) Q->alpha[k] = src[o++];
    for (int d = 0; d < P->dim; ++d) Q->vel[d] = src[o++];
    Q->p = src[o++];
}

static int check_params(const orc_params* P)
{
    if (P->nmat < 1 || P->nmat > ORC_MAX_MAT) return fail("oracle: nmat out of range");
    if (P->dim < 1 || P->dim > ORC_MAX_DIM) return fail("oracle: dim out of range");
    for (int k = 0; k < P->nmat; ++k)
        if (P->gamma[k] <= 1.0) return fail("dim::EOSParams: gamma must be > 1");
    return 0;
}

/* ---- exported function-level entry points ----------------------------- */
int orc_cons_to_prim(const orc_params* P, int count, const double* U, double* out)
{
    if (check_params(P)) return 1;
    const int nv = orc_nv(P), np = prim_len(P);
    for (int i = 0; i < count; ++i) {
        st_t s;
        pr_t q;
        unpack(P, U + (size_t)i * nv, &s);
        cons_to_prim(P, &s, &q);
        pack_prim(P, &q, out + (size_t)i * np);
    }
    return 0;
}

int orc_prim_to_cons(const orc_params* P, int count, const double* in, double* U)
{
    if (check_params(P)) return 1;
    const int nv = orc_nv(P), np = prim_len(P);
    for (int i = 0; i < count; ++i) {
        pr_t q;
        st_t s;
        unpack_prim(P, in + (size_t)i * np, &q);
        prim_to_cons(P, &q, &s);
        pack(P, &s, U + (size_t)i * nv);
    }
    return 0;
}

int orc_repair_state(const orc_params* P, int count, double* U)
{
    if (check_params(P)) return 1;
    const int nv = orc_nv(P);
    for (int i = 0; i < count; ++i) {
        st_t s;
        unpack(P, U + (size_t)i * nv, &s);
        repair_state(P, &s);
        pack(P, &s, U + (size_t)i * nv);
    }
    return 0;
}

int orc_pressure_from_internal_energy(const orc_params* P, int count, const double* rho_total,
                                      const double* e, const double* alpha, const double* rho,
                                      double* out)
{
    if (check_params(P)) return 1;
    for (int i = 0; i < count; ++i)
        out[i] = pressure_from_internal_energy(P, rho_total[i], e[i], alpha + (size_t)i * P->nmat,
                                               rho + (size_t)i * P->nmat);
    return 0;
}

int orc_mixture_sound_speed(const orc_params* P, int count, const double* rho_total,
                            const double* pr, const double* alpha, const double* rho, double* out)
{
    if (check_params(P)) return 1;
    for (int i = 0; i < count; ++i)
        out[i] = generic_mixture_sound_speed(P, rho_total[i], pr[i], alpha + (size_t)i * P->nmat,
                                             rho + (size_t)i * P->nmat);
    return 0;
}

int orc_alpha_rhs(const orc_params* P, int count, const double* in, double* rhs)
{
    if (check_params(P)) return 1;
    const int np = prim_len(P);
    for (int i = 0; i < count; ++i) {
        pr_t q;
        unpack_prim(P, in + (size_t)i * np, &q);
        alpha_rhs_coefficients(P, &q, rhs + (size_t)i * P->nmat);
    }
    return 0;
}

int orc_hllc(const orc_params* P, int dir, int count, const double* UL, const double* UR,
             double* F, double* fv)
{
    if (check_params(P)) return 1;
    const int nv = orc_nv(P), nf = P->nmat + P->dim + 1;
    for (int i = 0; i < count; ++i) {
        st_t l, r;
        pr_t pl, pr;
        fx_t f;
        unpack(P, UL + (size_t)i * nv, &l);
        unpack(P, UR + (size_t)i * nv, &r);
        cons_to_prim(P, &l, &pl);
        cons_to_prim(P, &r, &pr);
        if (hllc_flux(P, &l, &r, &pl, &pr, dir, &f, &fv[i])) return 1;
        int o = 0;
        for (int k = 0; k < P->nmat; ++k) F[(size_t)i * nf + o++] = f.m[k];
        for (int d = 0; d < P->dim; ++d) F[(size_t)i * nf + o++] = f.mom[d];
        F[(size_t)i * nf + o++] = f.E;
    }
    return 0;
}

int orc_reconstruct_line(const orc_params* P, int dir, int n, const double* U, double dt,
                         double dx, double* UL_face, double* UR_face)
{
    if (check_params(P)) return 1;
    if (n < 2) return fail("dim::reconstruct_line_interfaces: need at least two cells");
    const int nv = orc_nv(P);
    st_t* u = (st_t*)malloc(sizeof(st_t) * (size_t)n);
    st_t* l = (st_t*)malloc(sizeof(st_t) * (size_t)n);
    st_t* r = (st_t*)malloc(sizeof(st_t) * (size_t)n);
    for (int i = 0; i < n; ++i) unpack(P, U + (size_t)i * nv, &u[i]);
    int rc = reconstruct_line_interfaces(P, dir, n, u, dt, dx, l, r);
    if (!rc)
        for (int f = 0; f < n - 1; ++f) {
            pack(P, &l[f], UL_face + (size_t)f * nv);
            pack(P, &r[f], UR_face + (size_t)f * nv);
        }
    free(u);
    free(l);
    free(r);
    return rc;
}

int orc_solve_line(const orc_params* P, int dir, int n, const double* Uin, double* Uout,
                   double dx, double dt)
{
    if (check_params(P)) return 1;
    const int nv = orc_nv(P);
    st_t* a = (st_t*)malloc(sizeof(st_t) * (size_t)(n > 0 ? n : 1));
    st_t* b = (st_t*)malloc(sizeof(st_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) unpack(P, Uin + (size_t)i * nv, &a[i]);
    int rc = solve_line(P, dir, n, a, b, dx, dt);
    if (!rc)
        for (int i = 0; i < n; ++i) pack(P, &b[i], Uout + (size_t)i * nv);
    free(a);
    free(b);
    return rc;
}

/* ---- whole grid -------------------------------------------------------- */
static long total_cells(const orc_params* P, const int* N)
{
    long t = 1;
    for (int d = 0; d < P->dim; ++d) t *= N[d];
    return t;
}

static void strides(const orc_params* P, const int* N, long* s)
{
    s[0] = 1;
    for (int d = 1; d < P->dim; ++d) s[d] = s[d - 1] * N[d - 1];
}

/* cfl.hpp:15-66 (split) and 69-119 (unsplit) */
int orc_dt_cfl(const orc_params* P, const int* N, const double* dx, const double* U, double cfl,
               double dt_max, int unsplit, double* dt_out)
{
    if (check_params(P)) return 1;
    const long total = total_cells(P, N);
    if (total <= 0) return fail("dim::compute_dt_cfl: empty state vector");
    if (cfl <= 0.0) return fail("dim::compute_dt_cfl: cfl must be positive");
    if (dt_max <= 0.0) return fail("dim::compute_dt_cfl: dt_max must be positive");
    if (!unsplit)
        for (int d = 0; d < P->dim; ++d)
            if (dx[d] <= 0.0) return fail("dim::compute_dt_cfl: grid spacing must be positive");
    const int nv = orc_nv(P);
    double dt = dt_max;
    for (long i = 0; i < total; ++i) {
        st_t s;
        pr_t q;
        unpack(P, U + (size_t)i * nv, &s);
        cons_to_prim(P, &s, &q);
        const double c = generic_mixture_sound_speed(P, total_density(P, &s), q.p, q.alpha, q.rho);
        if (!unsplit) {
            for (int d = 0; d < P->dim; ++d) {
                const double ws = fabs(q.vel[d]) + c;
                if (ws > 1e-14) dt = smin(dt, cfl * dx[d] / ws);
            }
        } else {
            double spectral = 0.0;
            for (int d = 0; d < P->dim; ++d) spectral += (fabs(q.vel[d]) + c) / dx[d];
            if (spectral > 1e-14) dt = smin(dt, cfl / spectral);
        }
    }
    *dt_out = dt;
    return 0;
}

/* boundary.hpp:150-218 (local_nonreflective_state) */
static void nonreflective_state(const orc_params* P, const double* U, const int* N, const long* s,
                                const int* idx, const int* src_idx, st_t* out)
{
    const int dim = P->dim, nmat = P->nmat, nv = orc_nv(P);
    long src_lin = 0, sec_lin = 0;
    int second[ORC_MAX_DIM];
    for (int d = 0; d < dim; ++d) {
        second[d] = src_idx[d];
        if (idx[d] == 0 && N[d] > 2) second[d] = 2;
        else if (idx[d] == N[d] - 1 && N[d] > 2) second[d] = N[d] - 3;
    }
    for (int d = 0; d < dim; ++d) {
        src_lin += (long)src_idx[d] * s[d];
        sec_lin += (long)second[d] * s[d];
    }
    st_t first, sec, state, ex;
    unpack(P, U + (size_t)src_lin * nv, &first);
    unpack(P, U + (size_t)sec_lin * nv, &sec);
    state = first;
    ex = first;
    for (int k = 0; k < nmat; ++k) ex.m[k] = first.m[k] + (first.m[k] - sec.m[k]);
    for (int d = 0; d < dim; ++d) ex.mom[d] = first.mom[d] + (first.mom[d] - sec.mom[d]);
    ex.E = first.E + (first.E - sec.E);
    for (int k = 0; k < nmat - 1; ++k) ex.a[k] = first.a[k] + (first.a[k] - sec.a[k]);

    int axis = 0;
    for (int d = 0; d < dim; ++d)
        if (idx[d] == 0 || idx[d] == N[d] - 1) {
            axis = d;
            break;
        }
    const int outward = (idx[axis] == 0) ? -1 : 1;
    const double ex_rho = smax(total_density(P, &ex), 1e-12);
    const double ex_un = outward * safe_div(ex.mom[axis], ex_rho, 1e-10);
    if (ex_un > 0.0) {
        const double rho = smax(total_density(P, &state), 1e-12);
        double m2b = 0.0, m2a = 0.0;
        for (int d = 0; d < dim; ++d) m2b += state.mom[d] * state.mom[d];
        state.mom[axis] = ex.mom[axis];
        for (int d = 0; d < dim; ++d) m2a += state.mom[d] * state.mom[d];
        state.E += 0.5 * (m2a - m2b) / rho;
    }
    repair_state(P, &state);
    *out = state;
}

/* boundary.hpp:165-226 (apply_boundary_conditions): serial, in place */
int orc_apply_bc(const orc_params* P, const int* N, const int* bc_lo, const int* bc_hi, double* U)
{
    if (check_params(P)) return 1;
    const int dim = P->dim, nv = orc_nv(P);
    const long total = total_cells(P, N);
    long s[ORC_MAX_DIM];
    strides(P, N, s);
    for (long lin = 0; lin < total; ++lin) {
        int idx[ORC_MAX_DIM], src[ORC_MAX_DIM];
        long st = 1;
        for (int d = 0; d < dim; ++d) {
            idx[d] = (int)((lin / st) % N[d]);
            st *= N[d];
        }
        int is_boundary = 0;
        for (int d = 0; d < dim; ++d) {
            src[d] = idx[d];
            if (N[d] <= 1) continue;
            if (idx[d] == 0) {
                src[d] = 1;
                is_boundary = 1;
            } else if (idx[d] == N[d] - 1) {
                src[d] = N[d] - 2;
                is_boundary = 1;
            }
        }
        if (!is_boundary) continue;
        long src_lin = 0;
        for (int d = 0; d < dim; ++d) src_lin += (long)src[d] * s[d];
        int bc = ORC_BC_COPY;
        for (int d = 0; d < dim; ++d) {
            if (idx[d] == 0) {
                bc = bc_lo[d];
                break;
            }
            if (idx[d] == N[d] - 1) {
                bc = bc_hi[d];
                break;
            }
        }
        if (bc == ORC_BC_NONREFLECTIVE) {
            st_t out;
            nonreflective_state(P, U, N, s, idx, src, &out);
            pack(P, &out, U + (size_t)lin * nv);
            continue;
        }
        memcpy(U + (size_t)lin * nv, U + (size_t)src_lin * nv, sizeof(double) * (size_t)nv);
        for (int d = 0; d < dim; ++d) {
            if ((idx[d] == 0 && bc_lo[d] == ORC_BC_REFLECTIVE) ||
                (idx[d] == N[d] - 1 && bc_hi[d] == ORC_BC_REFLECTIVE))
                U[(size_t)lin * nv + P->nmat + d] *= -1.0;
        }
    }
    return 0;
}

/* sweep_core.hpp:151-202 */
int orc_sweep(const orc_params* P, const int* N, const double* dx, int dir, double dt,
              const double* Uin, double* Uout)
{
    if (check_params(P)) return 1;
    const int dim = P->dim, nv = orc_nv(P);
    const long total = total_cells(P, N);
    long s[ORC_MAX_DIM];
    strides(P, N, s);
    const int n = N[dir];
    const long lines = total / n;
    double* lin = (double*)malloc(sizeof(double) * (size_t)n * nv);
    double* lout = (double*)malloc(sizeof(double) * (size_t)n * nv);
    int rc = 0;
    for (long line = 0; line < lines && !rc; ++line) {
        int idx[ORC_MAX_DIM] = {0, 0, 0};
        long tmp = line;
        for (int d = dim - 1; d >= 0; --d) {
            if (d == dir) continue;
            idx[d] = (int)(tmp % N[d]);
            tmp /= N[d];
        }
        long base = 0;
        for (int d = 0; d < dim; ++d)
            if (d != dir) base += (long)idx[d] * s[d];
        for (int i = 0; i < n; ++i)
            memcpy(lin + (size_t)i * nv, Uin + (size_t)(base + (long)i * s[dir]) * nv,
                   sizeof(double) * (size_t)nv);
        rc = orc_solve_line(P, dir, n, lin, lout, dx[dir], dt);
        for (int i = 0; i < n && !rc; ++i)
            memcpy(Uout + (size_t)(base + (long)i * s[dir]) * nv, lout + (size_t)i * nv,
                   sizeof(double) * (size_t)nv);
    }
    free(lin);
    free(lout);
    return rc;
}

/* advance_step.hpp:29-149 + core/fv/directional_update.hpp:7-42 */
int orc_advance(const orc_params* P, const int* N, const double* dx, const int* bc_lo,
                const int* bc_hi, int unsplit, double cfl, double dt_max, double* U,
                double* dt_out)
{
    if (check_params(P)) return 1;
    const int dim = P->dim, nv = orc_nv(P), nmat = P->nmat;
    const long total = total_cells(P, N);
    for (int d = 0; d < dim; ++d)
        if (N[d] <= 0) return fail("dim::total_cells: grid size must be positive");
    double dt;
    if (orc_dt_cfl(P, N, dx, U, cfl, dt_max, unsplit, &dt)) return 1;
    const size_t bytes = sizeof(double) * (size_t)total * nv;
    int rc = 0;
    if (!unsplit) {
        double* next = (double*)malloc(bytes);
        for (int dir = 0; dir < dim && !rc; ++dir) {
            rc = orc_sweep(P, N, dx, dir, dt, U, next);
            if (!rc) rc = orc_apply_bc(P, N, bc_lo, bc_hi, next);
            if (!rc) memcpy(U, next, bytes);
        }
        free(next);
    } else {
        double* acc = (double*)malloc(bytes);
        double* dirU = (double*)malloc(bytes);
        memcpy(acc, U, bytes);
        for (int dir = 0; dir < dim && !rc; ++dir) {
            rc = orc_sweep(P, N, dx, dir, dt, U, dirU);
            if (rc) break;
            for (long i = 0; i < total; ++i) {
                double* a = acc + (size_t)i * nv;
                const double* b = dirU + (size_t)i * nv;
                const double* u = U + (size_t)i * nv;
                for (int k = 0; k < nmat; ++k) a[k] += b[k] - u[k];
                for (int d = 0; d < dim; ++d) a[nmat + d] += b[nmat + d] - u[nmat + d];
                a[nmat + dim] += b[nmat + dim] - u[nmat + dim];
                for (int k = 0; k < nmat - 1; ++k)
                    a[nmat + dim + 1 + k] += b[nmat + dim + 1 + k] - u[nmat + dim + 1 + k];
                st_t s;
                unpack(P, a, &s);
                repair_state(P, &s);
                pack(P, &s, a);
            }
        }
        if (!rc) rc = orc_apply_bc(P, N, bc_lo, bc_hi, acc);
        if (!rc) memcpy(U, acc, bytes);
        free(acc);
        free(dirU);
    }
    *dt_out = dt;
    return rc;
}

/* conservation_report.hpp:188-227 */
int orc_conservation_totals(const orc_params* P, const int* N, const double* dx, const double* U,
                            double* out)
{
    if (check_params(P)) return 1;
    const int dim = P->dim, nmat = P->nmat, nv = orc_nv(P);
    double volume = 1.0;
    for (int d = 0; d < dim; ++d) volume *= dx[d];
    const long total = total_cells(P, N);
    double mk[ORC_MAX_MAT] = {0}, mom[ORC_MAX_DIM] = {0}, mass = 0.0, E = 0.0;
    for (long i = 0; i < total; ++i) {
        const double* c = U + (size_t)i * nv;
        double cell_mass = 0.0;
        for (int k = 0; k < nmat; ++k) {
            const double part = c[k] * volume;
            mk[k] += part;
            cell_mass += part;
        }
        mass += cell_mass;
        E += c[nmat + dim] * volume;
        for (int d = 0; d < dim; ++d) mom[d] += c[nmat + d] * volume;
    }
    int o = 0;
    for (int k = 0; k < nmat; ++k) out[o++] = mk[k];
    out[o++] = mass;
    for (int d = 0; d < dim; ++d) out[o++] = mom[d];
    out[o++] = E;
    return 0;
}
