// ic_kernels.hpp — initial conditions generated on the device.
//
// The primitive of every cell is built from the case geometry exactly as the
// reference's initialisers build it on the host
// (src/setup/dim_initial_conditions.hpp:16-628): region boxes and oblique
// planar slabs blended with a tanh profile, the shock-bubble and the coated
// shock-bubble.  The arithmetic is the same FP64 operations in the same
// order; the one difference is tanh, which here is the device library's and
// may differ from the host C library's in the last bit, so diffuse cells can
// differ from the reference's initial state at the 1e-16 level (sharp
// interfaces, thickness 0, are identical).
#pragma once

#include <hip/hip_runtime.h>

#include "dim_params.hpp"

namespace dimhip {

constexpr int kIcMaxRegions = 16;

struct IcRegion {
    double lower[3], upper[3];
    double rho, vel[3], p;
    int slot;
};

struct IcPhase {
    double rho, vel[3], p;
    int slot;
};

enum : int { IC_REGIONS = 0, IC_PLANAR_REGIONS = 1, IC_SHOCK_BUBBLE = 2, IC_COATED_SHOCK_BUBBLE = 3 };

struct IcDesc {
    int kind, nregions;
    double domain_min[3], domain_max[3], dx[3];
    double thickness, sharpness;
    double normal[3];  // normalised on the host (dim_initial_conditions.hpp:33-55)
    IcRegion regions[kIcMaxRegions];
    int shock_axis;
    double shock_position;
    IcPhase left, right, bubble, film;
    double centre[3], radius, film_radius;
};

enum : int { IC_ERR_UNCOVERED = 1 };

template<int D, int M>
struct IcPrim {
    double rho[M];
    double alpha[M];
    double vel[D];
    double p;
};

template<int D, int M>
__device__ __forceinline__ void ic_blank(IcPrim<D, M>& P)
{
#pragma unroll
    for (int k = 0; k < M; ++k) {
        P.rho[k] = 1.0;
        P.alpha[k] = 0.0;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) P.vel[d] = 0.0;
    P.p = 0.0;
}

// sanitise_alpha with floor 0 (primitives.hpp:45-70)
template<int M>
__device__ __forceinline__ void ic_sanitise(double* alpha)
{
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        double v = __builtin_isfinite(alpha[k]) ? alpha[k] : 0.0;
        v = (v < 0.0) ? 0.0 : v;
        alpha[k] = v;
        sum += v;
    }
    if (!__builtin_isfinite(sum) || sum <= 0.0) {
#pragma unroll
        for (int k = 0; k < M; ++k) alpha[k] = (k == 0) ? 1.0 : 0.0;
        return;
    }
#pragma unroll
    for (int k = 0; k < M; ++k) alpha[k] /= sum;
}

template<int D>
__device__ __forceinline__ double ic_planar(const IcDesc& c, const double* x)
{
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) s += c.normal[d] * x[d];
    return s;
}

template<int D>
__device__ __forceinline__ bool ic_inside(const IcDesc& c, const IcRegion& r, const double* x)
{
    if (c.kind == IC_PLANAR_REGIONS) {
        const double s = ic_planar<D>(c, x);
        return s >= r.lower[0] && s < r.upper[0];
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (x[d] < r.lower[d] || x[d] >= r.upper[d]) return false;
    return true;
}

// signed_distance_to_box (dim_initial_conditions.hpp:81-133)
template<int D>
__device__ __forceinline__ double ic_box_distance(const IcDesc& c, const IcRegion& r, const double* x)
{
    double outside_sq = 0.0;
    double inside_dist = 1.7976931348623157e308;
    bool in = true, any_axis = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const bool lo_edge = __builtin_fabs(r.lower[d] - c.domain_min[d]) <= 1e-14;
        const bool hi_edge = __builtin_fabs(r.upper[d] - c.domain_max[d]) <= 1e-14;
        if (lo_edge && hi_edge) continue;
        any_axis = true;
        if (x[d] < r.lower[d]) {
            const double gap = r.lower[d] - x[d];
            outside_sq += gap * gap;
            in = false;
        } else if (x[d] > r.upper[d]) {
            const double gap = x[d] - r.upper[d];
            outside_sq += gap * gap;
            in = false;
        } else {
            const double a = x[d] - r.lower[d];
            const double b = r.upper[d] - x[d];
            const double m = (b < a) ? b : a;
            inside_dist = (m < inside_dist) ? m : inside_dist;
        }
    }
    if (!any_axis) return -1.7976931348623157e308;
    return in ? -inside_dist : __builtin_sqrt(outside_sq);
}

// signed_distance_to_planar_interval (dim_initial_conditions.hpp:150-167)
template<int D>
__device__ __forceinline__ double ic_slab_distance(const IcDesc& c, const IcRegion& r, const double* x)
{
    const double s = ic_planar<D>(c, x);
    if (s < r.lower[0]) return r.lower[0] - s;
    if (s > r.upper[0]) return s - r.upper[0];
    const double a = s - r.lower[0], b = r.upper[0] - s;
    return -((b < a) ? b : a);
}

template<int D>
__device__ __forceinline__ double ic_sphere_distance(const double* x, const double* centre, double radius)
{
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double delta = x[d] - centre[d];
        r2 += delta * delta;
    }
    return __builtin_sqrt(r2) - radius;
}

// smooth_indicator (dim_initial_conditions.hpp:218-229)
__device__ __forceinline__ double ic_weight(const IcDesc& c, double sd)
{
    if (c.thickness <= 0.0) return (sd <= 0.0) ? 1.0 : 0.0;
    const double a = (c.sharpness < 1.0e-12) ? 1.0e-12 : c.sharpness;
    return 0.5 * (1.0 - tanh(a * sd / c.thickness));
}

// sharp_region_primitive (dim_initial_conditions.hpp:185-216)
template<int D, int M>
__device__ __forceinline__ bool ic_sharp_region(const IcDesc& c, const double* x, IcPrim<D, M>& P)
{
    for (int i = 0; i < c.nregions; ++i) {
        const IcRegion& r = c.regions[i];
        if (!ic_inside<D>(c, r, x)) continue;
        ic_blank<D, M>(P);
#pragma unroll
        for (int d = 0; d < D; ++d) P.vel[d] = r.vel[d];
        P.p = r.p;
#pragma unroll
        for (int k = 0; k < M; ++k)
            if (k == r.slot) {
                P.alpha[k] = 1.0;
                P.rho[k] = r.rho;
            }
        return true;
    }
    return false;
}

// diffuse_region_primitive (dim_initial_conditions.hpp:232-312)
template<int D, int M>
__device__ __forceinline__ bool ic_diffuse_region(const IcDesc& c, const double* x, IcPrim<D, M>& P)
{
    if (c.thickness <= 0.0) return ic_sharp_region<D, M>(c, x, P);
    double score[M], rho_sum[M];
#pragma unroll
    for (int k = 0; k < M; ++k) score[k] = rho_sum[k] = 0.0;
    double total = 0.0, p_sum = 0.0, vel_sum[D];
#pragma unroll
    for (int d = 0; d < D; ++d) vel_sum[d] = 0.0;
    for (int i = 0; i < c.nregions; ++i) {
        const IcRegion& r = c.regions[i];
        const double sd = (c.kind == IC_PLANAR_REGIONS) ? ic_slab_distance<D>(c, r, x) : ic_box_distance<D>(c, r, x);
        const double w = ic_weight(c, sd);
        if (w <= 1e-14) continue;
#pragma unroll
        for (int k = 0; k < M; ++k)
            if (k == r.slot) {
                score[k] += w;
                rho_sum[k] += w * r.rho;
            }
        p_sum += w * r.p;
#pragma unroll
        for (int d = 0; d < D; ++d) vel_sum[d] += w * r.vel[d];
        total += w;
    }
    if (total <= 1e-14) return ic_sharp_region<D, M>(c, x, P);
    ic_blank<D, M>(P);
    P.p = p_sum / total;
#pragma unroll
    for (int d = 0; d < D; ++d) P.vel[d] = vel_sum[d] / total;
    double alpha_sum = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        P.alpha[k] = score[k] / total;
        alpha_sum += P.alpha[k];
        if (score[k] > 1e-14) P.rho[k] = rho_sum[k] / score[k];
    }
    if (alpha_sum <= 0.0) return ic_sharp_region<D, M>(c, x, P);
    ic_sanitise<M>(P.alpha);
    return true;
}

template<int M>
__device__ __forceinline__ void ic_blend(double* alpha, double* rho_sum, double* w_sum, int slot, double w, double rho)
{
    if (w <= 1e-14) return;
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (k == slot) {
            alpha[k] += w;
            rho_sum[k] += w * rho;
            w_sum[k] += w;
        }
}

// diffuse_shock_bubble_primitive (dim_initial_conditions.hpp:315-403)
template<int D, int M>
__device__ __forceinline__ void ic_shock_bubble(const IcDesc& c, const double* x, IcPrim<D, M>& P)
{
    const double sd = ic_sphere_distance<D>(x, c.centre, c.radius);
    const IcPhase& gas = (x[c.shock_axis] >= c.shock_position) ? c.right : c.left;
    ic_blank<D, M>(P);
    if (c.thickness <= 0.0) {
        const bool in = sd <= 0.0;
        const IcPhase& s = in ? c.bubble : gas;
#pragma unroll
        for (int k = 0; k < M; ++k)
            if (k == s.slot) {
                P.alpha[k] = 1.0;
                P.rho[k] = s.rho;
            }
#pragma unroll
        for (int d = 0; d < D; ++d) P.vel[d] = s.vel[d];
        P.p = s.p;
        return;
    }
    const double wb = ic_weight(c, sd);
    const double wg = 1.0 - wb;
    double rho_sum[M], w_sum[M];
#pragma unroll
    for (int k = 0; k < M; ++k) rho_sum[k] = w_sum[k] = 0.0;
    ic_blend<M>(P.alpha, rho_sum, w_sum, gas.slot, wg, gas.rho);
    ic_blend<M>(P.alpha, rho_sum, w_sum, c.bubble.slot, wb, c.bubble.rho);
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (w_sum[k] > 1e-14) P.rho[k] = rho_sum[k] / w_sum[k];
#pragma unroll
    for (int d = 0; d < D; ++d) P.vel[d] = wg * gas.vel[d] + wb * c.bubble.vel[d];
    P.p = wg * gas.p + wb * c.bubble.p;
    ic_sanitise<M>(P.alpha);
}

// diffuse_coated_shock_bubble_primitive (dim_initial_conditions.hpp:406-522)
template<int D, int M>
__device__ __forceinline__ void ic_coated_shock_bubble(const IcDesc& c, const double* x, IcPrim<D, M>& P)
{
    const double sd_in = ic_sphere_distance<D>(x, c.centre, c.radius);
    const double sd_out = ic_sphere_distance<D>(x, c.centre, c.film_radius);
    const IcPhase& gas = (x[c.shock_axis] >= c.shock_position) ? c.right : c.left;
    ic_blank<D, M>(P);
    if (c.thickness <= 0.0) {
        const IcPhase& s = (sd_in <= 0.0) ? c.bubble : (sd_out <= 0.0) ? c.film : gas;
#pragma unroll
        for (int k = 0; k < M; ++k)
            if (k == s.slot) {
                P.rho[k] = s.rho;
                P.alpha[k] = 1.0;
            }
#pragma unroll
        for (int d = 0; d < D; ++d) P.vel[d] = s.vel[d];
        P.p = s.p;
        return;
    }
    const double wb = ic_weight(c, sd_in);
    const double wc = ic_weight(c, sd_out);
    const double wf0 = wc - wb, wg0 = 1.0 - wc;
    const double wf = (wf0 < 0.0) ? 0.0 : wf0;
    const double wg = (wg0 < 0.0) ? 0.0 : wg0;
    double rho_sum[M], w_sum[M];
#pragma unroll
    for (int k = 0; k < M; ++k) rho_sum[k] = w_sum[k] = 0.0;
    ic_blend<M>(P.alpha, rho_sum, w_sum, gas.slot, wg, gas.rho);
    ic_blend<M>(P.alpha, rho_sum, w_sum, c.film.slot, wf, c.film.rho);
    ic_blend<M>(P.alpha, rho_sum, w_sum, c.bubble.slot, wb, c.bubble.rho);
    const double wt0 = wg + wf + wb;
    const double wt = (wt0 < 1e-14) ? 1e-14 : wt0;
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (w_sum[k] > 1e-14) P.rho[k] = rho_sum[k] / w_sum[k];
#pragma unroll
    for (int d = 0; d < D; ++d) P.vel[d] = (wg * gas.vel[d] + wf * c.film.vel[d] + wb * c.bubble.vel[d]) / wt;
    P.p = (wg * gas.p + wf * c.film.p + wb * c.bubble.p) / wt;
    ic_sanitise<M>(P.alpha);
}

// Primitive rows rho[M], alpha[M], vel[D], p of cells [first, first+count).
template<int D, int M>
__global__ void __launch_bounds__(256)
ic_prim_kernel(IcDesc c, int n0, int n1, int n2, long long first, long long count, double* __restrict__ prim,
               int* __restrict__ err)
{
    constexpr int NP = 2 * M + D + 1;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (long long)gridDim.x * blockDim.x) {
        const long long lin = first + t;
        // unflatten_index + compute_cell_center (geometry.hpp:57-89)
        const long long idx[3] = {lin % n0, (lin / n0) % n1, lin / ((long long)n0 * n1)};
        double x[D];
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = c.domain_min[d] + ((int)idx[d] + 0.5) * c.dx[d];
        IcPrim<D, M> P;
        bool ok = true;
        if (c.kind == IC_SHOCK_BUBBLE) ic_shock_bubble<D, M>(c, x, P);
        else if (c.kind == IC_COATED_SHOCK_BUBBLE) ic_coated_shock_bubble<D, M>(c, x, P);
        else ok = ic_diffuse_region<D, M>(c, x, P);
        if (!ok) {
            atomicOr(err, IC_ERR_UNCOVERED);
            ic_blank<D, M>(P);
        }
        double* row = prim + t * NP;
        int o = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) row[o++] = P.rho[k];
#pragma unroll
        for (int k = 0; k < M; ++k) row[o++] = P.alpha[k];
#pragma unroll
        for (int d = 0; d < D; ++d) row[o++] = P.vel[d];
        row[o++] = P.p;
    }
}

}  // namespace dimhip
