// io_kernels.hpp — device side of the driver's input and output: initial
// states from primitives, the per-cell values of a solution CSV row, and the
// physical flux through the domain boundary.  All arithmetic is the FP64
// path's (dim_fp64.hpp), whatever the storage form of the state.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_fp64.hpp"

namespace dimhip {

// Initial conditions build a primitive per cell and store
// prim_to_cons(P) (src/setup/dim_initial_conditions.hpp:531-628).
// prim rows: rho[M], alpha[M], vel[D], p  ->  packed state rows.
template<int D, int M>
__global__ void prim_to_packed_kernel(EosParamsNA P, const double* __restrict__ prim, double* __restrict__ packed,
                                      long long count)
{
    using LY = Layout<D, M>;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (long long)gridDim.x * blockDim.x) {
        Prim<D, M> Q;
        State<D, M> U;
        load_packed_prim<D, M>(prim + t * LY::NP, Q);
        prim_to_cons<D, M>(P, Q, U);
        store_packed<D, M>(packed + t * LY::NV, U);
    }
}

// Number of values per output row: rho, vel[D], p, e, alpha[M], rho_mat[M], mass[M].
template<int D, int M>
struct RowLayout {
    static constexpr int NR = 1 + D + 2 + 3 * M;
};

// The values of one solution CSV row (src/io/dim_write_csv.hpp:78-121):
// P = cons_to_prim(U), rho = total_density(U), e = max(safe_div(E - K, rho), 0)
// with K = 0.5 * rho * sum(P.vel^2).
template<int D, int M>
__global__ void output_rows_kernel(EosParamsNA P, const double* __restrict__ packed, double* __restrict__ rows,
                                   long long count)
{
    using LY = Layout<D, M>;
    constexpr int NR = RowLayout<D, M>::NR;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (long long)gridDim.x * blockDim.x) {
        State<D, M> U;
        Prim<D, M> Q;
        load_packed<D, M>(packed + t * LY::NV, U);
        cons_to_prim<D, M>(P, U, Q);
        const double rho = total_density<D, M>(U);
        double v2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) v2 += Q.vel[d] * Q.vel[d];
        const double kinetic = 0.5 * rho * v2;
        const double e = smax(safe_div(U.E - kinetic, rho), 0.0);
        double* r = rows + t * NR;
        int o = 0;
        r[o++] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) r[o++] = Q.vel[d];
        r[o++] = Q.p;
        r[o++] = e;
#pragma unroll
        for (int k = 0; k < M; ++k) r[o++] = Q.alpha[k];
#pragma unroll
        for (int k = 0; k < M; ++k) r[o++] = Q.rho[k];
#pragma unroll
        for (int k = 0; k < M; ++k) r[o++] = U.m[k];
    }
}

// Boundary flux totals: [mass_k (M), mass, mom (D), E].
template<int D, int M>
struct FluxTotals {
    static constexpr int NO = M + 1 + D + 1;
};

template<int D, int M, int DIR>
__device__ __forceinline__ void add_scaled_flux(double* acc, const State<D, M>& U, const EosParamsNA& P, double scale)
{
    Prim<D, M> Q;
    Flux<D, M> F;
    cons_to_prim<D, M>(P, U, Q);
    flux_normal<D, M, DIR>(U, Q, F);
    // add_dim_flux_to_totals (conservation_report.hpp:135-153)
    double mass_flux = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double v = scale * F.m[k];
        acc[k] += v;
        mass_flux += v;
    }
    acc[M] += mass_flux;
#pragma unroll
    for (int d = 0; d < D; ++d) acc[M + 1 + d] += scale * F.mom[d];
    acc[M + 1 + D] += scale * F.E;
}

template<typename Pol, int D, int M>
__device__ __forceinline__ void load_fp64_cell(const void* q, long long pitch, long long c, const typename Pol::Eos& PE,
                                               State<D, M>& U)
{
    typename Pol::template Cell<D, M> C;
    double packed[Layout<D, M>::NV];
    Pol::template load<D, M>(q, pitch, c, C);
    Pol::template to_packed<D, M>(PE, C, packed);
    load_packed<D, M>(packed, U);
}

// compute_dim_boundary_flux (conservation_report.hpp:253-305): every line
// normal to each face contributes +area * F(last cell) - area * F(first cell).
// Lines of all directions form one index space [0, line_start[D]); block b
// sums its lines in a fixed tree and writes partial[b].
template<typename Pol, int D, int M>
__global__ void __launch_bounds__(256)
boundary_flux_kernel(const void* __restrict__ q, Grid g, typename Pol::Eos PE, EosParamsNA P, long long line_start0,
                     long long line_start1, long long line_start2, long long line_start3, double* __restrict__ partial)
{
    constexpr int NO = FluxTotals<D, M>::NO;
    __shared__ double red[NO][256];
    const long long ls[4] = {line_start0, line_start1, line_start2, line_start3};
    double acc[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) acc[o] = 0.0;
    for (long long l = (long long)blockIdx.x * 256 + threadIdx.x; l < ls[D]; l += (long long)gridDim.x * 256) {
        int dir = 0;
        while (dir + 1 < D && l >= ls[dir + 1]) ++dir;
        long long rem = l - ls[dir];
        long long base = 0;
        for (int d = 0; d < D; ++d) {
            if (d == dir) continue;
            base += (rem % g.N[d]) * g.stride[d];
            rem /= g.N[d];
        }
        double area = 1.0;
        for (int d = 0; d < D; ++d)
            if (d != dir) area *= g.dx[d];
        const long long last = base + (long long)(g.N[dir] - 1) * g.stride[dir];
        State<D, M> UL, UR;
        load_fp64_cell<Pol, D, M>(q, g.pitch, base, PE, UL);
        load_fp64_cell<Pol, D, M>(q, g.pitch, last, PE, UR);
        if (dir == 0) {
            add_scaled_flux<D, M, 0>(acc, UR, P, area);
            add_scaled_flux<D, M, 0>(acc, UL, P, -area);
        }
        if constexpr (D >= 2)
            if (dir == 1) {
                add_scaled_flux<D, M, 1>(acc, UR, P, area);
                add_scaled_flux<D, M, 1>(acc, UL, P, -area);
            }
        if constexpr (D >= 3)
            if (dir == 2) {
                add_scaled_flux<D, M, 2>(acc, UR, P, area);
                add_scaled_flux<D, M, 2>(acc, UL, P, -area);
            }
    }
#pragma unroll
    for (int o = 0; o < NO; ++o) red[o][threadIdx.x] = acc[o];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int o = 0; o < NO; ++o) red[o][threadIdx.x] += red[o][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < NO) partial[(long long)blockIdx.x * NO + threadIdx.x] = red[threadIdx.x][0];
}

// Sums the block partials in order into out.  With integral != nullptr it
// also performs the report's accumulate_fluxes (conservation_report.hpp:506-520):
// integral += dt * flux, for a step that was taken (ctl->active).
static __global__ void finish_boundary_flux_kernel(const double* __restrict__ partial, int nblocks, int no,
                                            double* __restrict__ out, double* __restrict__ integral,
                                            const StepCtl* __restrict__ ctl)
{
    const int o = threadIdx.x;
    if (o >= no) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[(long long)b * no + o];
    out[o] = s;
    if (integral && ctl->active) integral[o] += ctl->dt * s;
}

}  // namespace dimhip
