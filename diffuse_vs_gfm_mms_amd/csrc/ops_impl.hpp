// ops_impl.hpp — definitions of the launcher templates of context.hpp.
// Included only by the per-policy translation units, which instantiate them.
#pragma once

#include "context.hpp"
#include "io_kernels.hpp"
#include "kernels_fp64.hpp"

namespace dimhip {

template<typename Pol, int D, int M, int DIR, int MODE, bool FUSE>
void launch_sweep(dimhip_ctx* c, const void* in, void* out, const double* pin, double* pout, const typename Pol::Eos& P,
                  const CflArgs& ca)
{
    const Grid& g = c->grid;
    if constexpr (DIR == 0) {
        constexpr int L = Pol::kSweepL0, W = Pol::kSweepW0, T = L - 6;
        dim3 blk(L, W, 1);
        dim3 grd((g.N[0] + T - 1) / T, (g.N[1] + W - 1) / W, g.N[2]);
        sweep_kernel<Pol, D, M, 0, L, W, MODE, FUSE><<<grd, blk, 0, c->stream>>>(in, out, g, P, c->ctl, ca, pin, pout);
    } else {
        constexpr int L = Pol::kSweepL1, W = Pol::kSweepW1, T = L - 6;
        dim3 blk(W, L, 1);
        dim3 grd((g.N[0] + W - 1) / W, (g.N[DIR] + T - 1) / T, DIR == 1 ? g.N[2] : g.N[1]);
        CflArgs a = ca;
        if (Pol::kXcdStrips && D == 2) a.band = (int)((grd.x + 7) / 8);
        sweep_kernel<Pol, D, M, DIR, L, W, MODE, FUSE><<<grd, blk, 0, c->stream>>>(in, out, g, P, c->ctl, a, pin, pout);
    }
}

template<typename Pol, int D, int M, int DIR>
void sweep_modes(dimhip_ctx* c, int mode, bool fuse, const void* in, void* out, const double* pin, double* pout,
                 const typename Pol::Eos& P)
{
    CflArgs ca = make_cfl_args(c, c->dt_ready_cfl, c->desc.time_update == DIMHIP_UNSPLIT);
    if (mode == 0 && c->xring_now) {
        ca.xring = 1;
        ca.xflip = (c->bc.bc_lo[0] == BC_REFLECTIVE ? 1 : 0) | (c->bc.bc_hi[0] == BC_REFLECTIVE ? 2 : 0);
    }
    if (mode == 0 && c->yring_now && D == 2) {
        ca.yring = 1;
        ca.yflip = (c->bc.bc_lo[1] == BC_REFLECTIVE ? 1 : 0) | (c->bc.bc_hi[1] == BC_REFLECTIVE ? 2 : 0);
    }
    if (mode == 0 && !fuse) launch_sweep<Pol, D, M, DIR, 0, false>(c, in, out, pin, nullptr, P, ca);
    else if (mode == 0 && fuse) launch_sweep<Pol, D, M, DIR, 0, true>(c, in, out, pin, pout, P, ca);
    else launch_sweep<Pol, D, M, DIR, 1, false>(c, in, out, pin, nullptr, P, ca);
}

template<typename Pol, int D, int M>
void PolOps<Pol, D, M>::sweep(dimhip_ctx* c, int dir, int mode, bool fuse, const void* in, void* out,
                              const double* pin, double* pout, const Eos& P)
{
    if (dir == 0) sweep_modes<Pol, D, M, 0>(c, mode, fuse, in, out, pin, pout, P);
    if constexpr (D >= 2)
        if (dir == 1) sweep_modes<Pol, D, M, 1>(c, mode, fuse, in, out, pin, pout, P);
    if constexpr (D >= 3)
        if (dir == 2) sweep_modes<Pol, D, M, 2>(c, mode, fuse, in, out, pin, pout, P);
}

template<typename Pol, int D, int M>
void PolOps<Pol, D, M>::bc(dimhip_ctx* c, void* q, int end_step, double* pstore, const Eos& P)
{
    const long long total = c->bc.face_start[2 * D];
    bool copy_kind = true;
    for (int d = 0; d < D; ++d)
        if (c->bc.bc_lo[d] == BC_NONREFLECTIVE || c->bc.bc_hi[d] == BC_NONREFLECTIVE) copy_kind = false;
    if (copy_kind) {
        // one value per thread: x blocks over the ring cells, y over the fields
        const int bx = grid_blocks(total, 256);
        const dim3 grd(bx < 2048 ? bx : 2048, 2 * M + D + (pstore ? 1 : 0), 1);
        bc_copy_kernel<Pol, D, M><<<grd, 256, 0, c->stream>>>(q, c->grid, c->bc, c->ctl, end_step, pstore,
                                                              c->xring_now ? 1 : 0);
        return;
    }
    const int blocks = grid_blocks(total, 256);
    bc_kernel<Pol, D, M><<<blocks, 256, 0, c->stream>>>(q, c->grid, P, c->bc, c->ctl, end_step, pstore);
}

template<typename Pol, int D, int M>
void PolOps<Pol, D, M>::cfl(dimhip_ctx* c, const void* q, bool unsplit, double cfl, double* pout, const Eos& P)
{
    const CflArgs ca = make_cfl_args(c, cfl, unsplit);
    const int blocks = grid_blocks(c->grid.ncells, 256 * 4);
    cfl_kernel<Pol, D, M><<<blocks, 256, 0, c->stream>>>(q, c->grid, P, c->ctl, ca, pout);
}

template<typename Pol, int D, int M>
void PolOps<Pol, D, M>::totals(dimhip_ctx* c, const void* q, const Eos& P)
{
    double volume = 1.0;
    for (int d = 0; d < D; ++d) volume *= c->grid.dx[d];
    const long long chunk = (c->grid.ncells + c->partial_blocks - 1) / c->partial_blocks;
    totals_kernel<Pol, D, M><<<c->partial_blocks, 256, 0, c->stream>>>(q, c->grid, P, volume, chunk, c->partial);
    sum_partials_kernel<<<1, 64, 0, c->stream>>>(c->partial, c->partial_blocks, M + 1 + D + 1, c->totals);
}

template<typename Pol, int D, int M>
void PolOps<Pol, D, M>::pack(dimhip_ctx* c, bool in, double* staging, void* q, long long first, long long count,
                             const Eos& P)
{
    const int blocks = grid_blocks(count, 256);
    if (in) pack_in_kernel<Pol, D, M><<<blocks, 256, 0, c->stream>>>(P, staging, q, first, count, c->grid.pitch);
    else pack_out_kernel<Pol, D, M><<<blocks, 256, 0, c->stream>>>(P, q, staging, first, count, c->grid.pitch);
}

template<typename Pol, int D, int M>
void PolOps<Pol, D, M>::boundary_flux(dimhip_ctx* c, const void* q, double* integral, const Eos& PE)
{
    const long long* L = c->flux_lines;
    boundary_flux_kernel<Pol, D, M><<<c->flux_blocks, 256, 0, c->stream>>>(q, c->grid, PE, c->eos_na, L[0], L[1], L[2],
                                                                           L[3], c->flux_partial);
    finish_boundary_flux_kernel<<<1, 64, 0, c->stream>>>(c->flux_partial, c->flux_blocks, FluxTotals<D, M>::NO,
                                                          c->flux_out, integral, c->ctl);
}

template<int D, int M, typename EP>
int KatOps<D, M, EP>::kat(const EP& P, int which, int dir, int count, const double* in0, const double* in1,
                          double* out0, double* out1, int* err, hipStream_t s)
{
    const int blocks = (count + 127) / 128;
    if (dir == 0) {
        kat64_kernel<D, M, 0, EP><<<blocks, 128, 0, s>>>(P, which, count, in0, in1, out0, out1, err);
        return 0;
    }
    if constexpr (D >= 2)
        if (dir == 1) {
            kat64_kernel<D, M, 1, EP><<<blocks, 128, 0, s>>>(P, which, count, in0, in1, out0, out1, err);
            return 0;
        }
    if constexpr (D >= 3)
        if (dir == 2) {
            kat64_kernel<D, M, 2, EP><<<blocks, 128, 0, s>>>(P, which, count, in0, in1, out0, out1, err);
            return 0;
        }
    return 1;
}

template<int D, int M>
void IoOps<D, M>::prim_to_packed(dimhip_ctx* c, const double* prim, double* packed, long long count)
{
    prim_to_packed_kernel<D, M><<<grid_blocks(count, 256), 256, 0, c->stream>>>(c->eos_na, prim, packed, count);
}

template<int D, int M>
void IoOps<D, M>::output_rows(dimhip_ctx* c, const double* packed, double* rows, long long count)
{
    output_rows_kernel<D, M><<<grid_blocks(count, 256), 256, 0, c->stream>>>(c->eos_na, packed, rows, count);
}

}  // namespace dimhip

// Explicit instantiation of one policy's launchers for every (dim, nmat).
#define DIMHIP_INSTANTIATE_POLICY(POL)          \
    template struct dimhip::PolOps<POL, 1, 1>; \
    template struct dimhip::PolOps<POL, 2, 1>; \
    template struct dimhip::PolOps<POL, 3, 1>; \
    template struct dimhip::PolOps<POL, 1, 2>; \
    template struct dimhip::PolOps<POL, 2, 2>; \
    template struct dimhip::PolOps<POL, 3, 2>; \
    template struct dimhip::PolOps<POL, 1, 3>; \
    template struct dimhip::PolOps<POL, 2, 3>; \
    template struct dimhip::PolOps<POL, 3, 3>;

#define DIMHIP_INSTANTIATE_KAT(EP)                 \
    template struct dimhip::KatOps<1, 1, EP>; \
    template struct dimhip::KatOps<2, 1, EP>; \
    template struct dimhip::KatOps<3, 1, EP>; \
    template struct dimhip::KatOps<1, 2, EP>; \
    template struct dimhip::KatOps<2, 2, EP>; \
    template struct dimhip::KatOps<3, 2, EP>; \
    template struct dimhip::KatOps<1, 3, EP>; \
    template struct dimhip::KatOps<2, 3, EP>; \
    template struct dimhip::KatOps<3, 3, EP>;
