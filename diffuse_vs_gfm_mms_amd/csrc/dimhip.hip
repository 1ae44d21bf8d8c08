// dimhip.hip — the C ABI of include/dimhip.h on top of the gfx950 kernels.
//
// One context = one grid, one HIP stream, SoA state double-buffered in HBM
// plus a device-resident step-control block (CFL minimum slots, dt, time).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dimhip.h"
#include "batch.hpp"
#include "context.hpp"
#include "ic_kernels.hpp"
#include "kernels_fp64.hpp"

using namespace dimhip;

namespace {

thread_local std::string g_global_err;

}  // namespace

// ---- fixed scratch of dimhip_output_rows_noalloc ---------------------------
// One array per GPU, held for the life of the process: the packed FP64 cells
// (2 nmat + dim <= 9 doubles) and the CSV rows (3 + dim + 3 nmat <= 15
// doubles) of at most kNoAllocMaxCells cells, 192 B per cell, 12 MiB.  Every
// call on a GPU reuses it; each call writes every element it reads, so
// nothing in it needs to start at zero; calls that overlap on one GPU race.
constexpr long long kNoAllocMaxCells = 65536;
constexpr size_t kNoAllocDoubles = (size_t)kNoAllocMaxCells * (9 + 15);
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch array (the array's name is a host
// stand-in, not a device address): looked up once per device with the device
// current, then reused.  Returns nonzero on failure.
int scratch_address(int device, double** out)
{
    if (device < 0 || device >= kMaxDevices) return 1;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return 1;
        g_scratch_ptr[device] = (double*)p;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

}  // namespace


namespace {

int fail(dimhip_ctx* ctx, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    g_global_err = msg;
    return 1;
}

#define HIP_OK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail((ctx), std::string("HIP error: ") + hipGetErrorString(e_) +    \
                                   " at " #expr);                                      \
    } while (0)

// ---- event timing ---------------------------------------------------------
int ev_begin(dimhip_ctx* c)
{
    if (!c->timing_now) return -1;
    if (c->ev_next + 2 > c->ev_pool.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return -1;
            c->ev_pool.push_back(e);
        }
    }
    const int i = (int)c->ev_next;
    c->ev_next += 2;
    hipEventRecord(c->ev_pool[i], c->stream);
    return i;
}

void ev_end(dimhip_ctx* c, int i, int cls, int cls2 = -1)
{
    if (i < 0) return;
    hipEventRecord(c->ev_pool[i + 1], c->stream);
    c->ev_pending.push_back({cls, i});
    if (cls2 >= 0) c->ev_pending.push_back({cls2, i});
}

void ev_collect(dimhip_ctx* c)
{
    for (auto& pr : c->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_pool[pr.second], c->ev_pool[pr.second + 1]) == hipSuccess) {
            c->kt_sum[pr.first] += ms;
            c->kt_cnt[pr.first] += 1;
        }
    }
    c->ev_pending.clear();
    c->ev_next = 0;
}

// ---- per-(D, M) launch table ---------------------------------------------
// One (dim, nmat): dispatch on the context's precision form (and, in FP64,
// on whether a material is noble-Abel) to the policy launchers.
template<int D, int M>
struct OpsImpl final : Ops {
    void sweep(dimhip_ctx* c, int dir, int mode, bool fuse, const void* in, void* out, const double* pin,
               double* pout) const override
    {
        if (c->desc.precision != DIMHIP_FP64 && c->fp32_ideal)
            PolOps<Fp32Ideal, D, M>::sweep(c, dir, mode, fuse, in, out, nullptr, nullptr, c->eosmi);
        else if (c->desc.precision != DIMHIP_FP64)
            PolOps<Fp32Mixed, D, M>::sweep(c, dir, mode, fuse, in, out, nullptr, nullptr, c->eosm);
        else if (c->eos.noble_abel) PolOps<Fp64NA, D, M>::sweep(c, dir, mode, fuse, in, out, pin, pout, c->eos_na);
        else PolOps<Fp64, D, M>::sweep(c, dir, mode, fuse, in, out, pin, pout, c->eos);
    }

    bool xring_fits(const dimhip_ctx* c) const override
    {
        if (c->desc.precision != DIMHIP_FP64) return xring_fits_tiles<Fp32Mixed>(c->grid, D);
        return xring_fits_tiles<Fp64>(c->grid, D);
    }

    void bc(dimhip_ctx* c, void* q, int end_step, double* pstore) const override
    {
        if (c->desc.precision != DIMHIP_FP64 && c->fp32_ideal)
            PolOps<Fp32Ideal, D, M>::bc(c, q, end_step, nullptr, c->eosmi);
        else if (c->desc.precision != DIMHIP_FP64) PolOps<Fp32Mixed, D, M>::bc(c, q, end_step, nullptr, c->eosm);
        else if (c->eos.noble_abel) PolOps<Fp64NA, D, M>::bc(c, q, end_step, pstore, c->eos_na);
        else PolOps<Fp64, D, M>::bc(c, q, end_step, pstore, c->eos);
    }

    void cfl(dimhip_ctx* c, const void* q, bool unsplit, double cfl, double* pout) const override
    {
        if (c->desc.precision != DIMHIP_FP64 && c->fp32_ideal)
            PolOps<Fp32Ideal, D, M>::cfl(c, q, unsplit, cfl, nullptr, c->eosmi);
        else if (c->desc.precision != DIMHIP_FP64) PolOps<Fp32Mixed, D, M>::cfl(c, q, unsplit, cfl, nullptr, c->eosm);
        else if (c->eos.noble_abel) PolOps<Fp64NA, D, M>::cfl(c, q, unsplit, cfl, pout, c->eos_na);
        else PolOps<Fp64, D, M>::cfl(c, q, unsplit, cfl, pout, c->eos);
    }

    // the FP64 totals and packing touch no material law
    void totals(dimhip_ctx* c, const void* q) const override
    {
        if (c->desc.precision != DIMHIP_FP64) PolOps<Fp32Mixed, D, M>::totals(c, q, c->eosm);
        else PolOps<Fp64, D, M>::totals(c, q, c->eos);
    }

    void pack(dimhip_ctx* c, bool in, double* staging, void* q, long long first, long long count) const override
    {
        if (c->desc.precision != DIMHIP_FP64) PolOps<Fp32Mixed, D, M>::pack(c, in, staging, q, first, count, c->eosm);
        else PolOps<Fp64, D, M>::pack(c, in, staging, q, first, count, c->eos);
    }

    void prim_to_packed(dimhip_ctx* c, const double* prim, double* packed, long long count) const override
    {
        IoOps<D, M>::prim_to_packed(c, prim, packed, count);
    }

    void output_rows(dimhip_ctx* c, const double* packed, double* rows, long long count) const override
    {
        IoOps<D, M>::output_rows(c, packed, rows, count);
    }

    void ic_prims(dimhip_ctx* c, const IcDesc& ic, long long first, long long count, double* prim,
                  int* err) const override
    {
        ic_prim_kernel<D, M><<<grid_blocks(count, 256), 256, 0, c->stream>>>(ic, c->grid.N[0], c->grid.N[1],
                                                                             c->grid.N[2], first, count, prim, err);
    }

    void boundary_flux(dimhip_ctx* c, const void* q, double* integral) const override
    {
        if (c->desc.precision != DIMHIP_FP64) PolOps<Fp32Mixed, D, M>::boundary_flux(c, q, integral, c->eosm);
        else PolOps<Fp64, D, M>::boundary_flux(c, q, integral, c->eos);
    }

    int kat(const EosParams& P, int which, int dir, int count, const double* in0, const double* in1, double* out0,
            double* out1, int* err, hipStream_t s) const override
    {
        if (P.noble_abel)
            return KatOps<D, M, EosParamsNA>::kat(EosParamsNA(P), which, dir, count, in0, in1, out0, out1, err, s);
        return KatOps<D, M, EosParams>::kat(P, which, dir, count, in0, in1, out0, out1, err, s);
    }
};

const Ops* select_ops(int dim, int nmat)
{
    static const OpsImpl<1, 2> o12;
    static const OpsImpl<2, 2> o22;
    static const OpsImpl<3, 2> o32;
    static const OpsImpl<1, 1> o11;
    static const OpsImpl<2, 1> o21;
    static const OpsImpl<3, 1> o31;
    static const OpsImpl<1, 3> o13;
    static const OpsImpl<2, 3> o23;
    static const OpsImpl<3, 3> o33;
    if (nmat == 2) return dim == 1 ? (const Ops*)&o12 : dim == 2 ? (const Ops*)&o22 : (const Ops*)&o32;
    if (nmat == 1) return dim == 1 ? (const Ops*)&o11 : dim == 2 ? (const Ops*)&o21 : (const Ops*)&o31;
    if (nmat == 3) return dim == 1 ? (const Ops*)&o13 : dim == 2 ? (const Ops*)&o23 : (const Ops*)&o33;
    return nullptr;
}

// Material constants of a descriptor (unused slots: an inert ideal gas).
// gm1 and gpinf are formed with the reference's own double operations.
void fill_eos(EosParams& P, const dimhip_desc& d)
{
    P = EosParams{};
    for (int k = 0; k < kMaxMat; ++k) {
        const bool on = k < d.nmat;
        P.kind[k] = on ? d.eos_kind[k] : 0;
        P.gamma[k] = on ? d.gamma[k] : 1.4;
        P.pinf[k] = on ? d.p_inf[k] : 0.0;
        P.covol[k] = (on && P.kind[k] == EOS_NOBLE_ABEL) ? d.covolume[k] : 0.0;
        P.gm1[k] = P.gamma[k] - 1.0;
        P.gpinf[k] = P.gamma[k] * P.pinf[k];
        const double num = P.kind[k] == EOS_STIFFENED ? 1e-10 + P.gpinf[k] : 1e-10;
        const bool bounded = P.gm1[k] > 0.0 && P.gpinf[k] >= 0.0 && P.covol[k] >= 0.0;
        P.emin_ub[k] = bounded ? num / P.gm1[k] * (1.0 + 1e-13) : HUGE_VAL;
        if (!(P.emin_ub[k] < HUGE_VAL)) P.emin_ub[k] = HUGE_VAL;
        if (on && P.kind[k] == EOS_NOBLE_ABEL) P.noble_abel = 1;
    }
    P.lambda_model = d.lambda_model;
    P.alpha_source_floor = d.alpha_source_floor;
}

int validate_desc(const dimhip_desc* d, std::string& msg)
{
    if (d->dim < 1 || d->dim > 3) return msg = "dimhip: dim must be 1, 2 or 3", 1;
    if (d->nmat < 1 || d->nmat > 3) return msg = "dimhip: nmat must be 1, 2 or 3", 1;
    for (int k = 0; k < d->nmat; ++k) {
        if (!(d->gamma[k] > 1.0)) return msg = "dim::EOSParams: gamma must be > 1", 1;
        if (d->eos_kind[k] != DIMHIP_EOS_IDEAL && d->eos_kind[k] != DIMHIP_EOS_STIFFENED &&
            d->eos_kind[k] != DIMHIP_EOS_NOBLE_ABEL)
            return msg = "dimhip: unsupported EOS kind", 1;
        if (d->eos_kind[k] == DIMHIP_EOS_NOBLE_ABEL && d->precision == DIMHIP_FP32_MIXED)
            return msg = "dimhip: the FP32-mixed form supports ideal and stiffened gases only; "
                         "use DIMHIP_FP64 for noble-Abel materials", 1;
    }
    for (int a = 0; a < d->dim; ++a) {
        if (d->N[a] <= 0) return msg = "dim::total_cells: grid size must be positive", 1;
        if (d->N[a] == 2) return msg = "dimhip: grid extent 2 is not supported (boundary sources overlap)", 1;
        if (!(d->dx[a] > 0.0)) return msg = "dim::compute_dt_cfl: grid spacing must be positive", 1;
        for (int side = 0; side < 2; ++side) {
            const int b = side ? d->bc_hi[a] : d->bc_lo[a];
            if (b < 0 || b > 3) return msg = "Invalid boundary condition", 1;
            if (b == DIMHIP_BC_NONREFLECTIVE && d->N[a] < 4)
                return msg = "dimhip: nonreflective boundaries need at least 4 cells", 1;
        }
    }
    // the x and y sweeps address a tile's cells with 32-bit byte offsets:
    // up to 128 rows of N[0] doubles
    if (d->dim >= 2 && (long long)d->N[0] * 128 * 8 >= (1ll << 32))
        return msg = "dimhip: grid extent N[0] above 4194303 is not supported", 1;
    if (d->lambda_model < 0 || d->lambda_model > 2) return msg = "dim::alpha_rhs_coefficients: unknown lambda model", 1;
    if (d->time_update != DIMHIP_SPLIT && d->time_update != DIMHIP_UNSPLIT) return msg = "dimhip: unknown time_update", 1;
    if (d->precision != DIMHIP_FP64 && d->precision != DIMHIP_FP32_MIXED) return msg = "dimhip: unknown precision", 1;
    return 0;
}

// 2 * kDtSlots threads: the header and every dt slot of both parities
__global__ void init_ctl_kernel(StepCtl* ctl)
{
    if (threadIdx.x < 2 * kDtSlots)
        dt_slots(ctl, threadIdx.x / kDtSlots)[(threadIdx.x % kDtSlots) * kDtStride] = kInfBits;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->dtmin[0] = kInfBits;
        ctl->dtmin[1] = kInfBits;
        ctl->time = 0.0;
        ctl->dt = 0.0;
        ctl->dt_max = 0.0;
        ctl->t_end = 0.0;
        ctl->step = 0;
        ctl->slot = 0;
        ctl->cfl_parity = 1;
        ctl->active = 1;
        ctl->use_t_end = 0;
        ctl->err = 0;
        ctl->err_key = ~0ull;
    }
}

__global__ void set_mode_kernel(StepCtl* ctl, int use_t_end, double t_end, double time, int set_time)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->use_t_end = use_t_end;
        ctl->t_end = t_end;
        if (set_time) ctl->time = time;
        ctl->err = 0;
        ctl->err_key = ~0ull;
    }
}

__global__ void copy_kernel(const double* __restrict__ a, double* __restrict__ b, long long n)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        b[t] = a[t];
}

// a plain streaming copy of known size: the calibration kernel for HBM byte counters
__global__ void copy_probe_kernel(const double* __restrict__ a, double* __restrict__ b, long long n)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        b[t] = a[t];
}

void* state_ptr(dimhip_ctx* c, int which) { return c->q[which]; }

void copy_state(dimhip_ctx* c, const void* a, void* b)
{
    // storage sizes are whole doubles: pitch is a multiple of 64
    const long long n = (long long)(c->grid.pitch * c->cell_bytes / sizeof(double));
    copy_kernel<<<grid_blocks(n, 1024), 256, 0, c->stream>>>((const double*)a, (double*)b, n);
}

bool fusable(const dimhip_ctx* c)
{
    if (c->desc.time_update != DIMHIP_SPLIT) return false;
    for (int d = 0; d < c->desc.dim; ++d)
        if (c->desc.bc_lo[d] == DIMHIP_BC_NONREFLECTIVE || c->desc.bc_hi[d] == DIMHIP_BC_NONREFLECTIVE) return false;
    return true;
}

// Enqueue one advance_one_step.  When dt_max_host < 0 the step takes
// dt_max = t_end - time on the device.
void enqueue_step(dimhip_ctx* c, double cfl, double dt_max_host)
{
    const bool unsplit = c->desc.time_update == DIMHIP_UNSPLIT;
    const int D = c->desc.dim;
    c->timing_now = c->timing > 0 && c->timing_step++ % c->timing == 0;
    // The CFL pass of the step's input state (the previous step's fused
    // epilogue, or the stand-alone kernel below) leaves its recovered
    // pressures in pstore (FP64, dim >= 2): the sweeps that read that state
    // take them.
    double* ps = c->pstore;
    if (!c->dt_ready || c->dt_ready_cfl != cfl) {
        const int e = ev_begin(c);
        reset_dtmin_kernel<<<1, kDtSlots, 0, c->stream>>>(c->ctl, 0);
        c->ops->cfl(c, state_ptr(c, c->cur), unsplit, cfl, ps);
        ev_end(c, e, KC_CFL);
    }
    begin_step_kernel<<<1, kDtSlots, 0, c->stream>>>(c->ctl, dt_max_host);
    const bool fuse = fusable(c);
    c->dt_ready_cfl = cfl;
    c->xring_now = fuse && D == 3 && c->ops->xring_fits(c);
    c->yring_now = fuse && D == 2 && c->grid.N[0] >= 3 && c->grid.N[1] >= 3;
    if (!unsplit) {
        for (int dir = 0; dir < D; ++dir) {
            void* in = state_ptr(c, c->cur);
            void* out = state_ptr(c, c->cur ^ 1);
            const bool last = dir == D - 1;
            int e = ev_begin(c);
            c->ops->sweep(c, dir, 0, fuse && last, in, out, dir == 0 ? ps : nullptr, fuse && last ? ps : nullptr);
            ev_end(c, e, KC_SWEEP, KC_SWEEP_X + dir);
            // 2D, copy-kind boundaries: the x sweep wrote the y-face ring rows
            // the y sweep reads, and the pass after the y sweep restores the
            // rest of the ring
            if (!(c->yring_now && dir == 0)) {
                e = ev_begin(c);
                c->ops->bc(c, out, last, fuse && last ? ps : nullptr);
                ev_end(c, e, KC_BC);
            }
            c->cur ^= 1;
        }
    } else {
        void* base = state_ptr(c, c->cur);
        void* acc = state_ptr(c, c->cur ^ 1);
        copy_state(c, base, acc);
        for (int dir = 0; dir < D; ++dir) {
            const int e = ev_begin(c);
            c->ops->sweep(c, dir, 1, false, base, acc, ps, nullptr);
            ev_end(c, e, KC_SWEEP, KC_SWEEP_X + dir);
        }
        const int e = ev_begin(c);
        c->ops->bc(c, acc, 1, nullptr);
        ev_end(c, e, KC_BC);
        c->cur ^= 1;
    }
    c->xring_now = false;
    c->yring_now = false;
    c->dt_ready = fuse;
    if (c->track_flux) c->ops->boundary_flux(c, state_ptr(c, c->cur), c->flux_integral);
}

// The reference's message for a device error word.  A failing HLLC face
// throws from its first failing require_positive (hllc.hpp:88-91, the message
// and " (non-finite)", numerical_safety.hpp:67-78); of all failing faces the
// one the reference meets first is the one with the smallest order key
// (hllc_err_key), whose low two bits name the check.  Otherwise the caller's
// dt check (run_dim.hpp:105-107).
const char* device_error_text(int err, unsigned long long key)
{
    static const char* const hllc[4] = {
        "dim::hllc_flux_normal: invalid left density (non-finite)",
        "dim::hllc_flux_normal: invalid right density (non-finite)",
        "dim::hllc_flux_normal: invalid left pressure (non-finite)",
        "dim::hllc_flux_normal: invalid right pressure (non-finite)",
    };
    if (err & (ERR_HLLC_LEFT_DENSITY | ERR_HLLC_RIGHT_DENSITY | ERR_HLLC_LEFT_PRESSURE | ERR_HLLC_RIGHT_PRESSURE))
        return hllc[key & 3u];
    if (err & ERR_NONPOSITIVE_DT) return "run_dim_case: non-positive timestep";
    return nullptr;
}

int check_device_error(dimhip_ctx* c)
{
    if (const char* msg = device_error_text(c->ctl_host->err, c->ctl_host->err_key)) return fail(c, msg);
    return 0;
}

int fetch_ctl(dimhip_ctx* c)
{
    HIP_OK(c, hipMemcpyAsync(c->ctl_host, c->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    ev_collect(c);
    return 0;
}

void drop_graph(dimhip_ctx* c)
{
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_cfl = -1.0;
    c->graph_cur = -1;
}

int ensure_staging(dimhip_ctx* c, long long cells)
{
    if (c->staging_cells >= cells) return 0;
    hipFree(c->staging);
    c->staging = nullptr;
    c->staging_cells = 0;
    HIP_OK(c, hipMalloc(&c->staging, sizeof(double) * (size_t)cells * c->nv));
    c->staging_cells = cells;
    return 0;
}

int ensure_io(dimhip_ctx* c, size_t doubles)
{
    if (c->io_doubles >= doubles) return 0;
    hipFree(c->io_buf);
    c->io_buf = nullptr;
    c->io_doubles = 0;
    HIP_OK(c, hipMalloc(&c->io_buf, sizeof(double) * doubles));
    c->io_doubles = doubles;
    return 0;
}

}  // namespace

extern "C" {

int dimhip_version(void) { return DIMHIP_VERSION; }

const char* dimhip_last_error(const dimhip_ctx* ctx) { return ctx ? ctx->err.c_str() : g_global_err.c_str(); }

int dimhip_create(const dimhip_desc* desc, dimhip_ctx** out)
{
    if (!desc || !out) return fail(nullptr, "dimhip_create: null argument");
    std::string msg;
    if (validate_desc(desc, msg)) return fail(nullptr, msg);
    // an early return releases whatever was allocated so far
    std::unique_ptr<dimhip_ctx, int (*)(dimhip_ctx*)> c(new dimhip_ctx, dimhip_destroy);
    c->desc = *desc;
    for (int d = desc->dim; d < 3; ++d) {
        c->desc.N[d] = 1;
        c->desc.dx[d] = 1.0;
    }
    c->ops = select_ops(desc->dim, desc->nmat);
    if (!c->ops) return fail(nullptr, "dimhip: no kernels for this dim/nmat");
    HIP_OK(nullptr, hipSetDevice(desc->device));
    Grid& g = c->grid;
    for (int d = 0; d < 3; ++d) {
        g.N[d] = c->desc.N[d];
        g.dx[d] = c->desc.dx[d];
    }
    g.stride[0] = 1;
    g.stride[1] = g.N[0];
    g.stride[2] = (long long)g.N[0] * g.N[1];
    g.ncells = (long long)g.N[0] * g.N[1] * g.N[2];
    g.pitch = (g.ncells + 63) / 64 * 64;
    c->nv = 2 * desc->nmat + desc->dim;
    fill_eos(c->eos, *desc);
    c->eos_na = EosParamsNA(c->eos);
    for (int k = 0; k < kMaxMat; ++k) {
        c->eosm.kind[k] = c->eos.kind[k];
        c->eosm.gamma[k] = (float)c->eos.gamma[k];
        c->eosm.pinf[k] = (float)(c->eos.kind[k] == EOS_STIFFENED ? c->eos.pinf[k] : 0.0);
        c->eosm.gpinf[k] = (float)(c->eos.kind[k] == EOS_STIFFENED ? c->eos.gpinf[k] : 0.0);
        c->eosm.inv_gm1_d[k] = 1.0 / c->eos.gm1[k];
        c->eosm.b_d[k] = (c->eos.kind[k] == EOS_STIFFENED ? c->eos.gpinf[k] : 0.0) / c->eos.gm1[k];
        c->eosm.inv_gm1[k] = (float)c->eosm.inv_gm1_d[k];
        c->eosm.b[k] = (float)c->eosm.b_d[k];
        c->eosm.bmin[k] = (float)(1e-10 / c->eos.gm1[k]);
        c->eosm.gm1[k] = (float)c->eos.gm1[k];
        c->eosm.coef_lo[k] = std::min(1.0f, c->eosm.inv_gm1[k]);
        c->eosm.b_ghost[k] = c->eos.gamma[k] < 2.0 ? c->eosm.b[k] : 0.0f;
    }
    c->eosm.lambda_model = desc->lambda_model;
    c->eosm.stiff = 0;
    c->eosm.stiff_ghost = 0;
    for (int k = 0; k < desc->nmat; ++k) {
        if (desc->eos_kind[k] == DIMHIP_EOS_STIFFENED) c->eosm.stiff = 1;
        if (c->eosm.b_ghost[k] != 0.0f) c->eosm.stiff_ghost = 1;
    }
    c->eosm.alpha_source_floor = (float)desc->alpha_source_floor;
    c->eosmi = EosMixedIdeal(c->eosm);
    // the ideal-gas kernels also take min(1, 1/(gamma-1)) = 1 as a constant:
    // every material has gamma <= 2
    bool unit_coef_lo = true;
    for (int k = 0; k < desc->nmat; ++k) unit_coef_lo = unit_coef_lo && c->eosm.coef_lo[k] == 1.0f;
    c->fp32_ideal = !c->eosm.stiff && desc->lambda_model == DIMHIP_LAMBDA_KAPILA && unit_coef_lo;
    // ring cells per face (d, side), owned by the face of their lowest boundary axis
    long long acc = 0;
    for (int f = 0; f < 6; ++f) {
        c->bc.face_start[f] = acc;
        const int fd = f >> 1;
        if (fd >= desc->dim || g.N[fd] <= 1) continue;
        long long cnt = 1;
        for (int d = 0; d < desc->dim; ++d) {
            if (d == fd) continue;
            if (d < fd && g.N[d] > 1) cnt *= (g.N[d] - 2);
            else cnt *= g.N[d];
        }
        acc += cnt;
    }
    c->bc.face_start[6] = acc;
    for (int f = 2 * desc->dim; f <= 6; ++f) c->bc.face_start[f] = acc;
    for (int d = 0; d < 3; ++d) {
        c->bc.bc_lo[d] = d < desc->dim ? desc->bc_lo[d] : 0;
        c->bc.bc_hi[d] = d < desc->dim ? desc->bc_hi[d] : 0;
    }
    HIP_OK(nullptr, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->cell_bytes = desc->precision == DIMHIP_FP64
                        ? sizeof(double) * c->nv
                        : sizeof(double) * desc->nmat + sizeof(float) * (desc->dim + desc->nmat);
    const size_t bytes = (size_t)g.pitch * c->cell_bytes;
    HIP_OK(nullptr, hipMalloc(&c->q[0], bytes));
    HIP_OK(nullptr, hipMalloc(&c->q[1], bytes));
    HIP_OK(nullptr, hipMemsetAsync(c->q[0], 0, bytes, c->stream));
    HIP_OK(nullptr, hipMemsetAsync(c->q[1], 0, bytes, c->stream));
    // the recovered-pressure plane: one double per cell (not in 1D, where the
    // step's only sweep would read and write it)
    if (desc->precision == DIMHIP_FP64 && desc->dim >= 2)
        HIP_OK(nullptr, hipMalloc(&c->pstore, sizeof(double) * (size_t)g.pitch));
    HIP_OK(nullptr, hipMalloc(&c->ctl, kCtlAllocBytes));
    HIP_OK(nullptr, hipHostMalloc(&c->ctl_host, sizeof(StepCtl), hipHostMallocDefault));
    HIP_OK(nullptr, hipHostMalloc(&c->probe, sizeof(StepCtl), hipHostMallocDefault));
    HIP_OK(nullptr, hipEventCreateWithFlags(&c->probe_event, hipEventDisableTiming));
    c->partial_blocks = (int)std::min<long long>(1024, std::max<long long>(1, g.ncells / 4096));
    HIP_OK(nullptr, hipMalloc(&c->partial, sizeof(double) * (size_t)c->partial_blocks * 16));
    HIP_OK(nullptr, hipMalloc(&c->totals, sizeof(double) * 16));
    {
        long long acc_lines = 0;
        for (int d = 0; d < desc->dim; ++d) {
            c->flux_lines[d] = acc_lines;
            long long cnt = 1;
            for (int e = 0; e < desc->dim; ++e)
                if (e != d) cnt *= g.N[e];
            acc_lines += cnt;
        }
        for (int d = desc->dim; d < 4; ++d) c->flux_lines[d] = acc_lines;
        c->flux_blocks = (int)std::min<long long>(64, std::max<long long>(1, (acc_lines + 255) / 256));
    }
    HIP_OK(nullptr, hipMalloc(&c->flux_partial, sizeof(double) * (size_t)c->flux_blocks * 16));
    HIP_OK(nullptr, hipMalloc(&c->flux_out, sizeof(double) * 16));
    HIP_OK(nullptr, hipMalloc(&c->flux_integral, sizeof(double) * 16));
    HIP_OK(nullptr, hipMemsetAsync(c->flux_integral, 0, sizeof(double) * 16, c->stream));
    init_ctl_kernel<<<1, 2 * kDtSlots, 0, c->stream>>>(c->ctl);
    HIP_OK(nullptr, hipGetLastError());
    HIP_OK(nullptr, hipStreamSynchronize(c->stream));
    *out = c.release();
    return 0;
}

int dimhip_destroy(dimhip_ctx* c)
{
    if (!c) return 0;
    hipSetDevice(c->desc.device);
    if (c->stream) hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (auto e : c->ev_pool) hipEventDestroy(e);
    hipFree(c->q[0]);
    hipFree(c->q[1]);
    hipFree(c->pstore);
    hipFree(c->ctl);
    hipHostFree(c->ctl_host);
    hipHostFree(c->probe);
    if (c->probe_event) hipEventDestroy(c->probe_event);
    hipFree(c->staging);
    hipFree(c->partial);
    hipFree(c->totals);
    hipFree(c->io_buf);
    hipFree(c->flux_partial);
    hipFree(c->flux_out);
    hipFree(c->flux_integral);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int dimhip_upload(dimhip_ctx* c, const double* packed, size_t ncells)
{
    if (!c || !packed) return fail(c, "dimhip_upload: null argument");
    if ((long long)ncells != c->grid.ncells) return fail(c, "dim::advance_one_step: grid size mismatch");
    HIP_OK(c, hipSetDevice(c->desc.device));
    const long long chunk = std::min<long long>(c->grid.ncells, 1ll << 22);
    if (ensure_staging(c, chunk)) return 1;
    for (long long first = 0; first < c->grid.ncells; first += chunk) {
        const long long cnt = std::min(chunk, c->grid.ncells - first);
        HIP_OK(c, hipMemcpyAsync(c->staging, packed + first * c->nv, sizeof(double) * (size_t)(cnt * c->nv),
                                 hipMemcpyHostToDevice, c->stream));
        c->ops->pack(c, true, c->staging, c->q[c->cur], first, cnt);
        HIP_OK(c, hipGetLastError());
        HIP_OK(c, hipStreamSynchronize(c->stream));
    }
    c->dt_ready = false;
    return 0;
}

int dimhip_download(dimhip_ctx* c, double* packed, size_t ncells)
{
    if (!c || !packed) return fail(c, "dimhip_download: null argument");
    if ((long long)ncells != c->grid.ncells) return fail(c, "dim::advance_one_step: grid size mismatch");
    HIP_OK(c, hipSetDevice(c->desc.device));
    const long long chunk = std::min<long long>(c->grid.ncells, 1ll << 22);
    if (ensure_staging(c, chunk)) return 1;
    for (long long first = 0; first < c->grid.ncells; first += chunk) {
        const long long cnt = std::min(chunk, c->grid.ncells - first);
        c->ops->pack(c, false, c->staging, c->q[c->cur], first, cnt);
        HIP_OK(c, hipGetLastError());
        HIP_OK(c, hipMemcpyAsync(packed + first * c->nv, c->staging, sizeof(double) * (size_t)(cnt * c->nv),
                                 hipMemcpyDeviceToHost, c->stream));
        HIP_OK(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int dimhip_step(dimhip_ctx* c, double cfl, double dt_max, double* dt_out, dimhip_timings* timings)
{
    if (!c) return fail(c, "dimhip_step: null context");
    if (cfl <= 0.0) return fail(c, "dim::compute_dt_cfl: cfl must be positive");
    if (dt_max <= 0.0) return fail(c, "dim::compute_dt_cfl: dt_max must be positive");
    HIP_OK(c, hipSetDevice(c->desc.device));
    set_mode_kernel<<<1, 1, 0, c->stream>>>(c->ctl, 0, 0.0, 0.0, 0);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timings) {
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        hipEventRecord(e0, c->stream);
    }
    enqueue_step(c, cfl, dt_max);
    HIP_OK(c, hipGetLastError());
    if (timings) hipEventRecord(e1, c->stream);
    if (fetch_ctl(c)) return 1;
    if (timings) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        timings->cfl_seconds = 0.0;
        timings->sweep_seconds = ms * 1e-3;
        timings->boundary_seconds = 0.0;
        hipEventDestroy(e0);
        hipEventDestroy(e1);
    }
    if (check_device_error(c)) return 1;
    if (dt_out) *dt_out = c->ctl_host->dt;
    return 0;
}

int dimhip_advance(dimhip_ctx* c, int nsteps, double cfl, double t_end, int use_graph, double* time_inout,
                   long long* steps_done)
{
    if (!c || !time_inout) return fail(c, "dimhip_advance: null argument");
    if (cfl <= 0.0) return fail(c, "dim::compute_dt_cfl: cfl must be positive");
    HIP_OK(c, hipSetDevice(c->desc.device));
    set_mode_kernel<<<1, 1, 0, c->stream>>>(c->ctl, 1, t_end, *time_inout, 1);
    HIP_OK(c, hipMemcpyAsync(c->ctl_host, c->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    const long long step0 = c->ctl_host->step;
    if (nsteps > 0) {
        // the first step computes the CFL minimum of the uploaded state
        if (!c->dt_ready || c->dt_ready_cfl != cfl) {
            enqueue_step(c, cfl, -1.0);
            --nsteps;
        }
        const bool graphable = use_graph && !c->timing && c->dt_ready;
        if (graphable) {
            if (!c->graph_exec || c->graph_cfl != cfl || c->graph_cur != c->cur) {
                drop_graph(c);
                // one graph = two steps, so the double buffer returns to `cur`
                HIP_OK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                const int cur0 = c->cur;
                enqueue_step(c, cfl, -1.0);
                enqueue_step(c, cfl, -1.0);
                const bool round_trip = (c->cur == cur0);
                HIP_OK(c, hipStreamEndCapture(c->stream, &c->graph));
                HIP_OK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
                c->graph_cfl = cfl;
                c->graph_cur = c->cur;
                if (!round_trip) return fail(c, "dimhip_advance: graph does not return to its buffer");
            }
        }
        // Steps go out in chunks; after each chunk the step-control block is
        // copied back asynchronously, and once the device time has reached
        // t_end (or an error was flagged) no further chunk is enqueued.  The
        // check lags one chunk so the queue never drains: up to two chunks of
        // inactive steps (less than 2 * kChunk) can be issued past t_end.
        // They are launches that return at once: a 2D split step's inactive
        // sweeps copy nothing (sweep_tile).
        constexpr int kChunk = 256;
        bool probing = false;
        while (nsteps > 0) {
            int n = nsteps < kChunk ? nsteps : kChunk;
            nsteps -= n;
            if (graphable)
                for (; n >= 2; n -= 2) HIP_OK(c, hipGraphLaunch(c->graph_exec, c->stream));
            for (; n > 0; --n) enqueue_step(c, cfl, -1.0);
            HIP_OK(c, hipGetLastError());
            if (nsteps == 0) break;
            if (probing) {
                HIP_OK(c, hipEventSynchronize(c->probe_event));
                if (!(c->probe->time < t_end - 1e-14) || c->probe->err) break;
            }
            HIP_OK(c, hipMemcpyAsync(c->probe, c->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, c->stream));
            HIP_OK(c, hipEventRecord(c->probe_event, c->stream));
            probing = true;
        }
    }
    if (fetch_ctl(c)) return 1;
    if (check_device_error(c)) return 1;
    *time_inout = c->ctl_host->time;
    if (steps_done) *steps_done = c->ctl_host->step - step0;
    return 0;
}

int dimhip_compute_dt(dimhip_ctx* c, double cfl, double dt_max, double* dt_out)
{
    if (!c || !dt_out) return fail(c, "dimhip_compute_dt: null argument");
    if (cfl <= 0.0) return fail(c, "dim::compute_dt_cfl: cfl must be positive");
    if (dt_max <= 0.0) return fail(c, "dim::compute_dt_cfl: dt_max must be positive");
    HIP_OK(c, hipSetDevice(c->desc.device));
    reset_dtmin_kernel<<<1, kDtSlots, 0, c->stream>>>(c->ctl, 0);
    c->ops->cfl(c, state_ptr(c, c->cur), c->desc.time_update == DIMHIP_UNSPLIT, cfl, c->pstore);
    reduce_dtmin_kernel<<<1, kDtSlots, 0, c->stream>>>(c->ctl);
    HIP_OK(c, hipGetLastError());
    if (fetch_ctl(c)) return 1;
    const double m = __builtin_bit_cast(double, c->ctl_host->dtmin[c->ctl_host->slot]);
    *dt_out = (m < dt_max) ? m : dt_max;
    c->dt_ready = true;
    c->dt_ready_cfl = cfl;
    return 0;
}

int dimhip_sweep(dimhip_ctx* c, int dir, double dt)
{
    if (!c) return fail(c, "dimhip_sweep: null context");
    if (dir < 0 || dir >= c->desc.dim) return fail(c, "dimhip_sweep: direction out of range");
    HIP_OK(c, hipSetDevice(c->desc.device));
    set_active_kernel<<<1, 1, 0, c->stream>>>(c->ctl, dt);
    c->ops->sweep(c, dir, 0, false, state_ptr(c, c->cur), state_ptr(c, c->cur ^ 1), nullptr, nullptr);
    HIP_OK(c, hipGetLastError());
    c->cur ^= 1;
    c->dt_ready = false;
    if (fetch_ctl(c)) return 1;
    return check_device_error(c);
}

int dimhip_apply_bc(dimhip_ctx* c)
{
    if (!c) return fail(c, "dimhip_apply_bc: null context");
    HIP_OK(c, hipSetDevice(c->desc.device));
    set_active_kernel<<<1, 1, 0, c->stream>>>(c->ctl, 0.0);
    c->ops->bc(c, state_ptr(c, c->cur), 0, nullptr);
    HIP_OK(c, hipGetLastError());
    c->dt_ready = false;
    return fetch_ctl(c);
}

int dimhip_conservation(dimhip_ctx* c, double* out)
{
    if (!c || !out) return fail(c, "dimhip_conservation: null argument");
    HIP_OK(c, hipSetDevice(c->desc.device));
    c->ops->totals(c, state_ptr(c, c->cur));
    HIP_OK(c, hipGetLastError());
    const int no = c->desc.nmat + 1 + c->desc.dim + 1;
    HIP_OK(c, hipMemcpyAsync(out, c->totals, sizeof(double) * no, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dimhip_upload_primitive(dimhip_ctx* c, const double* prim, size_t ncells)
{
    if (!c || !prim) return fail(c, "dimhip_upload_primitive: null argument");
    if ((long long)ncells != c->grid.ncells) return fail(c, "dim::advance_one_step: grid size mismatch");
    HIP_OK(c, hipSetDevice(c->desc.device));
    const int np = c->nv + 1;
    const long long chunk = std::min<long long>(c->grid.ncells, 1ll << 22);
    if (ensure_staging(c, chunk) || ensure_io(c, (size_t)chunk * np)) return 1;
    for (long long first = 0; first < c->grid.ncells; first += chunk) {
        const long long cnt = std::min(chunk, c->grid.ncells - first);
        HIP_OK(c, hipMemcpyAsync(c->io_buf, prim + first * np, sizeof(double) * (size_t)(cnt * np),
                                 hipMemcpyHostToDevice, c->stream));
        c->ops->prim_to_packed(c, c->io_buf, c->staging, cnt);
        c->ops->pack(c, true, c->staging, c->q[c->cur], first, cnt);
        HIP_OK(c, hipGetLastError());
        HIP_OK(c, hipStreamSynchronize(c->stream));
    }
    c->dt_ready = false;
    return 0;
}

int dimhip_initialise(dimhip_ctx* c, const dimhip_ic* ic)
{
    if (!c || !ic) return fail(c, "dimhip_initialise: null argument");
    const int D = c->desc.dim, M = c->desc.nmat;
    if (ic->kind < DIMHIP_IC_REGIONS || ic->kind > DIMHIP_IC_COATED_SHOCK_BUBBLE)
        return fail(c, "dim::initialise_dim_from_config: unsupported initial_condition");
    IcDesc d{};
    d.kind = ic->kind;
    for (int a = 0; a < 3; ++a) {
        d.domain_min[a] = ic->domain_min[a];
        d.domain_max[a] = ic->domain_max[a];
        d.dx[a] = a < D ? (ic->domain_max[a] - ic->domain_min[a]) / static_cast<double>(c->grid.N[a]) : 1.0;
        d.centre[a] = ic->bubble_center[a];
    }
    d.thickness = ic->interface_thickness;
    d.sharpness = ic->sharpness_alpha;
    const bool region_ic = ic->kind == DIMHIP_IC_REGIONS || ic->kind == DIMHIP_IC_PLANAR_REGIONS;
    if (region_ic) {
        if (ic->nregions <= 0) return fail(c, "dim::diffuse_region_primitive: no regions defined");
        if (ic->nregions > DIMHIP_MAX_REGIONS) return fail(c, "dimhip_initialise: too many regions");
        d.nregions = ic->nregions;
        for (int i = 0; i < ic->nregions; ++i) {
            const dimhip_ic_region& r = ic->regions[i];
            if (r.material < 0 || r.material >= M)
                return fail(c, "dim::material_slot_from_id: material id not found");
            IcRegion& o = d.regions[i];
            for (int a = 0; a < 3; ++a) {
                o.lower[a] = r.lower[a];
                o.upper[a] = r.upper[a];
                o.vel[a] = r.vel[a];
            }
            o.rho = r.rho;
            o.p = r.p;
            o.slot = r.material;
        }
    }
    if (ic->kind == DIMHIP_IC_PLANAR_REGIONS) {
        // normalised_planar_normal (dim_initial_conditions.hpp:33-55)
        double n2 = 0.0;
        for (int a = 0; a < D; ++a) n2 += ic->planar_normal[a] * ic->planar_normal[a];
        if (n2 <= 0.0) return fail(c, "dim::normalised_planar_normal: zero planar normal");
        const double inv = 1.0 / std::sqrt(n2);
        for (int a = 0; a < D; ++a) d.normal[a] = ic->planar_normal[a] * inv;
    }
    if (!region_ic) {
        if (ic->shock_axis < 0 || ic->shock_axis >= D) return fail(c, "Invalid shock axis");
        if (!(ic->bubble_radius > 0.0)) return fail(c, "Invalid bubble radius");
        const dimhip_ic_state* src[4] = {&ic->left, &ic->right, &ic->bubble, &ic->film};
        IcPhase* dst[4] = {&d.left, &d.right, &d.bubble, &d.film};
        const int used = ic->kind == DIMHIP_IC_COATED_SHOCK_BUBBLE ? 4 : 3;
        for (int i = 0; i < 4; ++i) {
            if (i < used && (src[i]->material < 0 || src[i]->material >= M))
                return fail(c, "dim::material_slot_from_id: material id not found");
            dst[i]->rho = src[i]->rho;
            dst[i]->p = src[i]->p;
            dst[i]->slot = src[i]->material;
            for (int a = 0; a < 3; ++a) dst[i]->vel[a] = src[i]->vel[a];
        }
        d.shock_axis = ic->shock_axis;
        d.shock_position = ic->shock_position;
        d.radius = ic->bubble_radius;
        d.film_radius = ic->film_radius;
    }
    HIP_OK(c, hipSetDevice(c->desc.device));
    const int np = c->nv + 1;
    const long long chunk = std::min<long long>(c->grid.ncells, 1ll << 22);
    if (ensure_staging(c, chunk) || ensure_io(c, (size_t)chunk * np + 1)) return 1;
    int* derr = (int*)(c->io_buf + (size_t)chunk * np);
    HIP_OK(c, hipMemsetAsync(derr, 0, sizeof(int), c->stream));
    for (long long first = 0; first < c->grid.ncells; first += chunk) {
        const long long cnt = std::min(chunk, c->grid.ncells - first);
        c->ops->ic_prims(c, d, first, cnt, c->io_buf, derr);
        c->ops->prim_to_packed(c, c->io_buf, c->staging, cnt);
        c->ops->pack(c, true, c->staging, c->q[c->cur], first, cnt);
        HIP_OK(c, hipGetLastError());
    }
    int herr = 0;
    HIP_OK(c, hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    c->dt_ready = false;
    if (herr & IC_ERR_UNCOVERED) return fail(c, "dim::sharp_region_primitive: point not covered by any region");
    return 0;
}

int dimhip_output_rows(dimhip_ctx* c, double* rows, size_t ncells)
{
    if (!c || !rows) return fail(c, "dimhip_output_rows: null argument");
    if ((long long)ncells != c->grid.ncells) return fail(c, "dim::write_csv: grid size mismatch");
    HIP_OK(c, hipSetDevice(c->desc.device));
    const int nr = 3 + c->desc.dim + 3 * c->desc.nmat;
    const long long chunk = std::min<long long>(c->grid.ncells, 1ll << 22);
    if (ensure_staging(c, chunk) || ensure_io(c, (size_t)chunk * nr)) return 1;
    for (long long first = 0; first < c->grid.ncells; first += chunk) {
        const long long cnt = std::min(chunk, c->grid.ncells - first);
        c->ops->pack(c, false, c->staging, c->q[c->cur], first, cnt);
        c->ops->output_rows(c, c->staging, c->io_buf, cnt);
        HIP_OK(c, hipGetLastError());
        HIP_OK(c, hipMemcpyAsync(rows + first * nr, c->io_buf, sizeof(double) * (size_t)(cnt * nr),
                                 hipMemcpyDeviceToHost, c->stream));
        HIP_OK(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int dimhip_output_rows_noalloc(dimhip_ctx* c, double* rows, size_t ncells)
{
    if (!c || !rows) return fail(c, "dimhip_output_rows_noalloc: null argument");
    if ((long long)ncells != c->grid.ncells) return fail(c, "dim::write_csv: grid size mismatch");
    if (ncells > (size_t)kNoAllocMaxCells)
        return fail(c, "dimhip_output_rows_noalloc: " + std::to_string(ncells) + " cells exceed the " +
                           std::to_string(kNoAllocMaxCells) + " of its fixed scratch; use dimhip_output_rows");
    HIP_OK(c, hipSetDevice(c->desc.device));
    double* base = nullptr;
    if (scratch_address(c->desc.device, &base)) return fail(c, "dimhip_output_rows_noalloc: hipGetSymbolAddress failed");
    const int nr = 3 + c->desc.dim + 3 * c->desc.nmat;
    const long long cnt = c->grid.ncells;
    // the packed FP64 cells, then the rows, both inside the scratch array
    double* staging = base;
    double* out = base + (size_t)cnt * c->nv;
    c->ops->pack(c, false, staging, c->q[c->cur], 0, cnt);
    c->ops->output_rows(c, staging, out, cnt);
    HIP_OK(c, hipGetLastError());
    HIP_OK(c, hipMemcpyAsync(rows, out, sizeof(double) * (size_t)(cnt * nr), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dimhip_boundary_flux(dimhip_ctx* c, double* out)
{
    if (!c || !out) return fail(c, "dimhip_boundary_flux: null argument");
    HIP_OK(c, hipSetDevice(c->desc.device));
    c->ops->boundary_flux(c, state_ptr(c, c->cur), nullptr);
    HIP_OK(c, hipGetLastError());
    const int no = c->desc.nmat + 1 + c->desc.dim + 1;
    HIP_OK(c, hipMemcpyAsync(out, c->flux_out, sizeof(double) * no, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dimhip_track_boundary_flux(dimhip_ctx* c, int enable)
{
    if (!c) return fail(c, "dimhip_track_boundary_flux: null context");
    HIP_OK(c, hipSetDevice(c->desc.device));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    drop_graph(c);
    HIP_OK(c, hipMemsetAsync(c->flux_integral, 0, sizeof(double) * 16, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    c->track_flux = enable != 0;
    return 0;
}

int dimhip_boundary_flux_integral(dimhip_ctx* c, double* out)
{
    if (!c || !out) return fail(c, "dimhip_boundary_flux_integral: null argument");
    HIP_OK(c, hipSetDevice(c->desc.device));
    const int no = c->desc.nmat + 1 + c->desc.dim + 1;
    HIP_OK(c, hipMemcpyAsync(out, c->flux_integral, sizeof(double) * no, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dimhip_synchronize(dimhip_ctx* c)
{
    if (!c) return fail(c, "dimhip_synchronize: null context");
    HIP_OK(c, hipStreamSynchronize(c->stream));
    ev_collect(c);
    return 0;
}

int dimhip_set_timing(dimhip_ctx* c, int every)
{
    if (!c) return fail(c, "dimhip_set_timing: null context");
    HIP_OK(c, hipStreamSynchronize(c->stream));
    ev_collect(c);
    c->timing = every > 0 ? every : 0;
    c->timing_step = 0;
    c->timing_now = false;
    for (int k = 0; k < KC_COUNT; ++k) {
        c->kt_sum[k] = 0.0;
        c->kt_cnt[k] = 0;
    }
    return 0;
}

int dimhip_kernel_times(dimhip_ctx* c, double* mean_ms, long long* counts)
{
    if (!c || !mean_ms) return fail(c, "dimhip_kernel_times: null argument");
    HIP_OK(c, hipStreamSynchronize(c->stream));
    ev_collect(c);
    for (int k = 0; k < KC_SWEEP_X; ++k) {
        mean_ms[k] = c->kt_cnt[k] ? c->kt_sum[k] / (double)c->kt_cnt[k] : 0.0;
        if (counts) counts[k] = c->kt_cnt[k];
    }
    return 0;
}

int dimhip_sweep_times(dimhip_ctx* c, double* mean_ms, long long* counts)
{
    if (!c || !mean_ms) return fail(c, "dimhip_sweep_times: null argument");
    HIP_OK(c, hipStreamSynchronize(c->stream));
    ev_collect(c);
    for (int d = 0; d < 3; ++d) {
        const int k = KC_SWEEP_X + d;
        mean_ms[d] = c->kt_cnt[k] ? c->kt_sum[k] / (double)c->kt_cnt[k] : 0.0;
        if (counts) counts[d] = c->kt_cnt[k];
    }
    return 0;
}

int dimhip_kat(const dimhip_desc* desc, int which, int dir, int count, const double* in0, const double* in1,
               double* out0, double* out1)
{
    if (!desc || count < 0) return fail(nullptr, "dimhip_kat: bad argument");
    if (count == 0) return 0;
    std::string msg;
    dimhip_desc d = *desc;
    for (int a = 0; a < d.dim; ++a) {
        if (d.N[a] <= 0 || d.N[a] == 2) d.N[a] = 3;
        if (!(d.dx[a] > 0.0)) d.dx[a] = 1.0;
    }
    if (validate_desc(&d, msg)) return fail(nullptr, msg);
    const int M = d.nmat, D = d.dim;
    const int nv = 2 * M + D, np = 2 * M + D + 1, nf = M + D + 1;
    size_t n_in0 = 0, n_in1 = 0, n_out0 = 0, n_out1 = 0;
    switch (which) {
        case 0: n_in0 = (size_t)count * nv; n_out0 = (size_t)count * np; break;
        case 1: n_in0 = (size_t)count * np; n_out0 = (size_t)count * nv; break;
        case 2: n_in0 = (size_t)count * nv; n_out0 = (size_t)count * nv; break;
        case 3: n_in0 = n_in1 = (size_t)count * nv; n_out0 = (size_t)count * nf; n_out1 = count; break;
        case 4: n_in0 = (size_t)count * np; n_out0 = (size_t)count * M; break;
        case 5: n_in0 = (size_t)count * 2; n_out0 = n_out1 = count; break;
        case 6: n_in0 = (size_t)count * (2 + 2 * M); n_out0 = n_out1 = count; break;
        case 7: n_in0 = n_out0 = (size_t)count; break;
        default: return fail(nullptr, "dimhip_kat: unknown function");
    }
    if (which == 3 && (dir < 0 || dir >= D)) return fail(nullptr, "dimhip_kat: direction out of range");
    HIP_OK(nullptr, hipSetDevice(d.device));
    EosParams P{};
    fill_eos(P, d);
    double *di0 = nullptr, *di1 = nullptr, *do0 = nullptr, *do1 = nullptr;
    int* derr = nullptr;
    auto cleanup = [&] {
        hipFree(di0);
        hipFree(di1);
        hipFree(do0);
        hipFree(do1);
        hipFree(derr);
    };
    auto alloc = [&](double** p, size_t n) { return n ? hipMalloc(p, n * sizeof(double)) : hipSuccess; };
    if (alloc(&di0, n_in0) != hipSuccess || alloc(&di1, n_in1) != hipSuccess || alloc(&do0, n_out0) != hipSuccess ||
        alloc(&do1, n_out1) != hipSuccess || hipMalloc(&derr, 4 * sizeof(int)) != hipSuccess) {
        cleanup();
        return fail(nullptr, "dimhip_kat: device allocation failed");
    }
    // derr[0]: error bits; derr[2..3]: the smallest order key of a failing item
    const int herr0[4] = {0, 0, -1, -1};
    hipMemcpy(derr, herr0, sizeof(herr0), hipMemcpyHostToDevice);
    hipMemcpy(di0, in0, n_in0 * sizeof(double), hipMemcpyHostToDevice);
    if (n_in1) hipMemcpy(di1, in1, n_in1 * sizeof(double), hipMemcpyHostToDevice);
    int rc = 0;
    if (which == 5) {
        kat_ieee_kernel<<<(count + 127) / 128, 128>>>(count, di0, do0, do1);
    } else if (which == 7) {
        copy_probe_kernel<<<grid_blocks(count, 1024), 256>>>(di0, do0, count);
    } else {
        const Ops* ops = select_ops(D, M);
        rc = ops ? ops->kat(P, which, dir, count, di0, di1, do0, do1, derr, nullptr) : 1;
    }
    hipError_t e = hipDeviceSynchronize();
    int herr = 0;
    unsigned long long hkey = ~0ull;
    if (e == hipSuccess) {
        hipMemcpy(out0, do0, n_out0 * sizeof(double), hipMemcpyDeviceToHost);
        if (n_out1) hipMemcpy(out1, do1, n_out1 * sizeof(double), hipMemcpyDeviceToHost);
        int h[4];
        hipMemcpy(h, derr, sizeof(h), hipMemcpyDeviceToHost);
        herr = h[0];
        std::memcpy(&hkey, h + 2, sizeof(hkey));
    }
    cleanup();
    if (e != hipSuccess) return fail(nullptr, std::string("HIP error: ") + hipGetErrorString(e));
    if (rc) return fail(nullptr, "dimhip_kat: launch failed");
    if (herr) return fail(nullptr, device_error_text(herr, hkey) ? device_error_text(herr, hkey) : "dimhip_kat: device error");
    return 0;
}

}  // extern "C"

// ===========================================================================
// Batched form: many 2D two-material FP32-mixed instances of uneven size,
// one launch per stage over an (instance, tile) work table.
// ===========================================================================
namespace {

// The instance table is built once as BatchInst<Fp32Mixed>; when every
// instance is ideal-gas Kapila the launches take the Fp32Ideal kernels over
// the same table (its parameter block is layout-identical).
using BPol = Fp32Mixed;
using BInst = BatchInst<BPol>;
static_assert(sizeof(BatchInst<Fp32Ideal>) == sizeof(BInst), "layout-identical instance tables");

// tile-shape variants of the batched sweeps (autotuned)
constexpr int kBatchXL[2] = {128, 256};  // x sweep: positions per line; 256/L lines per block
constexpr int kBatchYW[3] = {8, 16, 4};  // y sweep: lines (x) per block
constexpr int kBatchYL[3] = {64, 32, 128};

}  // namespace

struct dimhip_batch {
    std::vector<dimhip_ctx*> ctx;
    hipStream_t stream = nullptr;
    std::vector<BInst> h_insts;
    BInst* d_insts = nullptr;
    Work* d_work[4] = {nullptr, nullptr, nullptr, nullptr};  // x, y, bc, cfl
    int nwork[4] = {0, 0, 0, 0};
    int xv = 0, yv = 0, schedule = 1, blocks_per_cu = 4;
    bool ideal = false;  // every instance ideal-gas Kapila: the Fp32Ideal kernels
    int n_cu = 256;
    int cur = 0;
    bool dt_ready = false;
    double dt_ready_cfl = 0.0;
    bool fusable = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    double graph_cfl = -1.0;
    int graph_cur = -1;
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_pending;
    size_t ev_next = 0;
    double kt_sum[KC_COUNT] = {};
    long long kt_cnt[KC_COUNT] = {};
    std::string err;
};

namespace {

int bfail(dimhip_batch* b, const std::string& msg)
{
    if (b) b->err = msg;
    g_global_err = msg;
    return 1;
}

#define BHIP_OK(b, expr)                                                                            \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return bfail((b), std::string("HIP error: ") + hipGetErrorString(e_) + \
                                                     " at " #expr);                                 \
    } while (0)

int bev_begin(dimhip_batch* b)
{
    if (!b->timing) return -1;
    if (b->ev_next + 2 > b->ev_pool.size())
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return -1;
            b->ev_pool.push_back(e);
        }
    const int i = (int)b->ev_next;
    b->ev_next += 2;
    hipEventRecord(b->ev_pool[i], b->stream);
    return i;
}

void bev_end(dimhip_batch* b, int i, int cls, int cls2 = -1)
{
    if (i < 0) return;
    hipEventRecord(b->ev_pool[i + 1], b->stream);
    b->ev_pending.push_back({cls, i});
    if (cls2 >= 0) b->ev_pending.push_back({cls2, i});
}

void bev_collect(dimhip_batch* b)
{
    for (auto& pr : b->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, b->ev_pool[pr.second], b->ev_pool[pr.second + 1]) == hipSuccess) {
            b->kt_sum[pr.first] += ms;
            b->kt_cnt[pr.first] += 1;
        }
    }
    b->ev_pending.clear();
    b->ev_next = 0;
}

void batch_drop_graph(dimhip_batch* b)
{
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->graph_cfl = -1.0;
    b->graph_cur = -1;
}

// work tables for the current variant choice
int batch_build_work(dimhip_batch* b)
{
    std::vector<Work> w[4];
    const int Lx = kBatchXL[b->xv], Wx = 256 / Lx, Tx = Lx - 6;
    const int Wy = kBatchYW[b->yv], Ly = kBatchYL[b->yv], Ty = Ly - 6;
    for (int i = 0; i < (int)b->ctx.size(); ++i) {
        const Grid& g = b->ctx[i]->grid;
        // a tile whose outputs [lo0, hi0] x [lo1, hi1] reach index 0, 1, N-2
        // or N-1 on some axis: the only tiles that hold ring cells or their
        // sources, where the sweeps' fused ring treatment acts
        const auto edge = [&](int lo0, int hi0, int lo1, int hi1) {
            return (lo0 <= 1 || hi0 >= g.N[0] - 2 || lo1 <= 1 || hi1 >= g.N[1] - 2) ? 1 : 0;
        };
        for (int by = 0; by < (g.N[1] + Wx - 1) / Wx; ++by)
            for (int bx = 0; bx < (g.N[0] + Tx - 1) / Tx; ++bx)
                w[0].push_back({i, bx, by, edge(bx * Tx, bx * Tx + Tx - 1, by * Wx, by * Wx + Wx - 1)});
        for (int by = 0; by < (g.N[1] + Ty - 1) / Ty; ++by)
            for (int bx = 0; bx < (g.N[0] + Wy - 1) / Wy; ++bx)
                w[1].push_back({i, bx, by, edge(bx * Wy, bx * Wy + Wy - 1, by * Ty, by * Ty + Ty - 1)});
        const long long ring = b->ctx[i]->bc.face_start[4];
        for (long long a = 0; a < ring; a += 1024) w[2].push_back({i, (int)a, (int)std::min<long long>(ring, a + 1024), 0});
        for (long long a = 0; a < g.ncells; a += 16384)
            w[3].push_back({i, (int)a, (int)std::min<long long>(g.ncells, a + 16384), 0});
    }
    for (int k = 0; k < 4; ++k) {
        hipFree(b->d_work[k]);
        b->d_work[k] = nullptr;
        b->nwork[k] = (int)w[k].size();
        BHIP_OK(b, hipMalloc(&b->d_work[k], sizeof(Work) * std::max<size_t>(1, w[k].size())));
        BHIP_OK(b, hipMemcpy(b->d_work[k], w[k].data(), sizeof(Work) * w[k].size(), hipMemcpyHostToDevice));
    }
    return 0;
}

int batch_sync_insts(dimhip_batch* b, double cfl)
{
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        dimhip_ctx* c = b->ctx[i];
        BInst& I = b->h_insts[i];
        I.g = c->grid;
        I.P = c->eosm;
        I.bc = c->bc;
        I.cfl = make_cfl_args(c, cfl, false);
        I.q[0] = c->q[0];
        I.q[1] = c->q[1];
        I.ctl = c->ctl;
    }
    BHIP_OK(b, hipMemcpyAsync(b->d_insts, b->h_insts.data(), sizeof(BInst) * b->h_insts.size(),
                              hipMemcpyHostToDevice, b->stream));
    return 0;
}

int sweep_blocks(const dimhip_batch* b, int nwork)
{
    return std::max(1, std::min(nwork, b->n_cu * b->blocks_per_cu));
}

template<typename Pol>
const BatchInst<Pol>* insts_as(const dimhip_batch* b)
{
    return reinterpret_cast<const BatchInst<Pol>*>(b->d_insts);
}

// RING: the instances' ring treatment runs inside the sweep (copy-kind
// boundaries); with FUSE the sweep also closes each instance's step
template<int DIR, int L, int W, bool FUSE, bool RING>
void launch_bsweep(dimhip_batch* b, int k)
{
    const dim3 blk = DIR == 0 ? dim3(L, W, 1) : dim3(W, L, 1);
    if (b->schedule == 2) {
        if (b->ideal)
            batch_tile_kernel<Fp32Ideal, DIR, L, W, FUSE, RING><<<b->nwork[k], blk, 0, b->stream>>>(
                insts_as<Fp32Ideal>(b), b->d_work[k], b->cur);
        else
            batch_tile_kernel<BPol, DIR, L, W, FUSE, RING><<<b->nwork[k], blk, 0, b->stream>>>(
                b->d_insts, b->d_work[k], b->cur);
        return;
    }
    const int g = sweep_blocks(b, b->nwork[k]);
    if (b->ideal)
        batch_sweep_kernel<Fp32Ideal, DIR, L, W, FUSE, RING><<<g, blk, 0, b->stream>>>(
            insts_as<Fp32Ideal>(b), b->d_work[k], b->nwork[k], b->schedule, b->cur);
    else
        batch_sweep_kernel<BPol, DIR, L, W, FUSE, RING><<<g, blk, 0, b->stream>>>(
            b->d_insts, b->d_work[k], b->nwork[k], b->schedule, b->cur);
}

template<typename Pol>
void launch_bcfl(dimhip_batch* b)
{
    batch_cfl_kernel<Pol><<<std::max(1, std::min(b->nwork[3], 4096)), 256, 0, b->stream>>>(
        insts_as<Pol>(b), b->d_work[3], b->nwork[3], b->cur);
}

template<typename Pol>
void launch_bbc(dimhip_batch* b, int end_step)
{
    const int bcb = std::max(1, std::min(b->nwork[2], 2048));
    batch_bc_kernel<Pol><<<bcb, 256, 0, b->stream>>>(insts_as<Pol>(b), b->d_work[2], b->nwork[2], b->cur, end_step);
}

// copy-kind boundaries on every instance (fuse): ring treatment inside the
// sweeps, the y sweep with the CFL epilogue and the step's close
void batch_sweep_x(dimhip_batch* b, bool fuse)
{
    switch (b->xv * 2 + (fuse ? 1 : 0)) {
        case 0: launch_bsweep<0, 128, 2, false, false>(b, 0); break;
        case 1: launch_bsweep<0, 128, 2, false, true>(b, 0); break;
        case 2: launch_bsweep<0, 256, 1, false, false>(b, 0); break;
        default: launch_bsweep<0, 256, 1, false, true>(b, 0); break;
    }
}

void batch_sweep_y(dimhip_batch* b, bool fuse)
{
    switch (b->yv * 2 + (fuse ? 1 : 0)) {
        case 0: launch_bsweep<1, 64, 8, false, false>(b, 1); break;
        case 1: launch_bsweep<1, 64, 8, true, true>(b, 1); break;
        case 2: launch_bsweep<1, 32, 16, false, false>(b, 1); break;
        case 3: launch_bsweep<1, 32, 16, true, true>(b, 1); break;
        case 4: launch_bsweep<1, 128, 4, false, false>(b, 1); break;
        default: launch_bsweep<1, 128, 4, true, true>(b, 1); break;
    }
}

void batch_enqueue_step(dimhip_batch* b, double cfl)
{
    const int n = (int)b->ctx.size();
    if (!b->dt_ready || b->dt_ready_cfl != cfl) {
        const int e = bev_begin(b);
        batch_begin_step_kernel<BPol><<<n, kDtSlots, 0, b->stream>>>(b->d_insts, n, 1);
        if (b->ideal) launch_bcfl<Fp32Ideal>(b);
        else launch_bcfl<BPol>(b);
        bev_end(b, e, KC_CFL);
    }
    batch_begin_step_kernel<BPol><<<n, kDtSlots, 0, b->stream>>>(b->d_insts, n, 0);
    // copy-kind boundaries: the ring treatment runs inside the sweeps and the
    // y sweep closes each instance's step; otherwise boundary passes follow
    int e = bev_begin(b);
    batch_sweep_x(b, b->fusable);
    bev_end(b, e, KC_SWEEP, KC_SWEEP_X);
    b->cur ^= 1;
    if (!b->fusable) {
        e = bev_begin(b);
        if (b->ideal) launch_bbc<Fp32Ideal>(b, 0);
        else launch_bbc<BPol>(b, 0);
        bev_end(b, e, KC_BC);
    }
    e = bev_begin(b);
    batch_sweep_y(b, b->fusable);
    bev_end(b, e, KC_SWEEP, KC_SWEEP_X + 1);
    b->cur ^= 1;
    if (!b->fusable) {
        e = bev_begin(b);
        if (b->ideal) launch_bbc<Fp32Ideal>(b, 1);
        else launch_bbc<BPol>(b, 1);
        bev_end(b, e, KC_BC);
    }
    b->dt_ready = b->fusable;
    b->dt_ready_cfl = cfl;
    for (auto* c : b->ctx) c->cur = b->cur;
}

}  // namespace

extern "C" {

const char* dimhip_batch_last_error(const dimhip_batch* b) { return b ? b->err.c_str() : g_global_err.c_str(); }

int dimhip_batch_create(const dimhip_desc* descs, int count, dimhip_batch** out)
{
    if (!descs || !out || count <= 0) return bfail(nullptr, "dimhip_batch_create: bad argument");
    // an early return releases the instances and buffers created so far
    std::unique_ptr<dimhip_batch, int (*)(dimhip_batch*)> b(new dimhip_batch, dimhip_batch_destroy);
    for (int i = 0; i < count; ++i) {
        const dimhip_desc& d = descs[i];
        if (d.dim != 2 || d.nmat != 2 || d.precision != DIMHIP_FP32_MIXED || d.time_update != DIMHIP_SPLIT)
            return bfail(nullptr, "dimhip_batch_create: the batched form runs 2D, two-material, FP32-mixed, split steps");
        if (d.device != descs[0].device) return bfail(nullptr, "dimhip_batch_create: instances must share a device");
        for (int a = 0; a < 2; ++a)
            if (d.bc_lo[a] == DIMHIP_BC_NONREFLECTIVE || d.bc_hi[a] == DIMHIP_BC_NONREFLECTIVE) b->fusable = false;
        dimhip_ctx* c = nullptr;
        if (dimhip_create(&d, &c))
            return bfail(nullptr, std::string("instance ") + std::to_string(i) + ": " + g_global_err);
        b->ctx.push_back(c);
    }
    b->ideal = true;
    for (auto* c : b->ctx) b->ideal = b->ideal && c->fp32_ideal;
    BHIP_OK(nullptr, hipSetDevice(descs[0].device));
    hipDeviceProp_t prop;
    BHIP_OK(nullptr, hipGetDeviceProperties(&prop, descs[0].device));
    b->n_cu = prop.multiProcessorCount;
    BHIP_OK(nullptr, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->h_insts.resize(count);
    BHIP_OK(nullptr, hipMalloc(&b->d_insts, sizeof(BInst) * count));
    if (batch_build_work(b.get()) || batch_sync_insts(b.get(), 0.8)) return 1;
    BHIP_OK(nullptr, hipStreamSynchronize(b->stream));
    *out = b.release();
    return 0;
}

int dimhip_batch_destroy(dimhip_batch* b)
{
    if (!b) return 0;
    if (b->stream) hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    for (auto e : b->ev_pool) hipEventDestroy(e);
    for (int k = 0; k < 4; ++k) hipFree(b->d_work[k]);
    hipFree(b->d_insts);
    for (auto* c : b->ctx) dimhip_destroy(c);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int dimhip_batch_upload(dimhip_batch* b, int i, const double* packed, size_t ncells)
{
    if (!b || i < 0 || i >= (int)b->ctx.size()) return bfail(b, "dimhip_batch_upload: bad instance");
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    if (dimhip_upload(b->ctx[i], packed, ncells)) return bfail(b, b->ctx[i]->err);
    if (b->ctx[i]->cur != b->cur) {
        // keep every instance on the same buffer parity
        copy_state(b->ctx[i], b->ctx[i]->q[b->ctx[i]->cur], b->ctx[i]->q[b->cur]);
        b->ctx[i]->cur = b->cur;
        BHIP_OK(b, hipStreamSynchronize(b->ctx[i]->stream));
    }
    b->dt_ready = false;
    return 0;
}

int dimhip_batch_download(dimhip_batch* b, int i, double* packed, size_t ncells)
{
    if (!b || i < 0 || i >= (int)b->ctx.size()) return bfail(b, "dimhip_batch_download: bad instance");
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    if (dimhip_download(b->ctx[i], packed, ncells)) return bfail(b, b->ctx[i]->err);
    return 0;
}

int dimhip_batch_configure(dimhip_batch* b, int x_variant, int y_variant, int schedule, int blocks_per_cu)
{
    if (!b) return bfail(b, "dimhip_batch_configure: null batch");
    if (x_variant < 0 || x_variant > 1 || y_variant < 0 || y_variant > 2 || schedule < 0 || schedule > 2 ||
        blocks_per_cu < 1 || blocks_per_cu > 64)
        return bfail(b, "dimhip_batch_configure: variant out of range");
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    b->xv = x_variant;
    b->yv = y_variant;
    b->schedule = schedule;
    b->blocks_per_cu = blocks_per_cu;
    batch_drop_graph(b);
    return batch_build_work(b);
}

int dimhip_batch_advance(dimhip_batch* b, int nsteps, double cfl, double t_end, int use_graph, double* times,
                         long long* steps)
{
    if (!b || !times) return bfail(b, "dimhip_batch_advance: null argument");
    if (cfl <= 0.0) return bfail(b, "dim::compute_dt_cfl: cfl must be positive");
    BHIP_OK(b, hipSetDevice(b->ctx[0]->desc.device));
    const int n = (int)b->ctx.size();
    std::vector<long long> step0(n);
    for (int i = 0; i < n; ++i) {
        set_mode_kernel<<<1, 1, 0, b->stream>>>(b->ctx[i]->ctl, 1, t_end, times[i], 1);
    }
    if (batch_sync_insts(b, cfl)) return 1;
    for (int i = 0; i < n; ++i)
        BHIP_OK(b, hipMemcpyAsync(b->ctx[i]->ctl_host, b->ctx[i]->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, b->stream));
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    for (int i = 0; i < n; ++i) step0[i] = b->ctx[i]->ctl_host->step;
    if (nsteps > 0 && (!b->dt_ready || b->dt_ready_cfl != cfl)) {
        batch_enqueue_step(b, cfl);
        --nsteps;
    }
    if (use_graph && !b->timing && b->dt_ready && nsteps >= 2) {
        if (!b->graph_exec || b->graph_cfl != cfl || b->graph_cur != b->cur) {
            batch_drop_graph(b);
            const int cur0 = b->cur;
            BHIP_OK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            batch_enqueue_step(b, cfl);
            batch_enqueue_step(b, cfl);
            BHIP_OK(b, hipStreamEndCapture(b->stream, &b->graph));
            BHIP_OK(b, hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
            b->graph_cfl = cfl;
            b->graph_cur = b->cur;
            if (b->cur != cur0) return bfail(b, "dimhip_batch_advance: graph does not return to its buffer");
        }
        while (nsteps >= 2) {
            BHIP_OK(b, hipGraphLaunch(b->graph_exec, b->stream));
            nsteps -= 2;
        }
    }
    while (nsteps > 0) {
        batch_enqueue_step(b, cfl);
        --nsteps;
    }
    BHIP_OK(b, hipGetLastError());
    for (int i = 0; i < n; ++i)
        BHIP_OK(b, hipMemcpyAsync(b->ctx[i]->ctl_host, b->ctx[i]->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, b->stream));
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    bev_collect(b);
    for (int i = 0; i < n; ++i) {
        if (check_device_error(b->ctx[i])) return bfail(b, std::string("instance ") + std::to_string(i) + ": " + b->ctx[i]->err);
        times[i] = b->ctx[i]->ctl_host->time;
        if (steps) steps[i] = b->ctx[i]->ctl_host->step - step0[i];
    }
    return 0;
}

int dimhip_batch_set_timing(dimhip_batch* b, int enable)
{
    if (!b) return bfail(b, "dimhip_batch_set_timing: null batch");
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    bev_collect(b);
    b->timing = enable != 0;
    for (int k = 0; k < KC_COUNT; ++k) {
        b->kt_sum[k] = 0.0;
        b->kt_cnt[k] = 0;
    }
    return 0;
}

int dimhip_batch_kernel_times(dimhip_batch* b, double* mean_ms, long long* counts)
{
    if (!b || !mean_ms) return bfail(b, "dimhip_batch_kernel_times: null argument");
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    bev_collect(b);
    for (int k = 0; k < KC_SWEEP_X; ++k) {
        mean_ms[k] = b->kt_cnt[k] ? b->kt_sum[k] / (double)b->kt_cnt[k] : 0.0;
        if (counts) counts[k] = b->kt_cnt[k];
    }
    return 0;
}

int dimhip_batch_sweep_times(dimhip_batch* b, double* mean_ms, long long* counts)
{
    if (!b || !mean_ms) return bfail(b, "dimhip_batch_sweep_times: null argument");
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    bev_collect(b);
    for (int d = 0; d < 3; ++d) {
        const int k = KC_SWEEP_X + d;
        mean_ms[d] = b->kt_cnt[k] ? b->kt_sum[k] / (double)b->kt_cnt[k] : 0.0;
        if (counts) counts[d] = b->kt_cnt[k];
    }
    return 0;
}

// Times every (x tile, y tile, schedule, blocks per CU) choice (schedule 2:
// one launch configuration, its block count fixed by the work table) over
// trial_steps steps from the current state, restores the state, keeps the
// fastest choice and describes the trials as JSON in report.
int dimhip_batch_autotune(dimhip_batch* b, double cfl, int trial_steps, char* report, size_t report_len)
{
    if (!b || trial_steps < 1) return bfail(b, "dimhip_batch_autotune: bad argument");
    const int n = (int)b->ctx.size();
    BHIP_OK(b, hipStreamSynchronize(b->stream));
    // back up every instance's state and control block
    std::vector<void*> backup(n, nullptr);
    std::vector<StepCtl> ctl_backup(n);
    std::vector<std::vector<char>> slot_backup(n, std::vector<char>(kCtlAllocBytes));
    for (int i = 0; i < n; ++i) {
        dimhip_ctx* c = b->ctx[i];
        const size_t bytes = (size_t)c->grid.pitch * c->cell_bytes;
        BHIP_OK(b, hipMalloc(&backup[i], bytes));
        BHIP_OK(b, hipMemcpy(backup[i], c->q[b->cur], bytes, hipMemcpyDeviceToDevice));
        BHIP_OK(b, hipMemcpy(slot_backup[i].data(), c->ctl, kCtlAllocBytes, hipMemcpyDeviceToHost));
        std::memcpy(&ctl_backup[i], slot_backup[i].data(), sizeof(StepCtl));
    }
    const int cur0 = b->cur;
    const bool ready0 = b->dt_ready;
    std::vector<double> times(n, 0.0);
    hipEvent_t e0, e1;
    BHIP_OK(b, hipEventCreate(&e0));
    BHIP_OK(b, hipEventCreate(&e1));
    std::string js = "{\"trials\": [";
    double best = 1e300;
    int bx = 0, by = 0, bs = 1, bb = 4;
    const int bpcs[3] = {2, 4, 8};
    bool first = true;
    for (int xv = 0; xv < 2; ++xv)
        for (int yv = 0; yv < 3; ++yv)
            for (int sc = 0; sc < 3; ++sc)
                for (int bi = 0; bi < 3; ++bi) {
                    // one tile per block: the block count is the work table's
                    if (sc == 2 && bi > 0) continue;
                    if (dimhip_batch_configure(b, xv, yv, sc, bpcs[bi])) return 1;
                    for (int i = 0; i < n; ++i) times[i] = ctl_backup[i].time;
                    // one untimed step settles dt, then the timed trial
                    if (dimhip_batch_advance(b, 1, cfl, 1e300, 0, times.data(), nullptr)) return 1;
                    hipEventRecord(e0, b->stream);
                    for (int s = 0; s < trial_steps; ++s) batch_enqueue_step(b, cfl);
                    hipEventRecord(e1, b->stream);
                    BHIP_OK(b, hipEventSynchronize(e1));
                    float ms = 0.f;
                    hipEventElapsedTime(&ms, e0, e1);
                    const double per = ms / trial_steps;
                    char buf[160];
                    snprintf(buf, sizeof buf, "%s{\"x\": %d, \"y\": %d, \"schedule\": %d, \"blocks_per_cu\": %d, \"ms\": %.5f}",
                             first ? "" : ", ", xv, yv, sc, bpcs[bi], per);
                    js += buf;
                    first = false;
                    if (per < best) {
                        best = per;
                        bx = xv;
                        by = yv;
                        bs = sc;
                        bb = bpcs[bi];
                    }
                    // restore
                    BHIP_OK(b, hipStreamSynchronize(b->stream));
                    b->cur = cur0;
                    for (int i = 0; i < n; ++i) {
                        dimhip_ctx* c = b->ctx[i];
                        BHIP_OK(b, hipMemcpy(c->q[cur0], backup[i], (size_t)c->grid.pitch * c->cell_bytes,
                                             hipMemcpyDeviceToDevice));
                        BHIP_OK(b, hipMemcpy(c->ctl, slot_backup[i].data(), kCtlAllocBytes, hipMemcpyHostToDevice));
                        c->cur = cur0;
                    }
                    b->dt_ready = ready0;
                }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    for (auto* p : backup) hipFree(p);
    if (dimhip_batch_configure(b, bx, by, bs, bb)) return 1;
    char tail[256];
    snprintf(tail, sizeof tail,
             "], \"best\": {\"x\": %d, \"y\": %d, \"schedule\": %d, \"blocks_per_cu\": %d, \"ms\": %.5f}, "
             "\"x_tiles\": [[128, 2], [256, 1]], \"y_tiles\": [[8, 64], [16, 32], [4, 128]]}",
             bx, by, bs, bb, best);
    js += tail;
    if (report && report_len) {
        strncpy(report, js.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return 0;
}

}  // extern "C"
