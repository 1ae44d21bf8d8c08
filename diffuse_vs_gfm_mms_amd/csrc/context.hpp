// context.hpp — the library context (one grid, one stream, double-buffered
// SoA state, device step control) and the launch interface shared by the
// library's translation units.  The kernels of each arithmetic policy are
// instantiated in a translation unit of their own (ops_fp64.hip,
// ops_fp64na.hip, ops_fp32.hip) through the launcher templates declared here
// and defined in ops_impl.hpp; dimhip.hip holds the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/dimhip.h"
#include "ic_kernels.hpp"
#include "policy_fp32.hpp"
#include "policy_fp64.hpp"
#include "sweep.hpp"

namespace dimhip {

// sweep tiles come from the policy: Pol::kSweepL0 positions x Pol::kSweepW0
// lines for x, Pol::kSweepW1 lines (along x) x Pol::kSweepL1 positions for y/z

// KC_SWEEP covers every sweep launch; KC_SWEEP_X + d the launches along axis d
enum KernelClass { KC_SWEEP = 0, KC_BC = 1, KC_CFL = 2, KC_SWEEP_X = 3, KC_COUNT = 6 };

struct Ops;

}  // namespace dimhip

struct dimhip_ctx {
    dimhip_desc desc{};
    dimhip::Grid grid{};
    dimhip::EosParams eos{};
    dimhip::EosParamsNA eos_na{};  // the same constants, for kernels compiled with the noble-Abel law
    dimhip::EosMixed eosm{};
    dimhip::EosMixedIdeal eosmi{};  // eosm, for the ideal-gas Kapila kernels
    bool fp32_ideal = false;        // FP32-mixed with ideal gases only and the Kapila lambda model
    dimhip::BcArgs bc{};
    int nv = 0;
    size_t cell_bytes = 0;  // storage bytes per cell
    hipStream_t stream = nullptr;
    void* q[2] = {nullptr, nullptr};
    int cur = 0;
    dimhip::StepCtl* ctl = nullptr;
    dimhip::StepCtl* ctl_host = nullptr;  // pinned
    dimhip::StepCtl* probe = nullptr;     // pinned: step control read back while steps are in flight
    hipEvent_t probe_event = nullptr;
    double* staging = nullptr;
    // FP64 form, dim >= 2: the recovered pressure of every cell of the
    // current state, written by the CFL pass (the fused epilogue of a step's
    // last sweep plus the boundary copies, or the stand-alone CFL kernel) and
    // read by the next step's first sweep; valid whenever dt_ready is
    double* pstore = nullptr;
    long long staging_cells = 0;
    double* partial = nullptr;
    double* totals = nullptr;
    int partial_blocks = 0;
    // driver I/O: primitive rows in, CSV row values out (grown on demand)
    double* io_buf = nullptr;
    size_t io_doubles = 0;
    // boundary flux: per-block partials, last result, running integral
    long long flux_lines[4] = {0, 0, 0, 0};
    int flux_blocks = 0;
    double* flux_partial = nullptr;
    double* flux_out = nullptr;
    double* flux_integral = nullptr;
    bool track_flux = false;
    bool dt_ready = false;       // ctl->dtmin[slot] holds the CFL minimum of the current state
    // set by enqueue_step for its sweeps and ring passes: the sweeps give the
    // x-face ring cells (other indices interior) their state and the ring
    // pass leaves them out (3D, copy-kind boundaries, split steps, and a grid
    // whose x-face sources share a wave with their ring cells)
    bool xring_now = false;
    // set by enqueue_step for a 2D split step with copy-kind boundaries: the
    // x sweep writes the y-face ring rows and no ring pass follows it
    bool yring_now = false;
    double dt_ready_cfl = 0.0;   // the cfl it was computed with
    std::string err;
    const dimhip::Ops* ops = nullptr;
    // graph for dimhip_advance
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    double graph_cfl = -1.0;
    int graph_cur = -1;
    // event timing: 0 off; k > 0 records events around the kernels of one
    // step in k (timing_step counts the steps enqueued since it was set)
    int timing = 0;
    long long timing_step = 0;
    bool timing_now = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_pending;  // (class, index of start event)
    size_t ev_next = 0;
    double kt_sum[dimhip::KC_COUNT] = {};
    long long kt_cnt[dimhip::KC_COUNT] = {};
};

namespace dimhip {

// Precision- and shape-specific work of a context, one implementation per
// (dim, nmat) (dimhip.hip), dispatching on the context's precision form.
struct Ops {
    virtual ~Ops() = default;
    // pin / pout / pstore: the recovered-pressure plane of the FP64 form
    // (dimhip_ctx::pstore) or nullptr
    virtual void sweep(dimhip_ctx* c, int dir, int mode, bool fuse, const void* in, void* out, const double* pin,
                       double* pout) const = 0;
    virtual void bc(dimhip_ctx* c, void* q, int end_step, double* pstore) const = 0;
    virtual void cfl(dimhip_ctx* c, const void* q, bool unsplit, double cfl, double* pout) const = 0;
    virtual void totals(dimhip_ctx* c, const void* q) const = 0;
    virtual void pack(dimhip_ctx* c, bool in, double* staging, void* q, long long first, long long count) const = 0;
    virtual int kat(const EosParams& P, int which, int dir, int count, const double* in0, const double* in1,
                    double* out0, double* out1, int* err, hipStream_t s) const = 0;
    virtual void prim_to_packed(dimhip_ctx* c, const double* prim, double* packed, long long count) const = 0;
    virtual void output_rows(dimhip_ctx* c, const double* packed, double* rows, long long count) const = 0;
    virtual void boundary_flux(dimhip_ctx* c, const void* q, double* integral) const = 0;
    // the sweep tiles of the context's policy keep every x-face ring cell's
    // source in the ring cell's wave
    virtual bool xring_fits(const dimhip_ctx* c) const = 0;
    virtual void ic_prims(dimhip_ctx* c, const IcDesc& ic, long long first, long long count, double* prim,
                          int* err) const = 0;
};

// Tiles of policy Pol hold each x-face ring cell's source (its x-neighbour)
// as an own cell of the same wave: x tiles (positions along x, 64 to a wave)
// unless x = N0-1 is the first output of its tile or sits at the start of a
// wave; y/z tiles (kSweepW1 consecutive x lines per position, lanes of one
// wave) unless x = N0-1 starts a tile.
template<typename Pol>
inline bool xring_fits_tiles(const Grid& g, int dim)
{
    const int n0 = g.N[0];
    if (n0 < 3) return false;
    const int T0 = Pol::kSweepL0 - 6;
    const int gx = (n0 + T0 - 1) / T0;
    const int pr = (n0 - 1) - (gx - 1) * T0 + 3;  // tile position of x = N0-1
    if (pr < 4 || pr % 64 == 0) return false;
    if (dim >= 2 && (n0 - 1) % Pol::kSweepW1 == 0) return false;
    return true;
}

inline CflArgs make_cfl_args(const dimhip_ctx* c, double cfl, bool unsplit)
{
    CflArgs a{};
    for (int d = 0; d < 3; ++d) {
        a.dx[d] = c->grid.dx[d];
        a.cfl_dx[d] = cfl * c->grid.dx[d];
        a.inv_cfl_dx[d] = 1.0 / a.cfl_dx[d];
        a.inv_dx[d] = 1.0 / c->grid.dx[d];
    }
    a.cfl = cfl;
    a.inv_cfl = 1.0 / cfl;
    a.unsplit = unsplit ? 1 : 0;
    return a;
}

inline int grid_blocks(long long work, int per)
{
    long long b = (work + per - 1) / per;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (int)b;
}


// Launchers of one arithmetic policy's kernels for one (dim, nmat):
// explicitly instantiated in that policy's translation unit.
template<typename Pol, int D, int M>
struct PolOps {
    using Eos = typename Pol::Eos;
    static void sweep(dimhip_ctx* c, int dir, int mode, bool fuse, const void* in, void* out, const double* pin,
                      double* pout, const Eos& P);
    static void bc(dimhip_ctx* c, void* q, int end_step, double* pstore, const Eos& P);
    static void cfl(dimhip_ctx* c, const void* q, bool unsplit, double cfl, double* pout, const Eos& P);
    static void totals(dimhip_ctx* c, const void* q, const Eos& P);
    static void pack(dimhip_ctx* c, bool in, double* staging, void* q, long long first, long long count, const Eos& P);
    static void boundary_flux(dimhip_ctx* c, const void* q, double* integral, const Eos& PE);
};

// The FP64 device functions one by one (dimhip_kat), with either parameter block.
template<int D, int M, typename EP>
struct KatOps {
    static int kat(const EP& P, int which, int dir, int count, const double* in0, const double* in1, double* out0,
                   double* out1, int* err, hipStream_t s);
};

// Driver input and output rows, FP64 with every supported material law.
template<int D, int M>
struct IoOps {
    static void prim_to_packed(dimhip_ctx* c, const double* prim, double* packed, long long count);
    static void output_rows(dimhip_ctx* c, const double* packed, double* rows, long long count);
};

}  // namespace dimhip
