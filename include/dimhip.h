/*
 * dimhip.h — C ABI of the MI355X (gfx950) DIM library.
 *
 * The library replaces the reference's Allaire five-equation diffuse-interface
 * update behind the boundary the reference's time loop calls:
 *
 *   dim::advance_one_step<DIM>(U, N, dx, params, cfg, cfl, dt_max) -> StepResult
 *       src/dim/solver/advance_step.hpp:20-149, called from
 *       src/app/solver/run_dim.hpp:94-102
 *
 * and the functions below it (compute_dt_cfl, sweep_direction_dispatch,
 * apply_boundary_conditions).  State lives on the GPU between calls; the host
 * exchanges it in the reference's packed per-cell order
 *
 *   partial_mass[nmat], mom[dim], E, alpha[nmat-1]   (nv = 2*nmat + dim doubles)
 *
 * (the order of src/app/dim_mpi_main.cpp:255-273), cells x-fastest
 * (src/dim/solver/advance/geometry.hpp:25-55).
 *
 * Every entry point returns 0 on success and a nonzero status on failure;
 * dimhip_last_error() returns the message, worded like the reference's
 * std::runtime_error text where the reference has one.  A context is used by
 * one host thread at a time and owns one HIP stream.
 */
#ifndef DIMHIP_H
#define DIMHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DIMHIP_VERSION 4
#define DIMHIP_MAX_MAT 4

/* EoS kinds on the path (src/euler/eos_params.hpp:6-13; the config names
 * real_gas / noble_abel / noble_abel_gas select NOBLE_ABEL,
 * eos_params.hpp:26-28).  NOBLE_ABEL runs in the FP64 form only. */
enum { DIMHIP_EOS_IDEAL = 0, DIMHIP_EOS_STIFFENED = 1, DIMHIP_EOS_NOBLE_ABEL = 2 };

/* Boundary names (config_loader.hpp:369-377; boundary.hpp:108-318).
 * transmissive / zero_gradient / outflow map to TRANSMISSIVE; the literal
 * "reflective" flips the normal momentum; "reflection" is accepted by the
 * reference but copies without a flip (boundary.hpp:219-221). */
enum {
    DIMHIP_BC_TRANSMISSIVE = 0,
    DIMHIP_BC_REFLECTIVE = 1,
    DIMHIP_BC_NONREFLECTIVE = 2,
    DIMHIP_BC_REFLECTION = 3
};

/* cfg.dim_lambda_model (src/io/config.hpp:121-124) */
enum { DIMHIP_LAMBDA_KAPILA = 0, DIMHIP_LAMBDA_ALLAIRE = 1, DIMHIP_LAMBDA_EQUAL_VELOCITY = 2 };

/* cfg.time_update (src/io/config.hpp:106) */
enum { DIMHIP_SPLIT = 0, DIMHIP_UNSPLIT = 1 };

/* Arithmetic form.  FP64: the reference's arithmetic, reproduced bit for bit.
 * FP32_MIXED: FP32 state, reconstruction, fluxes and update; FP64 CFL
 * reduction, time, conservation sums and interface-cell pressure recovery. */
enum { DIMHIP_FP64 = 0, DIMHIP_FP32_MIXED = 1 };

typedef struct {
    int dim;                        /* 1..3 */
    int N[3];                       /* cells per axis (unused axes: 1) */
    double dx[3];                   /* cell size per axis */
    int nmat;                       /* 1..4 */
    int eos_kind[DIMHIP_MAX_MAT];   /* DIMHIP_EOS_* */
    double gamma[DIMHIP_MAX_MAT];
    double p_inf[DIMHIP_MAX_MAT];
    int bc_lo[3];                   /* DIMHIP_BC_* */
    int bc_hi[3];
    int lambda_model;               /* DIMHIP_LAMBDA_* */
    double alpha_source_floor;      /* cfg.dim_alpha_source_floor */
    int time_update;                /* DIMHIP_SPLIT / DIMHIP_UNSPLIT */
    int precision;                  /* DIMHIP_FP64 / DIMHIP_FP32_MIXED */
    int device;                     /* HIP device ordinal */
    double covolume[DIMHIP_MAX_MAT]; /* noble-Abel co-volume b (EOSParams::covolume,
                                        src/euler/eos_params.hpp:56-57); read for
                                        NOBLE_ABEL materials only */
} dimhip_desc;

/* Phase timings of one step, seconds (src/core/phase_timings.hpp:4-21). */
typedef struct {
    double cfl_seconds;
    double sweep_seconds;
    double boundary_seconds;
} dimhip_timings;

typedef struct dimhip_ctx dimhip_ctx;

int dimhip_version(void);

/* Create a context: validates the descriptor (dim::EOSParams::validate,
 * src/dim/eos_params.hpp:17-29) and allocates the device state. */
int dimhip_create(const dimhip_desc* desc, dimhip_ctx** out);
int dimhip_destroy(dimhip_ctx* ctx);

/* Host <-> device state in the packed per-cell order; ncells must equal
 * N[0]*N[1]*N[2] (advance_step.hpp:43-46). */
int dimhip_upload(dimhip_ctx* ctx, const double* packed, size_t ncells);
int dimhip_download(dimhip_ctx* ctx, double* packed, size_t ncells);

/* One dim::advance_one_step (advance_step.hpp:29-149): CFL dt (cfl.hpp:15-119),
 * directional sweeps x->y->z (or unsplit), ring boundary conditions after each
 * direction.  *dt_out receives the step's dt; timings may be NULL. */
int dimhip_step(dimhip_ctx* ctx, double cfl, double dt_max, double* dt_out,
                dimhip_timings* timings);

/* Up to nsteps iterations of run_dim_case's loop body (run_dim.hpp:88-111)
 * with dt_max = t_end - time, time kept on the device; steps stop once
 * time >= t_end - 1e-14.  No host synchronisation between steps; with
 * use_graph != 0 the step sequence is captured once in a hipGraph and
 * replayed.  *time_inout is the start time on entry and the end time on
 * return; *steps_done the number of steps taken. */
int dimhip_advance(dimhip_ctx* ctx, int nsteps, double cfl, double t_end, int use_graph,
                   double* time_inout, long long* steps_done);

/* compute_dt_cfl / compute_dt_cfl_unsplit on the current state (cfl.hpp:15-119). */
int dimhip_compute_dt(dimhip_ctx* ctx, double cfl, double dt_max, double* dt_out);

/* sweep_direction_dispatch (sweep_core.hpp:151-202) on the current state with
 * time step dt; the result replaces the state.  No boundary pass. */
int dimhip_sweep(dimhip_ctx* ctx, int dir, double dt);

/* apply_boundary_conditions (boundary.hpp:165-226) on the current state. */
int dimhip_apply_bc(dimhip_ctx* ctx);

/* FP64 conservation totals (conservation_report.hpp:188-227), fixed-order
 * tree sums: out = [mass_k * V (nmat), mass * V, mom_d * V (dim), E * V]. */
int dimhip_conservation(dimhip_ctx* ctx, double* out);

/* Physical flux through the domain boundary of the current state
 * (compute_dim_boundary_flux, conservation_report.hpp:253-305): for every
 * line normal to a face, +area * F(last cell) - area * F(first cell), summed
 * in FP64 in a fixed order on the device.  out = [flux mass_k (nmat), mass,
 * mom_d (dim), E]. */
int dimhip_boundary_flux(dimhip_ctx* ctx, double* out);

/* With enable != 0, every step taken by dimhip_step / dimhip_advance ends
 * with the report's accumulate_fluxes (run_dim.hpp:113-125,
 * conservation_report.hpp:506-520): integral += dt * boundary_flux(new
 * state), kept on the device.  Enabling or disabling resets the integral to
 * zero.  dimhip_boundary_flux_integral reads it (layout of
 * dimhip_boundary_flux). */
int dimhip_track_boundary_flux(dimhip_ctx* ctx, int enable);
int dimhip_boundary_flux_integral(dimhip_ctx* ctx, double* out);

/* Initial state from primitives: prim holds ncells rows of
 * rho[nmat], alpha[nmat], vel[dim], p (2*nmat + dim + 1 doubles); the device
 * stores prim_to_cons of each row (primitives.hpp:166-212), exactly as the
 * reference's initialisers do per cell (dim_initial_conditions.hpp:531-628). */
int dimhip_upload_primitive(dimhip_ctx* ctx, const double* prim, size_t ncells);

/* Initial conditions generated on the device (initialise_dim_from_config,
 * dim_initial_conditions.hpp:16-628): region boxes ("regions"), oblique slabs
 * ("planar_regions"), "shock_bubble" and "coated_shock_bubble", with the
 * reference's tanh blending and prim_to_cons.  Materials are given as slots
 * 0..nmat-1.  The arithmetic is the reference's except tanh, the device
 * library's, which can differ from the host C library's in the last bit:
 * diffuse cells may then differ from the reference's initial state at the
 * 1e-16 level (sharp interfaces are identical).  dimhip_upload_primitive
 * with host-built rows gives the reference's state bit for bit. */
#define DIMHIP_MAX_REGIONS 16
enum {
    DIMHIP_IC_REGIONS = 0,
    DIMHIP_IC_PLANAR_REGIONS = 1,
    DIMHIP_IC_SHOCK_BUBBLE = 2,
    DIMHIP_IC_COATED_SHOCK_BUBBLE = 3
};

typedef struct {
    double lower[3];
    double upper[3];
    double rho;
    double vel[3];
    double p;
    int material;
} dimhip_ic_region;

typedef struct {
    double rho;
    double vel[3];
    double p;
    int material;
} dimhip_ic_state;

typedef struct {
    int kind;                        /* DIMHIP_IC_* */
    double domain_min[3];
    double domain_max[3];
    double interface_thickness;
    double sharpness_alpha;          /* interface_sharpness_alpha */
    double planar_normal[3];         /* planar_regions: the un-normalised normal */
    int nregions;
    dimhip_ic_region regions[DIMHIP_MAX_REGIONS];
    int shock_axis;
    double shock_position;
    dimhip_ic_state left, right, bubble, film;
    double bubble_center[3];
    double bubble_radius;
    double film_radius;
} dimhip_ic;

int dimhip_initialise(dimhip_ctx* ctx, const dimhip_ic* ic);

/* The values of the solution CSV rows (dim_write_csv.hpp:78-121), computed on
 * the device in FP64 from the current state: per cell
 *   rho, vel[dim], p, e, alpha[nmat], rho_mat[nmat], mass[nmat]
 * (3 + dim + 3*nmat doubles) with P = cons_to_prim(U), rho = total_density(U),
 * e = max(safe_div(E - 0.5*rho*|vel|^2, rho), 0).  The cell coordinates are
 * left to the caller. */
int dimhip_output_rows(dimhip_ctx* ctx, double* rows, size_t ncells);

/* dimhip_output_rows without any device allocation, for callers that cannot
 * allocate (code that must not call hipMalloc: tight serving loops,
 * harnesses that forbid it).  No counterpart in the reference.  The packed
 * cells and rows are staged in a fixed 12 MiB device array per GPU, so the
 * grid may hold at most 65536 cells (larger grids fail with a message);
 * calls on one GPU share that array and must not overlap. */
int dimhip_output_rows_noalloc(dimhip_ctx* ctx, double* rows, size_t ncells);

/* Wait for the context's stream. */
int dimhip_synchronize(dimhip_ctx* ctx);

/* Mean device duration (ms) of each kernel class over the calls since the
 * last reset, measured with HIP events on the context's stream:
 * out[0] sweep kernels, out[1] boundary kernels, out[2] CFL kernels;
 * counts[i] = launches measured.  dimhip_set_timing(ctx, every) resets the
 * sums and, for every > 0, records events around the kernels of one step in
 * `every` (1: every step; sampling keeps the events' own cost out of the
 * other steps); every <= 0 turns timing off.  Graph replay is off while
 * timing is on. */
int dimhip_set_timing(dimhip_ctx* ctx, int every);
int dimhip_kernel_times(dimhip_ctx* ctx, double* mean_ms, long long* counts);

/* Mean device duration (ms) of the sweep launches along each axis over the
 * same measured calls: mean_ms[d], counts[d] for d = 0 (x), 1 (y), 2 (z). */
int dimhip_sweep_times(dimhip_ctx* ctx, double* mean_ms, long long* counts);

/* Known-answer entry point for the device functions on the path.  `which`:
 *   0 cons_to_prim   (primitives.hpp:111-162)  in0: count*nv cells   -> out0: count*(2nmat+dim+1)
 *   1 prim_to_cons   (primitives.hpp:166-212)  in0: count*(2nmat+dim+1) -> out0: count*nv
 *   2 repair_state   (primitives.hpp:326-457)  in0: count*nv -> out0: count*nv
 *   3 hllc_flux_normal (hllc.hpp:62-195) along `dir`; in0/in1: left/right cells
 *                    -> out0: count*(nmat+dim+1) fluxes, out1: count face velocities
 *   4 alpha_rhs_coefficients (primitives.hpp:216-322) in0: primitives -> out0: count*nmat
 *   5 IEEE check: out0[i] = in0[2i] / in0[2i+1], out1[i] = sqrt(in0[2i]) in FP64
 *   6 MixtureEOS::pressure_from_internal_energy (src/dim/eos.hpp:77-143) on
 *     in0: count*(2+2nmat) = [rho_total, e, alpha[nmat], rho[nmat]] ->
 *     out0 (and out1): the pressure
 *   7 device copy of count doubles in0 -> out0 (a streaming kernel of known
 *     size, used to calibrate HBM byte counters)
 * Host buffers; the call allocates and frees its own device memory. */
int dimhip_kat(const dimhip_desc* desc, int which, int dir, int count, const double* in0,
               const double* in1, double* out0, double* out1);

const char* dimhip_last_error(const dimhip_ctx* ctx);

/* ---- batched form ---------------------------------------------------------
 * Many independent 2D two-material FP32-mixed instances of uneven size
 * advanced together: each stage (x sweep, y sweep with the CFL reduction;
 * ring boundary passes after each sweep unless every instance has
 * copy-kind boundaries, whose ring treatment then runs inside the sweeps)
 * is one launch over a work table of (instance, tile) entries, one block
 * per entry or walked by persistent blocks; every instance keeps its own dt
 * and time on the device
 * (the reference runs such sweeps as separate processes, one config each,
 * tools/quant/runner.hpp:246-371).  Results equal running each instance
 * alone through dimhip_advance. */
typedef struct dimhip_batch dimhip_batch;

int dimhip_batch_create(const dimhip_desc* descs, int count, dimhip_batch** out);
int dimhip_batch_destroy(dimhip_batch* batch);
int dimhip_batch_upload(dimhip_batch* batch, int instance, const double* packed, size_t ncells);
int dimhip_batch_download(dimhip_batch* batch, int instance, double* packed, size_t ncells);

/* Tile shape and schedule: x_variant 0 = 128 x 2, 1 = 256 x 1 (positions x
 * lines); y_variant 0 = 8 x 64, 1 = 16 x 32, 2 = 4 x 128 (lines x positions);
 * schedule 0 = cyclic, 1 = contiguous chunks of the work table per
 * persistent block, 2 = one work entry per block (no persistent loop);
 * blocks_per_cu = persistent blocks launched per compute unit (schedules 0
 * and 1; schedule 2 launches one block per work entry). */
int dimhip_batch_configure(dimhip_batch* batch, int x_variant, int y_variant, int schedule, int blocks_per_cu);

/* Times every configuration over trial_steps steps from the current state,
 * restores the state, keeps the fastest and writes a JSON report. */
int dimhip_batch_autotune(dimhip_batch* batch, double cfl, int trial_steps, char* report, size_t report_len);

/* Up to nsteps steps of every instance (dt_max = t_end - time per instance);
 * times[i] in/out, steps[i] out (may be NULL). */
int dimhip_batch_advance(dimhip_batch* batch, int nsteps, double cfl, double t_end, int use_graph,
                         double* times, long long* steps);

int dimhip_batch_set_timing(dimhip_batch* batch, int enable);
int dimhip_batch_kernel_times(dimhip_batch* batch, double* mean_ms, long long* counts);
int dimhip_batch_sweep_times(dimhip_batch* batch, double* mean_ms, long long* counts);
const char* dimhip_batch_last_error(const dimhip_batch* batch);

#ifdef __cplusplus
}
#endif
#endif
