// ref_harness.cpp — C-ABI wrapper around the reference's own DIM functions.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference's headers where they lie (-I<reference root>) with the
// reference's own flags (g++ -std=c++17 -O2 -fopenmp, README.md:82-88) into
// oracle/_ref/libdimref.so.  It is used (a) to pin the C restatement in
// dim_oracle.c bit for bit, (b) to write the golden fixtures under
// tests/golden/, and (c) as the timed CPU baseline (the reference's own
// OpenMP path).  The entry points mirror dim_oracle.h with a ref_ prefix and
// take the same packed layouts.
#include <omp.h>

#include <array>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "src/app/material/material_builder.hpp"
#include "src/dim/solver/advance_step.hpp"
#include "src/io/config_loader.hpp"
#include "src/setup/dim_initial_conditions.hpp"

namespace {

std::string g_err;

struct RefParams {
    int dim;
    int nmat;
    int kind[4];
    double gamma[4];
    double pinf[4];
    int lambda_model;  // 0 kapila, 1 allaire, 2 equal_velocity
    double alpha_source_floor;
    double covolume[4];  // noble-Abel b
};

dim::EOSParams make_params(const RefParams* p)
{
    dim::EOSParams out;
    out.material.resize(p->nmat);
    for (int k = 0; k < p->nmat; ++k) {
        out.material[k].kind = (p->kind[k] == 2)   ? EOSKind::noble_abel
                               : (p->kind[k] == 1) ? EOSKind::stiffened_gas
                                                   : EOSKind::ideal_gas;
        out.material[k].gamma = p->gamma[k];
        out.material[k].p_inf = p->pinf[k];
        out.material[k].covolume = p->covolume[k];
    }
    return out;
}

const char* lambda_name(const RefParams* p)
{
    switch (p->lambda_model) {
        case 1: return "allaire";
        case 2: return "equal_velocity";
        default: return "kapila";
    }
}

const char* bc_name(int code)
{
    switch (code) {
        case 1: return "reflective";
        case 2: return "nonreflective";
        case 3: return "reflection";
        default: return "transmissive";
    }
}

template<int DIM>
dim::State<DIM> unpack(const RefParams* p, const double* src)
{
    dim::State<DIM> U = dim::make_state<DIM>(p->nmat);
    int o = 0;
    for (int k = 0; k < p->nmat; ++k) U.partial_mass[k] = src[o++];
    for (int d = 0; d < DIM; ++d) U.mom[d] = src[o++];
    U.E = src[o++];
    for (int k = 0; k < p->nmat - 1; ++k) U.alpha[k] = src[o++];
    return U;
}

template<int DIM>
void pack(const RefParams* p, const dim::State<DIM>& U, double* dst)
{
    int o = 0;
    for (int k = 0; k < p->nmat; ++k) dst[o++] = U.partial_mass[k];
    for (int d = 0; d < DIM; ++d) dst[o++] = U.mom[d];
    dst[o++] = U.E;
    for (int k = 0; k < p->nmat - 1; ++k) dst[o++] = U.alpha[k];
}

template<int DIM>
dim::Primitive<DIM> unpack_prim(const RefParams* p, const double* src)
{
    dim::Primitive<DIM> P;
    P.rho.assign(p->nmat, 0.0);
    P.alpha.assign(p->nmat, 0.0);
    int o = 0;
    for (int k = 0; k < p->nmat; ++k) P.rho[k] = src[o++];
    for (int k = 0; k < p->nmat; ++k) P.alpha[k] = src[o++];
    for (int d = 0; d < DIM; ++d) P.vel[d] = src[o++];
    P.p = src[o++];
    return P;
}

template<int DIM>
void pack_prim(const RefParams* p, const dim::Primitive<DIM>& P, double* dst)
{
    int o = 0;
    for (int k = 0; k < p->nmat; ++k) dst[o++] = P.rho[k];
    for (int k = 0; k < p->nmat; ++k) dst[o++] = P.alpha[k];
    for (int d = 0; d < DIM; ++d) dst[o++] = P.vel[d];
    dst[o++] = P.p;
}

int nv_of(const RefParams* p) { return 2 * p->nmat + p->dim; }

template<int DIM>
Config<DIM> make_config(const RefParams* p, const int* bc_lo, const int* bc_hi, int unsplit)
{
    Config<DIM> cfg;
    for (int d = 0; d < DIM; ++d) {
        cfg.bc_lo[d] = bc_name(bc_lo[d]);
        cfg.bc_hi[d] = bc_name(bc_hi[d]);
    }
    cfg.time_update = unsplit ? "unsplit" : "split";
    cfg.dim_lambda_model = lambda_name(p);
    cfg.dim_alpha_source_floor = p->alpha_source_floor;
    return cfg;
}

template<int DIM>
std::vector<dim::State<DIM>> unpack_grid(const RefParams* p, long total, const double* U)
{
    std::vector<dim::State<DIM>> out(total);
    const int nv = nv_of(p);
    for (long i = 0; i < total; ++i) out[i] = unpack<DIM>(p, U + (size_t)i * nv);
    return out;
}

template<int DIM>
void pack_grid(const RefParams* p, const std::vector<dim::State<DIM>>& in, double* U)
{
    const int nv = nv_of(p);
    for (size_t i = 0; i < in.size(); ++i) pack<DIM>(p, in[i], U + i * nv);
}

template<typename F>
int guarded(F&& f)
{
    try {
        f();
        return 0;
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

#define DISPATCH_DIM(p, FN, ...)                                              \
    ((p)->dim == 1 ? FN<1>(__VA_ARGS__)                                      \
                   : ((p)->dim == 2 ? FN<2>(__VA_ARGS__) : FN<3>(__VA_ARGS__)))

template<int DIM>
int cons_to_prim_t(const RefParams* p, int count, const double* U, double* P)
{
    return guarded([&] {
        const auto params = make_params(p);
        const int nv = nv_of(p), np = 2 * p->nmat + DIM + 1;
        for (int i = 0; i < count; ++i)
            pack_prim<DIM>(p, dim::cons_to_prim<DIM>(unpack<DIM>(p, U + (size_t)i * nv), params),
                           P + (size_t)i * np);
    });
}

template<int DIM>
int prim_to_cons_t(const RefParams* p, int count, const double* P, double* U)
{
    return guarded([&] {
        const auto params = make_params(p);
        const int nv = nv_of(p), np = 2 * p->nmat + DIM + 1;
        for (int i = 0; i < count; ++i)
            pack<DIM>(p, dim::prim_to_cons<DIM>(unpack_prim<DIM>(p, P + (size_t)i * np), params),
                      U + (size_t)i * nv);
    });
}

template<int DIM>
int repair_t(const RefParams* p, int count, double* U)
{
    return guarded([&] {
        const auto params = make_params(p);
        const int nv = nv_of(p);
        for (int i = 0; i < count; ++i) {
            auto s = unpack<DIM>(p, U + (size_t)i * nv);
            dim::repair_state<DIM>(s, params);
            pack<DIM>(p, s, U + (size_t)i * nv);
        }
    });
}

template<int DIM>
int alpha_rhs_t(const RefParams* p, int count, const double* P, double* rhs)
{
    return guarded([&] {
        const auto params = make_params(p);
        const int np = 2 * p->nmat + DIM + 1;
        for (int i = 0; i < count; ++i) {
            const auto r = dim::alpha_rhs_coefficients<DIM>(
                unpack_prim<DIM>(p, P + (size_t)i * np), params, 1e-12, p->alpha_source_floor,
                lambda_name(p));
            for (int k = 0; k < p->nmat; ++k) rhs[(size_t)i * p->nmat + k] = r[k];
        }
    });
}

template<int DIM>
int hllc_t(const RefParams* p, int dir, int count, const double* UL, const double* UR, double* F,
           double* fv)
{
    return guarded([&] {
        const auto params = make_params(p);
        const int nv = nv_of(p), nf = p->nmat + DIM + 1;
        std::array<double, DIM> normal{};
        normal[dir] = 1.0;
        for (int i = 0; i < count; ++i) {
            const auto r = dim::hllc_flux_normal<DIM>(unpack<DIM>(p, UL + (size_t)i * nv),
                                                      unpack<DIM>(p, UR + (size_t)i * nv), params,
                                                      normal, lambda_name(p));
            int o = 0;
            for (int k = 0; k < p->nmat; ++k) F[(size_t)i * nf + o++] = r.flux.partial_mass[k];
            for (int d = 0; d < DIM; ++d) F[(size_t)i * nf + o++] = r.flux.mom[d];
            F[(size_t)i * nf + o++] = r.flux.E;
            fv[i] = r.face_velocity;
        }
    });
}

template<int DIM>
int reconstruct_t(const RefParams* p, int dir, int n, const double* U, double dt, double dx,
                  double* UL, double* UR)
{
    return guarded([&] {
        const auto params = make_params(p);
        auto line = unpack_grid<DIM>(p, n, U);
        std::vector<dim::State<DIM>> l, r;
        dim::reconstruct_line_interfaces<DIM>(line, params, dir, dt, dx, p->alpha_source_floor,
                                              lambda_name(p), l, r);
        pack_grid<DIM>(p, l, UL);
        pack_grid<DIM>(p, r, UR);
    });
}

template<int DIM>
int solve_line_t(const RefParams* p, int dir, int n, const double* Uin, double* Uout, double dx,
                 double dt)
{
    return guarded([&] {
        const auto params = make_params(p);
        auto line = unpack_grid<DIM>(p, n, Uin);
        std::vector<dim::State<DIM>> out;
        dim::solve_line<DIM>(line, out, dx, dt, params, dir, p->alpha_source_floor,
                             lambda_name(p));
        pack_grid<DIM>(p, out, Uout);
    });
}

template<int DIM>
long total_of(const int* N)
{
    long t = 1;
    for (int d = 0; d < DIM; ++d) t *= N[d];
    return t;
}

template<int DIM>
int dt_cfl_t(const RefParams* p, const int* N, const double* dx, const double* U, double cfl,
             double dt_max, int unsplit, double* dt)
{
    return guarded([&] {
        const auto params = make_params(p);
        const auto grid = unpack_grid<DIM>(p, total_of<DIM>(N), U);
        std::array<double, DIM> h{};
        for (int d = 0; d < DIM; ++d) h[d] = dx[d];
        *dt = unsplit ? dim::compute_dt_cfl_unsplit<DIM>(grid, h, params, lambda_name(p), cfl, dt_max)
                      : dim::compute_dt_cfl<DIM>(grid, h, params, lambda_name(p), cfl, dt_max);
    });
}

template<int DIM>
int apply_bc_t(const RefParams* p, const int* N, const int* bc_lo, const int* bc_hi, double* U)
{
    return guarded([&] {
        const auto params = make_params(p);
        auto grid = unpack_grid<DIM>(p, total_of<DIM>(N), U);
        std::array<int, DIM> n{};
        for (int d = 0; d < DIM; ++d) n[d] = N[d];
        const auto cfg = make_config<DIM>(p, bc_lo, bc_hi, 0);
        dim::apply_boundary_conditions<DIM>(grid, n, cfg, params);
        pack_grid<DIM>(p, grid, U);
    });
}

template<int DIM>
int advance_t(const RefParams* p, const int* N, const double* dx, const int* bc_lo,
              const int* bc_hi, int unsplit, double cfl, double dt_max, int nsteps, int threads,
              double* U, double* dt_out, double* seconds)
{
    return guarded([&] {
        const auto params = make_params(p);
        auto grid = unpack_grid<DIM>(p, total_of<DIM>(N), U);
        std::array<int, DIM> n{};
        std::array<double, DIM> h{};
        for (int d = 0; d < DIM; ++d) {
            n[d] = N[d];
            h[d] = dx[d];
        }
        const auto cfg = make_config<DIM>(p, bc_lo, bc_hi, unsplit);
        if (threads > 0) omp_set_num_threads(threads);
        // the reference benchmark environment: OMP_SCHEDULE=guided,1 (tools/quant/runner.hpp:323-333)
        omp_set_schedule(omp_sched_guided, 1);
        double wall = 0.0;
        for (int s = 0; s < nsteps; ++s) {
            // the timed region matches run_dim.hpp:94-110: advance_one_step plus the state copy
            const auto t0 = std::chrono::steady_clock::now();
            const auto result = dim::advance_one_step<DIM>(grid, n, h, params, cfg, cfl, dt_max);
            grid = result.U_new;
            const auto t1 = std::chrono::steady_clock::now();
            wall += std::chrono::duration<double>(t1 - t0).count();
            dt_out[s] = result.dt;
        }
        if (seconds) *seconds = wall;
        pack_grid<DIM>(p, grid, U);
    });
}

// The reference's initial state of a case file: load_config,
// build_dim_material_params and initialise_dim_from_config for resolution
// N_list[n_index], packed.  U == nullptr only reports the cell count and nmat.
template<int DIM>
int initial_state_t(const char* path, int n_index, double* U, long long* ncells, int* nmat)
{
    try {
        const Config<DIM> cfg = load_config<DIM>(path);
        if (n_index < 0 || n_index >= (int)cfg.N_list.size()) {
            g_err = "ref_initial_state: resolution index out of range";
            return 1;
        }
        const auto N = cfg.N_list[n_index];
        const dim::EOSParams params = build_dim_material_params(cfg.materials);
        long long total = 1;
        for (int d = 0; d < DIM; ++d) total *= N[d];
        *ncells = total;
        *nmat = params.nmat();
        if (!U) return 0;
        std::vector<dim::State<DIM>> grid;
        dim::initialise_dim_from_config<DIM>(grid, cfg, N, params);
        RefParams rp{};
        rp.dim = DIM;
        rp.nmat = params.nmat();
        for (long long i = 0; i < total; ++i) pack<DIM>(&rp, grid[i], U + i * nv_of(&rp));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

}  // namespace

extern "C" {

int ref_initial_state(const char* path, int dim, int n_index, double* U, long long* ncells, int* nmat)
{
    if (dim == 1) return initial_state_t<1>(path, n_index, U, ncells, nmat);
    if (dim == 2) return initial_state_t<2>(path, n_index, U, ncells, nmat);
    if (dim == 3) return initial_state_t<3>(path, n_index, U, ncells, nmat);
    g_err = "ref_initial_state: dim must be 1, 2 or 3";
    return 1;
}

const char* ref_last_error() { return g_err.c_str(); }

int ref_max_threads() { return omp_get_max_threads(); }

int ref_cons_to_prim(const RefParams* p, int count, const double* U, double* P)
{
    return DISPATCH_DIM(p, cons_to_prim_t, p, count, U, P);
}

int ref_prim_to_cons(const RefParams* p, int count, const double* P, double* U)
{
    return DISPATCH_DIM(p, prim_to_cons_t, p, count, P, U);
}

int ref_repair_state(const RefParams* p, int count, double* U)
{
    return DISPATCH_DIM(p, repair_t, p, count, U);
}

int ref_alpha_rhs(const RefParams* p, int count, const double* P, double* rhs)
{
    return DISPATCH_DIM(p, alpha_rhs_t, p, count, P, rhs);
}

int ref_pressure_from_internal_energy(const RefParams* p, int count, const double* rho_total,
                                      const double* e, const double* alpha, const double* rho,
                                      double* out)
{
    return guarded([&] {
        const auto params = make_params(p);
        for (int i = 0; i < count; ++i) {
            std::vector<double> a(alpha + (size_t)i * p->nmat, alpha + (size_t)(i + 1) * p->nmat);
            std::vector<double> r(rho + (size_t)i * p->nmat, rho + (size_t)(i + 1) * p->nmat);
            out[i] = dim::MixtureEOS::pressure_from_internal_energy(rho_total[i], e[i], a, r, params);
        }
    });
}

int ref_mixture_sound_speed(const RefParams* p, int count, const double* rho_total,
                            const double* pr, const double* alpha, const double* rho, double* out)
{
    return guarded([&] {
        const auto params = make_params(p);
        for (int i = 0; i < count; ++i) {
            std::vector<double> a(alpha + (size_t)i * p->nmat, alpha + (size_t)(i + 1) * p->nmat);
            std::vector<double> r(rho + (size_t)i * p->nmat, rho + (size_t)(i + 1) * p->nmat);
            out[i] = dim::MixtureEOS::generic_mixture_sound_speed(rho_total[i], pr[i], a, r, params,
                                                                  lambda_name(p));
        }
    });
}

int ref_hllc(const RefParams* p, int dir, int count, const double* UL, const double* UR,
             double* F, double* fv)
{
    return DISPATCH_DIM(p, hllc_t, p, dir, count, UL, UR, F, fv);
}

int ref_reconstruct_line(const RefParams* p, int dir, int n, const double* U, double dt,
                         double dx, double* UL, double* UR)
{
    return DISPATCH_DIM(p, reconstruct_t, p, dir, n, U, dt, dx, UL, UR);
}

int ref_solve_line(const RefParams* p, int dir, int n, const double* Uin, double* Uout,
                   double dx, double dt)
{
    return DISPATCH_DIM(p, solve_line_t, p, dir, n, Uin, Uout, dx, dt);
}

int ref_dt_cfl(const RefParams* p, const int* N, const double* dx, const double* U, double cfl,
               double dt_max, int unsplit, double* dt)
{
    return DISPATCH_DIM(p, dt_cfl_t, p, N, dx, U, cfl, dt_max, unsplit, dt);
}

int ref_apply_bc(const RefParams* p, const int* N, const int* bc_lo, const int* bc_hi, double* U)
{
    return DISPATCH_DIM(p, apply_bc_t, p, N, bc_lo, bc_hi, U);
}

// nsteps calls of dim::advance_one_step; dt_out has nsteps entries; seconds
// receives the wall time of the stepping loop only.
int ref_advance(const RefParams* p, const int* N, const double* dx, const int* bc_lo,
                const int* bc_hi, int unsplit, double cfl, double dt_max, int nsteps, int threads,
                double* U, double* dt_out, double* seconds)
{
    return DISPATCH_DIM(p, advance_t, p, N, dx, bc_lo, bc_hi, unsplit, cfl, dt_max, nsteps,
                        threads, U, dt_out, seconds);
}

}  // extern "C"
