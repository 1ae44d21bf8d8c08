/*
 * dim_oracle.h — CPU restatement of the reference's Allaire five-equation
 * diffuse-interface (DIM) update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file and dim_oracle.c are the checker the
 * parity tests, the smoke check and the benchmark's CPU-baseline leg compare
 * the HIP library against.  Nothing in the product links or calls it.
 *
 * Cell layout ("packed", the reference's own per-cell order,
 * src/app/dim_mpi_main.cpp:255-273):
 *     partial_mass[nmat], mom[dim], E, alpha[nmat-1]      (nv = 2*nmat + dim)
 * Grid order is x-fastest row-major (src/dim/solver/advance/geometry.hpp:25-55).
 *
 * Parity status: every function below is pinned against the reference built
 * from its own sources (oracle/_ref/libdimref.so) on seeded random and
 * edge-case inputs, bit for bit (tests/test_oracle.py), and against the
 * committed golden fixtures (tests/golden/).  Closed-form / invariant checks
 * add independent pins for the EoS, the bisection, HLLC consistency and
 * uniform-state preservation.
 */
#ifndef DIM_ORACLE_H
#define DIM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_MAT 4
#define ORC_MAX_DIM 3

enum { ORC_EOS_IDEAL = 0, ORC_EOS_STIFFENED = 1, ORC_EOS_NOBLE_ABEL = 2 };
enum { ORC_LAMBDA_KAPILA = 0, ORC_LAMBDA_ALLAIRE = 1, ORC_LAMBDA_EQUAL_VELOCITY = 2 };
/* boundary names collapse to three behaviours (boundary.hpp:108-318):
 * copy (transmissive, zero_gradient, outflow, reflection), the literal
 * "reflective" (copy + normal-momentum flip), nonreflective. */
enum { ORC_BC_COPY = 0, ORC_BC_REFLECTIVE = 1, ORC_BC_NONREFLECTIVE = 2 };

typedef struct {
    int dim;
    int nmat;
    int kind[ORC_MAX_MAT];
    double gamma[ORC_MAX_MAT];
    double pinf[ORC_MAX_MAT];
    int lambda_model;
    double alpha_source_floor;
    double covolume[ORC_MAX_MAT];  /* noble-Abel b; read for ORC_EOS_NOBLE_ABEL only */
} orc_params;

int orc_nv(const orc_params* p);

/* Function-level entry points (each works on `count` independent items). */
int orc_cons_to_prim(const orc_params* p, int count, const double* U, double* P);
int orc_prim_to_cons(const orc_params* p, int count, const double* P, double* U);
int orc_repair_state(const orc_params* p, int count, double* U);
int orc_pressure_from_internal_energy(const orc_params* p, int count, const double* rho_total,
                                      const double* e, const double* alpha, const double* rho,
                                      double* out);
int orc_mixture_sound_speed(const orc_params* p, int count, const double* rho_total,
                            const double* pr, const double* alpha, const double* rho,
                            double* out);
int orc_alpha_rhs(const orc_params* p, int count, const double* P, double* rhs);
int orc_hllc(const orc_params* p, int dir, int count, const double* UL, const double* UR,
             double* F, double* fv);
int orc_reconstruct_line(const orc_params* p, int dir, int n, const double* U, double dt,
                         double dx, double* UL_face, double* UR_face);
int orc_solve_line(const orc_params* p, int dir, int n, const double* Uin, double* Uout,
                   double dx, double dt);

/* Whole-grid entry points. */
int orc_dt_cfl(const orc_params* p, const int* N, const double* dx, const double* U,
               double cfl, double dt_max, int unsplit, double* dt);
int orc_apply_bc(const orc_params* p, const int* N, const int* bc_lo, const int* bc_hi, double* U);
int orc_sweep(const orc_params* p, const int* N, const double* dx, int dir, double dt,
              const double* Uin, double* Uout);
/* One dim::advance_one_step: CFL dt, split (x->y->z) or unsplit sweeps, ring BCs. */
int orc_advance(const orc_params* p, const int* N, const double* dx, const int* bc_lo,
                const int* bc_hi, int unsplit, double cfl, double dt_max, double* U,
                double* dt_out);
/* FP64 conservation totals (conservation_report.hpp:188-227):
 * out = [mass_k * V ..., mom_d * V ..., E * V]. */
int orc_conservation_totals(const orc_params* p, const int* N, const double* dx,
                            const double* U, double* out);

const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
