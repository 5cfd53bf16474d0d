/*
 * oracle.h -- C API of the CPU FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a line-by-line FP64 restatement of the reference's
 * dust drag path (AthenaK dust module) used to check the HIP library; nothing in the product
 * (athenak_amd/) may link, load or call it.  Only tests/, the smoke check of the package entry
 * and the CPU-baseline leg of the benchmark use it, always as the checker.
 *
 * The data layouts and the option structs are the library's (include/dustk.h); all pointers
 * here are host pointers.
 */
#ifndef ORACLE_ORACLE_H_
#define ORACLE_ORACLE_H_

#include <stdint.h>
#include "../include/dustk.h"
#include "../include/dustk_setup.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_sim or_sim;

or_sim *or_create(const dk_mesh *mesh, const dk_block *blocks, const dk_physics *phys,
                  int64_t capacity);
void or_destroy(or_sim *s);
/* OpenMP threads for the particle and cell loops (1 = serial, the deterministic order of
 * Kokkos::Serial; >1 = omp atomic deposits as Kokkos::OpenMP). */
void or_set_threads(or_sim *s, int n);

/* storage (host pointers owned by the oracle) */
double *or_rdata(or_sim *s);
int32_t *or_idata(or_sim *s);
int64_t or_capacity(or_sim *s);
int64_t or_npart(or_sim *s);
void or_set_npart(or_sim *s, int64_t n);
double *or_u0(or_sim *s);
double *or_u1(or_sim *s);
/* 0 qdep(5) 1 ustar(3) 2 dmom(3) 3 ap(3) 4 resid(4) */
double *or_field(or_sim *s, int which, int *nvar);
int64_t or_cells_per_block(or_sim *s);   /* n1*n2*n3 with ghosts */

/* the synthetic problem of dk_setup (dust_setup.hpp draws restated) */
int or_setup(or_sim *s, const dk_setup *su);

/* the task methods (same meaning as the dk_* entry points) */
int or_first_two_imp_rk(or_sim *s, const dk_stage *st);
int or_explicit_push(or_sim *s, const dk_stage *st);
int or_update_particle_gids(or_sim *s, const dk_stage *st);
int or_deposit_drag(or_sim *s, const dk_stage *st);
int or_fold_deposit_qp(or_sim *s, const dk_stage *st);
int or_gas_implicit_solve(or_sim *s, const dk_stage *st);
int or_fill_ustar(or_sim *s, const dk_stage *st);
int or_gather_kick_pmbr(or_sim *s, const dk_stage *st);
int or_fold_pmbr(or_sim *s, const dk_stage *st);
int or_apply_pmbr(or_sim *s, const dk_stage *st);
int or_add_drag_history_gas(or_sim *s, const dk_stage *st);
int or_gas_hook(or_sim *s, const dk_stage *st, const dk_gas_src *src);
int or_fill_gas_ghosts(or_sim *s, const dk_stage *st);
int or_stage_run(or_sim *s, const dk_stage *st, const dk_gas_src *src);
int or_sort_particles(or_sim *s);
int or_new_time_step(or_sim *s, double dt_cfl, double *dt_out);

/* operator pieces for brute-force tests: y = A x with A = diag(rho) + G^T diag(m c/V) G on
 * active cells (field layout of ustar, 3 components); uses qdep-free u0 densities. */
int or_apply_operator(or_sim *s, double a_dt, double sh, const double *x, double *y);

/* refinable-cloud pieces for closed-form tests (same-level plain clouds): the cells under the
 * cloud at x centred in block m (padded cell index (k*n2 + j)*n1 + i, weight, first moments
 * m1x, m1y, m1z; at most 27), returned count or -1; and the linear gather at x of the
 * 3-component padded field vel (the layout of ustar). */
int or_cloud_cells(or_sim *s, int m, const double x[3], int64_t *cell, double *w,
                   double *m1 /* [27][3] */);
int or_cloud_gather(or_sim *s, const double *vel, int m, const double x[3], double out[3]);
/* the same on a statically refined mesh (every block a leaf): the leaf cells under the cloud of
 * level lev at x (block, padded cell index, weight, first moments; at most cap), returned count
 * or -1; the linear gather there; the leaf holding x (-1 on a uniform mesh). */
int or_leaf_cells(or_sim *s, const double x[3], int lev, int32_t *blk, int64_t *cell,
                  double *w, double *m1, int cap);
int or_leaf_gather(or_sim *s, const double *vel, const double x[3], int lev, double out[3]);
int or_leaf_at(or_sim *s, const double x[3]);

int or_diag_sums(or_sim *s, double out[9]);
/* the strict solver's parameters (<dust> drag_rtol, drag_atol, drag_iter_max,
 * drag_adaptive_order_c, drag_adaptive_rtol_max, drag_adaptive_tref, drag_adaptive_state_floor)
 * and the last implicit solve's iterations, fast accept (0/1), true residual, error bound,
 * acceptance target, state scale and max Q/rho (-1 where the path did not form them) */
void or_set_solver(or_sim *s, double rtol, double atol, int iter_max, double order_c,
                   double rtol_max, double tref, double state_floor);
void or_solver_stats(or_sim *s, double out[7]);
/* UpdateStiffness records (dust_pm.cpp:288-342): chi of the last pc2 cycle, its running max,
 * the rate max */
void or_stiffness(or_sim *s, double out[3]);
int or_last_error(or_sim *s, int64_t *particle);
/* RemoveOutflowParticles (dust_tasks.cpp:365-420): out[0..1] mass and count removed now,
 * out[2..3] the run totals; or_x3_removed: the ledger since the last removal */
int or_remove_outflow_particles(or_sim *s, double out[4]);
void or_x3_removed(or_sim *s, double out[2]);

/* imex2+ (gamma = 1+1/sqrt(2)) and rk2 tableaux of the reference driver */
void or_imex2p_coef(double time, double dt, dk_stage *st_out /* [3] */);
void or_rk2_coef(double time, double dt, dk_stage *st_out /* [2] */);

#ifdef __cplusplus
}
#endif

#endif  /* ORACLE_ORACLE_H_ */
