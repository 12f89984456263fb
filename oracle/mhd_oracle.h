/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 *
 * CPU oracle: a plain C++ restatement of ppkMHD's deterministic MUSCL-Hancock + HLLD + CT MHD
 * step (implementation version 0), 2D and 3D, used only by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, and only as the checker.  The reference itself cannot be built
 * here (its only parallel runtime, Kokkos, is an empty submodule), so this restatement follows
 * the reference source files line by line (citations per function in mhd_oracle.cpp) and is
 * pinned by the mathematical checks under tests/ (textbook HLLD, consistency, symmetry,
 * div B, convergence order).
 *
 * Arrays use the reference layout of include/ppkmhd_amd.h:
 *     U[((v*ksize + k)*jsize + j)*isize + i], v in {ID, IP, IU, IV, IW, IA, IB, IC}.
 */
#ifndef PPKMHD_ORACLE_H_
#define PPKMHD_ORACLE_H_

#include "../include/ppkmhd_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_solver ora_solver;

int ora_params_size(void);
int ora_problem_size(void);

/* stage level (arrays over the padded box) */
void ora_fill_boundaries(const ppk_mhd_params *p, double *U);
void ora_primitives(const ppk_mhd_params *p, const double *U, double *Q);
double ora_inv_dt(const ppk_mhd_params *p, const double *Q);
void ora_init_problem_array(const ppk_mhd_params *p, const ppk_mhd_problem *prob, double *U);

/* point-wise pieces (8-component primitive states) */
void ora_riemann(const ppk_mhd_params *p, const double *qleft, const double *qright,
                 double *flux);
/* qedge: 4 states in the order RT, RB, LT, LB (enums IRT..ILB); emf_dir 0 = X, 1 = Y, 2 = Z */
double ora_compute_emf(const ppk_mhd_params *p, int emf_dir, const double *qedge);
double ora_mag_riemann2d_hlld(const ppk_mhd_params *p, const double *qLLRR,
                              const double *eLLRR);
double ora_find_speed_fast(const ppk_mhd_params *p, const double *q, int dir);
double ora_slope(const ppk_mhd_params *p, double q, double qplus, double qminus);

/* time loop (SolverMHDMuscl + SolverBase), owns U / U2 / Q and the impl-0 scratch */
ora_solver *ora_create(const ppk_mhd_params *p);
void ora_destroy(ora_solver *s);
void ora_init_problem(ora_solver *s, const ppk_mhd_problem *prob);
void ora_set_state(ora_solver *s, const double *U);
void ora_get_state(const ora_solver *s, double *U);
/* dt_mode: 0 fixed (dt argument), 1 CFL (reference).  Returns the dt used. */
double ora_step(ora_solver *s, double dt, int dt_mode);
void ora_get_time(const ora_solver *s, double *t, double *dt, long long *iteration);
void ora_set_time(ora_solver *s, double t, long long iteration);
int ora_finished(const ora_solver *s);
/* intermediate arrays of the last step, for stage-level checks: which = 0 Q, 1 Fluxes_x,
 * 2 Fluxes_y, 3 Fluxes_z, 4 Emf (3 comps in 3D: I_EMFZ, I_EMFY, I_EMFX; 1 in 2D) */
const double *ora_scratch(const ora_solver *s, int which);

/* operation counts of one step (instrumented build only; zeros otherwise):
 * counts[0] add/sub, [1] mul, [2] div, [3] sqrt, [4] min/max/abs/compare-select */
void ora_op_counts(long long counts[5]);
/* same, for one stage of the step: 0 boundaries+copy, 1 primitives+dt, 2 E field (3D),
 * 3 magnetic slopes (3D), 4 trace, 5 fluxes, 6 EMFs, 7 update + CT */
void ora_op_counts_stage(int stage, long long counts[5]);
void ora_op_counts_reset(void);
int ora_is_counting_build(void);

#ifdef __cplusplus
}
#endif

#endif
