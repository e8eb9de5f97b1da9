// PETSc-free time integration of a HoneeMiProblem on the GPU: the solver loop HONEE delegates to PETSc TS/SNES/KSP
// (TSSolve_NS, src/setupts.c:334-420, with the callbacks RHS_NS :31-75, IFunction_NS :77-119, FormIJacobian_NS
// :121-156 and MatGetDiagonal_Ceed, src/mat-ceed.c:1603-1630), restated over the device operators:
//   * explicit Runge-Kutta (TSRK; TSSSP third order) with the "mass_" KSP: preonly + point-Jacobi (row sum or diagonal)
//     or GMRES + point-Jacobi (src/setuplibceed.c:56-95),
//   * implicit BDF of order 1 or 2 (TSBDF), the generalized-alpha method (TSALPHA) and backward Euler (TSBEULER) with
//     Newton (SNESNEWTONLS, backtracking) and GMRES(restart) left-preconditioned by point-Jacobi from the IJacobian
//     diagonal or point blocks; the Jacobian action runs in FP32 when the problem's resource has
//     jacobian_precision=fp32.
// The solution, stage and Krylov vectors live in device memory; every operator apply, vector update and reduction is
// a HIP launch on the backend's stream, and the host reads back only the scalars the algorithms branch on.
#pragma once

#include "honee_setup.h"

#ifdef __cplusplus
extern "C" {
#endif

// TSALPHA (-ts_type alpha, examples/gaussianwave.yaml) and TSBEULER (-ts_type beuler, examples/channel.yaml) are the
// implicit integrators of HONEE's examples beside the default TSBDF (src/setupts.c:343)
typedef enum { HONEE_MI_TS_RK = 0, HONEE_MI_TS_BDF = 1, HONEE_MI_TS_ALPHA = 2, HONEE_MI_TS_BEULER = 3 } HoneeMiTSType;
// Explicit schemes: classic 4-stage RK ("4"), Fehlberg 5(4) propagating the 5th-order solution ("5f", HONEE's explicit
// default), optimal third-order SSP with n^2 stages (TSSSP "rks3", stages given by ssp_stages), Shu-Osher SSP(3,3),
// forward Euler
typedef enum { HONEE_MI_RK_4 = 0, HONEE_MI_RK_5F = 1, HONEE_MI_RK_SSP_S3 = 2, HONEE_MI_RK_SSP33 = 3, HONEE_MI_RK_EULER = 4 } HoneeMiRKType;

typedef struct {
  int    type;        // HoneeMiTSType
  int    rk_type;     // HoneeMiRKType
  int    ssp_stages;  // stages of HONEE_MI_RK_SSP_S3 (a square >= 4)
  int    bdf_order;   // 1 or 2
  double dt;
  // explicit mass solve ("mass_" KSP)
  int    mass_gmres;        // 0: preonly, 1: GMRES
  int    mass_jacobi_rowsum;  // Jacobi type: 1 row sum (HONEE's lumped default), 0 diagonal
  double mass_rtol;
  int    mass_restart, mass_max_it;
  // Newton (SNES) and its linear solve (KSP)
  double snes_rtol, snes_atol, snes_stol;
  int    snes_max_it;
  double ksp_rtol, ksp_atol;
  int    ksp_restart, ksp_max_it;
  // preconditioners: 0 point-Jacobi (PCJACOBI), 1 point-block Jacobi on the 5x5 node blocks (PCVPBJACOBI)
  int    ksp_pc, mass_pc;
  // rebuild the Jacobian preconditioner every pc_lag-th Newton iteration of a solve (the first always; PETSc's
  // -snes_lag_preconditioner; 1 rebuilds every iteration)
  int    pc_lag;
  // PETSc's -snes_mf_operator: the Krylov operator is the finite-difference product of the whole IFunction (MatMFFD,
  // Walker-Pernice step), the preconditioner still comes from the IJacobian.  Terms the IJacobian leaves out (the
  // data-driven SGS residual) then enter the Newton direction.
  int    snes_mf_operator;
  // TSALPHA's -ts_alpha_radius: spectral radius rho at infinity, alpha_m = (3 - rho) / (2 (1 + rho)),
  // alpha_f = 1 / (1 + rho), gamma = 1/2 + alpha_m - alpha_f (default 1)
  double alpha_radius;
  // PETSc's -snes_lag_jacobian L and -snes_lag_jacobian_persists: the Jacobian is evaluated (the preconditioner rebuilt
  // and the operator's shift updated) only when (Newton iteration + iterations of earlier solves, if persistent) is a
  // multiple of L.  As with HONEE's MATCEED Jacobian, the operator between evaluations acts with the latest IFunction's
  // Jacobian data and the shift of the last evaluation (FormIJacobian_NS, src/setupts.c:129-157).  Default 1, 0.
  int    snes_lag_jacobian;
  int    snes_lag_jacobian_persists;
  // Line search of the Newton solves: 0 PETSc's SNESLINESEARCHBT (cubic order; the default of SNESNEWTONLS and so of
  // HONEE), 1 SNESLINESEARCHBASIC (the full step), 2 backtracking by quadratic fits with the slope -||F||^2 (the slope
  // along an exact Newton direction; no Jacobian product, no minimum lambda, at most 40 cuts; not a PETSc type).
  // Default 0.
  int    snes_linesearch_type;
} HoneeMiTSOptions;

typedef struct {
  int       steps;
  double    time;
  long long rhs_evals, ifunction_evals, jacobian_applies, diagonal_assemblies, mass_applies;
  long long newton_its, krylov_its, mass_krylov_its, line_search_backtracks;
  int       last_newton_its, last_krylov_its;
  double    last_fnorm0, last_fnorm;
} HoneeMiTSStats;

typedef struct HoneeMiTS_ *HoneeMiTS;

// Defaults of PETSc and HONEE: RK "5f" / BDF2, lumped mass (preonly, row sum), SNES rtol 1e-8 atol 1e-50 stol 1e-8
// max_it 50, KSP rtol 1e-5 atol 1e-50 GMRES(30) max_it 10000, mass KSP rtol 1e-5.
void        HoneeMiTSOptionsDefault(HoneeMiTSOptions *o);
// Butcher tableau (A row-major [s][s], b[s], c[s]) of an explicit scheme; host-only
int         HoneeMiTSGetTableau(int rk_type, int ssp_stages, int max_stages, double *A, double *b, double *c, int *num_stages);
int         HoneeMiTSCreate(HoneeMiProblem p, const HoneeMiTSOptions *o, HoneeMiTS *ts);
// Set the solution (copied) and time; restarts the multistep history
int         HoneeMiTSSetSolution(HoneeMiTS ts, int mem_type, const double *q, double time);
int         HoneeMiTSGetSolution(HoneeMiTS ts, int mem_type, double *q);
// Device pointer of the solution (valid until the next step or destroy)
const double *HoneeMiTSGetSolutionDevice(HoneeMiTS ts);
// Advance nsteps steps; fails (nonzero) when a nonlinear solve does not converge or the solution is not finite
int         HoneeMiTSStep(HoneeMiTS ts, int nsteps);
int         HoneeMiTSGetStats(HoneeMiTS ts, HoneeMiTSStats *stats);
const char *HoneeMiTSError(HoneeMiTS ts);
int         HoneeMiTSDestroy(HoneeMiTS *ts);

#ifdef __cplusplus
}
#endif
