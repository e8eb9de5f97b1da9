// CPU ORACLE -- test infrastructure only (see ns_pointwise.h).
//
// Public C interface of the oracle shared library liboracle.so.  The oracle
// restates libCEED's /cpu/self operator semantics (restriction gather, tensor
// basis apply, QFunction per quadrature point, basis transpose, scatter-add)
// for the HONEE Newtonian operators, in plain FP64 C.
//
// Layouts:
//   L-vector       double[nnodes][5]       node-major, components interleaved
//   offsets        int[nelem][P^3]         node index of each element node (lexicographic, x fastest)
//   qdata          double[nelem][10][Q^3]  wdetJ, dXdx[3][3] row-major (setupgeo.h:42-56)
//   jac_data       double[nelem][njd][Q^3] q[5], strain rate (KM)[6], Tau_d[3] (+ sigma) (newtonian.h:251-254)
//   B, D           double[Q][P]            1D interpolation / derivative on [-1, 1]
#pragma once

#include "ns_pointwise.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int           nelem, P, Q, Px;
  int           nnodes, nxnodes;
  const int    *offsets;   // [nelem][P^3]
  const int    *xoffsets;  // [nelem][Px^3]
  const double *xcoord;    // [nxnodes][3]
  const double *B, *D;     // [Q][P]
  const double *Bx, *Dx;   // [Q][Px]
  const double *w1d;       // [Q]
} OMesh;

// Flat view of struct NewtonianIdealGasContext_ (qfunctions/newtonian_types.h:29-47)
typedef struct {
  OGas     gas;
  double   g[3];
  double   time, ijacobian_time_shift;
  double   idl_pressure, idl_amplitude, idl_start, idl_length;
  OTauCoef tau;
  double   dt;
  int      is_implicit, state_var, stabilization, idl_enable;
} ONewt;

// Freestream context: Newtonian context plus the exterior state (bc_freestream_type.h:9-12)
typedef struct {
  ONewt  newt;
  double S_inf_U[5], S_inf_Y[5];
  int    riemann;  // 0 HLL, 1 HLLC
} OFreestream;

// Outflow context (bc_freestream_type.h:14-21) and the outflow variant: 0 Riemann (bc_outflow.h:37-164), 1 weak
// pressure (bc_outflow.h:185-293)
typedef struct {
  ONewt  newt;
  double recirc, softplus_velocity, pressure, temperature;
  int    pressure_outflow;
} OOutflow;

// Boundary faces of a hex mesh: each face is described by its volume element and local face id
// (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z).
typedef struct {
  int        nfaces;
  const int *face_elem;
  const int *face_id;
} OFaces;

int    oracle_setup_qdata(const OMesh *m, double *qdata);
int    oracle_ifunction(const OMesh *m, const ONewt *c, const double *qdata, const double *q, const double *qdot, double *v, double *jac);
int    oracle_ijacobian(const OMesh *m, const ONewt *c, const double *qdata, const double *jac, const double *dq, double *v);
int    oracle_ijacobian_diagonal(const OMesh *m, const ONewt *c, const double *qdata, const double *jac, double *diag);
int    oracle_ijacobian_point_block_diagonal(const OMesh *m, const ONewt *c, const double *qdata, const double *jac, double *pb);
int    oracle_rhs(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *v);
int    oracle_mass_newtonian(const OMesh *m, const ONewt *c, const double *qdata, const double *q, const double *qdot, double *v);
int    oracle_mass_newtonian_diagonal(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *diag);
int    oracle_monitor_ke(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double sums[5]);
void   oracle_ic_taylorgreen(const ONewt *c, const double ref_Y[5], const double bg_u[3], const double L[3], int n, const double *xyz, double *q);
double oracle_volume(const OMesh *m, const double *qdata);

// Time integration (honee_oracle_solve.c)
typedef struct {
  int    gmres;          // 0: KSPPREONLY, 1: GMRES
  int    jacobi_rowsum;  // PCJACOBI type: 1 row sum, 0 diagonal
  double rtol;
  int    restart, max_it;
} ORKMass;

typedef struct {
  int    order;  // 1 or 2
  double dt;
  double snes_rtol, snes_atol, snes_stol;
  int    snes_max_it;
  double ksp_rtol, ksp_atol;
  int    ksp_restart, ksp_max_it;
  int    pc_pbjacobi;  // 0 point-Jacobi, 1 point-block Jacobi
  int    linesearch;   // 0 SNESLINESEARCHBT (cubic), 1 SNESLINESEARCHBASIC (full step), 2 quadratic fits with slope -||F||^2
} OBDFOptions;

typedef struct {
  int       steps;
  double    time;
  long long ifunction_evals, newton_its, krylov_its, line_search_backtracks;
  int       last_newton_its, last_krylov_its;
  double    last_fnorm0, last_fnorm;
} OBDFStats;

// Explicit Runge-Kutta (Butcher tableau A[s][s], b[s], c[s]) of the RHS with the mass KSP; optional kinetic-energy
// rows (step, time, 5 monitor sums / volume) every record_every steps
int oracle_rk(const OMesh *m, ONewt *c, const double *qdata, double *q, double t0, double dt, int nsteps, int s, const double *A, const double *b,
              const double *cs, const ORKMass *mo, int record_every, double *ke_rows, int *nrows, long long *mass_its);
// The RHS composite's extra parts: freestream faces (explicit sign) and essential boundary conditions as the box-mesh
// mirror restates HONEE's DMPlex constraints (src/misc.c:51-95): operator inputs q .* mask + qbc, RHS rows masked, the
// mass matrix the identity on the constrained entries.  Any member may be null.
typedef struct {
  const OFaces      *faces;
  const OFreestream *freestream;
  const double      *qdata_sur;
  const double      *mask, *qbc;  // [5 nnodes]: 0 / the boundary value at constrained entries, 1 / 0 elsewhere
} ORKExtra;
int oracle_rk_ex(const OMesh *m, ONewt *c, const double *qdata, double *q, double t0, double dt, int nsteps, int s, const double *A, const double *b,
                 const double *cs, const ORKMass *mo, const ORKExtra *ex, int record_every, double *ke_rows, int *nrows, long long *mass_its);
// BDF(order) with Newton-GMRES and point-Jacobi; returns nonzero when a nonlinear solve fails
int oracle_bdf(const OMesh *m, ONewt *c, const double *qdata, double *q, double t0, int nsteps, const OBDFOptions *o, OBDFStats *st);
// The backtracking line search's next lambda after a rejected trial with a finite residual (SNESLineSearchApply_BT,
// cubic order): quadratic fit when nfit == 0, else cubic fit through the last two trials, clamped to [0.1, 0.5] lambda
double oracle_bt_next_lambda(double fnorm, double initslope, double lambda, double gnorm, int nfit, double lamprev, double gnormprev);

// Explicit time stepping of the RHS with the SUPG-stabilized consistent mass (problems/newtonian.c:212-265) solved by
// left-Jacobi-preconditioned GMRES; used to reproduce the reference's Taylor-Green kinetic-energy fixture.
int oracle_rk4_stabilized(const OMesh *m, ONewt *c, const double *qdata, double *q, double dt, int nsteps, double ksp_rtol, int ksp_restart,
                          int record_every, double *ke_rows /* [(nsteps/record_every+2)][7] */, int *nrows);

// Data-driven SGS closure (honee_oracle.c; pointwise restatement in sgs_pointwise.h)
int oracle_velocity_gradient_rhs(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *r);
int oracle_grid_anisotropy_rhs(const OMesh *m, const double *qdata, double *r);
int oracle_mass_n(const OMesh *m, const double *qdata, int ncomp, const double *u, double *v);
int oracle_mass_n_diagonal(const OMesh *m, const double *qdata, int ncomp, double *d);
int oracle_velocity_gradient_projection(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *G);
int oracle_grid_anisotropy_projection(const OMesh *m, const double *qdata, double *X);
int oracle_ifunction_nodal_sgs(const OMesh *m, const ONewt *c, const double *qdata, const double *q, const double *km_sgs, double *v);
// Differential filter options (struct DifferentialFilterContext_, qfunctions/differential_filter.h:16-25)
typedef struct {
  int    grid_based_width;
  double width_scaling[3], kernel_scaling, friction_length;
  int    damping;  // 0 none, 1 van Driest, 2 MMS
  double damping_constant;
} ODiffFilter;
int oracle_anisotropy_collocate(const OMesh *m, const double *qdata, double *A);
int oracle_diff_filter_rhs(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *r);
int oracle_diff_filter_lhs(const OMesh *m, const ODiffFilter *f, const double *qdata, const double *Acol, int ncomp, const double *u, double *v);
int oracle_diff_filter_lhs_diagonal(const OMesh *m, const ODiffFilter *f, const double *qdata, const double *Acol, int ncomp, double *d);
int oracle_diff_filter_solve(const OMesh *m, const ODiffFilter *f, const double *qdata, const double *Acol, int ncomp, const double *R, double rtol,
                             double atol, int max_it, double *X);
int oracle_sgs_training_nodal(const OGas *g, int nnodes, const double *Yf, const double *G, const double *A, double *out);
int oracle_diagonalize3(const double *A, int max_sweeps, double *eval, double *evec);
void oracle_anisotropy_tensor(const double *dXdx, double *km_A, double *delta);

// Faces
int oracle_setup_qdata_boundary(const OMesh *m, const OFaces *f, double *qdata_sur);
int oracle_freestream(const OMesh *m, const OFaces *f, const OFreestream *c, const double *qdata_sur, const double *q, double *v, double *jac_sur);
int oracle_freestream_jacobian(const OMesh *m, const OFaces *f, const OFreestream *c, const double *qdata_sur, const double *jac_sur,
                               const double *dq, double *v);
int oracle_outflow(const OMesh *m, const OFaces *f, const OOutflow *c, const double *qdata_sur, const double *q, double *v, double *jac_sur);
int oracle_outflow_jacobian(const OMesh *m, const OFaces *f, const OOutflow *c, const double *qdata_sur, const double *jac_sur, const double *dq,
                            double *v);
int oracle_slip(const OMesh *m, const OFaces *f, const ONewt *c, const double *qdata_sur, const double *q, double *v, double *jac_sur);
int oracle_slip_jacobian(const OMesh *m, const OFaces *f, const ONewt *c, const double *qdata_sur, const double *jac_sur, const double *dq,
                         double *v);
int oracle_boundary_integral(const OMesh *m, const OFaces *f, const ONewt *c, const double *qdata_sur, const double *q, double *v,
                             double *jac_sur);
int oracle_boundary_integral_jacobian(const OMesh *m, const OFaces *f, const ONewt *c, const double *qdata_sur, const double *jac_sur,
                                      const double *dq, double *v);

// Pointwise entry points used by the unit tests
void oracle_state_from_q(const OGas *g, const double *q, int var, double *U, double *Y);
void oracle_state_from_q_fwd(const OGas *g, const double *q, const double *dq, int var, double *dU, double *dY);
void oracle_state_to_q(const OGas *g, const double *U, const double *Y, int var, double *q);
void oracle_state_to_q_fwd(const OGas *g, const double *U, const double *Y, const double *dU, const double *dY, int var, double *dq);
void oracle_riemann_flux(const OGas *g, int hllc, const double *qL, const double *qR, int var, const double *n, double *F);
void oracle_riemann_flux_fwd(const OGas *g, int hllc, const double *qL, const double *dqL, const double *qR, const double *dqR, int var,
                             const double *n, double *dF);
void oracle_hll_speeds(const OGas *g, const double *qL, const double *qR, int var, const double *n, double *sL, double *sR);
void oracle_hll_speeds_fwd(const OGas *g, const double *qL, const double *dqL, const double *qR, const double *dqR, int var, const double *n,
                           double *sL, double *dsL, double *sR, double *dsR);
void oracle_set_num_threads(int n);
int  oracle_max_threads(void);
void oracle_qf_ifunction(const ONewt *c, const double *q, const double *gref, const double *qdot, double wdetJ, const double *dXdx,
                         const double *x, double *v, double *gv, double *jd);
void oracle_qf_ijacobian(const ONewt *c, const double *dq, const double *gref, double wdetJ, const double *dXdx, const double *jd, double *v,
                         double *gv);
void oracle_qf_rhs(const ONewt *c, const double *U, const double *gref, double wdetJ, const double *dXdx, const double *x, double *v,
                   double *gv);
void oracle_tau_diag(const OGas *g, const OTauCoef *tc, const double *q, int var, const double *dXdx, double dt, double *Tau);

#ifdef __cplusplus
}
#endif
