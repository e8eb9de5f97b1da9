// PETSc-free Newtonian problem driver over the libCEED API (mirrors HONEE's SetupLibceed wiring); used by the tests,
// the benchmark and the time integrator of this package.
#pragma once

#include <hip/hip_runtime_api.h>

#include "ceed.h"
#include "honee_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

// Mesh and 1D tabulation supplied in place of DMPlex/PetscFE.
typedef struct {
  int           nelem, P, Q, nnodes, nxnodes;
  const int    *offsets;   // [nelem][P^3] node index of each element node (x fastest)
  const int    *xoffsets;  // [nelem][8] coordinate node index
  const double *xcoord;    // [nxnodes][3]
  const double *B, *D;     // [Q][P]
  const double *Bx, *Dx;   // [Q][2]
  const double *qref, *qweight;  // [Q]
} HoneeMiMesh;

typedef struct HoneeMiProblem_ *HoneeMiProblem;

int          HoneeMiProblemCreate(const char *resource, const HoneeMiMesh *mesh, const CeedMiNewtonianContext *ctx, int state_var,
                                  HoneeMiProblem *problem);
int          HoneeMiProblemDestroy(HoneeMiProblem *problem);
Ceed         HoneeMiProblemGetCeed(HoneeMiProblem p);
CeedOperator HoneeMiProblemGetOperator(HoneeMiProblem p, int which);  // 0: IFunction composite, 1: IJacobian composite
int          HoneeMiApplyIFunction(HoneeMiProblem p, int mem_type, double time, double dt, const double *q, const double *q_dot, double *g);
int          HoneeMiApplyIJacobian(HoneeMiProblem p, int mem_type, double shift, const double *dq, double *v);
int          HoneeMiIJacobianDiagonal(HoneeMiProblem p, int mem_type, double shift, double *diag);
// Point-block diagonals [nnodes][5 (c_out)][5 (c_in)] (CeedOperatorLinearAssemblePointBlockDiagonal) of the shifted
// IJacobian and of the explicit mass at state q
int          HoneeMiIJacobianPointBlockDiagonal(HoneeMiProblem p, int mem_type, double shift, double *pb);
int          HoneeMiMassPointBlockDiagonal(HoneeMiProblem p, int mem_type, const double *q, double *pb);
// Boundary faces of the problem (HONEE's BC definitions, problems/newtonian.c:555-569, problems/bc_freestream.c:30-180,
// problems/bc_outflow.c:114-180, problems/bc_slip.c:59-73): kind 0: consistent boundary integral; 1: freestream with
// HLL flux; 2: freestream with HLLC flux (S_inf_Y is the exterior primitive state); 3: Riemann outflow; 4: weak-pressure
// outflow (S_inf_Y holds {pressure, temperature, recirc, softplus_velocity}); 5: slip wall (S_inf_Y unused).  face_nodes [nfaces][P^2] are the face nodes in face-local (X1 fastest) order with
// outward dx/dX1 x dx/dX2, face_xnodes [nfaces][4] the coordinate nodes likewise.  The face operators are added to the
// IFunction and IJacobian composites.
int          HoneeMiProblemAddBoundary(HoneeMiProblem p, int kind, int nfaces, const int *face_nodes, const int *face_xnodes, const double *S_inf_Y);
// Essential (strong) boundary conditions on n L-vector entries dofs[i] = 5 node + component, with the boundary values
// values[i] (NULL: zero), as HONEE's DM_BC_ESSENTIAL boundaries with values from the initial condition or prescribed
// constants (SetBCsFromICs, src/misc.c:51-95; essential_comps, src/bc_definition.c:249-280).  From then on every
// operator input gets the boundary values inserted (and a zero time derivative) on those entries, their residual rows
// are zero, and the Jacobian, mass matrix, diagonals and point blocks are the identity on them.  n = 0 removes them.
int          HoneeMiProblemSetEssential(HoneeMiProblem p, int n, const int *dofs, const double *values);
int          HoneeMiGetBoundaryData(HoneeMiProblem p, int boundary, int which, double *out);  // which 0: surface qdata, 1: jac data
// Explicit problems (is_implicit false): RHS composite apply (RHS_NS, src/setupts.c:31-52), mass matrix action and
// diagonal at state q (the "mass_" KSP operator, src/setuplibceed.c:56-95; SUPG-stabilized when stabilization is SUPG)
int          HoneeMiApplyRHS(HoneeMiProblem p, int mem_type, double time, double dt, const double *q, double *r);
int          HoneeMiApplyMass(HoneeMiProblem p, int mem_type, const double *q, const double *x, double *y);
int          HoneeMiMassDiagonal(HoneeMiProblem p, int mem_type, const double *q, double *diag);
int          HoneeMiProblemIsExplicit(HoneeMiProblem p);
double       HoneeMiProblemVolume(HoneeMiProblem p);
// Total kinetic energy monitor sums (TotalKineticEnergy, StrainDissipation, DivergenceDissipation, VolumeExpansion,
// MuVorticity2), not divided by the volume (src/monitor_totalkineticenergy.c:113-140)
int          HoneeMiMonitorKE(HoneeMiProblem p, int mem_type, const double *q, double sums[5]);
int          HoneeMiGetQData(HoneeMiProblem p, double *out);
int          HoneeMiGetJacData(HoneeMiProblem p, double *out);
int          HoneeMiGetJacData32(HoneeMiProblem p, float *out);
// Data-driven SGS model (problems/sgs_dd_model.c:652-737): the parameter arrays as the reference's .dat files hold
// them (b1.dat [20], b2.dat [6], w1.dat [6 inputs][20 neurons], w2.dat [20 neurons][6 outputs], OutScaling.dat [6][2])
// and -sgs_model_dd_leakyrelu_alpha.  Setup projects the grid anisotropy tensor (CG) and adds the SGS residual to
// every HoneeMiApplyIFunction (implicit problems).
// An optional second hidden layer (an extension; null for the reference's model): bh [20] and wh [20 first-layer
// neurons][20 second-layer neurons], the .dat orientation; w2 then holds [second-layer neuron][output].
typedef struct {
  double        alpha;
  const double *b1, *b2, *w1, *w2, *out_scaling;
  const double *bh, *wh;
} HoneeMiSgsDDParams;
int HoneeMiProblemSetupSgsDD(HoneeMiProblem p, const HoneeMiSgsDDParams *params);
int HoneeMiSgsGridAnisotropy(HoneeMiProblem p, double *out, int *cg_iterations);  // [nnodes][7], host
int HoneeMiVelocityGradientProjection(HoneeMiProblem p, int mem_type, const double *q, double *grad_velo);  // [nnodes][9]
int HoneeMiSgsNodalStress(HoneeMiProblem p, int mem_type, const double *q, const double *grad_velo, double *km_sgs);  // [nnodes][6]
int HoneeMiApplySgsResidual(HoneeMiProblem p, int mem_type, const double *q, const double *km_sgs, double *g);  // g += SGS residual
// Differential filter (src/differential_filter.c:177-275): the options of DifferentialFilterSetup (damping_function
// 0 none, 1 van Driest, 2 MMS) and of its KSP (0 selects PETSc's defaults: rtol 1e-10 as HONEE sets it, atol 1e-50,
// max_it 10000).  The filtered field is [nnodes][11]: primitive state, then velocity products XX YY ZZ YZ XZ XY.
typedef struct {
  int    grid_based_width;
  double width_scaling[3], kernel_scaling, friction_length;
  int    damping_function;
  double damping_constant;
  double ksp_rtol, ksp_atol;
  int    ksp_max_it;
} HoneeMiDiffFilterOptions;
int HoneeMiProblemSetupDiffFilter(HoneeMiProblem p, const HoneeMiDiffFilterOptions *options);
int HoneeMiSetFilterWidthScaling(HoneeMiProblem p, const double scaling[3]);
int HoneeMiDiffFilterRHS(HoneeMiProblem p, int mem_type, const double *q, double *rhs);
int HoneeMiDiffFilterLHS(HoneeMiProblem p, const double *u, double *v, double *diag);  // host arrays [nnodes][11]
int HoneeMiDiffFilterApply(HoneeMiProblem p, int mem_type, const double *q, double *filtered, int *iterations);
// SGS training pairs (src/smartsim/sgs_dd_training.c): [nnodes][12] per filter width scale
int HoneeMiProblemSetupSgsTraining(HoneeMiProblem p);
int HoneeMiSgsTrainingData(HoneeMiProblem p, int mem_type, const double *q, double width_scale, double *out, int *filter_iterations);
int HoneeMiSgsTrainingFields(HoneeMiProblem p, double *filtered, double *grad_velo, double *aniso);  // host [nnodes][11], [9], [7]
const char  *HoneeMiProblemError(HoneeMiProblem p);
const char  *HoneeMiLastError(void);

#ifdef __cplusplus
}
#endif
