// PETSc-free restatement of HONEE's operator wiring for the Newtonian ideal-gas problem, written against the public
// libCEED API only (it links to libceed.so like HONEE does).  It builds exactly the fields, composites and context
// registrations of SetupLibceed (src/setuplibceed.c:97-318), QDataGet (src/qdata.c:133-190) and NS_NEWTONIAN_IG's
// context setup (problems/newtonian.c:723-757), with the mesh, basis tabulation and offsets supplied by the caller in
// place of DMPlex/PetscFE.  Vectors are handed to the operators with CeedVectorSetArray/TakeArray around every
// apply, as VecReadPetscToCeed/VecReadCeedToPetsc do (include/petsc-ceed-utils.h:127-160).
#include "honee_setup.h"

#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

// QFunction bodies stay on the host side of HONEE; the MI355X backend identifies them by location string and runs its
// fused kernels instead, so the host bodies here only report that they are not executed.
static int HoneeMiQFunctionNotOnHost(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  return CEED_ERROR_UNSUPPORTED;
}

static const char *const kIFunctionLoc[3] = {"qfunctions/newtonian.h:IFunction_Newtonian_Conserv", "qfunctions/newtonian.h:IFunction_Newtonian_Prim",
                                             "qfunctions/newtonian.h:IFunction_Newtonian_Entropy"};
static const char *const kIJacobianLoc[3] = {"qfunctions/newtonian.h:IJacobian_Newtonian_Conserv", "qfunctions/newtonian.h:IJacobian_Newtonian_Prim",
                                             "qfunctions/newtonian.h:IJacobian_Newtonian_Entropy"};
static const char *const kRHSLoc          = "qfunctions/newtonian.h:RHSFunction_Newtonian";
static const char *const kMassStabLoc     = "qfunctions/newtonian.h:MassFunction_Newtonian_Conserv";
static const char *const kMass5Loc        = "qfunctions/mass.h:Mass_5";
static const char *const kMonitorKELoc[3] = {"qfunctions/monitor_totalkineticenergy.h:MonitorTotalKineticEnergy_Conserv",
                                             "qfunctions/monitor_totalkineticenergy.h:MonitorTotalKineticEnergy_Prim",
                                             "qfunctions/monitor_totalkineticenergy.h:MonitorTotalKineticEnergy_Entropy"};
static const char *const kSetupLoc        = "qfunctions/setupgeo.h:Setup";
static const char *const kSetupBoundaryLoc = "qfunctions/setupgeo.h:SetupBoundary";
static const char *const kVarName[3]       = {"Conserv", "Prim", "Entropy"};

#define HONEE_MI_MAX_BOUNDARIES 8

// One boundary definition: its face restrictions, bases, quadrature data and Jacobian data
typedef struct {
  int                  kind, nfaces, njac;
  CeedElemRestriction  restr_q, restr_x, restr_qd, restr_jd;
  CeedBasis            basis_q, basis_x;
  CeedVector           x_coord, q_data, jac_data;
  CeedQFunctionContext fs_qfctx;
} HoneeMiBoundary;

struct HoneeMiProblem_ {
  Ceed                  ceed;
  CeedElemRestriction   elem_restr_q, elem_restr_x, elem_restr_qd, elem_restr_jd;
  CeedBasis             basis_q, basis_x;
  CeedVector            q_ceed, q_dot_ceed, g_ceed, x_coord, q_data, jac_data;
  CeedOperator          op_ifunction, op_ijacobian;
  CeedOperator          op_rhs, op_mass, op_monitor;  // explicit problems: RHS composite and mass; KE monitor
  CeedElemRestriction   elem_restr_ke;
  CeedVector            ke_values, pb_ceed;
  // essential (strong) boundary conditions: mask (0 at constrained entries), its complement, boundary values, the
  // point-block mask and identity, and work copies of the operator inputs
  int                   ess_enabled;
  CeedVector            ess_mask, ess_cmask, ess_qbc, ess_pb_mask, ess_pb_eye, ess_q, ess_qdot;
  CeedContextFieldLabel rhs_time_label, rhs_dt_label;
  int                   mass_stabilized;
  double                volume;
  CeedContextFieldLabel solution_time_label, timestep_size_label, ijacobian_shift_label;
  CeedQFunctionContext  newt_qfctx;
  CeedMiNewtonianContext *ctx;
  int                   nelem, Q3, nnodes, num_comps_jac_data, state_var;
  HoneeMiMesh           mesh;  // tabulation pointers stay owned by the caller for the problem's lifetime
  HoneeMiBoundary       bnd[HONEE_MI_MAX_BOUNDARIES];
  int                   nbnd;
  // data-driven SGS model (problems/sgs_dd_model.c): nodal fields, projections and operators
  int                   sgs_enabled, grid_aniso_its;
  CeedElemRestriction   elem_restr_grad_velo, elem_restr_grid_aniso, elem_restr_sgs, elem_restr_inv_mult;
  CeedBasis             basis_grad_velo, basis_grid_aniso, basis_sgs, basis_x_to_q;
  CeedVector            grad_velo, grad_velo_inv_rowsum, grid_aniso, sgs_nodal, inv_mult;
  CeedOperator          op_grad_velo_rhs, op_sgs_nodal, op_sgs_apply;
  CeedQFunctionContext  sgsdd_qfctx;
  // differential filter (src/differential_filter.c) and SGS training pairs (src/smartsim/sgs_dd_training.c)
  int                   filter_enabled, filter_pc_ready, train_enabled, filter_its;
  double                filter_rtol, filter_atol;
  int                   filter_max_it;
  CeedElemRestriction   elem_restr_filter[2], elem_restr_aniso_col, elem_restr_train;
  CeedBasis             basis_filter[2];
  CeedVector            filter_rhs, filtered, filter_dinv, aniso_col, train_data, cg_work[3];
  CeedOperator          op_filter_rhs, op_filter_lhs, op_filtered_grad_velo_rhs, op_sgs_train;
  CeedQFunctionContext  diff_filter_qfctx, train_qfctx;
  CeedContextFieldLabel filter_width_label;
  char                  err[512];
};

#define CHK(p, expr)                                                                                         \
  do {                                                                                                       \
    int ierr_ = (expr);                                                                                      \
    if (ierr_) {                                                                                             \
      const char *msg_ = "";                                                                                 \
      if ((p)->ceed) CeedGetErrorMessage((p)->ceed, &msg_);                                                  \
      snprintf((p)->err, sizeof((p)->err), "%s failed (%d): %s", #expr, ierr_, msg_);                       \
      return ierr_;                                                                                          \
    }                                                                                                        \
  } while (0)

static char g_last_error[512];

const char *HoneeMiLastError(void) { return g_last_error; }

static int FreeContext(void *data) {
  free(data);
  return 0;
}

static int Setup(HoneeMiProblem p, const char *resource, const HoneeMiMesh *m, const CeedMiNewtonianContext *ctx_in, int state_var) {
  const int P = m->P, Q = m->Q, P3 = P * P * P;
  p->mesh      = *m;
  p->state_var = state_var;
  p->nelem     = m->nelem;
  p->Q3     = Q * Q * Q;
  p->nnodes = m->nnodes;
  CHK(p, CeedInit(resource, &p->ceed));
  CHK(p, CeedSetErrorHandler(p->ceed, CeedErrorStore));

  // Bases: solution (5 comps) and trilinear coordinates (3 comps), DMPlexCeedBasisCreate (src/dm-utils.c:844-850)
  CHK(p, CeedBasisCreateTensorH1(p->ceed, 3, 5, P, Q, m->B, m->D, m->qref, m->qweight, &p->basis_q));
  CHK(p, CeedBasisCreateTensorH1(p->ceed, 3, 3, 2, Q, m->Bx, m->Dx, m->qref, m->qweight, &p->basis_x));

  // Restrictions with L-vector offsets of component 0 (src/dm-utils.c:108-112)
  {
    CeedInt *off = malloc(sizeof(CeedInt) * (size_t)m->nelem * P3);
    for (size_t i = 0; i < (size_t)m->nelem * P3; i++) off[i] = 5 * m->offsets[i];
    CHK(p, CeedElemRestrictionCreate(p->ceed, m->nelem, P3, 5, 1, 5 * (CeedSize)m->nnodes, CEED_MEM_HOST, CEED_COPY_VALUES, off, &p->elem_restr_q));
    for (size_t i = 0; i < (size_t)m->nelem * 8; i++) off[i] = 3 * m->xoffsets[i];
    CHK(p, CeedElemRestrictionCreate(p->ceed, m->nelem, 8, 3, 1, 3 * (CeedSize)m->nxnodes, CEED_MEM_HOST, CEED_COPY_VALUES, off, &p->elem_restr_x));
    free(off);
  }
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_x, &p->x_coord, NULL));
  CHK(p, CeedVectorSetArray(p->x_coord, CEED_MEM_HOST, CEED_COPY_VALUES, (CeedScalar *)m->xcoord));

  // QDataGet: Setup QFunction applied to the coordinates (src/qdata.c:133-190)
  {
    CeedQFunction qf_setup;
    CeedOperator  op_setup;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kSetupLoc, &qf_setup));
    CHK(p, CeedQFunctionSetUserFlopsEstimate(qf_setup, 0));
    CHK(p, CeedQFunctionAddInput(qf_setup, "dx", 9, CEED_EVAL_GRAD));
    CHK(p, CeedQFunctionAddInput(qf_setup, "weight", 1, CEED_EVAL_WEIGHT));
    CHK(p, CeedQFunctionAddOutput(qf_setup, "surface qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedElemRestrictionCreateStrided(p->ceed, m->nelem, p->Q3, 10, (CeedSize)m->nelem * p->Q3 * 10, CEED_STRIDES_BACKEND, &p->elem_restr_qd));
    CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_qd, &p->q_data, NULL));
    CHK(p, CeedOperatorCreate(p->ceed, qf_setup, NULL, NULL, &op_setup));
    CHK(p, CeedOperatorSetField(op_setup, "dx", p->elem_restr_x, p->basis_x, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_setup, "weight", CEED_ELEMRESTRICTION_NONE, p->basis_x, CEED_VECTOR_NONE));
    CHK(p, CeedOperatorSetField(op_setup, "surface qdata", p->elem_restr_qd, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorApply(op_setup, p->x_coord, p->q_data, CEED_REQUEST_IMMEDIATE));
    CHK(p, CeedOperatorDestroy(&op_setup));
    CHK(p, CeedQFunctionDestroy(&qf_setup));
  }

  // Newtonian QFunction context with the registered fields of problems/newtonian.c:744-753
  p->ctx  = malloc(sizeof(*p->ctx));
  *p->ctx = *ctx_in;
  CHK(p, CeedQFunctionContextCreate(p->ceed, &p->newt_qfctx));
  CHK(p, CeedQFunctionContextSetData(p->newt_qfctx, CEED_MEM_HOST, CEED_USE_POINTER, sizeof(*p->ctx), p->ctx));
  CHK(p, CeedQFunctionContextSetDataDestroy(p->newt_qfctx, CEED_MEM_HOST, FreeContext));
  CHK(p, CeedQFunctionContextRegisterDouble(p->newt_qfctx, "timestep size", offsetof(CeedMiNewtonianContext, dt), 1, "Size of timestep, delta t"));
  CHK(p, CeedQFunctionContextRegisterDouble(p->newt_qfctx, "ijacobian time shift", offsetof(CeedMiNewtonianContext, ijacobian_time_shift), 1,
                                            "Shift for mass matrix in IJacobian"));
  CHK(p, CeedQFunctionContextRegisterDouble(p->newt_qfctx, "solution time", offsetof(CeedMiNewtonianContext, time), 1, "Current solution time"));

  p->num_comps_jac_data = 14 + (ctx_in->idl_enable ? 1 : 0);  // problems/newtonian.c:756-757
  CHK(p, CeedElemRestrictionCreateStrided(p->ceed, m->nelem, p->Q3, p->num_comps_jac_data, (CeedSize)m->nelem * p->Q3 * p->num_comps_jac_data,
                                          CEED_STRIDES_BACKEND, &p->elem_restr_jd));
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_jd, &p->jac_data, NULL));
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_q, &p->q_ceed, NULL));
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_q, &p->q_dot_ceed, NULL));
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_q, &p->g_ceed, NULL));

  CeedOperator op_ifunction_vol, op_ijacobian_vol;
  {  // IFunction volume operator (src/setuplibceed.c:209-243)
    CeedQFunction qf;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kIFunctionLoc[state_var], &qf));
    CHK(p, CeedQFunctionSetContext(qf, p->newt_qfctx));
    CHK(p, CeedQFunctionSetUserFlopsEstimate(qf, 0));
    CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "Grad_q", 15, CEED_EVAL_GRAD));
    CHK(p, CeedQFunctionAddInput(qf, "q dot", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "x", 3, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddOutput(qf, "v", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddOutput(qf, "Grad_v", 15, CEED_EVAL_GRAD));
    CHK(p, CeedQFunctionAddOutput(qf, "jac_data", p->num_comps_jac_data, CEED_EVAL_NONE));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &op_ifunction_vol));
    CHK(p, CeedOperatorSetName(op_ifunction_vol, "IFunction Domain Operator"));
    CHK(p, CeedOperatorSetField(op_ifunction_vol, "q", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_ifunction_vol, "Grad_q", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_ifunction_vol, "q dot", p->elem_restr_q, p->basis_q, p->q_dot_ceed));
    CHK(p, CeedOperatorSetField(op_ifunction_vol, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
    CHK(p, CeedOperatorSetField(op_ifunction_vol, "x", p->elem_restr_x, p->basis_x, p->x_coord));
    CHK(p, CeedOperatorSetField(op_ifunction_vol, "v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_ifunction_vol, "Grad_v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_ifunction_vol, "jac_data", p->elem_restr_jd, CEED_BASIS_NONE, p->jac_data));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  {  // IJacobian volume operator (src/setuplibceed.c:245-269)
    CeedQFunction qf;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kIJacobianLoc[state_var], &qf));
    CHK(p, CeedQFunctionSetContext(qf, p->newt_qfctx));
    CHK(p, CeedQFunctionSetUserFlopsEstimate(qf, 0));
    CHK(p, CeedQFunctionAddInput(qf, "dq", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "Grad_dq", 15, CEED_EVAL_GRAD));
    CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "jac_data", p->num_comps_jac_data, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddOutput(qf, "v", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddOutput(qf, "Grad_v", 15, CEED_EVAL_GRAD));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &op_ijacobian_vol));
    CHK(p, CeedOperatorSetName(op_ijacobian_vol, "IJacobian Domain Operator"));
    CHK(p, CeedOperatorSetField(op_ijacobian_vol, "dq", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_ijacobian_vol, "Grad_dq", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_ijacobian_vol, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
    CHK(p, CeedOperatorSetField(op_ijacobian_vol, "jac_data", p->elem_restr_jd, CEED_BASIS_NONE, p->jac_data));
    CHK(p, CeedOperatorSetField(op_ijacobian_vol, "v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_ijacobian_vol, "Grad_v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  // Composites (src/setuplibceed.c:291-306)
  CHK(p, CeedOperatorCreateComposite(p->ceed, &p->op_ifunction));
  CHK(p, CeedOperatorSetName(p->op_ifunction, "IFunction Composite Operator"));
  CHK(p, CeedOperatorCompositeAddSub(p->op_ifunction, op_ifunction_vol));
  CHK(p, CeedOperatorCreateComposite(p->ceed, &p->op_ijacobian));
  CHK(p, CeedOperatorSetName(p->op_ijacobian, "IJacobian Composite Operator"));
  CHK(p, CeedOperatorCompositeAddSub(p->op_ijacobian, op_ijacobian_vol));
  CHK(p, CeedOperatorGetContextFieldLabel(p->op_ifunction, "solution time", &p->solution_time_label));
  CHK(p, CeedOperatorGetContextFieldLabel(p->op_ifunction, "timestep size", &p->timestep_size_label));
  CHK(p, CeedOperatorGetContextFieldLabel(p->op_ijacobian, "ijacobian time shift", &p->ijacobian_shift_label));
  CHK(p, CeedOperatorDestroy(&op_ifunction_vol));
  CHK(p, CeedOperatorDestroy(&op_ijacobian_vol));

  if (!ctx_in->is_implicit) {  // explicit problems: RHS composite (src/setuplibceed.c:176-203, 272-287) and the mass operator
    CeedQFunction qf;
    CeedOperator  op_rhs_vol;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kRHSLoc, &qf));
    CHK(p, CeedQFunctionSetContext(qf, p->newt_qfctx));
    CHK(p, CeedQFunctionSetUserFlopsEstimate(qf, 0));
    CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "Grad_q", 15, CEED_EVAL_GRAD));
    CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "x", 3, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddOutput(qf, "v", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddOutput(qf, "Grad_v", 15, CEED_EVAL_GRAD));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &op_rhs_vol));
    CHK(p, CeedOperatorSetField(op_rhs_vol, "q", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_rhs_vol, "Grad_q", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_rhs_vol, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
    CHK(p, CeedOperatorSetField(op_rhs_vol, "x", p->elem_restr_x, p->basis_x, p->x_coord));
    CHK(p, CeedOperatorSetField(op_rhs_vol, "v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_rhs_vol, "Grad_v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedQFunctionDestroy(&qf));
    CHK(p, CeedOperatorCreateComposite(p->ceed, &p->op_rhs));
    CHK(p, CeedOperatorCompositeAddSub(p->op_rhs, op_rhs_vol));
    CHK(p, CeedOperatorDestroy(&op_rhs_vol));
    CHK(p, CeedOperatorGetContextFieldLabel(p->op_rhs, "solution time", &p->rhs_time_label));
    CHK(p, CeedOperatorGetContextFieldLabel(p->op_rhs, "timestep size", &p->rhs_dt_label));

    // Mass: SUPG-stabilized for explicit SUPG (problems/newtonian.c:795, 212-265), else Mass_5 (src/setuplibceed.c:18-50).
    // The stabilized mass reads the current state through the passive field "q", bound to q_ceed as in the reference.
    p->mass_stabilized = ctx_in->stabilization == 2;
    if (p->mass_stabilized) {
      CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kMassStabLoc, &qf));
      CHK(p, CeedQFunctionSetContext(qf, p->newt_qfctx));
      CHK(p, CeedQFunctionSetUserFlopsEstimate(qf, 0));
      CHK(p, CeedQFunctionAddInput(qf, "q_dot", 5, CEED_EVAL_INTERP));
      CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_INTERP));
      CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
      CHK(p, CeedQFunctionAddOutput(qf, "v", 5, CEED_EVAL_INTERP));
      CHK(p, CeedQFunctionAddOutput(qf, "Grad_v", 15, CEED_EVAL_GRAD));
      CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &p->op_mass));
      CHK(p, CeedOperatorSetName(p->op_mass, "RHS Mass Operator, Newtonian Stabilized"));
      CHK(p, CeedOperatorSetField(p->op_mass, "q_dot", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
      CHK(p, CeedOperatorSetField(p->op_mass, "q", p->elem_restr_q, p->basis_q, p->q_ceed));
      CHK(p, CeedOperatorSetField(p->op_mass, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
      CHK(p, CeedOperatorSetField(p->op_mass, "v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
      CHK(p, CeedOperatorSetField(p->op_mass, "Grad_v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    } else {
      CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kMass5Loc, &qf));
      CHK(p, CeedQFunctionAddInput(qf, "u", 5, CEED_EVAL_INTERP));
      CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
      CHK(p, CeedQFunctionAddOutput(qf, "v", 5, CEED_EVAL_INTERP));
      CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &p->op_mass));
      CHK(p, CeedOperatorSetName(p->op_mass, "RHS Mass Operator"));
      CHK(p, CeedOperatorSetField(p->op_mass, "u", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
      CHK(p, CeedOperatorSetField(p->op_mass, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
      CHK(p, CeedOperatorSetField(p->op_mass, "v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    }
    CHK(p, CeedQFunctionDestroy(&qf));
  }

  {  // Total kinetic energy monitor (src/monitor_totalkineticenergy.c:66-110), values at quadrature points
    CeedQFunction qf;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kMonitorKELoc[state_var], &qf));
    CHK(p, CeedQFunctionSetContext(qf, p->newt_qfctx));
    CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "Grad_q", 15, CEED_EVAL_GRAD));
    CHK(p, CeedQFunctionAddInput(qf, "q_data", 10, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddOutput(qf, "v", 5, CEED_EVAL_NONE));
    CHK(p, CeedElemRestrictionCreateStrided(p->ceed, m->nelem, p->Q3, 5, (CeedSize)m->nelem * p->Q3 * 5, CEED_STRIDES_BACKEND, &p->elem_restr_ke));
    CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_ke, &p->ke_values, NULL));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &p->op_monitor));
    CHK(p, CeedOperatorSetField(p->op_monitor, "q", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(p->op_monitor, "Grad_q", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(p->op_monitor, "q_data", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
    CHK(p, CeedOperatorSetField(p->op_monitor, "v", p->elem_restr_ke, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  {  // domain volume: sum of w det J (HoneeCalculateDomainSize, src/misc.c)
    const CeedScalar *a;
    CHK(p, CeedVectorGetArrayRead(p->q_data, CEED_MEM_HOST, &a));
    double vol = 0;
    for (long long e = 0; e < p->nelem; e++)
      for (long long i = 0; i < p->Q3; i++) vol += a[e * 10 * p->Q3 + i];
    CHK(p, CeedVectorRestoreArrayRead(p->q_data, &a));
    p->volume = vol;
  }
  return 0;
}

int HoneeMiProblemCreate(const char *resource, const HoneeMiMesh *mesh, const CeedMiNewtonianContext *ctx, int state_var, HoneeMiProblem *problem) {
  HoneeMiProblem p = calloc(1, sizeof(*p));
  if (state_var < 0 || state_var > 2) {
    snprintf(g_last_error, sizeof(g_last_error), "invalid state variable %d", state_var);
    free(p);
    return CEED_ERROR_MAJOR;
  }
  int ierr = Setup(p, resource, mesh, ctx, state_var);
  if (ierr) {
    snprintf(g_last_error, sizeof(g_last_error), "%s", p->err);
    HoneeMiProblemDestroy(&p);
    return ierr;
  }
  *problem = p;
  return 0;
}

Ceed HoneeMiProblemGetCeed(HoneeMiProblem p) { return p->ceed; }

CeedOperator HoneeMiProblemGetOperator(HoneeMiProblem p, int which) { return which == 0 ? p->op_ifunction : p->op_ijacobian; }

// StateFromPrimitive (newtonian_state.h:175-180) for the exterior state of the freestream context
static void StateFromPrimitive(const CeedMiGas *gas, const double Y[5], CeedMiState *s) {
  const double Rgas = gas->cp - gas->cv;
  for (int i = 0; i < 5; i++) s->Y[i] = Y[i];
  s->U[0]         = Y[0] / (Rgas * Y[4]);
  for (int i = 0; i < 3; i++) s->U[1 + i] = s->U[0] * Y[1 + i];
  const double ei = gas->cv * Y[4];
  const double ek = .5 * (Y[1] * Y[1] + Y[2] * Y[2] + Y[3] * Y[3]);
  s->U[4]         = s->U[0] * (ei + ek);
}

// A boundary definition (HoneeBCAddIFunctionOp / HoneeBCAddIJacobianOp, src/boundary_condition.c:290-490)
static int AddBoundary(HoneeMiProblem p, int kind, int nfaces, const int *face_nodes, const int *face_xnodes, const double *S_inf_Y) {
  if (p->nbnd >= HONEE_MI_MAX_BOUNDARIES || kind < 0 || kind > 5 || nfaces <= 0) return CEED_ERROR_DIMENSION;
  HoneeMiBoundary *b = &p->bnd[p->nbnd++];
  const HoneeMiMesh *m = &p->mesh;
  const int P = m->P, Q = m->Q, P2 = P * P, Q2 = Q * Q;
  b->kind   = kind;
  b->nfaces = nfaces;
  b->njac   = (kind == 0 || kind == 3 || kind == 4) ? 11 : 5;  // boundary integral and outflows store {q, strain rate}
  CHK(p, CeedBasisCreateTensorH1(p->ceed, 2, 5, P, Q, m->B, m->D, m->qref, m->qweight, &b->basis_q));
  CHK(p, CeedBasisCreateTensorH1(p->ceed, 2, 3, 2, Q, m->Bx, m->Dx, m->qref, m->qweight, &b->basis_x));
  {
    CeedInt *off = malloc(sizeof(CeedInt) * (size_t)nfaces * P2);
    for (size_t i = 0; i < (size_t)nfaces * P2; i++) off[i] = 5 * face_nodes[i];
    CHK(p, CeedElemRestrictionCreate(p->ceed, nfaces, P2, 5, 1, 5 * (CeedSize)m->nnodes, CEED_MEM_HOST, CEED_COPY_VALUES, off, &b->restr_q));
    for (size_t i = 0; i < (size_t)nfaces * 4; i++) off[i] = 3 * face_xnodes[i];
    CHK(p, CeedElemRestrictionCreate(p->ceed, nfaces, 4, 3, 1, 3 * (CeedSize)m->nxnodes, CEED_MEM_HOST, CEED_COPY_VALUES, off, &b->restr_x));
    free(off);
  }
  CHK(p, CeedElemRestrictionCreateVector(b->restr_x, &b->x_coord, NULL));
  CHK(p, CeedVectorSetArray(b->x_coord, CEED_MEM_HOST, CEED_COPY_VALUES, (CeedScalar *)m->xcoord));
  {  // boundary quadrature data (QDataBoundaryGet, src/qdata.c)
    CeedQFunction qf_setup;
    CeedOperator  op_setup;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kSetupBoundaryLoc, &qf_setup));
    CHK(p, CeedQFunctionAddInput(qf_setup, "dx", 6, CEED_EVAL_GRAD));
    CHK(p, CeedQFunctionAddInput(qf_setup, "weight", 1, CEED_EVAL_WEIGHT));
    CHK(p, CeedQFunctionAddOutput(qf_setup, "surface qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedElemRestrictionCreateStrided(p->ceed, nfaces, Q2, 10, (CeedSize)nfaces * Q2 * 10, CEED_STRIDES_BACKEND, &b->restr_qd));
    CHK(p, CeedElemRestrictionCreateVector(b->restr_qd, &b->q_data, NULL));
    CHK(p, CeedOperatorCreate(p->ceed, qf_setup, NULL, NULL, &op_setup));
    CHK(p, CeedOperatorSetField(op_setup, "dx", b->restr_x, b->basis_x, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op_setup, "weight", CEED_ELEMRESTRICTION_NONE, b->basis_x, CEED_VECTOR_NONE));
    CHK(p, CeedOperatorSetField(op_setup, "surface qdata", b->restr_qd, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorApply(op_setup, b->x_coord, b->q_data, CEED_REQUEST_IMMEDIATE));
    CHK(p, CeedOperatorDestroy(&op_setup));
    CHK(p, CeedQFunctionDestroy(&qf_setup));
  }
  const int implicit = p->ctx->is_implicit;
  if (implicit) {
    CHK(p, CeedElemRestrictionCreateStrided(p->ceed, nfaces, Q2, b->njac, (CeedSize)nfaces * Q2 * b->njac, CEED_STRIDES_BACKEND, &b->restr_jd));
    CHK(p, CeedElemRestrictionCreateVector(b->restr_jd, &b->jac_data, NULL));
  }
  CeedQFunctionContext qfctx = p->newt_qfctx;
  if (kind == 3 || kind == 4) {  // OutflowContext_ (bc_freestream_type.h:14-21), filled as problems/bc_outflow.c:154-160 does
    CeedMiOutflowContext *oc = malloc(sizeof(*oc));
    oc->newt_ctx             = *p->ctx;
    oc->pressure             = S_inf_Y[0];
    oc->temperature          = S_inf_Y[1];
    oc->recirc               = S_inf_Y[2];
    oc->softplus_velocity    = S_inf_Y[3];
    CHK(p, CeedQFunctionContextCreate(p->ceed, &b->fs_qfctx));
    CHK(p, CeedQFunctionContextSetData(b->fs_qfctx, CEED_MEM_HOST, CEED_USE_POINTER, sizeof(*oc), oc));
    CHK(p, CeedQFunctionContextSetDataDestroy(b->fs_qfctx, CEED_MEM_HOST, FreeContext));
    qfctx = b->fs_qfctx;
  } else if (kind == 1 || kind == 2) {  // FreestreamContext_ (bc_freestream_type.h:9-12), filled as problems/bc_freestream.c:156-163 does
    CeedMiFreestreamContext *fs = malloc(sizeof(*fs));
    fs->newt_ctx                = *p->ctx;
    StateFromPrimitive(&p->ctx->gas, S_inf_Y, &fs->S_infty);
    CHK(p, CeedQFunctionContextCreate(p->ceed, &b->fs_qfctx));
    CHK(p, CeedQFunctionContextSetData(b->fs_qfctx, CEED_MEM_HOST, CEED_USE_POINTER, sizeof(*fs), fs));
    CHK(p, CeedQFunctionContextSetDataDestroy(b->fs_qfctx, CEED_MEM_HOST, FreeContext));
    qfctx = b->fs_qfctx;
  }
  char loc_f[128], loc_j[128];
  if (kind == 0) {
    snprintf(loc_f, sizeof(loc_f), "qfunctions/newtonian.h:BoundaryIntegral_%s", kVarName[p->state_var]);
    snprintf(loc_j, sizeof(loc_j), "qfunctions/newtonian.h:BoundaryIntegral_Jacobian_%s", kVarName[p->state_var]);
  } else if (kind == 3 || kind == 4) {  // problems/bc_outflow.c:26-112
    const char *name = kind == 3 ? "RiemannOutflow" : "PressureOutflow";
    snprintf(loc_f, sizeof(loc_f), "qfunctions/bc_outflow.h:%s_%s", name, kVarName[p->state_var]);
    snprintf(loc_j, sizeof(loc_j), "qfunctions/bc_outflow.h:%s_Jacobian_%s", name, kVarName[p->state_var]);
  } else if (kind == 5) {  // problems/bc_slip.c:16-57
    snprintf(loc_f, sizeof(loc_f), "qfunctions/bc_slip.h:Slip_%s", kVarName[p->state_var]);
    snprintf(loc_j, sizeof(loc_j), "qfunctions/bc_slip.h:Slip_Jacobian_%s", kVarName[p->state_var]);
  } else {
    const char *flux = kind == 1 ? "HLL" : "HLLC";
    snprintf(loc_f, sizeof(loc_f), "qfunctions/bc_freestream.h:Freestream_%s_%s", kVarName[p->state_var], flux);
    snprintf(loc_j, sizeof(loc_j), "qfunctions/bc_freestream.h:Freestream_Jacobian_%s_%s", kVarName[p->state_var], flux);
  }
  {  // IFunction face operator (src/boundary_condition.c:300-310, 400-410)
    CeedQFunction qf;
    CeedOperator  op;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, loc_f, &qf));
    CHK(p, CeedQFunctionSetContext(qf, qfctx));
    CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "Grad_q", 10, CEED_EVAL_GRAD));
    CHK(p, CeedQFunctionAddInput(qf, "surface qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "x", 3, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddOutput(qf, "v", 5, CEED_EVAL_INTERP));
    if (implicit) CHK(p, CeedQFunctionAddOutput(qf, "surface jacobian data", b->njac, CEED_EVAL_NONE));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &op));
    CHK(p, CeedOperatorSetField(op, "q", b->restr_q, b->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op, "Grad_q", b->restr_q, b->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op, "surface qdata", b->restr_qd, CEED_BASIS_NONE, b->q_data));
    CHK(p, CeedOperatorSetField(op, "x", b->restr_x, b->basis_x, b->x_coord));
    CHK(p, CeedOperatorSetField(op, "v", b->restr_q, b->basis_q, CEED_VECTOR_ACTIVE));
    if (implicit) CHK(p, CeedOperatorSetField(op, "surface jacobian data", b->restr_jd, CEED_BASIS_NONE, b->jac_data));
    CHK(p, CeedOperatorCompositeAddSub(p->op_ifunction, op));
    if (p->op_rhs) CHK(p, CeedOperatorCompositeAddSub(p->op_rhs, op));  // explicit: BCDefinitionAddOperators(op_rhs)
    CHK(p, CeedOperatorDestroy(&op));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  if (implicit) {  // IJacobian face operator (src/boundary_condition.c:350-360, 475-485)
    CeedQFunction qf;
    CeedOperator  op;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, loc_j, &qf));
    CHK(p, CeedQFunctionSetContext(qf, qfctx));
    CHK(p, CeedQFunctionAddInput(qf, "dq", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "Grad_dq", 10, CEED_EVAL_GRAD));
    CHK(p, CeedQFunctionAddInput(qf, "surface qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "x", 3, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "surface jacobian data", b->njac, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddOutput(qf, "v", 5, CEED_EVAL_INTERP));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &op));
    CHK(p, CeedOperatorSetField(op, "dq", b->restr_q, b->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op, "Grad_dq", b->restr_q, b->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(op, "surface qdata", b->restr_qd, CEED_BASIS_NONE, b->q_data));
    CHK(p, CeedOperatorSetField(op, "x", b->restr_x, b->basis_x, b->x_coord));
    CHK(p, CeedOperatorSetField(op, "surface jacobian data", b->restr_jd, CEED_BASIS_NONE, b->jac_data));
    CHK(p, CeedOperatorSetField(op, "v", b->restr_q, b->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorCompositeAddSub(p->op_ijacobian, op));
    CHK(p, CeedOperatorDestroy(&op));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  return 0;
}

int HoneeMiProblemAddBoundary(HoneeMiProblem p, int kind, int nfaces, const int *face_nodes, const int *face_xnodes, const double *S_inf_Y) {
  if (kind > 0 && kind != 5 && !S_inf_Y) return CEED_ERROR_INCOMPLETE;
  return AddBoundary(p, kind, nfaces, face_nodes, face_xnodes, S_inf_Y);
}

int HoneeMiGetBoundaryData(HoneeMiProblem p, int boundary, int which, double *out) {
  if (boundary < 0 || boundary >= p->nbnd) return CEED_ERROR_DIMENSION;
  HoneeMiBoundary *b = &p->bnd[boundary];
  CeedVector       v = which == 0 ? b->q_data : b->jac_data;
  if (!v) return CEED_ERROR_INCOMPLETE;
  const CeedScalar *a;
  CeedSize          n;
  CHK(p, CeedVectorGetLength(v, &n));
  CHK(p, CeedVectorGetArrayRead(v, CEED_MEM_HOST, &a));
  memcpy(out, a, sizeof(double) * (size_t)n);  // backend layout [face][comp][qpt]
  CHK(p, CeedVectorRestoreArrayRead(v, &a));
  return 0;
}

// ---------------------------------------------------------------------------------------------------------------------
// Data-driven subgrid-stress model (problems/sgs_dd_model.c, src/velocity_gradient_projection.c,
// src/grid_anisotropy_tensor.c, src/inverse_multiplicity.c).  The SGS DM of the reference is a clone of the solution
// DM with 6 components; on one rank its nodes are the solution nodes, so each nodal field here has the solution
// offsets scaled by its component count and the solution basis tables.
// ---------------------------------------------------------------------------------------------------------------------
static const char *const kVelocityGradientRHSLoc[3] = {"qfunctions/velocity_gradient_projection.h:VelocityGradientProjectionRHS_Conserv",
                                                       "qfunctions/velocity_gradient_projection.h:VelocityGradientProjectionRHS_Prim",
                                                       "qfunctions/velocity_gradient_projection.h:VelocityGradientProjectionRHS_Entropy"};
static const char *const kSgsNodalLoc[3]            = {"qfunctions/sgs_dd_model.h:ComputeSgsDDNodal_Conserv", "qfunctions/sgs_dd_model.h:ComputeSgsDDNodal_Prim",
                                                       "qfunctions/sgs_dd_model.h:ComputeSgsDDNodal_Entropy"};
static const char *const kSgsApplyLoc[3]            = {"qfunctions/sgs_dd_model.h:IFunction_NodalSgs_Conserv", "qfunctions/sgs_dd_model.h:IFunction_NodalSgs_Prim",
                                                       "qfunctions/sgs_dd_model.h:IFunction_NodalSgs_Entropy"};

// Nodal restriction and basis of an ncomp-component field on the solution nodes (DMPlexCeedElemRestrictionCreate and
// DMPlexCeedBasisCreate of a cloned DM)
static int NodalField(HoneeMiProblem p, int ncomp, CeedElemRestriction *restr, CeedBasis *basis) {
  const HoneeMiMesh *m  = &p->mesh;
  const int          P3 = m->P * m->P * m->P;
  CeedInt           *off = malloc(sizeof(CeedInt) * (size_t)m->nelem * P3);
  for (size_t i = 0; i < (size_t)m->nelem * P3; i++) off[i] = ncomp * m->offsets[i];
  int ierr = CeedElemRestrictionCreate(p->ceed, m->nelem, P3, ncomp, 1, ncomp * (CeedSize)m->nnodes, CEED_MEM_HOST, CEED_COPY_VALUES, off, restr);
  free(off);
  CHK(p, ierr);
  if (basis) CHK(p, CeedBasisCreateTensorH1(p->ceed, 3, ncomp, m->P, m->Q, m->B, m->D, m->qref, m->qweight, basis));
  return 0;
}

// HoneeMassQFunctionCreate (src/setuplibceed.c) with Mass_N on a nodal field
static int MassOperator(HoneeMiProblem p, int ncomp, CeedElemRestriction restr, CeedBasis basis, CeedOperator *op) {
  static const char *const kMassLoc[10] = {NULL, "qfunctions/mass.h:Mass_1", "qfunctions/mass.h:Mass_2", "qfunctions/mass.h:Mass_3",
                                           "qfunctions/mass.h:Mass_4", "qfunctions/mass.h:Mass_5", "qfunctions/mass.h:Mass_6",
                                           "qfunctions/mass.h:Mass_7", NULL, "qfunctions/mass.h:Mass_9"};
  CeedQFunction qf;
  if (ncomp < 1 || ncomp > 9 || !kMassLoc[ncomp]) return CEED_ERROR_UNSUPPORTED;
  CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kMassLoc[ncomp], &qf));
  CHK(p, CeedQFunctionAddInput(qf, "u", ncomp, CEED_EVAL_INTERP));
  CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
  CHK(p, CeedQFunctionAddOutput(qf, "v", ncomp, CEED_EVAL_INTERP));
  CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, op));
  CHK(p, CeedOperatorSetField(*op, "u", restr, basis, CEED_VECTOR_ACTIVE));
  CHK(p, CeedOperatorSetField(*op, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
  CHK(p, CeedOperatorSetField(*op, "v", restr, basis, CEED_VECTOR_ACTIVE));
  CHK(p, CeedQFunctionDestroy(&qf));
  return 0;
}

// PCJACOBI's inverse diagonal.  PETSc maps a zero diagonal entry to 1; the mass diagonals and row sums here are
// strictly positive (positive w detJ), where that rule and CeedVectorReciprocal agree.
static int JacobiInverse(HoneeMiProblem p, CeedVector d) {
  CHK(p, CeedVectorReciprocal(d));
  return 0;
}

// VelocityGradientProjectionSetup (src/velocity_gradient_projection.c:36-130): RHS operator and the lumped mass
// (PCJACOBI row sum of Mass_9, KSPPREONLY), i.e. G = R ./ (M 1)
// The velocity-gradient RHS operator on a 5-component state field (restriction/basis), state variables var
static int VelocityGradientRHSOperator(HoneeMiProblem p, int var, CeedElemRestriction restr, CeedBasis basis, CeedOperator *op) {
  CeedQFunction qf;
  CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kVelocityGradientRHSLoc[var], &qf));
  CHK(p, CeedQFunctionSetContext(qf, p->newt_qfctx));
  CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_INTERP));
  CHK(p, CeedQFunctionAddInput(qf, "Grad_q", 15, CEED_EVAL_GRAD));
  CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
  CHK(p, CeedQFunctionAddInput(qf, "x", 3, CEED_EVAL_INTERP));
  CHK(p, CeedQFunctionAddOutput(qf, "velocity gradient", 9, CEED_EVAL_INTERP));
  CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, op));
  CHK(p, CeedOperatorSetField(*op, "q", restr, basis, CEED_VECTOR_ACTIVE));
  CHK(p, CeedOperatorSetField(*op, "Grad_q", restr, basis, CEED_VECTOR_ACTIVE));
  CHK(p, CeedOperatorSetField(*op, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
  CHK(p, CeedOperatorSetField(*op, "x", p->elem_restr_x, p->basis_x, p->x_coord));
  CHK(p, CeedOperatorSetField(*op, "velocity gradient", p->elem_restr_grad_velo, p->basis_grad_velo, CEED_VECTOR_ACTIVE));
  CHK(p, CeedQFunctionDestroy(&qf));
  return 0;
}

static int VelocityGradientProjectionSetup(HoneeMiProblem p) {
  if (p->grad_velo) return 0;
  CHK(p, NodalField(p, 9, &p->elem_restr_grad_velo, &p->basis_grad_velo));
  CHK(p, VelocityGradientRHSOperator(p, p->state_var, p->elem_restr_q, p->basis_q, &p->op_grad_velo_rhs));
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_grad_velo, &p->grad_velo, NULL));
  {  // PCJacobiSetType(PC_JACOBI_ROWSUM): MatGetRowSum = M 1
    CeedOperator op_mass;
    CeedVector   ones;
    CHK(p, MassOperator(p, 9, p->elem_restr_grad_velo, p->basis_grad_velo, &op_mass));
    CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_grad_velo, &ones, NULL));
    CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_grad_velo, &p->grad_velo_inv_rowsum, NULL));
    CHK(p, CeedVectorSetValue(ones, 1.0));
    CHK(p, CeedOperatorApply(op_mass, ones, p->grad_velo_inv_rowsum, CEED_REQUEST_IMMEDIATE));
    CHK(p, JacobiInverse(p, p->grad_velo_inv_rowsum));
    CHK(p, CeedVectorDestroy(&ones));
    CHK(p, CeedOperatorDestroy(&op_mass));
  }
  return 0;
}

// KSPCG with PCJACOBI (inverse diagonal dinv) and KSP_NORM_NATURAL from a zero initial guess: converged when
// sqrt(r.z) <= max(rtol sqrt(r0.z0), atol), at most max_it iterations (KSPSolve_CG recurrences).  work holds three
// vectors of the system size.
static int CgJacobi(HoneeMiProblem p, CeedOperator op, CeedVector dinv, CeedVector R, CeedVector X, CeedVector work[3], CeedVector Ap, double rtol,
                    double atol, int max_it, int *iterations) {
  CeedVector r = work[0], z = work[1], pv = work[2];
  double     rz, pAp, rz_new;
  CHK(p, CeedVectorSetValue(X, 0.0));
  CHK(p, CeedVectorCopy(R, r));
  CHK(p, CeedVectorPointwiseMult(z, dinv, r));
  CHK(p, CeedVectorCopy(z, pv));
  CHK(p, CeedMiVectorDot(r, z, &rz));
  const double ttol = fmax(rtol * sqrt(fabs(rz)), atol);
  int          its  = 0;
  while (sqrt(fabs(rz)) > ttol && its < max_it) {
    CHK(p, CeedOperatorApply(op, pv, Ap, CEED_REQUEST_IMMEDIATE));
    CHK(p, CeedMiVectorDot(pv, Ap, &pAp));
    const double alpha = rz / pAp;
    CHK(p, CeedVectorAXPY(X, alpha, pv));
    CHK(p, CeedVectorAXPY(r, -alpha, Ap));
    CHK(p, CeedVectorPointwiseMult(z, dinv, r));
    CHK(p, CeedMiVectorDot(r, z, &rz_new));
    const double beta = rz_new / rz;
    CHK(p, CeedVectorAXPBY(pv, 1.0, beta, z));
    rz = rz_new;
    its++;
  }
  *iterations = its;
  return 0;
}

// GridAnisotropyTensorProjectionSetupApply (src/grid_anisotropy_tensor.c:14-111): R = AnisotropyTensorProjection,
// KSPCG with PCJACOBI (diagonal of Mass_7), natural norm, rtol 1e-10, zero initial guess.
static int GridAnisotropyTensorProjection(HoneeMiProblem p) {
  if (p->grid_aniso) return 0;
  CeedQFunction qf;
  CeedOperator  op_rhs, op_mass;
  CeedVector    R, dinv, Ap, work[3];
  CHK(p, NodalField(p, 7, &p->elem_restr_grid_aniso, &p->basis_grid_aniso));
  CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, "qfunctions/grid_anisotropy_tensor.h:AnisotropyTensorProjection", &qf));
  CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
  CHK(p, CeedQFunctionAddOutput(qf, "v", 7, CEED_EVAL_INTERP));
  CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &op_rhs));
  CHK(p, CeedOperatorSetField(op_rhs, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
  CHK(p, CeedOperatorSetField(op_rhs, "v", p->elem_restr_grid_aniso, p->basis_grid_aniso, CEED_VECTOR_ACTIVE));
  CHK(p, CeedQFunctionDestroy(&qf));
  CHK(p, MassOperator(p, 7, p->elem_restr_grid_aniso, p->basis_grid_aniso, &op_mass));
  CeedVector *vs[7] = {&R, &dinv, &Ap, &work[0], &work[1], &work[2], &p->grid_aniso};
  for (int i = 0; i < 7; i++) CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_grid_aniso, vs[i], NULL));
  CHK(p, CeedOperatorApply(op_rhs, CEED_VECTOR_NONE, R, CEED_REQUEST_IMMEDIATE));
  CHK(p, CeedOperatorLinearAssembleDiagonal(op_mass, dinv, CEED_REQUEST_IMMEDIATE));
  CHK(p, JacobiInverse(p, dinv));
  CHK(p, CgJacobi(p, op_mass, dinv, R, p->grid_aniso, work, Ap, 1e-10, 1e-50, 10000, &p->grid_aniso_its));
  for (int i = 0; i < 6; i++) CHK(p, CeedVectorDestroy(vs[i]));
  CHK(p, CeedOperatorDestroy(&op_rhs));
  CHK(p, CeedOperatorDestroy(&op_mass));
  return 0;
}

// GetInverseMultiplicity (src/inverse_multiplicity.c:24-80): element-node (collocated) 1 / multiplicity of the nodes
// of restr_src (every nodal field here shares the solution nodes, so one vector serves all of them)
static int InverseMultiplicity(HoneeMiProblem p, CeedElemRestriction restr_src) {
  if (p->inv_mult) return 0;
  const int     P3 = p->mesh.P * p->mesh.P * p->mesh.P;
  CeedQFunction qf;
  CeedOperator  op;
  CeedVector    mult;
  CeedInt       ncomp;
  CHK(p, CeedElemRestrictionGetNumComponents(restr_src, &ncomp));
  CHK(p, CeedElemRestrictionCreateStrided(p->ceed, p->nelem, P3, 1, (CeedSize)p->nelem * P3, CEED_STRIDES_BACKEND, &p->elem_restr_inv_mult));
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_inv_mult, &p->inv_mult, NULL));
  CHK(p, CeedElemRestrictionCreateVector(restr_src, &mult, NULL));
  CHK(p, CeedElemRestrictionGetMultiplicity(restr_src, mult));
  CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, "qfunctions/inverse_multiplicity.h:InverseMultiplicity", &qf));
  CHK(p, CeedQFunctionAddInput(qf, "multiplicity", ncomp, CEED_EVAL_NONE));
  CHK(p, CeedQFunctionAddOutput(qf, "inverse multiplicity", 1, CEED_EVAL_NONE));
  CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &op));
  CHK(p, CeedOperatorSetName(op, "InverseMultiplicity"));
  CHK(p, CeedOperatorSetField(op, "multiplicity", restr_src, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
  CHK(p, CeedOperatorSetField(op, "inverse multiplicity", p->elem_restr_inv_mult, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
  CHK(p, CeedOperatorApply(op, mult, p->inv_mult, CEED_REQUEST_IMMEDIATE));
  CHK(p, CeedVectorDestroy(&mult));
  CHK(p, CeedOperatorDestroy(&op));
  CHK(p, CeedQFunctionDestroy(&qf));
  return 0;
}

static int FreeSgsContext(void *data) {
  free(data);
  return 0;
}

// SgsDDContextFill (problems/sgs_dd_model.c:602-650): parameters at the reference's offsets, with w1.dat
// [input][neuron] and w2.dat [neuron][output] transposed to [neuron][input] and [output][neuron]
static int SgsDDContextCreate(HoneeMiProblem p, const HoneeMiSgsDDParams *prm) {
  if ((prm->wh == NULL) != (prm->bh == NULL)) return CEED_ERROR_MINOR;  // a second hidden layer needs both arrays
  const int    ni = 6, no = 6, nn = 20, layers = prm->wh ? 2 : 1;
  const size_t bias1 = 0, bias2 = bias1 + nn, weight1 = bias2 + nn, weight2 = weight1 + (size_t)nn * ni, out_scaling = weight2 + (size_t)ni * nn;
  const size_t hidden2       = out_scaling + 2 * (size_t)no;  // second hidden layer (extension): biases, then weights
  const size_t total_scalars = hidden2 + (layers == 2 ? (size_t)nn + (size_t)nn * nn : 0);
  const size_t total_bytes   = sizeof(CeedMiSgsDDContext) + total_scalars * sizeof(double);
  CeedMiSgsDDContext *c      = calloc(1, total_bytes);
  c->num_inputs              = ni;
  c->num_outputs             = no;
  c->num_layers              = layers;
  c->num_neurons             = nn;
  c->alpha                   = prm->alpha;
  c->newt_ctx                = *p->ctx;
  c->offsets.bias1           = bias1;
  c->offsets.bias2           = bias2;
  c->offsets.weight1         = weight1;
  c->offsets.weight2         = weight2;
  c->offsets.out_scaling     = out_scaling;
  c->total_bytes             = total_bytes;
  memcpy(&c->data[bias1], prm->b1, sizeof(double) * nn);
  memcpy(&c->data[bias2], prm->b2, sizeof(double) * no);
  memcpy(&c->data[out_scaling], prm->out_scaling, sizeof(double) * 2 * no);
  for (int i = 0; i < ni; i++)  // TransposeMatrix(w1, N = num_inputs, M = num_neurons)
    for (int j = 0; j < nn; j++) c->data[weight1 + (size_t)j * ni + i] = prm->w1[(size_t)i * nn + j];
  for (int i = 0; i < nn; i++)  // TransposeMatrix(w2, N = num_neurons, M = num_outputs)
    for (int j = 0; j < no; j++) c->data[weight2 + (size_t)j * nn + i] = prm->w2[(size_t)i * no + j];
  if (layers == 2) {
    memcpy(&c->data[hidden2], prm->bh, sizeof(double) * nn);
    for (int i = 0; i < nn; i++)  // [first-layer neuron][second-layer neuron] -> [second][first]
      for (int j = 0; j < nn; j++) c->data[hidden2 + nn + (size_t)j * nn + i] = prm->wh[(size_t)i * nn + j];
  }
  CHK(p, CeedQFunctionContextCreate(p->ceed, &p->sgsdd_qfctx));
  CHK(p, CeedQFunctionContextSetData(p->sgsdd_qfctx, CEED_MEM_HOST, CEED_USE_POINTER, total_bytes, c));
  CHK(p, CeedQFunctionContextSetDataDestroy(p->sgsdd_qfctx, CEED_MEM_HOST, FreeSgsContext));
  return 0;
}

// SgsDDSetup (problems/sgs_dd_model.c:652-737) with the fused nodal evaluation (SgsDDSetupNodalEvaluation_Fused,
// :114-203) and the nodal IFunction (SgsSetupNodalIFunction, :498-558)
int HoneeMiProblemSetupSgsDD(HoneeMiProblem p, const HoneeMiSgsDDParams *prm) {
  if (p->sgs_enabled) return CEED_ERROR_MINOR;
  CHK(p, VelocityGradientProjectionSetup(p));
  CHK(p, SgsDDContextCreate(p, prm));
  CHK(p, GridAnisotropyTensorProjection(p));
  CHK(p, NodalField(p, 6, &p->elem_restr_sgs, &p->basis_sgs));
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_sgs, &p->sgs_nodal, NULL));
  CHK(p, InverseMultiplicity(p, p->elem_restr_sgs));
  CHK(p, CeedBasisCreateProjection(p->basis_x, p->basis_q, &p->basis_x_to_q));
  {
    CeedQFunction qf;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kSgsNodalLoc[p->state_var], &qf));
    CHK(p, CeedQFunctionSetContext(qf, p->sgsdd_qfctx));
    CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "x", 3, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "gradient velocity", 9, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "anisotropy tensor", 7, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "inverse multiplicity", 1, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddOutput(qf, "km_sgs", 6, CEED_EVAL_NONE));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &p->op_sgs_nodal));
    CHK(p, CeedOperatorSetField(p->op_sgs_nodal, "q", p->elem_restr_q, CEED_BASIS_NONE, p->q_ceed));
    CHK(p, CeedOperatorSetField(p->op_sgs_nodal, "x", p->elem_restr_x, p->basis_x_to_q, p->x_coord));
    CHK(p, CeedOperatorSetField(p->op_sgs_nodal, "gradient velocity", p->elem_restr_grad_velo, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(p->op_sgs_nodal, "anisotropy tensor", p->elem_restr_grid_aniso, CEED_BASIS_NONE, p->grid_aniso));
    CHK(p, CeedOperatorSetField(p->op_sgs_nodal, "inverse multiplicity", p->elem_restr_inv_mult, CEED_BASIS_NONE, p->inv_mult));
    CHK(p, CeedOperatorSetField(p->op_sgs_nodal, "km_sgs", p->elem_restr_sgs, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  {
    CeedQFunction qf;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kSgsApplyLoc[p->state_var], &qf));
    CHK(p, CeedQFunctionSetContext(qf, p->newt_qfctx));
    CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "km_sgs", 6, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddOutput(qf, "Grad_v", 15, CEED_EVAL_GRAD));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &p->op_sgs_apply));
    CHK(p, CeedOperatorSetField(p->op_sgs_apply, "q", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(p->op_sgs_apply, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
    CHK(p, CeedOperatorSetField(p->op_sgs_apply, "km_sgs", p->elem_restr_sgs, p->basis_sgs, p->sgs_nodal));
    CHK(p, CeedOperatorSetField(p->op_sgs_apply, "Grad_v", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  p->sgs_enabled = 1;
  return 0;
}

static int CopyVectorOut(HoneeMiProblem p, CeedVector v, int mem_type, double *out);

// ---- Differential filter (src/differential_filter.c:17-305) ----
static const char *const kFilterRHSLoc[3] = {"qfunctions/differential_filter.h:DifferentialFilter_RHS_Conserv",
                                             "qfunctions/differential_filter.h:DifferentialFilter_RHS_Prim",
                                             "qfunctions/differential_filter.h:DifferentialFilter_RHS_Entropy"};

// DifferentialFilterSetup + DifferentialFilterCreateOperators: the filtered field holds per node the primitive state
// (field 0, 5 components) then the velocity products (field 1, 6 components); the LHS is the composite of
// DifferentialFilter_LHS_5 and _6 over those fields with the collocated grid anisotropy tensor; the solve is KSPCG +
// PCJACOBI (diagonal), natural norm.
int HoneeMiProblemSetupDiffFilter(HoneeMiProblem p, const HoneeMiDiffFilterOptions *o) {
  if (p->filter_enabled) return CEED_ERROR_MINOR;
  const HoneeMiMesh *m = &p->mesh;
  const int          P3 = m->P * m->P * m->P, ncomps[2] = {5, 6}, field_off[2] = {0, 5};
  for (int f = 0; f < 2; f++) {
    CeedInt *off = malloc(sizeof(CeedInt) * (size_t)m->nelem * P3);
    for (size_t i = 0; i < (size_t)m->nelem * P3; i++) off[i] = 11 * m->offsets[i] + field_off[f];
    int ierr = CeedElemRestrictionCreate(p->ceed, m->nelem, P3, ncomps[f], 1, 11 * (CeedSize)m->nnodes, CEED_MEM_HOST, CEED_COPY_VALUES, off,
                                         &p->elem_restr_filter[f]);
    free(off);
    CHK(p, ierr);
    CHK(p, CeedBasisCreateTensorH1(p->ceed, 3, ncomps[f], m->P, m->Q, m->B, m->D, m->qref, m->qweight, &p->basis_filter[f]));
  }
  {  // context (src/differential_filter.c:236-275), units of 1 meter
    CeedMiDiffFilterContext *c = calloc(1, sizeof(*c));
    c->grid_based_width        = o->grid_based_width != 0;
    for (int i = 0; i < 3; i++) c->width_scaling[i] = o->width_scaling[i];
    c->kernel_scaling   = o->kernel_scaling;
    c->friction_length  = o->friction_length;
    c->damping_function = o->damping_function;
    c->damping_constant = o->damping_constant;
    c->newt_ctx         = *p->ctx;
    CHK(p, CeedQFunctionContextCreate(p->ceed, &p->diff_filter_qfctx));
    CHK(p, CeedQFunctionContextSetData(p->diff_filter_qfctx, CEED_MEM_HOST, CEED_USE_POINTER, sizeof(*c), c));
    CHK(p, CeedQFunctionContextSetDataDestroy(p->diff_filter_qfctx, CEED_MEM_HOST, FreeContext));
    CHK(p, CeedQFunctionContextRegisterDouble(p->diff_filter_qfctx, "filter width scaling", offsetof(CeedMiDiffFilterContext, width_scaling), 3,
                                              "Filter width scaling"));
  }
  {  // RHS operator
    CeedQFunction qf;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kFilterRHSLoc[p->state_var], &qf));
    CHK(p, CeedQFunctionSetContext(qf, p->diff_filter_qfctx));
    CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "x", 3, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddOutput(qf, "v0", 5, CEED_EVAL_INTERP));
    CHK(p, CeedQFunctionAddOutput(qf, "v1", 6, CEED_EVAL_INTERP));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &p->op_filter_rhs));
    CHK(p, CeedOperatorSetField(p->op_filter_rhs, "q", p->elem_restr_q, p->basis_q, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(p->op_filter_rhs, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
    CHK(p, CeedOperatorSetField(p->op_filter_rhs, "x", p->elem_restr_x, p->basis_x, p->x_coord));
    CHK(p, CeedOperatorSetField(p->op_filter_rhs, "v0", p->elem_restr_filter[0], p->basis_filter[0], CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(p->op_filter_rhs, "v1", p->elem_restr_filter[1], p->basis_filter[1], CEED_VECTOR_ACTIVE));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  {  // GridAnisotropyTensorCalculateCollocatedVector (src/grid_anisotropy_tensor.c:114-160)
    CeedQFunction qf;
    CeedOperator  op;
    CHK(p, CeedElemRestrictionCreateStrided(p->ceed, p->nelem, p->Q3, 7, (CeedSize)p->nelem * p->Q3 * 7, CEED_STRIDES_BACKEND, &p->elem_restr_aniso_col));
    CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_aniso_col, &p->aniso_col, NULL));
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, "qfunctions/grid_anisotropy_tensor.h:AnisotropyTensorCollocate", &qf));
    CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddOutput(qf, "v", 7, CEED_EVAL_NONE));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &op));
    CHK(p, CeedOperatorSetField(op, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
    CHK(p, CeedOperatorSetField(op, "v", p->elem_restr_aniso_col, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorApply(op, CEED_VECTOR_NONE, p->aniso_col, CEED_REQUEST_IMMEDIATE));
    CHK(p, CeedOperatorDestroy(&op));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  {  // LHS composite
    static const char *const kLHSLoc[2] = {"qfunctions/differential_filter.h:DifferentialFilter_LHS_5",
                                           "qfunctions/differential_filter.h:DifferentialFilter_LHS_6"};
    CHK(p, CeedOperatorCreateComposite(p->ceed, &p->op_filter_lhs));
    for (int f = 0; f < 2; f++) {
      CeedQFunction qf;
      CeedOperator  sub;
      const int     N = ncomps[f];
      CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, kLHSLoc[f], &qf));
      CHK(p, CeedQFunctionSetContext(qf, p->diff_filter_qfctx));
      CHK(p, CeedQFunctionSetUserFlopsEstimate(qf, 0));
      CHK(p, CeedQFunctionAddInput(qf, "q", N, CEED_EVAL_INTERP));
      CHK(p, CeedQFunctionAddInput(qf, "Grad_q", 3 * N, CEED_EVAL_GRAD));
      CHK(p, CeedQFunctionAddInput(qf, "anisotropy tensor", 7, CEED_EVAL_NONE));
      CHK(p, CeedQFunctionAddInput(qf, "x", 3, CEED_EVAL_INTERP));
      CHK(p, CeedQFunctionAddInput(qf, "qdata", 10, CEED_EVAL_NONE));
      CHK(p, CeedQFunctionAddOutput(qf, "v", N, CEED_EVAL_INTERP));
      CHK(p, CeedQFunctionAddOutput(qf, "Grad_v", 3 * N, CEED_EVAL_GRAD));
      CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &sub));
      CHK(p, CeedOperatorSetField(sub, "q", p->elem_restr_filter[f], p->basis_filter[f], CEED_VECTOR_ACTIVE));
      CHK(p, CeedOperatorSetField(sub, "Grad_q", p->elem_restr_filter[f], p->basis_filter[f], CEED_VECTOR_ACTIVE));
      CHK(p, CeedOperatorSetField(sub, "anisotropy tensor", p->elem_restr_aniso_col, CEED_BASIS_NONE, p->aniso_col));
      CHK(p, CeedOperatorSetField(sub, "x", p->elem_restr_x, p->basis_x, p->x_coord));
      CHK(p, CeedOperatorSetField(sub, "qdata", p->elem_restr_qd, CEED_BASIS_NONE, p->q_data));
      CHK(p, CeedOperatorSetField(sub, "v", p->elem_restr_filter[f], p->basis_filter[f], CEED_VECTOR_ACTIVE));
      CHK(p, CeedOperatorSetField(sub, "Grad_v", p->elem_restr_filter[f], p->basis_filter[f], CEED_VECTOR_ACTIVE));
      CHK(p, CeedOperatorCompositeAddSub(p->op_filter_lhs, sub));
      CHK(p, CeedOperatorDestroy(&sub));
      CHK(p, CeedQFunctionDestroy(&qf));
    }
    CHK(p, CeedOperatorGetContextFieldLabel(p->op_filter_lhs, "filter width scaling", &p->filter_width_label));
  }
  CeedVector *vs[6] = {&p->filter_rhs, &p->filtered, &p->filter_dinv, &p->cg_work[0], &p->cg_work[1], &p->cg_work[2]};
  for (int i = 0; i < 6; i++) CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_filter[0], vs[i], NULL));
  p->filter_rtol    = o->ksp_rtol > 0 ? o->ksp_rtol : 1e-10;
  p->filter_atol    = o->ksp_atol > 0 ? o->ksp_atol : 1e-50;
  p->filter_max_it  = o->ksp_max_it > 0 ? o->ksp_max_it : 10000;
  p->filter_enabled = 1;
  return 0;
}

// DifferentialFilterApply (src/differential_filter.c:283-305): q_ceed holds the state; result in p->filtered.  The
// Jacobi preconditioner is set up at the first solve and kept: a later change of the filter width through the
// operator context does not set up the KSP's preconditioner again.
static int DiffFilterApply(HoneeMiProblem p) {
  CHK(p, CeedOperatorApply(p->op_filter_rhs, p->q_ceed, p->filter_rhs, CEED_REQUEST_IMMEDIATE));
  if (!p->filter_pc_ready) {
    CHK(p, CeedOperatorLinearAssembleDiagonal(p->op_filter_lhs, p->filter_dinv, CEED_REQUEST_IMMEDIATE));
    CHK(p, JacobiInverse(p, p->filter_dinv));
    p->filter_pc_ready = 1;
  }
  CeedVector ApV;
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_filter[0], &ApV, NULL));
  int ierr = CgJacobi(p, p->op_filter_lhs, p->filter_dinv, p->filter_rhs, p->filtered, p->cg_work, ApV, p->filter_rtol, p->filter_atol,
                      p->filter_max_it, &p->filter_its);
  CeedVectorDestroy(&ApV);
  return ierr;
}

int HoneeMiSetFilterWidthScaling(HoneeMiProblem p, const double scaling[3]) {
  if (!p->filter_enabled) return CEED_ERROR_INCOMPATIBLE;
  CHK(p, CeedOperatorSetContextDouble(p->op_filter_lhs, p->filter_width_label, (double *)scaling));
  return 0;
}

int HoneeMiDiffFilterApply(HoneeMiProblem p, int mem_type, const double *q, double *filtered, int *iterations) {
  CeedScalar *arr;
  if (!p->filter_enabled) return CEED_ERROR_INCOMPATIBLE;
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, (CeedScalar *)q));
  int ierr = DiffFilterApply(p);
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mem_type, &arr));
  CHK(p, ierr);
  if (iterations) *iterations = p->filter_its;
  return CopyVectorOut(p, p->filtered, mem_type, filtered);
}

int HoneeMiDiffFilterRHS(HoneeMiProblem p, int mem_type, const double *q, double *rhs) {
  CeedScalar *arr;
  if (!p->filter_enabled) return CEED_ERROR_INCOMPATIBLE;
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, (CeedScalar *)q));
  int ierr = CeedOperatorApply(p->op_filter_rhs, p->q_ceed, p->filter_rhs, CEED_REQUEST_IMMEDIATE);
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mem_type, &arr));
  CHK(p, ierr);
  return CopyVectorOut(p, p->filter_rhs, mem_type, rhs);
}

// The filter LHS operator and its diagonal on a host array [nnodes][11] (tests)
int HoneeMiDiffFilterLHS(HoneeMiProblem p, const double *u, double *v, double *diag) {
  if (!p->filter_enabled) return CEED_ERROR_INCOMPATIBLE;
  CeedVector uv, vv;
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_filter[0], &uv, NULL));
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_filter[0], &vv, NULL));
  CHK(p, CeedVectorSetArray(uv, CEED_MEM_HOST, CEED_COPY_VALUES, (CeedScalar *)u));
  CHK(p, CeedOperatorApply(p->op_filter_lhs, uv, vv, CEED_REQUEST_IMMEDIATE));
  CHK(p, CopyVectorOut(p, vv, CEED_MEM_HOST, v));
  CHK(p, CeedOperatorLinearAssembleDiagonal(p->op_filter_lhs, vv, CEED_REQUEST_IMMEDIATE));
  CHK(p, CopyVectorOut(p, vv, CEED_MEM_HOST, diag));
  CHK(p, CeedVectorDestroy(&uv));
  CHK(p, CeedVectorDestroy(&vv));
  return 0;
}

// SGS_DD_TrainingSetup (src/smartsim/sgs_dd_training.c:183-282) and SetupTrainingDataCalculation (:93-180): the
// projected grid anisotropy, the velocity-gradient projection of the filtered primitive state, and the nodal
// training-pair operator (12 components per node)
int HoneeMiProblemSetupSgsTraining(HoneeMiProblem p) {
  if (!p->filter_enabled) return CEED_ERROR_INCOMPATIBLE;
  if (p->train_enabled) return CEED_ERROR_MINOR;
  CHK(p, VelocityGradientProjectionSetup(p));
  CHK(p, GridAnisotropyTensorProjection(p));
  CHK(p, NodalField(p, 12, &p->elem_restr_train, NULL));
  CHK(p, CeedElemRestrictionCreateVector(p->elem_restr_train, &p->train_data, NULL));
  CHK(p, InverseMultiplicity(p, p->elem_restr_train));
  CHK(p, VelocityGradientRHSOperator(p, 1, p->elem_restr_filter[0], p->basis_filter[0], &p->op_filtered_grad_velo_rhs));
  {
    CeedMiSgsTrainingContext *c = calloc(1, sizeof(*c));
    c->newt_ctx                 = *p->ctx;
    CHK(p, CeedQFunctionContextCreate(p->ceed, &p->train_qfctx));
    CHK(p, CeedQFunctionContextSetData(p->train_qfctx, CEED_MEM_HOST, CEED_USE_POINTER, sizeof(*c), c));
    CHK(p, CeedQFunctionContextSetDataDestroy(p->train_qfctx, CEED_MEM_HOST, FreeContext));
  }
  {
    CeedQFunction qf;
    CHK(p, CeedQFunctionCreateInterior(p->ceed, 1, HoneeMiQFunctionNotOnHost, "qfunctions/sgs_dd_training.h:ComputeSGS_DDAnisotropicTrainingDataNodal_Prim",
                                       &qf));
    CHK(p, CeedQFunctionSetContext(qf, p->train_qfctx));
    CHK(p, CeedQFunctionAddInput(qf, "q", 5, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "velocity product", 6, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "gradient velocity", 9, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "anisotropy tensor", 7, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddInput(qf, "inverse multiplicity", 1, CEED_EVAL_NONE));
    CHK(p, CeedQFunctionAddOutput(qf, "training data", 12, CEED_EVAL_NONE));
    CHK(p, CeedOperatorCreate(p->ceed, qf, NULL, NULL, &p->op_sgs_train));
    CHK(p, CeedOperatorSetField(p->op_sgs_train, "q", p->elem_restr_filter[0], CEED_BASIS_NONE, p->filtered));
    CHK(p, CeedOperatorSetField(p->op_sgs_train, "velocity product", p->elem_restr_filter[1], CEED_BASIS_NONE, p->filtered));
    CHK(p, CeedOperatorSetField(p->op_sgs_train, "gradient velocity", p->elem_restr_grad_velo, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
    CHK(p, CeedOperatorSetField(p->op_sgs_train, "anisotropy tensor", p->elem_restr_grid_aniso, CEED_BASIS_NONE, p->grid_aniso));
    CHK(p, CeedOperatorSetField(p->op_sgs_train, "inverse multiplicity", p->elem_restr_inv_mult, CEED_BASIS_NONE, p->inv_mult));
    CHK(p, CeedOperatorSetField(p->op_sgs_train, "training data", p->elem_restr_train, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE));
    CHK(p, CeedQFunctionDestroy(&qf));
  }
  p->train_enabled = 1;
  return 0;
}

// TSMonitor_SGS_DD_Training's computation for one filter width (src/smartsim/sgs_dd_training.c:313-349): filter the
// state with the width scaled by width_scale in every direction, project the filtered velocity gradient, evaluate the
// training pairs [nnodes][12] (six model inputs, six filtered SGS stress components)
int HoneeMiSgsTrainingData(HoneeMiProblem p, int mem_type, const double *q, double width_scale, double *out, int *filter_iterations) {
  CeedScalar *arr;
  if (!p->train_enabled) return CEED_ERROR_INCOMPATIBLE;
  const double scaling[3] = {width_scale, width_scale, width_scale};
  CHK(p, HoneeMiSetFilterWidthScaling(p, scaling));
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, (CeedScalar *)q));
  int ierr = DiffFilterApply(p);
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mem_type, &arr));
  CHK(p, ierr);
  if (filter_iterations) *filter_iterations = p->filter_its;
  CHK(p, CeedOperatorApply(p->op_filtered_grad_velo_rhs, p->filtered, p->grad_velo, CEED_REQUEST_IMMEDIATE));
  CHK(p, CeedVectorPointwiseMult(p->grad_velo, p->grad_velo, p->grad_velo_inv_rowsum));
  CHK(p, CeedOperatorApply(p->op_sgs_train, p->grad_velo, p->train_data, CEED_REQUEST_IMMEDIATE));
  return CopyVectorOut(p, p->train_data, mem_type, out);
}

// Fields of the last HoneeMiSgsTrainingData call (host copies): filtered [nnodes][11], projected filtered velocity
// gradient [nnodes][9] and the projected grid anisotropy [nnodes][7]
int HoneeMiSgsTrainingFields(HoneeMiProblem p, double *filtered, double *grad_velo, double *aniso) {
  if (!p->train_enabled) return CEED_ERROR_INCOMPATIBLE;
  if (filtered) CHK(p, CopyVectorOut(p, p->filtered, CEED_MEM_HOST, filtered));
  if (grad_velo) CHK(p, CopyVectorOut(p, p->grad_velo, CEED_MEM_HOST, grad_velo));
  if (aniso) CHK(p, CopyVectorOut(p, p->grid_aniso, CEED_MEM_HOST, aniso));
  return 0;
}

// VelocityGradientProjectionApply (src/velocity_gradient_projection.c:145-155): q_ceed holds the state
static int VelocityGradientProjectionApply(HoneeMiProblem p) {
  CHK(p, CeedOperatorApply(p->op_grad_velo_rhs, p->q_ceed, p->grad_velo, CEED_REQUEST_IMMEDIATE));
  CHK(p, CeedVectorPointwiseMult(p->grad_velo, p->grad_velo, p->grad_velo_inv_rowsum));
  return 0;
}

// SgsDDApplyIFunction (problems/sgs_dd_model.c:561-588): q_ceed holds the state, g_ceed the residual
static int SgsDDApplyIFunction(HoneeMiProblem p) {
  CHK(p, VelocityGradientProjectionApply(p));
  CHK(p, CeedOperatorApply(p->op_sgs_nodal, p->grad_velo, p->sgs_nodal, CEED_REQUEST_IMMEDIATE));
  CHK(p, CeedOperatorApplyAdd(p->op_sgs_apply, p->q_ceed, p->g_ceed, CEED_REQUEST_IMMEDIATE));
  return 0;
}

static int CopyVectorOut(HoneeMiProblem p, CeedVector v, int mem_type, double *out) {
  CeedSize          n;
  const CeedScalar *a;
  CHK(p, CeedVectorGetLength(v, &n));
  CHK(p, CeedVectorGetArrayRead(v, (CeedMemType)mem_type, &a));
  const hipError_t herr = hipMemcpy(out, a, sizeof(double) * (size_t)n, mem_type == CEED_MEM_HOST ? hipMemcpyHostToHost : hipMemcpyDeviceToDevice);
  CHK(p, CeedVectorRestoreArrayRead(v, &a));
  if (herr != hipSuccess) {
    snprintf(p->err, sizeof(p->err), "hipMemcpy failed: %d", (int)herr);
    return CEED_ERROR_BACKEND;
  }
  return 0;
}

int HoneeMiSgsGridAnisotropy(HoneeMiProblem p, double *out, int *cg_iterations) {
  if (!p->sgs_enabled) return CEED_ERROR_INCOMPATIBLE;
  if (cg_iterations) *cg_iterations = p->grid_aniso_its;
  return CopyVectorOut(p, p->grid_aniso, CEED_MEM_HOST, out);
}

int HoneeMiVelocityGradientProjection(HoneeMiProblem p, int mem_type, const double *q, double *grad_velo) {
  CeedScalar *arr;
  if (!p->sgs_enabled) return CEED_ERROR_INCOMPATIBLE;
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, (CeedScalar *)q));
  CHK(p, VelocityGradientProjectionApply(p));
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mem_type, &arr));
  return CopyVectorOut(p, p->grad_velo, mem_type, grad_velo);
}

int HoneeMiSgsNodalStress(HoneeMiProblem p, int mem_type, const double *q, const double *grad_velo, double *km_sgs) {
  CeedScalar *arr;
  if (!p->sgs_enabled) return CEED_ERROR_INCOMPATIBLE;
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, (CeedScalar *)q));
  CHK(p, CeedVectorSetArray(p->grad_velo, (CeedMemType)mem_type, CEED_COPY_VALUES, (CeedScalar *)grad_velo));
  CHK(p, CeedOperatorApply(p->op_sgs_nodal, p->grad_velo, p->sgs_nodal, CEED_REQUEST_IMMEDIATE));
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mem_type, &arr));
  return CopyVectorOut(p, p->sgs_nodal, mem_type, km_sgs);
}

int HoneeMiApplySgsResidual(HoneeMiProblem p, int mem_type, const double *q, const double *km_sgs, double *g) {
  CeedScalar *arr;
  if (!p->sgs_enabled) return CEED_ERROR_INCOMPATIBLE;
  CHK(p, CeedVectorSetArray(p->sgs_nodal, (CeedMemType)mem_type, CEED_COPY_VALUES, (CeedScalar *)km_sgs));
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, (CeedScalar *)q));
  CHK(p, CeedVectorSetArray(p->g_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, g));
  CHK(p, CeedOperatorApplyAdd(p->op_sgs_apply, p->q_ceed, p->g_ceed, CEED_REQUEST_IMMEDIATE));
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mem_type, &arr));
  CHK(p, CeedVectorTakeArray(p->g_ceed, (CeedMemType)mem_type, &arr));
  return 0;
}

// ---------------------------------------------------------------------------------------------------------------------
// Essential (strong) boundary conditions.  HONEE constrains them through the DMPlex section: the solver's vectors
// exclude the constrained entries, every operator input is a local vector into which DMPlexInsertBoundaryValues
// inserts the boundary values (SetBCsFromICs and DMPlexInsertBoundaryValues_FromICs, src/misc.c:51-95: zero the
// constrained entries with the boundary mask, then add Qbc), the time derivative's constrained entries stay zero, and
// local-to-global drops the constrained residual rows.  Over the full L-vector the same system is: inputs
// q .* mask + Qbc and q_dot .* mask, residual rows masked, and the Jacobian (and the mass matrix of explicit problems)
// the identity on the constrained entries, with diagonals and point blocks to match.
// ---------------------------------------------------------------------------------------------------------------------
static int HostToVector(HoneeMiProblem p, CeedSize n, const double *h, CeedVector *v) {
  if (!*v) CHK(p, CeedVectorCreate(p->ceed, n, v));
  CHK(p, CeedVectorSetArray(*v, CEED_MEM_HOST, CEED_COPY_VALUES, (CeedScalar *)h));
  return 0;
}

int HoneeMiProblemSetEssential(HoneeMiProblem p, int n, const int *dofs, const double *values) {
  const CeedSize N = 5 * (CeedSize)p->nnodes;
  if (n < 0 || (n > 0 && !dofs)) return CEED_ERROR_INCOMPLETE;
  for (int i = 0; i < n; i++)
    if (dofs[i] < 0 || dofs[i] >= N) return CEED_ERROR_DIMENSION;
  double *mask = malloc(sizeof(double) * (size_t)N), *qbc = calloc((size_t)N, sizeof(double));
  double *pbm = malloc(sizeof(double) * 25 * (size_t)p->nnodes), *pbe = calloc(25 * (size_t)p->nnodes, sizeof(double));
  if (!mask || !qbc || !pbm || !pbe) {
    free(mask);
    free(qbc);
    free(pbm);
    free(pbe);
    return CEED_ERROR_MAJOR;
  }
  for (CeedSize i = 0; i < N; i++) mask[i] = 1.0;
  for (int i = 0; i < n; i++) {
    mask[dofs[i]] = 0.0;
    qbc[dofs[i]]  = values ? values[i] : 0.0;
  }
  for (CeedSize node = 0; node < p->nnodes; node++)
    for (int r = 0; r < 5; r++)
      for (int c = 0; c < 5; c++) {
        pbm[node * 25 + r * 5 + c] = mask[node * 5 + r] * mask[node * 5 + c];
        pbe[node * 25 + r * 5 + c] = (r == c) ? 1.0 - mask[node * 5 + r] : 0.0;
      }
  int ierr = HostToVector(p, N, mask, &p->ess_mask);
  for (CeedSize i = 0; i < N; i++) mask[i] = 1.0 - mask[i];
  if (!ierr) ierr = HostToVector(p, N, mask, &p->ess_cmask);
  if (!ierr) ierr = HostToVector(p, N, qbc, &p->ess_qbc);
  if (!ierr) ierr = HostToVector(p, 25 * (CeedSize)p->nnodes, pbm, &p->ess_pb_mask);
  if (!ierr) ierr = HostToVector(p, 25 * (CeedSize)p->nnodes, pbe, &p->ess_pb_eye);
  if (!ierr && !p->ess_q) ierr = CeedVectorCreate(p->ceed, N, &p->ess_q);
  if (!ierr && !p->ess_qdot) ierr = CeedVectorCreate(p->ceed, N, &p->ess_qdot);
  free(mask);
  free(qbc);
  free(pbm);
  free(pbe);
  if (ierr) return ierr;
  p->ess_enabled = n > 0;
  return 0;
}

// The local vector of an operator input: work = x .* mask (+ Qbc when insert); *dev is its device array
static int EssentialLocal(HoneeMiProblem p, int mem_type, const double *x, CeedVector work, int insert, const double **dev) {
  CHK(p, CeedVectorSetArray(work, (CeedMemType)mem_type, CEED_COPY_VALUES, (CeedScalar *)x));
  CHK(p, CeedVectorPointwiseMult(work, work, p->ess_mask));
  if (insert) CHK(p, CeedVectorAXPY(work, 1.0, p->ess_qbc));
  const double *tmp;
  CHK(p, CeedVectorGetArrayRead(work, CEED_MEM_DEVICE, &tmp));
  *dev = tmp;
  CHK(p, CeedVectorRestoreArrayRead(work, &tmp));
  return 0;
}

// y = y .* mask + cmask .* x (identity on the constrained entries); x given as the caller's array
static int EssentialIdentityRows(HoneeMiProblem p, int mem_type, CeedVector y, const double *x) {
  CHK(p, CeedVectorPointwiseMult(y, y, p->ess_mask));
  CHK(p, CeedVectorSetArray(p->ess_qdot, (CeedMemType)mem_type, CEED_COPY_VALUES, (CeedScalar *)x));
  CHK(p, CeedVectorPointwiseMult(p->ess_qdot, p->ess_qdot, p->ess_cmask));
  CHK(p, CeedVectorAXPY(y, 1.0, p->ess_qdot));
  return 0;
}

// IFunction_NS's operator section (src/setupts.c:90-111): update context, hand over arrays, apply, take arrays back.
int HoneeMiApplyIFunction(HoneeMiProblem p, int mem_type, double time, double dt, const double *q, const double *q_dot, double *g) {
  CeedScalar *arr;
  int         mt_in = mem_type;
  if (p->solution_time_label) CHK(p, CeedOperatorSetContextDouble(p->op_ifunction, p->solution_time_label, &time));
  if (p->timestep_size_label) CHK(p, CeedOperatorSetContextDouble(p->op_ifunction, p->timestep_size_label, &dt));
  if (p->ess_enabled) {  // UpdateBoundaryValues and the local time derivative (src/setupts.c:86-101)
    CHK(p, EssentialLocal(p, mem_type, q, p->ess_q, 1, &q));
    CHK(p, EssentialLocal(p, mem_type, q_dot, p->ess_qdot, 0, &q_dot));
    mt_in = CEED_MEM_DEVICE;
  }
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mt_in, CEED_USE_POINTER, (CeedScalar *)q));
  CHK(p, CeedVectorSetArray(p->q_dot_ceed, (CeedMemType)mt_in, CEED_USE_POINTER, (CeedScalar *)q_dot));
  CHK(p, CeedVectorSetArray(p->g_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, g));
  CHK(p, CeedOperatorApply(p->op_ifunction, p->q_ceed, p->g_ceed, CEED_REQUEST_IMMEDIATE));
  if (p->sgs_enabled) CHK(p, SgsDDApplyIFunction(p));  // src/setupts.c:118-120
  if (p->ess_enabled) CHK(p, CeedVectorPointwiseMult(p->g_ceed, p->g_ceed, p->ess_mask));  // constrained rows dropped
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mt_in, &arr));
  CHK(p, CeedVectorTakeArray(p->q_dot_ceed, (CeedMemType)mt_in, &arr));
  CHK(p, CeedVectorTakeArray(p->g_ceed, (CeedMemType)mem_type, &arr));
  return 0;
}

// MatMult_Ceed's operator section (src/mat-ceed.c:1643-1694) with FormIJacobian_NS's shift (src/setupts.c:141)
int HoneeMiApplyIJacobian(HoneeMiProblem p, int mem_type, double shift, const double *dq, double *v) {
  CeedScalar   *arr;
  const double *dq_in = dq;
  int           mt_in = mem_type;
  if (p->ijacobian_shift_label) CHK(p, CeedOperatorSetContextDouble(p->op_ijacobian, p->ijacobian_shift_label, &shift));
  if (p->ess_enabled) {
    CHK(p, EssentialLocal(p, mem_type, dq, p->ess_q, 0, &dq_in));
    mt_in = CEED_MEM_DEVICE;
  }
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mt_in, CEED_USE_POINTER, (CeedScalar *)dq_in));
  CHK(p, CeedVectorSetArray(p->g_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, v));
  CHK(p, CeedOperatorApply(p->op_ijacobian, p->q_ceed, p->g_ceed, CEED_REQUEST_IMMEDIATE));
  if (p->ess_enabled) CHK(p, EssentialIdentityRows(p, mem_type, p->g_ceed, dq));
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mt_in, &arr));
  CHK(p, CeedVectorTakeArray(p->g_ceed, (CeedMemType)mem_type, &arr));
  return 0;
}

// Point-Jacobi diagonal of the shifted IJacobian (MatGetDiagonal_Ceed, src/mat-ceed.c:1603-1630)
int HoneeMiIJacobianDiagonal(HoneeMiProblem p, int mem_type, double shift, double *diag) {
  CeedScalar *arr;
  if (p->ijacobian_shift_label) CHK(p, CeedOperatorSetContextDouble(p->op_ijacobian, p->ijacobian_shift_label, &shift));
  CHK(p, CeedVectorSetArray(p->g_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, diag));
  CHK(p, CeedOperatorLinearAssembleDiagonal(p->op_ijacobian, p->g_ceed, CEED_REQUEST_IMMEDIATE));
  if (p->ess_enabled) {  // unit diagonal on the constrained entries
    CHK(p, CeedVectorPointwiseMult(p->g_ceed, p->g_ceed, p->ess_mask));
    CHK(p, CeedVectorAXPY(p->g_ceed, 1.0, p->ess_cmask));
  }
  CHK(p, CeedVectorTakeArray(p->g_ceed, (CeedMemType)mem_type, &arr));
  return 0;
}

// RHS_NS's operator section (src/setupts.c:31-52): update time and dt, apply the RHS composite
int HoneeMiApplyRHS(HoneeMiProblem p, int mem_type, double time, double dt, const double *q, double *r) {
  CeedScalar *arr;
  int         mt_in = mem_type;
  if (!p->op_rhs) return CEED_ERROR_INCOMPATIBLE;
  if (p->rhs_time_label) CHK(p, CeedOperatorSetContextDouble(p->op_rhs, p->rhs_time_label, &time));
  if (p->rhs_dt_label) CHK(p, CeedOperatorSetContextDouble(p->op_rhs, p->rhs_dt_label, &dt));
  if (p->ess_enabled) {  // UpdateBoundaryValues (src/setupts.c:40)
    CHK(p, EssentialLocal(p, mem_type, q, p->ess_q, 1, &q));
    mt_in = CEED_MEM_DEVICE;
  }
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mt_in, CEED_USE_POINTER, (CeedScalar *)q));
  CHK(p, CeedVectorSetArray(p->g_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, r));
  CHK(p, CeedOperatorApply(p->op_rhs, p->q_ceed, p->g_ceed, CEED_REQUEST_IMMEDIATE));
  if (p->ess_enabled) CHK(p, CeedVectorPointwiseMult(p->g_ceed, p->g_ceed, p->ess_mask));
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mt_in, &arr));
  CHK(p, CeedVectorTakeArray(p->g_ceed, (CeedMemType)mem_type, &arr));
  return 0;
}

// Mass matrix action y = M(q) x of the explicit mass KSP (MatMult of the "mass_" KSP's MatCeed, src/setuplibceed.c:56-95);
// q is the state the SUPG-stabilized mass is evaluated at (ignored by Mass_5).
int HoneeMiApplyMass(HoneeMiProblem p, int mem_type, const double *q, const double *x, double *y) {
  CeedScalar   *arr;
  const double *x_in  = x;
  int           mt_q  = mem_type, mt_x = mem_type;
  if (!p->op_mass) return CEED_ERROR_INCOMPATIBLE;
  if (p->ess_enabled) {
    if (p->mass_stabilized) {
      CHK(p, EssentialLocal(p, mem_type, q, p->ess_q, 1, &q));
      mt_q = CEED_MEM_DEVICE;
    }
    CHK(p, EssentialLocal(p, mem_type, x, p->ess_qdot, 0, &x_in));
    mt_x = CEED_MEM_DEVICE;
  }
  if (p->mass_stabilized) CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mt_q, CEED_USE_POINTER, (CeedScalar *)q));
  CHK(p, CeedVectorSetArray(p->q_dot_ceed, (CeedMemType)mt_x, CEED_USE_POINTER, (CeedScalar *)x_in));
  CHK(p, CeedVectorSetArray(p->g_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, y));
  CHK(p, CeedOperatorApply(p->op_mass, p->q_dot_ceed, p->g_ceed, CEED_REQUEST_IMMEDIATE));
  if (p->mass_stabilized) CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mt_q, &arr));
  CHK(p, CeedVectorTakeArray(p->q_dot_ceed, (CeedMemType)mt_x, &arr));
  if (p->ess_enabled) CHK(p, EssentialIdentityRows(p, mem_type, p->g_ceed, x));
  CHK(p, CeedVectorTakeArray(p->g_ceed, (CeedMemType)mem_type, &arr));
  return 0;
}

// Diagonal of the mass matrix (PCJACOBI "diagonal" of the mass KSP)
int HoneeMiMassDiagonal(HoneeMiProblem p, int mem_type, const double *q, double *diag) {
  CeedScalar *arr;
  int         mt_q = mem_type;
  if (!p->op_mass) return CEED_ERROR_INCOMPATIBLE;
  if (p->ess_enabled && p->mass_stabilized) {
    CHK(p, EssentialLocal(p, mem_type, q, p->ess_q, 1, &q));
    mt_q = CEED_MEM_DEVICE;
  }
  if (p->mass_stabilized) CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mt_q, CEED_USE_POINTER, (CeedScalar *)q));
  CHK(p, CeedVectorSetArray(p->g_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, diag));
  CHK(p, CeedOperatorLinearAssembleDiagonal(p->op_mass, p->g_ceed, CEED_REQUEST_IMMEDIATE));
  if (p->ess_enabled) {
    CHK(p, CeedVectorPointwiseMult(p->g_ceed, p->g_ceed, p->ess_mask));
    CHK(p, CeedVectorAXPY(p->g_ceed, 1.0, p->ess_cmask));
  }
  if (p->mass_stabilized) CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mt_q, &arr));
  CHK(p, CeedVectorTakeArray(p->g_ceed, (CeedMemType)mem_type, &arr));
  return 0;
}

int HoneeMiProblemIsExplicit(HoneeMiProblem p) { return p->op_rhs != NULL; }

double HoneeMiProblemVolume(HoneeMiProblem p) { return p->volume; }

// TSMonitor_TotalKineticEnergy (src/monitor_totalkineticenergy.c:113-140): sums over the quadrature-point values of
// the monitor QFunction; sums[] are not divided by the domain volume.
int HoneeMiMonitorKE(HoneeMiProblem p, int mem_type, const double *q, double sums[5]) {
  CeedScalar       *arr;
  const CeedScalar *a;
  CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, (CeedScalar *)q));
  CHK(p, CeedOperatorApply(p->op_monitor, p->q_ceed, p->ke_values, CEED_REQUEST_IMMEDIATE));
  CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mem_type, &arr));
  CHK(p, CeedVectorGetArrayRead(p->ke_values, CEED_MEM_HOST, &a));
  const long long Q3 = p->Q3;
  for (int c = 0; c < 5; c++) sums[c] = 0;
  for (long long e = 0; e < p->nelem; e++)
    for (int c = 0; c < 5; c++)
      for (long long i = 0; i < Q3; i++) sums[c] += a[(e * 5 + c) * Q3 + i];
  CHK(p, CeedVectorRestoreArrayRead(p->ke_values, &a));
  return 0;
}

// Point-block diagonal [node][c_out][c_in] of the shifted IJacobian (MatCeedAssemblePointBlockDiagonalCOO's values,
// src/mat-ceed.c:118-141), as PCVPBJACOBI uses it
int HoneeMiIJacobianPointBlockDiagonal(HoneeMiProblem p, int mem_type, double shift, double *pb) {
  CeedScalar *arr;
  if (!p->pb_ceed) CHK(p, CeedVectorCreate(p->ceed, 25 * (CeedSize)p->nnodes, &p->pb_ceed));
  if (p->ijacobian_shift_label) CHK(p, CeedOperatorSetContextDouble(p->op_ijacobian, p->ijacobian_shift_label, &shift));
  CHK(p, CeedVectorSetArray(p->pb_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, pb));
  CHK(p, CeedOperatorLinearAssemblePointBlockDiagonal(p->op_ijacobian, p->pb_ceed, CEED_REQUEST_IMMEDIATE));
  if (p->ess_enabled) {  // constrained rows and columns of each block replaced by the identity
    CHK(p, CeedVectorPointwiseMult(p->pb_ceed, p->pb_ceed, p->ess_pb_mask));
    CHK(p, CeedVectorAXPY(p->pb_ceed, 1.0, p->ess_pb_eye));
  }
  CHK(p, CeedVectorTakeArray(p->pb_ceed, (CeedMemType)mem_type, &arr));
  return 0;
}

// Point-block diagonal of the explicit mass matrix at state q
int HoneeMiMassPointBlockDiagonal(HoneeMiProblem p, int mem_type, const double *q, double *pb) {
  CeedScalar *arr;
  int         mt_q = mem_type;
  if (!p->op_mass) return CEED_ERROR_INCOMPATIBLE;
  if (!p->pb_ceed) CHK(p, CeedVectorCreate(p->ceed, 25 * (CeedSize)p->nnodes, &p->pb_ceed));
  if (p->ess_enabled && p->mass_stabilized) {
    CHK(p, EssentialLocal(p, mem_type, q, p->ess_q, 1, &q));
    mt_q = CEED_MEM_DEVICE;
  }
  if (p->mass_stabilized) CHK(p, CeedVectorSetArray(p->q_ceed, (CeedMemType)mt_q, CEED_USE_POINTER, (CeedScalar *)q));
  CHK(p, CeedVectorSetArray(p->pb_ceed, (CeedMemType)mem_type, CEED_USE_POINTER, pb));
  CHK(p, CeedOperatorLinearAssemblePointBlockDiagonal(p->op_mass, p->pb_ceed, CEED_REQUEST_IMMEDIATE));
  if (p->ess_enabled) {
    CHK(p, CeedVectorPointwiseMult(p->pb_ceed, p->pb_ceed, p->ess_pb_mask));
    CHK(p, CeedVectorAXPY(p->pb_ceed, 1.0, p->ess_pb_eye));
  }
  if (p->mass_stabilized) CHK(p, CeedVectorTakeArray(p->q_ceed, (CeedMemType)mt_q, &arr));
  CHK(p, CeedVectorTakeArray(p->pb_ceed, (CeedMemType)mem_type, &arr));
  return 0;
}

// Copy backend-strided quadrature data ([elem][comp][qpt], CEED_STRIDES_BACKEND of this backend) to host
static int CopyStrided(HoneeMiProblem p, CeedVector v, CeedElemRestriction r, int ncomp, double *out) {
  const CeedScalar *a;
  CHK(p, CeedVectorGetArrayRead(v, CEED_MEM_HOST, &a));
  const long long Q3 = p->Q3;
  for (long long e = 0; e < p->nelem; e++)
    for (int c = 0; c < ncomp; c++)
      for (long long i = 0; i < Q3; i++) out[(e * ncomp + c) * Q3 + i] = a[(e * ncomp + c) * Q3 + i];
  CHK(p, CeedVectorRestoreArrayRead(v, &a));
  return 0;
}

int HoneeMiGetQData(HoneeMiProblem p, double *out) { return CopyStrided(p, p->q_data, p->elem_restr_qd, 10, out); }
int HoneeMiGetJacData(HoneeMiProblem p, double *out) { return CopyStrided(p, p->jac_data, p->elem_restr_jd, p->num_comps_jac_data, out); }

int HoneeMiGetJacData32(HoneeMiProblem p, float *out) {
  const float *d;
  CeedSize     n;
  CHK(p, CeedMiVectorGetFp32ShadowRead(p->jac_data, &d, &n));
  float *h = malloc(sizeof(float) * n);
  if (hipMemcpy(h, d, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess) {
    free(h);
    return CEED_ERROR_BACKEND;
  }
  const long long Q3 = p->Q3;
  for (long long e = 0; e < p->nelem; e++)
    for (int c = 0; c < p->num_comps_jac_data; c++)
      for (long long i = 0; i < Q3; i++) out[(e * p->num_comps_jac_data + c) * Q3 + i] = h[(e * p->num_comps_jac_data + c) * Q3 + i];
  free(h);
  return 0;
}

const char *HoneeMiProblemError(HoneeMiProblem p) { return p->err; }

int HoneeMiProblemDestroy(HoneeMiProblem *problem) {
  HoneeMiProblem p = *problem;
  if (!p) return 0;
  CeedVectorDestroy(&p->ess_mask);
  CeedVectorDestroy(&p->ess_cmask);
  CeedVectorDestroy(&p->ess_qbc);
  CeedVectorDestroy(&p->ess_pb_mask);
  CeedVectorDestroy(&p->ess_pb_eye);
  CeedVectorDestroy(&p->ess_q);
  CeedVectorDestroy(&p->ess_qdot);
  CeedOperatorDestroy(&p->op_ifunction);
  CeedOperatorDestroy(&p->op_ijacobian);
  CeedOperatorDestroy(&p->op_grad_velo_rhs);
  CeedOperatorDestroy(&p->op_sgs_nodal);
  CeedOperatorDestroy(&p->op_sgs_apply);
  CeedVectorDestroy(&p->grad_velo);
  CeedVectorDestroy(&p->grad_velo_inv_rowsum);
  CeedVectorDestroy(&p->grid_aniso);
  CeedVectorDestroy(&p->sgs_nodal);
  CeedVectorDestroy(&p->inv_mult);
  CeedElemRestrictionDestroy(&p->elem_restr_grad_velo);
  CeedElemRestrictionDestroy(&p->elem_restr_grid_aniso);
  CeedElemRestrictionDestroy(&p->elem_restr_sgs);
  CeedElemRestrictionDestroy(&p->elem_restr_inv_mult);
  CeedBasisDestroy(&p->basis_grad_velo);
  CeedBasisDestroy(&p->basis_grid_aniso);
  CeedBasisDestroy(&p->basis_sgs);
  CeedBasisDestroy(&p->basis_x_to_q);
  if (p->sgsdd_qfctx) CeedQFunctionContextDestroy(&p->sgsdd_qfctx);
  CeedOperatorDestroy(&p->op_filter_rhs);
  CeedOperatorDestroy(&p->op_filter_lhs);
  CeedOperatorDestroy(&p->op_filtered_grad_velo_rhs);
  CeedOperatorDestroy(&p->op_sgs_train);
  for (int i = 0; i < 2; i++) {
    CeedElemRestrictionDestroy(&p->elem_restr_filter[i]);
    CeedBasisDestroy(&p->basis_filter[i]);
  }
  CeedElemRestrictionDestroy(&p->elem_restr_aniso_col);
  CeedElemRestrictionDestroy(&p->elem_restr_train);
  CeedVector *fv[9] = {&p->filter_rhs, &p->filtered, &p->filter_dinv, &p->aniso_col, &p->train_data, &p->cg_work[0], &p->cg_work[1], &p->cg_work[2],
                       NULL};
  for (int i = 0; fv[i]; i++) CeedVectorDestroy(fv[i]);
  if (p->diff_filter_qfctx) CeedQFunctionContextDestroy(&p->diff_filter_qfctx);
  if (p->train_qfctx) CeedQFunctionContextDestroy(&p->train_qfctx);
  CeedOperatorDestroy(&p->op_rhs);
  CeedOperatorDestroy(&p->op_mass);
  CeedOperatorDestroy(&p->op_monitor);
  CeedVectorDestroy(&p->ke_values);
  CeedVectorDestroy(&p->pb_ceed);
  CeedElemRestrictionDestroy(&p->elem_restr_ke);
  CeedVectorDestroy(&p->q_ceed);
  CeedVectorDestroy(&p->q_dot_ceed);
  CeedVectorDestroy(&p->g_ceed);
  CeedVectorDestroy(&p->x_coord);
  CeedVectorDestroy(&p->q_data);
  CeedVectorDestroy(&p->jac_data);
  CeedElemRestrictionDestroy(&p->elem_restr_q);
  CeedElemRestrictionDestroy(&p->elem_restr_x);
  CeedElemRestrictionDestroy(&p->elem_restr_qd);
  CeedElemRestrictionDestroy(&p->elem_restr_jd);
  CeedBasisDestroy(&p->basis_q);
  CeedBasisDestroy(&p->basis_x);
  for (int i = 0; i < p->nbnd; i++) {
    HoneeMiBoundary *b = &p->bnd[i];
    CeedVectorDestroy(&b->x_coord);
    CeedVectorDestroy(&b->q_data);
    CeedVectorDestroy(&b->jac_data);
    CeedElemRestrictionDestroy(&b->restr_q);
    CeedElemRestrictionDestroy(&b->restr_x);
    CeedElemRestrictionDestroy(&b->restr_qd);
    CeedElemRestrictionDestroy(&b->restr_jd);
    CeedBasisDestroy(&b->basis_q);
    CeedBasisDestroy(&b->basis_x);
    if (b->fs_qfctx) CeedQFunctionContextDestroy(&b->fs_qfctx);
  }
  if (p->newt_qfctx) CeedQFunctionContextDestroy(&p->newt_qfctx);
  else free(p->ctx);
  CeedDestroy(&p->ceed);
  free(p);
  *problem = NULL;
  return 0;
}
