// Operator binding for the MI355X backend.
//
// A non-composite CeedOperator is bound, on its first application, to one fused kernel chosen by the identity of its
// QFunction (the "<file>:<name>" location HONEE registers, e.g. newtonian.h:IFunction_Newtonian_Prim).  Binding checks
// that the operator's fields have exactly the layout HONEE's SetupLibceed gives them (src/setuplibceed.c:209-269,
// src/qdata.c:154-169) and records the restriction/basis/vector objects; device pointers and the QFunction context are
// read again at every application, because HONEE swaps vector arrays and context values between calls.
// QFunctions without a fused kernel are rejected with CEED_ERROR_UNSUPPORTED: there is no host execution path.
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ceed_impl.hpp"

struct OpPlan {
  QfKind              kind = QfKind::Unknown;
  int                 var  = 0;
  int                 P = 0, Q = 0, Px = 0, nelem = 0;
  int                 njac       = 0;
  CeedElemRestriction r_q        = nullptr, r_x = nullptr, r_qd = nullptr, r_jd = nullptr;
  CeedBasis           b_q        = nullptr, b_x = nullptr;
  CeedVector          v_qdot     = nullptr, v_qdata = nullptr, v_x = nullptr, v_jac = nullptr;
  QDataLayout         qd{}, jd{};
  bool                face = false;  // boundary-face operator: 2D tensor basis, Q^2 points per face
  bool                hllc = false;
  std::string         kernel_name;
  // field operators of the SGS closure (own output restriction, second inputs) and the model parameters
  int                 ncomp_out = 0;
  CeedElemRestriction r_out = nullptr, r_in2 = nullptr, r_aux = nullptr, r_aux2 = nullptr, r_im = nullptr;
  CeedBasis           b_out = nullptr, b_in2 = nullptr;
  CeedVector          v_in2 = nullptr, v_aux = nullptr, v_aux2 = nullptr, v_im = nullptr;
  double             *d_model = nullptr;  // packed sgs model parameters (device)
  std::vector<double> model_key;          // host copy of what d_model holds
};

void op_plan_destroy(OpPlan *plan) {
  if (plan && plan->d_model) (void)hipFree(plan->d_model);
  delete plan;
}

static int plan_bind_fields(CeedOperator op);

// The fused kernels address strided quadrature data in the backend layout [elem][comp][qpt], i.e. strides
// {1, Q^3, ncomp Q^3} with one point per Q^3 slot; a restriction with other strides is not bound.
static int check_point_layout(CeedOperator op, CeedElemRestriction r, int Q, bool face) {
  if (!r) return 0;
  const CeedInt Q3 = face ? Q * Q : Q * Q * Q;
  if (!(r->elem_size == Q3 && r->strides[0] == 1 && r->strides[1] == Q3 && r->strides[2] == (CeedInt)r->num_comp * Q3))
    return MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "strided quadrature data must use CEED_STRIDES_BACKEND (or strides {1, Q^3, ncomp Q^3}) "
                                                    "with %d points per element",
                  Q3);
  return 0;
}

// Quadrature-data fields read by the fused kernels: 10 components ({wdetJ, dXdx}) per point of every element, held in
// a vector of at least that length (the kernels copy whole elements' qdata blocks).
static int check_qdata_field(CeedOperator op, const OpPlan *p) {
  if (p->kind == QfKind::Setup || p->kind == QfKind::SetupBoundary || !p->r_qd) return 0;
  const CeedSize npts = p->face ? (CeedSize)p->Q * p->Q : (CeedSize)p->Q * p->Q * p->Q;
  const CeedSize need = (CeedSize)p->nelem * 10 * npts;
  if (p->r_qd->num_comp != 10 || p->r_qd->num_elem != p->nelem || !p->v_qdata || p->v_qdata->length < need)
    return MI_ERR(op->ceed, CEED_ERROR_DIMENSION, "qdata must be a 10-component strided field over the operator's %d elements", p->nelem);
  if (p->r_jd && (p->r_jd->num_elem != p->nelem || (p->v_jac && p->v_jac->length < (CeedSize)p->nelem * p->r_jd->num_comp * npts)))
    return MI_ERR(op->ceed, CEED_ERROR_DIMENSION, "jac_data must be a strided field over the operator's %d elements", p->nelem);
  return 0;
}

static int plan_bind(CeedOperator op) {
  MI_CALL(plan_bind_fields(op));
  OpPlan *p    = op->plan;
  int     ierr = check_point_layout(op, p->r_qd, p->Q, p->face);
  if (!ierr) ierr = check_point_layout(op, p->r_jd, p->Q, p->face);
  if (!ierr) ierr = check_qdata_field(op, p);
  if (ierr) {
    op_plan_destroy(p);
    op->plan = nullptr;
  }
  return ierr;
}

static QDataLayout strided_layout(CeedElemRestriction r) {
  return QDataLayout{r->strides[0], r->strides[1], r->strides[2]};
}

static bool is_active(CeedVector v) { return v == CEED_VECTOR_ACTIVE; }
static bool is_strided(CeedElemRestriction r) { return r && r != CEED_ELEMRESTRICTION_NONE && r->strided; }
static bool is_offsets(CeedElemRestriction r) { return r && r != CEED_ELEMRESTRICTION_NONE && !r->strided; }
static bool is_tensor3(CeedBasis b) { return b && b != CEED_BASIS_NONE && b->tensor && b->dim == 3; }
static bool is_tensor2(CeedBasis b) { return b && b != CEED_BASIS_NONE && b->tensor && b->dim == 2; }

#define BIND_CHECK(cond, ...)                                                       \
  do {                                                                              \
    if (!(cond)) return MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, __VA_ARGS__);      \
  } while (0)

// Whether every two element nodes that share an offset of `out` also read the same offset of each restriction in `ins`
// (all with out's element nodes): nodal kernels then evaluate each output node once.
static bool shared_outputs_read_same_inputs(CeedElemRestriction out, std::initializer_list<CeedElemRestriction> ins) {
  std::vector<int> rowptr, index, toff;
  group_by_offset(out->h_offsets, rowptr, index, toff);
  for (size_t r = 0; r + 1 < rowptr.size(); r++)
    for (int j = rowptr[r] + 1; j < rowptr[r + 1]; j++)
      for (CeedElemRestriction in : ins)
        if (in->h_offsets[index[j - 1]] != in->h_offsets[index[j]]) return false;
  return true;
}

static int check_field(CeedOperator op, size_t idx, bool input, const char *name, CeedEvalMode emode, CeedInt size) {
  const auto &fields = input ? op->qf->inputs : op->qf->outputs;
  BIND_CHECK(idx < fields.size(), "QFunction %s: missing %s field %zu (%s)", op->qf->name.c_str(), input ? "input" : "output", idx, name);
  const QFField &f = fields[idx];
  BIND_CHECK(f.emode == emode, "QFunction %s: field '%s' has eval mode %d, expected %d", op->qf->name.c_str(), f.name.c_str(), (int)f.emode,
             (int)emode);
  BIND_CHECK(size < 0 || f.size == size, "QFunction %s: field '%s' has size %d, expected %d", op->qf->name.c_str(), f.name.c_str(), f.size, size);
  return 0;
}

static int bind_solution_basis(CeedOperator op, OpPlan *p, CeedOperatorField_private *f) {
  BIND_CHECK(is_offsets(f->rstr) && is_tensor3(f->basis), "field '%s' needs an offset restriction and a 3D tensor basis", f->name.c_str());
  BIND_CHECK(f->rstr->num_comp == 5 && f->basis->num_comp == 5, "field '%s' must have 5 components", f->name.c_str());
  p->r_q   = f->rstr;
  p->b_q   = f->basis;
  p->P     = f->basis->P_1d;
  p->Q     = f->basis->Q_1d;
  p->nelem = f->rstr->num_elem;
  BIND_CHECK(volume_pq_supported(p->P, p->Q), "no fused kernel for P_1d=%d, Q_1d=%d", p->P, p->Q);
  BIND_CHECK(f->rstr->elem_size == p->P * p->P * p->P, "restriction element size does not match basis");
  MI_CALL(basis_device_tables(f->basis));
  return 0;
}

// Face solution field: 5-component field on the boundary faces with a 2D tensor basis (Q_1d = P_1d + q_extra)
static int bind_face_basis(CeedOperator op, OpPlan *p, CeedOperatorField_private *f) {
  BIND_CHECK(is_offsets(f->rstr) && is_tensor2(f->basis), "face field '%s' needs an offset restriction and a 2D tensor basis", f->name.c_str());
  BIND_CHECK(f->rstr->num_comp == 5 && f->basis->num_comp == 5, "face field '%s' must have 5 components", f->name.c_str());
  p->face  = true;
  p->r_q   = f->rstr;
  p->b_q   = f->basis;
  p->P     = f->basis->P_1d;
  p->Q     = f->basis->Q_1d;
  p->nelem = f->rstr->num_elem;
  BIND_CHECK(face_pq_supported(p->P, p->Q), "no fused face kernel for P_1d=%d, Q_1d=%d", p->P, p->Q);
  BIND_CHECK(f->rstr->elem_size == p->P * p->P, "face restriction element size does not match the face basis");
  MI_CALL(basis_device_tables(f->basis));
  return 0;
}

// Face operators (boundary_condition.c:300-360): q/dq (INTERP), Grad_q/Grad_dq (GRAD, 2 reference directions), surface
// qdata, x; output v (INTERP) and, for the residual of an implicit problem, the surface Jacobian data
static int bind_face_op(CeedOperator op, OpPlan *p, bool jacobian, int njac) {
  CeedQFunction qf       = op->qf;
  const char   *names[5] = {jacobian ? "dq" : "q", jacobian ? "Grad_dq" : "Grad_q", "surface qdata", "x", "surface jacobian data"};
  CeedEvalMode  modes[5] = {CEED_EVAL_INTERP, CEED_EVAL_GRAD, CEED_EVAL_NONE, CEED_EVAL_INTERP, CEED_EVAL_NONE};
  CeedInt       sizes[5] = {5, 10, 10, 3, njac};
  const size_t  nin      = jacobian ? 5 : 4;
  BIND_CHECK(qf->inputs.size() == nin, "%s: unexpected input fields", qf->name.c_str());
  for (size_t i = 0; i < nin; i++) MI_CALL(check_field(op, i, true, names[i], modes[i], sizes[i]));
  MI_CALL(check_field(op, 0, false, "v", CEED_EVAL_INTERP, 5));
  auto **fi = op->in_fields.data(), **fo = op->out_fields.data();
  MI_CALL(bind_face_basis(op, p, fi[0]));
  BIND_CHECK(is_active(fi[0]->vec) && is_active(fi[1]->vec) && fi[1]->rstr == p->r_q && fi[1]->basis == p->b_q,
             "%s: the solution value and gradient must share the active face restriction/basis", qf->name.c_str());
  BIND_CHECK(is_strided(fi[2]->rstr) && !is_active(fi[2]->vec), "%s: surface qdata must be a passive strided field", qf->name.c_str());
  p->r_qd    = fi[2]->rstr;
  p->v_qdata = fi[2]->vec;
  p->qd      = strided_layout(p->r_qd);
  BIND_CHECK(is_active(fo[0]->vec) && fo[0]->rstr == p->r_q, "%s: v must be active on the face space", qf->name.c_str());
  if (jacobian) {
    BIND_CHECK(is_strided(fi[4]->rstr) && !is_active(fi[4]->vec), "%s: surface jacobian data must be a passive strided field", qf->name.c_str());
    p->r_jd  = fi[4]->rstr;
    p->v_jac = fi[4]->vec;
    p->jd    = strided_layout(p->r_jd);
    p->njac  = njac;
  } else if (qf->outputs.size() == 2) {
    MI_CALL(check_field(op, 1, false, "surface jacobian data", CEED_EVAL_NONE, njac));
    BIND_CHECK(is_strided(fo[1]->rstr) && !is_active(fo[1]->vec), "%s: surface jacobian data must be a passive strided field", qf->name.c_str());
    p->r_jd  = fo[1]->rstr;
    p->v_jac = fo[1]->vec;
    p->jd    = strided_layout(p->r_jd);
    p->njac  = njac;
  }
  return 0;
}

static int bind_coords(CeedOperator op, OpPlan *p, CeedOperatorField_private *f) {
  BIND_CHECK(is_offsets(f->rstr) && is_tensor3(f->basis) && f->basis->P_1d == 2 && f->rstr->num_comp == 3,
             "coordinate field '%s' must be a 3-component trilinear (P_1d = 2) field", f->name.c_str());
  BIND_CHECK(!is_active(f->vec) && f->vec != CEED_VECTOR_NONE, "coordinate field must be passive");
  BIND_CHECK(f->rstr->num_elem == p->nelem && f->basis->Q_1d == p->Q, "coordinate field does not match the solution field");
  p->r_x = f->rstr;
  p->b_x = f->basis;
  p->v_x = f->vec;
  MI_CALL(basis_device_tables(f->basis));
  return 0;
}

// A field with an offset restriction and a 3D tensor basis of ncomp components (ncomp < 0: any), interpolating to
// the operator's quadrature points; the first such field fixes P, Q and the element count of the plan.
static int bind_field_basis(CeedOperator op, OpPlan *p, CeedOperatorField_private *f, int ncomp) {
  BIND_CHECK(is_offsets(f->rstr) && is_tensor3(f->basis), "field '%s' needs an offset restriction and a 3D tensor basis", f->name.c_str());
  BIND_CHECK(ncomp < 0 || (f->rstr->num_comp == ncomp && f->basis->num_comp == ncomp), "field '%s' must have %d components", f->name.c_str(), ncomp);
  BIND_CHECK(f->rstr->num_comp == f->basis->num_comp, "field '%s': restriction and basis component counts differ", f->name.c_str());
  if (!p->b_q && !p->b_out) {
    p->P     = f->basis->P_1d;
    p->Q     = f->basis->Q_1d;
    p->nelem = f->rstr->num_elem;
    BIND_CHECK(sgs_pq_supported(p->P, p->Q), "no fused kernel for P_1d=%d, Q_1d=%d", p->P, p->Q);
  } else {
    const CeedBasis b0 = p->b_q ? p->b_q : p->b_out;
    BIND_CHECK(f->basis->P_1d == p->P && f->basis->Q_1d == p->Q && f->rstr->num_elem == p->nelem && f->basis->interp_1d == b0->interp_1d &&
                   f->basis->grad_1d == b0->grad_1d,
               "field '%s' must use the same 1D basis tables and elements as the operator's other fields", f->name.c_str());
  }
  BIND_CHECK(f->rstr->elem_size == p->P * p->P * p->P, "restriction element size does not match basis");
  MI_CALL(basis_device_tables(f->basis));
  return 0;
}

static int bind_qdata_in(CeedOperator op, OpPlan *p, CeedOperatorField_private *f) {
  BIND_CHECK(is_strided(f->rstr) && !is_active(f->vec), "%s: qdata must be a passive strided field", op->qf->name.c_str());
  p->r_qd    = f->rstr;
  p->v_qdata = f->vec;
  p->qd      = strided_layout(p->r_qd);
  return 0;
}

// Fields of the SGS closure and its projections (problems/sgs_dd_model.c:114-203, 498-558;
// src/velocity_gradient_projection.c:60-130; src/grid_anisotropy_tensor.c:44-88; src/inverse_multiplicity.c:24-80)
static int bind_sgs_op(CeedOperator op, OpPlan *p) {
  CeedQFunction qf = op->qf;
  auto        **fi = op->in_fields.data(), **fo = op->out_fields.data();
  switch (qf->kind) {
    case QfKind::VelocityGradientRHS: {
      const char  *names[4] = {"q", "Grad_q", "qdata", "x"};
      CeedEvalMode modes[4] = {CEED_EVAL_INTERP, CEED_EVAL_GRAD, CEED_EVAL_NONE, CEED_EVAL_INTERP};
      CeedInt      sizes[4] = {5, 15, 10, 3};
      BIND_CHECK(qf->inputs.size() == 4 && qf->outputs.size() == 1, "%s: unexpected fields", qf->name.c_str());
      for (int i = 0; i < 4; i++) MI_CALL(check_field(op, i, true, names[i], modes[i], sizes[i]));
      MI_CALL(check_field(op, 0, false, "velocity gradient", CEED_EVAL_INTERP, 9));
      MI_CALL(bind_field_basis(op, p, fi[0], 5));
      p->r_q = fi[0]->rstr;
      p->b_q = fi[0]->basis;
      BIND_CHECK(is_active(fi[0]->vec) && is_active(fi[1]->vec) && fi[1]->rstr == p->r_q && fi[1]->basis == p->b_q,
                 "%s: q and Grad_q must share the active restriction/basis", qf->name.c_str());
      MI_CALL(bind_qdata_in(op, p, fi[2]));
      MI_CALL(bind_coords(op, p, fi[3]));
      BIND_CHECK(is_active(fo[0]->vec), "%s: the velocity gradient must be active", qf->name.c_str());
      MI_CALL(bind_field_basis(op, p, fo[0], 9));
      p->r_out       = fo[0]->rstr;
      p->b_out       = fo[0]->basis;
      p->ncomp_out   = 9;
      p->kernel_name = "velocity_gradient_rhs";
      return 0;
    }
    case QfKind::AnisotropyRHS: {
      BIND_CHECK(qf->inputs.size() == 1 && qf->outputs.size() == 1, "%s: unexpected fields", qf->name.c_str());
      MI_CALL(check_field(op, 0, true, "qdata", CEED_EVAL_NONE, 10));
      MI_CALL(check_field(op, 0, false, "v", CEED_EVAL_INTERP, 7));
      BIND_CHECK(is_active(fo[0]->vec), "%s: the output must be active", qf->name.c_str());
      MI_CALL(bind_field_basis(op, p, fo[0], 7));
      p->r_out     = fo[0]->rstr;
      p->b_out     = fo[0]->basis;
      p->ncomp_out = 7;
      MI_CALL(bind_qdata_in(op, p, fi[0]));
      p->kernel_name = "anisotropy_rhs";
      return 0;
    }
    case QfKind::MassN: {
      const int nc = qf->ncomp;
      BIND_CHECK(mass_n_supported(nc), "%s: no fused kernel for %d components", qf->name.c_str(), nc);
      BIND_CHECK(qf->inputs.size() == 2 && qf->outputs.size() == 1, "%s: unexpected fields", qf->name.c_str());
      MI_CALL(check_field(op, 0, true, "u", CEED_EVAL_INTERP, nc));
      MI_CALL(check_field(op, 1, true, "qdata", CEED_EVAL_NONE, -1));
      MI_CALL(check_field(op, 0, false, "v", CEED_EVAL_INTERP, nc));
      MI_CALL(bind_field_basis(op, p, fi[0], nc));
      p->r_q = p->r_out = fi[0]->rstr;
      p->b_q = p->b_out = fi[0]->basis;
      BIND_CHECK(is_active(fi[0]->vec) && is_active(fo[0]->vec) && fo[0]->rstr == p->r_q && fo[0]->basis == p->b_q,
                 "%s: u and v must be active on the same restriction/basis", qf->name.c_str());
      MI_CALL(bind_qdata_in(op, p, fi[1]));
      p->ncomp_out   = nc;
      p->kernel_name = "mass_n";
      return 0;
    }
    case QfKind::InverseMultiplicity: {
      BIND_CHECK(qf->inputs.size() == 1 && qf->outputs.size() == 1, "%s: unexpected fields", qf->name.c_str());
      MI_CALL(check_field(op, 0, true, "multiplicity", CEED_EVAL_NONE, -1));
      MI_CALL(check_field(op, 0, false, "inverse multiplicity", CEED_EVAL_NONE, 1));
      BIND_CHECK(is_offsets(fi[0]->rstr) && is_active(fi[0]->vec), "%s: multiplicity must be active with an offset restriction", qf->name.c_str());
      BIND_CHECK(is_strided(fo[0]->rstr) && is_active(fo[0]->vec) && fo[0]->rstr->num_comp == 1, "%s: output must be an active strided field",
                 qf->name.c_str());
      BIND_CHECK(fo[0]->rstr->num_elem == fi[0]->rstr->num_elem && fo[0]->rstr->elem_size == fi[0]->rstr->elem_size,
                 "%s: input and output restrictions must have the same elements", qf->name.c_str());
      p->r_q         = fi[0]->rstr;
      p->r_out       = fo[0]->rstr;
      p->nelem       = fi[0]->rstr->num_elem;
      p->kernel_name = "inverse_multiplicity";
      return 0;
    }
    case QfKind::SgsNodal: {
      const char  *names[5] = {"q", "x", "gradient velocity", "anisotropy tensor", "inverse multiplicity"};
      CeedEvalMode modes[5] = {CEED_EVAL_NONE, CEED_EVAL_INTERP, CEED_EVAL_NONE, CEED_EVAL_NONE, CEED_EVAL_NONE};
      CeedInt      sizes[5] = {5, 3, 9, 7, 1};
      BIND_CHECK(qf->inputs.size() == 5 && qf->outputs.size() == 1, "%s: unexpected fields", qf->name.c_str());
      for (int i = 0; i < 5; i++) MI_CALL(check_field(op, i, true, names[i], modes[i], sizes[i]));
      MI_CALL(check_field(op, 0, false, "km_sgs", CEED_EVAL_NONE, 6));
      BIND_CHECK(is_offsets(fi[0]->rstr) && !is_active(fi[0]->vec) && fi[0]->vec, "%s: q must be a passive nodal field", qf->name.c_str());
      BIND_CHECK(!is_active(fi[1]->vec), "%s: x must be passive", qf->name.c_str());  // not read by the closure
      BIND_CHECK(is_offsets(fi[2]->rstr) && is_active(fi[2]->vec), "%s: the velocity gradient must be the active nodal input", qf->name.c_str());
      BIND_CHECK(is_offsets(fi[3]->rstr) && !is_active(fi[3]->vec) && fi[3]->vec, "%s: the anisotropy tensor must be a passive nodal field",
                 qf->name.c_str());
      BIND_CHECK(is_strided(fi[4]->rstr) && !is_active(fi[4]->vec) && fi[4]->vec, "%s: the inverse multiplicity must be a passive strided field",
                 qf->name.c_str());
      BIND_CHECK(is_offsets(fo[0]->rstr) && is_active(fo[0]->vec), "%s: km_sgs must be the active nodal output", qf->name.c_str());
      p->r_out  = fo[0]->rstr;
      p->r_q    = fi[2]->rstr;
      p->r_aux  = fi[0]->rstr;
      p->v_aux  = fi[0]->vec;
      p->r_aux2 = fi[3]->rstr;
      p->v_aux2 = fi[3]->vec;
      p->r_im   = fi[4]->rstr;
      p->v_im   = fi[4]->vec;
      p->nelem  = p->r_out->num_elem;
      const CeedInt            esz = p->r_out->elem_size;
      const CeedElemRestriction ins[4] = {p->r_q, p->r_aux, p->r_aux2, p->r_im};
      for (CeedElemRestriction r : ins)
        BIND_CHECK(r->num_elem == p->nelem && r->elem_size == esz, "%s: nodal fields must share the element nodes of km_sgs", qf->name.c_str());
      // one closure evaluation per node: the element nodes of a km_sgs node must read the same input nodes
      BIND_CHECK(shared_outputs_read_same_inputs(p->r_out, {p->r_q, p->r_aux, p->r_aux2}),
                 "%s: element nodes sharing a km_sgs node read different input nodes", qf->name.c_str());
      p->kernel_name = "sgs_dd_nodal";
      return 0;
    }
    case QfKind::NodalSgsResidual: {
      const char  *names[3] = {"q", "qdata", "km_sgs"};
      CeedEvalMode modes[3] = {CEED_EVAL_INTERP, CEED_EVAL_NONE, CEED_EVAL_INTERP};
      CeedInt      sizes[3] = {5, 10, 6};
      BIND_CHECK(qf->inputs.size() == 3 && qf->outputs.size() == 1, "%s: unexpected fields", qf->name.c_str());
      for (int i = 0; i < 3; i++) MI_CALL(check_field(op, i, true, names[i], modes[i], sizes[i]));
      MI_CALL(check_field(op, 0, false, "Grad_v", CEED_EVAL_GRAD, 15));
      MI_CALL(bind_field_basis(op, p, fi[0], 5));
      p->r_q = fi[0]->rstr;
      p->b_q = fi[0]->basis;
      BIND_CHECK(is_active(fi[0]->vec), "%s: q must be active", qf->name.c_str());
      MI_CALL(bind_qdata_in(op, p, fi[1]));
      BIND_CHECK(!is_active(fi[2]->vec) && fi[2]->vec, "%s: km_sgs must be a passive field", qf->name.c_str());
      MI_CALL(bind_field_basis(op, p, fi[2], 6));
      p->r_in2 = fi[2]->rstr;
      p->b_in2 = fi[2]->basis;
      p->v_in2 = fi[2]->vec;
      BIND_CHECK(is_active(fo[0]->vec) && fo[0]->rstr == p->r_q && fo[0]->basis == p->b_q, "%s: Grad_v must be active on the solution space",
                 qf->name.c_str());
      p->r_out       = p->r_q;
      p->b_out       = p->b_q;
      p->ncomp_out   = 5;
      p->kernel_name = "ifunction_nodal_sgs";
      return 0;
    }
    case QfKind::DiffFilterRHS: {  // src/differential_filter.c:51-88
      const char  *names[3] = {"q", "qdata", "x"};
      CeedEvalMode modes[3] = {CEED_EVAL_INTERP, CEED_EVAL_NONE, CEED_EVAL_INTERP};
      CeedInt      sizes[3] = {5, 10, 3};
      BIND_CHECK(qf->inputs.size() == 3 && qf->outputs.size() == 2, "%s: unexpected fields (filtered state and velocity products)", qf->name.c_str());
      for (int i = 0; i < 3; i++) MI_CALL(check_field(op, i, true, names[i], modes[i], sizes[i]));
      MI_CALL(check_field(op, 0, false, "v0", CEED_EVAL_INTERP, 5));
      MI_CALL(check_field(op, 1, false, "v1", CEED_EVAL_INTERP, 6));
      MI_CALL(bind_field_basis(op, p, fi[0], 5));
      p->r_q = fi[0]->rstr;
      p->b_q = fi[0]->basis;
      BIND_CHECK(is_active(fi[0]->vec), "%s: q must be active", qf->name.c_str());
      MI_CALL(bind_qdata_in(op, p, fi[1]));
      MI_CALL(bind_coords(op, p, fi[2]));
      BIND_CHECK(is_active(fo[0]->vec) && is_active(fo[1]->vec), "%s: v0 and v1 must be active", qf->name.c_str());
      MI_CALL(bind_field_basis(op, p, fo[0], 5));
      MI_CALL(bind_field_basis(op, p, fo[1], 6));
      {  // the two filtered fields interleave per node: [5 state | 6 products], one 11-component output
        CeedElemRestriction r0 = fo[0]->rstr, r1 = fo[1]->rstr;
        bool                ok = r0->comp_stride == 1 && r1->comp_stride == 1 && r0->l_size == r1->l_size && r0->h_offsets.size() == r1->h_offsets.size();
        for (size_t i = 0; ok && i < r0->h_offsets.size(); i++) ok = r1->h_offsets[i] == r0->h_offsets[i] + 5;
        BIND_CHECK(ok, "%s: v1 must follow v0 in each node of one interleaved vector (offsets + 5, component stride 1)", qf->name.c_str());
      }
      p->r_out       = fo[0]->rstr;
      p->b_out       = fo[0]->basis;
      p->ncomp_out   = 11;
      p->kernel_name = "diff_filter_rhs";
      return 0;
    }
    case QfKind::DiffFilterLHS: {  // src/differential_filter.c:90-149
      const int    N        = qf->ncomp;
      const char  *names[5] = {"q", "Grad_q", "anisotropy tensor", "x", "qdata"};
      CeedEvalMode modes[5] = {CEED_EVAL_INTERP, CEED_EVAL_GRAD, CEED_EVAL_NONE, CEED_EVAL_INTERP, CEED_EVAL_NONE};
      CeedInt      sizes[5] = {N, 3 * N, 7, 3, 10};
      BIND_CHECK(diff_filter_ncomp_supported(N), "%s: no fused kernel for %d components", qf->name.c_str(), N);
      BIND_CHECK(qf->inputs.size() == 5 && qf->outputs.size() == 2, "%s: unexpected fields", qf->name.c_str());
      for (int i = 0; i < 5; i++) MI_CALL(check_field(op, i, true, names[i], modes[i], sizes[i]));
      MI_CALL(check_field(op, 0, false, "v", CEED_EVAL_INTERP, N));
      MI_CALL(check_field(op, 1, false, "Grad_v", CEED_EVAL_GRAD, 3 * N));
      MI_CALL(bind_field_basis(op, p, fi[0], N));
      p->r_q = p->r_out = fi[0]->rstr;
      p->b_q = p->b_out = fi[0]->basis;
      BIND_CHECK(is_active(fi[0]->vec) && is_active(fi[1]->vec) && fi[1]->rstr == p->r_q && fi[1]->basis == p->b_q,
                 "%s: q and Grad_q must share the active restriction/basis", qf->name.c_str());
      BIND_CHECK(is_strided(fi[2]->rstr) && !is_active(fi[2]->vec) && fi[2]->vec, "%s: the anisotropy tensor must be a passive strided field",
                 qf->name.c_str());
      MI_CALL(check_point_layout(op, fi[2]->rstr, p->Q, false));
      BIND_CHECK(fi[2]->rstr->num_elem == p->nelem && fi[2]->rstr->num_comp == 7 && fi[2]->vec->length >= (CeedSize)p->nelem * 7 * p->Q * p->Q * p->Q,
                 "%s: the anisotropy tensor must hold 7 components at every point", qf->name.c_str());
      p->r_aux = fi[2]->rstr;
      p->v_aux = fi[2]->vec;
      MI_CALL(bind_coords(op, p, fi[3]));
      MI_CALL(bind_qdata_in(op, p, fi[4]));
      BIND_CHECK(is_active(fo[0]->vec) && is_active(fo[1]->vec) && fo[0]->rstr == p->r_q && fo[1]->rstr == p->r_q && fo[0]->basis == p->b_q &&
                     fo[1]->basis == p->b_q,
                 "%s: v and Grad_v must be active on the filtered field", qf->name.c_str());
      p->ncomp_out   = N;
      p->kernel_name = "diff_filter_lhs";
      return 0;
    }
    case QfKind::AnisotropyCollocate: {  // src/grid_anisotropy_tensor.c:114-160
      BIND_CHECK(qf->inputs.size() == 1 && qf->outputs.size() == 1, "%s: unexpected fields", qf->name.c_str());
      MI_CALL(check_field(op, 0, true, "qdata", CEED_EVAL_NONE, 10));
      MI_CALL(check_field(op, 0, false, "v", CEED_EVAL_NONE, 7));
      BIND_CHECK(is_strided(fi[0]->rstr) && !is_active(fi[0]->vec), "%s: qdata must be a passive strided field", qf->name.c_str());
      BIND_CHECK(is_strided(fo[0]->rstr) && is_active(fo[0]->vec) && fo[0]->rstr->num_comp == 7, "%s: output must be an active strided field",
                 qf->name.c_str());
      int q = 1;
      while (q * q * q < fi[0]->rstr->elem_size) q++;
      BIND_CHECK(q * q * q == fi[0]->rstr->elem_size && fo[0]->rstr->elem_size == fi[0]->rstr->elem_size &&
                     fo[0]->rstr->num_elem == fi[0]->rstr->num_elem,
                 "%s: qdata and output must share the element points", qf->name.c_str());
      p->Q     = q;
      p->nelem = fi[0]->rstr->num_elem;
      MI_CALL(bind_qdata_in(op, p, fi[0]));
      p->r_out       = fo[0]->rstr;
      p->kernel_name = "anisotropy_collocate";
      return 0;
    }
    case QfKind::SgsTrainingNodal: {  // src/smartsim/sgs_dd_training.c:93-180
      const char  *names[5] = {"q", "velocity product", "gradient velocity", "anisotropy tensor", "inverse multiplicity"};
      CeedInt      sizes[5] = {5, 6, 9, 7, 1};
      BIND_CHECK(qf->inputs.size() == 5 && qf->outputs.size() == 1, "%s: unexpected fields", qf->name.c_str());
      for (int i = 0; i < 5; i++) MI_CALL(check_field(op, i, true, names[i], CEED_EVAL_NONE, sizes[i]));
      MI_CALL(check_field(op, 0, false, "training data", CEED_EVAL_NONE, 12));
      BIND_CHECK(is_offsets(fi[0]->rstr) && is_offsets(fi[1]->rstr) && !is_active(fi[0]->vec) && fi[0]->vec && fi[1]->vec == fi[0]->vec,
                 "%s: q and the velocity products must be passive nodal fields of one filtered vector", qf->name.c_str());
      BIND_CHECK(is_offsets(fi[2]->rstr) && is_active(fi[2]->vec), "%s: the velocity gradient must be the active nodal input", qf->name.c_str());
      BIND_CHECK(is_offsets(fi[3]->rstr) && !is_active(fi[3]->vec) && fi[3]->vec, "%s: the anisotropy tensor must be a passive nodal field",
                 qf->name.c_str());
      BIND_CHECK(is_strided(fi[4]->rstr) && !is_active(fi[4]->vec) && fi[4]->vec, "%s: the inverse multiplicity must be a passive strided field",
                 qf->name.c_str());
      BIND_CHECK(is_offsets(fo[0]->rstr) && is_active(fo[0]->vec), "%s: the training data must be the active nodal output", qf->name.c_str());
      p->r_out  = fo[0]->rstr;
      p->r_q    = fi[2]->rstr;
      p->r_aux  = fi[0]->rstr;
      p->v_aux  = fi[0]->vec;
      p->r_in2  = fi[1]->rstr;
      p->r_aux2 = fi[3]->rstr;
      p->v_aux2 = fi[3]->vec;
      p->r_im   = fi[4]->rstr;
      p->v_im   = fi[4]->vec;
      p->nelem  = p->r_out->num_elem;
      const CeedInt             esz    = p->r_out->elem_size;
      const CeedElemRestriction ins[5] = {p->r_q, p->r_aux, p->r_in2, p->r_aux2, p->r_im};
      for (CeedElemRestriction r : ins)
        BIND_CHECK(r->num_elem == p->nelem && r->elem_size == esz, "%s: nodal fields must share the element nodes of the output", qf->name.c_str());
      BIND_CHECK(shared_outputs_read_same_inputs(p->r_out, {p->r_q, p->r_aux, p->r_in2, p->r_aux2}),
                 "%s: element nodes sharing an output node read different input nodes", qf->name.c_str());
      p->kernel_name = "sgs_training_nodal";
      return 0;
    }
    default:
      break;
  }
  return MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: not an SGS field operator", qf->name.c_str());
}

static int plan_bind_fields(CeedOperator op) {
  CeedQFunction qf = op->qf;
  auto          p  = new OpPlan();
  p->kind          = qf->kind;
  p->var           = qf->state_var;
  op->plan         = p;
  auto fail        = [&](int ierr) {
    op_plan_destroy(p);
    op->plan = nullptr;
    return ierr;
  };
  int ierr = 0;
  switch (qf->kind) {
    case QfKind::Setup: {  // setupgeo.h:42-56 wired by QDataGet (src/qdata.c:154-169)
      if ((ierr = check_field(op, 0, true, "dx", CEED_EVAL_GRAD, 9))) return fail(ierr);
      if ((ierr = check_field(op, 1, true, "weight", CEED_EVAL_WEIGHT, 1))) return fail(ierr);
      if ((ierr = check_field(op, 0, false, "qdata", CEED_EVAL_NONE, 10))) return fail(ierr);
      auto *fx = op->in_fields[0], *fq = op->out_fields[0];
      if (!(is_offsets(fx->rstr) && is_tensor3(fx->basis) && fx->basis->P_1d == 2 && fx->rstr->num_comp == 3) || !is_active(fx->vec))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "Setup: 'dx' must be the active trilinear coordinate field"));
      if (!is_strided(fq->rstr) || !is_active(fq->vec) || fq->rstr->num_comp != 10)
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "Setup: qdata output must be an active strided 10-component field"));
      p->r_x   = fx->rstr;
      p->b_x   = fx->basis;
      p->Q     = fx->basis->Q_1d;
      p->Px    = 2;
      p->nelem = fx->rstr->num_elem;
      p->r_qd  = fq->rstr;
      p->qd    = strided_layout(fq->rstr);
      if ((ierr = basis_device_tables(fx->basis))) return fail(ierr);
      p->kernel_name = "setup_qdata";
      return 0;
    }
    case QfKind::IFunction: {  // newtonian.h:205-257 wired by src/setuplibceed.c:209-243
      const char *in_names[5] = {"q", "Grad_q", "q dot", "qdata", "x"};
      CeedEvalMode in_modes[5] = {CEED_EVAL_INTERP, CEED_EVAL_GRAD, CEED_EVAL_INTERP, CEED_EVAL_NONE, CEED_EVAL_INTERP};
      CeedInt      in_sizes[5] = {5, 15, 5, 10, 3};
      if (qf->inputs.size() != 5)
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: only div_diff_flux_projection_method none is supported", qf->name.c_str()));
      for (int i = 0; i < 5; i++)
        if ((ierr = check_field(op, i, true, in_names[i], in_modes[i], in_sizes[i]))) return fail(ierr);
      if ((ierr = check_field(op, 0, false, "v", CEED_EVAL_INTERP, 5))) return fail(ierr);
      if ((ierr = check_field(op, 1, false, "Grad_v", CEED_EVAL_GRAD, 15))) return fail(ierr);
      auto **fi = op->in_fields.data(), **fo = op->out_fields.data();
      if ((ierr = bind_solution_basis(op, p, fi[0]))) return fail(ierr);
      if (!(is_active(fi[0]->vec) && is_active(fi[1]->vec) && fi[1]->rstr == p->r_q && fi[1]->basis == p->b_q))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: q and Grad_q must share the active restriction/basis", qf->name.c_str()));
      if (!(fi[2]->rstr == p->r_q && fi[2]->basis == p->b_q && !is_active(fi[2]->vec)))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: 'q dot' must be a passive field on the solution space", qf->name.c_str()));
      p->v_qdot = fi[2]->vec;
      if (!is_strided(fi[3]->rstr) || is_active(fi[3]->vec))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: qdata must be a passive strided field", qf->name.c_str()));
      p->r_qd    = fi[3]->rstr;
      p->v_qdata = fi[3]->vec;
      p->qd      = strided_layout(p->r_qd);
      if ((ierr = bind_coords(op, p, fi[4]))) return fail(ierr);
      if (!(is_active(fo[0]->vec) && is_active(fo[1]->vec) && fo[0]->rstr == p->r_q && fo[1]->rstr == p->r_q))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: v and Grad_v must be active on the solution space", qf->name.c_str()));
      if (qf->outputs.size() == 3) {
        if ((ierr = check_field(op, 2, false, "jac_data", CEED_EVAL_NONE, -1))) return fail(ierr);
        if (!is_strided(fo[2]->rstr) || is_active(fo[2]->vec))
          return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: jac_data must be a passive strided field", qf->name.c_str()));
        p->njac  = qf->outputs[2].size;
        p->r_jd  = fo[2]->rstr;
        p->v_jac = fo[2]->vec;
        p->jd    = strided_layout(p->r_jd);
      }
      p->kernel_name = "ifunction_newtonian";
      return 0;
    }
    case QfKind::IJacobian: {  // newtonian.h:272-311 wired by src/setuplibceed.c:245-269
      const char  *in_names[4] = {"dq", "Grad_dq", "qdata", "jac_data"};
      CeedEvalMode in_modes[4] = {CEED_EVAL_INTERP, CEED_EVAL_GRAD, CEED_EVAL_NONE, CEED_EVAL_NONE};
      CeedInt      in_sizes[4] = {5, 15, 10, -1};
      if (qf->inputs.size() != 4) return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: unexpected input fields", qf->name.c_str()));
      for (int i = 0; i < 4; i++)
        if ((ierr = check_field(op, i, true, in_names[i], in_modes[i], in_sizes[i]))) return fail(ierr);
      if ((ierr = check_field(op, 0, false, "v", CEED_EVAL_INTERP, 5))) return fail(ierr);
      if ((ierr = check_field(op, 1, false, "Grad_v", CEED_EVAL_GRAD, 15))) return fail(ierr);
      auto **fi = op->in_fields.data(), **fo = op->out_fields.data();
      if ((ierr = bind_solution_basis(op, p, fi[0]))) return fail(ierr);
      if (!(is_active(fi[0]->vec) && is_active(fi[1]->vec) && fi[1]->rstr == p->r_q && fi[1]->basis == p->b_q))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: dq and Grad_dq must share the active restriction/basis", qf->name.c_str()));
      if (!is_strided(fi[2]->rstr) || is_active(fi[2]->vec) || !is_strided(fi[3]->rstr) || is_active(fi[3]->vec))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: qdata and jac_data must be passive strided fields", qf->name.c_str()));
      p->r_qd    = fi[2]->rstr;
      p->v_qdata = fi[2]->vec;
      p->qd      = strided_layout(p->r_qd);
      p->r_jd    = fi[3]->rstr;
      p->v_jac   = fi[3]->vec;
      p->jd      = strided_layout(p->r_jd);
      p->njac    = qf->inputs[3].size;
      if (!(is_active(fo[0]->vec) && is_active(fo[1]->vec) && fo[0]->rstr == p->r_q && fo[1]->rstr == p->r_q))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: v and Grad_v must be active on the solution space", qf->name.c_str()));
      p->kernel_name = op->ceed->jac_fp32 ? "ijacobian_newtonian_fp32" : "ijacobian_newtonian";
      return 0;
    }
    case QfKind::RHSFunction: {  // newtonian.h:133-202 wired by src/setuplibceed.c:176-203
      const char  *in_names[4] = {"q", "Grad_q", "qdata", "x"};
      CeedEvalMode in_modes[4] = {CEED_EVAL_INTERP, CEED_EVAL_GRAD, CEED_EVAL_NONE, CEED_EVAL_INTERP};
      CeedInt      in_sizes[4] = {5, 15, 10, 3};
      if (qf->inputs.size() != 4)
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: only div_diff_flux_projection_method none is supported", qf->name.c_str()));
      for (int i = 0; i < 4; i++)
        if ((ierr = check_field(op, i, true, in_names[i], in_modes[i], in_sizes[i]))) return fail(ierr);
      if ((ierr = check_field(op, 0, false, "v", CEED_EVAL_INTERP, 5))) return fail(ierr);
      if ((ierr = check_field(op, 1, false, "Grad_v", CEED_EVAL_GRAD, 15))) return fail(ierr);
      auto **fi = op->in_fields.data(), **fo = op->out_fields.data();
      if ((ierr = bind_solution_basis(op, p, fi[0]))) return fail(ierr);
      if (!(is_active(fi[0]->vec) && is_active(fi[1]->vec) && fi[1]->rstr == p->r_q && fi[1]->basis == p->b_q))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: q and Grad_q must share the active restriction/basis", qf->name.c_str()));
      if (!is_strided(fi[2]->rstr) || is_active(fi[2]->vec))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: qdata must be a passive strided field", qf->name.c_str()));
      p->r_qd    = fi[2]->rstr;
      p->v_qdata = fi[2]->vec;
      p->qd      = strided_layout(p->r_qd);
      if ((ierr = bind_coords(op, p, fi[3]))) return fail(ierr);
      if (!(is_active(fo[0]->vec) && is_active(fo[1]->vec) && fo[0]->rstr == p->r_q && fo[1]->rstr == p->r_q))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: v and Grad_v must be active on the solution space", qf->name.c_str()));
      p->kernel_name = "rhs_newtonian";
      return 0;
    }
    case QfKind::MassNewtonian:  // newtonian.h:44-82 wired by problems/newtonian.c:240-256
    case QfKind::Mass5: {        // mass.h Mass_N wired by src/setuplibceed.c:36-45
      const bool stab = qf->kind == QfKind::MassNewtonian;
      if (stab) {
        if ((ierr = check_field(op, 0, true, "q_dot", CEED_EVAL_INTERP, 5))) return fail(ierr);
        if ((ierr = check_field(op, 1, true, "q", CEED_EVAL_INTERP, 5))) return fail(ierr);
        if ((ierr = check_field(op, 2, true, "qdata", CEED_EVAL_NONE, -1))) return fail(ierr);
        if ((ierr = check_field(op, 1, false, "Grad_v", CEED_EVAL_GRAD, 15))) return fail(ierr);
      } else {
        if ((ierr = check_field(op, 0, true, "u", CEED_EVAL_INTERP, 5))) return fail(ierr);
        if ((ierr = check_field(op, 1, true, "qdata", CEED_EVAL_NONE, -1))) return fail(ierr);
      }
      if ((ierr = check_field(op, 0, false, "v", CEED_EVAL_INTERP, 5))) return fail(ierr);
      auto **fi = op->in_fields.data(), **fo = op->out_fields.data();
      if ((ierr = bind_solution_basis(op, p, fi[0]))) return fail(ierr);
      if (!is_active(fi[0]->vec)) return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: first input must be active", qf->name.c_str()));
      if (stab) {
        if (!(fi[1]->rstr == p->r_q && fi[1]->basis == p->b_q && !is_active(fi[1]->vec)))
          return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: 'q' must be a passive field on the solution space", qf->name.c_str()));
        p->v_qdot = fi[1]->vec;  // the state the stabilization is evaluated at
      }
      auto *fqd = fi[stab ? 2 : 1];
      if (!is_strided(fqd->rstr) || is_active(fqd->vec))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: qdata must be a passive strided field", qf->name.c_str()));
      p->r_qd    = fqd->rstr;
      p->v_qdata = fqd->vec;
      p->qd      = strided_layout(p->r_qd);
      for (auto *f : op->out_fields)
        if (!(is_active(f->vec) && f->rstr == p->r_q))
          return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: outputs must be active on the solution space", qf->name.c_str()));
      p->kernel_name = stab ? "mass_newtonian" : "mass";
      return 0;
    }
    case QfKind::MonitorKE: {  // monitor_totalkineticenergy.h wired by src/monitor_totalkineticenergy.c:95-104
      if ((ierr = check_field(op, 0, true, "q", CEED_EVAL_INTERP, 5))) return fail(ierr);
      if ((ierr = check_field(op, 1, true, "Grad_q", CEED_EVAL_GRAD, 15))) return fail(ierr);
      if ((ierr = check_field(op, 2, true, "q_data", CEED_EVAL_NONE, 10))) return fail(ierr);
      if ((ierr = check_field(op, 0, false, "v", CEED_EVAL_NONE, 5))) return fail(ierr);
      auto **fi = op->in_fields.data(), **fo = op->out_fields.data();
      if ((ierr = bind_solution_basis(op, p, fi[0]))) return fail(ierr);
      if (!(is_active(fi[0]->vec) && is_active(fi[1]->vec) && fi[1]->rstr == p->r_q))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: q and Grad_q must be active", qf->name.c_str()));
      if (!is_strided(fi[2]->rstr) || is_active(fi[2]->vec))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: q_data must be a passive strided field", qf->name.c_str()));
      p->r_qd    = fi[2]->rstr;
      p->v_qdata = fi[2]->vec;
      p->qd      = strided_layout(p->r_qd);
      if (!is_strided(fo[0]->rstr) || !is_active(fo[0]->vec))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: output must be an active strided field", qf->name.c_str()));
      p->r_jd        = fo[0]->rstr;
      p->jd          = strided_layout(p->r_jd);
      p->kernel_name = "monitor_total_ke";
      return 0;
    }
    case QfKind::SetupBoundary: {  // setupgeo.h:100-117 wired by the boundary qdata setup (src/qdata.c)
      if ((ierr = check_field(op, 0, true, "dx", CEED_EVAL_GRAD, 6))) return fail(ierr);
      if ((ierr = check_field(op, 1, true, "weight", CEED_EVAL_WEIGHT, 1))) return fail(ierr);
      if ((ierr = check_field(op, 0, false, "surface qdata", CEED_EVAL_NONE, 10))) return fail(ierr);
      auto *fx = op->in_fields[0], *fq = op->out_fields[0];
      if (!(is_offsets(fx->rstr) && is_tensor2(fx->basis) && fx->basis->P_1d == 2 && fx->rstr->num_comp == 3 && fx->rstr->elem_size == 4) ||
          !is_active(fx->vec))
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "SetupBoundary: 'dx' must be the active bilinear face coordinate field"));
      if (!is_strided(fq->rstr) || !is_active(fq->vec) || fq->rstr->num_comp != 10)
        return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "SetupBoundary: surface qdata must be an active strided 10-component field"));
      p->face  = true;
      p->r_x   = fx->rstr;
      p->b_x   = fx->basis;
      p->Q     = fx->basis->Q_1d;
      p->Px    = 2;
      p->nelem = fx->rstr->num_elem;
      p->r_qd  = fq->rstr;
      p->qd    = strided_layout(fq->rstr);
      if (p->Q < 2 || p->Q > 6) return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "SetupBoundary: Q_1d=%d not supported", p->Q));
      if ((ierr = basis_device_tables(fx->basis))) return fail(ierr);
      p->kernel_name = "setup_boundary";
      return 0;
    }
    case QfKind::BoundaryIntegral:  // newtonian.h:328-375 wired by problems/newtonian.c:513-569
    case QfKind::BoundaryIntegralJacobian:
      if ((ierr = bind_face_op(op, p, qf->kind == QfKind::BoundaryIntegralJacobian, 11))) return fail(ierr);
      p->kernel_name = qf->kind == QfKind::BoundaryIntegral ? "boundary_integral" : "boundary_integral_jacobian";
      return 0;
    case QfKind::Freestream:  // bc_freestream.h:14-114 wired by problems/bc_freestream.c:30-180
    case QfKind::FreestreamJacobian:
      if ((ierr = bind_face_op(op, p, qf->kind == QfKind::FreestreamJacobian, 5))) return fail(ierr);
      p->hllc        = qf->hllc;
      p->kernel_name = qf->kind == QfKind::Freestream ? "freestream" : "freestream_jacobian";
      return 0;
    case QfKind::RiemannOutflow:  // bc_outflow.h:37-305 wired by problems/bc_outflow.c:114-180
    case QfKind::RiemannOutflowJacobian:
    case QfKind::PressureOutflow:
    case QfKind::PressureOutflowJacobian: {
      const bool jac = qf->kind == QfKind::RiemannOutflowJacobian || qf->kind == QfKind::PressureOutflowJacobian;
      if ((ierr = bind_face_op(op, p, jac, 11))) return fail(ierr);
      const bool pressure = qf->kind == QfKind::PressureOutflow || qf->kind == QfKind::PressureOutflowJacobian;
      p->kernel_name      = std::string(pressure ? "pressure_outflow" : "riemann_outflow") + (jac ? "_jacobian" : "");
      return 0;
    }
    case QfKind::Slip:  // bc_slip.h:11-111 wired by problems/bc_slip.c:59-73
    case QfKind::SlipJacobian:
      if ((ierr = bind_face_op(op, p, qf->kind == QfKind::SlipJacobian, 5))) return fail(ierr);
      p->kernel_name = qf->kind == QfKind::Slip ? "slip" : "slip_jacobian";
      return 0;
    case QfKind::VelocityGradientRHS:
    case QfKind::AnisotropyRHS:
    case QfKind::MassN:
    case QfKind::InverseMultiplicity:
    case QfKind::SgsNodal:
    case QfKind::NodalSgsResidual:
    case QfKind::DiffFilterRHS:
    case QfKind::DiffFilterLHS:
    case QfKind::AnisotropyCollocate:
    case QfKind::SgsTrainingNodal:
      if ((ierr = bind_sgs_op(op, p))) return fail(ierr);
      return 0;
    default:
      break;
  }
  return fail(MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED,
                     "QFunction '%s' (%s) has no fused MI355X kernel; operators of this type cannot run on /gpu/hip/mi355x", qf->name.c_str(),
                     qf->source.c_str()));
}

static int newtonian_context(CeedOperator op, CeedMiNewtonianContext *ctx) {
  CeedQFunctionContext c = op->qf->ctx;
  if (!c || !c->data) return MI_ERR(op->ceed, CEED_ERROR_INCOMPLETE, "%s: missing QFunction context", op->qf->name.c_str());
  if (c->size != sizeof(CeedMiNewtonianContext))
    return MI_ERR(op->ceed, CEED_ERROR_INCOMPATIBLE, "%s: context size %zu does not match struct NewtonianIdealGasContext_ (%zu)",
                  op->qf->name.c_str(), c->size, sizeof(CeedMiNewtonianContext));
  memcpy(ctx, c->data, sizeof(*ctx));
  if (ctx->divFdiff_method != 0)
    return MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: divergence of diffusive flux projection is not supported", op->qf->name.c_str());
  return 0;
}

// Face sub-operators: whether the body is a Jacobian, and the context it reads (0 NewtonianIdealGasContext_,
// 1 FreestreamContext_, 2 OutflowContext_)
static bool face_kind(QfKind k, bool *jacobian, int *ctx_type) {
  switch (k) {
    case QfKind::BoundaryIntegral: *jacobian = false, *ctx_type = 0; return true;
    case QfKind::BoundaryIntegralJacobian: *jacobian = true, *ctx_type = 0; return true;
    case QfKind::Slip: *jacobian = false, *ctx_type = 0; return true;
    case QfKind::SlipJacobian: *jacobian = true, *ctx_type = 0; return true;
    case QfKind::Freestream: *jacobian = false, *ctx_type = 1; return true;
    case QfKind::FreestreamJacobian: *jacobian = true, *ctx_type = 1; return true;
    case QfKind::RiemannOutflow:
    case QfKind::PressureOutflow: *jacobian = false, *ctx_type = 2; return true;
    case QfKind::RiemannOutflowJacobian:
    case QfKind::PressureOutflowJacobian: *jacobian = true, *ctx_type = 2; return true;
    default: return false;
  }
}

template <class C>
static int sized_context(CeedOperator op, const char *what, C *out) {
  CeedQFunctionContext c = op->qf->ctx;
  if (!c || !c->data) return MI_ERR(op->ceed, CEED_ERROR_INCOMPLETE, "%s: missing QFunction context", op->qf->name.c_str());
  if (c->size != sizeof(C))
    return MI_ERR(op->ceed, CEED_ERROR_INCOMPATIBLE, "%s: context size %zu does not match struct %s (%zu)", op->qf->name.c_str(), c->size, what,
                  sizeof(C));
  memcpy(out, c->data, sizeof(C));
  return 0;
}

struct FaceContexts {
  CeedMiNewtonianContext  nctx{};
  CeedMiFreestreamContext fctx{};
  CeedMiOutflowContext    octx{};
  bool                    implicit = false;
};

static int face_contexts(CeedOperator op, int ctx_type, FaceContexts *fc) {
  if (ctx_type == 1) {
    MI_CALL(sized_context(op, "FreestreamContext_", &fc->fctx));
    fc->implicit = fc->fctx.newt_ctx.is_implicit;
  } else if (ctx_type == 2) {
    MI_CALL(sized_context(op, "OutflowContext_", &fc->octx));
    fc->implicit = fc->octx.newt_ctx.is_implicit;
  } else {
    MI_CALL(newtonian_context(op, &fc->nctx));
    fc->implicit = fc->nctx.is_implicit;
  }
  return 0;
}

// Recomputed quadrature data (qdata=recompute): the kernel rebuilds {w detJ, dXdx} from the element vertex coordinates
// when the operator's qdata vector is still the Setup operator's output for these elements, in this strided layout,
// with no write since (GeoRecord).  P = Q only; otherwise the kernel reads the qdata vector.
static void geo_args(Ceed ceed, const OpPlan *p, VolumeArgs &a) {
  a.xe   = nullptr;
  a.gtab = nullptr;
  if (!ceed->qdata_recompute || !p->v_qdata || p->P != p->Q) return;
  const GeoRecord *g = p->v_qdata->geo.get();
  if (!g || g->nelem != p->nelem || g->Q != p->Q || g->qd.sn != p->qd.sn || g->qd.sc != p->qd.sc || g->qd.se != p->qd.se) return;
  a.xe   = g->d_xe;
  a.gtab = g->d_tab;
}

// Output binding of a scattering volume kernel.  Scatter mode 0 adds into the L-vector with FP64 atomics; mode 1
// (deterministic two-pass) writes the element-node contributions to the Ceed's E-vector workspace, and
// scatter_finish sums them per node in a fixed order with launch_node_gather.
// E-vector layout of the two-pass scatter: node-sorted (the fused kernel writes each element node to its slot in
// node order; the gather streams) or element-ordered (coalesced fused-kernel writes; the gather reads each node's
// contributions by index, whole).  Measured at 32^3 p=3 (IFunction apply): element-ordered 176 us (fused kernel
// 136 us + gather 50 us, per-kernel events), node-sorted 197 us (166 + 30).
constexpr bool kSortedEvec = false;

// fp32: the FP32 Jacobian apply writes its contributions in FP32 (element order), halving the E-vector's round trip;
// the node gather sums them in FP64.  Measured at 32^3 p=3: see the FP32 IJacobian figures of the bench.
static int scatter_prepare(OpPlan *p, Ceed ceed, VolumeArgs &a, bool fp32 = false) {
  a.ev   = nullptr;
  a.ev32 = nullptr;
  a.perm = nullptr;
  if (ceed->scatter != 1) return 0;
  MI_CALL(rstr_transpose_map(p->r_q));
  double *ws;
  MI_CALL(ceed_workspace(ceed, (size_t)p->nelem * p->r_q->elem_size * 5, &ws));
  if (fp32 && !kSortedEvec) {
    a.ev32 = (float *)ws;
    return 0;
  }
  a.ev   = ws;
  a.perm = kSortedEvec ? p->r_q->d_t_perm : nullptr;
  return 0;
}

static int scatter_finish(OpPlan *p, Ceed ceed, const VolumeArgs &a, bool overwrite) {
  if (!a.ev && !a.ev32) return 0;
  CeedElemRestriction r = p->r_q;
  hipEvent_t          ev;
  MI_CALL(timing_begin(ceed, "scatter_node_gather", &ev));
  if (a.ev32)
    MI_HIP(ceed, launch_node_gather_rows_f32(r->t_nrows, r->comp_stride, r->d_t_rowptr, r->d_t_index, r->d_t_offset, a.ev32, a.v, overwrite,
                                             ceed->stream));
  else
    MI_HIP(ceed, launch_node_gather(r->t_nrows, 5, r->comp_stride, r->d_t_rowptr, a.perm ? nullptr : r->d_t_index, r->d_t_offset, a.ev, a.v,
                                    overwrite, ceed->stream));
  MI_CALL(timing_end(ceed, "scatter_node_gather", ev));
  return 0;
}

// True when op_plan_apply(op, in, out, true) writes every entry of out, so CeedOperatorApply can skip zeroing it:
// a scattering volume operator in two-pass mode whose restriction rows cover the whole L-vector.
bool op_plan_overwrites(CeedOperator op, CeedVector out) {
  if (op->composite || op->ceed->scatter != 1) return false;
  if (!op->plan && (g_no_alloc || plan_bind(op))) return false;
  const OpPlan *p = op->plan;
  switch (p->kind) {
    case QfKind::IFunction:
    case QfKind::IJacobian:
    case QfKind::RHSFunction:
    case QfKind::MassNewtonian:
    case QfKind::Mass5:
      break;
    default:
      return false;
  }
  if (rstr_transpose_map(p->r_q)) return false;
  return p->r_q->t_covers_all && out->length == p->r_q->l_size;
}

// Packed parameters of the data-driven SGS model from its QFunction context (struct SgsDDContext_,
// qfunctions/sgs_dd_model.h:15-31): b1[20], b2[6], W1[20][6], W2[6][20], bounds[6][2], alpha.  Uploaded when they
// differ from what the plan's device copy holds.
static int sgs_model_upload(CeedOperator op, OpPlan *p, CeedMiNewtonianContext *newt) {
  Ceed                 ceed = op->ceed;
  CeedQFunctionContext c    = op->qf->ctx;
  if (!c || !c->data) return MI_ERR(ceed, CEED_ERROR_INCOMPLETE, "%s: missing SgsDDContext", op->qf->name.c_str());
  if (c->size < offsetof(CeedMiSgsDDContext, data))
    return MI_ERR(ceed, CEED_ERROR_INCOMPATIBLE, "%s: context is smaller than struct SgsDDContext_", op->qf->name.c_str());
  const auto *sc = (const CeedMiSgsDDContext *)c->data;
  if (sc->num_inputs != 6 || sc->num_outputs != 6 || sc->num_neurons != 20 || (sc->num_layers != 1 && sc->num_layers != 2))
    return MI_ERR(ceed, CEED_ERROR_UNSUPPORTED, "%s: only the 6-20-6 and 6-20-20-6 models have a fused kernel", op->qf->name.c_str());
  // a second hidden layer (num_layers 2) follows out_scaling: its biases [20], then its weights [20][20]
  const size_t hidden2 = sc->offsets.out_scaling + 12;
  const size_t ends[6] = {sc->offsets.bias1 + 20, sc->offsets.bias2 + 6, sc->offsets.weight1 + 120, sc->offsets.weight2 + 120,
                          sc->offsets.out_scaling + 12, sc->num_layers == 2 ? hidden2 + 420 : 0};
  for (size_t e : ends)
    if (offsetof(CeedMiSgsDDContext, data) + e * sizeof(double) > c->size)
      return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: model parameters extend past the context", op->qf->name.c_str());
  if (sc->newt_ctx.divFdiff_method != 0)
    return MI_ERR(ceed, CEED_ERROR_UNSUPPORTED, "%s: divergence of diffusive flux projection is not supported", op->qf->name.c_str());
  *newt = sc->newt_ctx;
  std::vector<double> packed;
  packed.reserve(sgs_model_doubles());
  auto put = [&](size_t off, size_t n) { packed.insert(packed.end(), sc->data + off, sc->data + off + n); };
  put(sc->offsets.bias1, 20);
  put(sc->offsets.bias2, 6);
  put(sc->offsets.weight1, 120);
  put(sc->offsets.weight2, 120);
  put(sc->offsets.out_scaling, 12);
  packed.push_back(sc->alpha);
  packed.push_back((double)sc->num_layers);
  if (sc->num_layers == 2) put(hidden2, 420);
  else packed.insert(packed.end(), 420, 0.0);
  if (packed.size() != sgs_model_doubles()) return MI_ERR(ceed, CEED_ERROR_BACKEND, "sgs model packing mismatch");
  if (p->d_model && packed == p->model_key) return 0;
  if (!p->d_model && g_no_alloc) return MI_ERR(ceed, CEED_ERROR_UNSUPPORTED, "allocation-free apply: the SGS model is not on the device yet");
  if (!p->d_model) MI_HIP(ceed, hipMalloc((void **)&p->d_model, sizeof(double) * packed.size()));
  MI_HIP(ceed, hipMemcpyAsync(p->d_model, packed.data(), sizeof(double) * packed.size(), hipMemcpyHostToDevice, ceed->stream));
  MI_HIP(ceed, hipStreamSynchronize(ceed->stream));  // the host buffer is released on return
  p->model_key = std::move(packed);
  return 0;
}

// struct DifferentialFilterContext_ (qfunctions/differential_filter.h:16-25)
static int diff_filter_context(CeedOperator op, CeedMiDiffFilterContext *ctx) {
  CeedQFunctionContext c = op->qf->ctx;
  if (!c || !c->data) return MI_ERR(op->ceed, CEED_ERROR_INCOMPLETE, "%s: missing QFunction context", op->qf->name.c_str());
  if (c->size != sizeof(CeedMiDiffFilterContext))
    return MI_ERR(op->ceed, CEED_ERROR_INCOMPATIBLE, "%s: context size %zu does not match struct DifferentialFilterContext_ (%zu)",
                  op->qf->name.c_str(), c->size, sizeof(CeedMiDiffFilterContext));
  memcpy(ctx, c->data, sizeof(*ctx));
  if (ctx->damping_function < 0 || ctx->damping_function > 2)
    return MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "%s: unknown wall damping function %d", op->qf->name.c_str(), ctx->damping_function);
  return 0;
}

// Collocated anisotropy tensor and coordinate field of the filter operator
static int filter_aux_args(OpPlan *p, FieldArgs &a) {
  MI_CALL(vec_device_read(p->v_aux, &a.aux));
  a.auxl = strided_layout(p->r_aux);
  MI_CALL(rstr_device_offsets(p->r_x, &a.xoffsets));
  a.xcomp_stride = p->r_x->comp_stride;
  MI_CALL(vec_device_read(p->v_x, &a.x));
  a.Bx = p->b_x->d_interp_1d;
  return 0;
}

static int sgs_field_apply(CeedOperator op, OpPlan *p, CeedVector in, CeedVector out, bool overwrite) {
  Ceed        ceed = op->ceed;
  hipStream_t s    = ceed->stream;
  hipEvent_t  ev;
  if (p->kind == QfKind::InverseMultiplicity) {
    if (in->length != p->r_q->l_size || out->length < (CeedSize)p->nelem * p->r_out->elem_size)
      return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: vector sizes do not match the restrictions", op->qf->name.c_str());
    const int    *off;
    const double *d_in;
    double       *d_out;
    MI_CALL(rstr_device_offsets(p->r_q, &off));
    MI_CALL(vec_device_read(in, &d_in));
    MI_CALL(vec_device_write(out, &d_out, true));
    MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
    MI_HIP(ceed, launch_inverse_multiplicity(off, p->nelem, p->r_out->elem_size, d_in, d_out, strided_layout(p->r_out), s));
    MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
    return 0;
  }
  if (p->kind == QfKind::AnisotropyCollocate) {
    if (out->length < (CeedSize)p->nelem * 7 * p->Q * p->Q * p->Q)
      return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: output vector too short", op->qf->name.c_str());
    const double *qd;
    double       *d_out;
    MI_CALL(vec_device_read(p->v_qdata, &qd));
    MI_CALL(vec_device_write(out, &d_out, true));
    MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
    MI_HIP(ceed, launch_anisotropy_collocate(p->Q, p->nelem, qd, p->qd, d_out, strided_layout(p->r_out), s));
    MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
    return 0;
  }
  if (p->kind == QfKind::SgsTrainingNodal) {
    CeedQFunctionContext c = op->qf->ctx;
    if (!c || !c->data || c->size != sizeof(CeedMiSgsTrainingContext))
      return MI_ERR(ceed, CEED_ERROR_INCOMPATIBLE, "%s: context must be struct SGS_DD_TrainingContext_", op->qf->name.c_str());
    CeedMiNewtonianContext ctx = ((const CeedMiSgsTrainingContext *)c->data)->newt_ctx;
    if (in->length != p->r_q->l_size || out->length != p->r_out->l_size)
      return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: active vectors must have the restrictions' L-vector sizes", op->qf->name.c_str());
    CeedElemRestriction r = p->r_out;
    MI_CALL(rstr_transpose_map(r));
    SgsNodalArgs a{};
    a.nrows           = r->t_nrows;
    a.rowptr          = r->d_t_rowptr;
    a.index           = r->d_t_index;
    a.row_offset      = r->d_t_offset;
    a.out_comp_stride = r->comp_stride;
    MI_CALL(vec_device_write(out, &a.v, true));
    MI_CALL(rstr_device_offsets(p->r_aux, &a.q_offsets));
    MI_CALL(rstr_device_offsets(p->r_in2, &a.p_offsets));
    MI_CALL(rstr_device_offsets(p->r_q, &a.g_offsets));
    MI_CALL(rstr_device_offsets(p->r_aux2, &a.a_offsets));
    a.q_comp_stride = p->r_aux->comp_stride;
    a.p_comp_stride = p->r_in2->comp_stride;
    a.g_comp_stride = p->r_q->comp_stride;
    a.a_comp_stride = p->r_aux2->comp_stride;
    MI_CALL(vec_device_read(p->v_aux, &a.q));
    MI_CALL(vec_device_read(in, &a.g));
    MI_CALL(vec_device_read(p->v_aux2, &a.aniso));
    MI_CALL(vec_device_read(p->v_im, &a.inv_mult));
    a.im        = strided_layout(p->r_im);
    a.elem_size = r->elem_size;
    MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
    MI_HIP(ceed, launch_sgs_training_nodal(a, ctx, s));
    MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
    return 0;
  }
  if (p->kind == QfKind::SgsNodal) {
    CeedMiNewtonianContext ctx;
    MI_CALL(sgs_model_upload(op, p, &ctx));
    if (in->length != p->r_q->l_size || out->length != p->r_out->l_size)
      return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: active vectors must have the restrictions' L-vector sizes", op->qf->name.c_str());
    CeedElemRestriction r = p->r_out;
    MI_CALL(rstr_transpose_map(r));
    SgsNodalArgs a{};
    a.nrows      = r->t_nrows;
    a.rowptr     = r->d_t_rowptr;
    a.index      = r->d_t_index;
    a.row_offset = r->d_t_offset;
    a.out_comp_stride = r->comp_stride;
    MI_CALL(vec_device_write(out, &a.v, true));
    MI_CALL(rstr_device_offsets(p->r_aux, &a.q_offsets));
    MI_CALL(rstr_device_offsets(p->r_q, &a.g_offsets));
    MI_CALL(rstr_device_offsets(p->r_aux2, &a.a_offsets));
    a.q_comp_stride = p->r_aux->comp_stride;
    a.g_comp_stride = p->r_q->comp_stride;
    a.a_comp_stride = p->r_aux2->comp_stride;
    MI_CALL(vec_device_read(p->v_aux, &a.q));
    MI_CALL(vec_device_read(in, &a.g));
    MI_CALL(vec_device_read(p->v_aux2, &a.aniso));
    MI_CALL(vec_device_read(p->v_im, &a.inv_mult));
    a.im        = strided_layout(p->r_im);
    a.elem_size = r->elem_size;
    a.model     = p->d_model;
    MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
    MI_HIP(ceed, launch_sgs_nodal(p->var, a, ctx, s));
    MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
    return 0;
  }
  CeedMiNewtonianContext  ctx{};
  CeedMiDiffFilterContext fctx{};
  if (p->kind == QfKind::VelocityGradientRHS || p->kind == QfKind::NodalSgsResidual) MI_CALL(newtonian_context(op, &ctx));
  if (p->kind == QfKind::DiffFilterRHS || p->kind == QfKind::DiffFilterLHS) MI_CALL(diff_filter_context(op, &fctx));
  FieldArgs a{};
  a.nelem = p->nelem;
  if (p->kind == QfKind::DiffFilterLHS) MI_CALL(filter_aux_args(p, a));
  if (p->kind != QfKind::AnisotropyRHS) {
    if (in->length != p->r_q->l_size) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: input must have the L-vector size", op->qf->name.c_str());
    MI_CALL(rstr_device_offsets(p->r_q, &a.in_offsets));
    a.in_comp_stride = p->r_q->comp_stride;
    MI_CALL(vec_device_read(in, &a.u));
  }
  if (p->kind == QfKind::NodalSgsResidual) {
    MI_CALL(rstr_device_offsets(p->r_in2, &a.in2_offsets));
    a.in2_comp_stride = p->r_in2->comp_stride;
    MI_CALL(vec_device_read(p->v_in2, &a.u2));
  }
  CeedElemRestriction r = p->r_out;
  if (out->length != r->l_size) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: output must have the L-vector size", op->qf->name.c_str());
  MI_CALL(rstr_device_offsets(r, &a.out_offsets));
  a.out_comp_stride = r->comp_stride;
  MI_CALL(vec_device_write(out, &a.v, true));
  if (ceed->scatter == 1) {
    MI_CALL(rstr_transpose_map(r));
    MI_CALL(ceed_workspace(ceed, (size_t)p->nelem * r->elem_size * p->ncomp_out, &a.ev));
    a.perm = r->d_t_perm;
  }
  MI_CALL(vec_device_read(p->v_qdata, &a.qdata));
  a.qd         = p->qd;
  const int PQ = p->P * p->Q;
  CeedBasis b  = p->b_out;
  a.bt         = BasisTables{b->d_interp_1d, b->d_collo, b->d_prod, b->d_prod + PQ, b->d_prod + 2 * PQ};
  MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
  switch (p->kind) {
    case QfKind::VelocityGradientRHS:
      MI_HIP(ceed, launch_velocity_gradient_rhs(p->P, p->Q, p->var, a, ctx, s));
      break;
    case QfKind::AnisotropyRHS:
      MI_HIP(ceed, launch_anisotropy_rhs(p->P, p->Q, a, s));
      break;
    case QfKind::MassN:
      MI_HIP(ceed, launch_mass_n(p->P, p->Q, p->ncomp_out, a, s));
      break;
    case QfKind::DiffFilterRHS:
      MI_HIP(ceed, launch_diff_filter_rhs(p->P, p->Q, p->var, a, fctx, s));
      break;
    case QfKind::DiffFilterLHS:
      MI_HIP(ceed, launch_diff_filter_lhs(p->P, p->Q, p->ncomp_out, a, fctx, s));
      break;
    default:
      MI_HIP(ceed, launch_ifunction_nodal_sgs(p->P, p->Q, p->var, a, ctx, s));
      break;
  }
  MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
  if (a.ev) {
    MI_CALL(timing_begin(ceed, "scatter_node_gather", &ev));
    MI_HIP(ceed, launch_node_gather(r->t_nrows, p->ncomp_out, r->comp_stride, r->d_t_rowptr, nullptr, r->d_t_offset, a.ev, a.v, overwrite, s));
    MI_CALL(timing_end(ceed, "scatter_node_gather", ev));
  }
  return 0;
}

int op_plan_apply(CeedOperator op, CeedVector in, CeedVector out, bool overwrite) {
  if (!op->plan) {
    if (g_no_alloc) return MI_ERR(op->ceed, CEED_ERROR_UNSUPPORTED, "allocation-free apply: apply the operator once with CeedOperatorApply first");
    MI_CALL(plan_bind(op));
  }
  OpPlan      *p    = op->plan;
  Ceed         ceed = op->ceed;
  hipStream_t  s    = ceed->stream;
  hipEvent_t   ev;
  switch (p->kind) {
    case QfKind::Setup: {
      SetupArgs a{};
      a.nelem = p->nelem;
      MI_CALL(rstr_device_offsets(p->r_x, &a.xoffsets));
      a.xcomp_stride = p->r_x->comp_stride;
      MI_CALL(vec_device_read(in, &a.x));
      a.Bx  = p->b_x->d_interp_1d;
      a.Dx  = p->b_x->d_grad_1d;
      a.w1d = p->b_x->d_weight_1d;
      if (out->length < (CeedSize)p->nelem * p->Q * p->Q * p->Q * 10) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "qdata vector too short");
      MI_CALL(vec_device_write(out, &a.qdata, true));
      a.qd = p->qd;
      MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
      MI_HIP(ceed, launch_setup_qdata(2, p->Q, a, s));
      MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
      if (ceed->qdata_recompute && !g_no_alloc) {  // the geometry the values came from, for the kernels that recompute them
        auto g   = std::make_shared<GeoRecord>();
        g->nelem = p->nelem;
        g->Q     = p->Q;
        g->qd    = p->qd;
        const int Q = p->Q;
        MI_HIP(ceed, hipMalloc((void **)&g->d_xe, sizeof(double) * 24 * std::max<size_t>((size_t)p->nelem, 1)));
        MI_HIP(ceed, hipMalloc((void **)&g->d_tab, sizeof(double) * 5 * Q));
        MI_HIP(ceed, hipMemcpyAsync(g->d_tab, a.Bx, sizeof(double) * 2 * Q, hipMemcpyDeviceToDevice, s));
        MI_HIP(ceed, hipMemcpyAsync(g->d_tab + 2 * Q, a.Dx, sizeof(double) * 2 * Q, hipMemcpyDeviceToDevice, s));
        MI_HIP(ceed, hipMemcpyAsync(g->d_tab + 4 * Q, a.w1d, sizeof(double) * Q, hipMemcpyDeviceToDevice, s));
        if (p->nelem) MI_HIP(ceed, launch_gather_vertex_coords(p->nelem, a.xoffsets, a.x, a.xcomp_stride, g->d_xe, s));
        out->geo = g;
      }
      return 0;
    }
    case QfKind::IFunction:
    case QfKind::IJacobian: {
      CeedMiNewtonianContext ctx;
      MI_CALL(newtonian_context(op, &ctx));
      if (in->length != p->r_q->l_size || out->length != p->r_q->l_size)
        return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: active vectors must have the restriction's L-vector size", op->qf->name.c_str());
      VolumeArgs a{};
      a.nelem = p->nelem;
      MI_CALL(rstr_device_offsets(p->r_q, &a.offsets));
      a.comp_stride = p->r_q->comp_stride;
      MI_CALL(vec_device_read(in, &a.u));
      MI_CALL(vec_device_write(out, &a.v, true));
      MI_CALL(scatter_prepare(p, ceed, a, p->kind == QfKind::IJacobian && ceed->jac_fp32));
      MI_CALL(vec_device_read(p->v_qdata, &a.qdata));
      a.qd = p->qd;
      a.jd = p->jd;
      geo_args(ceed, p, a);
      const int PQ = p->P * p->Q;
      a.bt = BasisTables{p->b_q->d_interp_1d, p->b_q->d_collo, p->b_q->d_prod, p->b_q->d_prod + PQ, p->b_q->d_prod + 2 * PQ, p->b_q->d_grad_1d};
      if (p->kind == QfKind::IFunction) {
        const int need = 14 + (ctx.idl_enable ? 1 : 0);
        if (p->v_jac && p->njac < need) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "jac_data needs %d components", need);
        MI_CALL(vec_device_read(p->v_qdot, &a.u2));
        MI_CALL(rstr_device_offsets(p->r_x, &a.xoffsets));
        MI_CALL(vec_device_read(p->v_x, &a.x));
        a.Bx = p->b_x->d_interp_1d;
        if (p->v_jac) {
          if (ceed->jac_fp32) MI_CALL(vec_device_shadow32(p->v_jac, &a.jac32, true));
          else MI_CALL(vec_device_write(p->v_jac, &a.jac, true));
        }
        MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
        MI_HIP(ceed, launch_ifunction(p->P, p->Q, p->var, a, ctx, s));
        MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
        MI_CALL(scatter_finish(p, ceed, a, overwrite));
      } else {
        const int need = 14 + (ctx.idl_enable ? 1 : 0);
        if (p->njac < need) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "jac_data needs %d components", need);
        if (ceed->jac_fp32) MI_CALL(vec_device_shadow32(p->v_jac, &a.jac32, false));
        else MI_CALL(vec_device_read(p->v_jac, &a.jacr));
        MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
        MI_HIP(ceed, launch_ijacobian(p->P, p->Q, p->var, ceed->jac_fp32, a, ctx, s));
        MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
        MI_CALL(scatter_finish(p, ceed, a, overwrite));
      }
      return 0;
    }
    case QfKind::RHSFunction:
    case QfKind::MassNewtonian:
    case QfKind::Mass5:
    case QfKind::MonitorKE: {
      CeedMiNewtonianContext ctx{};
      if (p->kind != QfKind::Mass5) MI_CALL(newtonian_context(op, &ctx));
      if (in->length != p->r_q->l_size) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: input must have the L-vector size", op->qf->name.c_str());
      VolumeArgs a{};
      a.nelem = p->nelem;
      MI_CALL(rstr_device_offsets(p->r_q, &a.offsets));
      a.comp_stride = p->r_q->comp_stride;
      MI_CALL(vec_device_read(in, &a.u));
      MI_CALL(vec_device_read(p->v_qdata, &a.qdata));
      a.qd         = p->qd;
      const int PQ = p->P * p->Q;
      a.bt         = BasisTables{p->b_q->d_interp_1d, p->b_q->d_collo, p->b_q->d_prod, p->b_q->d_prod + PQ, p->b_q->d_prod + 2 * PQ, p->b_q->d_grad_1d};
      if (p->kind == QfKind::MonitorKE) {
        if (out->length < p->r_jd->l_size) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "monitor output too short");
        MI_CALL(vec_device_write(out, &a.jac, true));
        a.jd = p->jd;
      } else {
        if (out->length != p->r_q->l_size) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: output must have the L-vector size", op->qf->name.c_str());
        MI_CALL(vec_device_write(out, &a.v, true));
        MI_CALL(scatter_prepare(p, ceed, a));
      }
      if (p->kind == QfKind::RHSFunction) {
        MI_CALL(rstr_device_offsets(p->r_x, &a.xoffsets));
        MI_CALL(vec_device_read(p->v_x, &a.x));
        a.Bx = p->b_x->d_interp_1d;
        geo_args(ceed, p, a);
      }
      if (p->kind == QfKind::MassNewtonian) MI_CALL(vec_device_read(p->v_qdot, &a.u2));
      MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
      if (p->kind == QfKind::RHSFunction) MI_HIP(ceed, launch_rhs(p->P, p->Q, a, ctx, s));
      else if (p->kind == QfKind::MonitorKE) MI_HIP(ceed, launch_monitor_ke(p->P, p->Q, p->var, a, ctx, s));
      else MI_HIP(ceed, launch_mass(p->P, p->Q, p->kind == QfKind::MassNewtonian, a, ctx, s));
      MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
      MI_CALL(scatter_finish(p, ceed, a, overwrite));
      return 0;
    }
    case QfKind::SetupBoundary: {
      SetupArgs a{};
      a.nelem = p->nelem;
      MI_CALL(rstr_device_offsets(p->r_x, &a.xoffsets));
      a.xcomp_stride = p->r_x->comp_stride;
      MI_CALL(vec_device_read(in, &a.x));
      a.Bx  = p->b_x->d_interp_1d;
      a.Dx  = p->b_x->d_grad_1d;
      a.w1d = p->b_x->d_weight_1d;
      if (out->length < (CeedSize)p->nelem * p->Q * p->Q * 10) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "surface qdata vector too short");
      MI_CALL(vec_device_write(out, &a.qdata, true));
      a.qd = p->qd;
      MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
      MI_HIP(ceed, launch_setup_boundary(p->Q, a, s));
      MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
      return 0;
    }
    case QfKind::BoundaryIntegral:
    case QfKind::BoundaryIntegralJacobian:
    case QfKind::Freestream:
    case QfKind::FreestreamJacobian:
    case QfKind::RiemannOutflow:
    case QfKind::RiemannOutflowJacobian:
    case QfKind::PressureOutflow:
    case QfKind::PressureOutflowJacobian:
    case QfKind::Slip:
    case QfKind::SlipJacobian: {
      bool jacobian = false;
      int  ctx_type = 0;
      face_kind(p->kind, &jacobian, &ctx_type);
      if (in->length != p->r_q->l_size || out->length != p->r_q->l_size)
        return MI_ERR(ceed, CEED_ERROR_DIMENSION, "%s: active vectors must have the face restriction's L-vector size", op->qf->name.c_str());
      FaceContexts fc;
      MI_CALL(face_contexts(op, ctx_type, &fc));
      const bool implicit = fc.implicit;
      VolumeArgs a{};
      a.nelem = p->nelem;
      MI_CALL(rstr_device_offsets(p->r_q, &a.offsets));
      a.comp_stride = p->r_q->comp_stride;
      MI_CALL(vec_device_read(in, &a.u));
      MI_CALL(vec_device_write(out, &a.v, true));
      MI_CALL(scatter_prepare(p, ceed, a));
      MI_CALL(vec_device_read(p->v_qdata, &a.qdata));
      a.qd = p->qd;
      a.jd = p->jd;
      a.bt = BasisTables{p->b_q->d_interp_1d, p->b_q->d_collo, nullptr, nullptr, nullptr};
      if (p->v_jac) {
        if (jacobian) {
          if (ceed->jac_fp32) MI_CALL(vec_device_shadow32(p->v_jac, &a.jac32, false));
          else MI_CALL(vec_device_read(p->v_jac, &a.jacr));
        } else if (implicit) {
          if (ceed->jac_fp32) MI_CALL(vec_device_shadow32(p->v_jac, &a.jac32, true));
          else MI_CALL(vec_device_write(p->v_jac, &a.jac, true));
        }
      } else if (jacobian) {
        return MI_ERR(ceed, CEED_ERROR_INCOMPLETE, "%s: surface jacobian data is required", op->qf->name.c_str());
      }
      const bool fp32 = ceed->jac_fp32;
      MI_CALL(timing_begin(ceed, p->kernel_name.c_str(), &ev));
      switch (p->kind) {
        case QfKind::Freestream:
        case QfKind::FreestreamJacobian: MI_HIP(ceed, launch_face_freestream(p->P, p->Q, p->var, p->hllc, jacobian, fp32, a, fc.fctx, s)); break;
        case QfKind::RiemannOutflow:
        case QfKind::RiemannOutflowJacobian: MI_HIP(ceed, launch_face_outflow(p->P, p->Q, p->var, false, jacobian, fp32, a, fc.octx, s)); break;
        case QfKind::PressureOutflow:
        case QfKind::PressureOutflowJacobian: MI_HIP(ceed, launch_face_outflow(p->P, p->Q, p->var, true, jacobian, fp32, a, fc.octx, s)); break;
        case QfKind::Slip:
        case QfKind::SlipJacobian: MI_HIP(ceed, launch_face_slip(p->P, p->Q, p->var, jacobian, fp32, a, fc.nctx, s)); break;
        default: MI_HIP(ceed, launch_face_boundary_integral(p->P, p->Q, p->var, jacobian, fp32, a, fc.nctx, s)); break;
      }
      MI_CALL(timing_end(ceed, p->kernel_name.c_str(), ev));
      MI_CALL(scatter_finish(p, ceed, a, overwrite));
      return 0;
    }
    case QfKind::VelocityGradientRHS:
    case QfKind::AnisotropyRHS:
    case QfKind::MassN:
    case QfKind::InverseMultiplicity:
    case QfKind::SgsNodal:
    case QfKind::NodalSgsResidual:
    case QfKind::DiffFilterRHS:
    case QfKind::DiffFilterLHS:
    case QfKind::AnisotropyCollocate:
    case QfKind::SgsTrainingNodal:
      return sgs_field_apply(op, p, in, out, overwrite);
    default:
      return MI_ERR(ceed, CEED_ERROR_UNSUPPORTED, "operator kind not executable");
  }
}

// The point blocks go through the element-ordered E-vector (25 doubles per element node of workspace) and a per-node
// gather that adds them to the output in a fixed order: deterministic, and measured 2.4 -> 1.0 ms at 32^3 p=3 against
// 25 FP64 atomics per element node.  The Jacobi diagonal of the volume operators (5 atomics per element node) adds
// atomically (measured 0.62 ms atomic, 0.68 ms two-pass).
static int diag_evec_begin(Ceed ceed, OpPlan *p, VolumeArgs &a, bool point_block) {
  a.ev = nullptr;
  if (!point_block) return 0;
  MI_CALL(rstr_transpose_map(p->r_q));
  return ceed_workspace(ceed, (size_t)p->nelem * p->r_q->elem_size * 25, &a.ev);
}
static int diag_evec_finish(Ceed ceed, OpPlan *p, const VolumeArgs &a) {
  if (!a.ev) return 0;
  const CeedElemRestriction r = p->r_q;
  MI_HIP(ceed, launch_node_gather_blocks(r->t_nrows, r->d_t_rowptr, r->d_t_index, r->d_t_offset, a.ev, a.v, ceed->stream));
  return 0;
}

// The diagonals of the projection and filter operators (Mass_N, the filter LHS) through the node-sorted E-vector and
// the fixed-order node gather, as their applies do: the Jacobi preconditioners of those CG solves are then bitwise
// reproducible, and so are the solves.
static int field_diag_evec_begin(Ceed ceed, OpPlan *p, CeedElemRestriction r, FieldArgs &a) {
  MI_CALL(rstr_transpose_map(r));
  MI_CALL(ceed_workspace(ceed, (size_t)p->nelem * r->elem_size * p->ncomp_out, &a.ev));
  a.perm = r->d_t_perm;
  return 0;
}
static int field_diag_evec_finish(Ceed ceed, OpPlan *p, CeedElemRestriction r, const FieldArgs &a) {
  MI_HIP(ceed, launch_node_gather(r->t_nrows, p->ncomp_out, r->comp_stride, r->d_t_rowptr, nullptr, r->d_t_offset, a.ev, a.v, false, ceed->stream));
  return 0;
}

// Point-Jacobi diagonal (MatGetDiagonal_Ceed -> CeedOperatorLinearAssembleDiagonal, src/mat-ceed.c:1603-1630) or, with
// point_block, the [node][c_out][c_in] point-block diagonal (PCVPBJACOBI's MatInvertVariableBlockDiagonal input,
// CeedOperatorLinearAssemblePointBlockDiagonal) of an interleaved (comp_stride 1) 5-component operator.
int op_plan_diagonal(CeedOperator op, CeedVector diag, bool point_block) {
  if (!op->plan) MI_CALL(plan_bind(op));
  OpPlan *p    = op->plan;
  Ceed    ceed = op->ceed;
  if (p->kind == QfKind::MassN) {  // Mass_N of the SGS projections: the scalar mass diagonal in every component
    if (point_block) return MI_ERR(ceed, CEED_ERROR_UNSUPPORTED, "point-block diagonal of '%s' is not provided", op->qf->name.c_str());
    if (diag->length != p->r_out->l_size) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "diagonal vector has the wrong size");
    FieldArgs a{};
    a.nelem = p->nelem;
    MI_CALL(rstr_device_offsets(p->r_out, &a.out_offsets));
    a.out_comp_stride = p->r_out->comp_stride;
    MI_CALL(vec_device_write(diag, &a.v, true));
    MI_CALL(vec_device_read(p->v_qdata, &a.qdata));
    a.qd         = p->qd;
    const int PQ = p->P * p->Q;
    a.bt         = BasisTables{p->b_out->d_interp_1d, p->b_out->d_collo, p->b_out->d_prod, p->b_out->d_prod + PQ, p->b_out->d_prod + 2 * PQ};
    MI_CALL(field_diag_evec_begin(ceed, p, p->r_out, a));
    hipEvent_t ev;
    MI_CALL(timing_begin(ceed, "mass_n_diagonal", &ev));
    MI_HIP(ceed, launch_mass_n_diagonal(p->P, p->Q, p->ncomp_out, a, ceed->stream));
    MI_CALL(field_diag_evec_finish(ceed, p, p->r_out, a));
    MI_CALL(timing_end(ceed, "mass_n_diagonal", ev));
    return 0;
  }
  if (p->kind == QfKind::DiffFilterLHS) {  // Jacobi diagonal of the filter solve (PCJACOBI, src/differential_filter.c:178-181)
    if (point_block) return MI_ERR(ceed, CEED_ERROR_UNSUPPORTED, "point-block diagonal of '%s' is not provided", op->qf->name.c_str());
    if (diag->length != p->r_q->l_size) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "diagonal vector has the wrong size");
    CeedMiDiffFilterContext fctx;
    MI_CALL(diff_filter_context(op, &fctx));
    FieldArgs a{};
    a.nelem = p->nelem;
    MI_CALL(filter_aux_args(p, a));
    MI_CALL(rstr_device_offsets(p->r_q, &a.out_offsets));
    a.out_comp_stride = p->r_q->comp_stride;
    MI_CALL(vec_device_write(diag, &a.v, true));
    MI_CALL(vec_device_read(p->v_qdata, &a.qdata));
    a.qd         = p->qd;
    const int PQ = p->P * p->Q;
    a.bt         = BasisTables{p->b_q->d_interp_1d, p->b_q->d_collo, p->b_q->d_prod, p->b_q->d_prod + PQ, p->b_q->d_prod + 2 * PQ};
    MI_CALL(field_diag_evec_begin(ceed, p, p->r_q, a));
    hipEvent_t ev;
    MI_CALL(timing_begin(ceed, "diff_filter_lhs_diagonal", &ev));
    MI_HIP(ceed, launch_diff_filter_lhs_diagonal(p->P, p->Q, p->ncomp_out, a, fctx, ceed->stream));
    MI_CALL(field_diag_evec_finish(ceed, p, p->r_q, a));
    MI_CALL(timing_end(ceed, "diff_filter_lhs_diagonal", ev));
    return 0;
  }
  const CeedSize want = point_block ? 5 * p->r_q->l_size : p->r_q->l_size;
  if (point_block && p->r_q->comp_stride != 1)
    return MI_ERR(ceed, CEED_ERROR_UNSUPPORTED, "point-block diagonal assembly needs an interleaved (comp_stride 1) restriction");
  if (p->kind == QfKind::MassNewtonian || p->kind == QfKind::Mass5) {  // mass matrices for the explicit path
    CeedMiNewtonianContext ctx{};
    if (p->kind == QfKind::MassNewtonian) MI_CALL(newtonian_context(op, &ctx));
    if (diag->length != want) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "diagonal vector has the wrong size");
    VolumeArgs a{};
    a.nelem = p->nelem;
    MI_CALL(rstr_device_offsets(p->r_q, &a.offsets));
    a.comp_stride = p->r_q->comp_stride;
    MI_CALL(vec_device_write(diag, &a.v, true));
    MI_CALL(vec_device_read(p->v_qdata, &a.qdata));
    if (p->kind == QfKind::MassNewtonian) MI_CALL(vec_device_read(p->v_qdot, &a.u2));
    a.qd         = p->qd;
    const int PQ = p->P * p->Q;
    a.bt         = BasisTables{p->b_q->d_interp_1d, p->b_q->d_collo, p->b_q->d_prod, p->b_q->d_prod + PQ, p->b_q->d_prod + 2 * PQ};
    hipEvent_t ev;
    MI_CALL(timing_begin(ceed, "mass_diagonal", &ev));
    MI_CALL(diag_evec_begin(ceed, p, a, point_block));
    if (point_block) MI_HIP(ceed, launch_mass_point_block_diagonal(p->P, p->Q, p->kind == QfKind::MassNewtonian, a, ctx, ceed->stream));
    else MI_HIP(ceed, launch_mass_diagonal(p->P, p->Q, p->kind == QfKind::MassNewtonian, a, ctx, ceed->stream));
    MI_CALL(diag_evec_finish(ceed, p, a));
    MI_CALL(timing_end(ceed, "mass_diagonal", ev));
    return 0;
  }
  bool face_jac = false;
  int  face_ctx = 0;
  if (face_kind(p->kind, &face_jac, &face_ctx) && face_jac) {  // face sub-operators
    FaceContexts fc;
    MI_CALL(face_contexts(op, face_ctx, &fc));
    if (diag->length != want) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "diagonal vector has the wrong size");
    VolumeArgs a{};
    a.nelem = p->nelem;
    MI_CALL(rstr_device_offsets(p->r_q, &a.offsets));
    a.comp_stride = p->r_q->comp_stride;
    MI_CALL(vec_device_write(diag, &a.v, true));
    MI_CALL(vec_device_read(p->v_qdata, &a.qdata));
    a.qd         = p->qd;
    a.jd         = p->jd;
    const int PQ = p->P * p->Q;
    a.bt         = BasisTables{p->b_q->d_interp_1d, p->b_q->d_collo, p->b_q->d_prod, p->b_q->d_prod + PQ, p->b_q->d_prod + 2 * PQ};
    if (ceed->jac_fp32) MI_CALL(vec_device_shadow32(p->v_jac, &a.jac32, false));
    else MI_CALL(vec_device_read(p->v_jac, &a.jacr));
    hipEvent_t ev;
    MI_CALL(timing_begin(ceed, "face_jacobian_diagonal", &ev));
    switch (p->kind) {
      case QfKind::RiemannOutflowJacobian:
        MI_HIP(ceed, launch_face_bc_jacobian_diagonal(point_block, p->P, p->Q, p->var, kFaceRiemannOutflow, ceed->jac_fp32, a, fc.nctx, fc.octx,
                                                      ceed->stream));
        break;
      case QfKind::PressureOutflowJacobian:
        MI_HIP(ceed, launch_face_bc_jacobian_diagonal(point_block, p->P, p->Q, p->var, kFacePressureOutflow, ceed->jac_fp32, a, fc.nctx, fc.octx,
                                                      ceed->stream));
        break;
      case QfKind::SlipJacobian:
        MI_HIP(ceed, launch_face_bc_jacobian_diagonal(point_block, p->P, p->Q, p->var, kFaceSlip, ceed->jac_fp32, a, fc.nctx, fc.octx, ceed->stream));
        break;
      default:
        MI_HIP(ceed, launch_face_jacobian_diagonal(point_block, p->P, p->Q, p->var, p->kind == QfKind::FreestreamJacobian, p->hllc, ceed->jac_fp32, a,
                                                   fc.nctx, fc.fctx, ceed->stream));
        break;
    }
    MI_CALL(timing_end(ceed, "face_jacobian_diagonal", ev));
    return 0;
  }
  if (p->kind != QfKind::IJacobian)
    return MI_ERR(ceed, CEED_ERROR_UNSUPPORTED, "diagonal assembly of '%s' is not provided", op->qf->name.c_str());
  CeedMiNewtonianContext ctx;
  MI_CALL(newtonian_context(op, &ctx));
  if (diag->length != want) return MI_ERR(ceed, CEED_ERROR_DIMENSION, "diagonal vector has the wrong size");
  VolumeArgs a{};
  a.nelem = p->nelem;
  MI_CALL(rstr_device_offsets(p->r_q, &a.offsets));
  a.comp_stride = p->r_q->comp_stride;
  MI_CALL(vec_device_write(diag, &a.v, true));
  MI_CALL(vec_device_read(p->v_qdata, &a.qdata));
  a.qd         = p->qd;
  a.jd         = p->jd;
  const int PQ = p->P * p->Q;
  a.bt         = BasisTables{p->b_q->d_interp_1d, p->b_q->d_collo, p->b_q->d_prod, p->b_q->d_prod + PQ, p->b_q->d_prod + 2 * PQ};
  if (ceed->jac_fp32) MI_CALL(vec_device_shadow32(p->v_jac, &a.jac32, false));
  else MI_CALL(vec_device_read(p->v_jac, &a.jacr));
  hipEvent_t ev;
  MI_CALL(timing_begin(ceed, "ijacobian_diagonal", &ev));
  MI_CALL(diag_evec_begin(ceed, p, a, point_block));
  if (point_block) MI_HIP(ceed, launch_ijacobian_point_block_diagonal(p->P, p->Q, p->var, ceed->jac_fp32, a, ctx, ceed->stream));
  else MI_HIP(ceed, launch_ijacobian_diagonal(p->P, p->Q, p->var, ceed->jac_fp32, a, ctx, ceed->stream));
  MI_CALL(diag_evec_finish(ceed, p, a));
  MI_CALL(timing_end(ceed, "ijacobian_diagonal", ev));
  return 0;
}
