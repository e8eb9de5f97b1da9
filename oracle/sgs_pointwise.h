// CPU ORACLE -- test infrastructure only (see ns_pointwise.h).
//
// Pointwise restatement of HONEE's data-driven subgrid-stress (SGS) closure:
//   * the max-pivot Jacobi eigenvalue algorithm for symmetric 3x3 matrices with decreasing-eigenvalue row sorting
//     (qfunctions/utils_eigensolver_jacobi.h:29-335, itself after jacobi_pd),
//   * the element anisotropy tensor A_ij = Delta_ij / ||Delta||, Delta = g^(-1/2) (qfunctions/grid_anisotropy_tensor.h:14-31),
//   * the model inputs (eigenframe strain rates, vorticity, viscous scale) and outputs (de-normalization, rotation back to
//     the physical frame) (qfunctions/sgs_dd_utils.h:17-127),
//   * the 6-20-6 MLP with leaky ReLU (qfunctions/sgs_dd_model.h:33-53), optionally with a second hidden layer,
//   * the SGS flux added to the residual (qfunctions/sgs_dd_model.h:199-213).
// Kelvin-Mandel packing as in qfunctions/utils.h:239-258.
#pragma once

#include <math.h>

#include "ns_pointwise.h"

#define O_SGS_EPSILON 0x1p-52  // CEED_EPSILON

typedef struct {
  int           num_inputs, num_outputs, num_neurons;
  double        alpha;
  const double *b1, *b2;  // [num_neurons], [num_outputs]
  const double *W1;       // [num_neurons][num_inputs]
  const double *W2;       // [num_outputs][num_neurons]
  const double *bounds;   // [num_outputs][2] new (min, max) of the de-normalization
  // second hidden layer (an extension of the reference's model; Wh null for one hidden layer): W2 then maps it to the
  // outputs
  const double *bh;  // [num_neurons]
  const double *Wh;  // [num_neurons][num_neurons]
} OSgsModel;

static inline void o_km_pack(const double A[3][3], double v[6]) {
  const double w = sqrt(2.);
  v[0]           = A[0][0];
  v[1]           = A[1][1];
  v[2]           = A[2][2];
  v[3]           = A[2][1] * w;
  v[4]           = A[2][0] * w;
  v[5]           = A[1][0] * w;
}

// ---- Jacobi eigenvalue algorithm (3x3, symmetric) ----
// Column of the largest |A[i][j]|, j > i (the first one on ties)
static inline int o_jac_max_row(const double A[3][3], int i) {
  int jm = i + 1;
  for (int j = i + 2; j < 3; j++)
    if (fabs(A[i][j]) > fabs(A[i][jm])) jm = j;
  return jm;
}

// Rotation in the (i, j) plane that zeroes A[i][j]: c, s, t = cos, sin, tan of the angle
static inline void o_jac_rotation(const double A[3][3], int i, int j, double *c, double *s, double *t) {
  *t               = 1.0;
  const double djj = A[j][j] - A[i][i];
  if (djj != 0.0) {
    *t = 0.0;
    if (A[i][j] != 0.0) {
      const double kappa = djj / (2.0 * A[i][j]);
      *t                 = 1.0 / (sqrt(1 + kappa * kappa) + fabs(kappa));
      if (kappa < 0.0) *t = -*t;
    }
  }
  *c = 1.0 / sqrt(1 + *t * *t);
  *s = *c * *t;
}

// A <- R^T A R for the rotation above (upper triangle kept; A[i][j] becomes 0)
static inline void o_jac_apply(double A[3][3], int i, int j, double c, double s, double t) {
  A[i][i] -= t * A[i][j];
  A[j][j] += t * A[i][j];
  A[i][j] = 0.0;
  for (int w = 0; w < 3; w++) {
    if (w == i || w == j) continue;
    // upper-triangle entries (min, max) of the pairs (w, i) and (w, j)
    double     *wi  = w < i ? &A[w][i] : &A[i][w];
    double     *wj  = w < j ? &A[w][j] : &A[j][w];
    const double ai = *wi, aj = *wj;
    *wi             = c * ai - s * aj;
    *wj             = s * ai + c * aj;
  }
}

// Eigenvalues (decreasing) and eigenvectors (rows of E) of the symmetric matrix A (overwritten); at most
// max_sweeps * 3 rotations.  Returns the number of rotations attempted, 0 if the limit was reached.
static inline int o_diagonalize3(double A[3][3], double eval[3], double E[3][3], int max_sweeps) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) E[i][j] = i == j ? 1.0 : 0.0;
  int       mr[2]    = {o_jac_max_row(A, 0), o_jac_max_row(A, 1)};
  const int max_iter = max_sweeps * 3;
  int       it;
  for (it = 1; it <= max_iter; it++) {
    int    i = 0, j = mr[0];
    double amax = fabs(A[0][mr[0]]);
    if (fabs(A[1][mr[1]]) > amax) i = 1, j = mr[1];
    if ((A[i][i] + A[i][j] == A[i][i]) && (A[j][j] + A[i][j] == A[j][j])) {
      A[i][j] = 0.0;
      mr[i]   = o_jac_max_row(A, i);
    }
    if (A[i][j] == 0.0) break;
    double c, s, t;
    o_jac_rotation(A, i, j, &c, &s, &t);
    o_jac_apply(A, i, j, c, s, t);
    for (int v = 0; v < 3; v++) {  // E <- R^T E (rows i, j)
      const double ei = E[i][v];
      E[i][v]         = c * E[i][v] - s * E[j][v];
      E[j][v]         = s * ei + c * E[j][v];
    }
    mr[0] = o_jac_max_row(A, 0);
    mr[1] = o_jac_max_row(A, 1);
  }
  for (int i = 0; i < 3; i++) eval[i] = A[i][i];
  for (int i = 0; i < 2; i++) {  // selection sort, decreasing, rows of E follow
    int im = i;
    for (int j = i + 1; j < 3; j++)
      if (eval[j] > eval[im]) im = j;
    double tv = eval[i];
    eval[i] = eval[im], eval[im] = tv;
    for (int k = 0; k < 3; k++) {
      double te = E[i][k];
      E[i][k] = E[im][k], E[im][k] = te;
    }
  }
  return it > max_iter ? 0 : it;
}

// ---- anisotropy tensor ----
// km_g = KM(dXdx^T dXdx); A_ij = E^T diag(1/sqrt(eval)) E / ||1/sqrt(eval)||, delta = 2 ||1/sqrt(eval)||
static inline void o_anisotropy_tensor(const double dXdx[3][3], double km_A[6], double *delta) {
  double g[3][3] = {{0}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) g[i][j] += dXdx[k][i] * dXdx[k][j];
  double km_g[6], gu[3][3];
  o_km_pack(g, km_g);
  o_km_unpack(km_g, gu);
  double ev[3], E[3][3];
  o_diagonalize3(gu, ev, E, 15);
  for (int i = 0; i < 3; i++) ev[i] = 1 / sqrt(ev[i]);
  double A[3][3] = {{0}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) A[i][j] += E[k][i] * (ev[k] * E[k][j]);
  const double d = sqrt(ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) A[i][j] *= 1 / d;
  o_km_pack(A, km_A);
  *delta = 2 * d;
}

// ---- model inputs / inference / outputs ----
static inline void o_orient_basis(double E[3][3], const double v[3]) {
  double al[3] = {0, 0, 0};
  for (int i = 0; i < 3; i++) al[i] += E[i][0] * v[0] + E[i][1] * v[1] + E[i][2] * v[2];
  if (al[0] < 0)
    for (int k = 0; k < 3; k++) E[0][k] *= -1;
  if (al[2] < 0)
    for (int k = 0; k < 3; k++) E[2][k] *= -1;
  const double cr[3] = {E[2][1] * E[0][2] - E[2][2] * E[0][1], E[2][2] * E[0][0] - E[2][0] * E[0][2], E[2][0] * E[0][1] - E[2][1] * E[0][0]};
  if (o_dot3(cr, E[1]) < 0)
    for (int k = 0; k < 3; k++) E[1][k] *= -1;
}

static inline void o_sgs_inputs(const double gu_aniso[3][3], const double km_A[6], double delta, double nu, double E[3][3], double in[6],
                                double *gmag) {
  double A[3][3], gi[3][3] = {{0}};
  o_km_unpack(km_A, A);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) gi[i][j] += gu_aniso[i][k] * A[k][j];
  const double w      = 1 / sqrt(2.);
  const double kms[6] = {gi[0][0], gi[1][1], gi[2][2], w * (gi[1][2] + gi[2][1]), w * (gi[0][2] + gi[2][0]), w * (gi[0][1] + gi[1][0])};
  double       S[3][3], sev[3];
  o_km_unpack(kms, S);
  o_diagonalize3(S, sev, E, 5);
  const double r12 = 0.5 * (gi[1][2] - gi[2][1]), r02 = 0.5 * (gi[0][2] - gi[2][0]), r01 = 0.5 * (gi[0][1] - gi[1][0]);
  const double om[3] = {-2 * r12, 2 * r02, -2 * r01};
  o_orient_basis(E, om);
  double vs[3] = {0, 0, 0};
  for (int i = 0; i < 3; i++) vs[i] += E[i][0] * om[0] + E[i][1] * om[1] + E[i][2] * om[2];
  *gmag = sqrt(o_dot3(sev, sev) + 0.5 * o_dot3(vs, vs));
  in[0] = sev[0];
  in[1] = sev[1];
  in[2] = sev[2];
  in[3] = vs[0];
  in[4] = vs[1];
  in[5] = nu / (delta * delta);
  for (int k = 0; k < 6; k++) in[k] *= 1 / (*gmag + O_SGS_EPSILON);
}

static inline void o_dd_inference(const OSgsModel *m, const double *in, double *out) {
  double V[64];
  for (int i = 0; i < m->num_neurons; i++) {
    double acc = 0;
    for (int j = 0; j < m->num_inputs; j++) acc += m->W1[i * m->num_inputs + j] * in[j];
    V[i] = m->b1[i] + acc;
    V[i] *= V[i] < 0 ? m->alpha : 1.;
  }
  if (m->Wh) {  // second hidden layer, same form as the first
    double V2[64];
    for (int i = 0; i < m->num_neurons; i++) {
      double acc = 0;
      for (int j = 0; j < m->num_neurons; j++) acc += m->Wh[i * m->num_neurons + j] * V[j];
      V2[i] = m->bh[i] + acc;
      V2[i] *= V2[i] < 0 ? m->alpha : 1.;
    }
    for (int i = 0; i < m->num_neurons; i++) V[i] = V2[i];
  }
  for (int i = 0; i < m->num_outputs; i++) {
    double acc = 0;
    for (int j = 0; j < m->num_neurons; j++) acc += m->W2[i * m->num_neurons + j] * V[j];
    out[i] = m->b2[i] + acc;
  }
}

static inline void o_sgs_outputs(const OSgsModel *m, double out[6], double delta, const double E[3][3], double gmag, double km_sgs[6]) {
  for (int i = 0; i < 6; i++) {  // min-max de-normalization from [0, 1] to bounds[i]
    const double lo = m->bounds[2 * i], hi = m->bounds[2 * i + 1];
    out[i]          = (hi - lo) / (1.0 - 0.0) * (out[i] - 1.0) + hi;
  }
  for (int i = 0; i < 6; i++) out[i] *= delta * delta * (gmag * gmag);
  const double S[3][3] = {{out[0], out[3], out[4]}, {out[3], out[1], out[5]}, {out[4], out[5], out[2]}};
  double       ES[3][3] = {{0}}, T[3][3] = {{0}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) ES[i][j] += E[k][i] * S[k][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) T[i][j] += ES[i][k] * E[k][j];
  o_km_pack(T, km_sgs);
}

// Nodal SGS stress (ComputeSgsDD_Fused): inputs at a node -> Kelvin-Mandel SGS stress
static inline void o_sgs_nodal(const OSgsModel *m, const double gu_aniso[3][3], const double km_A[6], double delta, double nu, double km_sgs[6]) {
  double E[3][3], in[6], out[6], gmag;
  o_sgs_inputs(gu_aniso, km_A, delta, nu, E, in, &gmag);
  o_dd_inference(m, in, out);
  o_sgs_outputs(m, out, delta, E, gmag, km_sgs);
}

// SGS flux (FluxSubgridStress): F[0] = 0, F[1+k][j] = sgs[k][j], F[4][j] = u . sgs[:, j]
static inline void o_flux_sgs(const double Y[5], const double km_sgs[6], double F[5][3]) {
  double sgs[3][3];
  o_km_unpack(km_sgs, sgs);
  for (int j = 0; j < 3; j++) {
    F[0][j] = 0.;
    for (int k = 0; k < 3; k++) F[k + 1][j] = sgs[k][j];
    F[4][j] = Y[1] * sgs[0][j] + Y[2] * sgs[1][j] + Y[3] * sgs[2][j];
  }
}
