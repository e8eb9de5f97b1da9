// Pointwise device code of the data-driven subgrid-stress closure (FP64):
//   * max-pivot Jacobi eigen solver for symmetric 3x3 matrices, eigenvalues decreasing, eigenvectors as rows
//     (qfunctions/utils_eigensolver_jacobi.h:29-335),
//   * element anisotropy tensor from dX/dx (qfunctions/grid_anisotropy_tensor.h:14-31),
//   * model inputs, 6-20-6 (or 6-20-20-6) leaky-ReLU MLP, de-normalization and rotation back to the physical frame
//     (qfunctions/sgs_dd_utils.h:17-127, qfunctions/sgs_dd_model.h:33-63),
//   * the SGS flux of IFunction_NodalSgs (qfunctions/sgs_dd_model.h:199-213).
// Kelvin-Mandel packing as in qfunctions/utils.h:239-258.  Every 3x3 index is a compile-time constant (the Jacobi
// pivot is selected with predicated swaps), so the matrices stay in registers.
#pragma once

#include <hip/hip_runtime.h>

namespace sgs {

constexpr int kInputs = 6, kOutputs = 6, kNeurons = 20;

// Model parameters in device memory: b1[20], b2[6], W1[20][6], W2[6][20], bounds[6][2], alpha, then the number of
// hidden layers (1, the reference's model, or 2) and the second hidden layer bh[20], Wh[20][20] (used when 2): with two
// hidden layers W2 maps the second hidden layer to the outputs.
struct Model {
  double b1[kNeurons], b2[kOutputs], W1[kNeurons][kInputs], W2[kOutputs][kNeurons], bounds[kOutputs][2], alpha;
  double layers, bh[kNeurons], Wh[kNeurons][kNeurons];
};

__device__ __forceinline__ void km_pack(const double A[3][3], double v[6]) {
  const double w = sqrt(2.);
  v[0]           = A[0][0];
  v[1]           = A[1][1];
  v[2]           = A[2][2];
  v[3]           = A[2][1] * w;
  v[4]           = A[2][0] * w;
  v[5]           = A[1][0] * w;
}

__device__ __forceinline__ void km_unpack(const double v[6], double A[3][3]) {
  const double w = 1 / sqrt(2.);
  A[0][0]        = v[0];
  A[1][1]        = v[1];
  A[2][2]        = v[2];
  A[2][1] = A[1][2] = w * v[3];
  A[2][0] = A[0][2] = w * v[4];
  A[1][0] = A[0][1] = w * v[5];
}

// Off-diagonal entries of the upper triangle: (0,1), (0,2), (1,2) as a[0], a[1], a[2]
__device__ __forceinline__ int jac_row0_max(double a01, double a02) { return fabs(a02) > fabs(a01) ? 2 : 1; }

// Symmetric 3x3 eigen decomposition.  A: d[3] diagonal and the three upper off-diagonals o01, o02, o12 (overwritten);
// returns eigenvalues ev (decreasing) and eigenvector rows E.
__device__ __forceinline__ void diagonalize3(double d[3], double o01, double o02, double o12, int max_sweeps, double ev[3], double E[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) E[i][j] = i == j ? 1.0 : 0.0;
  const int max_iter = 3 * max_sweeps;
  for (int it = 1; it <= max_iter; it++) {
    // pivot: the first largest of |o01|, |o02| (row 0), then |o12| (row 1) if strictly larger
    const int m0   = jac_row0_max(o01, o02);
    int       piv  = m0 == 1 ? 0 : 1;  // 0: (0,1) 1: (0,2) 2: (1,2)
    double    amax = fabs(m0 == 1 ? o01 : o02);
    if (fabs(o12) > amax) piv = 2;
    const int    i = piv == 2 ? 1 : 0, j = piv == 0 ? 1 : 2;
    double       aij = piv == 0 ? o01 : (piv == 1 ? o02 : o12);
    const double aii = i == 0 ? d[0] : d[1], ajj = j == 1 ? d[1] : d[2];
    if ((aii + aij == aii) && (ajj + aij == ajj)) aij = 0.0;
    if (aij == 0.0) break;
    // rotation
    double       t = 1.0;
    const double djj = ajj - aii;
    if (djj != 0.0) {
      const double kappa = djj / (2.0 * aij);
      t                  = 1.0 / (sqrt(1 + kappa * kappa) + fabs(kappa));
      if (kappa < 0.0) t = -t;
    }
    const double c = 1.0 / sqrt(1 + t * t), s = c * t;
    // diagonal
    const double nii = aii - t * aij, njj = ajj + t * aij;
    // the remaining index w and its entries (w,i) and (w,j) of the upper triangle
    double wi, wj;
    if (piv == 0) {  // i=0, j=1, w=2: (0,2)=o02 is (w,i), (1,2)=o12 is (w,j)
      wi = o02, wj = o12;
    } else if (piv == 1) {  // i=0, j=2, w=1: (0,1)=o01 is (w,i), (1,2)=o12 is (w,j)
      wi = o01, wj = o12;
    } else {  // i=1, j=2, w=0: (0,1)=o01 is (w,i), (0,2)=o02 is (w,j)
      wi = o01, wj = o02;
    }
    const double nwi = c * wi - s * wj, nwj = s * wi + c * wj;
    if (piv == 0) {
      d[0] = nii, d[1] = njj, o01 = 0.0, o02 = nwi, o12 = nwj;
    } else if (piv == 1) {
      d[0] = nii, d[2] = njj, o02 = 0.0, o01 = nwi, o12 = nwj;
    } else {
      d[1] = nii, d[2] = njj, o12 = 0.0, o01 = nwi, o02 = nwj;
    }
    // E <- R^T E on rows i, j
#pragma unroll
    for (int v = 0; v < 3; v++) {
      const double ei = i == 0 ? E[0][v] : E[1][v], ej = j == 1 ? E[1][v] : E[2][v];
      const double ni = c * ei - s * ej, nj = s * ei + c * ej;
      if (i == 0) E[0][v] = ni;
      else E[1][v] = ni;
      if (j == 1) E[1][v] = nj;
      else E[2][v] = nj;
    }
  }
  ev[0] = d[0], ev[1] = d[1], ev[2] = d[2];
  // selection sort, decreasing (the first maximum wins), rows of E follow
#pragma unroll
  for (int i = 0; i < 2; i++) {
    int im = i;
#pragma unroll
    for (int j = i + 1; j < 3; j++)
      if (ev[j] > (im == 0 ? ev[0] : (im == 1 ? ev[1] : ev[2]))) im = j;
#pragma unroll
    for (int k = i + 1; k < 3; k++) {
      if (k != im) continue;
      const double tv = ev[i];
      ev[i]           = ev[k];
      ev[k]           = tv;
#pragma unroll
      for (int v = 0; v < 3; v++) {
        const double te = E[i][v];
        E[i][v]         = E[k][v];
        E[k][v]         = te;
      }
    }
  }
}

__device__ __forceinline__ void diagonalize3(const double A[3][3], int max_sweeps, double ev[3], double E[3][3]) {
  double d[3] = {A[0][0], A[1][1], A[2][2]};
  diagonalize3(d, A[0][1], A[0][2], A[1][2], max_sweeps, ev, E);
}

// A_ij = E^T diag(1/sqrt(eval)) E / ||1/sqrt(eval)||, delta = 2 ||1/sqrt(eval)|| of g = dXdx^T dXdx
__device__ __forceinline__ void anisotropy_tensor(const double dXdx[3][3], double km_A[6], double &delta) {
  double g[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) g[i][j] = dXdx[0][i] * dXdx[0][j] + dXdx[1][i] * dXdx[1][j] + dXdx[2][i] * dXdx[2][j];
  double km_g[6], gu[3][3], ev[3], E[3][3];
  km_pack(g, km_g);
  km_unpack(km_g, gu);
  diagonalize3(gu, 15, ev, E);
#pragma unroll
  for (int i = 0; i < 3; i++) ev[i] = 1 / sqrt(ev[i]);
  double A[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) A[i][j] = E[0][i] * (ev[0] * E[0][j]) + E[1][i] * (ev[1] * E[1][j]) + E[2][i] * (ev[2] * E[2][j]);
  const double nrm = sqrt(ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2]);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) A[i][j] *= 1 / nrm;
  km_pack(A, km_A);
  delta = 2 * nrm;
}

__device__ __forceinline__ void orient_basis(double E[3][3], const double v[3]) {
  const double al0 = E[0][0] * v[0] + E[0][1] * v[1] + E[0][2] * v[2];
  const double al2 = E[2][0] * v[0] + E[2][1] * v[1] + E[2][2] * v[2];
  if (al0 < 0)
#pragma unroll
    for (int k = 0; k < 3; k++) E[0][k] *= -1;
  if (al2 < 0)
#pragma unroll
    for (int k = 0; k < 3; k++) E[2][k] *= -1;
  const double cr[3] = {E[2][1] * E[0][2] - E[2][2] * E[0][1], E[2][2] * E[0][0] - E[2][0] * E[0][2], E[2][0] * E[0][1] - E[2][1] * E[0][0]};
  if (cr[0] * E[1][0] + cr[1] * E[1][1] + cr[2] * E[1][2] < 0)
#pragma unroll
    for (int k = 0; k < 3; k++) E[1][k] *= -1;
}

// ComputeSgsDDInputs (qfunctions/sgs_dd_utils.h:58-90): eigenframe strain rates, vorticity components and the viscous
// scale, normalized by |grad u| + epsilon; E holds the oriented eigenvectors (rows), gmag |grad u|
__device__ __forceinline__ void sgs_inputs(const double gu[3][3], const double km_A[6], double delta, double nu, double E[3][3], double in[6],
                                           double &gmag) {
  double A[3][3], gi[3][3];
  km_unpack(km_A, A);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) gi[i][j] = gu[i][0] * A[0][j] + gu[i][1] * A[1][j] + gu[i][2] * A[2][j];
  const double w      = 1 / sqrt(2.);
  const double kms[6] = {gi[0][0], gi[1][1], gi[2][2], w * (gi[1][2] + gi[2][1]), w * (gi[0][2] + gi[2][0]), w * (gi[0][1] + gi[1][0])};
  double       S[3][3], sev[3];
  km_unpack(kms, S);
  diagonalize3(S, 5, sev, E);
  const double r12 = 0.5 * (gi[1][2] - gi[2][1]), r02 = 0.5 * (gi[0][2] - gi[2][0]), r01 = 0.5 * (gi[0][1] - gi[1][0]);
  const double om[3] = {-2 * r12, 2 * r02, -2 * r01};
  orient_basis(E, om);
  double vs[3];
#pragma unroll
  for (int i = 0; i < 3; i++) vs[i] = E[i][0] * om[0] + E[i][1] * om[1] + E[i][2] * om[2];
  gmag = sqrt((sev[0] * sev[0] + sev[1] * sev[1] + sev[2] * sev[2]) + 0.5 * (vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2]));
  in[0] = sev[0], in[1] = sev[1], in[2] = sev[2], in[3] = vs[0], in[4] = vs[1], in[5] = nu / (delta * delta);
  const double sc = 1 / (gmag + 0x1p-52);
#pragma unroll
  for (int k = 0; k < 6; k++) in[k] *= sc;
}

// MLP inference (DataDrivenInference, qfunctions/sgs_dd_model.h:37-53), de-normalization and rotation back to the
// physical frame (ComputeSgsDDOutputs, qfunctions/sgs_dd_utils.h:92-127)
__device__ __forceinline__ void sgs_finish(const Model &m, const double in[6], double delta, const double E[3][3], double gmag, double km_sgs[6]) {
  double V[kNeurons];
#pragma unroll
  for (int i = 0; i < kNeurons; i++) {
    double acc = 0;
#pragma unroll
    for (int j = 0; j < kInputs; j++) acc += m.W1[i][j] * in[j];
    V[i] = m.b1[i] + acc;
    V[i] *= V[i] < 0 ? m.alpha : 1.;
  }
  if (m.layers > 1.5) {  // second hidden layer (an extension of the reference's single hidden layer)
    double V2[kNeurons];
#pragma unroll
    for (int i = 0; i < kNeurons; i++) {
      double acc = 0;
#pragma unroll
      for (int j = 0; j < kNeurons; j++) acc += m.Wh[i][j] * V[j];
      V2[i] = m.bh[i] + acc;
      V2[i] *= V2[i] < 0 ? m.alpha : 1.;
    }
#pragma unroll
    for (int i = 0; i < kNeurons; i++) V[i] = V2[i];
  }
  double out[kOutputs];
#pragma unroll
  for (int i = 0; i < kOutputs; i++) {
    double acc = 0;
#pragma unroll
    for (int j = 0; j < kNeurons; j++) acc += m.W2[i][j] * V[j];
    out[i] = m.b2[i] + acc;
  }
  // de-normalize from [0, 1], re-dimensionalize, rotate back: E^T S E
#pragma unroll
  for (int i = 0; i < kOutputs; i++) {
    const double lo = m.bounds[i][0], hi = m.bounds[i][1];
    out[i]          = (hi - lo) / (1.0 - 0.0) * (out[i] - 1.0) + hi;
  }
  const double scale = delta * delta * (gmag * gmag);
#pragma unroll
  for (int i = 0; i < kOutputs; i++) out[i] *= scale;
  const double Ss[3][3] = {{out[0], out[3], out[4]}, {out[3], out[1], out[5]}, {out[4], out[5], out[2]}};
  double       ES[3][3], T[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) ES[i][j] = E[0][i] * Ss[0][j] + E[1][i] * Ss[1][j] + E[2][i] * Ss[2][j];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) T[i][j] = ES[i][0] * E[0][j] + ES[i][1] * E[1][j] + ES[i][2] * E[2][j];
  km_pack(T, km_sgs);
}

// ComputeSgsDD_Fused (qfunctions/sgs_dd_model.h:55-63): nodal inputs -> Kelvin-Mandel SGS stress
__device__ __forceinline__ void sgs_stress(const Model &m, const double gu[3][3], const double km_A[6], double delta, double nu, double km_sgs[6]) {
  double E[3][3], in[6], gmag;
  sgs_inputs(gu, km_A, delta, nu, E, in, gmag);
  sgs_finish(m, in, delta, E, gmag, km_sgs);
}

// FluxSubgridStress: F[0] = 0, F[1+k][j] = sgs[k][j], F[4][j] = u . sgs[:, j]
__device__ __forceinline__ void flux_sgs(const double Y[5], const double km_sgs[6], double F[5][3]) {
  double s[3][3];
  km_unpack(km_sgs, s);
#pragma unroll
  for (int j = 0; j < 3; j++) {
    F[0][j] = 0.;
#pragma unroll
    for (int k = 0; k < 3; k++) F[k + 1][j] = s[k][j];
    F[4][j] = Y[1] * s[0][j] + Y[2] * s[1][j] + Y[3] * s[2][j];
  }
}

}  // namespace sgs
