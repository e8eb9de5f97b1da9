// CPU ORACLE -- test infrastructure only (see ns_pointwise.h).
//
// libCEED /cpu/self-style operator application for the HONEE Newtonian volume operators.  Per element: gather the
// L-vector through the offsets, apply the tensor-product basis (interp with B, gradient with D in one direction and B
// in the others, contracting x first as libCEED's CeedTensorContractApply does), run the QFunction restatement at each
// quadrature point, apply the basis transpose, and scatter-add.  Element outputs are computed in parallel (OpenMP) into
// a private E-vector and assembled serially in element order, so results do not depend on the thread count.
#include "honee_oracle.h"
#include "ns_qfunctions.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

static inline int ipow3(int n) { return n * n * n; }

// Thread count of the element loops (results do not depend on it).
void oracle_set_num_threads(int n) { omp_set_num_threads(n); }
int  oracle_max_threads(void) { return omp_get_max_threads(); }

// ----------------------------------------------------------------------------------------------------------------------
// Tensor contractions.  u is [nz][ny][nx] with x fastest.  A is [nrow][ncol] row-major.
// ----------------------------------------------------------------------------------------------------------------------
// out = A applied along direction dir (A: [nout][nin]); ext = input extents {nx, ny, nz}
static void contract(const double *A, int nout, int nin, int dir, const int ext[3], const double *in, double *out) {
  int stride = 1, pre = 1;
  for (int d = 0; d < dir; d++) stride *= ext[d];
  for (int d = dir + 1; d < 3; d++) pre *= ext[d];
  for (int a = 0; a < pre; a++)
    for (int o = 0; o < nout; o++)
      for (int s = 0; s < stride; s++) {
        double acc = 0;
        for (int i = 0; i < nin; i++) acc += A[o * nin + i] * in[(a * nin + i) * stride + s];
        out[(a * nout + o) * stride + s] = acc;
      }
}

// out = A^T applied along direction dir (A: [nin][nout], i.e. contract over A's row index)
static void contract_t(const double *A, int nin, int nout, int dir, const int ext[3], const double *in, double *out) {
  int stride = 1, pre = 1;
  for (int d = 0; d < dir; d++) stride *= ext[d];
  for (int d = dir + 1; d < 3; d++) pre *= ext[d];
  for (int a = 0; a < pre; a++)
    for (int o = 0; o < nout; o++)
      for (int s = 0; s < stride; s++) {
        double acc = 0;
        for (int i = 0; i < nin; i++) acc += A[i * nout + o] * in[(a * nin + i) * stride + s];
        out[(a * nout + o) * stride + s] = acc;
      }
}

// v[Q^3] = (Az (x) Ay (x) Ax) u[P^3], matrices [Q][P]
static void tensor3(int P, int Q, const double *Ax, const double *Ay, const double *Az, const double *u, double *v) {
  double t1[512], t2[512];
  int    e0[3] = {P, P, P}, e1[3] = {Q, P, P}, e2[3] = {Q, Q, P};
  contract(Ax, Q, P, 0, e0, u, t1);
  contract(Ay, Q, P, 1, e1, t1, t2);
  contract(Az, Q, P, 2, e2, t2, v);
}

// u[P^3] += (Az (x) Ay (x) Ax)^T v[Q^3]
static void tensor3_t_add(int P, int Q, const double *Ax, const double *Ay, const double *Az, const double *v, double *u) {
  double t1[512], t2[512], t3[512];
  int    e0[3] = {Q, Q, Q}, e1[3] = {P, Q, Q}, e2[3] = {P, P, Q};
  contract_t(Ax, Q, P, 0, e0, v, t1);
  contract_t(Ay, Q, P, 1, e1, t1, t2);
  contract_t(Az, Q, P, 2, e2, t2, t3);
  for (int i = 0; i < ipow3(P); i++) u[i] += t3[i];
}

// Interp + grad of ncomp components of an element (u_e: [ncomp][P^3]) -> uq [ncomp][Q^3], duq [3][ncomp][Q^3]
static void interp_grad(int P, int Q, const double *B, const double *D, int ncomp, const double *u_e, double *uq, double *duq) {
  const int Q3 = ipow3(Q), P3 = ipow3(P);
  for (int c = 0; c < ncomp; c++) {
    if (uq) tensor3(P, Q, B, B, B, &u_e[c * P3], &uq[c * Q3]);
    if (duq) {
      tensor3(P, Q, D, B, B, &u_e[c * P3], &duq[(0 * ncomp + c) * Q3]);
      tensor3(P, Q, B, D, B, &u_e[c * P3], &duq[(1 * ncomp + c) * Q3]);
      tensor3(P, Q, B, B, D, &u_e[c * P3], &duq[(2 * ncomp + c) * Q3]);
    }
  }
}

// Transpose: r_e[ncomp][P^3] = B^T v + sum_d G_d^T gv[d]
static void interp_grad_t(int P, int Q, const double *B, const double *D, int ncomp, const double *vq, const double *gvq, double *r_e) {
  const int Q3 = ipow3(Q), P3 = ipow3(P);
  memset(r_e, 0, sizeof(double) * ncomp * P3);
  for (int c = 0; c < ncomp; c++) {
    if (vq) tensor3_t_add(P, Q, B, B, B, &vq[c * Q3], &r_e[c * P3]);
    if (gvq) {
      tensor3_t_add(P, Q, D, B, B, &gvq[(0 * ncomp + c) * Q3], &r_e[c * P3]);
      tensor3_t_add(P, Q, B, D, B, &gvq[(1 * ncomp + c) * Q3], &r_e[c * P3]);
      tensor3_t_add(P, Q, B, B, D, &gvq[(2 * ncomp + c) * Q3], &r_e[c * P3]);
    }
  }
}

// Restriction gather for an L-vector with 5 interleaved components (comp_stride = 1, src/dm-utils.c:108-112)
static void gather5(const OMesh *m, int e, const double *L, double *u_e) {
  const int P3 = ipow3(m->P);
  for (int n = 0; n < P3; n++) {
    const int node = m->offsets[e * P3 + n];
    for (int c = 0; c < 5; c++) u_e[c * P3 + n] = L[node * 5 + c];
  }
}

// Serial scatter-add of E-vector [nelem][5][P^3] in element order
static void scatter5(const OMesh *m, const double *E, double *L) {
  const int P3 = ipow3(m->P);
  memset(L, 0, sizeof(double) * 5 * m->nnodes);
  for (int e = 0; e < m->nelem; e++)
    for (int n = 0; n < P3; n++) {
      const int node = m->offsets[e * P3 + n];
      for (int c = 0; c < 5; c++) L[node * 5 + c] += E[(e * 5 + c) * P3 + n];
    }
}

// Coordinates of the element's quadrature points: x[3][Q^3] (coordinate basis, INTERP), and dxdX[3 dirs][3 comps][Q^3] (GRAD)
static void elem_coords(const OMesh *m, int e, double *xq, double *dxq) {
  const int Px3 = ipow3(m->Px);
  double    xe[3 * 64];
  for (int n = 0; n < Px3; n++) {
    const int node = m->xoffsets[e * Px3 + n];
    for (int c = 0; c < 3; c++) xe[c * Px3 + n] = m->xcoord[node * 3 + c];
  }
  interp_grad(m->Px, m->Q, m->Bx, m->Dx, 3, xe, xq, dxq);
}

// ----------------------------------------------------------------------------------------------------------------------
// Setup (qfunctions/setupgeo.h:42-56, setupgeo_helpers.h:26-32, utils.h:201-216)
// ----------------------------------------------------------------------------------------------------------------------
static void mat_inv3(const double A[3][3], double Ai[3][3], double *detJ) {
  Ai[0][0]       = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  Ai[0][1]       = A[0][2] * A[2][1] - A[0][1] * A[2][2];
  Ai[0][2]       = A[0][1] * A[1][2] - A[0][2] * A[1][1];
  Ai[1][0]       = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  Ai[1][1]       = A[0][0] * A[2][2] - A[0][2] * A[2][0];
  Ai[1][2]       = A[0][2] * A[1][0] - A[0][0] * A[1][2];
  Ai[2][0]       = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  Ai[2][1]       = A[0][1] * A[2][0] - A[0][0] * A[2][1];
  Ai[2][2]       = A[0][0] * A[1][1] - A[0][1] * A[1][0];
  const double d = A[0][0] * Ai[0][0] + A[1][0] * Ai[0][1] + A[2][0] * Ai[0][2];
  const double s = 1 / d;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Ai[i][j] *= s;
  *detJ = d;
}

int oracle_setup_qdata(const OMesh *m, double *qdata) {
  const int Q = m->Q, Q3 = ipow3(Q);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double dxq[9 * 512];
    elem_coords(m, e, NULL, dxq);
    for (int i = 0; i < Q3; i++) {
      const int    qx = i % Q, qy = (i / Q) % Q, qz = i / (Q * Q);
      const double w  = m->w1d[qx] * m->w1d[qy] * m->w1d[qz];
      double       J[3][3], Ji[3][3], detJ;
      for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) J[c][d] = dxq[(d * 3 + c) * Q3 + i];  // dx_c / dX_d (GradUnpack3D)
      mat_inv3(J, Ji, &detJ);
      double *qd        = &qdata[(size_t)e * 10 * Q3];
      qd[0 * Q3 + i]    = w * detJ;
      for (int k = 0; k < 9; k++) qd[(1 + k) * Q3 + i] = (&Ji[0][0])[k];
    }
  }
  return 0;
}

double oracle_volume(const OMesh *m, const double *qdata) {
  const int Q3  = ipow3(m->Q);
  double    vol = 0;
  for (int e = 0; e < m->nelem; e++)
    for (int i = 0; i < Q3; i++) vol += qdata[(size_t)e * 10 * Q3 + i];
  return vol;
}

static inline void unpack_qdata(const double *qd, int Q3, int i, double *wdetJ, double dXdx[3][3]) {
  *wdetJ = qd[i];
  for (int k = 0; k < 9; k++) (&dXdx[0][0])[k] = qd[(1 + k) * Q3 + i];
}


int oracle_ifunction(const OMesh *m, const ONewt *c, const double *qdata, const double *q, const double *qdot, double *v, double *jac) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q), nj = njac(c);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 5 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], uq[5 * 512], duq[15 * 512], udq[5 * 512], vq[5 * 512], gvq[15 * 512], xq[3 * 512];
    gather5(m, e, q, ue);
    interp_grad(m->P, m->Q, m->B, m->D, 5, ue, uq, duq);
    gather5(m, e, qdot, ue);
    interp_grad(m->P, m->Q, m->B, m->D, 5, ue, udq, NULL);
    if (c->idl_enable) elem_coords(m, e, xq, NULL);
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int i = 0; i < Q3; i++) {
      double qi[5], qdi[5], gref[3][5], wdetJ, dXdx[3][3], x[3] = {0}, vi[5], gvi[3][5], jd[15];
      for (int k = 0; k < 5; k++) {
        qi[k]  = uq[k * Q3 + i];
        qdi[k] = udq[k * Q3 + i];
        for (int d = 0; d < 3; d++) gref[d][k] = duq[(d * 5 + k) * Q3 + i];
      }
      if (c->idl_enable)
        for (int d = 0; d < 3; d++) x[d] = xq[d * Q3 + i];
      unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
      qf_ifunction(c, qi, gref, qdi, wdetJ, dXdx, x, vi, gvi, jd);
      for (int k = 0; k < 5; k++) {
        vq[k * Q3 + i] = vi[k];
        for (int d = 0; d < 3; d++) gvq[(d * 5 + k) * Q3 + i] = gvi[d][k];
      }
      for (int k = 0; k < nj; k++) jac[((size_t)e * nj + k) * Q3 + i] = jd[k];
    }
    interp_grad_t(m->P, m->Q, m->B, m->D, 5, vq, gvq, &E[(size_t)e * 5 * P3]);
  }
  scatter5(m, E, v);
  free(E);
  return 0;
}

// Element action of the IJacobian on an element vector ue [5][P^3] -> re [5][P^3]
static void elem_ijacobian(const OMesh *m, const ONewt *c, const double *qd, const double *jde, const double *ue, double *re) {
  const int Q3 = ipow3(m->Q), nj = njac(c);
  double    uq[5 * 512], duq[15 * 512], vq[5 * 512], gvq[15 * 512];
  interp_grad(m->P, m->Q, m->B, m->D, 5, ue, uq, duq);
  for (int i = 0; i < Q3; i++) {
    double dqi[5], gref[3][5], wdetJ, dXdx[3][3], vi[5], gvi[3][5], jd[15];
    for (int k = 0; k < 5; k++) {
      dqi[k] = uq[k * Q3 + i];
      for (int d = 0; d < 3; d++) gref[d][k] = duq[(d * 5 + k) * Q3 + i];
    }
    for (int k = 0; k < nj; k++) jd[k] = jde[k * Q3 + i];
    unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
    qf_ijacobian(c, dqi, gref, wdetJ, dXdx, jd, vi, gvi);
    for (int k = 0; k < 5; k++) {
      vq[k * Q3 + i] = vi[k];
      for (int d = 0; d < 3; d++) gvq[(d * 5 + k) * Q3 + i] = gvi[d][k];
    }
  }
  interp_grad_t(m->P, m->Q, m->B, m->D, 5, vq, gvq, re);
}

int oracle_ijacobian(const OMesh *m, const ONewt *c, const double *qdata, const double *jac, const double *dq, double *v) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q), nj = njac(c);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 5 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512];
    gather5(m, e, dq, ue);
    elem_ijacobian(m, c, &qdata[(size_t)e * 10 * Q3], &jac[(size_t)e * nj * Q3], ue, &E[(size_t)e * 5 * P3]);
  }
  scatter5(m, E, v);
  free(E);
  return 0;
}

// Diagonal of the assembled operator (CeedOperatorLinearAssembleDiagonal semantics): sum over elements of the element
// matrix diagonal, obtained here by applying the element operator to each local unit vector.
int oracle_ijacobian_diagonal(const OMesh *m, const ONewt *c, const double *qdata, const double *jac, double *diag) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q), nj = njac(c);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 5 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], re[5 * 512];
    for (int k = 0; k < 5 * P3; k++) {
      memset(ue, 0, sizeof(double) * 5 * P3);
      ue[k] = 1.0;
      elem_ijacobian(m, c, &qdata[(size_t)e * 10 * Q3], &jac[(size_t)e * nj * Q3], ue, re);
      E[(size_t)e * 5 * P3 + k] = re[k];
    }
  }
  scatter5(m, E, diag);
  free(E);
  return 0;
}

// Point-block diagonal (CeedOperatorLinearAssemblePointBlockDiagonal semantics): pb[node][c_out][c_in] is the sum over
// elements of the element matrix entries coupling (node, c_in) to (node, c_out), by local unit vectors.
int oracle_ijacobian_point_block_diagonal(const OMesh *m, const ONewt *c, const double *qdata, const double *jac, double *pb) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q), nj = njac(c);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 25 * P3);  // [e][node][c_out][c_in]
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], re[5 * 512];
    for (int n = 0; n < P3; n++)
      for (int ci = 0; ci < 5; ci++) {
        memset(ue, 0, sizeof(double) * 5 * P3);
        ue[ci * P3 + n] = 1.0;
        elem_ijacobian(m, c, &qdata[(size_t)e * 10 * Q3], &jac[(size_t)e * nj * Q3], ue, re);
        for (int co = 0; co < 5; co++) E[((size_t)e * P3 + n) * 25 + co * 5 + ci] = re[co * P3 + n];
      }
  }
  memset(pb, 0, sizeof(double) * 25 * m->nnodes);
  for (int e = 0; e < m->nelem; e++)
    for (int n = 0; n < P3; n++) {
      const int node = m->offsets[e * P3 + n];
      for (int k = 0; k < 25; k++) pb[node * 25 + k] += E[((size_t)e * P3 + n) * 25 + k];
    }
  free(E);
  return 0;
}

int oracle_rhs(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *v) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 5 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], uq[5 * 512], duq[15 * 512], vq[5 * 512], gvq[15 * 512], xq[3 * 512];
    gather5(m, e, q, ue);
    interp_grad(m->P, m->Q, m->B, m->D, 5, ue, uq, duq);
    if (c->idl_enable) elem_coords(m, e, xq, NULL);
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int i = 0; i < Q3; i++) {
      double qi[5], gref[3][5], wdetJ, dXdx[3][3], x[3] = {0}, vi[5], gvi[3][5];
      for (int k = 0; k < 5; k++) {
        qi[k] = uq[k * Q3 + i];
        for (int d = 0; d < 3; d++) gref[d][k] = duq[(d * 5 + k) * Q3 + i];
      }
      if (c->idl_enable)
        for (int d = 0; d < 3; d++) x[d] = xq[d * Q3 + i];
      unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
      qf_rhs(c, qi, gref, wdetJ, dXdx, x, vi, gvi);
      for (int k = 0; k < 5; k++) {
        vq[k * Q3 + i] = vi[k];
        for (int d = 0; d < 3; d++) gvq[(d * 5 + k) * Q3 + i] = gvi[d][k];
      }
    }
    interp_grad_t(m->P, m->Q, m->B, m->D, 5, vq, gvq, &E[(size_t)e * 5 * P3]);
  }
  scatter5(m, E, v);
  free(E);
  return 0;
}

// ----------------------------------------------------------------------------------------------------------------------
// MassFunction_Newtonian (newtonian.h:44-82): stabilized consistent mass applied to q_dot, state q (conservative)
// ----------------------------------------------------------------------------------------------------------------------
static void qf_mass(const ONewt *c, const double qdot[5], const double q[5], const double wdetJ, const double dXdx[3][3], double v[5],
                    double gv[3][5]) {
  const OGas   g  = c->gas;
  const OState s  = o_state_from_q(g, q, OVAR_CONS);
  const OState sd = o_state_from_q(g, qdot, OVAR_CONS);
  for (int f = 0; f < 5; f++) v[f] = wdetJ * qdot[f];
  OState gs[3];
  memset(gs, 0, sizeof(gs));
  double Tau[3], stab[5][3], bf[5] = {0}, divF[5] = {0}, Udot[5];
  memcpy(Udot, sd.U, sizeof(Udot));
  o_tau_diag_prim(c->tau, g, &s, dXdx, c->dt, Tau);
  o_stabilization(c->stabilization, g, &s, Tau, gs, Udot, bf, divF, stab);
  for (int j = 0; j < 5; j++)
    for (int k = 0; k < 3; k++) gv[k][j] = wdetJ * (stab[j][0] * dXdx[k][0] + stab[j][1] * dXdx[k][1] + stab[j][2] * dXdx[k][2]);
}

static void elem_mass(const OMesh *m, const ONewt *c, const double *qd, const double *qq /* [5][Q3] state at qpts */, const double *ue,
                      double *re) {
  const int Q3 = ipow3(m->Q);
  double    uq[5 * 512], vq[5 * 512], gvq[15 * 512];
  interp_grad(m->P, m->Q, m->B, m->D, 5, ue, uq, NULL);
  for (int i = 0; i < Q3; i++) {
    double qdi[5], qi[5], wdetJ, dXdx[3][3], vi[5], gvi[3][5];
    for (int k = 0; k < 5; k++) {
      qdi[k] = uq[k * Q3 + i];
      qi[k]  = qq[k * Q3 + i];
    }
    unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
    qf_mass(c, qdi, qi, wdetJ, dXdx, vi, gvi);
    for (int k = 0; k < 5; k++) {
      vq[k * Q3 + i] = vi[k];
      for (int d = 0; d < 3; d++) gvq[(d * 5 + k) * Q3 + i] = gvi[d][k];
    }
  }
  interp_grad_t(m->P, m->Q, m->B, m->D, 5, vq, gvq, re);
}

int oracle_mass_newtonian(const OMesh *m, const ONewt *c, const double *qdata, const double *q, const double *qdot, double *v) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 5 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], qq[5 * 512];
    gather5(m, e, q, ue);
    interp_grad(m->P, m->Q, m->B, m->D, 5, ue, qq, NULL);
    gather5(m, e, qdot, ue);
    elem_mass(m, c, &qdata[(size_t)e * 10 * Q3], qq, ue, &E[(size_t)e * 5 * P3]);
  }
  scatter5(m, E, v);
  free(E);
  return 0;
}

int oracle_mass_newtonian_diagonal(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *diag) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 5 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], qq[5 * 512], re[5 * 512];
    gather5(m, e, q, ue);
    interp_grad(m->P, m->Q, m->B, m->D, 5, ue, qq, NULL);
    for (int k = 0; k < 5 * P3; k++) {
      memset(ue, 0, sizeof(double) * 5 * P3);
      ue[k] = 1.0;
      elem_mass(m, c, &qdata[(size_t)e * 10 * Q3], qq, ue, re);
      E[(size_t)e * 5 * P3 + k] = re[k];
    }
  }
  scatter5(m, E, diag);
  free(E);
  return 0;
}

// ----------------------------------------------------------------------------------------------------------------------
// MonitorTotalKineticEnergy (qfunctions/monitor_totalkineticenergy.h:8-39); sums over all quadrature points
// ----------------------------------------------------------------------------------------------------------------------
int oracle_monitor_ke(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double sums[5]) {
  const int Q3 = ipow3(m->Q);
  const OGas g  = c->gas;
  double   *vals = malloc(sizeof(double) * (size_t)m->nelem * 5 * Q3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], uq[5 * 512], duq[15 * 512];
    gather5(m, e, q, ue);
    interp_grad(m->P, m->Q, m->B, m->D, 5, ue, uq, duq);
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int i = 0; i < Q3; i++) {
      double qi[5], gref[3][5], wdetJ, dXdx[3][3];
      for (int k = 0; k < 5; k++) {
        qi[k] = uq[k * Q3 + i];
        for (int d = 0; d < 3; d++) gref[d][k] = duq[(d * 5 + k) * Q3 + i];
      }
      unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
      const OState s = o_state_from_q(g, qi, c->state_var);
      OState       gs[3];
      o_phys_gradient(g, &s, c->state_var, gref, dXdx, gs);
      double *out = &vals[(size_t)e * 5 * Q3];
      out[0 * Q3 + i] = wdetJ * 0.5 * s.U[0] * o_dot3(&s.Y[1], &s.Y[1]);
      double kmsr[6], sr[3][3];
      o_km_strain_rate(gs, kmsr);
      double div_u = kmsr[0] + kmsr[1] + kmsr[2];
      o_km_unpack(kmsr, sr);
      out[1 * Q3 + i] = wdetJ * -2 * g.mu * o_dotn(&sr[0][0], &sr[0][0], 9);
      out[2 * Q3 + i] = wdetJ * -g.lambda * g.mu * o_sq(div_u);
      out[3 * Q3 + i] = wdetJ * s.Y[0] * div_u;
      double gu[3][3], w[3];
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) gu[a][b] = gs[b].Y[1 + a];
      w[0]            = gu[2][1] - gu[1][2];
      w[1]            = gu[0][2] - gu[2][0];
      w[2]            = gu[1][0] - gu[0][1];
      out[4 * Q3 + i] = wdetJ * g.mu * o_dot3(w, w);
    }
  }
  for (int k = 0; k < 5; k++) sums[k] = 0;
  for (int e = 0; e < m->nelem; e++)
    for (int k = 0; k < 5; k++)
      for (int i = 0; i < Q3; i++) sums[k] += vals[((size_t)e * 5 + k) * Q3 + i];
  free(vals);
  return 0;
}

// ----------------------------------------------------------------------------------------------------------------------
// ICsTaylorGreen (qfunctions/taylorgreen.h:23-51) evaluated at n points xyz[n][3] -> q[n][5]
// ----------------------------------------------------------------------------------------------------------------------
void oracle_ic_taylorgreen(const ONewt *c, const double ref_Y[5], const double bg_u[3], const double L[3], int n, const double *xyz,
                           double *q) {
  const OGas   g    = c->gas;
  const double R    = o_rgas(g);
  const double V0   = sqrt(ref_Y[1] * ref_Y[1] + ref_Y[2] * ref_Y[2] + ref_Y[3] * ref_Y[3]);
  const double rho0 = ref_Y[0] / (ref_Y[4] * R);
  for (int i = 0; i < n; i++) {
    double x[3];
    const double scale = 2 * M_PI / L[0];  // all three directions use lx (taylorgreen.h:40)
    for (int d = 0; d < 3; d++) x[d] = xyz[i * 3 + d] * scale;
    double Y[5];
    Y[0]     = ref_Y[0] + (rho0 * o_sq(V0) / 16) * (cos(2 * x[0]) + cos(2 * x[1])) * (cos(2 * x[2] + 2));
    Y[1]     = bg_u[0] + V0 * sin(x[0]) * cos(x[1]) * cos(x[2]);
    Y[2]     = bg_u[1] - V0 * cos(x[0]) * sin(x[1]) * cos(x[2]);
    Y[3]     = bg_u[2];
    Y[4]     = ref_Y[4];
    OState s = o_state_from_q(g, Y, OVAR_PRIM);
    o_state_to_q(g, &s, &q[i * 5], c->state_var);
  }
}

// ----------------------------------------------------------------------------------------------------------------------
// Pointwise entry points for unit tests
// ----------------------------------------------------------------------------------------------------------------------
void oracle_state_from_q(const OGas *g, const double *q, int var, double *U, double *Y) {
  OState s = o_state_from_q(*g, q, var);
  memcpy(U, s.U, sizeof(s.U));
  memcpy(Y, s.Y, sizeof(s.Y));
}
void oracle_state_from_q_fwd(const OGas *g, const double *q, const double *dq, int var, double *dU, double *dY) {
  OState s  = o_state_from_q(*g, q, var);
  OState ds = o_state_from_q_fwd(*g, &s, dq, var);
  memcpy(dU, ds.U, sizeof(ds.U));
  memcpy(dY, ds.Y, sizeof(ds.Y));
}
void oracle_state_to_q(const OGas *g, const double *U, const double *Y, int var, double *q) {
  OState s;
  memcpy(s.U, U, sizeof(s.U));
  memcpy(s.Y, Y, sizeof(s.Y));
  o_state_to_q(*g, &s, q, var);
}
void oracle_state_to_q_fwd(const OGas *g, const double *U, const double *Y, const double *dU, const double *dY, int var, double *dq) {
  OState s, ds;
  memcpy(s.U, U, sizeof(s.U));
  memcpy(s.Y, Y, sizeof(s.Y));
  memcpy(ds.U, dU, sizeof(ds.U));
  memcpy(ds.Y, dY, sizeof(ds.Y));
  o_state_to_q_fwd(*g, &s, &ds, dq, var);
}
void oracle_tau_diag(const OGas *g, const OTauCoef *tc, const double *q, int var, const double *dXdx, double dt, double *Tau) {
  OState s = o_state_from_q(*g, q, var);
  double J[3][3];
  memcpy(J, dXdx, sizeof(J));
  o_tau_diag_prim(*tc, *g, &s, J, dt, Tau);
}

// Pointwise IFunction / IJacobian / RHS QFunction bodies (one quadrature point), for the oracle's own consistency tests.
void oracle_qf_ifunction(const ONewt *c, const double *q, const double *gref, const double *qdot, double wdetJ, const double *dXdx,
                         const double *x, double *v, double *gv, double *jd) {
  double G[3][5], J[3][3], GV[3][5];
  memcpy(G, gref, sizeof(G));
  memcpy(J, dXdx, sizeof(J));
  qf_ifunction(c, q, G, qdot, wdetJ, J, x, v, GV, jd);
  memcpy(gv, GV, sizeof(GV));
}

void oracle_qf_ijacobian(const ONewt *c, const double *dq, const double *gref, double wdetJ, const double *dXdx, const double *jd, double *v,
                         double *gv) {
  double G[3][5], J[3][3], GV[3][5];
  memcpy(G, gref, sizeof(G));
  memcpy(J, dXdx, sizeof(J));
  qf_ijacobian(c, dq, G, wdetJ, J, jd, v, GV);
  memcpy(gv, GV, sizeof(GV));
}

void oracle_qf_rhs(const ONewt *c, const double *U, const double *gref, double wdetJ, const double *dXdx, const double *x, double *v,
                   double *gv) {
  double G[3][5], J[3][3], GV[3][5];
  memcpy(G, gref, sizeof(G));
  memcpy(J, dXdx, sizeof(J));
  qf_rhs(c, U, G, wdetJ, J, x, v, GV);
  memcpy(gv, GV, sizeof(GV));
}

// ----------------------------------------------------------------------------------------------------------------------
// Data-driven SGS closure and its nodal inputs (problems/sgs_dd_model.c, src/velocity_gradient_projection.c,
// src/grid_anisotropy_tensor.c).  Nodal fields are interleaved L-vectors [nnodes][ncomp] on the solution nodes.
// ----------------------------------------------------------------------------------------------------------------------
#include "sgs_pointwise.h"

static void gatherN(const OMesh *m, int e, const double *L, int ncomp, double *u_e) {
  const int P3 = ipow3(m->P);
  for (int n = 0; n < P3; n++) {
    const int node = m->offsets[e * P3 + n];
    for (int c = 0; c < ncomp; c++) u_e[c * P3 + n] = L[(size_t)node * ncomp + c];
  }
}

// Serial scatter-add of an E-vector [nelem][ncomp][P^3] in element order
static void scatterN_add(const OMesh *m, const double *E, int ncomp, double *L) {
  const int P3 = ipow3(m->P);
  for (int e = 0; e < m->nelem; e++)
    for (int n = 0; n < P3; n++) {
      const int node = m->offsets[e * P3 + n];
      for (int c = 0; c < ncomp; c++) L[(size_t)node * ncomp + c] += E[((size_t)e * ncomp + c) * P3 + n];
    }
}

// VelocityGradientProjectionRHS (qfunctions/velocity_gradient_projection.h:9-36): r = sum_e B^T (wdetJ du_j/dx_k)
int oracle_velocity_gradient_rhs(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *r) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 9 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], uq[5 * 512], duq[15 * 512], vq[9 * 512];
    gather5(m, e, q, ue);
    interp_grad(m->P, m->Q, m->B, m->D, 5, ue, uq, duq);
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int i = 0; i < Q3; i++) {
      double qi[5], gref[3][5], wdetJ, dXdx[3][3];
      for (int k = 0; k < 5; k++) {
        qi[k] = uq[k * Q3 + i];
        for (int d = 0; d < 3; d++) gref[d][k] = duq[(d * 5 + k) * Q3 + i];
      }
      unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
      const OState s = o_state_from_q(c->gas, qi, c->state_var);
      OState       gs[3];
      o_phys_gradient(c->gas, &s, c->state_var, gref, dXdx, gs);
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) vq[(j * 3 + k) * Q3 + i] = wdetJ * gs[k].Y[1 + j];  // grad_velocity[j][k] = d u_j / d x_k
    }
    interp_grad_t(m->P, m->Q, m->B, m->D, 9, vq, NULL, &E[(size_t)e * 9 * P3]);
  }
  memset(r, 0, sizeof(double) * 9 * (size_t)m->nnodes);
  scatterN_add(m, E, 9, r);
  free(E);
  return 0;
}

// AnisotropyTensorProjection (qfunctions/grid_anisotropy_tensor.h:34-54): r = sum_e B^T (wdetJ [KM(A_ij), delta])
int oracle_grid_anisotropy_rhs(const OMesh *m, const double *qdata, double *r) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 7 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double        vq[7 * 512];
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int i = 0; i < Q3; i++) {
      double wdetJ, dXdx[3][3], kmA[6], delta;
      unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
      o_anisotropy_tensor(dXdx, kmA, &delta);
      for (int k = 0; k < 6; k++) vq[k * Q3 + i] = wdetJ * kmA[k];
      vq[6 * Q3 + i] = wdetJ * delta;
    }
    interp_grad_t(m->P, m->Q, m->B, m->D, 7, vq, NULL, &E[(size_t)e * 7 * P3]);
  }
  memset(r, 0, sizeof(double) * 7 * (size_t)m->nnodes);
  scatterN_add(m, E, 7, r);
  free(E);
  return 0;
}

// Mass_N (qfunctions/mass.h): v = sum_e B^T (wdetJ B u), ncomp interleaved components
int oracle_mass_n(const OMesh *m, const double *qdata, int ncomp, const double *u, double *v) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * ncomp * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[9 * 512], uq[9 * 512];
    gatherN(m, e, u, ncomp, ue);
    interp_grad(m->P, m->Q, m->B, m->D, ncomp, ue, uq, NULL);
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int k = 0; k < ncomp; k++)
      for (int i = 0; i < Q3; i++) uq[k * Q3 + i] *= qd[i];
    interp_grad_t(m->P, m->Q, m->B, m->D, ncomp, uq, NULL, &E[(size_t)e * ncomp * P3]);
  }
  memset(v, 0, sizeof(double) * ncomp * (size_t)m->nnodes);
  scatterN_add(m, E, ncomp, v);
  free(E);
  return 0;
}

int oracle_mass_n_diagonal(const OMesh *m, const double *qdata, int ncomp, double *d) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * ncomp * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double        ue[512], uq[512], re[512];
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int n = 0; n < P3; n++) {
      memset(ue, 0, sizeof(double) * P3);
      ue[n] = 1.0;
      interp_grad(m->P, m->Q, m->B, m->D, 1, ue, uq, NULL);
      for (int i = 0; i < Q3; i++) uq[i] *= qd[i];
      interp_grad_t(m->P, m->Q, m->B, m->D, 1, uq, NULL, re);
      for (int k = 0; k < ncomp; k++) E[((size_t)e * ncomp + k) * P3 + n] = re[n];
    }
  }
  memset(d, 0, sizeof(double) * ncomp * (size_t)m->nnodes);
  scatterN_add(m, E, ncomp, d);
  free(E);
  return 0;
}

// Velocity gradient projection as HONEE configures it (src/velocity_gradient_projection.c:122-155): KSPPREONLY with
// PCJACOBI row sum, i.e. G = R / (M 1) per entry.
int oracle_velocity_gradient_projection(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *G) {
  const size_t n    = 9 * (size_t)m->nnodes;
  double      *ones = calloc(n, sizeof(double)), *ML = calloc(n, sizeof(double));
  for (size_t i = 0; i < n; i++) ones[i] = 1.0;
  oracle_mass_n(m, qdata, 9, ones, ML);
  oracle_velocity_gradient_rhs(m, c, qdata, q, G);
  for (size_t i = 0; i < n; i++) G[i] *= ML[i] != 0 ? 1.0 / ML[i] : 1.0;
  free(ones);
  free(ML);
  return 0;
}

// Grid anisotropy projection (src/grid_anisotropy_tensor.c:84-110): KSPCG with PCJACOBI (diagonal), natural norm,
// rtol 1e-10, zero initial guess, on the 7-component system M X = R.  Returns the CG iteration count.
int oracle_grid_anisotropy_projection(const OMesh *m, const double *qdata, double *X) {
  const size_t n = 7 * (size_t)m->nnodes;
  double *R = malloc(sizeof(double) * n), *dinv = malloc(sizeof(double) * n), *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n),
         *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  oracle_grid_anisotropy_rhs(m, qdata, R);
  oracle_mass_n_diagonal(m, qdata, 7, dinv);
  for (size_t i = 0; i < n; i++) dinv[i] = dinv[i] != 0 ? 1.0 / dinv[i] : 1.0;
  memset(X, 0, sizeof(double) * n);
  memcpy(r, R, sizeof(double) * n);
  for (size_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
  memcpy(p, z, sizeof(double) * n);
  double rz = 0;
  for (size_t i = 0; i < n; i++) rz += r[i] * z[i];
  const double ttol = fmax(1e-10 * sqrt(fabs(rz)), 1e-50);
  int          its  = 0;
  while (sqrt(fabs(rz)) > ttol && its < 10000) {
    oracle_mass_n(m, qdata, 7, p, Ap);
    double pAp = 0;
    for (size_t i = 0; i < n; i++) pAp += p[i] * Ap[i];
    const double alpha = rz / pAp;
    for (size_t i = 0; i < n; i++) {
      X[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
      z[i] = dinv[i] * r[i];
    }
    double rz_new = 0;
    for (size_t i = 0; i < n; i++) rz_new += r[i] * z[i];
    const double beta = rz_new / rz;
    for (size_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    rz = rz_new;
    its++;
  }
  free(R);
  free(dinv);
  free(r);
  free(z);
  free(p);
  free(Ap);
  return its;
}

// ComputeSgsDDNodal_Fused (qfunctions/sgs_dd_model.h:65-96) at every node: q [nnodes][5], G [nnodes][9] (du_j/dx_k at
// j*3+k), A [nnodes][7] (KM anisotropy tensor, delta) -> km_sgs [nnodes][6].  The reference evaluates it at each
// element node and sums with the inverse multiplicity; every element sees the same nodal inputs, so that sum is the
// nodal value itself.
int oracle_sgs_nodal(const OGas *g, int state_var, const OSgsModel *model, int nnodes, const double *q, const double *G, const double *A,
                     double *km_sgs) {
#pragma omp parallel for schedule(static)
  for (int nd = 0; nd < nnodes; nd++) {
    const OState s = o_state_from_q(*g, &q[(size_t)nd * 5], state_var);
    double       gu[3][3];
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) gu[j][k] = G[(size_t)nd * 9 + j * 3 + k];
    o_sgs_nodal(model, gu, &A[(size_t)nd * 7], A[(size_t)nd * 7 + 6], g->mu / s.U[0], &km_sgs[(size_t)nd * 6]);
  }
  return 0;
}

// IFunction_NodalSgs (qfunctions/sgs_dd_model.h:215-241): v += sum_e G^T (-wdetJ dXdx F_sgs(Y, B km_sgs))
int oracle_ifunction_nodal_sgs(const OMesh *m, const ONewt *c, const double *qdata, const double *q, const double *km_sgs, double *v) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 5 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], uq[5 * 512], ke[6 * 512], kq[6 * 512], gvq[15 * 512];
    gather5(m, e, q, ue);
    gatherN(m, e, km_sgs, 6, ke);
    interp_grad(m->P, m->Q, m->B, m->D, 5, ue, uq, NULL);
    interp_grad(m->P, m->Q, m->B, m->D, 6, ke, kq, NULL);
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int i = 0; i < Q3; i++) {
      double qi[5], ki[6], wdetJ, dXdx[3][3], F[5][3];
      for (int k = 0; k < 5; k++) qi[k] = uq[k * Q3 + i];
      for (int k = 0; k < 6; k++) ki[k] = kq[k * Q3 + i];
      unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
      const OState s = o_state_from_q(c->gas, qi, c->state_var);
      o_flux_sgs(s.Y, ki, F);
      for (int d = 0; d < 3; d++)
        for (int j = 0; j < 5; j++) gvq[(d * 5 + j) * Q3 + i] = -wdetJ * (dXdx[d][0] * F[j][0] + dXdx[d][1] * F[j][1] + dXdx[d][2] * F[j][2]);
    }
    interp_grad_t(m->P, m->Q, m->B, m->D, 5, NULL, gvq, &E[(size_t)e * 5 * P3]);
  }
  scatterN_add(m, E, 5, v);
  free(E);
  return 0;
}

// ----------------------------------------------------------------------------------------------------------------------
// Differential filter (qfunctions/differential_filter.h, src/differential_filter.c) and SGS training pairs
// (qfunctions/sgs_dd_training.h, src/smartsim/sgs_dd_training.c).  The filtered field has 11 interleaved components
// per node: the primitive state (5, DIFF_FILTER_STATE_NUM) then the velocity products (6, order XX YY ZZ YZ XZ XY).
// ----------------------------------------------------------------------------------------------------------------------
// AnisotropyTensorCollocate (qfunctions/grid_anisotropy_tensor.h:57-77): A [nelem][7][Q^3] = [KM(A_ij), delta]
int oracle_anisotropy_collocate(const OMesh *m, const double *qdata, double *A) {
  const int Q3 = ipow3(m->Q);
  for (int e = 0; e < m->nelem; e++) {
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int i = 0; i < Q3; i++) {
      double wdetJ, dXdx[3][3], kmA[6], delta;
      unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
      o_anisotropy_tensor(dXdx, kmA, &delta);
      for (int k = 0; k < 6; k++) A[((size_t)e * 7 + k) * Q3 + i] = kmA[k];
      A[((size_t)e * 7 + 6) * Q3 + i] = delta;
    }
  }
  return 0;
}

static inline double fmax_ref(double a, double b) { return a < b ? b : a; }  // Max (qfunctions/utils.h:14)

// alpha_ij of DifferentialFilter_LHS_N (qfunctions/differential_filter.h:91-125) at one point
static void o_filter_alpha(const ODiffFilter *f, const double *A7, const double x[3], double alpha[3][3]) {
  double D[3][3] = {{0}};
  if (f->grid_based_width) {
    double km[6];
    for (int k = 0; k < 6; k++) km[k] = A7[k] * A7[6];  // dimensionalize the normalized anisotropy tensor
    o_km_unpack(km, D);
  } else {
    D[0][0] = D[1][1] = D[2][2] = 1;
  }
  double S[3][3] = {{0}};
  if (f->damping == 1) {  // van Driest: -expm1(-y+ / A+)
    const double dc = -expm1(-(x[1] / f->friction_length) / f->damping_constant);
    S[0][0]         = fmax_ref(1, dc * f->width_scaling[0]);
    S[1][1]         = dc * f->width_scaling[1];
    S[2][2]         = fmax_ref(1, dc * f->width_scaling[2]);
  } else if (f->damping == 0) {
    S[0][0] = f->width_scaling[0];
    S[1][1] = f->width_scaling[1];
    S[2][2] = f->width_scaling[2];
  } else {
    S[0][0] = 1;
    S[1][1] = tanh(60 * x[1]);
    S[2][2] = 1;
  }
  double SD[3][3] = {{0}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) SD[i][j] += S[i][k] * D[k][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      alpha[i][j] = 0;
      for (int k = 0; k < 3; k++) alpha[i][j] += SD[i][k] * SD[k][j];
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) alpha[i][j] *= f->kernel_scaling;
}

// DifferentialFilter_RHS (qfunctions/differential_filter.h:28-52): r [nnodes][11] = sum_e B^T wdetJ [Y, u_i u_j]
int oracle_diff_filter_rhs(const OMesh *m, const ONewt *c, const double *qdata, const double *q, double *r) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * 11 * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[5 * 512], uq[5 * 512], vq[11 * 512];
    gather5(m, e, q, ue);
    interp_grad(m->P, m->Q, m->B, m->D, 5, ue, uq, NULL);
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int i = 0; i < Q3; i++) {
      double qi[5];
      for (int k = 0; k < 5; k++) qi[k] = uq[k * Q3 + i];
      const double wdetJ = qd[i];
      const OState s     = o_state_from_q(c->gas, qi, c->state_var);
      const double *u    = &s.Y[1];
      for (int k = 0; k < 5; k++) vq[k * Q3 + i] = wdetJ * s.Y[k];
      vq[5 * Q3 + i]  = wdetJ * u[0] * u[0];
      vq[6 * Q3 + i]  = wdetJ * u[1] * u[1];
      vq[7 * Q3 + i]  = wdetJ * u[2] * u[2];
      vq[8 * Q3 + i]  = wdetJ * u[1] * u[2];
      vq[9 * Q3 + i]  = wdetJ * u[0] * u[2];
      vq[10 * Q3 + i] = wdetJ * u[0] * u[1];
    }
    interp_grad_t(m->P, m->Q, m->B, m->D, 11, vq, NULL, &E[(size_t)e * 11 * P3]);
  }
  memset(r, 0, sizeof(double) * 11 * (size_t)m->nnodes);
  scatterN_add(m, E, 11, r);
  free(E);
  return 0;
}

// DifferentialFilter_LHS_N (qfunctions/differential_filter.h:63-138) on an ncomp-component nodal field, the same
// operator for every component; Acol is AnisotropyTensorCollocate's [nelem][7][Q^3] (read when grid_based_width)
int oracle_diff_filter_lhs(const OMesh *m, const ODiffFilter *f, const double *qdata, const double *Acol, int ncomp, const double *u, double *v) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * ncomp * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double *ue = malloc(sizeof(double) * ncomp * P3), *uq = malloc(sizeof(double) * ncomp * Q3), *duq = malloc(sizeof(double) * 3 * ncomp * Q3);
    double *vq = malloc(sizeof(double) * ncomp * Q3), *gvq = malloc(sizeof(double) * 3 * ncomp * Q3), xq[3 * 512];
    gatherN(m, e, u, ncomp, ue);
    interp_grad(m->P, m->Q, m->B, m->D, ncomp, ue, uq, duq);
    elem_coords(m, e, xq, NULL);
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int i = 0; i < Q3; i++) {
      double wdetJ, dXdx[3][3], A7[7], x[3], alpha[3][3];
      unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
      for (int k = 0; k < 7; k++) A7[k] = Acol[((size_t)e * 7 + k) * Q3 + i];
      for (int d = 0; d < 3; d++) x[d] = xq[d * Q3 + i];
      o_filter_alpha(f, A7, x, alpha);
      for (int j = 0; j < ncomp; j++) {
        vq[j * Q3 + i] = wdetJ * uq[j * Q3 + i];
        double dq[3], dqX[3], dqa[3];
        for (int k = 0; k < 3; k++)
          dq[k] = duq[(0 * ncomp + j) * Q3 + i] * dXdx[0][k] + duq[(1 * ncomp + j) * Q3 + i] * dXdx[1][k] + duq[(2 * ncomp + j) * Q3 + i] * dXdx[2][k];
        for (int k = 0; k < 3; k++) dqX[k] = dXdx[k][0] * dq[0] + dXdx[k][1] * dq[1] + dXdx[k][2] * dq[2];
        for (int k = 0; k < 3; k++) dqa[k] = alpha[k][0] * dqX[0] + alpha[k][1] * dqX[1] + alpha[k][2] * dqX[2];
        for (int k = 0; k < 3; k++) gvq[(k * ncomp + j) * Q3 + i] = wdetJ * dqa[k];
      }
    }
    interp_grad_t(m->P, m->Q, m->B, m->D, ncomp, vq, gvq, &E[(size_t)e * ncomp * P3]);
    free(ue), free(uq), free(duq), free(vq), free(gvq);
  }
  memset(v, 0, sizeof(double) * ncomp * (size_t)m->nnodes);
  scatterN_add(m, E, ncomp, v);
  free(E);
  return 0;
}

// Diagonal of the filter operator (every component the same): element operator applied to unit node vectors
int oracle_diff_filter_lhs_diagonal(const OMesh *m, const ODiffFilter *f, const double *qdata, const double *Acol, int ncomp, double *d) {
  const int P3 = ipow3(m->P), Q3 = ipow3(m->Q);
  double   *E  = malloc(sizeof(double) * (size_t)m->nelem * ncomp * P3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->nelem; e++) {
    double ue[512], uq[512], duq[3 * 512], vq[512], gvq[3 * 512], re[512], xq[3 * 512];
    elem_coords(m, e, xq, NULL);
    const double *qd = &qdata[(size_t)e * 10 * Q3];
    for (int n = 0; n < P3; n++) {
      memset(ue, 0, sizeof(double) * P3);
      ue[n] = 1.0;
      interp_grad(m->P, m->Q, m->B, m->D, 1, ue, uq, duq);
      for (int i = 0; i < Q3; i++) {
        double wdetJ, dXdx[3][3], A7[7], x[3], alpha[3][3];
        unpack_qdata(qd, Q3, i, &wdetJ, dXdx);
        for (int k = 0; k < 7; k++) A7[k] = Acol[((size_t)e * 7 + k) * Q3 + i];
        for (int dd = 0; dd < 3; dd++) x[dd] = xq[dd * Q3 + i];
        o_filter_alpha(f, A7, x, alpha);
        vq[i] = wdetJ * uq[i];
        double dq[3], dqX[3];
        for (int k = 0; k < 3; k++) dq[k] = duq[0 * Q3 + i] * dXdx[0][k] + duq[1 * Q3 + i] * dXdx[1][k] + duq[2 * Q3 + i] * dXdx[2][k];
        for (int k = 0; k < 3; k++) dqX[k] = dXdx[k][0] * dq[0] + dXdx[k][1] * dq[1] + dXdx[k][2] * dq[2];
        for (int k = 0; k < 3; k++) gvq[k * Q3 + i] = wdetJ * (alpha[k][0] * dqX[0] + alpha[k][1] * dqX[1] + alpha[k][2] * dqX[2]);
      }
      interp_grad_t(m->P, m->Q, m->B, m->D, 1, vq, gvq, re);
      for (int k = 0; k < ncomp; k++) E[((size_t)e * ncomp + k) * P3 + n] = re[n];
    }
  }
  memset(d, 0, sizeof(double) * ncomp * (size_t)m->nnodes);
  scatterN_add(m, E, ncomp, d);
  free(E);
  return 0;
}

// KSPCG with PCJACOBI (diagonal) and KSP_NORM_NATURAL from a zero initial guess (src/differential_filter.c:176-185):
// converged when sqrt(r.z) <= max(rtol sqrt(r0.z0), atol); at most max_it iterations.  Returns the iteration count.
int oracle_diff_filter_solve(const OMesh *m, const ODiffFilter *f, const double *qdata, const double *Acol, int ncomp, const double *R, double rtol,
                             double atol, int max_it, double *X) {
  const size_t n    = (size_t)ncomp * m->nnodes;
  double      *dinv = malloc(sizeof(double) * n), *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n),
         *Ap = malloc(sizeof(double) * n);
  oracle_diff_filter_lhs_diagonal(m, f, qdata, Acol, ncomp, dinv);
  for (size_t i = 0; i < n; i++) dinv[i] = dinv[i] != 0 ? 1.0 / dinv[i] : 1.0;
  memset(X, 0, sizeof(double) * n);
  memcpy(r, R, sizeof(double) * n);
  for (size_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
  memcpy(p, z, sizeof(double) * n);
  double rz = 0;
  for (size_t i = 0; i < n; i++) rz += r[i] * z[i];
  const double ttol = fmax(rtol * sqrt(fabs(rz)), atol);
  int          its  = 0;
  while (sqrt(fabs(rz)) > ttol && its < max_it) {
    oracle_diff_filter_lhs(m, f, qdata, Acol, ncomp, p, Ap);
    double pAp = 0;
    for (size_t i = 0; i < n; i++) pAp += p[i] * Ap[i];
    const double alpha = rz / pAp;
    for (size_t i = 0; i < n; i++) {
      X[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
      z[i] = dinv[i] * r[i];
    }
    double rz_new = 0;
    for (size_t i = 0; i < n; i++) rz_new += r[i] * z[i];
    const double beta = rz_new / rz;
    for (size_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    rz = rz_new;
    its++;
  }
  free(dinv), free(r), free(z), free(p), free(Ap);
  return its;
}

// ComputeSGS_DDAnisotropicTrainingDataNodal_Prim (qfunctions/sgs_dd_training.h:23-60) at every node: Yf [nnodes][11]
// (filtered primitive state, filtered velocity products), G [nnodes][9], A [nnodes][7] -> [nnodes][12]: the six model
// inputs, then the filtered SGS stress components XX YY ZZ YZ XZ XY
int oracle_sgs_training_nodal(const OGas *g, int nnodes, const double *Yf, const double *G, const double *A, double *out) {
#pragma omp parallel for schedule(static)
  for (int nd = 0; nd < nnodes; nd++) {
    const double *yf = &Yf[(size_t)nd * 11], *vp = yf + 5;
    const OState  s  = o_state_from_q(*g, yf, 1);
    double        gu[3][3], E[3][3], in[6], gmag;
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) gu[j][k] = G[(size_t)nd * 9 + j * 3 + k];
    o_sgs_inputs(gu, &A[(size_t)nd * 7], A[(size_t)nd * 7 + 6], g->mu / s.U[0], E, in, &gmag);
    double *o = &out[(size_t)nd * 12];
    for (int k = 0; k < 6; k++) o[k] = in[k];
    const double *u = &s.Y[1];
    o[6]            = vp[0] - u[0] * u[0];
    o[7]            = vp[1] - u[1] * u[1];
    o[8]            = vp[2] - u[2] * u[2];
    o[9]            = vp[3] - u[1] * u[2];
    o[10]           = vp[4] - u[0] * u[2];
    o[11]           = vp[5] - u[0] * u[1];
  }
  return 0;
}

// Pointwise entry points (tests of the eigen solver, anisotropy tensor and model)
int oracle_diagonalize3(const double *A_in, int max_sweeps, double *eval, double *evec) {
  double A[3][3], E[3][3];
  memcpy(A, A_in, sizeof(A));
  const int it = o_diagonalize3(A, eval, E, max_sweeps);
  memcpy(evec, E, sizeof(E));
  return it;
}

void oracle_anisotropy_tensor(const double *dXdx, double *km_A, double *delta) { o_anisotropy_tensor((const double(*)[3])dXdx, km_A, delta); }

void oracle_sgs_point(const OSgsModel *model, const double *gu, const double *km_A, double delta, double nu, double *inputs, double *outputs,
                      double *evec, double *gmag, double *km_sgs) {
  double E[3][3], in[6], out[6];
  o_sgs_inputs((const double(*)[3])gu, km_A, delta, nu, E, in, gmag);
  o_dd_inference(model, in, out);
  for (int k = 0; k < 6; k++) inputs[k] = in[k], outputs[k] = out[k];
  memcpy(evec, E, sizeof(E));
  o_sgs_outputs(model, out, delta, E, *gmag, km_sgs);
}
