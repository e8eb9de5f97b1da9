// CPU ORACLE -- test infrastructure only (see ns_pointwise.h).
//
// Boundary-face operators of the HONEE Newtonian problem, restated with libCEED /cpu/self semantics on the faces of a
// hexahedral mesh: the consistent boundary integral (qfunctions/newtonian.h:328-443), the freestream Riemann
// boundary (qfunctions/bc_freestream.h:14-138), the Riemann and weak-pressure outflows (qfunctions/bc_outflow.h:37-305)
// and the slip wall (qfunctions/bc_slip.h:11-111), with their Jacobians and the boundary quadrature data
// (qfunctions/setupgeo.h:100-117).
//
// A face is given by its volume element and local face id (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z).  Face-local node (i, j),
// i along the first face coordinate X1, j along X2, is element node (a, b, c) with the face coordinates chosen so that
// dx/dX1 x dx/dX2 points out of the element:
//   -x: a = 0,   X1 = z, X2 = y      +x: a = P-1, X1 = y, X2 = z
//   -y: b = 0,   X1 = x, X2 = z      +y: b = P-1, X1 = z, X2 = x
//   -z: c = 0,   X1 = y, X2 = x      +z: c = P-1, X1 = x, X2 = y
// Face data layouts: qdata_sur [nfaces][10][Q^2] = {wdetJb, normal[3], dXdx[2][3]}; jac_sur [nfaces][ncomp][Q^2].
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "honee_oracle.h"

// Element node (a, b, c) of face-local node (i, j)
static void face_node(int fid, int P, int i, int j, int abc[3]) {
  const int L = P - 1;
  switch (fid) {
    case 0: abc[0] = 0, abc[2] = i, abc[1] = j; break;
    case 1: abc[0] = L, abc[1] = i, abc[2] = j; break;
    case 2: abc[1] = 0, abc[0] = i, abc[2] = j; break;
    case 3: abc[1] = L, abc[2] = i, abc[0] = j; break;
    case 4: abc[2] = 0, abc[1] = i, abc[0] = j; break;
    default: abc[2] = L, abc[0] = i, abc[1] = j; break;
  }
}

// Face node indices of face k (entries of an element offset table with Pe nodes per direction): out[P^2]
static void face_offsets(const OMesh *m, const OFaces *f, int k, int P, const int *offsets, int Pe, int *out) {
  for (int j = 0; j < P; j++)
    for (int i = 0; i < P; i++) {
      int abc[3];
      face_node(f->face_id[k], P, i, j, abc);
      out[i + P * j] = offsets[(long)f->face_elem[k] * Pe * Pe * Pe + abc[0] + Pe * (abc[1] + Pe * abc[2])];
    }
}

// v[Q][Q] = (A2 (x) A1) u[P][P] with the first index fastest: contract X1 then X2
static void tensor2(int P, int Q, const double *A1, const double *A2, const double *u, double *v) {
  double t[64];
  for (int j = 0; j < P; j++)
    for (int qi = 0; qi < Q; qi++) {
      double acc = 0;
      for (int i = 0; i < P; i++) acc += A1[qi * P + i] * u[i + P * j];
      t[qi + Q * j] = acc;
    }
  for (int qj = 0; qj < Q; qj++)
    for (int qi = 0; qi < Q; qi++) {
      double acc = 0;
      for (int j = 0; j < P; j++) acc += A2[qj * P + j] * t[qi + Q * j];
      v[qi + Q * qj] = acc;
    }
}

// u[P][P] += (A2 (x) A1)^T v[Q][Q]
static void tensor2_t_add(int P, int Q, const double *A1, const double *A2, const double *v, double *u) {
  double t[64];
  for (int qj = 0; qj < Q; qj++)
    for (int i = 0; i < P; i++) {
      double acc = 0;
      for (int qi = 0; qi < Q; qi++) acc += A1[qi * P + i] * v[qi + Q * qj];
      t[i + P * qj] = acc;
    }
  for (int j = 0; j < P; j++)
    for (int i = 0; i < P; i++) {
      double acc = 0;
      for (int qj = 0; qj < Q; qj++) acc += A2[qj * P + j] * t[i + P * qj];
      u[i + P * j] += acc;
    }
}

// Gather the 5 components of a face: u_f[5][P^2]
static void face_gather5(const OMesh *m, const OFaces *f, int k, const double *L, double *u_f) {
  const int P = m->P, P2 = P * P;
  int       off[64];
  face_offsets(m, f, k, P, m->offsets, P, off);
  for (int c = 0; c < 5; c++)
    for (int n = 0; n < P2; n++) u_f[c * P2 + n] = L[off[n] * 5 + c];
}

static void face_scatter5(const OMesh *m, const OFaces *f, int k, const double *r_f, double *L) {
  const int P = m->P, P2 = P * P;
  int       off[64];
  face_offsets(m, f, k, P, m->offsets, P, off);
  for (int c = 0; c < 5; c++)
    for (int n = 0; n < P2; n++) L[off[n] * 5 + c] += r_f[c * P2 + n];
}

// Interp and reference gradient (d/dX1, d/dX2) of the 5 components at the face quadrature points
static void face_interp_grad(const OMesh *m, const double *u_f, double *uq /*[5][Q2]*/, double *gq /*[2][5][Q2]*/) {
  const int P = m->P, Q = m->Q, P2 = P * P, Q2 = Q * Q;
  for (int c = 0; c < 5; c++) {
    tensor2(P, Q, m->B, m->B, &u_f[c * P2], &uq[c * Q2]);
    if (gq) {
      tensor2(P, Q, m->D, m->B, &u_f[c * P2], &gq[(0 * 5 + c) * Q2]);
      tensor2(P, Q, m->B, m->D, &u_f[c * P2], &gq[(1 * 5 + c) * Q2]);
    }
  }
}

// SetupBoundary (setupgeo.h:100-117) with NormalVectorFromdxdX_3D and InvertBoundaryMappingJacobian_3D
// (setupgeo_helpers.h:78-91, 143-171)
int oracle_setup_qdata_boundary(const OMesh *m, const OFaces *f, double *qdata_sur) {
  const int Q = m->Q, Q2 = Q * Q, Px = m->Px;
  for (int k = 0; k < f->nfaces; k++) {
    int xo[4];
    face_offsets(m, f, k, Px, m->xoffsets, Px, xo);
    double xv[3][4];
    for (int n = 0; n < 4; n++)
      for (int d = 0; d < 3; d++) xv[d][n] = m->xcoord[xo[n] * 3 + d];
    double dx1[3][64], dx2[3][64];
    for (int d = 0; d < 3; d++) {
      tensor2(Px, Q, m->Dx, m->Bx, xv[d], dx1[d]);
      tensor2(Px, Q, m->Bx, m->Dx, xv[d], dx2[d]);
    }
    for (int q = 0; q < Q2; q++) {
      const double dxdX[3][2] = {{dx1[0][q], dx2[0][q]}, {dx1[1][q], dx2[1][q]}, {dx1[2][q], dx2[2][q]}};
      double       normal[3];
      normal[0]          = dxdX[1][0] * dxdX[2][1] - dxdX[2][0] * dxdX[1][1];
      normal[1]          = dxdX[2][0] * dxdX[0][1] - dxdX[0][0] * dxdX[2][1];
      normal[2]          = dxdX[0][0] * dxdX[1][1] - dxdX[1][0] * dxdX[0][1];
      const double detJb = sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
      for (int d = 0; d < 3; d++) normal[d] *= 1 / detJb;
      double JTJ[2][2] = {{0.}};
      for (int j = 0; j < 2; j++)
        for (int l2 = 0; l2 < 2; l2++)
          for (int l = 0; l < 3; l++) JTJ[j][l2] += dxdX[l][j] * dxdX[l][l2];
      const double det      = JTJ[0][0] * JTJ[1][1] - JTJ[1][0] * JTJ[0][1];
      const double inv[2][2] = {{JTJ[1][1] / det, -JTJ[0][1] / det}, {-JTJ[1][0] / det, JTJ[0][0] / det}};
      double       dXdx[2][3];
      for (int j = 0; j < 2; j++)
        for (int kk = 0; kk < 3; kk++) {
          dXdx[j][kk] = 0;
          for (int l = 0; l < 2; l++) dXdx[j][kk] += inv[l][j] * dxdX[kk][l];
        }
      const double w  = m->w1d[q % Q] * m->w1d[q / Q];
      double      *qd = &qdata_sur[(long)k * 10 * Q2];
      qd[0 * Q2 + q]  = w * detJb;
      for (int d = 0; d < 3; d++) qd[(1 + d) * Q2 + q] = normal[d];
      for (int j = 0; j < 6; j++) qd[(4 + j) * Q2 + q] = (&dXdx[0][0])[j];
    }
  }
  return 0;
}

static void unpack_qdata_sur(const double *qd, int Q2, int q, double *wdetJb, double dXdx[2][3], double normal[3]) {
  *wdetJb = qd[q];
  for (int d = 0; d < 3; d++) normal[d] = qd[(1 + d) * Q2 + q];
  for (int j = 0; j < 6; j++) (&dXdx[0][0])[j] = qd[(4 + j) * Q2 + q];
}

// StatePhysicalGradientFromReference_Boundary (newtonian_state.h:595-605)
static void phys_gradient_boundary(OGas g, const OState *s, int var, const double *gq, int Q2, int q, const double dXdx[2][3], OState gs[3]) {
  for (int k = 0; k < 3; k++) {
    double dqi[5];
    for (int j = 0; j < 5; j++) dqi[j] = gq[(0 * 5 + j) * Q2 + q] * dXdx[0][k] + gq[(1 * 5 + j) * Q2 + q] * dXdx[1][k];
    gs[k] = o_state_from_q_fwd(g, s, dqi, var);
  }
}

// BoundaryIntegral (newtonian.h:328-375)
int oracle_boundary_integral(const OMesh *m, const OFaces *f, const ONewt *c, const double *qdata_sur, const double *q, double *v,
                             double *jac_sur) {
  const int P2 = m->P * m->P, Q2 = m->Q * m->Q;
  for (int k = 0; k < f->nfaces; k++) {
    double u_f[5 * 64], uq[5 * 64], gq[10 * 64], vq[5 * 64], r_f[5 * 64];
    face_gather5(m, f, k, q, u_f);
    face_interp_grad(m, u_f, uq, gq);
    const double *qd = &qdata_sur[(long)k * 10 * Q2];
    for (int p = 0; p < Q2; p++) {
      double qi[5];
      for (int j = 0; j < 5; j++) qi[j] = uq[j * Q2 + p];
      const OState s = o_state_from_q(c->gas, qi, c->state_var);
      double       wdetJb, dXdx[2][3], normal[3];
      unpack_qdata_sur(qd, Q2, p, &wdetJb, dXdx, normal);
      OState gs[3];
      phys_gradient_boundary(c->gas, &s, c->state_var, gq, Q2, p, dXdx, gs);
      double sr[6], kmst[6], st[3][3], Fe[3], Finv[3][5], Flux[5];
      o_km_strain_rate(gs, sr);
      o_newtonian_stress(c->gas, sr, kmst);
      o_km_unpack(kmst, st);
      o_viscous_energy_flux(c->gas, s.Y, gs, st, Fe);
      o_flux_inviscid(&s, Finv);
      o_flux_total_boundary(Finv, st, Fe, normal, Flux);
      for (int j = 0; j < 5; j++) vq[j * Q2 + p] = wdetJb * Flux[j];
      if (c->is_implicit) {
        double *jd = &jac_sur[(long)k * 11 * Q2];
        for (int j = 0; j < 5; j++) jd[j * Q2 + p] = qi[j];
        for (int j = 0; j < 6; j++) jd[(5 + j) * Q2 + p] = sr[j];
      } else {
        for (int j = 0; j < 5; j++) vq[j * Q2 + p] *= -1;
      }
    }
    memset(r_f, 0, sizeof(double) * 5 * P2);
    for (int cc = 0; cc < 5; cc++) tensor2_t_add(m->P, m->Q, m->B, m->B, &vq[cc * Q2], &r_f[cc * P2]);
    face_scatter5(m, f, k, r_f, v);
  }
  return 0;
}

// BoundaryIntegral_Jacobian (newtonian.h:392-431)
int oracle_boundary_integral_jacobian(const OMesh *m, const OFaces *f, const ONewt *c, const double *qdata_sur, const double *jac_sur,
                                      const double *dq, double *v) {
  const int P2 = m->P * m->P, Q2 = m->Q * m->Q;
  for (int k = 0; k < f->nfaces; k++) {
    double u_f[5 * 64], uq[5 * 64], gq[10 * 64], vq[5 * 64], r_f[5 * 64];
    face_gather5(m, f, k, dq, u_f);
    face_interp_grad(m, u_f, uq, gq);
    const double *qd = &qdata_sur[(long)k * 10 * Q2];
    const double *jd = &jac_sur[(long)k * 11 * Q2];
    for (int p = 0; p < Q2; p++) {
      double wdetJb, dXdx[2][3], normal[3], qi[5], sr[6], dqi[5];
      unpack_qdata_sur(qd, Q2, p, &wdetJb, dXdx, normal);
      for (int j = 0; j < 5; j++) qi[j] = jd[j * Q2 + p];
      for (int j = 0; j < 6; j++) sr[j] = jd[(5 + j) * Q2 + p];
      for (int j = 0; j < 5; j++) dqi[j] = uq[j * Q2 + p];
      const OState s  = o_state_from_q(c->gas, qi, c->state_var);
      const OState ds = o_state_from_q_fwd(c->gas, &s, dqi, c->state_var);
      OState       gds[3];
      phys_gradient_boundary(c->gas, &s, c->state_var, gq, Q2, p, dXdx, gds);
      double dsr[6], dkm[6], km[6], st[3][3], dst[3][3], dFe[3], dFinv[3][5], dFlux[5];
      o_km_strain_rate(gds, dsr);
      o_newtonian_stress(c->gas, dsr, dkm);
      o_newtonian_stress(c->gas, sr, km);
      o_km_unpack(dkm, dst);
      o_km_unpack(km, st);
      o_viscous_energy_flux_fwd(c->gas, s.Y, ds.Y, gds, st, dst, dFe);
      o_flux_inviscid_fwd(&s, &ds, dFinv);
      o_flux_total_boundary(dFinv, dst, dFe, normal, dFlux);
      for (int j = 0; j < 5; j++) vq[j * Q2 + p] = wdetJb * dFlux[j];
    }
    memset(r_f, 0, sizeof(double) * 5 * P2);
    for (int cc = 0; cc < 5; cc++) tensor2_t_add(m->P, m->Q, m->B, m->B, &vq[cc * Q2], &r_f[cc * P2]);
    face_scatter5(m, f, k, r_f, v);
  }
  return 0;
}

static OState exterior_state(const OFreestream *c) {
  OState s;
  memcpy(s.U, c->S_inf_U, sizeof(s.U));
  memcpy(s.Y, c->S_inf_Y, sizeof(s.Y));
  return s;
}

// Freestream (bc_freestream.h:14-50)
int oracle_freestream(const OMesh *m, const OFaces *f, const OFreestream *c, const double *qdata_sur, const double *q, double *v, double *jac_sur) {
  const int    P2 = m->P * m->P, Q2 = m->Q * m->Q;
  const OState S_inf = exterior_state(c);
  for (int k = 0; k < f->nfaces; k++) {
    double u_f[5 * 64], uq[5 * 64], vq[5 * 64], r_f[5 * 64];
    face_gather5(m, f, k, q, u_f);
    face_interp_grad(m, u_f, uq, NULL);
    const double *qd = &qdata_sur[(long)k * 10 * Q2];
    for (int p = 0; p < Q2; p++) {
      double qi[5], wdetJb, dXdx[2][3], normal[3], F[5];
      for (int j = 0; j < 5; j++) qi[j] = uq[j * Q2 + p];
      const OState s = o_state_from_q(c->newt.gas, qi, c->newt.state_var);
      unpack_qdata_sur(qd, Q2, p, &wdetJb, dXdx, normal);
      if (c->riemann == 0) o_riemann_hll(c->newt.gas, &s, &S_inf, normal, F);
      else o_riemann_hllc(c->newt.gas, &s, &S_inf, normal, F);
      for (int j = 0; j < 5; j++) vq[j * Q2 + p] = wdetJb * F[j];
      if (c->newt.is_implicit) {
        for (int j = 0; j < 5; j++) jac_sur[((long)k * 5 + j) * Q2 + p] = qi[j];
      } else {
        for (int j = 0; j < 5; j++) vq[j * Q2 + p] *= -1;
      }
    }
    memset(r_f, 0, sizeof(double) * 5 * P2);
    for (int cc = 0; cc < 5; cc++) tensor2_t_add(m->P, m->Q, m->B, m->B, &vq[cc * Q2], &r_f[cc * P2]);
    face_scatter5(m, f, k, r_f, v);
  }
  return 0;
}

// Freestream_Jacobian (bc_freestream.h:76-114): exterior state held fixed (dS_infty = 0)
int oracle_freestream_jacobian(const OMesh *m, const OFaces *f, const OFreestream *c, const double *qdata_sur, const double *jac_sur,
                               const double *dq, double *v) {
  const int    P2 = m->P * m->P, Q2 = m->Q * m->Q;
  const OState S_inf = exterior_state(c);
  OState       dS_inf;
  memset(&dS_inf, 0, sizeof(dS_inf));
  for (int k = 0; k < f->nfaces; k++) {
    double u_f[5 * 64], uq[5 * 64], vq[5 * 64], r_f[5 * 64];
    face_gather5(m, f, k, dq, u_f);
    face_interp_grad(m, u_f, uq, NULL);
    const double *qd = &qdata_sur[(long)k * 10 * Q2];
    for (int p = 0; p < Q2; p++) {
      double qi[5], dqi[5], wdetJb, dXdx[2][3], normal[3], dF[5];
      unpack_qdata_sur(qd, Q2, p, &wdetJb, dXdx, normal);
      for (int j = 0; j < 5; j++) qi[j] = jac_sur[((long)k * 5 + j) * Q2 + p];
      for (int j = 0; j < 5; j++) dqi[j] = uq[j * Q2 + p];
      const OState s  = o_state_from_q(c->newt.gas, qi, c->newt.state_var);
      const OState ds = o_state_from_q_fwd(c->newt.gas, &s, dqi, c->newt.state_var);
      if (c->riemann == 0) o_riemann_hll_fwd(c->newt.gas, &s, &ds, &S_inf, &dS_inf, normal, dF);
      else o_riemann_hllc_fwd(c->newt.gas, &s, &ds, &S_inf, &dS_inf, normal, dF);
      for (int j = 0; j < 5; j++) vq[j * Q2 + p] = wdetJb * dF[j];
    }
    memset(r_f, 0, sizeof(double) * 5 * P2);
    for (int cc = 0; cc < 5; cc++) tensor2_t_add(m->P, m->Q, m->B, m->B, &vq[cc * Q2], &r_f[cc * P2]);
    face_scatter5(m, f, k, r_f, v);
  }
  return 0;
}

// Softplus (bc_outflow.h:18-21) and the value its forward mode returns (bc_outflow.h:23-27): the derivative factor,
// not multiplied by dx
static double softplus(double x, double width) { return x > 40 * width ? x : width * log1p(exp(x / width)); }
static double softplus_fwd(double x, double dx, double width) {
  (void)dx;
  if (x > 40 * width) return 1;
  const double t = exp(x / width);
  return t / (1 + t);
}

// FluxTotal_RiemannBoundary (newtonian_state.h:498-509)
static void flux_total_riemann_boundary(const double Fn[5], const double st[3][3], const double Fe[3], const double n[3], double Flux[5]) {
  for (int j = 0; j < 5; j++) Flux[j] = Fn[j];
  for (int j = 0; j < 3; j++) {
    for (int k = 0; k < 3; k++) Flux[k + 1] -= st[k][j] * n[j];
    Flux[4] += Fe[j] * n[j];
  }
}

// RiemannOutflow (bc_outflow.h:37-89) and PressureOutflow (bc_outflow.h:185-229)
int oracle_outflow(const OMesh *m, const OFaces *f, const OOutflow *c, const double *qdata_sur, const double *q, double *v, double *jac_sur) {
  const int  P2 = m->P * m->P, Q2 = m->Q * m->Q;
  const OGas g   = c->newt.gas;
  const int  var = c->newt.state_var;
  for (int k = 0; k < f->nfaces; k++) {
    double u_f[5 * 64], uq[5 * 64], gq[10 * 64], vq[5 * 64], r_f[5 * 64];
    face_gather5(m, f, k, q, u_f);
    face_interp_grad(m, u_f, uq, gq);
    const double *qd = &qdata_sur[(long)k * 10 * Q2];
    for (int p = 0; p < Q2; p++) {
      double qi[5], wdetJb, dXdx[2][3], normal[3];
      for (int j = 0; j < 5; j++) qi[j] = uq[j * Q2 + p];
      unpack_qdata_sur(qd, Q2, p, &wdetJb, dXdx, normal);
      OState s = o_state_from_q(g, qi, var);
      OState gs[3];
      double sr[6], km[6], st[3][3], Fe[3], Flux[5];
      if (c->pressure_outflow) {
        s.Y[0] = c->pressure;
        wdetJb *= c->newt.is_implicit ? -1. : 1.;
        phys_gradient_boundary(g, &s, var, gq, Q2, p, dXdx, gs);
        o_km_strain_rate(gs, sr);
        o_newtonian_stress(g, sr, km);
        o_km_unpack(km, st);
        o_viscous_energy_flux(g, s.Y, gs, st, Fe);
        double Finv[3][5];
        o_flux_inviscid(&s, Finv);
        o_flux_total_boundary(Finv, st, Fe, normal, Flux);
        for (int j = 0; j < 5; j++) vq[j * Q2 + p] = -wdetJb * Flux[j];
      } else {
        double ye[5];
        memcpy(ye, s.Y, sizeof(ye));
        ye[0]              = c->pressure;
        ye[4]              = c->temperature;
        const double un    = o_dot3(&ye[1], normal);
        const double proj  = (1 - c->recirc) * softplus(-un, c->softplus_velocity);
        for (int j = 0; j < 3; j++) ye[1 + j] += normal[j] * proj;
        const OState s_ext = o_state_from_q(g, ye, OVAR_PRIM);
        phys_gradient_boundary(g, &s, var, gq, Q2, p, dXdx, gs);
        o_km_strain_rate(gs, sr);
        o_newtonian_stress(g, sr, km);
        o_km_unpack(km, st);
        o_viscous_energy_flux(g, s.Y, gs, st, Fe);
        double Fn[5];
        o_riemann_hllc(g, &s, &s_ext, normal, Fn);
        flux_total_riemann_boundary(Fn, st, Fe, normal, Flux);
        for (int j = 0; j < 5; j++) vq[j * Q2 + p] = wdetJb * Flux[j];
        if (!c->newt.is_implicit)
          for (int j = 0; j < 5; j++) vq[j * Q2 + p] *= -1;
      }
      if (c->newt.is_implicit) {
        double *jd = &jac_sur[(long)k * 11 * Q2];
        for (int j = 0; j < 5; j++) jd[j * Q2 + p] = qi[j];
        for (int j = 0; j < 6; j++) jd[(5 + j) * Q2 + p] = sr[j];
      }
    }
    memset(r_f, 0, sizeof(double) * 5 * P2);
    for (int cc = 0; cc < 5; cc++) tensor2_t_add(m->P, m->Q, m->B, m->B, &vq[cc * Q2], &r_f[cc * P2]);
    face_scatter5(m, f, k, r_f, v);
  }
  return 0;
}

// RiemannOutflow_Jacobian (bc_outflow.h:106-164) and PressureOutflow_Jacobian (bc_outflow.h:246-293)
int oracle_outflow_jacobian(const OMesh *m, const OFaces *f, const OOutflow *c, const double *qdata_sur, const double *jac_sur, const double *dq,
                            double *v) {
  const int  P2 = m->P * m->P, Q2 = m->Q * m->Q;
  const OGas g   = c->newt.gas;
  const int  var = c->newt.state_var;
  for (int k = 0; k < f->nfaces; k++) {
    double u_f[5 * 64], uq[5 * 64], gq[10 * 64], vq[5 * 64], r_f[5 * 64];
    face_gather5(m, f, k, dq, u_f);
    face_interp_grad(m, u_f, uq, gq);
    const double *qd = &qdata_sur[(long)k * 10 * Q2];
    const double *jd = &jac_sur[(long)k * 11 * Q2];
    for (int p = 0; p < Q2; p++) {
      double wdetJb, dXdx[2][3], normal[3], qi[5], sr[6], dqi[5];
      unpack_qdata_sur(qd, Q2, p, &wdetJb, dXdx, normal);
      for (int j = 0; j < 5; j++) qi[j] = jd[j * Q2 + p];
      for (int j = 0; j < 6; j++) sr[j] = jd[(5 + j) * Q2 + p];
      for (int j = 0; j < 5; j++) dqi[j] = uq[j * Q2 + p];
      double km[6], st[3][3], dsr[6], dkm[6], dst[3][3], dFe[3], dFlux[5];
      o_newtonian_stress(g, sr, km);
      OState s  = o_state_from_q(g, qi, var);
      OState ds = o_state_from_q_fwd(g, &s, dqi, var);
      OState gds[3];
      if (c->pressure_outflow) {
        wdetJb *= c->newt.is_implicit ? -1. : 1.;
        s.Y[0]  = c->pressure;
        ds.Y[0] = 0.;
        phys_gradient_boundary(g, &s, var, gq, Q2, p, dXdx, gds);
        o_km_strain_rate(gds, dsr);
        o_newtonian_stress(g, dsr, dkm);
        o_km_unpack(dkm, dst);
        o_km_unpack(km, st);
        o_viscous_energy_flux_fwd(g, s.Y, ds.Y, gds, st, dst, dFe);
        double dFinv[3][5];
        o_flux_inviscid_fwd(&s, &ds, dFinv);
        o_flux_total_boundary(dFinv, dst, dFe, normal, dFlux);
        for (int j = 0; j < 5; j++) vq[j * Q2 + p] = -wdetJb * dFlux[j];
      } else {
        double ye[5], dye[5];
        memcpy(ye, s.Y, sizeof(ye));
        memcpy(dye, ds.Y, sizeof(dye));
        ye[0] = c->pressure, ye[4] = c->temperature;
        dye[0] = 0, dye[4] = 0;
        const double un    = o_dot3(&s.Y[1], normal);
        const double dun   = o_dot3(&ds.Y[1], normal);
        const double proj  = (1 - c->recirc) * softplus(-un, c->softplus_velocity);
        const double dproj = (1 - c->recirc) * softplus_fwd(-un, -dun, c->softplus_velocity);
        for (int j = 0; j < 3; j++) {
          ye[1 + j] += normal[j] * proj;
          dye[1 + j] += normal[j] * dproj;
        }
        const OState s_ext  = o_state_from_q(g, ye, OVAR_PRIM);
        const OState ds_ext = o_state_from_y_fwd(g, &s_ext, dye);
        phys_gradient_boundary(g, &s, var, gq, Q2, p, dXdx, gds);
        o_km_strain_rate(gds, dsr);
        o_newtonian_stress(g, dsr, dkm);
        o_km_unpack(dkm, dst);
        o_km_unpack(km, st);
        o_viscous_energy_flux_fwd(g, s.Y, ds.Y, gds, st, dst, dFe);
        double dFn[5];
        o_riemann_hllc_fwd(g, &s, &ds, &s_ext, &ds_ext, normal, dFn);
        flux_total_riemann_boundary(dFn, dst, dFe, normal, dFlux);
        for (int j = 0; j < 5; j++) vq[j * Q2 + p] = wdetJb * dFlux[j];
      }
    }
    memset(r_f, 0, sizeof(double) * 5 * P2);
    for (int cc = 0; cc < 5; cc++) tensor2_t_add(m->P, m->Q, m->B, m->B, &vq[cc * Q2], &r_f[cc * P2]);
    face_scatter5(m, f, k, r_f, v);
  }
  return 0;
}

// Mirror state of the slip wall (bc_slip.h:29-33): the normal velocity component reflected
static void slip_mirror(const double Y[5], const double n[3], double Yr[5]) {
  const double vn = o_dot3(&Y[1], n);
  Yr[0]           = Y[0];
  for (int j = 0; j < 3; j++) Yr[1 + j] = Y[1 + j] - 2. * n[j] * vn;
  Yr[4] = Y[4];
}

// Slip (bc_slip.h:11-46)
int oracle_slip(const OMesh *m, const OFaces *f, const ONewt *c, const double *qdata_sur, const double *q, double *v, double *jac_sur) {
  const int P2 = m->P * m->P, Q2 = m->Q * m->Q;
  for (int k = 0; k < f->nfaces; k++) {
    double u_f[5 * 64], uq[5 * 64], vq[5 * 64], r_f[5 * 64];
    face_gather5(m, f, k, q, u_f);
    face_interp_grad(m, u_f, uq, NULL);
    const double *qd = &qdata_sur[(long)k * 10 * Q2];
    for (int p = 0; p < Q2; p++) {
      double qi[5], wdetJb, dXdx[2][3], normal[3], Yr[5], F[5];
      for (int j = 0; j < 5; j++) qi[j] = uq[j * Q2 + p];
      unpack_qdata_sur(qd, Q2, p, &wdetJb, dXdx, normal);
      const OState s = o_state_from_q(c->gas, qi, c->state_var);
      slip_mirror(s.Y, normal, Yr);
      const OState s_r = o_state_from_q(c->gas, Yr, OVAR_PRIM);
      o_riemann_hllc(c->gas, &s, &s_r, normal, F);
      for (int j = 0; j < 5; j++) vq[j * Q2 + p] = wdetJb * F[j];
      if (c->is_implicit) {
        for (int j = 0; j < 5; j++) jac_sur[((long)k * 5 + j) * Q2 + p] = qi[j];
      } else {
        for (int j = 0; j < 5; j++) vq[j * Q2 + p] *= -1;
      }
    }
    memset(r_f, 0, sizeof(double) * 5 * P2);
    for (int cc = 0; cc < 5; cc++) tensor2_t_add(m->P, m->Q, m->B, m->B, &vq[cc * Q2], &r_f[cc * P2]);
    face_scatter5(m, f, k, r_f, v);
  }
  return 0;
}

// Slip_Jacobian (bc_slip.h:60-99)
int oracle_slip_jacobian(const OMesh *m, const OFaces *f, const ONewt *c, const double *qdata_sur, const double *jac_sur, const double *dq,
                         double *v) {
  const int P2 = m->P * m->P, Q2 = m->Q * m->Q;
  for (int k = 0; k < f->nfaces; k++) {
    double u_f[5 * 64], uq[5 * 64], vq[5 * 64], r_f[5 * 64];
    face_gather5(m, f, k, dq, u_f);
    face_interp_grad(m, u_f, uq, NULL);
    const double *qd = &qdata_sur[(long)k * 10 * Q2];
    for (int p = 0; p < Q2; p++) {
      double qi[5], dqi[5], wdetJb, dXdx[2][3], normal[3], Yr[5], dYr[5], dF[5];
      unpack_qdata_sur(qd, Q2, p, &wdetJb, dXdx, normal);
      for (int j = 0; j < 5; j++) qi[j] = jac_sur[((long)k * 5 + j) * Q2 + p];
      for (int j = 0; j < 5; j++) dqi[j] = uq[j * Q2 + p];
      const OState s  = o_state_from_q(c->gas, qi, c->state_var);
      const OState ds = o_state_from_q_fwd(c->gas, &s, dqi, c->state_var);
      slip_mirror(s.Y, normal, Yr);
      slip_mirror(ds.Y, normal, dYr);
      const OState s_r  = o_state_from_q(c->gas, Yr, OVAR_PRIM);
      const OState ds_r = o_state_from_y_fwd(c->gas, &s_r, dYr);
      o_riemann_hllc_fwd(c->gas, &s, &ds, &s_r, &ds_r, normal, dF);
      for (int j = 0; j < 5; j++) vq[j * Q2 + p] = wdetJb * dF[j];
    }
    memset(r_f, 0, sizeof(double) * 5 * P2);
    for (int cc = 0; cc < 5; cc++) tensor2_t_add(m->P, m->Q, m->B, m->B, &vq[cc * Q2], &r_f[cc * P2]);
    face_scatter5(m, f, k, r_f, v);
  }
  return 0;
}

// Pointwise entry points (states given by their q in variable set var)
void oracle_riemann_flux(const OGas *g, int hllc, const double *qL, const double *qR, int var, const double *n, double *F) {
  const OState l = o_state_from_q(*g, qL, var), r = o_state_from_q(*g, qR, var);
  if (hllc) o_riemann_hllc(*g, &l, &r, n, F);
  else o_riemann_hll(*g, &l, &r, n, F);
}

void oracle_riemann_flux_fwd(const OGas *g, int hllc, const double *qL, const double *dqL, const double *qR, const double *dqR, int var,
                             const double *n, double *dF) {
  const OState l = o_state_from_q(*g, qL, var), r = o_state_from_q(*g, qR, var);
  const OState dl = o_state_from_q_fwd(*g, &l, dqL, var), dr = o_state_from_q_fwd(*g, &r, dqR, var);
  if (hllc) o_riemann_hllc_fwd(*g, &l, &dl, &r, &dr, n, dF);
  else o_riemann_hll_fwd(*g, &l, &dl, &r, &dr, n, dF);
}

void oracle_hll_speeds(const OGas *g, const double *qL, const double *qR, int var, const double *n, double *sL, double *sR) {
  const OState l = o_state_from_q(*g, qL, var), r = o_state_from_q(*g, qR, var);
  o_hll_speeds_roe(*g, &l, o_dot3(&l.Y[1], n), &r, o_dot3(&r.Y[1], n), sL, sR);
}

void oracle_hll_speeds_fwd(const OGas *g, const double *qL, const double *dqL, const double *qR, const double *dqR, int var, const double *n,
                           double *sL, double *dsL, double *sR, double *dsR) {
  const OState l = o_state_from_q(*g, qL, var), r = o_state_from_q(*g, qR, var);
  const OState dl = o_state_from_q_fwd(*g, &l, dqL, var), dr = o_state_from_q_fwd(*g, &r, dqR, var);
  o_hll_speeds_roe_fwd(*g, &l, &dl, o_dot3(&l.Y[1], n), o_dot3(&dl.Y[1], n), &r, &dr, o_dot3(&r.Y[1], n), o_dot3(&dr.Y[1], n), sL, dsL, sR, dsR);
}
