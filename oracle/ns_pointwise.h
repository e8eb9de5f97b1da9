// CPU ORACLE -- test infrastructure only.
//
// Plain-C, FP64 restatement of HONEE's pointwise Newtonian ideal-gas physics
// (the QFunction helpers under qfunctions/ in the reference).  It is the
// checker the HIP backend is compared against; nothing in the product links,
// loads or calls it.  Each function names the reference file:line it follows.
// Compile with -ffp-contract=off so every product/sum is rounded as written.
//
// A state carries both the conservative vector U = (rho, rho u[3], E) and the
// primitive vector Y = (p, u[3], T), mirroring `State`
// (qfunctions/newtonian_state.h:13-22).
#pragma once
#include <math.h>
#include <string.h>

typedef struct {
  double lambda, mu, k, cv, cp;  // NewtonianIGProperties (qfunctions/newtonian_types.h:20-26)
} OGas;

typedef struct {
  double U[5];  // rho, m0, m1, m2, E
  double Y[5];  // p, u0, u1, u2, T
} OState;

enum { OVAR_CONS = 0, OVAR_PRIM = 1, OVAR_ENTROPY = 2 };
enum { OSTAB_NONE = 0, OSTAB_SU = 1, OSTAB_SUPG = 2 };

static inline double o_sq(double x) { return x * x; }
static inline double o_dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
// DotN accumulates from zero (qfunctions/utils.h:43-47)
static inline double o_dotn(const double *a, const double *b, int n) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static inline double o_gamma(OGas g) { return g.cp / g.cv; }      // newtonian_state.h:42
static inline double o_rgas(OGas g) { return g.cp - g.cv; }       // newtonian_state.h:44

// newtonian_state.h:64-73
static inline void o_prim_from_cons(OGas g, const double U[5], double Y[5]) {
  for (int i = 0; i < 3; i++) Y[1 + i] = U[1 + i] / U[0];
  double ek = .5 * o_dot3(&Y[1], &Y[1]);
  double et = U[4] / U[0];
  double ei = et - ek;
  Y[4]      = ei / g.cv;
  Y[0]      = (o_gamma(g) - 1) * U[0] * ei;
}

// newtonian_state.h:75-89
static inline void o_prim_from_cons_fwd(OGas g, const OState *s, const double dU[5], double dY[5]) {
  for (int i = 0; i < 3; i++) dY[1 + i] = (dU[1 + i] - s->Y[1 + i] * dU[0]) / s->U[0];
  double ek  = .5 * o_dot3(&s->Y[1], &s->Y[1]);
  double dek = o_dot3(&dY[1], &s->Y[1]);
  double et  = s->U[4] / s->U[0];
  double det = (dU[4] - et * dU[0]) / s->U[0];
  double ei  = et - ek;
  double dei = det - dek;
  dY[4]      = dei / g.cv;
  dY[0]      = (o_gamma(g) - 1) * (dU[0] * ei + s->U[0] * dei);
}

// newtonian_state.h:91-103
static inline void o_entropy_from_prim(OGas g, const double Y[5], double V[5]) {
  const double gamma   = o_gamma(g);
  const double rho     = Y[0] / (o_rgas(g) * Y[4]);
  const double entropy = log(Y[0]) - gamma * log(rho);
  const double rdp     = rho / Y[0];
  const double ek      = 0.5 * o_dot3(&Y[1], &Y[1]);
  V[0]                 = (gamma - entropy) / (gamma - 1) - rdp * ek;
  for (int i = 0; i < 3; i++) V[1 + i] = rdp * Y[1 + i];
  V[4] = -rdp;
}

// newtonian_state.h:105-121
static inline void o_entropy_from_prim_fwd(OGas g, const OState *s, const double dY[5], double dV[5]) {
  const double gamma = o_gamma(g);
  double       drho  = (dY[0] * s->Y[4] - s->Y[0] * dY[4]) / (o_rgas(g) * s->Y[4] * s->Y[4]);
  const double ek    = .5 * o_dot3(&s->Y[1], &s->Y[1]);
  const double dek   = o_dot3(&dY[1], &s->Y[1]);
  const double rdp   = s->U[0] / s->Y[0];
  const double drdp  = (drho * s->Y[0] - s->U[0] * dY[0]) / o_sq(s->Y[0]);
  double       dent  = dY[0] / s->Y[0] - gamma * drho / s->U[0];
  dV[0]              = -dent / (gamma - 1) - dek * rdp - ek * drdp;
  for (int i = 0; i < 3; i++) dV[1 + i] = rdp * dY[1 + i] + drdp * s->Y[1 + i];
  dV[4] = -drdp;
}

// newtonian_state.h:123-133
static inline void o_prim_from_entropy(OGas g, const double V[5], double Y[5]) {
  for (int i = 0; i < 3; i++) Y[1 + i] = -V[1 + i] / V[4];
  Y[4]                 = -1 / (o_rgas(g) * V[4]);
  const double gamma   = o_gamma(g);
  const double ek      = 0.5 * o_dot3(&Y[1], &Y[1]);
  const double entropy = gamma - (gamma - 1) * (V[0] - ek * V[4]);
  const double log_p   = -(entropy + gamma * log(-V[4])) / (gamma - 1);
  Y[0]                 = exp(log_p);
}

// newtonian_state.h:135-146
static inline void o_prim_from_entropy_fwd(OGas g, const OState *s, const double dV[5], double dY[5]) {
  double V[5];
  o_entropy_from_prim(g, s->Y, V);
  for (int i = 0; i < 3; i++) dY[1 + i] = -(dV[1 + i] - V[1 + i] * dV[4] / V[4]) / V[4];
  dY[4]              = dV[4] / (o_rgas(g) * V[4] * V[4]);
  const double gamma = o_gamma(g);
  const double ek    = 0.5 * o_dot3(&s->Y[1], &s->Y[1]);
  const double dek   = o_dot3(&dY[1], &s->Y[1]);
  const double dent  = (1 - gamma) * (dV[0] - ek * dV[4] - dek * V[4]);
  dY[0]              = s->Y[0] * (-dent - gamma * dV[4] / V[4]) / (gamma - 1);
}

// newtonian_state.h:148-157
static inline void o_cons_from_prim(OGas g, const double Y[5], double U[5]) {
  U[0] = Y[0] / (o_rgas(g) * Y[4]);
  for (int i = 0; i < 3; i++) U[1 + i] = U[0] * Y[1 + i];
  double ei = g.cv * Y[4];
  double ek = .5 * o_dot3(&Y[1], &Y[1]);
  double et = ei + ek;
  U[4]      = U[0] * et;
}

// newtonian_state.h:159-173
static inline void o_cons_from_prim_fwd(OGas g, const OState *s, const double dY[5], double dU[5]) {
  dU[0] = (dY[0] * s->Y[4] - s->Y[0] * dY[4]) / (o_rgas(g) * s->Y[4] * s->Y[4]);
  for (int i = 0; i < 3; i++) dU[1 + i] = dU[0] * s->Y[1 + i] + s->U[0] * dY[1 + i];
  double ek  = .5 * o_dot3(&s->Y[1], &s->Y[1]);
  double dek = o_dot3(&dY[1], &s->Y[1]);
  double ei  = g.cv * s->Y[4];
  double dei = g.cv * dY[4];
  double et  = ei + ek;
  double det = dei + dek;
  dU[4]      = dU[0] * et + s->U[0] * det;
}

// newtonian_state.h:175-187
static inline void o_entropy_from_cons(OGas g, const double U[5], double V[5]) {
  const double gamma   = o_gamma(g);
  const double ek      = .5 * o_dot3(&U[1], &U[1]) / U[0];
  const double ei      = U[4] - ek;
  const double p       = (gamma - 1) * ei;
  const double entropy = log(p) - gamma * log(U[0]);
  V[0]                 = (gamma - entropy) / (gamma - 1) - ek / p;
  for (int i = 0; i < 3; i++) V[1 + i] = U[1 + i] / p;
  V[4] = -U[0] / p;
}

// newtonian_state.h:189-206
static inline void o_entropy_from_cons_fwd(OGas g, const OState *s, const double dU[5], double dV[5]) {
  const double gamma = o_gamma(g);
  const double ek    = .5 * o_dot3(&s->U[1], &s->U[1]) / s->U[0];
  const double dek   = (o_dot3(&s->U[1], &dU[1]) - ek * dU[0]) / s->U[0];
  const double dei   = dU[4] - dek;
  const double p     = s->Y[0];
  const double dp    = (gamma - 1) * dei;
  double       dent  = dp / p - gamma * dU[0] / s->U[0];
  dV[0]              = -dent / (gamma - 1) - dek / p + dp * ek / o_sq(p);
  for (int i = 0; i < 3; i++) dV[1 + i] = (dU[1 + i] - s->U[1 + i] * dp / p) / p;
  dV[4] = -(dU[0] - s->U[0] * dp / p) / p;
}

// newtonian_state.h:208-222
static inline void o_cons_from_entropy(OGas g, const double V[5], double U[5]) {
  double u[3];
  for (int i = 0; i < 3; i++) u[i] = -V[1 + i] / V[4];
  const double gamma   = o_gamma(g);
  const double ek      = 0.5 * o_dot3(u, u);
  const double entropy = gamma - (gamma - 1) * (V[0] - ek * V[4]);
  const double log_rho = -(entropy + log(-V[4])) / (gamma - 1);
  U[0]                 = exp(log_rho);
  for (int i = 0; i < 3; i++) U[1 + i] = U[0] * u[i];
  const double ei = -g.cv / (o_rgas(g) * V[4]);
  U[4]            = U[0] * (ei + ek);
}

// newtonian_state.h:224-244
static inline void o_cons_from_entropy_fwd(OGas g, const OState *s, const double dV[5], double dU[5]) {
  double V[5], du[3];
  o_entropy_from_prim(g, s->Y, V);
  for (int i = 0; i < 3; i++) du[i] = (-dV[1 + i] - s->Y[1 + i] * dV[4]) / V[4];
  const double gamma   = o_gamma(g);
  const double ek      = 0.5 * o_dot3(&s->Y[1], &s->Y[1]);
  const double dek     = o_dot3(du, &s->Y[1]);
  const double entropy = gamma - (gamma - 1) * (V[0] - ek * V[4]);
  const double dent    = -(gamma - 1) * (dV[0] - (dek * V[4] + ek * dV[4]));
  const double log_rho = -(entropy + log(-V[4])) / (gamma - 1);
  const double rho     = exp(log_rho);
  dU[0]                = -rho / (gamma - 1) * (dent + dV[4] / V[4]);
  for (int i = 0; i < 3; i++) dU[1 + i] = dU[0] * s->Y[1 + i] + s->U[0] * du[i];
  const double ei  = -g.cv / (o_rgas(g) * V[4]);
  const double dei = g.cv * dV[4] / (o_rgas(g) * V[4] * V[4]);
  const double et  = ei + ek;
  dU[4]            = dU[0] * et + s->U[0] * (dei + dek);
}

// StateFromQ (newtonian_state.h:393-407)
static inline OState o_state_from_q(OGas g, const double q[5], int var) {
  OState s;
  switch (var) {
    case OVAR_CONS:
      memcpy(s.U, q, sizeof(s.U));
      o_prim_from_cons(g, s.U, s.Y);
      break;
    case OVAR_PRIM:
      memcpy(s.Y, q, sizeof(s.Y));
      o_cons_from_prim(g, s.Y, s.U);
      break;
    default:
      o_cons_from_entropy(g, q, s.U);
      o_prim_from_entropy(g, q, s.Y);
      break;
  }
  return s;
}

// StateFromQ_fwd (newtonian_state.h:409-423)
static inline OState o_state_from_q_fwd(OGas g, const OState *s, const double dq[5], int var) {
  OState ds;
  switch (var) {
    case OVAR_CONS:
      memcpy(ds.U, dq, sizeof(ds.U));
      o_prim_from_cons_fwd(g, s, ds.U, ds.Y);
      break;
    case OVAR_PRIM:
      memcpy(ds.Y, dq, sizeof(ds.Y));
      o_cons_from_prim_fwd(g, s, ds.Y, ds.U);
      break;
    default:
      o_cons_from_entropy_fwd(g, s, dq, ds.U);
      o_prim_from_entropy_fwd(g, s, dq, ds.Y);
      break;
  }
  return ds;
}

// StateToQ (newtonian_state.h:291-306)
static inline void o_state_to_q(OGas g, const OState *s, double q[5], int var) {
  switch (var) {
    case OVAR_CONS: memcpy(q, s->U, 5 * sizeof(double)); break;
    case OVAR_PRIM: memcpy(q, s->Y, 5 * sizeof(double)); break;
    default: o_entropy_from_prim(g, s->Y, q); break;
  }
}

// StateToQ_fwd (newtonian_state.h:308-321)
static inline void o_state_to_q_fwd(OGas g, const OState *s, const OState *ds, double dq[5], int var) {
  switch (var) {
    case OVAR_CONS: memcpy(dq, ds->U, 5 * sizeof(double)); break;
    case OVAR_PRIM: memcpy(dq, ds->Y, 5 * sizeof(double)); break;
    default: o_entropy_from_prim_fwd(g, s, ds->Y, dq); break;
  }
}

// StateFromY_fwd (newtonian_state.h:356-365)
static inline OState o_state_from_y_fwd(OGas g, const OState *s, const double dY[5]) {
  OState ds;
  memcpy(ds.Y, dY, sizeof(ds.Y));
  o_cons_from_prim_fwd(g, s, ds.Y, ds.U);
  return ds;
}

// FluxInviscid (newtonian_state.h:425-431); F[i] is the flux in direction i as a conservative 5-vector
static inline void o_flux_inviscid(const OState *s, double F[3][5]) {
  for (int i = 0; i < 3; i++) {
    F[i][0] = s->U[1 + i];
    for (int j = 0; j < 3; j++) F[i][1 + j] = s->U[1 + i] * s->Y[1 + j] + s->Y[0] * (i == j);
    F[i][4] = (s->U[4] + s->Y[0]) * s->Y[1 + i];
  }
}

// FluxInviscid_fwd (newtonian_state.h:433-441)
static inline void o_flux_inviscid_fwd(const OState *s, const OState *ds, double dF[3][5]) {
  for (int i = 0; i < 3; i++) {
    dF[i][0] = ds->U[1 + i];
    for (int j = 0; j < 3; j++) dF[i][1 + j] = ds->U[1 + i] * s->Y[1 + j] + s->U[1 + i] * ds->Y[1 + j] + ds->Y[0] * (i == j);
    dF[i][4] = (ds->U[4] + ds->Y[0]) * s->Y[1 + i] + (s->U[4] + s->Y[0]) * ds->Y[1 + i];
  }
}

// FluxInviscidDotNormal (newtonian_state.h:443-452)
static inline void o_flux_inviscid_dot_normal(const OState *s, const double n[3], double Fn[5]) {
  double F[3][5];
  o_flux_inviscid(s, F);
  for (int j = 0; j < 5; j++) Fn[j] = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 5; j++) Fn[j] += F[i][j] * n[i];
}

// FluxInviscidDotNormal_fwd (newtonian_state.h:454-464)
static inline void o_flux_inviscid_dot_normal_fwd(const OState *s, const OState *ds, const double n[3], double dFn[5]) {
  double dF[3][5];
  o_flux_inviscid_fwd(s, ds, dF);
  for (int j = 0; j < 5; j++) dFn[j] = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 5; j++) dFn[j] += dF[i][j] * n[i];
}

// FluxInviscidStrong (newtonian_state.h:466-475)
static inline void o_flux_inviscid_strong(const OState *s, const OState ds[3], double sc[5]) {
  for (int j = 0; j < 5; j++) sc[j] = 0;
  for (int i = 0; i < 3; i++) {
    double dF[3][5];
    o_flux_inviscid_fwd(s, &ds[i], dF);
    for (int j = 0; j < 5; j++) sc[j] += dF[i][j];
  }
}

// FluxTotal (newtonian_state.h:477-484): Flux[var][dir]
static inline void o_flux_total(const double Finv[3][5], const double st[3][3], const double Fe[3], double Flux[5][3]) {
  for (int j = 0; j < 3; j++) {
    Flux[0][j] = Finv[j][0];
    for (int k = 0; k < 3; k++) Flux[k + 1][j] = Finv[j][1 + k] - st[k][j];
    Flux[4][j] = Finv[j][4] + Fe[j];
  }
}

// FluxTotal_Boundary (newtonian_state.h:486-496)
static inline void o_flux_total_boundary(const double Finv[3][5], const double st[3][3], const double Fe[3], const double n[3], double Flux[5]) {
  for (int j = 0; j < 5; j++) Flux[j] = 0.;
  for (int j = 0; j < 3; j++) {
    Flux[0] += Finv[j][0] * n[j];
    for (int k = 0; k < 3; k++) Flux[k + 1] += (Finv[j][1 + k] - st[k][j]) * n[j];
    Flux[4] += (Finv[j][4] + Fe[j]) * n[j];
  }
}

// KMStrainRate_State (newtonian_state.h:511-538): velocity gradient grad_u[i][j] = d u_i / d x_j
static inline void o_km_strain_rate(const OState gs[3], double sr[6]) {
  double gu[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) gu[i][j] = gs[j].Y[1 + i];
  const double w = 1 / sqrt(2.);
  sr[0]          = gu[0][0];
  sr[1]          = gu[1][1];
  sr[2]          = gu[2][2];
  sr[3]          = w * (gu[1][2] + gu[2][1]);
  sr[4]          = w * (gu[0][2] + gu[2][0]);
  sr[5]          = w * (gu[0][1] + gu[1][0]);
}

// NewtonianStress (newtonian_state.h:553-558)
static inline void o_newtonian_stress(OGas g, const double sr[6], double st[6]) {
  double div_u = sr[0] + sr[1] + sr[2];
  for (int i = 0; i < 6; i++) st[i] = g.mu * (2 * sr[i] + g.lambda * div_u * (i < 3));
}

// KMUnpack (utils.h:239-247)
static inline void o_km_unpack(const double v[6], double A[3][3]) {
  const double w = 1 / sqrt(2.);
  A[0][0]        = v[0];
  A[1][1]        = v[1];
  A[2][2]        = v[2];
  A[2][1] = A[1][2] = w * v[3];
  A[2][0] = A[0][2] = w * v[4];
  A[1][0] = A[0][1] = w * v[5];
}

// ViscousEnergyFlux (newtonian_state.h:560-565)
static inline void o_viscous_energy_flux(OGas g, const double Y[5], const OState gs[3], const double st[3][3], double Fe[3]) {
  for (int i = 0; i < 3; i++) Fe[i] = -Y[1] * st[0][i] - Y[2] * st[1][i] - Y[3] * st[2][i] - g.k * gs[i].Y[4];
}

// ViscousEnergyFlux_fwd (newtonian_state.h:567-573)
static inline void o_viscous_energy_flux_fwd(OGas g, const double Y[5], const double dY[5], const OState gds[3], const double st[3][3],
                                             const double dst[3][3], double dFe[3]) {
  for (int i = 0; i < 3; i++)
    dFe[i] = -Y[1] * dst[0][i] - dY[1] * st[0][i] - Y[2] * dst[1][i] - dY[2] * st[1][i] - Y[3] * dst[2][i] - dY[3] * st[2][i] - g.k * gds[i].Y[4];
}

// StatePhysicalGradientFromReference (newtonian_state.h:581-593); gref[d][c] = d q_c / d X_d
static inline void o_phys_gradient(OGas g, const OState *s, int var, const double gref[3][5], const double dXdx[3][3], OState gs[3]) {
  double dq[5][3];
  for (int c = 0; c < 5; c++)
    for (int j = 0; j < 3; j++) {
      double acc = 0;
      for (int k = 0; k < 3; k++) acc += gref[k][c] * dXdx[k][j];
      dq[c][j] = acc;
    }
  for (int j = 0; j < 3; j++) {
    double dqi[5];
    for (int c = 0; c < 5; c++) dqi[c] = dq[c][j];
    gs[j] = o_state_from_q_fwd(g, s, dqi, var);
  }
}

typedef struct {
  double Ctau_t, Ctau_v, Ctau_C, Ctau_M, Ctau_E;  // TauDiagCoefficients (stabilization_types.h:32-38)
} OTauCoef;

// Tau_diagPrim (stabilization.h:67-107)
static inline void o_tau_diag_prim(OTauCoef c, OGas g, const OState *s, const double dXdx[3][3], double dt, double Tau[3]) {
  const double rho = s->U[0];
  double       gij[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double acc = 0;
      for (int k = 0; k < 3; k++) acc += dXdx[k][i] * dXdx[k][j];
      gij[i][j] = acc;
    }
  double dts = c.Ctau_t / dt;
  double gu[3];
  for (int i = 0; i < 3; i++) gu[i] = 0 + o_dotn(gij[i], &s->Y[1], 3);
  double vt   = o_dot3(&s->Y[1], gu);
  double tau  = o_sq(rho) * (4. * o_sq(dts) + vt) + c.Ctau_v * o_sq(g.mu) * o_dotn(&gij[0][0], &gij[0][0], 9);
  double fact = sqrt(tau);
  Tau[0]      = c.Ctau_C * fact / (rho * (gij[0][0] + gij[1][1] + gij[2][2])) * 0.125;
  Tau[1]      = c.Ctau_M / fact;
  Tau[2]      = c.Ctau_E / (fact * g.cv);
}

// StabilizationMatrix with dYFromTau (stabilization.h:14-39); stab[var][dir]
static inline void o_stab_matrix(OGas g, const OState *s, const double Tau[3], const double r[5], double stab[5][3]) {
  double dY[5] = {Tau[0] * r[0], Tau[1] * r[1], Tau[1] * r[2], Tau[1] * r[3], Tau[2] * r[4]};
  OState ds    = o_state_from_y_fwd(g, s, dY);
  double dF[3][5];
  o_flux_inviscid_fwd(s, &ds, dF);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 5; j++) stab[j][i] = 0 + dF[i][j];
}

// Stabilization (stabilization.h:41-58)
static inline void o_stabilization(int type, OGas g, const OState *s, const double Tau[3], const OState ds[3], const double Udot[5],
                                   const double bf[5], const double divF[5], double stab[5][3]) {
  double r[5] = {0};
  if (type == OSTAB_SU || type == OSTAB_SUPG) o_flux_inviscid_strong(s, ds, r);
  if (type == OSTAB_SUPG)
    for (int j = 0; j < 5; j++) r[j] += Udot[j] - bf[j] + divF[j];
  o_stab_matrix(g, s, Tau, r, stab);
}

// LinearRampCoefficient (utils.h:287-295)
static inline double o_linear_ramp(double amplitude, double length, double start, double x) {
  if (x < start) return amplitude;
  else if (x < start + length) return amplitude * ((x - start) * (-1 / length) + 1);
  else return 0;
}

// InternalDampingLayer_Residual (newtonian_integrands.h:13-21): adds U(StateFromY_fwd(sigma * dampY)) into res
static inline void o_idl_residual(OGas g, const OState *s, double sigma, double dampY[5], double res[5]) {
  for (int i = 0; i < 5; i++) dampY[i] *= sigma;
  OState ds = o_state_from_y_fwd(g, s, dampY);
  for (int i = 0; i < 5; i++) res[i] += ds.U[i];
}

// ---------------------------------------------------------------------------------------------------------------------
// Riemann fluxes for the freestream boundary (qfunctions/riemann_solver.h).  Left = interior state, right = exterior
// state, normal = unit outward normal.
// ---------------------------------------------------------------------------------------------------------------------

// TotalSpecificEnthalpy (newtonian_state.h:52-56)
static inline double o_total_enthalpy(OGas g, const OState *s) {
  const double e_kinetic  = 0.5 * o_dot3(&s->Y[1], &s->Y[1]);
  const double e_internal = g.cv * s->Y[4];
  return e_internal + e_kinetic + s->Y[0] / s->U[0];
}

// TotalSpecificEnthalpy_fwd (newtonian_state.h:58-62)
static inline double o_total_enthalpy_fwd(OGas g, const OState *s, const OState *ds) {
  const double de_kinetic  = o_dot3(&ds->Y[1], &s->Y[1]);
  const double de_internal = g.cv * ds->Y[4];
  return de_internal + de_kinetic + ds->Y[0] / s->U[0] - s->Y[0] / o_sq(s->U[0]) * ds->U[0];
}

// RoeSetup / RoeSetup_fwd (riemann_solver.h:27-40): weights {left, right}
static inline void o_roe_setup(double rho_l, double rho_r, double w[2]) {
  const double sl = sqrt(rho_l), sr = sqrt(rho_r);
  w[0] = sl / (sl + sr);
  w[1] = sr / (sl + sr);
}
static inline void o_roe_setup_fwd(double rho_l, double rho_r, double drho_l, double drho_r, double dw[2]) {
  const double sl = sqrt(rho_l), sr = sqrt(rho_r);
  const double ss = o_sq(sl + sr);
  const double r_right = (sl / (2 * sr * ss)) * drho_r - (sr / (2 * sl * ss)) * drho_l;
  const double r_left  = (sr / (2 * sl * ss)) * drho_l - (sl / (2 * sr * ss)) * drho_r;
  dw[0] = r_left;
  dw[1] = r_right;
}

// RoeAverage / RoeAverage_fwd (riemann_solver.h:42-47)
static inline double o_roe_avg(const double w[2], double ql, double qr) { return w[0] * ql + w[1] * qr; }
static inline double o_roe_avg_fwd(const double w[2], const double dw[2], double ql, double qr, double dql, double dqr) {
  return qr * dw[1] + ql * dw[0] + w[1] * dqr + w[0] * dql;
}

// ComputeHLLSpeeds_Roe (riemann_solver.h:87-106)
static inline void o_hll_speeds_roe(OGas g, const OState *l, double ul, const OState *r, double ur, double *sl, double *sr) {
  const double gamma = o_gamma(g);
  double       w[2];
  o_roe_setup(l->U[0], r->U[0], w);
  const double u_roe = o_roe_avg(w, ul, ur);
  const double H_l = o_total_enthalpy(g, l), H_r = o_total_enthalpy(g, r);
  const double H_roe = o_roe_avg(w, H_l, H_r);
  const double a_roe = sqrt((gamma - 1) * (H_roe - 0.5 * o_sq(u_roe)));
  *sl                = u_roe - a_roe;
  *sr                = u_roe + a_roe;
}

// ComputeHLLSpeeds_Roe_fwd (riemann_solver.h:108-137)
static inline void o_hll_speeds_roe_fwd(OGas g, const OState *l, const OState *dl, double ul, double dul, const OState *r, const OState *dr,
                                        double ur, double dur, double *sl, double *dsl, double *sr, double *dsr) {
  const double gamma = o_gamma(g);
  double       w[2], dw[2];
  o_roe_setup(l->U[0], r->U[0], w);
  o_roe_setup_fwd(l->U[0], r->U[0], dl->U[0], dr->U[0], dw);
  const double u_roe  = o_roe_avg(w, ul, ur);
  const double du_roe = o_roe_avg_fwd(w, dw, ul, ur, dul, dur);
  const double H_l = o_total_enthalpy(g, l), H_r = o_total_enthalpy(g, r);
  const double dH_l = o_total_enthalpy_fwd(g, l, dl), dH_r = o_total_enthalpy_fwd(g, r, dr);
  const double H_roe  = o_roe_avg(w, H_l, H_r);
  const double dH_roe = o_roe_avg_fwd(w, dw, H_l, H_r, dH_l, dH_r);
  const double a_roe  = sqrt((gamma - 1) * (H_roe - 0.5 * o_sq(u_roe)));
  double       da_roe = 0.5 * sqrt((gamma - 1) / (H_roe - 0.5 * o_sq(u_roe))) * dH_roe;
  da_roe -= 0.5 * sqrt(gamma - 1) * u_roe / sqrt(H_roe - 0.5 * o_sq(u_roe)) * du_roe;
  *sl  = u_roe - a_roe;
  *dsl = du_roe - da_roe;
  *sr  = u_roe + a_roe;
  *dsr = du_roe + da_roe;
}

// Flux_HLL (riemann_solver.h:49-66)
static inline void o_flux_hll(const OState *l, const OState *r, const double Fl[5], const double Fr[5], double s_l, double s_r, double F[5]) {
  for (int i = 0; i < 5; i++) F[i] = (s_r * Fl[i] - s_l * Fr[i] + s_l * s_r * (r->U[i] - l->U[i])) / (s_r - s_l);
}

// Flux_HLL_fwd (riemann_solver.h:68-85)
static inline void o_flux_hll_fwd(const OState *l, const OState *r, const OState *dl, const OState *dr, const double Fl[5], const double Fr[5],
                                  const double dFl[5], const double dFr[5], double S_l, double S_r, double dS_l, double dS_r, double dF[5]) {
  for (int i = 0; i < 5; i++) {
    const double S_diff = S_r - S_l;
    dF[i]               = 0.;
    dF[i] += (S_l * (-Fl[i] + Fr[i] + S_l * l->U[i] - S_l * r->U[i]) / o_sq(S_diff)) * dS_r;
    dF[i] += (S_r * (Fl[i] - Fr[i] - S_r * l->U[i] + S_r * r->U[i]) / o_sq(S_diff)) * dS_l;
    dF[i] += (S_r * dFl[i] - S_l * dFr[i] + S_r * S_l * (dr->U[i] - dl->U[i])) / S_diff;
  }
}

// RiemannFlux_HLL (riemann_solver.h:151-169)
static inline void o_riemann_hll(OGas g, const OState *l, const OState *r, const double n[3], double F[5]) {
  double Fl[5], Fr[5];
  o_flux_inviscid_dot_normal(l, n, Fl);
  o_flux_inviscid_dot_normal(r, n, Fr);
  const double ul = o_dot3(&l->Y[1], n), ur = o_dot3(&r->Y[1], n);
  double       s_l, s_r;
  o_hll_speeds_roe(g, l, ul, r, ur, &s_l, &s_r);
  if (0 <= s_l) memcpy(F, Fl, sizeof(Fl));
  else if (s_r <= 0) memcpy(F, Fr, sizeof(Fr));
  else o_flux_hll(l, r, Fl, Fr, s_l, s_r, F);
}

// RiemannFlux_HLL_fwd (riemann_solver.h:184-205)
static inline void o_riemann_hll_fwd(OGas g, const OState *l, const OState *dl, const OState *r, const OState *dr, const double n[3], double dF[5]) {
  double Fl[5], Fr[5], dFl[5], dFr[5];
  o_flux_inviscid_dot_normal(l, n, Fl);
  o_flux_inviscid_dot_normal(r, n, Fr);
  o_flux_inviscid_dot_normal_fwd(l, dl, n, dFl);
  o_flux_inviscid_dot_normal_fwd(r, dr, n, dFr);
  const double ul = o_dot3(&l->Y[1], n), ur = o_dot3(&r->Y[1], n);
  const double dul = o_dot3(&dl->Y[1], n), dur = o_dot3(&dr->Y[1], n);
  double       s_l, ds_l, s_r, ds_r;
  o_hll_speeds_roe_fwd(g, l, dl, ul, dul, r, dr, ur, dur, &s_l, &ds_l, &s_r, &ds_r);
  if (0 <= s_l) memcpy(dF, dFl, sizeof(dFl));
  else if (s_r <= 0) memcpy(dF, dFr, sizeof(dFr));
  else o_flux_hll_fwd(l, r, dl, dr, Fl, Fr, dFl, dFr, s_l, s_r, ds_l, ds_r, dF);
}

// RiemannFlux_HLLC_Star (riemann_solver.h:207-224): StateConservativeAXPBYPCZ(1, F_side, s_side * fact, star, -s_side, U_side)
static inline void o_hllc_star(const OState *side, const double F_side[5], const double n[3], double u_side, double s_side, double s_star, double F[5]) {
  const double fact  = side->U[0] * (s_side - u_side) / (s_side - s_star);
  const double denom = side->U[0] * (s_side - u_side);
  const double star[5] = {1.0, side->Y[1] + (s_star - u_side) * n[0], side->Y[2] + (s_star - u_side) * n[1], side->Y[3] + (s_star - u_side) * n[2],
                          side->U[4] / side->U[0] + (s_star - u_side) * (s_star + side->Y[0] / denom)};
  const double b = s_side * fact, c = -s_side;
  for (int i = 0; i < 5; i++) F[i] = 1 * F_side[i] + b * star[i] + c * side->U[i];
}

// RiemannFlux_HLLC_Star_fwd (riemann_solver.h:226-256): StateConservativeMult over {dF_side, star, dstar, U, dU}
static inline void o_hllc_star_fwd(const OState *side, const OState *dside, const double F_side[5], const double dF_side[5], const double n[3],
                                   double u_side, double du_side, double s_side, double ds_side, double s_star, double ds_star, double dF[5]) {
  const double fact  = side->U[0] * (s_side - u_side) / (s_side - s_star);
  const double dfact = (side->U[0] * (ds_side - du_side) + dside->U[0] * (s_side - u_side)) / (s_side - s_star) -
                       fact / (s_side - s_star) * (ds_side - ds_star);
  const double denom  = side->U[0] * (s_side - u_side);
  const double ddenom = side->U[0] * (ds_side - du_side) + dside->U[0] * (s_side - u_side);
  const double star[5] = {1.0, side->Y[1] + (s_star - u_side) * n[0], side->Y[2] + (s_star - u_side) * n[1], side->Y[3] + (s_star - u_side) * n[2],
                          side->U[4] / side->U[0] + (s_star - u_side) * (s_star + side->Y[0] / denom)};
  const double dstar[5] = {0., dside->Y[1] + (ds_star - du_side) * n[0], dside->Y[2] + (ds_star - du_side) * n[1],
                           dside->Y[3] + (ds_star - du_side) * n[2],
                           dside->U[4] / side->U[0] - side->U[4] / o_sq(side->U[0]) * dside->U[0] + (ds_star - du_side) * (s_star + side->Y[0] / denom) +
                               (s_star - u_side) * (ds_star + dside->Y[0] / denom - side->Y[0] / o_sq(denom) * ddenom)};
  const double  a[5] = {1, ds_side * fact + s_side * dfact, s_side * fact, -ds_side, -s_side};
  const double *X[5] = {dF_side, star, dstar, side->U, dside->U};
  for (int i = 0; i < 5; i++) dF[i] = 0;
  for (int k = 0; k < 5; k++)
    for (int i = 0; i < 5; i++) dF[i] += a[k] * X[k][i];
}

// RiemannFlux_HLLC (riemann_solver.h:258-284)
static inline void o_riemann_hllc(OGas g, const OState *l, const OState *r, const double n[3], double F[5]) {
  double Fl[5], Fr[5];
  o_flux_inviscid_dot_normal(l, n, Fl);
  o_flux_inviscid_dot_normal(r, n, Fr);
  const double ul = o_dot3(&l->Y[1], n), ur = o_dot3(&r->Y[1], n);
  double       s_l, s_r;
  o_hll_speeds_roe(g, l, ul, r, ur, &s_l, &s_r);
  const double rhou_l = l->U[0] * ul, rhou_r = r->U[0] * ur;
  const double numer  = r->Y[0] - l->Y[0] + rhou_l * (s_l - ul) - rhou_r * (s_r - ur);
  const double denom  = l->U[0] * (s_l - ul) - r->U[0] * (s_r - ur);
  const double s_star = numer / denom;
  if (0 <= s_l) memcpy(F, Fl, sizeof(Fl));
  else if (0 <= s_star) o_hllc_star(l, Fl, n, ul, s_l, s_star, F);
  else if (0 <= s_r) o_hllc_star(r, Fr, n, ur, s_r, s_star, F);
  else memcpy(F, Fr, sizeof(Fr));
}

// RiemannFlux_HLLC_fwd (riemann_solver.h:286-338)
static inline void o_riemann_hllc_fwd(OGas g, const OState *l, const OState *dl, const OState *r, const OState *dr, const double n[3], double dF[5]) {
  double Fl[5], Fr[5], dFl[5], dFr[5];
  o_flux_inviscid_dot_normal(l, n, Fl);
  o_flux_inviscid_dot_normal(r, n, Fr);
  o_flux_inviscid_dot_normal_fwd(l, dl, n, dFl);
  o_flux_inviscid_dot_normal_fwd(r, dr, n, dFr);
  const double ul = o_dot3(&l->Y[1], n), ur = o_dot3(&r->Y[1], n);
  const double dul = o_dot3(&dl->Y[1], n), dur = o_dot3(&dr->Y[1], n);
  double       s_l, ds_l, s_r, ds_r;
  o_hll_speeds_roe_fwd(g, l, dl, ul, dul, r, dr, ur, dur, &s_l, &ds_l, &s_r, &ds_r);
  const double rhou_l = l->U[0] * ul, drhou_l = l->U[0] * dul + dl->U[0] * ul;
  const double rhou_r = r->U[0] * ur, drhou_r = r->U[0] * dur + dr->U[0] * ur;
  const double numer  = r->Y[0] - l->Y[0] + rhou_l * (s_l - ul) - rhou_r * (s_r - ur);
  const double dnumer = dr->Y[0] - dl->Y[0] + rhou_l * (ds_l - dul) + drhou_l * (s_l - ul) - rhou_r * (ds_r - dur) - drhou_r * (s_r - ur);
  const double denom  = l->U[0] * (s_l - ul) - r->U[0] * (s_r - ur);
  const double ddenom = l->U[0] * (ds_l - dul) + dl->U[0] * (s_l - ul) - r->U[0] * (ds_r - dur) - dr->U[0] * (s_r - ur);
  const double s_star  = numer / denom;
  const double ds_star = dnumer / denom - numer / o_sq(denom) * ddenom;
  if (0 <= s_l) memcpy(dF, dFl, sizeof(dFl));
  else if (0 <= s_star) o_hllc_star_fwd(l, dl, Fl, dFl, n, ul, dul, s_l, ds_l, s_star, ds_star, dF);
  else if (0 <= s_r) o_hllc_star_fwd(r, dr, Fr, dFr, n, ur, dur, s_r, ds_r, s_star, ds_star, dF);
  else memcpy(dF, dFr, sizeof(dFr));
}
