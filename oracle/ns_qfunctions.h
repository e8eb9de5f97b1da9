// CPU ORACLE -- test infrastructure only (see ns_pointwise.h).
//
// The Newtonian QFunction bodies at one quadrature point: IFunction, IJacobian and the explicit RHS.  honee_oracle.c
// runs them inside its element loops; opcount.cpp compiles the same text over an operation-counting scalar to count
// the arithmetic of one point exactly.
#pragma once

#include "honee_oracle.h"

// ----------------------------------------------------------------------------------------------------------------------
// IFunction_Newtonian (qfunctions/newtonian.h:205-257) + IFunction_Newtonian_Integrand (newtonian_integrands.h:87-115)
// ----------------------------------------------------------------------------------------------------------------------
static void qf_ifunction(const ONewt *c, const double q[5], const double gref[3][5], const double qdot[5], const double wdetJ,
                         const double dXdx[3][3], const double x[3], double v[5], double gv[3][5], double *jd) {
  const OGas   g   = c->gas;
  const int    var = c->state_var;
  const OState s   = o_state_from_q(g, q, var);
  const OState sd  = o_state_from_q_fwd(g, &s, qdot, var);
  OState       gs[3];
  o_phys_gradient(g, &s, var, gref, dXdx, gs);
  const double divF[5] = {0};

  double v_i[5] = {0}, gv_i[5][3] = {{0}}, sr[6], Tau[3], sigma = 0;
  {
    double kmst[6], st[3][3], Fe[3], Finv[3][5], Ft[5][3];
    o_km_strain_rate(gs, sr);
    o_newtonian_stress(g, sr, kmst);
    o_km_unpack(kmst, st);
    o_viscous_energy_flux(g, s.Y, gs, st, Fe);
    o_flux_inviscid(&s, Finv);
    o_flux_total(Finv, st, Fe, Ft);
    for (int j = 0; j < 5; j++)
      for (int k = 0; k < 3; k++) gv_i[j][k] += -1 * Ft[j][k];
    const double bf[5] = {0, s.U[0] * c->g[0], s.U[0] * c->g[1], s.U[0] * c->g[2], o_dot3(&s.U[1], c->g)};
    double       Udot[5];
    memcpy(Udot, sd.U, sizeof(Udot));
    for (int j = 0; j < 5; j++) v_i[j] += Udot[j] - bf[j];
    double stab[5][3];
    o_tau_diag_prim(c->tau, g, &s, dXdx, c->dt, Tau);
    o_stabilization(c->stabilization, g, &s, Tau, gs, Udot, bf, divF, stab);
    for (int j = 0; j < 5; j++)
      for (int k = 0; k < 3; k++) gv_i[j][k] += 1 * stab[j][k];
  }
  if (c->idl_enable) {  // InternalDampingLayer_IFunction_Integrand (newtonian_integrands.h:40-48)
    sigma             = o_linear_ramp(c->idl_amplitude, c->idl_length, c->idl_start, x[0]);
    double damp[5]    = {s.Y[0] - c->idl_pressure, 0, 0, 0, 0}, res[5] = {0};
    o_idl_residual(g, &s, sigma, damp, res);
    for (int j = 0; j < 5; j++) v_i[j] += 1 * res[j];
  }
  for (int j = 0; j < 5; j++) v[j] = wdetJ * v_i[j];
  for (int j = 0; j < 5; j++)
    for (int k = 0; k < 3; k++) gv[k][j] = wdetJ * (gv_i[j][0] * dXdx[k][0] + gv_i[j][1] * dXdx[k][1] + gv_i[j][2] * dXdx[k][2]);
  for (int j = 0; j < 5; j++) jd[j] = q[j];
  for (int j = 0; j < 6; j++) jd[5 + j] = sr[j];
  for (int j = 0; j < 3; j++) jd[11 + j] = Tau[j];
  if (c->idl_enable) jd[14] = sigma;
}

static inline int njac(const ONewt *c) { return 14 + (c->idl_enable ? 1 : 0); }

// ----------------------------------------------------------------------------------------------------------------------
// IJacobian_Newtonian (newtonian.h:272-311) + IJacobian_Newtonian_Integrand (newtonian_integrands.h:134-166)
// ----------------------------------------------------------------------------------------------------------------------
static void qf_ijacobian(const ONewt *c, const double dq[5], const double gref[3][5], const double wdetJ, const double dXdx[3][3],
                         const double *jd, double v[5], double gv[3][5]) {
  const OGas   g   = c->gas;
  const int    var = c->state_var;
  double       qi[5], sr[6], Tau[3];
  for (int j = 0; j < 5; j++) qi[j] = jd[j];
  for (int j = 0; j < 6; j++) sr[j] = jd[5 + j];
  for (int j = 0; j < 3; j++) Tau[j] = jd[11 + j];
  const OState s  = o_state_from_q(g, qi, var);
  const OState ds = o_state_from_q_fwd(g, &s, dq, var);
  OState       gds[3];
  o_phys_gradient(g, &s, var, gref, dXdx, gds);

  double v_i[5] = {0}, gv_i[5][3] = {{0}};
  {
    double dsr[6], dkmst[6], kmst[6], st[3][3], dst[3][3], dFe[3], dFinv[3][5], dFt[5][3];
    o_km_strain_rate(gds, dsr);
    o_newtonian_stress(g, dsr, dkmst);
    o_newtonian_stress(g, sr, kmst);
    o_km_unpack(dkmst, dst);
    o_km_unpack(kmst, st);
    o_viscous_energy_flux_fwd(g, s.Y, ds.Y, gds, st, dst, dFe);
    o_flux_inviscid_fwd(&s, &ds, dFinv);
    o_flux_total(dFinv, dst, dFe, dFt);
    for (int j = 0; j < 5; j++)
      for (int k = 0; k < 3; k++) gv_i[j][k] += -1 * dFt[j][k];
    const double dbf[5] = {0, ds.U[0] * c->g[0], ds.U[0] * c->g[1], ds.U[0] * c->g[2], o_dot3(&ds.U[1], c->g)};
    double       dUdot[5];
    for (int j = 0; j < 5; j++) {
      dUdot[j] = c->ijacobian_time_shift * ds.U[j];
      v_i[j]   = dUdot[j] - dbf[j];
    }
    double       dstab[5][3];
    const double zero[5] = {0};
    o_stabilization(c->stabilization, g, &s, Tau, gds, dUdot, dbf, zero, dstab);
    for (int j = 0; j < 5; j++)
      for (int k = 0; k < 3; k++) gv_i[j][k] += 1 * dstab[j][k];
  }
  if (c->idl_enable) {  // InternalDampingLayer_IJacobian_Integrand (newtonian_integrands.h:61-65)
    const double sigma   = jd[14];
    double       damp[5] = {ds.Y[0], 0, 0, 0, 0}, res[5] = {0};
    o_idl_residual(g, &s, sigma, damp, res);
    for (int j = 0; j < 5; j++) v_i[j] += 1 * res[j];
  }
  for (int j = 0; j < 5; j++) v[j] = wdetJ * v_i[j];
  for (int j = 0; j < 5; j++)
    for (int k = 0; k < 3; k++) gv[k][j] = wdetJ * (gv_i[j][0] * dXdx[k][0] + gv_i[j][1] * dXdx[k][1] + gv_i[j][2] * dXdx[k][2]);
}

// ----------------------------------------------------------------------------------------------------------------------
// RHSFunction_Newtonian (newtonian.h:133-202), conservative variables
// ----------------------------------------------------------------------------------------------------------------------
static void qf_rhs(const ONewt *c, const double U[5], const double gref[3][5], const double wdetJ, const double dXdx[3][3], const double x[3],
                   double v[5], double gv[3][5]) {
  const OGas   g = c->gas;
  const OState s = o_state_from_q(g, U, OVAR_CONS);
  OState       gs[3];
  o_phys_gradient(g, &s, OVAR_CONS, gref, dXdx, gs);
  double sr[6], kmst[6], st[3][3], Fe[3], Finv[3][5], F[5][3];
  o_km_strain_rate(gs, sr);
  o_newtonian_stress(g, sr, kmst);
  o_km_unpack(kmst, st);
  o_viscous_energy_flux(g, s.Y, gs, st, Fe);
  o_flux_inviscid(&s, Finv);
  o_flux_total(Finv, st, Fe, F);
  for (int j = 0; j < 5; j++)
    for (int k = 0; k < 3; k++) gv[k][j] = wdetJ * (dXdx[k][0] * F[j][0] + dXdx[k][1] * F[j][1] + dXdx[k][2] * F[j][2]);
  const double bf[5] = {0, s.U[0] * c->g[0], s.U[0] * c->g[1], s.U[0] * c->g[2], o_dot3(&s.U[1], c->g)};
  for (int j = 0; j < 5; j++) v[j] = wdetJ * bf[j];
  if (c->idl_enable) {
    const double sigma   = o_linear_ramp(c->idl_amplitude, c->idl_length, c->idl_start, x[0]);
    double       damp[5] = {s.Y[0] - c->idl_pressure, 0, 0, 0, 0}, res[5] = {0};
    o_idl_residual(g, &s, sigma, damp, res);
    for (int j = 0; j < 5; j++) v[j] -= wdetJ * res[j];
  }
  const double divF[5] = {0}, Udot[5] = {0};
  double       Tau[3], stab[5][3];
  o_tau_diag_prim(c->tau, g, &s, dXdx, c->dt, Tau);
  o_stabilization(c->stabilization, g, &s, Tau, gs, Udot, bf, divF, stab);
  for (int j = 0; j < 5; j++)
    for (int k = 0; k < 3; k++) gv[k][j] -= wdetJ * (stab[j][0] * dXdx[k][0] + stab[j][1] * dXdx[k][1] + stab[j][2] * dXdx[k][2]);
}

