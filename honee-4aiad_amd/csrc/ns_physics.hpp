// Pointwise Newtonian ideal-gas physics for the fused MI355X kernels.
//
// Device functions templated on the arithmetic type T (double for residuals and the FP64 Jacobian, float for the
// FP32 Jacobian variant).  They compute exactly the quantities of HONEE's QFunction helpers (file:line cited per
// function) with the same operation order, so FP64 results agree with the reference arithmetic to rounding.
// A state keeps both the conservative U = (rho, rho u, E) and primitive Y = (p, u, T) vectors, as `State` does
// (qfunctions/newtonian_state.h:13-22).
#pragma once

#include <hip/hip_runtime.h>

#include "honee_mi355x.h"

namespace ns {

enum : int { VAR_CONS = 0, VAR_PRIM = 1, VAR_ENTROPY = 2 };
enum : int { STAB_NONE = 0, STAB_SU = 1, STAB_SUPG = 2 };

template <typename T>
struct Gas {
  T lambda, mu, k, cv, cp, gamma, Rgas, rcv;
  __host__ __device__ static Gas from(const CeedMiGas &g) {
    Gas r;
    r.lambda = (T)g.lambda;
    r.mu     = (T)g.mu;
    r.k      = (T)g.k;
    r.cv     = (T)g.cv;
    r.cp     = (T)g.cp;
    r.gamma  = (T)(g.cp / g.cv);  // HeatCapacityRatio (newtonian_state.h:42)
    r.Rgas   = (T)(g.cp - g.cv);  // GasConstant (newtonian_state.h:44)
    r.rcv    = T(1) / r.cv;
    return r;
  }
};

// A state with the per-point quantities its tangent maps reuse.  The reference recomputes them in every _fwd call;
// here the energies are computed once with the same operations, and the divisions by the state's density or by
// Rgas T^2 become multiplications by a reciprocal (a rounding-level difference).
template <typename T>
struct St {
  T U[5];
  T Y[5];
  T rRT2, etot;       // 1 / (Rgas T^2) and cv T + |u|^2 / 2 (StateConservativeFromPrimitive_fwd)
  T rrho, et, ei_c;   // 1 / rho, E / rho, E / rho - |u|^2 / 2 (StatePrimitiveFromConservative_fwd)
};

template <typename T>
__device__ __forceinline__ T dot3(const T *a, const T *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <typename T>
__device__ __forceinline__ T sq(T x) {
  return x * x;
}

// StatePrimitiveFromConservative (newtonian_state.h:64-73)
template <typename T>
__device__ __forceinline__ void prim_from_cons(const Gas<T> &g, const T U[5], T Y[5]) {
  for (int i = 0; i < 3; i++) Y[1 + i] = U[1 + i] / U[0];
  T ek = T(.5) * dot3(&Y[1], &Y[1]);
  T et = U[4] / U[0];
  T ei = et - ek;
  Y[4] = ei / g.cv;
  Y[0] = (g.gamma - 1) * U[0] * ei;
}

// StatePrimitiveFromConservative_fwd (newtonian_state.h:75-89)
template <typename T>
__device__ __forceinline__ void prim_from_cons_fwd(const Gas<T> &g, const St<T> &s, const T dU[5], T dY[5]) {
  for (int i = 0; i < 3; i++) dY[1 + i] = (dU[1 + i] - s.Y[1 + i] * dU[0]) * s.rrho;
  T dek = dot3(&dY[1], &s.Y[1]);
  T det = (dU[4] - s.et * dU[0]) * s.rrho;
  T dei = det - dek;
  dY[4] = dei * g.rcv;
  dY[0] = (g.gamma - 1) * (dU[0] * s.ei_c + s.U[0] * dei);
}

// StateConservativeFromPrimitive (newtonian_state.h:148-157)
template <typename T>
__device__ __forceinline__ void cons_from_prim(const Gas<T> &g, const T Y[5], T U[5]) {
  U[0] = Y[0] / (g.Rgas * Y[4]);
  for (int i = 0; i < 3; i++) U[1 + i] = U[0] * Y[1 + i];
  T ei = g.cv * Y[4];
  T ek = T(.5) * dot3(&Y[1], &Y[1]);
  U[4] = U[0] * (ei + ek);
}

// StateConservativeFromPrimitive_fwd (newtonian_state.h:159-173)
template <typename T>
__device__ __forceinline__ void cons_from_prim_fwd(const Gas<T> &g, const St<T> &s, const T dY[5], T dU[5]) {
  dU[0] = (dY[0] * s.Y[4] - s.Y[0] * dY[4]) * s.rRT2;
  for (int i = 0; i < 3; i++) dU[1 + i] = dU[0] * s.Y[1 + i] + s.U[0] * dY[1 + i];
  T dek = dot3(&dY[1], &s.Y[1]);
  T dei = g.cv * dY[4];
  dU[4] = dU[0] * s.etot + s.U[0] * (dei + dek);
}

// StateEntropyFromPrimitive (newtonian_state.h:91-103)
template <typename T>
__device__ __forceinline__ void entropy_from_prim(const Gas<T> &g, const T Y[5], T V[5]) {
  const T rho     = Y[0] / (g.Rgas * Y[4]);
  const T entropy = log(Y[0]) - g.gamma * log(rho);
  const T rdp     = rho / Y[0];
  const T ek      = T(0.5) * dot3(&Y[1], &Y[1]);
  V[0]            = (g.gamma - entropy) / (g.gamma - 1) - rdp * ek;
  for (int i = 0; i < 3; i++) V[1 + i] = rdp * Y[1 + i];
  V[4] = -rdp;
}

// StatePrimitiveFromEntropy (newtonian_state.h:123-133)
template <typename T>
__device__ __forceinline__ void prim_from_entropy(const Gas<T> &g, const T V[5], T Y[5]) {
  for (int i = 0; i < 3; i++) Y[1 + i] = -V[1 + i] / V[4];
  Y[4]            = T(-1) / (g.Rgas * V[4]);
  const T ek      = T(0.5) * dot3(&Y[1], &Y[1]);
  const T entropy = g.gamma - (g.gamma - 1) * (V[0] - ek * V[4]);
  const T log_p   = -(entropy + g.gamma * log(-V[4])) / (g.gamma - 1);
  Y[0]            = exp(log_p);
}

// StateConservativeFromEntropy (newtonian_state.h:208-222)
template <typename T>
__device__ __forceinline__ void cons_from_entropy(const Gas<T> &g, const T V[5], T U[5]) {
  T u[3];
  for (int i = 0; i < 3; i++) u[i] = -V[1 + i] / V[4];
  const T ek      = T(0.5) * dot3(u, u);
  const T entropy = g.gamma - (g.gamma - 1) * (V[0] - ek * V[4]);
  const T log_rho = -(entropy + log(-V[4])) / (g.gamma - 1);
  U[0]            = exp(log_rho);
  for (int i = 0; i < 3; i++) U[1 + i] = U[0] * u[i];
  const T ei = -g.cv / (g.Rgas * V[4]);
  U[4]       = U[0] * (ei + ek);
}

// Entropy variables of the state s, recomputed by the _fwd maps (newtonian_state.h:137, 227)
// StatePrimitiveFromEntropy_fwd (newtonian_state.h:135-146) and StateConservativeFromEntropy_fwd (:224-244)
template <typename T>
__device__ __forceinline__ void state_from_entropy_fwd(const Gas<T> &g, const St<T> &s, const T dV[5], St<T> &ds) {
  T V[5];
  entropy_from_prim(g, s.Y, V);
  // primitive part
  for (int i = 0; i < 3; i++) ds.Y[1 + i] = -(dV[1 + i] - V[1 + i] * dV[4] / V[4]) / V[4];
  ds.Y[4] = dV[4] / (g.Rgas * V[4] * V[4]);
  {
    const T ek   = T(0.5) * dot3(&s.Y[1], &s.Y[1]);
    const T dek  = dot3(&ds.Y[1], &s.Y[1]);
    const T dent = (1 - g.gamma) * (dV[0] - ek * dV[4] - dek * V[4]);
    ds.Y[0]      = s.Y[0] * (-dent - g.gamma * dV[4] / V[4]) / (g.gamma - 1);
  }
  // conservative part
  T du[3];
  for (int i = 0; i < 3; i++) du[i] = (-dV[1 + i] - s.Y[1 + i] * dV[4]) / V[4];
  const T ek      = T(0.5) * dot3(&s.Y[1], &s.Y[1]);
  const T dek     = dot3(du, &s.Y[1]);
  const T entropy = g.gamma - (g.gamma - 1) * (V[0] - ek * V[4]);
  const T dent    = -(g.gamma - 1) * (dV[0] - (dek * V[4] + ek * dV[4]));
  const T log_rho = -(entropy + log(-V[4])) / (g.gamma - 1);
  const T rho     = exp(log_rho);
  ds.U[0]         = -rho / (g.gamma - 1) * (dent + dV[4] / V[4]);
  for (int i = 0; i < 3; i++) ds.U[1 + i] = ds.U[0] * s.Y[1 + i] + s.U[0] * du[i];
  const T ei  = -g.cv / (g.Rgas * V[4]);
  const T dei = g.cv * dV[4] / (g.Rgas * V[4] * V[4]);
  ds.U[4]     = ds.U[0] * (ei + ek) + s.U[0] * (dei + dek);
}

// StateFromQ (newtonian_state.h:393-407), plus the quantities the state's tangent maps reuse
template <int VAR, typename T>
__device__ __forceinline__ St<T> state_from_q(const Gas<T> &g, const T q[5]) {
  St<T> s;
  if constexpr (VAR == VAR_CONS) {
    for (int i = 0; i < 5; i++) s.U[i] = q[i];
    prim_from_cons(g, s.U, s.Y);
  } else if constexpr (VAR == VAR_PRIM) {
    for (int i = 0; i < 5; i++) s.Y[i] = q[i];
    cons_from_prim(g, s.Y, s.U);
  } else {
    cons_from_entropy(g, q, s.U);
    prim_from_entropy(g, q, s.Y);
  }
  const T ek = T(.5) * dot3(&s.Y[1], &s.Y[1]);
  s.rRT2     = T(1) / (g.Rgas * s.Y[4] * s.Y[4]);
  s.etot     = g.cv * s.Y[4] + ek;
  if constexpr (VAR == VAR_CONS) {
    s.rrho = T(1) / s.U[0];
    s.et   = s.U[4] / s.U[0];
    s.ei_c = s.et - ek;
  } else {
    s.rrho = s.et = s.ei_c = T(0);  // prim_from_cons_fwd is only used for conservative states
  }
  return s;
}

// StateFromQ_fwd (newtonian_state.h:409-423)
template <int VAR, typename T>
__device__ __forceinline__ St<T> state_from_q_fwd(const Gas<T> &g, const St<T> &s, const T dq[5]) {
  St<T> ds;
  if constexpr (VAR == VAR_CONS) {
    for (int i = 0; i < 5; i++) ds.U[i] = dq[i];
    prim_from_cons_fwd(g, s, ds.U, ds.Y);
  } else if constexpr (VAR == VAR_PRIM) {
    for (int i = 0; i < 5; i++) ds.Y[i] = dq[i];
    cons_from_prim_fwd(g, s, ds.Y, ds.U);
  } else {
    state_from_entropy_fwd(g, s, dq, ds);
  }
  return ds;
}

// FluxInviscid_fwd, direction i only (newtonian_state.h:433-441)
template <typename T>
__device__ __forceinline__ void flux_inviscid_fwd_dir(const St<T> &s, const St<T> &ds, int i, T dF[5]) {
  dF[0] = ds.U[1 + i];
  for (int j = 0; j < 3; j++) dF[1 + j] = ds.U[1 + i] * s.Y[1 + j] + s.U[1 + i] * ds.Y[1 + j] + ds.Y[0] * (T)(i == j);
  dF[4] = (ds.U[4] + ds.Y[0]) * s.Y[1 + i] + (s.U[4] + s.Y[0]) * ds.Y[1 + i];
}

// FluxInviscid, direction i (newtonian_state.h:425-431)
template <typename T>
__device__ __forceinline__ void flux_inviscid_dir(const St<T> &s, int i, T F[5]) {
  F[0] = s.U[1 + i];
  for (int j = 0; j < 3; j++) F[1 + j] = s.U[1 + i] * s.Y[1 + j] + s.Y[0] * (T)(i == j);
  F[4] = (s.U[4] + s.Y[0]) * s.Y[1 + i];
}

// KMStrainRate_State (newtonian_state.h:511-538) from the three gradient states
template <typename T>
__device__ __forceinline__ void km_strain_rate(const St<T> gs[3], T sr[6]) {
  const T w = T(1) / (T)sqrt(2.);
  sr[0]     = gs[0].Y[1];
  sr[1]     = gs[1].Y[2];
  sr[2]     = gs[2].Y[3];
  sr[3]     = w * (gs[2].Y[2] + gs[1].Y[3]);
  sr[4]     = w * (gs[2].Y[1] + gs[0].Y[3]);
  sr[5]     = w * (gs[1].Y[1] + gs[0].Y[2]);
}

// NewtonianStress + KMUnpack (newtonian_state.h:553-558, utils.h:239-247): full symmetric stress
template <typename T>
__device__ __forceinline__ void stress_from_strain(const Gas<T> &g, const T sr[6], T st[3][3]) {
  const T div_u = sr[0] + sr[1] + sr[2];
  T       km[6];
  for (int i = 0; i < 6; i++) km[i] = g.mu * (2 * sr[i] + g.lambda * div_u * (T)(i < 3));
  const T w = T(1) / (T)sqrt(2.);
  st[0][0]  = km[0];
  st[1][1]  = km[1];
  st[2][2]  = km[2];
  st[2][1] = st[1][2] = w * km[3];
  st[2][0] = st[0][2] = w * km[4];
  st[1][0] = st[0][1] = w * km[5];
}

// Tau_diagPrim (stabilization.h:67-107)
template <typename T>
__device__ __forceinline__ void tau_diag_prim(const CeedMiTauCoeffs &tc, const Gas<T> &g, const St<T> &s, const T dXdx[3][3], T dt, T Tau[3]) {
  const T rho = s.U[0];
  T       gij[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) gij[i][j] = dXdx[0][i] * dXdx[0][j] + dXdx[1][i] * dXdx[1][j] + dXdx[2][i] * dXdx[2][j];
  const T dts = (T)tc.Ctau_t / dt;
  T       gu[3];
  for (int i = 0; i < 3; i++) gu[i] = gij[i][0] * s.Y[1] + gij[i][1] * s.Y[2] + gij[i][2] * s.Y[3];
  const T vt = dot3(&s.Y[1], gu);
  T       gg = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) gg += gij[i][j] * gij[i][j];
  const T tau   = sq(rho) * (T(4.) * sq(dts) + vt) + (T)tc.Ctau_v * sq(g.mu) * gg;
  const T fact  = sqrt(tau);
  const T rfact = T(1) / fact;
  Tau[0]        = (T)tc.Ctau_C * fact / (rho * (gij[0][0] + gij[1][1] + gij[2][2])) * T(0.125);
  Tau[1]        = (T)tc.Ctau_M * rfact;
  Tau[2]        = (T)tc.Ctau_E * rfact * g.rcv;
}

// StabilizationMatrix with dYFromTau (stabilization.h:14-39): stab[var][dir] added into gv
template <typename T>
__device__ __forceinline__ void add_stab_matrix(const Gas<T> &g, const St<T> &s, const T Tau[3], const T r[5], T gv[5][3]) {
  const T dY[5] = {Tau[0] * r[0], Tau[1] * r[1], Tau[1] * r[2], Tau[1] * r[3], Tau[2] * r[4]};
  St<T>   ds;
  for (int i = 0; i < 5; i++) ds.Y[i] = dY[i];
  cons_from_prim_fwd(g, s, ds.Y, ds.U);
  for (int i = 0; i < 3; i++) {
    T dF[5];
    flux_inviscid_fwd_dir(s, ds, i, dF);
    for (int j = 0; j < 5; j++) gv[j][i] += dF[j];
  }
}

// The state perturbation of StabilizationMatrix: dY = Tau o r on the primitive slots (dYFromTau) mapped by
// StateFromY_fwd (stabilization.h:14-39); the stabilization flux of direction i is FluxInviscid_fwd(s, ds, i).
template <typename T>
__device__ __forceinline__ void stab_state(const Gas<T> &g, const St<T> &s, const T Tau[3], const T r[5], St<T> &ds) {
  const T dY[5] = {Tau[0] * r[0], Tau[1] * r[1], Tau[1] * r[2], Tau[1] * r[3], Tau[2] * r[4]};
  for (int i = 0; i < 5; i++) ds.Y[i] = dY[i];
  cons_from_prim_fwd(g, s, ds.Y, ds.U);
}

// Strong convective residual sum_i dF_i/dq . dq/dx_i (FluxInviscidStrong, newtonian_state.h:466-475)
template <typename T>
__device__ __forceinline__ void flux_inviscid_strong(const St<T> &s, const St<T> gs[3], T r[5]) {
  for (int j = 0; j < 5; j++) r[j] = 0;
  for (int i = 0; i < 3; i++) {
    T dF[5];
    flux_inviscid_fwd_dir(s, gs[i], i, dF);
    for (int j = 0; j < 5; j++) r[j] += dF[j];
  }
}

// LinearRampCoefficient (utils.h:287-295)
template <typename T>
__device__ __forceinline__ T linear_ramp(T amplitude, T length, T start, T x) {
  if (x < start) return amplitude;
  if (x < start + length) return amplitude * ((x - start) * (T(-1) / length) + 1);
  return 0;
}

// Physical gradient state in direction j from reference gradients gref[d][c] (one column of
// StatePhysicalGradientFromReference, newtonian_state.h:581-593)
template <int VAR, typename T>
__device__ __forceinline__ St<T> phys_gradient_dir(const Gas<T> &g, const St<T> &s, const T gref[3][5], const T dXdx[3][3], int j) {
  T dq[5];
  for (int c = 0; c < 5; c++) dq[c] = gref[0][c] * dXdx[0][j] + gref[1][c] * dXdx[1][j] + gref[2][c] * dXdx[2][j];
  return state_from_q_fwd<VAR>(g, s, dq);
}

// Physical gradient states from reference gradients gref[d][c] (StatePhysicalGradientFromReference,
// newtonian_state.h:581-593)
template <int VAR, typename T>
__device__ __forceinline__ void phys_gradient(const Gas<T> &g, const St<T> &s, const T gref[3][5], const T dXdx[3][3], St<T> gs[3]) {
  for (int j = 0; j < 3; j++) {
    T dq[5];
    for (int c = 0; c < 5; c++) dq[c] = gref[0][c] * dXdx[0][j] + gref[1][c] * dXdx[1][j] + gref[2][c] * dXdx[2][j];
    gs[j] = state_from_q_fwd<VAR>(g, s, dq);
  }
}

}  // namespace ns
