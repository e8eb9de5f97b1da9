// Approximate Riemann fluxes of the freestream boundary for the fused face kernels, templated on the arithmetic type
// (double for residuals, float for the FP32 Jacobian).  Same quantities and operation order as HONEE's
// qfunctions/riemann_solver.h (cited per function).  Left = interior state, right = exterior state, n = unit outward
// normal; a flux is a conservative 5-vector.
#pragma once

#include "ns_physics.hpp"

namespace ns {

// FluxInviscidDotNormal (newtonian_state.h:443-452)
template <typename T>
__device__ __forceinline__ void flux_dot_normal(const St<T> &s, const T n[3], T Fn[5]) {
  for (int j = 0; j < 5; j++) Fn[j] = 0;
  for (int i = 0; i < 3; i++) {
    T F[5];
    flux_inviscid_dir(s, i, F);
    for (int j = 0; j < 5; j++) Fn[j] += F[j] * n[i];
  }
}

// FluxInviscidDotNormal_fwd (newtonian_state.h:454-464)
template <typename T>
__device__ __forceinline__ void flux_dot_normal_fwd(const St<T> &s, const St<T> &ds, const T n[3], T dFn[5]) {
  for (int j = 0; j < 5; j++) dFn[j] = 0;
  for (int i = 0; i < 3; i++) {
    T dF[5];
    flux_inviscid_fwd_dir(s, ds, i, dF);
    for (int j = 0; j < 5; j++) dFn[j] += dF[j] * n[i];
  }
}

// TotalSpecificEnthalpy and its derivative (newtonian_state.h:52-62)
template <typename T>
__device__ __forceinline__ T total_enthalpy(const Gas<T> &g, const St<T> &s) {
  return g.cv * s.Y[4] + T(0.5) * dot3(&s.Y[1], &s.Y[1]) + s.Y[0] / s.U[0];
}
template <typename T>
__device__ __forceinline__ T total_enthalpy_fwd(const Gas<T> &g, const St<T> &s, const St<T> &ds) {
  return g.cv * ds.Y[4] + dot3(&ds.Y[1], &s.Y[1]) + ds.Y[0] / s.U[0] - s.Y[0] / sq(s.U[0]) * ds.U[0];
}

// ComputeHLLSpeeds_Roe (riemann_solver.h:87-106), with RoeSetup/RoeAverage (:27-47)
template <typename T>
__device__ __forceinline__ void hll_speeds(const Gas<T> &g, const St<T> &l, T ul, const St<T> &r, T ur, T &s_l, T &s_r) {
  const T sql = sqrt(l.U[0]), sqr = sqrt(r.U[0]);
  const T wl = sql / (sql + sqr), wr = sqr / (sql + sqr);
  const T u_roe = wl * ul + wr * ur;
  const T H_roe = wl * total_enthalpy(g, l) + wr * total_enthalpy(g, r);
  const T a_roe = sqrt((g.gamma - 1) * (H_roe - T(0.5) * sq(u_roe)));
  s_l           = u_roe - a_roe;
  s_r           = u_roe + a_roe;
}

// ComputeHLLSpeeds_Roe_fwd (riemann_solver.h:108-137), with RoeSetup_fwd/RoeAverage_fwd (:33-47)
template <typename T>
__device__ __forceinline__ void hll_speeds_fwd(const Gas<T> &g, const St<T> &l, const St<T> &dl, T ul, T dul, const St<T> &r, const St<T> &dr, T ur,
                                               T dur, T &s_l, T &ds_l, T &s_r, T &ds_r) {
  const T sql = sqrt(l.U[0]), sqr = sqrt(r.U[0]);
  const T wl = sql / (sql + sqr), wr = sqr / (sql + sqr);
  const T ss  = sq(sql + sqr);
  const T dwr = (sql / (2 * sqr * ss)) * dr.U[0] - (sqr / (2 * sql * ss)) * dl.U[0];
  const T dwl = (sqr / (2 * sql * ss)) * dl.U[0] - (sql / (2 * sqr * ss)) * dr.U[0];
  const T u_roe  = wl * ul + wr * ur;
  const T du_roe = ur * dwr + ul * dwl + wr * dur + wl * dul;
  const T H_l = total_enthalpy(g, l), H_r = total_enthalpy(g, r);
  const T dH_l = total_enthalpy_fwd(g, l, dl), dH_r = total_enthalpy_fwd(g, r, dr);
  const T H_roe  = wl * H_l + wr * H_r;
  const T dH_roe = H_r * dwr + H_l * dwl + wr * dH_r + wl * dH_l;
  const T a_roe  = sqrt((g.gamma - 1) * (H_roe - T(0.5) * sq(u_roe)));
  T       da_roe = T(0.5) * sqrt((g.gamma - 1) / (H_roe - T(0.5) * sq(u_roe))) * dH_roe;
  da_roe -= T(0.5) * sqrt(g.gamma - 1) * u_roe / sqrt(H_roe - T(0.5) * sq(u_roe)) * du_roe;
  s_l  = u_roe - a_roe;
  ds_l = du_roe - da_roe;
  s_r  = u_roe + a_roe;
  ds_r = du_roe + da_roe;
}

// RiemannFlux_HLL (riemann_solver.h:151-169) with Flux_HLL (:49-66)
template <typename T>
__device__ __forceinline__ void riemann_hll(const Gas<T> &g, const St<T> &l, const St<T> &r, const T n[3], T F[5]) {
  T Fl[5], Fr[5];
  flux_dot_normal(l, n, Fl);
  flux_dot_normal(r, n, Fr);
  const T ul = dot3(&l.Y[1], n), ur = dot3(&r.Y[1], n);
  T       s_l, s_r;
  hll_speeds(g, l, ul, r, ur, s_l, s_r);
  for (int i = 0; i < 5; i++) {
    if (0 <= s_l) F[i] = Fl[i];
    else if (s_r <= 0) F[i] = Fr[i];
    else F[i] = (s_r * Fl[i] - s_l * Fr[i] + s_l * s_r * (r.U[i] - l.U[i])) / (s_r - s_l);
  }
}

// RiemannFlux_HLL_fwd (riemann_solver.h:184-205) with Flux_HLL_fwd (:68-85)
template <typename T>
__device__ __forceinline__ void riemann_hll_fwd(const Gas<T> &g, const St<T> &l, const St<T> &dl, const St<T> &r, const St<T> &dr, const T n[3],
                                                T dF[5]) {
  T Fl[5], Fr[5], dFl[5], dFr[5];
  flux_dot_normal(l, n, Fl);
  flux_dot_normal(r, n, Fr);
  flux_dot_normal_fwd(l, dl, n, dFl);
  flux_dot_normal_fwd(r, dr, n, dFr);
  const T ul = dot3(&l.Y[1], n), ur = dot3(&r.Y[1], n);
  const T dul = dot3(&dl.Y[1], n), dur = dot3(&dr.Y[1], n);
  T       S_l, dS_l, S_r, dS_r;
  hll_speeds_fwd(g, l, dl, ul, dul, r, dr, ur, dur, S_l, dS_l, S_r, dS_r);
  for (int i = 0; i < 5; i++) {
    if (0 <= S_l) dF[i] = dFl[i];
    else if (S_r <= 0) dF[i] = dFr[i];
    else {
      const T S_diff = S_r - S_l;
      T       acc    = 0.;
      acc += (S_l * (-Fl[i] + Fr[i] + S_l * l.U[i] - S_l * r.U[i]) / sq(S_diff)) * dS_r;
      acc += (S_r * (Fl[i] - Fr[i] - S_r * l.U[i] + S_r * r.U[i]) / sq(S_diff)) * dS_l;
      acc += (S_r * dFl[i] - S_l * dFr[i] + S_r * S_l * (dr.U[i] - dl.U[i])) / S_diff;
      dF[i] = acc;
    }
  }
}

// RiemannFlux_HLLC_Star (riemann_solver.h:207-224): F_side + s_side fact star - s_side U_side
template <typename T>
__device__ __forceinline__ void hllc_star(const St<T> &side, const T F_side[5], const T n[3], T u_side, T s_side, T s_star, T F[5]) {
  const T fact    = side.U[0] * (s_side - u_side) / (s_side - s_star);
  const T denom   = side.U[0] * (s_side - u_side);
  const T star[5] = {T(1.0), side.Y[1] + (s_star - u_side) * n[0], side.Y[2] + (s_star - u_side) * n[1], side.Y[3] + (s_star - u_side) * n[2],
                     side.U[4] / side.U[0] + (s_star - u_side) * (s_star + side.Y[0] / denom)};
  const T b = s_side * fact, c = -s_side;
  for (int i = 0; i < 5; i++) F[i] = T(1) * F_side[i] + b * star[i] + c * side.U[i];
}

// RiemannFlux_HLLC_Star_fwd (riemann_solver.h:226-256)
template <typename T>
__device__ __forceinline__ void hllc_star_fwd(const St<T> &side, const St<T> &dside, const T F_side[5], const T dF_side[5], const T n[3], T u_side,
                                              T du_side, T s_side, T ds_side, T s_star, T ds_star, T dF[5]) {
  const T fact  = side.U[0] * (s_side - u_side) / (s_side - s_star);
  const T dfact = (side.U[0] * (ds_side - du_side) + dside.U[0] * (s_side - u_side)) / (s_side - s_star) -
                  fact / (s_side - s_star) * (ds_side - ds_star);
  const T denom  = side.U[0] * (s_side - u_side);
  const T ddenom = side.U[0] * (ds_side - du_side) + dside.U[0] * (s_side - u_side);
  const T star[5] = {T(1.0), side.Y[1] + (s_star - u_side) * n[0], side.Y[2] + (s_star - u_side) * n[1], side.Y[3] + (s_star - u_side) * n[2],
                     side.U[4] / side.U[0] + (s_star - u_side) * (s_star + side.Y[0] / denom)};
  const T dstar[5] = {T(0.), dside.Y[1] + (ds_star - du_side) * n[0], dside.Y[2] + (ds_star - du_side) * n[1], dside.Y[3] + (ds_star - du_side) * n[2],
                      dside.U[4] / side.U[0] - side.U[4] / sq(side.U[0]) * dside.U[0] + (ds_star - du_side) * (s_star + side.Y[0] / denom) +
                          (s_star - u_side) * (ds_star + dside.Y[0] / denom - side.Y[0] / sq(denom) * ddenom)};
  const T a0 = 1, a1 = ds_side * fact + s_side * dfact, a2 = s_side * fact, a3 = -ds_side, a4 = -s_side;
  for (int i = 0; i < 5; i++) {  // StateConservativeMult (newtonian_state.h:263-271): accumulate term by term from zero
    T acc = 0;
    acc += a0 * dF_side[i];
    acc += a1 * star[i];
    acc += a2 * dstar[i];
    acc += a3 * side.U[i];
    acc += a4 * dside.U[i];
    dF[i] = acc;
  }
}

// RiemannFlux_HLLC (riemann_solver.h:258-284)
template <typename T>
__device__ __forceinline__ void riemann_hllc(const Gas<T> &g, const St<T> &l, const St<T> &r, const T n[3], T F[5]) {
  T Fl[5], Fr[5];
  flux_dot_normal(l, n, Fl);
  flux_dot_normal(r, n, Fr);
  const T ul = dot3(&l.Y[1], n), ur = dot3(&r.Y[1], n);
  T       s_l, s_r;
  hll_speeds(g, l, ul, r, ur, s_l, s_r);
  const T rhou_l = l.U[0] * ul, rhou_r = r.U[0] * ur;
  const T numer  = r.Y[0] - l.Y[0] + rhou_l * (s_l - ul) - rhou_r * (s_r - ur);
  const T denom  = l.U[0] * (s_l - ul) - r.U[0] * (s_r - ur);
  const T s_star = numer / denom;
  if (0 <= s_l) {
    for (int i = 0; i < 5; i++) F[i] = Fl[i];
  } else if (0 <= s_star) {
    hllc_star(l, Fl, n, ul, s_l, s_star, F);
  } else if (0 <= s_r) {
    hllc_star(r, Fr, n, ur, s_r, s_star, F);
  } else {
    for (int i = 0; i < 5; i++) F[i] = Fr[i];
  }
}

// RiemannFlux_HLLC_fwd (riemann_solver.h:286-338)
template <typename T>
__device__ __forceinline__ void riemann_hllc_fwd(const Gas<T> &g, const St<T> &l, const St<T> &dl, const St<T> &r, const St<T> &dr, const T n[3],
                                                 T dF[5]) {
  T Fl[5], Fr[5], dFl[5], dFr[5];
  flux_dot_normal(l, n, Fl);
  flux_dot_normal(r, n, Fr);
  flux_dot_normal_fwd(l, dl, n, dFl);
  flux_dot_normal_fwd(r, dr, n, dFr);
  const T ul = dot3(&l.Y[1], n), ur = dot3(&r.Y[1], n);
  const T dul = dot3(&dl.Y[1], n), dur = dot3(&dr.Y[1], n);
  T       s_l, ds_l, s_r, ds_r;
  hll_speeds_fwd(g, l, dl, ul, dul, r, dr, ur, dur, s_l, ds_l, s_r, ds_r);
  const T rhou_l = l.U[0] * ul, drhou_l = l.U[0] * dul + dl.U[0] * ul;
  const T rhou_r = r.U[0] * ur, drhou_r = r.U[0] * dur + dr.U[0] * ur;
  const T numer   = r.Y[0] - l.Y[0] + rhou_l * (s_l - ul) - rhou_r * (s_r - ur);
  const T dnumer  = dr.Y[0] - dl.Y[0] + rhou_l * (ds_l - dul) + drhou_l * (s_l - ul) - rhou_r * (ds_r - dur) - drhou_r * (s_r - ur);
  const T denom   = l.U[0] * (s_l - ul) - r.U[0] * (s_r - ur);
  const T ddenom  = l.U[0] * (ds_l - dul) + dl.U[0] * (s_l - ul) - r.U[0] * (ds_r - dur) - dr.U[0] * (s_r - ur);
  const T s_star  = numer / denom;
  const T ds_star = dnumer / denom - numer / sq(denom) * ddenom;
  if (0 <= s_l) {
    for (int i = 0; i < 5; i++) dF[i] = dFl[i];
  } else if (0 <= s_star) {
    hllc_star_fwd(l, dl, Fl, dFl, n, ul, dul, s_l, ds_l, s_star, ds_star, dF);
  } else if (0 <= s_r) {
    hllc_star_fwd(r, dr, Fr, dFr, n, ur, dur, s_r, ds_r, s_star, ds_star, dF);
  } else {
    for (int i = 0; i < 5; i++) dF[i] = dFr[i];
  }
}

}  // namespace ns
