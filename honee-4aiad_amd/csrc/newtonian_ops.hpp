// Quadrature-point bodies of the fused Newtonian volume operators.
//
// Each Op describes one HONEE QFunction for the generic fused element kernel (kernels_volume.hip): which element
// fields it interpolates (active input with/without gradient, a second passive L-vector), and a device qf() that maps
// the interpolated values at one quadrature point to the test-function weights v (for B^T) and gw (for the reference
// gradient transpose), reading/writing its strided quadrature data directly.  gw[d][c] already includes
// wdetJ * dXdx, i.e. gw[d][c] = wdetJ * sum_j gv_i[c][j] dXdx[d][j] as the reference QFunctions store Grad_v.
#pragma once

#include "kernels.hpp"
#include "ns_physics.hpp"

namespace nsop {

// Index of component c at point i of element e in strided quadrature data of the backend layout [elem][comp][qpt]
// (the operator binding checks sn = 1 and sc = Q^3): the component offset is a compile-time constant.
template <int Q>
__device__ __forceinline__ long long qidx(const QDataLayout &l, int e, int c, int i) {
  return (long long)e * l.se + c * (Q * Q * Q) + i;
}

// The 10 quadrature-data values {wdetJ, dXdx} of point t of element e, read from the strided field
template <int Q>
__device__ __forceinline__ void fetch_qdata(const VolumeArgs &a, int e, int t, bool ok, double qd[10]) {
#pragma unroll
  for (int k = 0; k < 10; k++) qd[k] = ok ? a.qdata[qidx<Q>(a.qd, e, k, t)] : 0.0;
}

template <typename T>
__device__ __forceinline__ void unpack_qdata(const double qd[10], bool ok, T &wdetJ, T dXdx[3][3]) {
  if (ok) {
    wdetJ = (T)qd[0];
#pragma unroll
    for (int k = 0; k < 9; k++) (&dXdx[0][0])[k] = (T)qd[1 + k];
  } else {  // padding thread of a partial batch: harmless geometry, zero weight
    wdetJ = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) (&dXdx[0][0])[k] = (k % 4 == 0) ? T(1) : T(0);
  }
}

// A physical state for padding threads so the pointwise physics stays finite
template <int VAR, typename T>
__device__ __forceinline__ void padding_state(T q[5]) {
#pragma unroll
  for (int c = 0; c < 5; c++) q[c] = (VAR == ns::VAR_ENTROPY) ? (c == 4 ? T(-1) : T(0)) : T(1);
}

// gw[d][c] = wdetJ * (gv[c][0] dXdx[d][0] + gv[c][1] dXdx[d][1] + gv[c][2] dXdx[d][2])   (newtonian.h:245-249)
template <typename T>
__device__ __forceinline__ void map_grad_out(T wdetJ, const T dXdx[3][3], const T gv[5][3], double gw[3][5]) {
#pragma unroll
  for (int c = 0; c < 5; c++)
#pragma unroll
    for (int d = 0; d < 3; d++) gw[d][c] = (double)(wdetJ * (gv[c][0] * dXdx[d][0] + gv[c][1] * dXdx[d][1] + gv[c][2] * dXdx[d][2]));
}

// The same mapping formed one direction j at a time: acc[d][c] += gv_j[c] dXdx[d][j] for j = 0, 1, 2 in order, then
// gw = wdetJ acc (identical operation order to map_grad_out)
template <typename T>
__device__ __forceinline__ void accumulate_grad_out(int j, const T dXdx[3][3], const T gvj[5], T acc[3][5]) {
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int c = 0; c < 5; c++) acc[d][c] = (j == 0) ? gvj[c] * dXdx[d][0] : acc[d][c] + gvj[c] * dXdx[d][j];
}
template <typename T>
__device__ __forceinline__ void finish_grad_out(T wdetJ, const T acc[3][5], double gw[3][5]) {
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int c = 0; c < 5; c++) gw[d][c] = (double)(wdetJ * acc[d][c]);
}

// {w detJ, dXdx} of quadrature point t from an element's vertex coordinates xv [3][8] (lexicographic vertices) and the
// 1D tables of the trilinear coordinate basis (Setup, qfunctions/setupgeo.h:42-56, with MatInv3, qfunctions/utils.h:201-
// 216).  The setup kernel and the volume kernels that recompute qdata both evaluate this function, so the recomputed
// values are the stored ones.
template <int Q>
__device__ __forceinline__ void geo_point(const double *xv, const double *Bx, const double *Dx, const double *w1d, int t, double qd[10]) {
  const int qx = t % Q, qy = (t / Q) % Q, qz = t / (Q * Q);
  double    J[3][3] = {{0}};
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int aa = 0; aa < 2; aa++) {
        const int    n  = aa + 2 * (b + 2 * c);
        const double bx = Bx[qx * 2 + aa], by = Bx[qy * 2 + b], bz = Bx[qz * 2 + c];
        const double dx = Dx[qx * 2 + aa], dy = Dx[qy * 2 + b], dz = Dx[qz * 2 + c];
#pragma unroll
        for (int comp = 0; comp < 3; comp++) {
          const double x = xv[comp * 8 + n];
          J[comp][0] += dx * by * bz * x;
          J[comp][1] += bx * dy * bz * x;
          J[comp][2] += bx * by * dz * x;
        }
      }
  double Ai[3][3];
  Ai[0][0]       = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  Ai[0][1]       = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  Ai[0][2]       = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  Ai[1][0]       = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  Ai[1][1]       = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  Ai[1][2]       = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  Ai[2][0]       = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  Ai[2][1]       = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  Ai[2][2]       = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double d = J[0][0] * Ai[0][0] + J[1][0] * Ai[0][1] + J[2][0] * Ai[0][2];
  const double s = 1 / d;
  const double w = w1d[qx] * w1d[qy] * w1d[qz];
  qd[0]          = w * d;
#pragma unroll
  for (int k = 0; k < 9; k++) qd[1 + k] = (&Ai[0][0])[k] * s;
}

// x coordinate (first component) of the quadrature point from the trilinear coordinate field
template <int Q>
__device__ __forceinline__ double coord_x(const VolumeArgs &a, int e, int t) {
  const int qx = t % Q, qy = (t / Q) % Q, qz = t / (Q * Q);
  double    x0 = 0;
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const int xo = a.xoffsets[(long long)e * 8 + n];
    x0 += a.Bx[qx * 2 + (n & 1)] * a.Bx[qy * 2 + ((n >> 1) & 1)] * a.Bx[qz * 2 + (n >> 2)] * a.x[xo];
  }
  return x0;
}

// ---------------------------------------------------------------------------------------------------------------------
// IFunction_Newtonian (newtonian.h:205-257) with IFunction_Newtonian_Integrand (newtonian_integrands.h:87-115) and the
// internal damping layer (newtonian_integrands.h:40-48).  Writes jac_data: q, strain rate (KM), Tau_d (+ sigma).
// ---------------------------------------------------------------------------------------------------------------------
template <int VAR>
struct IFunction {
  static constexpr bool kGrad = true, kIn2 = true, kOutV = true, kOutG = true, kScatter = true;
  static constexpr bool kStageQData = true;  // prefetch quadrature data into LDS (measured faster for this body)
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int e, int t, bool ok, const double qd[10], double q[5], double gref[3][5],
                            double qdot[5], double v[5], double gw[3][5]) {
    const ns::Gas<double> g = ns::Gas<double>::from(ctx.gas);
    double                wdetJ, dXdx[3][3];
    unpack_qdata(qd, ok, wdetJ, dXdx);
    if (!ok) {
      padding_state<VAR>(q);
#pragma unroll
      for (int c = 0; c < 5; c++) {
        qdot[c] = 0;
        gref[0][c] = gref[1][c] = gref[2][c] = 0;
      }
    }
    // jac_data (newtonian.h:251-254), each value written as soon as it is known so that it is not held in registers;
    // streamed past the caches (nontemporal), measured faster: nothing in this apply reads it back
    auto put = [&](int j, double val) {
      if (ok) {
        if (a.jac32) __builtin_nontemporal_store((float)val, &a.jac32[qidx<Q>(a.jd, e, j, t)]);
        else if (a.jac) __builtin_nontemporal_store(val, &a.jac[qidx<Q>(a.jd, e, j, t)]);
      }
    };
    const ns::St<double> s = ns::state_from_q<VAR>(g, q);
#pragma unroll
    for (int j = 0; j < 5; j++) put(j, q[j]);
    // Gradient states one direction at a time: the strong convective residual (FluxInviscidStrong,
    // newtonian_state.h:466-475) is accumulated in the same direction order, and only the primitive gradients
    // (velocity, temperature) are kept for the viscous fluxes.
    double r[5] = {0, 0, 0, 0, 0}, gY[3][4];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const ns::St<double> gsj = ns::phys_gradient_dir<VAR>(g, s, gref, dXdx, j);
      double               dF[5];
      ns::flux_inviscid_fwd_dir(s, gsj, j, dF);
#pragma unroll
      for (int k = 0; k < 5; k++) r[k] += dF[k];
#pragma unroll
      for (int k = 0; k < 4; k++) gY[j][k] = gsj.Y[1 + k];
    }
    // KMStrainRate_State / NewtonianStress / ViscousEnergyFlux (newtonian_state.h:511-573)
    const double w     = 1 / sqrt(2.);
    const double sr[6] = {gY[0][0], gY[1][1], gY[2][2], w * (gY[2][1] + gY[1][2]), w * (gY[2][0] + gY[0][2]), w * (gY[1][0] + gY[0][1])};
#pragma unroll
    for (int j = 0; j < 6; j++) put(5 + j, sr[j]);
    double st[3][3], Fe[3];
    ns::stress_from_strain(g, sr, st);
#pragma unroll
    for (int i = 0; i < 3; i++) Fe[i] = -s.Y[1] * st[0][i] - s.Y[2] * st[1][i] - s.Y[3] * st[2][i] - g.k * gY[i][3];
    // U_dot - body force
    const ns::St<double> sd    = ns::state_from_q_fwd<VAR>(g, s, qdot);
    const double         bf[5] = {0, s.U[0] * ctx.g[0], s.U[0] * ctx.g[1], s.U[0] * ctx.g[2], s.U[1] * ctx.g[0] + s.U[2] * ctx.g[1] + s.U[3] * ctx.g[2]};
    double               vi[5];
#pragma unroll
    for (int j = 0; j < 5; j++) vi[j] = sd.U[j] - bf[j];
    double Tau[3];
    ns::tau_diag_prim<double>(ctx.tau_coeffs, g, s, dXdx, ctx.dt, Tau);
#pragma unroll
    for (int j = 0; j < 3; j++) put(11 + j, Tau[j]);
    // Stabilization (stabilization.h:14-58, divFdiff = 0): the state perturbation StateFromY_fwd(Tau o r)
    const bool     stabilized = ctx.stabilization != ns::STAB_NONE;
    ns::St<double> dst;
    if (stabilized) {
      if (ctx.stabilization == ns::STAB_SUPG)
#pragma unroll
        for (int j = 0; j < 5; j++) r[j] += sd.U[j] - bf[j];
      ns::stab_state(g, s, Tau, r, dst);
    }
    // grad_v_i[var][dir] = -F_total + stab, one direction at a time, mapped with wdetJ dXdx as it is formed
    double acc[3][5];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double F[5], sj[5] = {0, 0, 0, 0, 0}, gvj[5];
      ns::flux_inviscid_dir(s, j, F);
      if (stabilized) ns::flux_inviscid_fwd_dir(s, dst, j, sj);
      gvj[0] = -F[0] + sj[0];
#pragma unroll
      for (int k = 0; k < 3; k++) gvj[k + 1] = -(F[1 + k] - st[k][j]) + sj[k + 1];
      gvj[4] = -(F[4] + Fe[j]) + sj[4];
      accumulate_grad_out(j, dXdx, gvj, acc);
    }
    if (ctx.idl_enable) {
      const double sigma = ns::linear_ramp<double>(ctx.idl_amplitude, ctx.idl_length, ctx.idl_start, ok ? coord_x<Q>(a, e, t) : 0.0);
      put(14, sigma);
      ns::St<double> ds;
      ds.Y[0] = (s.Y[0] - ctx.idl_pressure) * sigma;
#pragma unroll
      for (int i = 1; i < 5; i++) ds.Y[i] = 0 * sigma;
      ns::cons_from_prim_fwd(g, s, ds.Y, ds.U);
#pragma unroll
      for (int j = 0; j < 5; j++) vi[j] += ds.U[j];
    }
#pragma unroll
    for (int j = 0; j < 5; j++) v[j] = wdetJ * vi[j];
    finish_grad_out(wdetJ, acc, gw);
  }
};

// ---------------------------------------------------------------------------------------------------------------------
// IJacobian_Newtonian (newtonian.h:272-311) with IJacobian_Newtonian_Integrand (newtonian_integrands.h:134-166), in T.
// ---------------------------------------------------------------------------------------------------------------------
template <int VAR, typename T>
struct IJacobian {
  static constexpr bool kGrad = true, kIn2 = false, kOutV = true, kOutG = true, kScatter = true;
  static constexpr bool kStageQData = false;
  static constexpr bool kZCrossLane = sizeof(T) == 4;  // measured: faster in FP32, slower in FP64

  // State-dependent part at one quadrature point (shared by apply and diagonal assembly)
  struct Point {
    ns::St<T> s;
    T         wdetJ, dXdx[3][3], Tau[3], st[3][3], sigma;
  };

  static constexpr int kInSlots = 4;  // linear inputs: value and three reference derivatives
  template <int Q>
  __device__ static void load_point(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int e, int t, bool ok, const double qd[10], const double *, Point &p) {
    const ns::Gas<T> g = ns::Gas<T>::from(ctx.gas);
    unpack_qdata(qd, ok, p.wdetJ, p.dXdx);
    T qi[5], sr[6];
    p.sigma = 0;
    if (ok) {
      if (a.jac32) {
#pragma unroll
        for (int j = 0; j < 5; j++) qi[j] = (T)a.jac32[qidx<Q>(a.jd, e, j, t)];
#pragma unroll
        for (int j = 0; j < 6; j++) sr[j] = (T)a.jac32[qidx<Q>(a.jd, e, 5 + j, t)];
#pragma unroll
        for (int j = 0; j < 3; j++) p.Tau[j] = (T)a.jac32[qidx<Q>(a.jd, e, 11 + j, t)];
        if (ctx.idl_enable) p.sigma = (T)a.jac32[qidx<Q>(a.jd, e, 14, t)];
      } else {
#pragma unroll
        for (int j = 0; j < 5; j++) qi[j] = (T)a.jacr[qidx<Q>(a.jd, e, j, t)];
#pragma unroll
        for (int j = 0; j < 6; j++) sr[j] = (T)a.jacr[qidx<Q>(a.jd, e, 5 + j, t)];
#pragma unroll
        for (int j = 0; j < 3; j++) p.Tau[j] = (T)a.jacr[qidx<Q>(a.jd, e, 11 + j, t)];
        if (ctx.idl_enable) p.sigma = (T)a.jacr[qidx<Q>(a.jd, e, 14, t)];
      }
    } else {
      padding_state<VAR>(qi);
#pragma unroll
      for (int j = 0; j < 6; j++) sr[j] = 0;
#pragma unroll
      for (int j = 0; j < 3; j++) p.Tau[j] = 0;
    }
    p.s = ns::state_from_q<VAR>(g, qi);
    ns::stress_from_strain(g, sr, p.st);
  }

  // Linear action at a quadrature point: (dq, gref) -> (v, gw)
  __device__ static void action(const CeedMiNewtonianContext &ctx, const Point &p, const T dq[5], const T gref[3][5], double v[5],
                                double gw[3][5]) {
    const ns::Gas<T> g     = ns::Gas<T>::from(ctx.gas);
    const ns::St<T> &s     = p.s;
    const ns::St<T>  ds    = ns::state_from_q_fwd<VAR>(g, s, dq);
    const T          shift = (T)ctx.ijacobian_time_shift;
    T                r[5]  = {0, 0, 0, 0, 0}, gY[3][4];
#pragma unroll
    for (int j = 0; j < 3; j++) {  // gradient states one direction at a time (see IFunction)
      const ns::St<T> gdsj = ns::phys_gradient_dir<VAR>(g, s, gref, p.dXdx, j);
      T               dFj[5];
      ns::flux_inviscid_fwd_dir(s, gdsj, j, dFj);
#pragma unroll
      for (int k = 0; k < 5; k++) r[k] += dFj[k];
#pragma unroll
      for (int k = 0; k < 4; k++) gY[j][k] = gdsj.Y[1 + k];
    }
    const T w      = T(1) / (T)sqrt(2.);
    const T dsr[6] = {gY[0][0], gY[1][1], gY[2][2], w * (gY[2][1] + gY[1][2]), w * (gY[2][0] + gY[0][2]), w * (gY[1][0] + gY[0][1])};
    T       dstr[3][3], dFe[3];
    ns::stress_from_strain(g, dsr, dstr);
#pragma unroll
    for (int i = 0; i < 3; i++)  // ViscousEnergyFlux_fwd (newtonian_state.h:567-573)
      dFe[i] = -s.Y[1] * dstr[0][i] - ds.Y[1] * p.st[0][i] - s.Y[2] * dstr[1][i] - ds.Y[2] * p.st[1][i] - s.Y[3] * dstr[2][i] -
               ds.Y[3] * p.st[2][i] - g.k * gY[i][3];
    const T gx = (T)ctx.g[0], gy = (T)ctx.g[1], gz = (T)ctx.g[2];
    const T dbf[5] = {0, ds.U[0] * gx, ds.U[0] * gy, ds.U[0] * gz, ds.U[1] * gx + ds.U[2] * gy + ds.U[3] * gz};
    T       vi[5];
#pragma unroll
    for (int j = 0; j < 5; j++) vi[j] = shift * ds.U[j] - dbf[j];
    const bool stabilized = ctx.stabilization != ns::STAB_NONE;
    ns::St<T>  dstab;
    if (stabilized) {
      if (ctx.stabilization == ns::STAB_SUPG)
#pragma unroll
        for (int j = 0; j < 5; j++) r[j] += shift * ds.U[j] - dbf[j];
      ns::stab_state(g, s, p.Tau, r, dstab);
    }
    T acc[3][5];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      T dF[5], sj[5] = {0, 0, 0, 0, 0}, gvj[5];
      ns::flux_inviscid_fwd_dir(s, ds, j, dF);
      if (stabilized) ns::flux_inviscid_fwd_dir(s, dstab, j, sj);
      gvj[0] = -dF[0] + sj[0];
#pragma unroll
      for (int k = 0; k < 3; k++) gvj[k + 1] = -(dF[1 + k] - dstr[k][j]) + sj[k + 1];
      gvj[4] = -(dF[4] + dFe[j]) + sj[4];
      accumulate_grad_out(j, p.dXdx, gvj, acc);
    }
    if (ctx.idl_enable) {  // InternalDampingLayer_IJacobian_Integrand (newtonian_integrands.h:61-65)
      ns::St<T> dd;
      dd.Y[0] = ds.Y[0] * p.sigma;
#pragma unroll
      for (int i = 1; i < 5; i++) dd.Y[i] = 0 * p.sigma;
      ns::cons_from_prim_fwd(g, s, dd.Y, dd.U);
#pragma unroll
      for (int j = 0; j < 5; j++) vi[j] += dd.U[j];
    }
#pragma unroll
    for (int j = 0; j < 5; j++) v[j] = (double)(p.wdetJ * vi[j]);
    finish_grad_out(p.wdetJ, acc, gw);
  }

  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int e, int t, bool ok, const double qd[10], double u[5], double gu[3][5],
                            double *, double v[5], double gw[3][5]) {
    Point p;
    load_point<Q>(a, ctx, e, t, ok, qd, nullptr, p);
    T dq[5], gref[3][5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
      dq[c] = ok ? (T)u[c] : T(0);
#pragma unroll
      for (int d = 0; d < 3; d++) gref[d][c] = ok ? (T)gu[d][c] : T(0);
    }
    action(ctx, p, dq, gref, v, gw);
  }
};

// ---------------------------------------------------------------------------------------------------------------------
// RHSFunction_Newtonian (newtonian.h:133-202): explicit right-hand side in conservative variables
// ---------------------------------------------------------------------------------------------------------------------
// STAB: whether the stabilization term is compiled in.  The variant without it serves ctx.stabilization == none with
// the same values and fewer registers (166 against 184 VGPRs at Q = 3; the p=2 explicit residual at 96^3: 1.32
// against 1.46 ms per fused launch)
template <bool STAB>
struct RHSFunctionT {
  static constexpr bool kGrad = true, kIn2 = false, kOutV = true, kOutG = true, kScatter = true;
  static constexpr bool kStageQData = false;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int e, int t, bool ok, const double qd[10], double q[5], double gref[3][5],
                            double *, double v[5], double gw[3][5]) {
    const ns::Gas<double> g = ns::Gas<double>::from(ctx.gas);
    double                wdetJ, dXdx[3][3];
    unpack_qdata(qd, ok, wdetJ, dXdx);
    if (!ok) {
      padding_state<ns::VAR_CONS>(q);
#pragma unroll
      for (int c = 0; c < 5; c++) gref[0][c] = gref[1][c] = gref[2][c] = 0;
    }
    const ns::St<double> s = ns::state_from_q<ns::VAR_CONS>(g, q);
    ns::St<double>       gs[3];
    ns::phys_gradient<ns::VAR_CONS>(g, s, gref, dXdx, gs);
    double sr[6], st[3][3], Fe[3];
    ns::km_strain_rate(gs, sr);
    ns::stress_from_strain(g, sr, st);
#pragma unroll
    for (int i = 0; i < 3; i++) Fe[i] = -s.Y[1] * st[0][i] - s.Y[2] * st[1][i] - s.Y[3] * st[2][i] - g.k * gs[i].Y[4];
    double F[5][3];  // FluxTotal
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double Fi[5];
      ns::flux_inviscid_dir(s, j, Fi);
      F[0][j] = Fi[0];
#pragma unroll
      for (int k = 0; k < 3; k++) F[k + 1][j] = Fi[1 + k] - st[k][j];
      F[4][j] = Fi[4] + Fe[j];
    }
    const double bf[5] = {0, s.U[0] * ctx.g[0], s.U[0] * ctx.g[1], s.U[0] * ctx.g[2], s.U[1] * ctx.g[0] + s.U[2] * ctx.g[1] + s.U[3] * ctx.g[2]};
#pragma unroll
    for (int j = 0; j < 5; j++) v[j] = wdetJ * bf[j];
    if (ctx.idl_enable) {
      const double   sigma = ns::linear_ramp<double>(ctx.idl_amplitude, ctx.idl_length, ctx.idl_start, ok ? coord_x<Q>(a, e, t) : 0.0);
      ns::St<double> ds;
      ds.Y[0] = (s.Y[0] - ctx.idl_pressure) * sigma;
#pragma unroll
      for (int i = 1; i < 5; i++) ds.Y[i] = 0 * sigma;
      ns::cons_from_prim_fwd(g, s, ds.Y, ds.U);
#pragma unroll
      for (int j = 0; j < 5; j++) v[j] -= wdetJ * ds.U[j];
    }
    if constexpr (!STAB) {
#pragma unroll
      for (int c = 0; c < 5; c++)
#pragma unroll
        for (int d = 0; d < 3; d++) gw[d][c] = wdetJ * (dXdx[d][0] * F[c][0] + dXdx[d][1] * F[c][1] + dXdx[d][2] * F[c][2]);
      return;
    }
    double stab[5][3] = {{0}};
    if (ctx.stabilization != ns::STAB_NONE) {  // U_dot = 0, divFdiff = 0
      double Tau[3], r[5];
      ns::tau_diag_prim<double>(ctx.tau_coeffs, g, s, dXdx, ctx.dt, Tau);
      ns::flux_inviscid_strong(s, gs, r);
      if (ctx.stabilization == ns::STAB_SUPG)
#pragma unroll
        for (int j = 0; j < 5; j++) r[j] += 0 - bf[j];
      ns::add_stab_matrix(g, s, Tau, r, stab);
    }
#pragma unroll
    for (int c = 0; c < 5; c++)
#pragma unroll
      for (int d = 0; d < 3; d++)
        gw[d][c] = wdetJ * (dXdx[d][0] * F[c][0] + dXdx[d][1] * F[c][1] + dXdx[d][2] * F[c][2]) -
                   wdetJ * (stab[c][0] * dXdx[d][0] + stab[c][1] * dXdx[d][1] + stab[c][2] * dXdx[d][2]);
  }
};

// ---------------------------------------------------------------------------------------------------------------------
// MassFunction_Newtonian_Conserv (newtonian.h:44-82): SUPG-stabilized consistent mass acting on q_dot (active), with
// the state q passed as the second field.
// ---------------------------------------------------------------------------------------------------------------------
struct MassNewtonian {
  static constexpr bool kGrad = false, kIn2 = true, kOutV = true, kOutG = true, kScatter = true;
  static constexpr bool kStageQData = false;
  static constexpr int  kInSlots = 1;  // linear in the value of q_dot only

  struct Point {
    ns::St<double> s;
    double         wdetJ, dXdx[3][3], Tau[3];
  };
  template <int Q>
  __device__ static void load_point(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int e, int t, bool ok, const double qd[10], const double *q2, Point &p) {
    const ns::Gas<double> g = ns::Gas<double>::from(ctx.gas);
    unpack_qdata(qd, ok, p.wdetJ, p.dXdx);
    double q[5];
#pragma unroll
    for (int c = 0; c < 5; c++) q[c] = q2[c];
    if (!ok) padding_state<ns::VAR_CONS>(q);
    p.s = ns::state_from_q<ns::VAR_CONS>(g, q);
    ns::tau_diag_prim<double>(ctx.tau_coeffs, g, p.s, p.dXdx, ctx.dt, p.Tau);
  }
  __device__ static void action(const CeedMiNewtonianContext &ctx, const Point &p, const double qdot[5], const double (*)[5], double v[5],
                                double gw[3][5]) {
    const ns::Gas<double> g = ns::Gas<double>::from(ctx.gas);
#pragma unroll
    for (int f = 0; f < 5; f++) v[f] = p.wdetJ * qdot[f];
    double stab[5][3] = {{0}};
    if (ctx.stabilization != ns::STAB_NONE) {
      double r[5];
#pragma unroll
      for (int j = 0; j < 5; j++) r[j] = (ctx.stabilization == ns::STAB_SUPG) ? qdot[j] : 0.0;
      ns::add_stab_matrix(g, p.s, p.Tau, r, stab);
    }
#pragma unroll
    for (int c = 0; c < 5; c++)
#pragma unroll
      for (int d = 0; d < 3; d++) gw[d][c] = p.wdetJ * (stab[c][0] * p.dXdx[d][0] + stab[c][1] * p.dXdx[d][1] + stab[c][2] * p.dXdx[d][2]);
  }
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int e, int t, bool ok, const double qd[10], double qdot[5], double (*)[5],
                            double q[5], double v[5], double gw[3][5]) {
    const ns::Gas<double> g = ns::Gas<double>::from(ctx.gas);
    double                wdetJ, dXdx[3][3];
    unpack_qdata(qd, ok, wdetJ, dXdx);
    if (!ok) {
      padding_state<ns::VAR_CONS>(q);
#pragma unroll
      for (int c = 0; c < 5; c++) qdot[c] = 0;
    }
    const ns::St<double> s = ns::state_from_q<ns::VAR_CONS>(g, q);
#pragma unroll
    for (int f = 0; f < 5; f++) v[f] = wdetJ * qdot[f];
    double stab[5][3] = {{0}};
    if (ctx.stabilization != ns::STAB_NONE) {  // zero gradient, zero body force: r = U_dot for SUPG, 0 for SU
      double Tau[3], r[5];
      ns::tau_diag_prim<double>(ctx.tau_coeffs, g, s, dXdx, ctx.dt, Tau);
#pragma unroll
      for (int j = 0; j < 5; j++) r[j] = (ctx.stabilization == ns::STAB_SUPG) ? qdot[j] : 0.0;
      ns::add_stab_matrix(g, s, Tau, r, stab);
    }
#pragma unroll
    for (int c = 0; c < 5; c++)
#pragma unroll
      for (int d = 0; d < 3; d++) gw[d][c] = wdetJ * (stab[c][0] * dXdx[d][0] + stab[c][1] * dXdx[d][1] + stab[c][2] * dXdx[d][2]);
  }
};

// Mass_5 (qfunctions/mass.h): v = wdetJ u
struct Mass5 {
  static constexpr bool kGrad = false, kIn2 = false, kOutV = true, kOutG = false, kScatter = true;
  static constexpr bool kStageQData = false;
  static constexpr int  kInSlots = 1;
  struct Point {
    double wdetJ;
  };
  template <int Q>
  __device__ static void load_point(const VolumeArgs &a, const CeedMiNewtonianContext &, int e, int t, bool ok, const double qd[10], const double *, Point &p) {
    p.wdetJ = ok ? qd[0] : 0.0;
  }
  __device__ static void action(const CeedMiNewtonianContext &, const Point &p, const double u[5], const double (*)[5], double v[5],
                                double gw[3][5]) {
#pragma unroll
    for (int c = 0; c < 5; c++) {
      v[c]     = p.wdetJ * u[c];
      gw[0][c] = gw[1][c] = gw[2][c] = 0;
    }
  }
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &, int e, int t, bool ok, const double qd[10], double u[5], double (*)[5], double *,
                            double v[5], double gw[3][5]) {
    const double wdetJ = ok ? qd[0] : 0.0;
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = wdetJ * u[c];
  }
};

// ---------------------------------------------------------------------------------------------------------------------
// MonitorTotalKineticEnergy (qfunctions/monitor_totalkineticenergy.h:8-39): five integrands per quadrature point,
// written to a strided field (no basis transpose).
// ---------------------------------------------------------------------------------------------------------------------
template <int VAR>
struct MonitorKE {
  static constexpr bool kGrad = true, kIn2 = false, kOutV = false, kOutG = false, kScatter = false;
  static constexpr bool kStageQData = false;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int e, int t, bool ok, const double qd[10], double q[5], double gref[3][5],
                            double *, double *, double (*)[5]) {
    if (!ok) return;
    const ns::Gas<double> g = ns::Gas<double>::from(ctx.gas);
    double                wdetJ, dXdx[3][3];
    unpack_qdata(qd, ok, wdetJ, dXdx);
    const ns::St<double> s = ns::state_from_q<VAR>(g, q);
    ns::St<double>       gs[3];
    ns::phys_gradient<VAR>(g, s, gref, dXdx, gs);
    double kmsr[6];
    ns::km_strain_rate(gs, kmsr);
    const double div_u = kmsr[0] + kmsr[1] + kmsr[2];
    const double w     = 1 / sqrt(2.);
    const double S[9]  = {kmsr[0], w * kmsr[5], w * kmsr[4], w * kmsr[5], kmsr[1], w * kmsr[3], w * kmsr[4], w * kmsr[3], kmsr[2]};
    double       SS    = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) SS += S[i] * S[i];
    // vorticity = curl u with grad_u[i][j] = d u_i / d x_j = gs[j].Y[1 + i]
    const double w0 = gs[1].Y[3] - gs[2].Y[2], w1 = gs[2].Y[1] - gs[0].Y[3], w2 = gs[0].Y[2] - gs[1].Y[1];
    double      *out = a.jac;  // strided output field
    out[qidx<Q>(a.jd, e, 0, t)] = wdetJ * 0.5 * s.U[0] * (s.Y[1] * s.Y[1] + s.Y[2] * s.Y[2] + s.Y[3] * s.Y[3]);
    out[qidx<Q>(a.jd, e, 1, t)] = wdetJ * -2 * g.mu * SS;
    out[qidx<Q>(a.jd, e, 2, t)] = wdetJ * -g.lambda * g.mu * (div_u * div_u);
    out[qidx<Q>(a.jd, e, 3, t)] = wdetJ * s.Y[0] * div_u;
    out[qidx<Q>(a.jd, e, 4, t)] = wdetJ * g.mu * (w0 * w0 + w1 * w1 + w2 * w2);
  }
};

}  // namespace nsop
