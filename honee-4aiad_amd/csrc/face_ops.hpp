// Quadrature-point bodies of the fused boundary-face operators: HONEE's consistent boundary integral
// (qfunctions/newtonian.h:328-431), the freestream Riemann boundary (qfunctions/bc_freestream.h:14-114), the Riemann
// and weak-pressure outflows (qfunctions/bc_outflow.h:37-305) and the slip wall (qfunctions/bc_slip.h:11-111), each
// with its Jacobian.  A body maps the interpolated face values (and, for the boundary integral, the two reference tangential
// derivatives) at one face quadrature point to the test-function weights v (interpolation transpose only), reading and
// writing its strided face data directly.
//
// Face quadrature data (SetupBoundary, setupgeo.h:100-117): {wdetJb, normal[3], dXdx[2][3]} per point, layout
// [face][10][Q^2]; face Jacobian data [face][ncomp][Q^2].
#pragma once

#include "kernels.hpp"
#include "ns_physics.hpp"
#include "riemann.hpp"

namespace fop {

template <int Q>
__device__ __forceinline__ long long fidx(const QDataLayout &l, int f, int c, int i) {
  return (long long)f * l.se + c * (Q * Q) + i;
}

// QdataBoundaryUnpack_3D (utils.h:374-388)
template <int Q, typename T>
__device__ __forceinline__ void unpack_face_qdata(const VolumeArgs &a, int f, int t, bool ok, T &wdetJb, T dXdx[2][3], T n[3]) {
  if (ok) {
    wdetJb = (T)a.qdata[fidx<Q>(a.qd, f, 0, t)];
#pragma unroll
    for (int d = 0; d < 3; d++) n[d] = (T)a.qdata[fidx<Q>(a.qd, f, 1 + d, t)];
#pragma unroll
    for (int j = 0; j < 6; j++) (&dXdx[0][0])[j] = (T)a.qdata[fidx<Q>(a.qd, f, 4 + j, t)];
  } else {  // padding thread of a partial batch: unit normal, zero weight
    wdetJb = 0;
    n[0] = 1, n[1] = 0, n[2] = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) (&dXdx[0][0])[j] = (j == 0 || j == 4) ? T(1) : T(0);
  }
}

template <int VAR, typename T>
__device__ __forceinline__ void padding_q(T q[5]) {
#pragma unroll
  for (int c = 0; c < 5; c++) q[c] = (VAR == ns::VAR_ENTROPY) ? (c == 4 ? T(-1) : T(0)) : T(1);
}

// StatePhysicalGradientFromReference_Boundary (newtonian_state.h:595-605), direction k
template <int VAR, typename T>
__device__ __forceinline__ ns::St<T> phys_gradient_boundary_dir(const ns::Gas<T> &g, const ns::St<T> &s, const T gref[2][5], const T dXdx[2][3],
                                                                int k) {
  T dq[5];
#pragma unroll
  for (int j = 0; j < 5; j++) dq[j] = gref[0][j] * dXdx[0][k] + gref[1][j] * dXdx[1][k];
  return ns::state_from_q_fwd<VAR>(g, s, dq);
}

// Viscous part shared by the boundary integral and its Jacobian: from the three gradient (tangent) states, the KM
// strain rate, the Newtonian stress and its unpacked 3x3 form
template <typename T>
__device__ __forceinline__ void strain_rate_from_states(const ns::St<T> gs[3], T sr[6]) {
  const T w = T(1) / (T)sqrt(2.);
  sr[0]     = gs[0].Y[1];
  sr[1]     = gs[1].Y[2];
  sr[2]     = gs[2].Y[3];
  sr[3]     = w * (gs[2].Y[2] + gs[1].Y[3]);
  sr[4]     = w * (gs[2].Y[1] + gs[0].Y[3]);
  sr[5]     = w * (gs[1].Y[1] + gs[0].Y[2]);
}

// FluxTotal_Boundary (newtonian_state.h:486-496) accumulated one direction at a time
template <typename T>
__device__ __forceinline__ void add_total_flux_dir(const T Finv[5], const T st[3][3], T Fe_j, T n_j, int j, T Flux[5]) {
  Flux[0] += Finv[0] * n_j;
#pragma unroll
  for (int k = 0; k < 3; k++) Flux[k + 1] += (Finv[1 + k] - st[k][j]) * n_j;
  Flux[4] += (Finv[4] + Fe_j) * n_j;
}

// BoundaryIntegral (newtonian.h:328-375): v = wdetJb F_total . n; jac_data_sur = {q, strain rate} when implicit,
// v negated when explicit
template <int VAR>
struct BoundaryIntegral {
  static constexpr bool kGrad = true;
  static constexpr int  kJacComps = 11;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int f, int t, bool ok, double q[5], double gref[2][5],
                            double v[5]) {
    const ns::Gas<double> g = ns::Gas<double>::from(ctx.gas);
    double                wdetJb, dXdx[2][3], n[3];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    if (!ok) {
      padding_q<VAR>(q);
#pragma unroll
      for (int c = 0; c < 5; c++) gref[0][c] = gref[1][c] = 0;
    }
    const ns::St<double> s = ns::state_from_q<VAR>(g, q);
    ns::St<double>       gs[3];
#pragma unroll
    for (int k = 0; k < 3; k++) gs[k] = phys_gradient_boundary_dir<VAR>(g, s, gref, dXdx, k);
    double sr[6], st[3][3];
    strain_rate_from_states(gs, sr);
    ns::stress_from_strain(g, sr, st);
    double Flux[5] = {0., 0., 0., 0., 0.};
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double F[5];
      ns::flux_inviscid_dir(s, j, F);
      const double Fe_j = -s.Y[1] * st[0][j] - s.Y[2] * st[1][j] - s.Y[3] * st[2][j] - g.k * gs[j].Y[4];  // ViscousEnergyFlux
      add_total_flux_dir(F, st, Fe_j, n[j], j, Flux);
    }
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = wdetJb * Flux[c];
    if (ctx.is_implicit) {
      if (ok) {
        if (a.jac32) {
#pragma unroll
          for (int c = 0; c < 5; c++) a.jac32[fidx<Q>(a.jd, f, c, t)] = (float)q[c];
#pragma unroll
          for (int c = 0; c < 6; c++) a.jac32[fidx<Q>(a.jd, f, 5 + c, t)] = (float)sr[c];
        } else if (a.jac) {
#pragma unroll
          for (int c = 0; c < 5; c++) a.jac[fidx<Q>(a.jd, f, c, t)] = q[c];
#pragma unroll
          for (int c = 0; c < 6; c++) a.jac[fidx<Q>(a.jd, f, 5 + c, t)] = sr[c];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < 5; c++) v[c] *= -1;
    }
  }
};

// BoundaryIntegral_Jacobian (newtonian.h:392-431), in T
template <int VAR, typename T>
struct BoundaryIntegralJacobian {
  static constexpr bool kGrad = true;
  static constexpr int  kJacComps = 11;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int f, int t, bool ok, double dqd[5], double grefd[2][5],
                            double v[5]) {
    const ns::Gas<T> g = ns::Gas<T>::from(ctx.gas);
    T                wdetJb, dXdx[2][3], n[3], qi[5], sr[6], dq[5], gref[2][5];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    if (ok) {
#pragma unroll
      for (int c = 0; c < 5; c++) qi[c] = a.jac32 ? (T)a.jac32[fidx<Q>(a.jd, f, c, t)] : (T)a.jacr[fidx<Q>(a.jd, f, c, t)];
#pragma unroll
      for (int c = 0; c < 6; c++) sr[c] = a.jac32 ? (T)a.jac32[fidx<Q>(a.jd, f, 5 + c, t)] : (T)a.jacr[fidx<Q>(a.jd, f, 5 + c, t)];
    } else {
      padding_q<VAR>(qi);
#pragma unroll
      for (int c = 0; c < 6; c++) sr[c] = 0;
    }
#pragma unroll
    for (int c = 0; c < 5; c++) {
      dq[c]      = ok ? (T)dqd[c] : T(0);
      gref[0][c] = ok ? (T)grefd[0][c] : T(0);
      gref[1][c] = ok ? (T)grefd[1][c] : T(0);
    }
    const ns::St<T> s  = ns::state_from_q<VAR>(g, qi);
    const ns::St<T> ds = ns::state_from_q_fwd<VAR>(g, s, dq);
    ns::St<T>       gds[3];
#pragma unroll
    for (int k = 0; k < 3; k++) gds[k] = phys_gradient_boundary_dir<VAR>(g, s, gref, dXdx, k);
    T dsr[6], dst[3][3], st[3][3];
    strain_rate_from_states(gds, dsr);
    ns::stress_from_strain(g, dsr, dst);
    ns::stress_from_strain(g, sr, st);
    T dFlux[5] = {0., 0., 0., 0., 0.};
#pragma unroll
    for (int j = 0; j < 3; j++) {
      T dF[5];
      ns::flux_inviscid_fwd_dir(s, ds, j, dF);
      const T dFe_j = -s.Y[1] * dst[0][j] - ds.Y[1] * st[0][j] - s.Y[2] * dst[1][j] - ds.Y[2] * st[1][j] - s.Y[3] * dst[2][j] -
                      ds.Y[3] * st[2][j] - g.k * gds[j].Y[4];  // ViscousEnergyFlux_fwd
      add_total_flux_dir(dF, dst, dFe_j, n[j], j, dFlux);
    }
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = (double)(wdetJb * dFlux[c]);
  }
};

// Freestream (bc_freestream.h:14-50): v = wdetJb F_Riemann(s, S_infty, n); jac_data_sur = q when implicit
template <int VAR, bool HLLC>
struct Freestream {
  static constexpr bool kGrad = false;
  static constexpr int  kJacComps = 5;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiFreestreamContext &fctx, int f, int t, bool ok, double q[5], double (*)[5], double v[5]) {
    const CeedMiNewtonianContext &ctx = fctx.newt_ctx;
    const ns::Gas<double>         g   = ns::Gas<double>::from(ctx.gas);
    double                        wdetJb, dXdx[2][3], n[3];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    if (!ok) padding_q<VAR>(q);
    const ns::St<double> s = ns::state_from_q<VAR>(g, q);
    ns::St<double>       r;
#pragma unroll
    for (int c = 0; c < 5; c++) {
      r.U[c] = fctx.S_infty.U[c];
      r.Y[c] = fctx.S_infty.Y[c];
    }
    double F[5];
    if constexpr (HLLC) ns::riemann_hllc(g, s, r, n, F);
    else ns::riemann_hll(g, s, r, n, F);
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = wdetJb * F[c];
    if (ctx.is_implicit) {
      if (ok) {
        if (a.jac32) {
#pragma unroll
          for (int c = 0; c < 5; c++) a.jac32[fidx<Q>(a.jd, f, c, t)] = (float)q[c];
        } else if (a.jac) {
#pragma unroll
          for (int c = 0; c < 5; c++) a.jac[fidx<Q>(a.jd, f, c, t)] = q[c];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < 5; c++) v[c] *= -1;
    }
  }
};

// Freestream_Jacobian (bc_freestream.h:76-114): exterior state held fixed (dS_infty = 0), in T
template <int VAR, bool HLLC, typename T>
struct FreestreamJacobian {
  static constexpr bool kGrad = false;
  static constexpr int  kJacComps = 5;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiFreestreamContext &fctx, int f, int t, bool ok, double dqd[5], double (*)[5],
                            double v[5]) {
    const ns::Gas<T> g = ns::Gas<T>::from(fctx.newt_ctx.gas);
    T                wdetJb, dXdx[2][3], n[3], qi[5], dq[5];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    if (ok) {
#pragma unroll
      for (int c = 0; c < 5; c++) qi[c] = a.jac32 ? (T)a.jac32[fidx<Q>(a.jd, f, c, t)] : (T)a.jacr[fidx<Q>(a.jd, f, c, t)];
    } else {
      padding_q<VAR>(qi);
    }
#pragma unroll
    for (int c = 0; c < 5; c++) dq[c] = ok ? (T)dqd[c] : T(0);
    const ns::St<T> s  = ns::state_from_q<VAR>(g, qi);
    const ns::St<T> ds = ns::state_from_q_fwd<VAR>(g, s, dq);
    ns::St<T>       r, dr;
#pragma unroll
    for (int c = 0; c < 5; c++) {
      r.U[c]  = (T)fctx.S_infty.U[c];
      r.Y[c]  = (T)fctx.S_infty.Y[c];
      dr.U[c] = dr.Y[c] = T(0);
    }
    T dF[5];
    if constexpr (HLLC) ns::riemann_hllc_fwd(g, s, ds, r, dr, n, dF);
    else ns::riemann_hll_fwd(g, s, ds, r, dr, n, dF);
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = (double)(wdetJb * dF[c]);
  }
};

// ViscousEnergyFlux (newtonian_state.h:560-565), direction j
template <typename T>
__device__ __forceinline__ T viscous_energy_flux_dir(const ns::Gas<T> &g, const ns::St<T> &s, const ns::St<T> gs[3], const T st[3][3], int j) {
  return -s.Y[1] * st[0][j] - s.Y[2] * st[1][j] - s.Y[3] * st[2][j] - g.k * gs[j].Y[4];
}

// ViscousEnergyFlux_fwd (newtonian_state.h:567-576), direction j
template <typename T>
__device__ __forceinline__ T viscous_energy_flux_fwd_dir(const ns::Gas<T> &g, const ns::St<T> &s, const ns::St<T> &ds, const ns::St<T> gds[3],
                                                         const T st[3][3], const T dst[3][3], int j) {
  return -s.Y[1] * dst[0][j] - ds.Y[1] * st[0][j] - s.Y[2] * dst[1][j] - ds.Y[2] * st[1][j] - s.Y[3] * dst[2][j] - ds.Y[3] * st[2][j] -
         g.k * gds[j].Y[4];
}

// FluxTotal_RiemannBoundary (newtonian_state.h:498-509): normal Riemann flux plus the viscous flux through n
template <typename T>
__device__ __forceinline__ void flux_total_riemann_boundary(const T Fn[5], const T st[3][3], const T Fe[3], const T n[3], T Flux[5]) {
  Flux[0] = Fn[0];
#pragma unroll
  for (int k = 0; k < 3; k++) Flux[k + 1] = Fn[k + 1];
  Flux[4] = Fn[4];
#pragma unroll
  for (int j = 0; j < 3; j++) {
#pragma unroll
    for (int k = 0; k < 3; k++) Flux[k + 1] -= st[k][j] * n[j];
    Flux[4] += Fe[j] * n[j];
  }
}

// Softplus and the factor its forward mode returns (bc_outflow.h:18-27; the reference's Softplus_fwd returns the
// derivative dsoftplus/dx without multiplying by dx, and the outflow Jacobian uses that factor as is)
template <typename T>
__device__ __forceinline__ T softplus(T x, T width) {
  if (x > 40 * width) return x;
  return width * (T)log1p(exp(x / width));
}
template <typename T>
__device__ __forceinline__ T softplus_fwd(T x, T /*dx*/, T width) {
  if (x > 40 * width) return T(1);
  const T t = (T)exp(x / width);
  return t / (1 + t);
}

template <int Q, typename T>
__device__ __forceinline__ void store_q_strain(const VolumeArgs &a, int f, int t, bool ok, const double q[5], const T sr[6]) {
  if (!ok) return;
  if (a.jac32) {
#pragma unroll
    for (int c = 0; c < 5; c++) a.jac32[fidx<Q>(a.jd, f, c, t)] = (float)q[c];
#pragma unroll
    for (int c = 0; c < 6; c++) a.jac32[fidx<Q>(a.jd, f, 5 + c, t)] = (float)sr[c];
  } else if (a.jac) {
#pragma unroll
    for (int c = 0; c < 5; c++) a.jac[fidx<Q>(a.jd, f, c, t)] = q[c];
#pragma unroll
    for (int c = 0; c < 6; c++) a.jac[fidx<Q>(a.jd, f, 5 + c, t)] = (double)sr[c];
  }
}

template <int Q, typename T>
__device__ __forceinline__ void load_q_strain(const VolumeArgs &a, int f, int t, bool ok, T qi[5], T sr[6], int var) {
  if (ok) {
#pragma unroll
    for (int c = 0; c < 5; c++) qi[c] = a.jac32 ? (T)a.jac32[fidx<Q>(a.jd, f, c, t)] : (T)a.jacr[fidx<Q>(a.jd, f, c, t)];
#pragma unroll
    for (int c = 0; c < 6; c++) sr[c] = a.jac32 ? (T)a.jac32[fidx<Q>(a.jd, f, 5 + c, t)] : (T)a.jacr[fidx<Q>(a.jd, f, 5 + c, t)];
  } else {
#pragma unroll
    for (int c = 0; c < 5; c++) qi[c] = (var == ns::VAR_ENTROPY) ? (c == 4 ? T(-1) : T(0)) : T(1);
#pragma unroll
    for (int c = 0; c < 6; c++) sr[c] = 0;
  }
}

// Exterior state of the Riemann outflow (bc_outflow.h:52-61): the interior velocity (plus the softplus-regularized
// normal correction) at the prescribed pressure and temperature
template <typename T>
__device__ __forceinline__ void outflow_exterior_prim(const CeedMiOutflowContext &o, const ns::St<T> &s_int, const T n[3], T ye[5]) {
#pragma unroll
  for (int c = 0; c < 5; c++) ye[c] = s_int.Y[c];
  ye[0]        = (T)o.pressure;
  ye[4]        = (T)o.temperature;
  const T un   = ye[1] * n[0] + ye[2] * n[1] + ye[3] * n[2];
  const T proj = (T)(1 - o.recirc) * softplus(-un, (T)o.softplus_velocity);
#pragma unroll
  for (int j = 0; j < 3; j++) ye[1 + j] += n[j] * proj;
}

// RiemannOutflow (bc_outflow.h:37-89): v = wdetJb (HLLC(s_int, s_ext) + viscous flux) . n; jac_data_sur = {q, strain
// rate} when implicit, v negated when explicit
template <int VAR>
struct RiemannOutflow {
  static constexpr bool kGrad     = true;
  static constexpr int  kJacComps = 11;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiOutflowContext &octx, int f, int t, bool ok, double q[5], double gref[2][5],
                            double v[5]) {
    const CeedMiNewtonianContext &ctx = octx.newt_ctx;
    const ns::Gas<double>         g   = ns::Gas<double>::from(ctx.gas);
    double                        wdetJb, dXdx[2][3], n[3];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    if (!ok) {
      padding_q<VAR>(q);
#pragma unroll
      for (int c = 0; c < 5; c++) gref[0][c] = gref[1][c] = 0;
    }
    const ns::St<double> s_int = ns::state_from_q<VAR>(g, q);
    double               ye[5];
    outflow_exterior_prim(octx, s_int, n, ye);
    const ns::St<double> s_ext = ns::state_from_q<ns::VAR_PRIM>(g, ye);
    ns::St<double>       gs[3];
#pragma unroll
    for (int k = 0; k < 3; k++) gs[k] = phys_gradient_boundary_dir<VAR>(g, s_int, gref, dXdx, k);
    double sr[6], st[3][3], Fe[3], Fn[5], Flux[5];
    strain_rate_from_states(gs, sr);
    ns::stress_from_strain(g, sr, st);
#pragma unroll
    for (int j = 0; j < 3; j++) Fe[j] = viscous_energy_flux_dir(g, s_int, gs, st, j);
    ns::riemann_hllc(g, s_int, s_ext, n, Fn);
    flux_total_riemann_boundary(Fn, st, Fe, n, Flux);
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = wdetJb * Flux[c];
    if (ctx.is_implicit) store_q_strain<Q>(a, f, t, ok, q, sr);
    else
#pragma unroll
      for (int c = 0; c < 5; c++) v[c] *= -1;
  }
};

// RiemannOutflow_Jacobian (bc_outflow.h:106-164), in T: exterior pressure and temperature held fixed
template <int VAR, typename T>
struct RiemannOutflowJacobian {
  static constexpr bool kGrad     = true;
  static constexpr int  kJacComps = 11;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiOutflowContext &octx, int f, int t, bool ok, double dqd[5], double grefd[2][5],
                            double v[5]) {
    const ns::Gas<T> g = ns::Gas<T>::from(octx.newt_ctx.gas);
    T                wdetJb, dXdx[2][3], n[3], qi[5], sr[6], dq[5], gref[2][5];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    load_q_strain<Q>(a, f, t, ok, qi, sr, VAR);
#pragma unroll
    for (int c = 0; c < 5; c++) {
      dq[c]      = ok ? (T)dqd[c] : T(0);
      gref[0][c] = ok ? (T)grefd[0][c] : T(0);
      gref[1][c] = ok ? (T)grefd[1][c] : T(0);
    }
    const ns::St<T> s_int  = ns::state_from_q<VAR>(g, qi);
    const ns::St<T> ds_int = ns::state_from_q_fwd<VAR>(g, s_int, dq);
    T               ye[5], dye[5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
      ye[c]  = s_int.Y[c];
      dye[c] = ds_int.Y[c];
    }
    ye[0] = (T)octx.pressure, ye[4] = (T)octx.temperature;
    dye[0] = 0, dye[4] = 0;
    const T un    = s_int.Y[1] * n[0] + s_int.Y[2] * n[1] + s_int.Y[3] * n[2];
    const T dun   = ds_int.Y[1] * n[0] + ds_int.Y[2] * n[1] + ds_int.Y[3] * n[2];
    const T proj  = (T)(1 - octx.recirc) * softplus(-un, (T)octx.softplus_velocity);
    const T dproj = (T)(1 - octx.recirc) * softplus_fwd(-un, -dun, (T)octx.softplus_velocity);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      ye[1 + j] += n[j] * proj;
      dye[1 + j] += n[j] * dproj;
    }
    const ns::St<T> s_ext  = ns::state_from_q<ns::VAR_PRIM>(g, ye);
    const ns::St<T> ds_ext = ns::state_from_q_fwd<ns::VAR_PRIM>(g, s_ext, dye);
    ns::St<T>       gds[3];
#pragma unroll
    for (int k = 0; k < 3; k++) gds[k] = phys_gradient_boundary_dir<VAR>(g, s_int, gref, dXdx, k);
    T dsr[6], dst[3][3], st[3][3], dFe[3], dFn[5], dFlux[5];
    strain_rate_from_states(gds, dsr);
    ns::stress_from_strain(g, dsr, dst);
    ns::stress_from_strain(g, sr, st);
#pragma unroll
    for (int j = 0; j < 3; j++) dFe[j] = viscous_energy_flux_fwd_dir(g, s_int, ds_int, gds, st, dst, j);
    ns::riemann_hllc_fwd(g, s_int, ds_int, s_ext, ds_ext, n, dFn);
    flux_total_riemann_boundary(dFn, dst, dFe, n, dFlux);
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = (double)(wdetJb * dFlux[c]);
  }
};

// PressureOutflow (bc_outflow.h:185-229): the interior state at the prescribed pressure, v = -wdetJb F_total . n with
// wdetJb negated when implicit; jac_data_sur = {q, strain rate} when implicit
template <int VAR>
struct PressureOutflow {
  static constexpr bool kGrad     = true;
  static constexpr int  kJacComps = 11;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiOutflowContext &octx, int f, int t, bool ok, double q[5], double gref[2][5],
                            double v[5]) {
    const CeedMiNewtonianContext &ctx = octx.newt_ctx;
    const ns::Gas<double>         g   = ns::Gas<double>::from(ctx.gas);
    double                        wdetJb, dXdx[2][3], n[3];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    if (!ok) {
      padding_q<VAR>(q);
#pragma unroll
      for (int c = 0; c < 5; c++) gref[0][c] = gref[1][c] = 0;
    }
    ns::St<double> s = ns::state_from_q<VAR>(g, q);
    s.Y[0]           = octx.pressure;
    wdetJb *= ctx.is_implicit ? -1. : 1.;
    ns::St<double> gs[3];
#pragma unroll
    for (int k = 0; k < 3; k++) gs[k] = phys_gradient_boundary_dir<VAR>(g, s, gref, dXdx, k);
    double sr[6], st[3][3];
    strain_rate_from_states(gs, sr);
    ns::stress_from_strain(g, sr, st);
    double Flux[5] = {0., 0., 0., 0., 0.};
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double F[5];
      ns::flux_inviscid_dir(s, j, F);
      add_total_flux_dir(F, st, viscous_energy_flux_dir(g, s, gs, st, j), n[j], j, Flux);
    }
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = -wdetJb * Flux[c];
    if (ctx.is_implicit) store_q_strain<Q>(a, f, t, ok, q, sr);
  }
};

// PressureOutflow_Jacobian (bc_outflow.h:246-293), in T: dq's pressure part dropped after the state tangent
template <int VAR, typename T>
struct PressureOutflowJacobian {
  static constexpr bool kGrad     = true;
  static constexpr int  kJacComps = 11;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiOutflowContext &octx, int f, int t, bool ok, double dqd[5], double grefd[2][5],
                            double v[5]) {
    const ns::Gas<T> g = ns::Gas<T>::from(octx.newt_ctx.gas);
    T                wdetJb, dXdx[2][3], n[3], qi[5], sr[6], dq[5], gref[2][5];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    wdetJb *= octx.newt_ctx.is_implicit ? T(-1) : T(1);
    load_q_strain<Q>(a, f, t, ok, qi, sr, VAR);
#pragma unroll
    for (int c = 0; c < 5; c++) {
      dq[c]      = ok ? (T)dqd[c] : T(0);
      gref[0][c] = ok ? (T)grefd[0][c] : T(0);
      gref[1][c] = ok ? (T)grefd[1][c] : T(0);
    }
    ns::St<T> s  = ns::state_from_q<VAR>(g, qi);
    ns::St<T> ds = ns::state_from_q_fwd<VAR>(g, s, dq);
    s.Y[0]       = (T)octx.pressure;
    ds.Y[0]      = 0;
    ns::St<T> gds[3];
#pragma unroll
    for (int k = 0; k < 3; k++) gds[k] = phys_gradient_boundary_dir<VAR>(g, s, gref, dXdx, k);
    T dsr[6], dst[3][3], st[3][3];
    strain_rate_from_states(gds, dsr);
    ns::stress_from_strain(g, dsr, dst);
    ns::stress_from_strain(g, sr, st);
    T dFlux[5] = {0., 0., 0., 0., 0.};
#pragma unroll
    for (int j = 0; j < 3; j++) {
      T dF[5];
      ns::flux_inviscid_fwd_dir(s, ds, j, dF);
      add_total_flux_dir(dF, dst, viscous_energy_flux_fwd_dir(g, s, ds, gds, st, dst, j), n[j], j, dFlux);
    }
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = (double)(-wdetJb * dFlux[c]);
  }
};

// Slip (bc_slip.h:11-46): HLLC flux against the mirror state (normal velocity reflected); jac_data_sur = q when
// implicit, v negated when explicit
template <int VAR>
struct Slip {
  static constexpr bool kGrad     = false;
  static constexpr int  kJacComps = 5;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int f, int t, bool ok, double q[5], double (*)[5], double v[5]) {
    const ns::Gas<double> g = ns::Gas<double>::from(ctx.gas);
    double                wdetJb, dXdx[2][3], n[3];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    if (!ok) padding_q<VAR>(q);
    const ns::St<double> s  = ns::state_from_q<VAR>(g, q);
    const double         vn = s.Y[1] * n[0] + s.Y[2] * n[1] + s.Y[3] * n[2];
    double               yr[5];
    yr[0] = s.Y[0];
#pragma unroll
    for (int j = 0; j < 3; j++) yr[1 + j] = s.Y[1 + j] - 2. * n[j] * vn;
    yr[4]                    = s.Y[4];
    const ns::St<double> s_r = ns::state_from_q<ns::VAR_PRIM>(g, yr);
    double               F[5];
    ns::riemann_hllc(g, s, s_r, n, F);
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = wdetJb * F[c];
    if (ctx.is_implicit) {
      if (ok) {
        if (a.jac32) {
#pragma unroll
          for (int c = 0; c < 5; c++) a.jac32[fidx<Q>(a.jd, f, c, t)] = (float)q[c];
        } else if (a.jac) {
#pragma unroll
          for (int c = 0; c < 5; c++) a.jac[fidx<Q>(a.jd, f, c, t)] = q[c];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < 5; c++) v[c] *= -1;
    }
  }
};

// Slip_Jacobian (bc_slip.h:60-99), in T
template <int VAR, typename T>
struct SlipJacobian {
  static constexpr bool kGrad     = false;
  static constexpr int  kJacComps = 5;
  template <int Q>
  __device__ static void qf(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, int f, int t, bool ok, double dqd[5], double (*)[5],
                            double v[5]) {
    const ns::Gas<T> g = ns::Gas<T>::from(ctx.gas);
    T                wdetJb, dXdx[2][3], n[3], qi[5], dq[5];
    unpack_face_qdata<Q>(a, f, t, ok, wdetJb, dXdx, n);
    if (ok) {
#pragma unroll
      for (int c = 0; c < 5; c++) qi[c] = a.jac32 ? (T)a.jac32[fidx<Q>(a.jd, f, c, t)] : (T)a.jacr[fidx<Q>(a.jd, f, c, t)];
    } else {
      padding_q<VAR>(qi);
    }
#pragma unroll
    for (int c = 0; c < 5; c++) dq[c] = ok ? (T)dqd[c] : T(0);
    const ns::St<T> s   = ns::state_from_q<VAR>(g, qi);
    const ns::St<T> ds  = ns::state_from_q_fwd<VAR>(g, s, dq);
    const T         vn  = s.Y[1] * n[0] + s.Y[2] * n[1] + s.Y[3] * n[2];
    const T         dvn = ds.Y[1] * n[0] + ds.Y[2] * n[1] + ds.Y[3] * n[2];
    T               yr[5], dyr[5];
    yr[0]  = s.Y[0];
    dyr[0] = ds.Y[0];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      yr[1 + j]  = s.Y[1 + j] - T(2) * n[j] * vn;
      dyr[1 + j] = ds.Y[1 + j] - T(2) * n[j] * dvn;
    }
    yr[4]                 = s.Y[4];
    dyr[4]                = ds.Y[4];
    const ns::St<T> s_r   = ns::state_from_q<ns::VAR_PRIM>(g, yr);
    const ns::St<T> ds_r  = ns::state_from_q_fwd<ns::VAR_PRIM>(g, s_r, dyr);
    T               dF[5];
    ns::riemann_hllc_fwd(g, s, ds, s_r, ds_r, n, dF);
#pragma unroll
    for (int c = 0; c < 5; c++) v[c] = (double)(wdetJb * dF[c]);
  }
};

}  // namespace fop
