// Fused element kernels for HONEE's Newtonian volume operators on MI355X (gfx950).
//
// One kernel launch applies a whole volume sub-operator: per element it gathers the L-vector through the element
// restriction, interpolates values and reference gradients to the quadrature points in LDS, evaluates the QFunction
// in registers, applies the basis transpose and scatter-adds into the output L-vector.  Quadrature data (qdata,
// jac_data) is streamed in the backend's strided layout.  Each element owns Q^3 threads (one per quadrature point) and
// a block holds EPB elements; blocks loop over element batches with a grid stride.
#include <hip/hip_runtime.h>

#include "basis3d.hpp"
#include "kernels.hpp"
#include "ns_physics.hpp"

namespace {

using basis3::Dims;
using basis3::Elem;

template <int Q>
constexpr int elems_per_block() {
  return (Q * Q * Q >= 64) ? 1 : 64 / (Q * Q * Q);
}

// Strided access to backend quadrature data
struct StridedQ {
  const double *p;
  float        *p32w;
  double       *pw;
  long long     sn, sc, se;
};

__device__ __forceinline__ long long qidx(const QDataLayout &l, int e, int c, int i) { return (long long)i * l.sn + (long long)c * l.sc + (long long)e * l.se; }

// Load block-shared basis tables into LDS
template <int P, int Q>
__device__ __forceinline__ void load_tables(const BasisTables &bt, double *sB, double *sDc) {
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) sB[i] = bt.B[i];
  for (int i = threadIdx.x; i < Q * Q; i += blockDim.x) sDc[i] = bt.Dc[i];
}

// ------------------------------------------------------------------------------------------------------------------
// Setup: qdata = {w detJ, dXdx} from the coordinate field (qfunctions/setupgeo.h:42-56).  The coordinate basis
// (Px nodes) is applied with B and its (non-collocated) derivative directly: coordinates are trilinear.
// ------------------------------------------------------------------------------------------------------------------
template <int PX, int Q>
__global__ void __launch_bounds__(256) k_setup_qdata(SetupArgs a) {
  constexpr int Q3 = Q * Q * Q;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)a.nelem * Q3) return;
  const int e = (int)(gid / Q3), i = (int)(gid % Q3);
  const int qx = i % Q, qy = (i / Q) % Q, qz = i / (Q * Q);
  double    J[3][3] = {{0}};
#pragma unroll
  for (int c = 0; c < PX; c++)
#pragma unroll
    for (int b = 0; b < PX; b++)
#pragma unroll
      for (int aa = 0; aa < PX; aa++) {
        const int    n   = aa + PX * (b + PX * c);
        const int    off = a.xoffsets[(long long)e * PX * PX * PX + n];
        const double bx = a.Bx[qx * PX + aa], by = a.Bx[qy * PX + b], bz = a.Bx[qz * PX + c];
        const double dx = a.Dx[qx * PX + aa], dy = a.Dx[qy * PX + b], dz = a.Dx[qz * PX + c];
#pragma unroll
        for (int comp = 0; comp < 3; comp++) {
          const double xv = a.x[off + comp * a.xcomp_stride];
          J[comp][0] += dx * by * bz * xv;
          J[comp][1] += bx * dy * bz * xv;
          J[comp][2] += bx * by * dz * xv;
        }
      }
  // MatInv3 (qfunctions/utils.h:201-216)
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
  const double w = a.w1d[qx] * a.w1d[qy] * a.w1d[qz];
  a.qdata[qidx(a.qd, e, 0, i)] = w * d;
#pragma unroll
  for (int k = 0; k < 9; k++) a.qdata[qidx(a.qd, e, 1 + k, i)] = (&Ai[0][0])[k] * s;
}

// ------------------------------------------------------------------------------------------------------------------
// IFunction_Newtonian (qfunctions/newtonian.h:205-257) with IFunction_Newtonian_Integrand
// (newtonian_integrands.h:87-115) and the internal damping layer (newtonian_integrands.h:40-48).
// ------------------------------------------------------------------------------------------------------------------
template <int P, int Q, int VAR>
__global__ void __launch_bounds__(elems_per_block<Q>() * Q * Q * Q) k_ifunction(VolumeArgs a, const CeedMiNewtonianContext ctx) {
  using Dm           = Dims<P, Q>;
  constexpr int EPB  = elems_per_block<Q>();
  constexpr int Q3   = Dm::Q3;
  __shared__ double sB[Q * P], sDc[Q * Q];
  __shared__ double s0[EPB][Dm::MAX3], s1[EPB][Dm::MAX3], s2[EPB][Dm::MAX3];
  load_tables<P, Q>(a.bt, sB, sDc);
  const int      le = threadIdx.x / Q3, t = threadIdx.x % Q3;
  Elem<P, Q>     el(sB, sDc, s0[le], s1[le], s2[le], t);
  const ns::Gas<double> g = ns::Gas<double>::from(ctx.gas);
  const long long nbatch  = ((long long)a.nelem + EPB - 1) / EPB;

  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long e_  = bt * EPB + le;
    const bool      ok  = e_ < a.nelem;
    const int       e   = ok ? (int)e_ : 0;
    const int       off = (t < Dm::P3) ? a.offsets[(long long)e * Dm::P3 + t] : 0;

    double q[5], gref[3][5], qdot[5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
      const double u = (t < Dm::P3) ? a.u[off + c * a.comp_stride] : 0.0;
      q[c]           = el.interp(u);
      double gg[3];
      el.grad(q[c], gg);
      gref[0][c] = gg[0];
      gref[1][c] = gg[1];
      gref[2][c] = gg[2];
    }
#pragma unroll
    for (int c = 0; c < 5; c++) {
      const double u = (t < Dm::P3) ? a.u2[off + c * a.comp_stride] : 0.0;
      qdot[c]        = el.interp(u);
    }
    double x0 = 0;
    if (ctx.idl_enable) {  // x coordinate at the quadrature point (trilinear coordinate field)
      constexpr int PX = 2;
      const int     qx = t % Q, qy = (t / Q) % Q, qz = t / (Q * Q);
#pragma unroll
      for (int n = 0; n < 8; n++) {
        const int xo = a.xoffsets[(long long)e * 8 + n];
        x0 += a.Bx[qx * PX + (n & 1)] * a.Bx[qy * PX + ((n >> 1) & 1)] * a.Bx[qz * PX + (n >> 2)] * a.x[xo];
      }
    }

    // ---- QFunction at quadrature point t
    double wdetJ, dXdx[3][3];
    wdetJ = ok ? a.qdata[qidx(a.qd, e, 0, t)] : 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) (&dXdx[0][0])[k] = ok ? a.qdata[qidx(a.qd, e, 1 + k, t)] : 0.0;
    if (!ok) {
      wdetJ = 0;
      for (int k = 0; k < 9; k++) (&dXdx[0][0])[k] = (k % 4 == 0) ? 1.0 : 0.0;
      for (int c = 0; c < 5; c++) {  // a harmless physical state for padding threads
        q[c] = (VAR == ns::VAR_ENTROPY) ? (c == 4 ? -1.0 : 0.0) : 1.0;
        qdot[c] = 0;
        for (int d = 0; d < 3; d++) gref[d][c] = 0;
      }
    }
    const ns::St<double> s  = ns::state_from_q<VAR>(g, q);
    const ns::St<double> sd = ns::state_from_q_fwd<VAR>(g, s, qdot);
    ns::St<double>       gs[3];
    ns::phys_gradient<VAR>(g, s, gref, dXdx, gs);

    double sr[6], st[3][3];
    ns::km_strain_rate(gs, sr);
    ns::stress_from_strain(g, sr, st);
    double Fe[3];
#pragma unroll
    for (int i = 0; i < 3; i++) Fe[i] = -s.Y[1] * st[0][i] - s.Y[2] * st[1][i] - s.Y[3] * st[2][i] - g.k * gs[i].Y[4];
    double gv[5][3];  // grad_v_i[var][dir]
#pragma unroll
    for (int j = 0; j < 3; j++) {  // -FluxTotal (newtonian_state.h:477-484)
      double F[5];
      ns::flux_inviscid_dir(s, j, F);
      gv[0][j] = -F[0];
#pragma unroll
      for (int k = 0; k < 3; k++) gv[k + 1][j] = -(F[1 + k] - st[k][j]);
      gv[4][j] = -(F[4] + Fe[j]);
    }
    const double bf[5] = {0, s.U[0] * ctx.g[0], s.U[0] * ctx.g[1], s.U[0] * ctx.g[2],
                          s.U[1] * ctx.g[0] + s.U[2] * ctx.g[1] + s.U[3] * ctx.g[2]};
    double       vi[5];
#pragma unroll
    for (int j = 0; j < 5; j++) vi[j] = sd.U[j] - bf[j];
    double Tau[3];
    ns::tau_diag_prim<double>(ctx.tau_coeffs, g, s, dXdx, ctx.dt, Tau);
    if (ctx.stabilization != ns::STAB_NONE) {  // Stabilization (stabilization.h:41-58); divFdiff = 0
      double r[5];
      ns::flux_inviscid_strong(s, gs, r);
      if (ctx.stabilization == ns::STAB_SUPG)
#pragma unroll
        for (int j = 0; j < 5; j++) r[j] += sd.U[j] - bf[j];
      ns::add_stab_matrix(g, s, Tau, r, gv);
    }
    double sigma = 0;
    if (ctx.idl_enable) {  // InternalDampingLayer_IFunction_Integrand
      sigma = ns::linear_ramp<double>(ctx.idl_amplitude, ctx.idl_length, ctx.idl_start, x0);
      ns::St<double> ds;
      ds.Y[0] = (s.Y[0] - ctx.idl_pressure) * sigma;
      for (int i = 1; i < 5; i++) ds.Y[i] = 0 * sigma;
      ns::cons_from_prim_fwd(g, s, ds.Y, ds.U);
#pragma unroll
      for (int j = 0; j < 5; j++) vi[j] += ds.U[j];
    }
    // jac_data: q, strain rate (KM), Tau_d (+ sigma)  (newtonian.h:251-254)
    if (ok) {
      if (a.jac32) {
        float *J = a.jac32;
#pragma unroll
        for (int j = 0; j < 5; j++) J[qidx(a.jd, e, j, t)] = (float)q[j];
#pragma unroll
        for (int j = 0; j < 6; j++) J[qidx(a.jd, e, 5 + j, t)] = (float)sr[j];
#pragma unroll
        for (int j = 0; j < 3; j++) J[qidx(a.jd, e, 11 + j, t)] = (float)Tau[j];
        if (ctx.idl_enable) J[qidx(a.jd, e, 14, t)] = (float)sigma;
      } else if (a.jac) {
        double *J = a.jac;
#pragma unroll
        for (int j = 0; j < 5; j++) J[qidx(a.jd, e, j, t)] = q[j];
#pragma unroll
        for (int j = 0; j < 6; j++) J[qidx(a.jd, e, 5 + j, t)] = sr[j];
#pragma unroll
        for (int j = 0; j < 3; j++) J[qidx(a.jd, e, 11 + j, t)] = Tau[j];
        if (ctx.idl_enable) J[qidx(a.jd, e, 14, t)] = sigma;
      }
    }
    // ---- transpose and scatter-add, one component at a time
#pragma unroll
    for (int c = 0; c < 5; c++) {
      double gw[3];
#pragma unroll
      for (int k = 0; k < 3; k++) gw[k] = wdetJ * (gv[c][0] * dXdx[k][0] + gv[c][1] * dXdx[k][1] + gv[c][2] * dXdx[k][2]);
      const double r = el.template transpose<true, true>(wdetJ * vi[c], gw);
      if (ok && t < Dm::P3) atomicAdd(&a.v[off + c * a.comp_stride], r);
    }
  }
}

// ------------------------------------------------------------------------------------------------------------------
// IJacobian_Newtonian (newtonian.h:272-311) with IJacobian_Newtonian_Integrand (newtonian_integrands.h:134-166),
// computed in T (double or float).  Stored jac_data supplies q, the strain rate and Tau_d; the IDL term uses the
// stored sigma (newtonian_integrands.h:61-65).
// ------------------------------------------------------------------------------------------------------------------
template <int P, int Q, int VAR, typename T>
__global__ void __launch_bounds__(elems_per_block<Q>() * Q * Q * Q) k_ijacobian(VolumeArgs a, const CeedMiNewtonianContext ctx) {
  using Dm          = Dims<P, Q>;
  constexpr int EPB = elems_per_block<Q>();
  constexpr int Q3  = Dm::Q3;
  __shared__ double sB[Q * P], sDc[Q * Q];
  __shared__ double s0[EPB][Dm::MAX3], s1[EPB][Dm::MAX3], s2[EPB][Dm::MAX3];
  load_tables<P, Q>(a.bt, sB, sDc);
  const int        le = threadIdx.x / Q3, t = threadIdx.x % Q3;
  Elem<P, Q>       el(sB, sDc, s0[le], s1[le], s2[le], t);
  const ns::Gas<T> g      = ns::Gas<T>::from(ctx.gas);
  const long long  nbatch = ((long long)a.nelem + EPB - 1) / EPB;
  const T          shift  = (T)ctx.ijacobian_time_shift;

  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long e_  = bt * EPB + le;
    const bool      ok  = e_ < a.nelem;
    const int       e   = ok ? (int)e_ : 0;
    const int       off = (t < Dm::P3) ? a.offsets[(long long)e * Dm::P3 + t] : 0;

    T dq[5], gref[3][5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
      const double u = (t < Dm::P3) ? a.u[off + c * a.comp_stride] : 0.0;
      const double v = el.interp(u);
      double       gg[3];
      el.grad(v, gg);
      dq[c]      = (T)v;
      gref[0][c] = (T)gg[0];
      gref[1][c] = (T)gg[1];
      gref[2][c] = (T)gg[2];
    }
    T wdetJ = 0, dXdx[3][3], qi[5], sr[6], Tau[3], sigma = 0;
    if (ok) {
      wdetJ = (T)a.qdata[qidx(a.qd, e, 0, t)];
#pragma unroll
      for (int k = 0; k < 9; k++) (&dXdx[0][0])[k] = (T)a.qdata[qidx(a.qd, e, 1 + k, t)];
      if (a.jac32) {
#pragma unroll
        for (int j = 0; j < 5; j++) qi[j] = (T)a.jac32[qidx(a.jd, e, j, t)];
#pragma unroll
        for (int j = 0; j < 6; j++) sr[j] = (T)a.jac32[qidx(a.jd, e, 5 + j, t)];
#pragma unroll
        for (int j = 0; j < 3; j++) Tau[j] = (T)a.jac32[qidx(a.jd, e, 11 + j, t)];
        if (ctx.idl_enable) sigma = (T)a.jac32[qidx(a.jd, e, 14, t)];
      } else {
#pragma unroll
        for (int j = 0; j < 5; j++) qi[j] = (T)a.jacr[qidx(a.jd, e, j, t)];
#pragma unroll
        for (int j = 0; j < 6; j++) sr[j] = (T)a.jacr[qidx(a.jd, e, 5 + j, t)];
#pragma unroll
        for (int j = 0; j < 3; j++) Tau[j] = (T)a.jacr[qidx(a.jd, e, 11 + j, t)];
        if (ctx.idl_enable) sigma = (T)a.jacr[qidx(a.jd, e, 14, t)];
      }
    } else {
      for (int k = 0; k < 9; k++) (&dXdx[0][0])[k] = (k % 4 == 0) ? T(1) : T(0);
      for (int j = 0; j < 5; j++) {
        qi[j] = (VAR == ns::VAR_ENTROPY) ? (j == 4 ? T(-1) : T(0)) : T(1);
        dq[j] = 0;
        for (int d = 0; d < 3; d++) gref[d][j] = 0;
      }
      for (int j = 0; j < 6; j++) sr[j] = 0;
      for (int j = 0; j < 3; j++) Tau[j] = 0;
    }
    const ns::St<T> s  = ns::state_from_q<VAR>(g, qi);
    const ns::St<T> ds = ns::state_from_q_fwd<VAR>(g, s, dq);
    ns::St<T>       gds[3];
    ns::phys_gradient<VAR>(g, s, gref, dXdx, gds);

    T dsr[6], st[3][3], dst[3][3];
    ns::km_strain_rate(gds, dsr);
    ns::stress_from_strain(g, dsr, dst);
    ns::stress_from_strain(g, sr, st);
    T dFe[3];
#pragma unroll
    for (int i = 0; i < 3; i++)  // ViscousEnergyFlux_fwd (newtonian_state.h:567-573)
      dFe[i] = -s.Y[1] * dst[0][i] - ds.Y[1] * st[0][i] - s.Y[2] * dst[1][i] - ds.Y[2] * st[1][i] - s.Y[3] * dst[2][i] - ds.Y[3] * st[2][i] -
               g.k * gds[i].Y[4];
    T gv[5][3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      T dF[5];
      ns::flux_inviscid_fwd_dir(s, ds, j, dF);
      gv[0][j] = -dF[0];
#pragma unroll
      for (int k = 0; k < 3; k++) gv[k + 1][j] = -(dF[1 + k] - dst[k][j]);
      gv[4][j] = -(dF[4] + dFe[j]);
    }
    const T gx = (T)ctx.g[0], gy = (T)ctx.g[1], gz = (T)ctx.g[2];
    const T dbf[5] = {0, ds.U[0] * gx, ds.U[0] * gy, ds.U[0] * gz, ds.U[1] * gx + ds.U[2] * gy + ds.U[3] * gz};
    T       vi[5], dUdot[5];
#pragma unroll
    for (int j = 0; j < 5; j++) {
      dUdot[j] = shift * ds.U[j];
      vi[j]    = dUdot[j] - dbf[j];
    }
    if (ctx.stabilization != ns::STAB_NONE) {
      T r[5];
      ns::flux_inviscid_strong(s, gds, r);
      if (ctx.stabilization == ns::STAB_SUPG)
#pragma unroll
        for (int j = 0; j < 5; j++) r[j] += dUdot[j] - dbf[j];
      ns::add_stab_matrix(g, s, Tau, r, gv);
    }
    if (ctx.idl_enable) {
      ns::St<T> dd;
      dd.Y[0] = ds.Y[0] * sigma;
      for (int i = 1; i < 5; i++) dd.Y[i] = 0 * sigma;
      ns::cons_from_prim_fwd(g, s, dd.Y, dd.U);
#pragma unroll
      for (int j = 0; j < 5; j++) vi[j] += dd.U[j];
    }
#pragma unroll
    for (int c = 0; c < 5; c++) {
      double gw[3];
#pragma unroll
      for (int k = 0; k < 3; k++) gw[k] = (double)(wdetJ * (gv[c][0] * dXdx[k][0] + gv[c][1] * dXdx[k][1] + gv[c][2] * dXdx[k][2]));
      const double r = el.template transpose<true, true>((double)(wdetJ * vi[c]), gw);
      if (ok && t < Dm::P3) atomicAdd(&a.v[off + c * a.comp_stride], r);
    }
  }
}

template <typename K>
int grid_for(K kernel, int block, long long nbatch) {
  int dev = 0, ncu = 0, per_cu = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
  long long g = (long long)ncu * (per_cu > 0 ? per_cu : 1) * 8;
  if (g > nbatch) g = nbatch;
  if (g < 1) g = 1;
  return (int)g;
}

template <int P, int Q, int VAR>
hipError_t launch_ifunction_pq(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream) {
  constexpr int EPB   = elems_per_block<Q>();
  const int     block = EPB * Q * Q * Q;
  auto          k     = k_ifunction<P, Q, VAR>;
  const int     grid  = grid_for(k, block, ((long long)a.nelem + EPB - 1) / EPB);
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, stream, a, ctx);
  return hipGetLastError();
}

template <int P, int Q, int VAR, typename T>
hipError_t launch_ijacobian_pq(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream) {
  constexpr int EPB   = elems_per_block<Q>();
  const int     block = EPB * Q * Q * Q;
  auto          k     = k_ijacobian<P, Q, VAR, T>;
  const int     grid  = grid_for(k, block, ((long long)a.nelem + EPB - 1) / EPB);
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, stream, a, ctx);
  return hipGetLastError();
}

template <int P, int Q>
hipError_t launch_ifunction_var(int var, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  switch (var) {
    case 0: return launch_ifunction_pq<P, Q, 0>(a, ctx, s);
    case 1: return launch_ifunction_pq<P, Q, 1>(a, ctx, s);
    default: return launch_ifunction_pq<P, Q, 2>(a, ctx, s);
  }
}

template <int P, int Q>
hipError_t launch_ijacobian_var(int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  if (fp32) {
    switch (var) {
      case 0: return launch_ijacobian_pq<P, Q, 0, float>(a, ctx, s);
      case 1: return launch_ijacobian_pq<P, Q, 1, float>(a, ctx, s);
      default: return launch_ijacobian_pq<P, Q, 2, float>(a, ctx, s);
    }
  }
  switch (var) {
    case 0: return launch_ijacobian_pq<P, Q, 0, double>(a, ctx, s);
    case 1: return launch_ijacobian_pq<P, Q, 1, double>(a, ctx, s);
    default: return launch_ijacobian_pq<P, Q, 2, double>(a, ctx, s);
  }
}

}  // namespace

#define HONEE_PQ_DISPATCH(P_, Q_, CALL) \
  if (P == P_ && Q == Q_) return CALL;

hipError_t launch_setup_qdata(int Px, int Q, const SetupArgs &a, hipStream_t stream) {
  const long long n     = (long long)a.nelem * Q * Q * Q;
  const int       block = 256;
  const int       grid  = (int)((n + block - 1) / block);
  if (Px != 2) return hipErrorInvalidValue;
  switch (Q) {
    case 2: hipLaunchKernelGGL((k_setup_qdata<2, 2>), dim3(grid), dim3(block), 0, stream, a); break;
    case 3: hipLaunchKernelGGL((k_setup_qdata<2, 3>), dim3(grid), dim3(block), 0, stream, a); break;
    case 4: hipLaunchKernelGGL((k_setup_qdata<2, 4>), dim3(grid), dim3(block), 0, stream, a); break;
    case 5: hipLaunchKernelGGL((k_setup_qdata<2, 5>), dim3(grid), dim3(block), 0, stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

bool volume_pq_supported(int P, int Q) {
  return (P == 2 && Q == 2) || (P == 3 && Q == 3) || (P == 4 && Q == 4) || (P == 5 && Q == 5);
}

hipError_t launch_ifunction(int P, int Q, int var, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  HONEE_PQ_DISPATCH(2, 2, (launch_ifunction_var<2, 2>(var, a, ctx, s)));
  HONEE_PQ_DISPATCH(3, 3, (launch_ifunction_var<3, 3>(var, a, ctx, s)));
  HONEE_PQ_DISPATCH(4, 4, (launch_ifunction_var<4, 4>(var, a, ctx, s)));
  HONEE_PQ_DISPATCH(5, 5, (launch_ifunction_var<5, 5>(var, a, ctx, s)));
  return hipErrorInvalidValue;
}

hipError_t launch_ijacobian(int P, int Q, int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  HONEE_PQ_DISPATCH(2, 2, (launch_ijacobian_var<2, 2>(var, fp32, a, ctx, s)));
  HONEE_PQ_DISPATCH(3, 3, (launch_ijacobian_var<3, 3>(var, fp32, a, ctx, s)));
  HONEE_PQ_DISPATCH(4, 4, (launch_ijacobian_var<4, 4>(var, fp32, a, ctx, s)));
  HONEE_PQ_DISPATCH(5, 5, (launch_ijacobian_var<5, 5>(var, fp32, a, ctx, s)));
  return hipErrorInvalidValue;
}
