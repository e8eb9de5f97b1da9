// Kernels of the data-driven subgrid-stress (SGS) closure and its nodal inputs on MI355X (gfx950).
//
// k_field<P, Q, Op> applies one volume operator whose fields have their own restrictions: an optional active input
// (interpolated, optionally with its reference gradient), an optional passive interpolated input, strided quadrature
// data, and one output field (interpolated value or gradient test weights) scattered through its own restriction.
// It serves the velocity-gradient projection right-hand side (qfunctions/velocity_gradient_projection.h:9-36), the
// anisotropy-tensor projection right-hand side (qfunctions/grid_anisotropy_tensor.h:34-54), Mass_N for any component
// count (qfunctions/mass.h:19-26) and the SGS residual IFunction_NodalSgs (qfunctions/sgs_dd_model.h:215-241).
// Each element owns Q^3 threads, one per quadrature point, and its basis passes are staged through LDS (basis3::Stage).
//
// k_sgs_nodal evaluates ComputeSgsDDNodal (qfunctions/sgs_dd_model.h:65-96) once per node of the SGS field: every
// element node of a mesh node sees the same nodal inputs, so the closure value is computed once and the element
// contributions inverse_multiplicity(e, n) * value are summed in ascending element-node order, the order of the
// restriction transpose.  The model parameters are staged in LDS per block.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "basis3d.hpp"
#include "kernels.hpp"
#include "newtonian_ops.hpp"
#include "sgs_ops.hpp"

namespace {

using basis3::Dims;

template <int Q>
constexpr int sgs_elems_per_block() {
  return (Q * Q * Q >= 64) ? 1 : 64 / (Q * Q * Q);
}

// ---- quadrature-point bodies ----
// Op::NU: components of the active interpolated input (0: none), Op::GRAD: its reference gradient is used,
// Op::NU2: components of the passive interpolated input, Op::NV: output components, Op::OUTG: the output is gradient
// test weights out[d * NV + c] (else values out[c]), Op::NST: LDS slots per element.

template <int VAR>
struct VelocityGradientRHS {  // v[j*3+k] = wdetJ du_j/dx_k
  using Ctx                       = CeedMiNewtonianContext;
  static constexpr int  NU = 5, NU2 = 0, NV = 9, NST = 27;
  static constexpr bool GRAD = true, OUTV = true, OUTG = false;
  __device__ static void qf(const FieldArgs &, const Ctx &ctx, int, int, const double qd[10], bool ok, const double (&u)[5], const double (&gu)[3][5],
                            const double (&)[1], double (&out)[9]) {
    const auto g = ns::Gas<double>::from(ctx.gas);
    double     wdetJ, dXdx[3][3], qi[5];
    nsop::unpack_qdata(qd, ok, wdetJ, dXdx);
#pragma unroll
    for (int c = 0; c < 5; c++) qi[c] = u[c];
    if (!ok) nsop::padding_state<VAR>(qi);
    const ns::St<double> s = ns::state_from_q<VAR>(g, qi);
    ns::St<double>       gs[3];
    ns::phys_gradient<VAR>(g, s, gu, dXdx, gs);
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int k = 0; k < 3; k++) out[j * 3 + k] = wdetJ * gs[k].Y[1 + j];
  }
};

struct AnisotropyRHS {  // v = wdetJ [KM(A_ij), delta]
  using Ctx                       = CeedMiNewtonianContext;
  static constexpr int  NU = 0, NU2 = 0, NV = 7, NST = 21;
  static constexpr bool GRAD = false, OUTV = true, OUTG = false;
  __device__ static void qf(const FieldArgs &, const Ctx &, int, int, const double qd[10], bool ok, const double (&)[1], const double (&)[3][1],
                            const double (&)[1], double (&out)[7]) {
    double wdetJ, dXdx[3][3], kmA[6], delta;
    nsop::unpack_qdata(qd, ok, wdetJ, dXdx);
    sgs::anisotropy_tensor(dXdx, kmA, delta);
#pragma unroll
    for (int j = 0; j < 6; j++) out[j] = wdetJ * kmA[j];
    out[6] = wdetJ * delta;
  }
};

template <int NC>
struct MassN {  // v = wdetJ u
  using Ctx                       = CeedMiNewtonianContext;
  static constexpr int  NU = NC, NU2 = 0, NV = NC, NST = 3 * NC;
  static constexpr bool GRAD = false, OUTV = true, OUTG = false;
  __device__ static void qf(const FieldArgs &, const Ctx &, int, int, const double qd[10], bool ok, const double (&u)[NC], const double (&)[3][NC],
                            const double (&)[1], double (&out)[NC]) {
    const double w = ok ? qd[0] : 0.0;
#pragma unroll
    for (int c = 0; c < NC; c++) out[c] = w * u[c];
  }
};

template <int VAR>
struct NodalSgsResidual {  // Grad_v[k][j] = -wdetJ (dXdx[k] . F_sgs[j])
  using Ctx                       = CeedMiNewtonianContext;
  static constexpr int  NU = 5, NU2 = 6, NV = 5, NST = 20;
  static constexpr bool GRAD = false, OUTV = false, OUTG = true;
  __device__ static void qf(const FieldArgs &, const Ctx &ctx, int, int, const double qd[10], bool ok, const double (&u)[5], const double (&)[3][5],
                            const double (&km)[6], double (&out)[15]) {
    const auto g = ns::Gas<double>::from(ctx.gas);
    double     wdetJ, dXdx[3][3], qi[5], F[5][3];
    nsop::unpack_qdata(qd, ok, wdetJ, dXdx);
#pragma unroll
    for (int c = 0; c < 5; c++) qi[c] = u[c];
    if (!ok) nsop::padding_state<VAR>(qi);
    const ns::St<double> s = ns::state_from_q<VAR>(g, qi);
    sgs::flux_sgs(s.Y, km, F);
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 5; j++) out[k * 5 + j] = -wdetJ * (dXdx[k][0] * F[j][0] + dXdx[k][1] * F[j][1] + dXdx[k][2] * F[j][2]);
  }
};

template <int VAR>
struct DiffFilterRHS {  // v0 = wdetJ Y, v1 = wdetJ u_i u_j (XX YY ZZ YZ XZ XY): one 11-component output
  using Ctx                       = CeedMiDiffFilterContext;
  static constexpr int  NU = 5, NU2 = 0, NV = 11, NST = 33;
  static constexpr bool GRAD = false, OUTV = true, OUTG = false;
  __device__ static void qf(const FieldArgs &, const Ctx &ctx, int, int, const double qd[10], bool ok, const double (&u)[5], const double (&)[3][5],
                            const double (&)[1], double (&out)[11]) {
    const auto   g     = ns::Gas<double>::from(ctx.newt_ctx.gas);
    const double wdetJ = ok ? qd[0] : 0.0;
    double       qi[5];
#pragma unroll
    for (int c = 0; c < 5; c++) qi[c] = u[c];
    if (!ok) nsop::padding_state<VAR>(qi);
    const ns::St<double> s = ns::state_from_q<VAR>(g, qi);
#pragma unroll
    for (int c = 0; c < 5; c++) out[c] = wdetJ * s.Y[c];
    const double *v = &s.Y[1];
    out[5]          = wdetJ * v[0] * v[0];
    out[6]          = wdetJ * v[1] * v[1];
    out[7]          = wdetJ * v[2] * v[2];
    out[8]          = wdetJ * v[1] * v[2];
    out[9]          = wdetJ * v[0] * v[2];
    out[10]         = wdetJ * v[0] * v[1];
  }
};

// alpha_ij of DifferentialFilter_LHS_N (qfunctions/differential_filter.h:91-125) at point t of element e: the
// collocated anisotropy tensor (aux, [elem][7][Q^3]) when the width is grid based, the coordinates for wall damping
template <int Q>
__device__ __forceinline__ void filter_alpha(const FieldArgs &a, const CeedMiDiffFilterContext &f, int e, int t, bool ok, double alpha[3][3]) {
  constexpr int Q3      = Q * Q * Q;
  double        D[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (f.grid_based_width && ok) {
    double km[6];
    const double *A = a.aux + (long long)e * a.auxl.se + t;
    const double  delta = A[6 * Q3];
#pragma unroll
    for (int k = 0; k < 6; k++) km[k] = A[k * Q3] * delta;
    sgs::km_unpack(km, D);
  }
  double S[3] = {f.width_scaling[0], f.width_scaling[1], f.width_scaling[2]};
  if (f.damping_function != 0) {
    double x1 = 0.0;  // y coordinate of the point (trilinear coordinate field)
    if (ok) {
      const int qx = t % Q, qy = (t / Q) % Q, qz = t / (Q * Q);
#pragma unroll
      for (int v = 0; v < 8; v++) {
        const int ia = v & 1, ib = (v >> 1) & 1, ic = v >> 2;
        x1 += a.Bx[qx * 2 + ia] * a.Bx[qy * 2 + ib] * a.Bx[qz * 2 + ic] * a.x[a.xoffsets[(long long)e * 8 + v] + a.xcomp_stride];
      }
    }
    if (f.damping_function == 1) {  // van Driest: -expm1(-y+ / A+)
      const double dc = -expm1(-(x1 / f.friction_length) / f.damping_constant);
      const double s0 = dc * f.width_scaling[0], s2 = dc * f.width_scaling[2];
      S[0]            = 1 < s0 ? s0 : 1;
      S[1]            = dc * f.width_scaling[1];
      S[2]            = 1 < s2 ? s2 : 1;
    } else {  // MMS
      S[0] = 1;
      S[1] = tanh(60 * x1);
      S[2] = 1;
    }
  }
  double SD[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) SD[i][j] = S[i] * D[i][j];  // diag(S) . D
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) alpha[i][j] = (SD[i][0] * SD[0][j] + SD[i][1] * SD[1][j] + SD[i][2] * SD[2][j]) * f.kernel_scaling;
}

template <int N>
struct DiffFilterLHS {  // v = wdetJ q, Grad_v[k] = wdetJ (alpha (dXdx grad_x q))[k], every component alike
  using Ctx                       = CeedMiDiffFilterContext;
  static constexpr int  NU = N, NU2 = 0, NV = N, NST = 4 * N;
  static constexpr bool GRAD = true, OUTV = true, OUTG = true;
  template <int Q>
  __device__ static void qfq(const FieldArgs &a, const Ctx &ctx, int e, int t, const double qd[10], bool ok, const double (&u)[N],
                             const double (&gu)[3][N], const double (&)[1], double (&out)[4 * N]) {
    double wdetJ, dXdx[3][3], alpha[3][3];
    nsop::unpack_qdata(qd, ok, wdetJ, dXdx);
    filter_alpha<Q>(a, ctx, e, t, ok, alpha);
#pragma unroll
    for (int j = 0; j < N; j++) {
      out[j] = wdetJ * u[j];
      double dq[3], dqX[3];
#pragma unroll
      for (int k = 0; k < 3; k++) dq[k] = gu[0][j] * dXdx[0][k] + gu[1][j] * dXdx[1][k] + gu[2][j] * dXdx[2][k];
#pragma unroll
      for (int k = 0; k < 3; k++) dqX[k] = dXdx[k][0] * dq[0] + dXdx[k][1] * dq[1] + dXdx[k][2] * dq[2];
#pragma unroll
      for (int k = 0; k < 3; k++) out[N + k * N + j] = wdetJ * (alpha[k][0] * dqX[0] + alpha[k][1] * dqX[1] + alpha[k][2] * dqX[2]);
    }
  }
};

// Slot map of the stage: active input u -> 0..NU-1 (interpolation scratch NU..3NU-1), its reference gradient ->
// NU + NU d + c, passive input -> 5..10 (interpolation scratch 11..16 and 5..10; only with NU = 5 and no gradient).
// Outputs: values -> 0..NV-1 (transpose scratch up to 3NV), gradient test weights -> NV + NV d + c.
// Op::qf(a, ctx, e, t, ...) or, for bodies that depend on Q, Op::qfq<Q>(...)
template <class Op, int Q, class = void>
struct HasQfq {
  static constexpr bool value = false;
};
template <class Op, int Q>
struct HasQfq<Op, Q, decltype((void)&Op::template qfq<Q>)> {
  static constexpr bool value = true;
};

template <int P, int Q, class Op>
__global__ void __launch_bounds__(sgs_elems_per_block<Q>() * Q * Q * Q) k_field(FieldArgs a, const typename Op::Ctx ctx) {
  using Dm          = Dims<P, Q>;
  constexpr int EPB = sgs_elems_per_block<Q>();
  constexpr int Q3  = Dm::Q3;
  constexpr int NU = Op::NU, NU2 = Op::NU2, NV = Op::NV;
  constexpr int NUX = NU > 0 ? NU : 1, NU2X = NU2 > 0 ? NU2 : 1, NOUT = (Op::OUTV ? NV : 0) + (Op::OUTG ? 3 * NV : 0);
  constexpr int G0  = Op::OUTV ? NV : 0;  // first gradient-weight entry of out[]
  static_assert(NU2 == 0 || (NU == 5 && !Op::GRAD && NU2 == 6), "passive input slot map assumes a 5-component value input");
  static_assert(Op::NST >= 3 * NU && Op::NST >= (Op::GRAD ? 4 * NU : 0) && Op::NST >= (NU2 ? 17 : 0), "stage too small for the inputs");
  static_assert(Op::NST >= (Op::OUTG ? 4 * NV : 3 * NV), "stage too small for the output transpose");
  static_assert(Op::OUTV || Op::OUTG, "no output");
  static_assert(Dm::P3 <= Q3, "P <= Q is required");
  __shared__ double sB[Q * P], sDc[Q * Q];
  __shared__ double sX[EPB][Op::NST][Q3];
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) sB[i] = a.bt.B[i];
  for (int i = threadIdx.x; i < Q * Q; i += blockDim.x) sDc[i] = a.bt.Dc[i];
  const int       le = threadIdx.x / Q3, t = threadIdx.x % Q3;
  const bool      nd = t < Dm::P3;
  double(*X)[Q3]     = sX[le];
  const long long nbatch = ((long long)a.nelem + EPB - 1) / EPB;

  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long e_    = bt * EPB + le;
    const bool      ok    = e_ < a.nelem;
    const int       e     = ok ? (int)e_ : 0;
    const long long enode = (long long)e * Dm::P3 + t;
    basis3::Stage<P, Q> st(sB, sDc, nullptr, X, t);
    if constexpr (NU > 0) {
      double un[NUX];
      if (ok && nd) {
        const int off = a.in_offsets[enode];
#pragma unroll
        for (int c = 0; c < NU; c++) un[c] = a.u[off + c * a.in_comp_stride];
      } else {
#pragma unroll
        for (int c = 0; c < NU; c++) un[c] = 0.0;
      }
      st.template interp<NUX>(un, NU, 2 * NU, 0);
      if constexpr (Op::GRAD) st.template grad<NUX>(0, NU);
    } else {
      __syncthreads();  // the previous batch's transpose may still read the stage
    }
    if constexpr (NU2 > 0) {
      double kn[NU2X];
      if (ok && nd) {
        const int off = a.in2_offsets[enode];
#pragma unroll
        for (int c = 0; c < NU2; c++) kn[c] = a.u2[off + c * a.in2_comp_stride];
      } else {
#pragma unroll
        for (int c = 0; c < NU2; c++) kn[c] = 0.0;
      }
      st.template interp<NU2X>(kn, 11, 5, 5);
    }
    double qd[10];
#pragma unroll
    for (int k = 0; k < 10; k++) qd[k] = ok ? a.qdata[(long long)e * a.qd.se + k * Q3 + t] : 0.0;
    double u[NUX], gu[3][NUX], u2[NU2X], out[NOUT];
#pragma unroll
    for (int c = 0; c < NUX; c++) {
      u[c] = NU > 0 ? X[c][t] : 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) gu[d][c] = Op::GRAD ? X[NU + NU * d + c][t] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < NU2X; c++) u2[c] = NU2 > 0 ? X[5 + c][t] : 0.0;
    if constexpr (HasQfq<Op, Q>::value) Op::template qfq<Q>(a, ctx, e, t, qd, ok, u, gu, u2, out);
    else Op::qf(a, ctx, e, t, qd, ok, u, gu, u2, out);
    __syncthreads();  // other threads' last interpolation pass may still read the stage
    double r[NV];
    if constexpr (Op::OUTV) {
#pragma unroll
      for (int c = 0; c < NV; c++) X[c][t] = out[c];
    }
    if constexpr (Op::OUTG) {
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int c = 0; c < NV; c++) X[NV + NV * d + c][t] = out[G0 + d * NV + c];
    }
    st.template transpose<NV, Op::OUTV, Op::OUTG>(0, NV, r);
    if (ok && nd) {
      if (a.ev) {  // two-pass scatter: node-sorted element-node contributions, summed per node by launch_node_gather
        double *dst = a.ev + (long long)a.perm[enode] * NV;
#pragma unroll
        for (int c = 0; c < NV; c++) dst[c] = r[c];
      } else {
        const int off = a.out_offsets[enode];
#pragma unroll
        for (int c = 0; c < NV; c++) atomicAdd(&a.v[off + c * a.out_comp_stride], r[c]);
      }
    }
  }
}

// A diagonal entry r of element node enode in each of the ncomp components: into the node-sorted E-vector (two-pass,
// summed per node in a fixed order by launch_node_gather) or added atomically
__device__ __forceinline__ void diag_out(const FieldArgs &a, long long enode, int ncomp, double r) {
  if (a.ev) {
    double *dst = a.ev + (long long)a.perm[enode] * ncomp;
    for (int c = 0; c < ncomp; c++) dst[c] = r;
  } else {
    const int off = a.out_offsets[enode];
    for (int c = 0; c < ncomp; c++) atomicAdd(&a.v[off + c * a.out_comp_stride], r);
  }
}

// Diagonal of Mass_N: node n of every element receives sum_q wdetJ(q) B3(q, n)^2 in each of its NC components
template <int P, int Q>
__global__ void __launch_bounds__(sgs_elems_per_block<Q>() * Q * Q * Q) k_mass_n_diagonal(FieldArgs a, int ncomp) {
  using Dm          = Dims<P, Q>;
  constexpr int EPB = sgs_elems_per_block<Q>();
  constexpr int Q3  = Dm::Q3;
  __shared__ double sBB[Q * P];
  __shared__ double sS[EPB][3][Dm::MAX3];
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) sBB[i] = a.bt.BB[i];
  const int       le = threadIdx.x / Q3, t = threadIdx.x % Q3;
  const long long nbatch = ((long long)a.nelem + EPB - 1) / EPB;
  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long e_ = bt * EPB + le;
    const bool      ok = e_ < a.nelem;
    const int       e  = ok ? (int)e_ : 0;
    const double    w  = ok ? a.qdata[(long long)e * a.qd.se + t] : 0.0;
    basis3::Elem<P, Q> el(nullptr, nullptr, sS[le][0], sS[le][1], sS[le][2], t);
    const double       r = el.contract_t(w, sBB, sBB, sBB);
    if (ok && t < Dm::P3) diag_out(a, (long long)e * Dm::P3 + t, ncomp, r);
    __syncthreads();
  }
}

// InverseMultiplicity (qfunctions/inverse_multiplicity.h:6-12): strided output 1 / multiplicity[component 0]
__global__ void k_inverse_multiplicity(const int *offsets, long long n, const double *mult, double *out, QDataLayout ol, int elem_size) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long e = i / elem_size, k = i % elem_size;
    out[k * ol.sn + e * ol.se] = 1.0 / mult[offsets[i]];
  }
}

constexpr int kNodalBlock = 256;

template <int VAR>
__global__ void __launch_bounds__(kNodalBlock) k_sgs_nodal(SgsNodalArgs a, const CeedMiNewtonianContext ctx) {
  __shared__ sgs::Model m;
  {
    const double *src = (const double *)a.model;
    double       *dst = (double *)&m;
    for (int i = threadIdx.x; i < (int)(sizeof(sgs::Model) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  // One row per thread and no grid-stride loop: in a loop the compiler would hoist the loop-invariant LDS reads of
  // the 240 model weights into registers (and spill them).
  const auto      g = ns::Gas<double>::from(ctx.gas);
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < a.nrows) {
    const int j0 = a.rowptr[k], j1 = a.rowptr[k + 1];
    const int i0 = a.index[j0];  // the first element node of this mesh node provides the nodal inputs
    double    qi[5], gu[3][3], kmA[6];
    {
      const int off = a.q_offsets[i0];
#pragma unroll
      for (int c = 0; c < 5; c++) qi[c] = a.q[off + c * a.q_comp_stride];
    }
    {
      const int off = a.g_offsets[i0];
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int kk = 0; kk < 3; kk++) gu[j][kk] = a.g[off + (j * 3 + kk) * a.g_comp_stride];
    }
    const int    aoff  = a.a_offsets[i0];
#pragma unroll
    for (int c = 0; c < 6; c++) kmA[c] = a.aniso[aoff + c * a.a_comp_stride];
    const double delta = a.aniso[aoff + 6 * a.a_comp_stride];
    const ns::St<double> s = ns::state_from_q<VAR>(g, qi);
    double               km[6];
    sgs::sgs_stress(m, gu, kmA, delta, g.mu / s.U[0], km);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int j = j0; j < j1; j++) {
      const int    ie = a.index[j];
      const double w  = a.inv_mult[(long long)(ie % a.elem_size) * a.im.sn + (long long)(ie / a.elem_size) * a.im.se];
#pragma unroll
      for (int c = 0; c < 6; c++) acc[c] += w * km[c];
    }
    const int off = a.row_offset[k];
#pragma unroll
    for (int c = 0; c < 6; c++) a.v[off + c * a.out_comp_stride] += acc[c];
  }
}

// Diagonal of the filter operator DifferentialFilter_LHS_N: per point the mass weight wdetJ and the gradient coupling
// K = wdetJ alpha (dXdx dXdx^T) (Grad_v[k] = sum_d K[k][d] dq/dX_d), contracted with products of the 1D tables
// (B.B, G.B, G.G) by sum factorization; the same value in each of the ncomp components of a node.
template <int P, int Q>
__global__ void __launch_bounds__(sgs_elems_per_block<Q>() * Q * Q * Q) k_filter_lhs_diagonal(FieldArgs a, const CeedMiDiffFilterContext ctx,
                                                                                             int ncomp) {
  using Dm          = Dims<P, Q>;
  constexpr int EPB = sgs_elems_per_block<Q>();
  constexpr int Q3  = Dm::Q3;
  __shared__ double sBB[Q * P], sGB[Q * P], sGG[Q * P];
  __shared__ double sS[EPB][3][Dm::MAX3];
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) {
    sBB[i] = a.bt.BB[i];
    sGB[i] = a.bt.GB[i];
    sGG[i] = a.bt.GG[i];
  }
  const int       le = threadIdx.x / Q3, t = threadIdx.x % Q3;
  const long long nbatch = ((long long)a.nelem + EPB - 1) / EPB;
  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long e_ = bt * EPB + le;
    const bool      ok = e_ < a.nelem;
    const int       e  = ok ? (int)e_ : 0;
    double          qd[10];
#pragma unroll
    for (int k = 0; k < 10; k++) qd[k] = ok ? a.qdata[(long long)e * a.qd.se + k * Q3 + t] : 0.0;
    double wdetJ, dXdx[3][3], alpha[3][3], G[3][3], K[3][3];
    nsop::unpack_qdata(qd, ok, wdetJ, dXdx);
    filter_alpha<Q>(a, ctx, e, t, ok, alpha);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int d = 0; d < 3; d++) G[i][d] = dXdx[i][0] * dXdx[d][0] + dXdx[i][1] * dXdx[d][1] + dXdx[i][2] * dXdx[d][2];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int d = 0; d < 3; d++) K[k][d] = wdetJ * (alpha[k][0] * G[0][d] + alpha[k][1] * G[1][d] + alpha[k][2] * G[2][d]);
    basis3::Elem<P, Q> el(nullptr, nullptr, sS[le][0], sS[le][1], sS[le][2], t);
    double             r = el.contract_t(wdetJ, sBB, sBB, sBB);
    r += el.contract_t(K[0][0], sGG, sBB, sBB);
    r += el.contract_t(K[1][1], sBB, sGG, sBB);
    r += el.contract_t(K[2][2], sBB, sBB, sGG);
    r += el.contract_t(K[0][1] + K[1][0], sGB, sGB, sBB);
    r += el.contract_t(K[0][2] + K[2][0], sGB, sBB, sGB);
    r += el.contract_t(K[1][2] + K[2][1], sBB, sGB, sGB);
    if (ok && t < Dm::P3) diag_out(a, (long long)e * Dm::P3 + t, ncomp, r);
    __syncthreads();
  }
}

// AnisotropyTensorCollocate (qfunctions/grid_anisotropy_tensor.h:57-77): [KM(A_ij), delta] at every point
__global__ void k_anisotropy_collocate(int Q3, long long n, const double *qdata, QDataLayout qd, double *out, QDataLayout ol) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long e = i / Q3, t = i % Q3;
    const double   *q = qdata + e * qd.se + t * qd.sn;
    double          dXdx[3][3], kmA[6], delta;
#pragma unroll
    for (int k = 0; k < 9; k++) (&dXdx[0][0])[k] = q[(1 + k) * qd.sc];
    sgs::anisotropy_tensor(dXdx, kmA, delta);
    double *o = out + e * ol.se + t * ol.sn;
#pragma unroll
    for (int k = 0; k < 6; k++) o[k * ol.sc] = kmA[k];
    o[6 * ol.sc] = delta;
  }
}

// ComputeSGS_DDAnisotropicTrainingDataNodal_Prim (qfunctions/sgs_dd_training.h:23-60), once per node of the training
// field, summed over the node's element nodes with their inverse multiplicities like k_sgs_nodal
__global__ void __launch_bounds__(kNodalBlock) k_sgs_training_nodal(SgsNodalArgs a, const CeedMiNewtonianContext ctx) {
  const auto      g = ns::Gas<double>::from(ctx.gas);
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nrows) return;
  const int j0 = a.rowptr[k], j1 = a.rowptr[k + 1];
  const int i0 = a.index[j0];
  double    yi[5], vp[6], gu[3][3], kmA[6];
  {
    const int off = a.q_offsets[i0];
#pragma unroll
    for (int c = 0; c < 5; c++) yi[c] = a.q[off + c * a.q_comp_stride];
  }
  {
    const int off = a.p_offsets[i0];
#pragma unroll
    for (int c = 0; c < 6; c++) vp[c] = a.q[off + c * a.p_comp_stride];
  }
  {
    const int off = a.g_offsets[i0];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int kk = 0; kk < 3; kk++) gu[j][kk] = a.g[off + (j * 3 + kk) * a.g_comp_stride];
  }
  const int aoff = a.a_offsets[i0];
#pragma unroll
  for (int c = 0; c < 6; c++) kmA[c] = a.aniso[aoff + c * a.a_comp_stride];
  const double         delta = a.aniso[aoff + 6 * a.a_comp_stride];
  const ns::St<double> s     = ns::state_from_q<ns::VAR_PRIM>(g, yi);
  double               E[3][3], val[12], gmag;
  sgs::sgs_inputs(gu, kmA, delta, g.mu / s.U[0], E, val, gmag);
  const double *u = &s.Y[1];
  val[6]          = vp[0] - u[0] * u[0];
  val[7]          = vp[1] - u[1] * u[1];
  val[8]          = vp[2] - u[2] * u[2];
  val[9]          = vp[3] - u[1] * u[2];
  val[10]         = vp[4] - u[0] * u[2];
  val[11]         = vp[5] - u[0] * u[1];
  double acc[12];
#pragma unroll
  for (int c = 0; c < 12; c++) acc[c] = 0;
  for (int j = j0; j < j1; j++) {
    const int    ie = a.index[j];
    const double w  = a.inv_mult[(long long)(ie % a.elem_size) * a.im.sn + (long long)(ie / a.elem_size) * a.im.se];
#pragma unroll
    for (int c = 0; c < 12; c++) acc[c] += val[c] * w;
  }
  const int off = a.row_offset[k];
#pragma unroll
  for (int c = 0; c < 12; c++) a.v[off + c * a.out_comp_stride] += acc[c];
}

template <typename K>
int grid_for_batches(K kernel, int block, long long nbatch) {
  int dev = 0, ncu = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
  if (ncu <= 0) ncu = 256;
  if (per_cu <= 0) per_cu = 1;
  const long long cap = (long long)ncu * per_cu * 4;
  return (int)(nbatch < cap ? (nbatch > 0 ? nbatch : 1) : cap);
}

template <int P, int Q, class Op>
hipError_t launch_field_pq(const FieldArgs &a, const typename Op::Ctx &ctx, hipStream_t s) {
  if (a.nelem <= 0) return hipSuccess;
  constexpr int EPB   = sgs_elems_per_block<Q>();
  constexpr int block = EPB * Q * Q * Q;
  auto          kern  = k_field<P, Q, Op>;
  const int     grid  = grid_for_batches(kern, block, ((long long)a.nelem + EPB - 1) / EPB);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, a, ctx);
  return hipGetLastError();
}

template <int P, int Q>
hipError_t launch_mass_n_diagonal_pq(const FieldArgs &a, int ncomp, hipStream_t s) {
  if (a.nelem <= 0) return hipSuccess;
  constexpr int EPB   = sgs_elems_per_block<Q>();
  constexpr int block = EPB * Q * Q * Q;
  auto          kern  = k_mass_n_diagonal<P, Q>;
  const int     grid  = grid_for_batches(kern, block, ((long long)a.nelem + EPB - 1) / EPB);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, a, ncomp);
  return hipGetLastError();
}

// (P, Q) pairs of the SGS kernels: the same set as the Newtonian volume kernels (Q_1d = P_1d)
#define SGS_PQ_LIST(X) X(2, 2) X(3, 3) X(4, 4) X(5, 5)

template <template <int, int, class> class F, class Op, class... Args>
hipError_t dispatch_pq(int P, int Q, Args &&...args) {
#define SGS_CASE(p, q) \
  if (P == p && Q == q) return F<p, q, Op>::run(args...);
  SGS_PQ_LIST(SGS_CASE)
#undef SGS_CASE
  return hipErrorInvalidValue;
}

template <int P, int Q, class Op>
struct RunField {
  static hipError_t run(const FieldArgs &a, const typename Op::Ctx &ctx, hipStream_t s) { return launch_field_pq<P, Q, Op>(a, ctx, s); }
};

template <template <int> class OpT, class... Args>
hipError_t dispatch_var(int var, int P, int Q, Args &&...args) {
  switch (var) {
    case 0:
      return dispatch_pq<RunField, OpT<0>>(P, Q, args...);
    case 1:
      return dispatch_pq<RunField, OpT<1>>(P, Q, args...);
    case 2:
      return dispatch_pq<RunField, OpT<2>>(P, Q, args...);
  }
  return hipErrorInvalidValue;
}

}  // namespace

bool sgs_pq_supported(int P, int Q) {
#define SGS_SUP(p, q) \
  if (P == p && Q == q) return true;
  SGS_PQ_LIST(SGS_SUP)
#undef SGS_SUP
  return false;
}

bool mass_n_supported(int ncomp) { return ncomp >= 1 && ncomp <= 9 && ncomp != 8; }

hipError_t launch_velocity_gradient_rhs(int P, int Q, int var, const FieldArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  return dispatch_var<VelocityGradientRHS>(var, P, Q, a, ctx, s);
}

hipError_t launch_anisotropy_rhs(int P, int Q, const FieldArgs &a, hipStream_t s) {
  const CeedMiNewtonianContext ctx{};
  return dispatch_pq<RunField, AnisotropyRHS>(P, Q, a, ctx, s);
}

hipError_t launch_mass_n(int P, int Q, int ncomp, const FieldArgs &a, hipStream_t s) {
  const CeedMiNewtonianContext ctx{};
  switch (ncomp) {
    case 1:
      return dispatch_pq<RunField, MassN<1>>(P, Q, a, ctx, s);
    case 2:
      return dispatch_pq<RunField, MassN<2>>(P, Q, a, ctx, s);
    case 3:
      return dispatch_pq<RunField, MassN<3>>(P, Q, a, ctx, s);
    case 4:
      return dispatch_pq<RunField, MassN<4>>(P, Q, a, ctx, s);
    case 5:
      return dispatch_pq<RunField, MassN<5>>(P, Q, a, ctx, s);
    case 6:
      return dispatch_pq<RunField, MassN<6>>(P, Q, a, ctx, s);
    case 7:
      return dispatch_pq<RunField, MassN<7>>(P, Q, a, ctx, s);
    case 9:
      return dispatch_pq<RunField, MassN<9>>(P, Q, a, ctx, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_mass_n_diagonal(int P, int Q, int ncomp, const FieldArgs &a, hipStream_t s) {
#define SGS_DIAG(p, q) \
  if (P == p && Q == q) return launch_mass_n_diagonal_pq<p, q>(a, ncomp, s);
  SGS_PQ_LIST(SGS_DIAG)
#undef SGS_DIAG
  return hipErrorInvalidValue;
}

hipError_t launch_ifunction_nodal_sgs(int P, int Q, int var, const FieldArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  return dispatch_var<NodalSgsResidual>(var, P, Q, a, ctx, s);
}

hipError_t launch_inverse_multiplicity(const int *offsets, int nelem, int elem_size, const double *mult, double *out, QDataLayout ol, hipStream_t s) {
  const long long n = (long long)nelem * elem_size;
  if (n <= 0) return hipSuccess;
  const int grid = (int)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_inverse_multiplicity, dim3(grid), dim3(256), 0, s, offsets, n, mult, out, ol, elem_size);
  return hipGetLastError();
}

hipError_t launch_sgs_nodal(int var, const SgsNodalArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  if (a.nrows <= 0) return hipSuccess;
  const long long nb = (a.nrows + kNodalBlock - 1) / kNodalBlock;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  const int grid = (int)nb;
  switch (var) {
    case 0:
      hipLaunchKernelGGL(k_sgs_nodal<0>, dim3(grid), dim3(kNodalBlock), 0, s, a, ctx);
      break;
    case 1:
      hipLaunchKernelGGL(k_sgs_nodal<1>, dim3(grid), dim3(kNodalBlock), 0, s, a, ctx);
      break;
    case 2:
      hipLaunchKernelGGL(k_sgs_nodal<2>, dim3(grid), dim3(kNodalBlock), 0, s, a, ctx);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t sgs_model_doubles() { return sizeof(sgs::Model) / sizeof(double); }

hipError_t launch_diff_filter_rhs(int P, int Q, int var, const FieldArgs &a, const CeedMiDiffFilterContext &ctx, hipStream_t s) {
  return dispatch_var<DiffFilterRHS>(var, P, Q, a, ctx, s);
}

bool diff_filter_ncomp_supported(int ncomp) { return ncomp == 1 || ncomp == 5 || ncomp == 6 || ncomp == 11; }

hipError_t launch_diff_filter_lhs(int P, int Q, int ncomp, const FieldArgs &a, const CeedMiDiffFilterContext &ctx, hipStream_t s) {
  switch (ncomp) {
    case 1:
      return dispatch_pq<RunField, DiffFilterLHS<1>>(P, Q, a, ctx, s);
    case 5:
      return dispatch_pq<RunField, DiffFilterLHS<5>>(P, Q, a, ctx, s);
    case 6:
      return dispatch_pq<RunField, DiffFilterLHS<6>>(P, Q, a, ctx, s);
    case 11:
      return dispatch_pq<RunField, DiffFilterLHS<11>>(P, Q, a, ctx, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_diff_filter_lhs_diagonal(int P, int Q, int ncomp, const FieldArgs &a, const CeedMiDiffFilterContext &ctx, hipStream_t s) {
  if (a.nelem <= 0) return hipSuccess;
#define FD_DIAG(p, q)                                                                        \
  if (P == p && Q == q) {                                                                    \
    constexpr int EPB   = sgs_elems_per_block<q>();                                          \
    constexpr int block = EPB * q * q * q;                                                   \
    auto          kern  = k_filter_lhs_diagonal<p, q>;                                       \
    const int     grid  = grid_for_batches(kern, block, ((long long)a.nelem + EPB - 1) / EPB); \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, a, ctx, ncomp);                  \
    return hipGetLastError();                                                                \
  }
  SGS_PQ_LIST(FD_DIAG)
#undef FD_DIAG
  return hipErrorInvalidValue;
}

hipError_t launch_anisotropy_collocate(int Q, int nelem, const double *qdata, QDataLayout qd, double *out, QDataLayout ol, hipStream_t s) {
  const long long n = (long long)nelem * Q * Q * Q;
  if (n <= 0) return hipSuccess;
  const int grid = (int)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_anisotropy_collocate, dim3(grid), dim3(256), 0, s, Q * Q * Q, n, qdata, qd, out, ol);
  return hipGetLastError();
}

hipError_t launch_sgs_training_nodal(const SgsNodalArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  if (a.nrows <= 0) return hipSuccess;
  const long long nb = (a.nrows + kNodalBlock - 1) / kNodalBlock;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_sgs_training_nodal, dim3((int)nb), dim3(kNodalBlock), 0, s, a, ctx);
  return hipGetLastError();
}
