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
  static constexpr int  NU = 5, NU2 = 0, NV = 9, NST = 27;
  static constexpr bool GRAD = true, OUTG = false;
  __device__ static void qf(const CeedMiNewtonianContext &ctx, const double qd[10], bool ok, const double (&u)[5], const double (&gu)[3][5],
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
  static constexpr int  NU = 0, NU2 = 0, NV = 7, NST = 21;
  static constexpr bool GRAD = false, OUTG = false;
  __device__ static void qf(const CeedMiNewtonianContext &, const double qd[10], bool ok, const double (&)[1], const double (&)[3][1],
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
  static constexpr int  NU = NC, NU2 = 0, NV = NC, NST = 3 * NC;
  static constexpr bool GRAD = false, OUTG = false;
  __device__ static void qf(const CeedMiNewtonianContext &, const double qd[10], bool ok, const double (&u)[NC], const double (&)[3][NC],
                            const double (&)[1], double (&out)[NC]) {
    const double w = ok ? qd[0] : 0.0;
#pragma unroll
    for (int c = 0; c < NC; c++) out[c] = w * u[c];
  }
};

template <int VAR>
struct NodalSgsResidual {  // Grad_v[k][j] = -wdetJ (dXdx[k] . F_sgs[j])
  static constexpr int  NU = 5, NU2 = 6, NV = 5, NST = 20;
  static constexpr bool GRAD = false, OUTG = true;
  __device__ static void qf(const CeedMiNewtonianContext &ctx, const double qd[10], bool ok, const double (&u)[5], const double (&)[3][5],
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

// Slot map of the stage: active input u -> 0..NU-1 (interpolation scratch NU..3NU-1), its reference gradient ->
// NU + NU d + c, passive input -> 5..10 (interpolation scratch 11..16 and 5..10; only with NU = 5 and no gradient).
// Outputs: values -> 0..NV-1 (transpose scratch up to 3NV), gradient test weights -> NV + NV d + c.
template <int P, int Q, class Op>
__global__ void __launch_bounds__(sgs_elems_per_block<Q>() * Q * Q * Q) k_field(FieldArgs a, const CeedMiNewtonianContext ctx) {
  using Dm          = Dims<P, Q>;
  constexpr int EPB = sgs_elems_per_block<Q>();
  constexpr int Q3  = Dm::Q3;
  constexpr int NU = Op::NU, NU2 = Op::NU2, NV = Op::NV;
  constexpr int NUX = NU > 0 ? NU : 1, NU2X = NU2 > 0 ? NU2 : 1, NOUT = Op::OUTG ? 3 * NV : NV;
  static_assert(NU2 == 0 || (NU == 5 && !Op::GRAD && NU2 == 6), "passive input slot map assumes a 5-component value input");
  static_assert(Op::NST >= 3 * NU && Op::NST >= (Op::GRAD ? 4 * NU : 0) && Op::NST >= (NU2 ? 17 : 0), "stage too small for the inputs");
  static_assert(Op::NST >= (Op::OUTG ? 4 * NV : 3 * NV), "stage too small for the output transpose");
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
    basis3::Stage<P, Q> st(sB, sDc, X, t);
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
    Op::qf(ctx, qd, ok, u, gu, u2, out);
    __syncthreads();  // other threads' last interpolation pass may still read the stage
    double r[NV];
    if constexpr (Op::OUTG) {
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int c = 0; c < NV; c++) X[NV + NV * d + c][t] = out[d * NV + c];
      st.template transpose<NV, false, true>(0, NV, r);
    } else {
#pragma unroll
      for (int c = 0; c < NV; c++) X[c][t] = out[c];
      st.template transpose<NV, true, false>(0, NV, r);
    }
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
    if (ok && t < Dm::P3) {
      const int off = a.out_offsets[(long long)e * Dm::P3 + t];
      for (int c = 0; c < ncomp; c++) atomicAdd(&a.v[off + c * a.out_comp_stride], r);
    }
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
hipError_t launch_field_pq(const FieldArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
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
  static hipError_t run(const FieldArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) { return launch_field_pq<P, Q, Op>(a, ctx, s); }
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
