// Fused element kernels for HONEE's Newtonian volume operators on MI355X (gfx950).
//
// k_volume<P, Q, Op> applies a whole volume sub-operator in one launch: per element it gathers the L-vector through
// the element restriction, interpolates values and reference gradients to the quadrature points in LDS, evaluates the
// operator's quadrature-point body (newtonian_ops.hpp) in registers, applies the basis transpose and scatter-adds into
// the output L-vector.  Quadrature data (qdata, jac_data) is streamed in the backend's strided layout.  Each element
// owns Q^3 threads (one per quadrature point) and a block holds EPB elements; blocks loop over element batches.
#include "volume_common.hpp"

namespace {

// ------------------------------------------------------------------------------------------------------------------
// Setup: qdata = {w detJ, dXdx} from the coordinate field (qfunctions/setupgeo.h:42-56).  The trilinear coordinate
// basis is applied directly with its 1D interpolation and derivative tables.
// ------------------------------------------------------------------------------------------------------------------
template <int PX, int Q>
__global__ void __launch_bounds__(256) k_setup_qdata(SetupArgs a) {
  static_assert(PX == 2, "trilinear coordinate field");
  constexpr int   Q3  = Q * Q * Q;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)a.nelem * Q3) return;
  const int e = (int)(gid / Q3), i = (int)(gid % Q3);
  double    xv[24], qd[10];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const int off = a.xoffsets[(long long)e * 8 + n];
#pragma unroll
    for (int comp = 0; comp < 3; comp++) xv[comp * 8 + n] = a.x[off + comp * a.xcomp_stride];
  }
  nsop::geo_point<Q>(xv, a.Bx, a.Dx, a.w1d, i, qd);
#pragma unroll
  for (int k = 0; k < 10; k++) a.qdata[nsop::qidx<Q>(a.qd, e, k, i)] = qd[k];
}

// Element vertex coordinates [nelem][3][8] of the trilinear coordinate field (the GeoRecord of a qdata vector)
__global__ void __launch_bounds__(256) k_gather_vertex_coords(int nelem, const int *xoffsets, const double *x, int xcs, double *xe) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nelem * 24) return;
  const long long e = i / 24;
  const int       r = (int)(i % 24), comp = r / 8, n = r % 8;
  xe[i] = x[xoffsets[e * 8 + n] + comp * xcs];
}

// ------------------------------------------------------------------------------------------------------------------
// Generic fused volume kernel
// ------------------------------------------------------------------------------------------------------------------
// Asynchronous copy of the vertex coordinates of element batch bt ([elem][3][8], 192 bytes each) into sXE through the
// same LDS-DMA path, for the kernels that recompute qdata
template <int EPB>
__device__ __forceinline__ void stage_coords_async(const VolumeArgs &a, long long bt, double (*sXE)[24]) {
  const long long e0    = bt * EPB;
  const long long ne    = (long long)a.nelem - e0 < EPB ? (long long)a.nelem - e0 : EPB;
  const int       nb    = (int)(ne * 24 * 8);
  const char     *src   = (const char *)(a.xe + e0 * 24);
  char           *dst   = (char *)&sXE[0][0];
  const int       tid   = threadIdx.x;
  const int       nt    = blockDim.x;
  const int       wave0 = tid - (tid & 63);
  for (int k = 0; k * nt * 16 < nb; k++) {
    const int chunk = k * nt + tid;
    if (chunk * 16 < nb)
      __builtin_amdgcn_global_load_lds((const void *)(src + chunk * 16), (__attribute__((address_space(3))) void *)(dst + (k * nt + wave0) * 16),
                                       16, 0, 0);
  }
}

// The interpolated inputs of every quadrature point and the quadrature-point outputs are staged in LDS (X), one slot
// per (field component, point).  The basis stages treat the five components of a field together (basis3::Stage), so
// a 1D coefficient row is read once per pass and five independent contractions are in flight; the quadrature-point
// body then runs with only its own registers live.
//
// Slot map of the stage: u -> 0..4, reference gradient of u -> 5 + 5 d + c (kGrad), u2 -> 20..24 (kGrad) or 5..9;
// outputs: v -> 0..4, gradient test weights -> 5 + 5 d + c.  Interpolation scratch is chosen from the slots that are
// not live at that point.
// Asynchronous copy of the quadrature data of element batch bt (EPB consecutive elements, [elem][10][Q^3] and
// contiguous) into sQD through the LDS-DMA path (global_load_lds_dwordx4, 16 bytes per lane): no registers are held
// while the loads are in flight.  The caller waits with s_waitcnt vmcnt(0) and a barrier before reading sQD.
template <int EPB, int Q3>
__device__ __forceinline__ void stage_qdata_async(const VolumeArgs &a, long long bt, double (*sQD)[10][Q3]) {
  constexpr int   NT   = EPB * Q3;                       // threads of the block
  constexpr int   NB   = EPB * 10 * Q3 * 8;              // bytes of a full batch
  constexpr int   NCH  = (NB + NT * 16 - 1) / (NT * 16);  // copy instructions per thread
  const long long e0   = bt * EPB;
  const long long ne   = (long long)a.nelem - e0 < EPB ? (long long)a.nelem - e0 : EPB;
  const int       nb   = (int)(ne * 10 * Q3 * 8);
  const char     *src  = (const char *)(a.qdata + e0 * a.qd.se);
  char           *dst  = (char *)&sQD[0][0][0];
  const int       tid  = threadIdx.x;
  const int       wave0 = tid - (tid & 63);  // first thread of this wave: the DMA writes lane i at base + 16 i
#pragma unroll
  for (int k = 0; k < NCH; k++) {
    const int chunk = k * NT + tid;
    if (chunk * 16 < nb)
      __builtin_amdgcn_global_load_lds((const void *)(src + chunk * 16), (__attribute__((address_space(3))) void *)(dst + (k * NT + wave0) * 16),
                                       16, 0, 0);
  }
}

// The interpolated inputs of every quadrature point and the quadrature-point outputs are staged in LDS (X), one slot
// per (field component, point).  The basis stages treat the five components of a field together (basis3::Stage), so
// a 1D coefficient row is read once per pass and five independent contractions are in flight; the quadrature-point
// body then runs with only its own registers live.
//
// Slot map of the stage: u -> 0..4, reference gradient of u -> 5 + 5 d + c (kGrad), u2 -> 20..24 (kGrad) or 5..9;
// outputs: v -> 0..4, gradient test weights -> 5 + 5 d + c.  Interpolation scratch is chosen from the slots that are
// not live at that point.
//
// Software pipeline over a block's element batches: the next batch's quadrature data is copied into sQD by LDS DMA
// while the current batch's quadrature-point body runs, and its node values are loaded into registers while the
// current batch's basis transpose runs, so neither waits on HBM latency at the start of the next batch.
//
// With GEO the quadrature data is recomputed from the element's vertex coordinates (192 bytes per element, staged by
// the same DMA path) instead of read (80 bytes per quadrature point).
template <int P, int Q, class Op, bool GEO>
__global__ void __launch_bounds__(elems_per_block<Q>() * Q * Q * Q, 2) k_volume(VolumeArgs a, const CeedMiNewtonianContext ctx) {
  using Dm           = Dims<P, Q>;
  constexpr int EPB  = elems_per_block<Q>();
  constexpr int Q3   = Dm::Q3;
  constexpr int NIN  = 5 + (Op::kGrad ? 15 : 0) + (Op::kIn2 ? 5 : 0);
  constexpr int NOUT = Op::kScatter ? 20 : 1;
  constexpr int NST0 = NIN > NOUT ? NIN : NOUT;
  constexpr int NST  = NST0 > 15 ? NST0 : 15;            // interpolation scratch needs slots up to 14
  constexpr int U2   = Op::kGrad ? 20 : 5;               // slot of u2
  constexpr int SU2  = Op::kGrad ? 0 : 10;               // scratch of the u2 interpolation
  constexpr int SU   = (Op::kIn2 && !Op::kGrad) ? 10 : 5;  // scratch of the u interpolation
  static_assert(!Op::kIn2 || Op::kGrad || NST >= 20, "stage too small for the u2 scratch");
  __shared__ double                 sB[Q * P], sDc[Q * Q], sG[Q * P];
  __shared__ double                 sX[EPB][NST][Q3];
  // quadrature data staged through LDS by DMA (Op::kStageQData) or read from global memory at the point body
  constexpr bool                    SQ = Op::kStageQData && !GEO;
  __shared__ __align__(16) double   sQD[SQ ? EPB : 1][10][SQ ? Q3 : 1];
  __shared__ __align__(16) double   sXE[GEO ? EPB : 1][24];
  __shared__ double                 sGT[GEO ? 5 * Q : 1];
  if constexpr (GEO)
    for (int i = threadIdx.x; i < 5 * Q; i += blockDim.x) sGT[i] = a.gtab[i];
  static_assert(Dm::P3 <= Q3, "P <= Q is required");
  using Stage = typename StageFor<P, Q, ZCross<Op>::value>::type;
  load_tables<P, Q>(a.bt, sB, sDc, sG);
  const int           le = threadIdx.x / Q3, t = threadIdx.x % Q3;
  const bool          nd = t < Dm::P3;
  double(*X)[Q3]         = sX[le];
  const BatchRange br    = batch_range(((long long)a.nelem + EPB - 1) / EPB);
  const int        cs    = a.comp_stride;

  // prologue: quadrature data and node values of the first batch
  int    off = 0;
  double un[5] = {0, 0, 0, 0, 0}, u2n[5] = {0, 0, 0, 0, 0};
  if (br.first < br.end) {
    if constexpr (SQ) stage_qdata_async<EPB, Q3>(a, br.first, sQD);
    if constexpr (GEO) stage_coords_async<EPB>(a, br.first, sXE);
    const long long e_ = br.first * EPB + le;
    if (nd && e_ < a.nelem) {
      off = a.offsets[e_ * Dm::P3 + t];
#pragma unroll
      for (int c = 0; c < 5; c++) un[c] = a.u[off + c * cs];
      if constexpr (Op::kIn2)
#pragma unroll
        for (int c = 0; c < 5; c++) u2n[c] = a.u2[off + c * cs];
    }
  }

  for (long long bt = br.first; bt < br.end; bt += br.stride) {
    const long long e_   = bt * EPB + le;
    const bool      ok   = e_ < a.nelem;
    const int       e    = ok ? (int)e_ : 0;
    const long long btn  = bt + br.stride;
    const bool      more = btn < br.end;
    const long long en_  = btn * EPB + le;
    const bool      okn  = more && en_ < a.nelem;
    const long long enode = (long long)e * Dm::P3 + t;

    // The basis stages index LDS with per-thread addresses derived from t.  An opaque copy of t per phase keeps the
    // compiler from hoisting those addresses out of the loop, where they would stay live (or spill) across the
    // quadrature-point body.
    int tb = t;
    asm volatile("" : "+v"(tb));
    Stage st(sB, sDc, sG, X, tb);
    if constexpr (Op::kIn2) st.template interp<5>(u2n, SU2, SU2 + 5, U2);
    if constexpr (Op::kGrad) st.template interp_grad<5>(un, SU, SU + 5, 0, 5);
    else st.template interp<5>(un, SU, SU + 5, 0);
    int off_next = 0;
    {
      double qd[10], u[5], gu[3][5], u2[5], v[5], gw[3][5];
      if constexpr (SQ) {
        __builtin_amdgcn_s_waitcnt(0);  // this batch's quadrature data has landed in sQD
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 10; k++) qd[k] = sQD[le][k][t];
        __syncthreads();  // all of sQD has been read: the next batch's copy may overwrite it
        if (more) stage_qdata_async<EPB, Q3>(a, btn, sQD);
      } else if constexpr (GEO) {
        __builtin_amdgcn_s_waitcnt(0);  // this batch's vertex coordinates have landed in sXE
        __syncthreads();
        nsop::geo_point<Q>(sXE[le], sGT, sGT + 2 * Q, sGT + 4 * Q, t, qd);
        __syncthreads();  // all of sXE has been read: the next batch's copy may overwrite it
        if (more) stage_coords_async<EPB>(a, btn, sXE);
      } else {
        nsop::fetch_qdata<Q>(a, e, t, ok, qd);
      }
      if (more && nd && okn) off_next = a.offsets[en_ * Dm::P3 + t];
#pragma unroll
      for (int c = 0; c < 5; c++) {
        u[c] = X[c][t];
        if constexpr (Op::kGrad)
#pragma unroll
          for (int d = 0; d < 3; d++) gu[d][c] = X[5 + 5 * d + c][t];
        if constexpr (Op::kIn2) u2[c] = X[U2 + c][t];
      }
      Op::template qf<Q>(a, ctx, e, t, ok, qd, u, gu, u2, v, gw);
      if constexpr (Op::kScatter) {
        __syncthreads();  // other threads' gradient passes read slots 0..4
#pragma unroll
        for (int c = 0; c < 5; c++) {
          if (Op::kOutV) X[c][t] = v[c];
          if (Op::kOutG)
#pragma unroll
            for (int d = 0; d < 3; d++) X[5 + 5 * d + c][t] = gw[d][c];
        }
      }
    }
    // Node values of the next batch, consumed by its interpolation, and this batch's E-vector slots: loaded while
    // the basis transpose runs.  The scheduling barrier keeps these loads (and their registers) out of the
    // quadrature-point body above.
    __builtin_amdgcn_sched_barrier(0);
    const int off_cur = off;
    const int slot    = (a.perm && ok && nd) ? a.perm[enode] : 0;
    if (more) {
      const bool ld = nd && okn;
#pragma unroll
      for (int c = 0; c < 5; c++) un[c] = ld ? a.u[off_next + c * cs] : 0.0;
      if constexpr (Op::kIn2)
#pragma unroll
        for (int c = 0; c < 5; c++) u2n[c] = ld ? a.u2[off_next + c * cs] : 0.0;
      off = off_next;
    }
    if constexpr (Op::kScatter) {
      int tt = t;
      asm volatile("" : "+v"(tt));
      Stage               sto(sB, sDc, sG, X, tt);
      double              r[5];
      sto.template transpose<5, Op::kOutV, Op::kOutG>(0, 5, r);
      if (ok && nd) {
        if (a.ev32) {  // two-pass scatter in FP32: element order
          float *dst = a.ev32 + enode * 5;
#pragma unroll
          for (int c = 0; c < 5; c++) dst[c] = (float)r[c];
        } else if (a.ev) {  // two-pass scatter: element-node contributions, summed per node by a second kernel
          double *dst = a.ev + (a.perm ? (long long)slot : enode) * 5;
#pragma unroll
          for (int c = 0; c < 5; c++) dst[c] = r[c];
        } else {
#pragma unroll
          for (int c = 0; c < 5; c++) atomicAdd(&a.v[off_cur + c * cs], r[c]);
        }
      }
    } else {
      __syncthreads();  // the next batch's first writes go to slots that the gradient pass read
    }
  }
}

// Operators and orders whose kernels recompute qdata when the backend passes the element vertices (P = Q); the
// others read it.  Measured against reading it (back-to-back applies, bench workloads):
//   * IFunction, Q = 4 (32^3 p=3): 175.9 -> 153.2 us per apply; Q = 5 (96^3 p=4): 19.0 -> 19.7 GDoF/s;
//   * FP64 IJacobian, Q = 4: 27.0 -> 29.1 GDoF/s; Q = 5: 22.9 -> 22.1 GDoF/s (not adopted);
//   * FP32 IJacobian, Q = 4: 28.7 -> 26.4 GDoF/s (not adopted): 136 -> 196 VGPRs, three waves per SIMD to two, and
//     forced to three waves it spills 31 registers;
//   * explicit RHS, Q = 3 (96^3 p=2 fine run of the LES study): 5.92 -> 6.43 GPU s (not adopted).
template <class Op, int Q>
struct RecomputeAt : std::false_type {};
template <int VAR, int Q>
struct RecomputeAt<nsop::IFunction<VAR>, Q> : std::integral_constant<bool, Q == 4 || Q == 5> {};
template <int VAR>
struct RecomputeAt<nsop::IJacobian<VAR, double>, 4> : std::true_type {};

template <int P, int Q, class Op>
hipError_t launch_op(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream) {
  constexpr int EPB   = elems_per_block<Q>();
  const int     block = EPB * Q * Q * Q;
  auto          k     = k_volume<P, Q, Op, false>;
  if constexpr (P == Q && RecomputeAt<Op, Q>::value)
    if (a.xe) k = k_volume<P, Q, Op, true>;
  const int     grid  = grid_for(k, block, ((long long)a.nelem + EPB - 1) / EPB);
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, stream, a, ctx);
  return hipGetLastError();
}

template <int P, int Q>
hipError_t launch_ifunction_var(int var, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  switch (var) {
    case 0: return launch_op<P, Q, nsop::IFunction<0>>(a, ctx, s);
    case 1: return launch_op<P, Q, nsop::IFunction<1>>(a, ctx, s);
    default: return launch_op<P, Q, nsop::IFunction<2>>(a, ctx, s);
  }
}

template <int P, int Q>
hipError_t launch_ijacobian_var(int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  if (fp32) {
    switch (var) {
      case 0: return launch_op<P, Q, nsop::IJacobian<0, float>>(a, ctx, s);
      case 1: return launch_op<P, Q, nsop::IJacobian<1, float>>(a, ctx, s);
      default: return launch_op<P, Q, nsop::IJacobian<2, float>>(a, ctx, s);
    }
  }
  switch (var) {
    case 0: return launch_op<P, Q, nsop::IJacobian<0, double>>(a, ctx, s);
    case 1: return launch_op<P, Q, nsop::IJacobian<1, double>>(a, ctx, s);
    default: return launch_op<P, Q, nsop::IJacobian<2, double>>(a, ctx, s);
  }
}

}  // namespace

hipError_t launch_gather_vertex_coords(int nelem, const int *xoffsets, const double *x, int xcs, double *xe, hipStream_t s) {
  const long long n = (long long)nelem * 24;
  hipLaunchKernelGGL(k_gather_vertex_coords, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nelem, xoffsets, x, xcs, xe);
  return hipGetLastError();
}

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
    case 6: hipLaunchKernelGGL((k_setup_qdata<2, 6>), dim3(grid), dim3(block), 0, stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

bool volume_pq_supported(int P, int Q) {
#define SUP(P_, Q_) (P == P_ && Q == Q_) ||
  return HONEE_PQ_LIST(SUP) false;
#undef SUP
}

hipError_t launch_ifunction(int P, int Q, int var, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
#define CALL_IF(P_, Q_) (launch_ifunction_var<P_, Q_>(var, a, ctx, s))
  HONEE_PQ_ALL(CALL_IF)
#undef CALL_IF
  return hipErrorInvalidValue;
}

hipError_t launch_ijacobian(int P, int Q, int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
#define CALL_IJ(P_, Q_) (launch_ijacobian_var<P_, Q_>(var, fp32, a, ctx, s))
  HONEE_PQ_ALL(CALL_IJ)
#undef CALL_IJ
  return hipErrorInvalidValue;
}

hipError_t launch_rhs(int P, int Q, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
#define CALL_RHS(P_, Q_)                                                                                     \
  (ctx.stabilization == ns::STAB_NONE ? launch_op<P_, Q_, nsop::RHSFunctionT<false>>(a, ctx, s) \
                                      : launch_op<P_, Q_, nsop::RHSFunctionT<true>>(a, ctx, s))
  HONEE_PQ_ALL(CALL_RHS)
#undef CALL_RHS
  return hipErrorInvalidValue;
}

hipError_t launch_mass(int P, int Q, bool stabilized, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
#define CALL_MS(P_, Q_) (stabilized ? launch_op<P_, Q_, nsop::MassNewtonian>(a, ctx, s) : launch_op<P_, Q_, nsop::Mass5>(a, ctx, s))
  HONEE_PQ_ALL(CALL_MS)
#undef CALL_MS
  return hipErrorInvalidValue;
}

hipError_t launch_monitor_ke(int P, int Q, int var, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
#define CALL_KE(P_, Q_)                                                                                       \
  (var == 0 ? launch_op<P_, Q_, nsop::MonitorKE<0>>(a, ctx, s)                                               \
            : (var == 1 ? launch_op<P_, Q_, nsop::MonitorKE<1>>(a, ctx, s) : launch_op<P_, Q_, nsop::MonitorKE<2>>(a, ctx, s)))
  HONEE_PQ_ALL(CALL_KE)
#undef CALL_KE
  return hipErrorInvalidValue;
}
