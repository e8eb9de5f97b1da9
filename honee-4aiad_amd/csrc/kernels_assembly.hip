// Diagonal assembly of the linear volume operators (CeedOperatorLinearAssembleDiagonal and
// CeedOperatorLinearAssemblePointBlockDiagonal) on MI355X (gfx950).
//
// k_diagonal<P, Q, Op> assembles the diagonal of a linear operator (CeedOperatorLinearAssembleDiagonal): at each
// quadrature point the 4x4 block coupling {value, d/dX0, d/dX1, d/dX2} of a component to itself is obtained by applying
// the linear quadrature-point action to unit inputs, then contracted against products of the 1D tables
// (B.B, G.B, G.G) with sum factorization.
//
// The point blocks are assembled by k_pblock_diagonal (LDS slot-pair weights) or, for the 4 x 4 x 4 element, by
// k_pblock_diagonal_mf (z passes on the FP64 matrix cores, no LDS weights).
#include "volume_common.hpp"

namespace {

// The unit input of slot s (value, or reference derivative s - 1) in component C; with s and C compile-time constants
// the action's products with the zero entries fold away (this file is built assuming finite values and no signed
// zeros).
template <typename T, int C>
__device__ __forceinline__ void unit_input(int s, T (&dq)[5], T (&gref)[3][5]) {
#pragma unroll
  for (int c = 0; c < 5; c++) {
    dq[c] = (s == 0 && c == C) ? T(1) : T(0);
#pragma unroll
    for (int d = 0; d < 3; d++) gref[d][c] = (s == 1 + d && c == C) ? T(1) : T(0);
  }
}

// Calls f(std::integral_constant<int, C>) for the runtime component c (one specialized body per component)
template <class F>
__device__ __forceinline__ void with_component(int c, F &&f) {
  switch (c) {
    case 0: f(std::integral_constant<int, 0>{}); break;
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    default: f(std::integral_constant<int, 4>{}); break;
  }
}

// ------------------------------------------------------------------------------------------------------------------
// Diagonal of a linear operator whose quadrature-point action is Op::action (IJacobian)
// ------------------------------------------------------------------------------------------------------------------
template <int P, int Q, class Op, typename T>
__global__ void __launch_bounds__(elems_per_block<Q>() * Q * Q * Q, 2) k_diagonal(VolumeArgs a, const CeedMiNewtonianContext ctx) {
  using Dm          = Dims<P, Q>;
  constexpr int EPB = elems_per_block<Q>();
  constexpr int Q3  = Dm::Q3;
  __shared__ double sBB[Q * P], sGB[Q * P], sGG[Q * P];
  __shared__ double s0[EPB][Dm::MAX3], s1[EPB][Dm::MAX3], s2[EPB][Dm::MAX3];
  __shared__ double sWd[EPB][10][1][Dm::MAX3];  // slot-pair weights of one component
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) {
    sBB[i] = a.bt.BB[i];
    sGB[i] = a.bt.GB[i];
    sGG[i] = a.bt.GG[i];
  }
  __shared__ double sB[Q * P], sDc[Q * Q];
  load_tables<P, Q>(a.bt, sB, sDc);
  const int       le = threadIdx.x / Q3, t = threadIdx.x % Q3;
  Elem<P, Q>      el(sB, sDc, s0[le], s1[le], s2[le], t);
  double(*Wd)[1][Dm::MAX3] = sWd[le];
  const long long nbatch = ((long long)a.nelem + EPB - 1) / EPB;

  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long e_  = bt * EPB + le;
    const bool      ok  = e_ < a.nelem;
    const int       e   = ok ? (int)e_ : 0;
    const int       off = (t < Dm::P3) ? a.offsets[(long long)e * Dm::P3 + t] : 0;
    double          u2[5];
    if constexpr (Op::kIn2) {
#pragma unroll
      for (int c = 0; c < 5; c++) u2[c] = el.interp((t < Dm::P3) ? a.u2[off + c * a.comp_stride] : 0.0);
    }
    typename Op::Point p;
    double             qd[10];
    nsop::fetch_qdata<Q>(a, e, t, ok, qd);
    Op::template load_point<Q>(a, ctx, e, t, ok, qd, u2, p);
    // One specialized body per component c: the block M[s][b] of c to itself from the actions on the unit inputs (s, c),
    // one action's temporaries live at a time
#pragma unroll 1
    for (int c = 0; c < 5; c++) {
      double M[4][4] = {{0}};
      with_component(c, [&](auto cc) {
        constexpr int C = decltype(cc)::value;
#pragma unroll
        for (int s = 0; s < Op::kInSlots; s++) {
          T dq[5], gref[3][5];
          unit_input<T, C>(s, dq, gref);
          double v[5], gw[3][5];
          Op::action(ctx, p, dq, gref, v, gw);
          M[s][0] = ok ? v[C] : 0.0;
#pragma unroll
          for (int d = 0; d < 3; d++) M[s][1 + d] = ok ? gw[d][C] : 0.0;
          __builtin_amdgcn_sched_barrier(0);
        }
      });
      // sum over input slot a and test slot b of M[a][b] phi_a phi_b: the weights of the unordered slot pairs, each with
      // its 1D factor triple, contracted together (basis3::ContractPairs); the contraction's coefficient addresses
      // derive from an opaque copy of t, so that its table loads are not hoisted out of the component loop
      int tc = t;
      asm volatile("" : "+v"(tc));
      __syncthreads();  // the previous component's contraction has read Wd
      Wd[0][0][t] = M[0][0], Wd[1][0][t] = M[1][1], Wd[2][0][t] = M[2][2], Wd[3][0][t] = M[3][3];
      Wd[4][0][t] = M[0][1] + M[1][0], Wd[5][0][t] = M[0][2] + M[2][0], Wd[6][0][t] = M[0][3] + M[3][0];
      Wd[7][0][t] = M[1][2] + M[2][1], Wd[8][0][t] = M[1][3] + M[3][1], Wd[9][0][t] = M[2][3] + M[3][2];
      double r[1] = {0};
      basis3::ContractPairs<P, Q, 1>{tc}.add(Wd, sBB, sGB, sGG, r);
      if (ok && t < Dm::P3) atomicAdd(&a.v[off + c * a.comp_stride], r[0]);
    }
  }
}

// ------------------------------------------------------------------------------------------------------------------
// Point-block diagonal of a linear operator whose quadrature-point action is Op::action: for each input component
// c_in and input slot s (value, three reference derivatives), one action gives the test-slot weights of all five
// output components; the weights of each unordered slot pair are summed (the diagonal's 1D products are symmetric in
// the pair) and contracted five components at a time, grouped by shared 1D factors (basis3::ContractPairs).  Output
// layout [node][c_out][c_in] at 25 * (offset / 5).
// ------------------------------------------------------------------------------------------------------------------
// The weight slot of the unordered input/test slot pair (s, b), as in k_diagonal: 0..3 the diagonal pairs, then
// (0,1) (0,2) (0,3) (1,2) (1,3) (2,3) -> 4..9
__device__ __forceinline__ int pair_slot(int s, int b) {
  const int i = s < b ? s : b, j = s < b ? b : s;
  return i == j ? i : (i == 0 ? 3 + j : (i == 1 ? 5 + j : 9));
}

template <int P, int Q, class Op, typename T>
__global__ void __launch_bounds__(elems_per_block<Q>() * Q * Q * Q, 2) k_pblock_diagonal(VolumeArgs a, const CeedMiNewtonianContext ctx) {
  using Dm          = Dims<P, Q>;
  constexpr int EPB = elems_per_block<Q>();
  constexpr int Q3  = Dm::Q3;
  constexpr int NS  = Op::kInSlots;
  constexpr int NP  = 10;  // slot pairs (with one input slot only pairs 0, 4, 5, 6 carry weights)
  __shared__ double sBB[Q * P], sGB[Q * P], sGG[Q * P];
  __shared__ double s0[EPB][Dm::MAX3], s1[EPB][Dm::MAX3], s2[EPB][Dm::MAX3];
  // pair weights W[pair][c_out][q] (each thread owns its q column, so the accumulation needs no barrier) and the
  // contraction's intermediate arrays
  __shared__ double sW[EPB][NP][5][Dm::MAX3];
  __shared__ double c1[EPB][5][Dm::MAX3], c2[EPB][5][Dm::MAX3];
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) {
    sBB[i] = a.bt.BB[i];
    sGB[i] = a.bt.GB[i];
    sGG[i] = a.bt.GG[i];
  }
  __shared__ double sB[Q * P], sDc[Q * Q];
  load_tables<P, Q>(a.bt, sB, sDc);
  const int               le = threadIdx.x / Q3, t = threadIdx.x % Q3;
  Elem<P, Q>              el(sB, sDc, s0[le], s1[le], s2[le], t);
  double(*W)[5][Dm::MAX3] = sW[le];
  const long long         nbatch = ((long long)a.nelem + EPB - 1) / EPB;

  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long e_  = bt * EPB + le;
    const bool      ok  = e_ < a.nelem;
    const int       e   = ok ? (int)e_ : 0;
    const int       off = (t < Dm::P3) ? a.offsets[(long long)e * Dm::P3 + t] : 0;
    double          u2[5];
    if constexpr (Op::kIn2) {
#pragma unroll
      for (int c = 0; c < 5; c++) u2[c] = el.interp((t < Dm::P3) ? a.u2[off + c * a.comp_stride] : 0.0);
    }
#pragma unroll 1
    for (int ci = 0; ci < 5; ci++) {
      __syncthreads();  // the previous component's contractions have read W
      for (int k = 0; k < NP; k++)
        if (NS > 1 || k == 0 || (k >= 4 && k <= 6))
#pragma unroll
          for (int co = 0; co < 5; co++) W[k][co][t] = 0;
      // The point data is reloaded for each input component (an opaque element index keeps the compiler from hoisting
      // it out of the loop), so that it is not live across the contractions below.
      int eo = e;
      asm volatile("" : "+v"(eo));
      typename Op::Point p;
      double             qd[10];
      nsop::fetch_qdata<Q>(a, eo, t, ok, qd);
      Op::template load_point<Q>(a, ctx, eo, t, ok, qd, u2, p);
      // one specialized body per input component, the slot loop unrolled (the four actions overlap)
      with_component(ci, [&](auto cc) {
        constexpr int C = decltype(cc)::value;
#pragma unroll
        for (int s = 0; s < NS; s++) {
          T dq[5], gref[3][5];
          unit_input<T, C>(s, dq, gref);
          double v[5], gw[3][5];
          Op::action(ctx, p, dq, gref, v, gw);
          if (!ok) continue;
#pragma unroll
          for (int co = 0; co < 5; co++) {
            W[pair_slot(s, 0)][co][t] += v[co];
#pragma unroll
            for (int d = 0; d < 3; d++) W[pair_slot(s, 1 + d)][co][t] += gw[d][co];
          }
        }
      });
      // the contraction's coefficient addresses derive from an opaque copy of t, so that the compiler does not hoist
      // the ten contractions' table loads out of the component loop (they would stay live across the actions)
      int tc = t;
      asm volatile("" : "+v"(tc));
      double r[5] = {0, 0, 0, 0, 0};
      if constexpr (NS > 1) {
        basis3::ContractPairs<P, Q, 5>{tc}.add(W, sBB, sGB, sGG, r);
      } else {
        basis3::Contract5<P, Q> pb{nullptr, c1[le], c2[le], tc};
        pb.add_staged(W[0], sBB, sBB, sBB, r);
        pb.add_staged(W[4], sGB, sBB, sBB, r);
        pb.add_staged(W[5], sBB, sGB, sBB, r);
        pb.add_staged(W[6], sBB, sBB, sGB, r);
      }
      if (ok && t < Dm::P3) {
        if (a.ev) {  // two-pass: the element node's 5 x 5 block in the element-ordered E-vector
          double *dst = a.ev + ((long long)e * Dm::P3 + t) * 25 + ci;
#pragma unroll
          for (int co = 0; co < 5; co++) dst[co * 5] = r[co];
        } else {
#pragma unroll
          for (int co = 0; co < 5; co++) atomicAdd(&a.v[(long long)off * 5 + co * 5 + ci], r[co]);
        }
      }
    }
  }
}

// The 4 x 4 x 4 element (one wave per element, lane = qx + 4 qy + 16 qz): each action's test-slot weights are contracted
// along z on the FP64 matrix cores as soon as the action is done, accumulating into the six (Fx, Fy) groups of
// basis3::ContractPairs (the matrix instruction's accumulator operand carries the sum over the actions and slot pairs),
// so the slot-pair weights never go through LDS; the y and x passes then run across lanes.  A pair (s, b) with s != b
// receives M[s][b] from action s and M[b][s] from action b, as in the LDS form.
// Pair -> (group, z factor): 0 (g0, BB) 1 (g3, BB) 2 (g4, BB) 3 (g0, GG) 4 (g1, BB) 5 (g2, BB) 6 (g0, GB) 7 (g5, BB)
// 8 (g1, GB) 9 (g2, GB).  With a.ev set, the blocks go to the element-ordered E-vector [element node][c_out][c_in]
// (no atomics; launch_node_gather_blocks sums them per node in a fixed order); otherwise they are added atomically.
__device__ __forceinline__ constexpr int pair_group(int pr) {
  return pr == 0 || pr == 3 || pr == 6 ? 0 : (pr == 4 || pr == 8 ? 1 : (pr == 5 || pr == 9 ? 2 : (pr == 1 ? 3 : (pr == 2 ? 4 : 5))));
}
__device__ __forceinline__ constexpr int pair_zfactor(int pr) { return pr == 3 ? 2 : (pr == 6 || pr == 8 || pr == 9 ? 1 : 0); }

template <class Op, typename T>
__global__ void __launch_bounds__(64, 2) k_pblock_diagonal_mf(VolumeArgs a, const CeedMiNewtonianContext ctx) {
  using Dm         = Dims<4, 4>;
  constexpr int NS = Op::kInSlots;
  __shared__ double sF[3][16];  // BB, GB, GG
  __shared__ double s0[1][64], s1[1][64], s2[1][64];
  __shared__ double sB[16], sDc[16];
  for (int i = threadIdx.x; i < 16; i += blockDim.x) {
    sF[0][i] = a.bt.BB[i];
    sF[1][i] = a.bt.GB[i];
    sF[2][i] = a.bt.GG[i];
  }
  load_tables<4, 4>(a.bt, sB, sDc);
  __syncthreads();
  const int  t = threadIdx.x, qx = t & 3, qy = (t >> 2) & 3;
  Elem<4, 4> el(sB, sDc, s0[0], s1[0], s2[0], t);
  // z-pass operands: factor F[z][c] at (i = node c, k = point z) of the matrix instruction's A operand
  double zf[3];
#pragma unroll
  for (int f = 0; f < 3; f++) zf[f] = sF[f][(t >> 4) * 4 + (t & 3)];

  for (long long e_ = blockIdx.x; e_ < a.nelem; e_ += gridDim.x) {
    const int e   = (int)e_;
    const int off = a.offsets[(long long)e * Dm::P3 + t];
    double    u2[5];
    if constexpr (Op::kIn2) {
#pragma unroll
      for (int c = 0; c < 5; c++) u2[c] = el.interp(a.u2[off + c * a.comp_stride]);
    }
#pragma unroll 1
    for (int ci = 0; ci < 5; ci++) {
      int eo = e;  // opaque: the point data is reloaded for each input component instead of living across the actions
      asm volatile("" : "+v"(eo));
      typename Op::Point p;
      double             qd[10];
      nsop::fetch_qdata<4>(a, eo, t, true, qd);
      Op::template load_point<4>(a, ctx, eo, t, true, qd, u2, p);
      double g[6][5];
#pragma unroll
      for (int m = 0; m < 6; m++)
#pragma unroll
        for (int co = 0; co < 5; co++) g[m][co] = 0;
      with_component(ci, [&](auto cc) {
        constexpr int C = decltype(cc)::value;
#pragma unroll
        for (int s = 0; s < NS; s++) {
          T dq[5], gref[3][5];
          unit_input<T, C>(s, dq, gref);
          double v[5], gw[3][5];
          Op::action(ctx, p, dq, gref, v, gw);
#pragma unroll
          for (int b = 0; b < 4; b++) {
            const int    pr = pair_slot(s, b), m = pair_group(pr);
            const double zF = zf[pair_zfactor(pr)];
#pragma unroll
            for (int co = 0; co < 5; co++) {
              const double x = b == 0 ? v[co] : gw[b - 1][co];
              if (__builtin_constant_p(x) && x == 0.0) continue;  // a weight that folded to zero adds nothing
              g[m][co] = basis3::mfma44(zF, x, g[m][co]);
            }
          }
          __builtin_amdgcn_sched_barrier(0);  // one action's temporaries live at a time
        }
      });
      double r[5] = {0, 0, 0, 0, 0};
      basis3::ContractPairs<4, 4, 5>::yx_groups<5>(g, sF[0], sF[1], sF[2], qx, qy, r);
      if (a.ev) {  // two-pass: the element node's 5 x 5 block in the element-ordered E-vector, summed per node afterwards
        double *dst = a.ev + ((long long)e * Dm::P3 + t) * 25 + ci;
#pragma unroll
        for (int co = 0; co < 5; co++) dst[co * 5] = r[co];
      } else {
#pragma unroll
        for (int co = 0; co < 5; co++) atomicAdd(&a.v[(long long)off * 5 + co * 5 + ci], r[co]);
      }
    }
  }
}

template <int P, int Q, class Op, typename T>
hipError_t launch_diag(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream) {
  constexpr int EPB   = elems_per_block<Q>();
  const int     block = EPB * Q * Q * Q;
  auto          k     = k_diagonal<P, Q, Op, T>;
  const int     grid  = grid_for(k, block, ((long long)a.nelem + EPB - 1) / EPB);
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, stream, a, ctx);
  return hipGetLastError();
}

template <int P, int Q, class Op, typename T>
hipError_t launch_pblock(const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream) {
  if constexpr (P == 4 && Q == 4) {
    auto      k    = k_pblock_diagonal_mf<Op, T>;
    const int grid = grid_for(k, 64, a.nelem);
    hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, stream, a, ctx);
    return hipGetLastError();
  }
  constexpr int EPB   = elems_per_block<Q>();
  const int     block = EPB * Q * Q * Q;
  auto          k     = k_pblock_diagonal<P, Q, Op, T>;
  const int     grid  = grid_for(k, block, ((long long)a.nelem + EPB - 1) / EPB);
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, stream, a, ctx);
  return hipGetLastError();
}

template <int P, int Q>
hipError_t launch_ijacobian_pblock_var(int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  if (fp32) {
    switch (var) {
      case 0: return launch_pblock<P, Q, nsop::IJacobian<0, float>, float>(a, ctx, s);
      case 1: return launch_pblock<P, Q, nsop::IJacobian<1, float>, float>(a, ctx, s);
      default: return launch_pblock<P, Q, nsop::IJacobian<2, float>, float>(a, ctx, s);
    }
  }
  switch (var) {
    case 0: return launch_pblock<P, Q, nsop::IJacobian<0, double>, double>(a, ctx, s);
    case 1: return launch_pblock<P, Q, nsop::IJacobian<1, double>, double>(a, ctx, s);
    default: return launch_pblock<P, Q, nsop::IJacobian<2, double>, double>(a, ctx, s);
  }
}

template <int P, int Q>
hipError_t launch_ijacobian_diag_var(int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  if (fp32) {
    switch (var) {
      case 0: return launch_diag<P, Q, nsop::IJacobian<0, float>, float>(a, ctx, s);
      case 1: return launch_diag<P, Q, nsop::IJacobian<1, float>, float>(a, ctx, s);
      default: return launch_diag<P, Q, nsop::IJacobian<2, float>, float>(a, ctx, s);
    }
  }
  switch (var) {
    case 0: return launch_diag<P, Q, nsop::IJacobian<0, double>, double>(a, ctx, s);
    case 1: return launch_diag<P, Q, nsop::IJacobian<1, double>, double>(a, ctx, s);
    default: return launch_diag<P, Q, nsop::IJacobian<2, double>, double>(a, ctx, s);
  }
}

}  // namespace

hipError_t launch_ijacobian_diagonal(int P, int Q, int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
#define CALL_ID(P_, Q_) (launch_ijacobian_diag_var<P_, Q_>(var, fp32, a, ctx, s))
  HONEE_PQ_ALL(CALL_ID)
#undef CALL_ID
  return hipErrorInvalidValue;
}

hipError_t launch_mass_diagonal(int P, int Q, bool stabilized, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
#define CALL_MD(P_, Q_) \
  (stabilized ? launch_diag<P_, Q_, nsop::MassNewtonian, double>(a, ctx, s) : launch_diag<P_, Q_, nsop::Mass5, double>(a, ctx, s))
  HONEE_PQ_ALL(CALL_MD)
#undef CALL_MD
  return hipErrorInvalidValue;
}

hipError_t launch_ijacobian_point_block_diagonal(int P, int Q, int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx,
                                                 hipStream_t s) {
#define CALL_PB(P_, Q_) (launch_ijacobian_pblock_var<P_, Q_>(var, fp32, a, ctx, s))
  HONEE_PQ_ALL(CALL_PB)
#undef CALL_PB
  return hipErrorInvalidValue;
}

hipError_t launch_mass_point_block_diagonal(int P, int Q, bool stabilized, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
#define CALL_MPB(P_, Q_) \
  (stabilized ? launch_pblock<P_, Q_, nsop::MassNewtonian, double>(a, ctx, s) : launch_pblock<P_, Q_, nsop::Mass5, double>(a, ctx, s))
  HONEE_PQ_ALL(CALL_MPB)
#undef CALL_MPB
  return hipErrorInvalidValue;
}
