// Helpers shared by the fused volume kernels (kernels_volume.hip) and the diagonal assembly kernels
// (kernels_assembly.hip): block shapes, the XCD-aware element-batch schedule, basis-table staging, the persistent grid
// size and the (P_1d, Q_1d) dispatch list.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "basis3d.hpp"
#include "kernels.hpp"
#include "newtonian_ops.hpp"

namespace {

using basis3::Dims;
using basis3::Elem;

template <int Q>
constexpr int elems_per_block() {
  return (Q * Q * Q >= 64) ? 1 : 64 / (Q * Q * Q);
}

// Element-batch schedule of the persistent blocks.  The workgroups of a dispatch are distributed round-robin over
// the 8 XCDs of an MI355X, each with its own L2 (block b runs on XCD b % 8).  Each XCD gets one contiguous range of
// element batches, which its blocks sweep together, so the nodes that neighbouring elements share are fetched into
// one L2 once instead of into several.
constexpr int kNumXcd = 8;

struct BatchRange {
  long long first, end, stride;
};

__device__ __forceinline__ BatchRange batch_range(long long nbatch) {
  const int G = gridDim.x, b = blockIdx.x;
  if (G % kNumXcd != 0) return {b, nbatch, G};
  const long long per = (nbatch + kNumXcd - 1) / kNumXcd;
  const int       x = b % kNumXcd, j = b / kNumXcd;
  const long long lo = x * per, hi = (x + 1) * per < nbatch ? (x + 1) * per : nbatch;
  return {lo + j, hi, G / kNumXcd};
}

// The basis stage of the fused kernel: for the 4 x 4 x 4 element the x and z passes run on the FP64 matrix cores and the
// y passes across lanes (basis3::StageMF), except for ops with Op::kZCrossLane (the FP32 Jacobian, measured faster with
// all passes across lanes); otherwise sum factorization through LDS (basis3::Stage, ZX: z passes across lanes)
template <int P, int Q, bool ZX>
struct StageFor {
  using type = basis3::Stage<P, Q, ZX>;
};
template <>
struct StageFor<4, 4, false> {
  using type = basis3::StageMF;
};

// Op::kZCrossLane selects the basis stage whose z passes also run across lanes (basis3::Stage's ZX); default off
template <class Op, class = void>
struct ZCross {
  static constexpr bool value = false;
};
template <class Op>
struct ZCross<Op, std::void_t<decltype(Op::kZCrossLane)>> {
  static constexpr bool value = Op::kZCrossLane;
};

template <int P, int Q>
__device__ __forceinline__ void load_tables(const BasisTables &bt, double *sB, double *sDc) {
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) sB[i] = bt.B[i];
  for (int i = threadIdx.x; i < Q * Q; i += blockDim.x) sDc[i] = bt.Dc[i];
}
// with the derivative table G [Q][P] (formed as Dc B when the basis does not provide it)
template <int P, int Q>
__device__ __forceinline__ void load_tables(const BasisTables &bt, double *sB, double *sDc, double *sG) {
  load_tables<P, Q>(bt, sB, sDc);
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) {
    if (bt.G) {
      sG[i] = bt.G[i];
    } else {
      const int q = i / P, n = i % P;
      double    g = 0;
      for (int m = 0; m < Q; m++) g += bt.Dc[q * Q + m] * bt.B[m * P + n];
      sG[i] = g;
    }
  }
}

template <typename K>
int grid_for(K kernel, int block, long long nbatch) {
  int dev = 0, ncu = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
  // one persistent block per resident slot: each XCD's blocks then sweep a contiguous window of its batch range
  long long g = (long long)ncu * (per_cu > 0 ? per_cu : 1);
  if (g > nbatch) g = nbatch >= kNumXcd ? nbatch / kNumXcd * kNumXcd : nbatch;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

#define HONEE_PQ_LIST(X) X(2, 2) X(3, 3) X(4, 4) X(5, 5) X(2, 3) X(2, 4) X(3, 4) X(3, 5) X(4, 5) X(4, 6)
#define HONEE_PQ_DISPATCH(P_, Q_, CALL) \
  if (P == P_ && Q == Q_) return CALL;
// P_1d = degree + 1 for degrees 1-4 with Q_1d = P_1d, and Q_1d = P_1d + q_extra for q_extra 1, 2 at degrees 1-3
// (HONEE's -q_extra, src/cloptions.c:152-153)
#define HONEE_PQ_ALL(CALL_T)          \
  HONEE_PQ_DISPATCH(2, 2, CALL_T(2, 2)) \
  HONEE_PQ_DISPATCH(3, 3, CALL_T(3, 3)) \
  HONEE_PQ_DISPATCH(4, 4, CALL_T(4, 4)) \
  HONEE_PQ_DISPATCH(5, 5, CALL_T(5, 5)) \
  HONEE_PQ_DISPATCH(2, 3, CALL_T(2, 3)) \
  HONEE_PQ_DISPATCH(2, 4, CALL_T(2, 4)) \
  HONEE_PQ_DISPATCH(3, 4, CALL_T(3, 4)) \
  HONEE_PQ_DISPATCH(3, 5, CALL_T(3, 5)) \
  HONEE_PQ_DISPATCH(4, 5, CALL_T(4, 5)) \
  HONEE_PQ_DISPATCH(4, 6, CALL_T(4, 6))
