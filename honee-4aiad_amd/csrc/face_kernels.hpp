// Kernel templates of the fused boundary-face operators on MI355X, shared by the translation units that instantiate
// them for the different face bodies (kernels_face.hip: boundary integral and freestream; kernels_face_bc.hip:
// outflow and slip).
//
// k_face<P, Q, Op, Ctx> applies one face sub-operator in one launch: per face it gathers the L-vector through the face
// restriction, interpolates values (and the two reference tangential derivatives, for bodies with kGrad) to the
// Q x Q face quadrature points in LDS, evaluates the face body (face_ops.hpp), applies the interpolation transpose and
// scatters (FP64 atomics, or the E-vector of the deterministic two-pass scatter).  Each face owns Q^2 consecutive
// threads; a 64-thread block holds FPB faces.  k_face_diagonal and k_face_pblock assemble the Jacobi diagonal and the
// 5x5 node blocks of a face Jacobian body.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "face_ops.hpp"
#include "kernels.hpp"

namespace {

template <int Q>
constexpr int faces_per_block() {
  return (Q * Q >= 64) ? 1 : 64 / (Q * Q);
}

// SetupBoundary with NormalVectorFromdxdX_3D (setupgeo_helpers.h:78-91) and InvertBoundaryMappingJacobian_3D
// (setupgeo_helpers.h:143-171); one thread per face quadrature point.
template <int Q>
__global__ void __launch_bounds__(256) k_setup_boundary(SetupArgs a) {
  constexpr int   Q2 = Q * Q;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)a.nelem * Q2) return;
  const int f = (int)(gid / Q2), i = (int)(gid % Q2), qx = i % Q, qy = i / Q;
  double    dxdX[3][2] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
  for (int n = 0; n < 4; n++) {
    const int    xo = a.xoffsets[(long long)f * 4 + n];
    const int    nx = n & 1, ny = n >> 1;
    const double b1 = a.Bx[qx * 2 + nx], d1 = a.Dx[qx * 2 + nx], b2 = a.Bx[qy * 2 + ny], d2 = a.Dx[qy * 2 + ny];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double xv = a.x[xo + c * a.xcomp_stride];
      dxdX[c][0] += d1 * b2 * xv;
      dxdX[c][1] += b1 * d2 * xv;
    }
  }
  double normal[3];
  normal[0]          = dxdX[1][0] * dxdX[2][1] - dxdX[2][0] * dxdX[1][1];
  normal[1]          = dxdX[2][0] * dxdX[0][1] - dxdX[0][0] * dxdX[2][1];
  normal[2]          = dxdX[0][0] * dxdX[1][1] - dxdX[1][0] * dxdX[0][1];
  const double detJb = sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
#pragma unroll
  for (int d = 0; d < 3; d++) normal[d] *= 1 / detJb;
  double JTJ[2][2] = {{0., 0.}, {0., 0.}};
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
      for (int l = 0; l < 3; l++) JTJ[j][k] += dxdX[l][j] * dxdX[l][k];
  const double det       = JTJ[0][0] * JTJ[1][1] - JTJ[1][0] * JTJ[0][1];
  const double inv[2][2] = {{JTJ[1][1] / det, -JTJ[0][1] / det}, {-JTJ[1][0] / det, JTJ[0][0] / det}};
  const double w         = a.w1d[qx] * a.w1d[qy];
  double      *qd        = a.qdata + (long long)f * a.qd.se + i;
  qd[0]                  = w * detJb;
#pragma unroll
  for (int d = 0; d < 3; d++) qd[(1 + d) * Q2] = normal[d];
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double acc = 0;
#pragma unroll
      for (int l = 0; l < 2; l++) acc += inv[l][j] * dxdX[k][l];
      qd[(4 + 3 * j + k) * Q2] = acc;
    }
}

// Per-face LDS stage: X[slot][Q^2].  Slot map: values 0..4, reference tangential derivatives 5 + 5 d + c (d = 0, 1);
// scratch 15..19.  Outputs (interpolation test weights) reuse slots 0..4.
template <int P, int Q, class Op, class Ctx>
__global__ void __launch_bounds__(faces_per_block<Q>() * Q * Q) k_face(VolumeArgs a, const Ctx ctx) {
  static_assert(!std::is_reference<Ctx>::value, "the context is passed to the kernel by value");
  constexpr int FPB = faces_per_block<Q>();
  constexpr int Q2 = Q * Q, P2 = P * P, PQ = P * Q;
  constexpr int NST = 20;
  __shared__ double sB[Q * P], sDc[Q * Q];
  __shared__ double sX[FPB][NST][Q2];
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) sB[i] = a.bt.B[i];
  for (int i = threadIdx.x; i < Q * Q; i += blockDim.x) sDc[i] = a.bt.Dc[i];
  const int le = threadIdx.x / Q2, t = threadIdx.x % Q2, qx = t % Q, qy = t / Q;
  const bool nd = t < P2;
  double(*X)[Q2] = sX[le];
  const long long nbatch = ((long long)a.nelem + FPB - 1) / FPB;

  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long f_  = bt * FPB + le;
    const bool      ok  = f_ < a.nelem;
    const int       f   = ok ? (int)f_ : 0;
    const int       off = (ok && nd) ? a.offsets[(long long)f * P2 + t] : 0;
    double          un[5];
#pragma unroll
    for (int c = 0; c < 5; c++) un[c] = (ok && nd) ? a.u[off + c * a.comp_stride] : 0.0;
    __syncthreads();  // the previous batch's transpose has finished reading the stage
    if (nd) {
#pragma unroll
      for (int c = 0; c < 5; c++) X[15 + c][t] = un[c];
    }
    __syncthreads();
    if (t < PQ) {  // x: X[5 + c][j][ix] = sum_i B[ix][i] u[j][i]
      const int ix = t % Q, j = t / Q;
#pragma unroll
      for (int c = 0; c < 5; c++) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < P; i++) acc += sB[ix * P + i] * X[15 + c][j * P + i];
        X[5 + c][t] = acc;
      }
    }
    __syncthreads();
    double u[5];
#pragma unroll
    for (int c = 0; c < 5; c++) {  // y: value at (qx, qy)
      double acc = 0;
#pragma unroll
      for (int j = 0; j < P; j++) acc += sB[qy * P + j] * X[5 + c][j * Q + qx];
      u[c] = acc;
    }
    double gref[2][5];
    if constexpr (Op::kGrad) {  // collocated tangential derivatives
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 5; c++) X[c][t] = u[c];
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 5; c++) {
        double g0 = 0, g1 = 0;
#pragma unroll
        for (int m = 0; m < Q; m++) {
          g0 += sDc[qx * Q + m] * X[c][qy * Q + m];
          g1 += sDc[qy * Q + m] * X[c][m * Q + qx];
        }
        gref[0][c] = g0;
        gref[1][c] = g1;
      }
    }
    double v[5];
    Op::template qf<Q>(a, ctx, f, t, ok, u, gref, v);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 5; c++) X[c][t] = v[c];
    __syncthreads();
    if (t < PQ) {  // y^T: X[5 + c][j][ix] = sum_qy B[qy][j] v[qy][ix]
      const int ix = t % Q, j = t / Q;
#pragma unroll
      for (int c = 0; c < 5; c++) {
        double acc = 0;
#pragma unroll
        for (int y = 0; y < Q; y++) acc += sB[y * P + j] * X[c][y * Q + ix];
        X[5 + c][t] = acc;
      }
    }
    __syncthreads();
    if (ok && nd) {  // x^T and scatter: node (i, j) = t
      const int i = t % P, j = t / P;
      double    r[5];
#pragma unroll
      for (int c = 0; c < 5; c++) {
        double acc = 0;
#pragma unroll
        for (int x = 0; x < Q; x++) acc += sB[x * P + i] * X[5 + c][j * Q + x];
        r[c] = acc;
      }
      if (a.ev) {
        const long long en  = (long long)f * P2 + t;
        double         *dst = a.ev + (a.perm ? (long long)a.perm[en] : en) * 5;
#pragma unroll
        for (int c = 0; c < 5; c++) dst[c] = r[c];
      } else {
#pragma unroll
        for (int c = 0; c < 5; c++) atomicAdd(&a.v[off + c * a.comp_stride], r[c]);
      }
    }
  }
}

// Diagonal of a face Jacobian (CeedOperatorLinearAssembleDiagonal for the face sub-operators).  At each point and for
// each component c, the linear body is applied to unit inputs in slot s (value, d/dX1, d/dX2 of component c), giving
// M_s = dv_c/d(slot s); the node diagonal is sum_q M_0 (B.B)(B.B) + M_1 (G.B)(B.B) + M_2 (B.B)(G.B) with the 1D
// products of the face basis, contracted in two LDS passes.
template <int P, int Q, class Op, class Ctx>
__global__ void __launch_bounds__(faces_per_block<Q>() * Q * Q) k_face_diagonal(VolumeArgs a, const Ctx ctx) {
  static_assert(!std::is_reference<Ctx>::value, "the context is passed to the kernel by value");
  constexpr int FPB = faces_per_block<Q>();
  constexpr int Q2 = Q * Q, P2 = P * P, PQ = P * Q;
  __shared__ double sBB[Q * P], sGB[Q * P];
  __shared__ double sM[FPB][3][Q2], sT[FPB][3][PQ];
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) {
    sBB[i] = a.bt.BB[i];
    sGB[i] = a.bt.GB[i];
  }
  const int  le = threadIdx.x / Q2, t = threadIdx.x % Q2;
  const bool nd = t < P2;
  const long long nbatch = ((long long)a.nelem + FPB - 1) / FPB;
  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long f_  = bt * FPB + le;
    const bool      ok  = f_ < a.nelem;
    const int       f   = ok ? (int)f_ : 0;
    const int       off = (ok && nd) ? a.offsets[(long long)f * P2 + t] : 0;
    for (int c = 0; c < 5; c++) {
      double M[3];
#pragma unroll
      for (int s = 0; s < 3; s++) {
        double u[5] = {0, 0, 0, 0, 0}, gref[2][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}}, v[5];
        if (s == 0) u[c] = 1;
        else gref[s - 1][c] = 1;
        if (s > 0 && !Op::kGrad) {
          M[s] = 0;
          continue;
        }
        Op::template qf<Q>(a, ctx, f, t, ok, u, gref, v);
        M[s] = ok ? v[c] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < 3; s++) sM[le][s][t] = M[s];
      __syncthreads();
      if (t < PQ) {  // contract X2: T_s[j][qx] = sum_qy F2_s[qy][j] M_s[qy][qx]
        const int qx = t % Q, j = t / Q;
#pragma unroll
        for (int s = 0; s < 3; s++) {
          const double *F2  = (s == 2) ? sGB : sBB;
          double        acc = 0;
#pragma unroll
          for (int y = 0; y < Q; y++) acc += F2[y * P + j] * sM[le][s][y * Q + qx];
          sT[le][s][t] = acc;
        }
      }
      __syncthreads();
      if (ok && nd) {  // contract X1 and sum the slots: node (i, j) = t
        const int i = t % P, j = t / P;
        double    r = 0;
#pragma unroll
        for (int s = 0; s < 3; s++) {
          const double *F1 = (s == 1) ? sGB : sBB;
#pragma unroll
          for (int x = 0; x < Q; x++) r += F1[x * P + i] * sT[le][s][j * Q + x];
        }
        atomicAdd(&a.v[off + c * a.comp_stride], r);
      }
    }
  }
}

// Point-block diagonal of a face Jacobian: as k_face_diagonal with the five output components of each unit-input action
// contracted together; output [node][c_out][c_in] at 25 * (offset / 5).
template <int P, int Q, class Op, class Ctx>
__global__ void __launch_bounds__(faces_per_block<Q>() * Q * Q) k_face_pblock(VolumeArgs a, const Ctx ctx) {
  static_assert(!std::is_reference<Ctx>::value, "the context is passed to the kernel by value");
  constexpr int FPB = faces_per_block<Q>();
  constexpr int Q2 = Q * Q, P2 = P * P, PQ = P * Q;
  __shared__ double sBB[Q * P], sGB[Q * P];
  __shared__ double sM[FPB][3][5][Q2], sT[FPB][3][5][PQ];
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) {
    sBB[i] = a.bt.BB[i];
    sGB[i] = a.bt.GB[i];
  }
  const int       le = threadIdx.x / Q2, t = threadIdx.x % Q2;
  const bool      nd = t < P2;
  const long long nbatch = ((long long)a.nelem + FPB - 1) / FPB;
  for (long long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long long f_  = bt * FPB + le;
    const bool      ok  = f_ < a.nelem;
    const int       f   = ok ? (int)f_ : 0;
    const int       off = (ok && nd) ? a.offsets[(long long)f * P2 + t] : 0;
    for (int ci = 0; ci < 5; ci++) {
      double M[3][5];
#pragma unroll
      for (int s = 0; s < 3; s++) {
        double u[5] = {0, 0, 0, 0, 0}, gref[2][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}}, v[5];
        if (s == 0) u[ci] = 1;
        else gref[s - 1][ci] = 1;
        if (s > 0 && !Op::kGrad) {
#pragma unroll
          for (int co = 0; co < 5; co++) M[s][co] = 0;
          continue;
        }
        Op::template qf<Q>(a, ctx, f, t, ok, u, gref, v);
#pragma unroll
        for (int co = 0; co < 5; co++) M[s][co] = ok ? v[co] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < 3; s++)
#pragma unroll
        for (int co = 0; co < 5; co++) sM[le][s][co][t] = M[s][co];
      __syncthreads();
      if (t < PQ) {  // contract X2
        const int qx = t % Q, j = t / Q;
#pragma unroll
        for (int s = 0; s < 3; s++) {
          const double *F2 = (s == 2) ? sGB : sBB;
#pragma unroll
          for (int co = 0; co < 5; co++) {
            double acc = 0;
#pragma unroll
            for (int y = 0; y < Q; y++) acc += F2[y * P + j] * sM[le][s][co][y * Q + qx];
            sT[le][s][co][t] = acc;
          }
        }
      }
      __syncthreads();
      if (ok && nd) {  // contract X1, sum the slots
        const int i = t % P, j = t / P;
#pragma unroll
        for (int co = 0; co < 5; co++) {
          double r = 0;
#pragma unroll
          for (int s = 0; s < 3; s++) {
            const double *F1 = (s == 1) ? sGB : sBB;
#pragma unroll
            for (int x = 0; x < Q; x++) r += F1[x * P + i] * sT[le][s][co][j * Q + x];
          }
          atomicAdd(&a.v[(long long)off * 5 + co * 5 + ci], r);
        }
      }
    }
  }
}

template <int P, int Q, class Op, class Ctx>
hipError_t launch_face_pblock(const VolumeArgs &a, const Ctx &ctx, hipStream_t stream) {
  constexpr int   FPB    = faces_per_block<Q>();
  const int       block  = FPB * Q * Q;
  const long long nbatch = ((long long)a.nelem + FPB - 1) / FPB;
  int             dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  long long grid = (long long)ncu * 8;
  if (grid > nbatch) grid = nbatch;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((k_face_pblock<P, Q, Op, Ctx>), dim3((unsigned)grid), dim3(block), 0, stream, a, ctx);
  return hipGetLastError();
}

template <int P, int Q, class Op, class Ctx>
hipError_t launch_face_diag(const VolumeArgs &a, const Ctx &ctx, hipStream_t stream) {
  constexpr int   FPB    = faces_per_block<Q>();
  const int       block  = FPB * Q * Q;
  const long long nbatch = ((long long)a.nelem + FPB - 1) / FPB;
  int             dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  long long grid = (long long)ncu * 8;
  if (grid > nbatch) grid = nbatch;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((k_face_diagonal<P, Q, Op, Ctx>), dim3((unsigned)grid), dim3(block), 0, stream, a, ctx);
  return hipGetLastError();
}

template <int P, int Q, class Op, class Ctx>
hipError_t launch_face_op(const VolumeArgs &a, const Ctx &ctx, hipStream_t stream) {
  constexpr int   FPB    = faces_per_block<Q>();
  const int       block  = FPB * Q * Q;
  const long long nbatch = ((long long)a.nelem + FPB - 1) / FPB;
  int             dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  long long grid = (long long)ncu * 8;
  if (grid > nbatch) grid = nbatch;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((k_face<P, Q, Op, Ctx>), dim3((unsigned)grid), dim3(block), 0, stream, a, ctx);
  return hipGetLastError();
}

}  // namespace

// the volume kernels' (P_1d, Q_1d) pairs: Q_1d = P_1d, and Q_1d = P_1d + q_extra (1, 2) at degrees 1-3
#define HONEE_FACE_PQ(CALL)                   \
  if (P == 2 && Q == 2) return CALL(2, 2);   \
  if (P == 3 && Q == 3) return CALL(3, 3);   \
  if (P == 4 && Q == 4) return CALL(4, 4);   \
  if (P == 5 && Q == 5) return CALL(5, 5);   \
  if (P == 2 && Q == 3) return CALL(2, 3);   \
  if (P == 2 && Q == 4) return CALL(2, 4);   \
  if (P == 3 && Q == 4) return CALL(3, 4);   \
  if (P == 3 && Q == 5) return CALL(3, 5);   \
  if (P == 4 && Q == 5) return CALL(4, 5);   \
  if (P == 4 && Q == 6) return CALL(4, 6);
