// Tensor-product element basis for the fused kernels: sum-factorized interpolation, collocated gradient and their
// transposes, staged through LDS.
//
// Thread layout: each element owns Q^3 consecutive threads of the block, one per quadrature point
// t = qx + Q*(qy + Q*qz); the element's P^3 nodes (P <= Q) use threads t < P^3 for gather and scatter.
// The gradient is computed in collocated form: values are first interpolated to the quadrature points with B, then
// differentiated with the Q x Q collocated derivative Dc (Dc B = G exactly for the element space), and the transpose
// applies Dc^T then B^T -- the same factorization libCEED's reference backend uses for Q >= P.
#pragma once

#include <hip/hip_runtime.h>

namespace basis3 {

template <int P, int Q>
struct Dims {
  static constexpr int P3   = P * P * P;
  static constexpr int Q3   = Q * Q * Q;
  static constexpr int PPQ  = P * P * Q;
  static constexpr int PQQ  = P * Q * Q;
  static constexpr int MAX3 = Q3 > P3 ? Q3 : P3;
};

// Per-element view of the block's LDS: coefficient tables shared by the block and three scratch arrays per element.
template <int P, int Q>
struct Elem {
  using D = Dims<P, Q>;
  const double *B;   // [Q][P]
  const double *Dc;  // [Q][Q]
  double       *s0, *s1, *s2;
  int           t, qx, qy, qz;

  __device__ Elem(const double *B_, const double *Dc_, double *s0_, double *s1_, double *s2_, int t_)
      : B(B_), Dc(Dc_), s0(s0_), s1(s1_), s2(s2_), t(t_) {
    qx = t % Q;
    qy = (t / Q) % Q;
    qz = t / (Q * Q);
  }

  // Interpolate the element's node values (u valid on threads t < P^3) to the quadrature point of this thread.
  // The interpolated value is left in s1 for a following grad(); all threads of the block must call this.
  __device__ double interp(double u) {
    if (t < D::P3) s0[t] = u;
    __syncthreads();
    if (t < D::PPQ) {  // x: s1[k][j][ix] = sum_a B[ix][a] u[k][j][a]
      const int ix = t % Q, kj = t / Q;
      double    acc = 0;
#pragma unroll
      for (int a = 0; a < P; a++) acc += B[ix * P + a] * s0[kj * P + a];
      s1[t] = acc;
    }
    __syncthreads();
    if (t < D::PQQ) {  // y: s2[k][iy][ix] = sum_b B[iy][b] s1[k][b][ix]
      const int ix = t % Q, iy = (t / Q) % Q, k = t / (Q * Q);
      double    acc = 0;
#pragma unroll
      for (int b = 0; b < P; b++) acc += B[iy * P + b] * s1[(k * P + b) * Q + ix];
      s2[t] = acc;
    }
    __syncthreads();
    double acc = 0;  // z: value at (qx, qy, qz)
#pragma unroll
    for (int c = 0; c < P; c++) acc += B[qz * P + c] * s2[(c * Q + qy) * Q + qx];
    return acc;
  }

  // Reference gradient at this thread's quadrature point from interpolated values (one value per thread).
  // Must directly follow interp(): interp's last pass reads only s2, so s1 is free to hold the values.
  __device__ void grad(double val, double g[3]) {
    s1[t] = val;
    __syncthreads();
    double gx = 0, gy = 0, gz = 0;
#pragma unroll
    for (int m = 0; m < Q; m++) {
      gx += Dc[qx * Q + m] * s1[(qz * Q + qy) * Q + m];
      gy += Dc[qy * Q + m] * s1[(qz * Q + m) * Q + qx];
      gz += Dc[qz * Q + m] * s1[(m * Q + qy) * Q + qx];
    }
    g[0] = gx;
    g[1] = gy;
    g[2] = gz;
    __syncthreads();
  }

  // Transpose: given the quadrature-point value v and reference-gradient test-function weights g[3] of this thread,
  // return the element node contribution sum_q B3(q,n) v(q) + sum_q sum_d G3_d(q,n) g_d(q) for node n = t (t < P^3).
  template <bool HAS_V, bool HAS_G>
  __device__ double transpose(double v, const double g[3]) {
    double w = HAS_V ? v : 0.0;
    __syncthreads();  // a preceding interp() may still be reading s2
    if (HAS_G) {
      s0[t]              = g[0];
      s1[t]              = g[1];
      s2[t]              = g[2];
      __syncthreads();
#pragma unroll
      for (int m = 0; m < Q; m++) {
        w += Dc[m * Q + qx] * s0[(qz * Q + qy) * Q + m];
        w += Dc[m * Q + qy] * s1[(qz * Q + m) * Q + qx];
        w += Dc[m * Q + qz] * s2[(m * Q + qy) * Q + qx];
      }
      __syncthreads();
    }
    s0[t] = w;
    __syncthreads();
    if (t < D::PQQ) {  // z^T: s1[c][iy][ix] = sum_qz B[qz][c] w[qz][iy][ix]
      const int ix = t % Q, iy = (t / Q) % Q, c = t / (Q * Q);
      double    acc = 0;
#pragma unroll
      for (int z = 0; z < Q; z++) acc += B[z * P + c] * s0[(z * Q + iy) * Q + ix];
      s1[t] = acc;
    }
    __syncthreads();
    if (t < D::PPQ) {  // y^T: s2[c][b][ix] = sum_qy B[qy][b] s1[c][qy][ix]
      const int ix = t % Q, b = (t / Q) % P, c = t / (Q * P);
      double    acc = 0;
#pragma unroll
      for (int y = 0; y < Q; y++) acc += B[y * P + b] * s1[(c * Q + y) * Q + ix];
      s2[t] = acc;
    }
    __syncthreads();
    double r = 0;
    if (t < D::P3) {  // x^T: r[c][b][a] = sum_qx B[qx][a] s2[c][b][qx]
      const int a = t % P, cb = t / P;
#pragma unroll
      for (int x = 0; x < Q; x++) r += B[x * P + a] * s2[cb * Q + x];
    }
    __syncthreads();
    return r;
  }

  // Node value sum_q Fz[qz][c] Fy[qy][b] Fx[qx][a] w(q) for node t = (a, b, c), with arbitrary 1D matrices [Q][P]
  // (used by diagonal assembly with products of B and G).  Every thread passes its quadrature-point value w.
  __device__ double contract_t(double w, const double *Fx, const double *Fy, const double *Fz) {
    __syncthreads();
    s0[t] = w;
    __syncthreads();
    if (t < D::PQQ) {
      const int ix = t % Q, iy = (t / Q) % Q, c = t / (Q * Q);
      double    acc = 0;
#pragma unroll
      for (int z = 0; z < Q; z++) acc += Fz[z * P + c] * s0[(z * Q + iy) * Q + ix];
      s1[t] = acc;
    }
    __syncthreads();
    if (t < D::PPQ) {
      const int ix = t % Q, b = (t / Q) % P, c = t / (Q * P);
      double    acc = 0;
#pragma unroll
      for (int y = 0; y < Q; y++) acc += Fy[y * P + b] * s1[(c * Q + y) * Q + ix];
      s2[t] = acc;
    }
    __syncthreads();
    double r = 0;
    if (t < D::P3) {
      const int a = t % P, cb = t / P;
#pragma unroll
      for (int x = 0; x < Q; x++) r += Fx[x * P + a] * s2[cb * Q + x];
    }
    return r;
  }
};

}  // namespace basis3
