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


// Five-component form of Elem::contract_t for point-block diagonal assembly: r[k] += sum_q Fz[qz][c] Fy[qy][b]
// Fx[qx][a] w[k](q) for node t = (a, b, c), the five components contracted together in each pass.  s0, s1, s2 are the
// element's [5][MAX3] scratch arrays.
template <int P, int Q>
struct Contract5 {
  using D = Dims<P, Q>;
  double (*s0)[D::MAX3];
  double (*s1)[D::MAX3];
  double (*s2)[D::MAX3];
  int t;

  __device__ void add(const double (&w)[5], const double *Fx, const double *Fy, const double *Fz, double (&r)[5]) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 5; k++) s0[k][t] = w[k];
    add_staged(s0, Fx, Fy, Fz, r);
  }

  // As add(), with the five weights already in src[k][q] (written by every thread before the call)
  __device__ void add_staged(const double (*src)[D::MAX3], const double *Fx, const double *Fy, const double *Fz, double (&r)[5]) {
    __syncthreads();
    if (t < D::PQQ) {
      const int ix = t % Q, iy = (t / Q) % Q, c = t / (Q * Q);
#pragma unroll
      for (int k = 0; k < 5; k++) {
        double acc = 0;
#pragma unroll
        for (int z = 0; z < Q; z++) acc += Fz[z * P + c] * src[k][(z * Q + iy) * Q + ix];
        s1[k][t] = acc;
      }
    }
    __syncthreads();
    if (t < D::PPQ) {
      const int ix = t % Q, b = (t / Q) % P, c = t / (Q * P);
#pragma unroll
      for (int k = 0; k < 5; k++) {
        double acc = 0;
#pragma unroll
        for (int y = 0; y < Q; y++) acc += Fy[y * P + b] * s1[k][(c * Q + y) * Q + ix];
        s2[k][t] = acc;
      }
    }
    __syncthreads();
    if (t < D::P3) {
      const int a = t % P, cb = t / P;
#pragma unroll
      for (int k = 0; k < 5; k++)
#pragma unroll
        for (int x = 0; x < Q; x++) r[k] += Fx[x * P + a] * s2[k][cb * Q + x];
    }
  }
};

// Multi-component basis stage: the element's LDS stage X[slot][Q^3] holds one slot per (field component, point);
// every pass treats NC components together, so each 1D coefficient row is read from LDS once per pass and the NC
// independent contractions overlap their LDS latency.  Barriers are block-wide; a pass reads other threads' slots only
// after a barrier that follows their writes.
// Cross-lane access for the 4 x 4 x 4 element (P = Q = 4, one 64-lane wave per element, lane = qx + 4 qy + 16 qz):
// the four lanes of an x line form a DPP quad and the sixteen lanes of an xy slab form a DPP row.  dpp_d<CTRL> returns
// the double held by the lane that the DPP control CTRL selects (quad_perm broadcast 0x00, 0x55, 0xAA, 0xFF of quad
// lane 0..3; row_ror 0x124, 0x128, 0x12C by 4, 8 and 12 lanes within the row), moving the two 32-bit halves.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b  = __double_as_longlong(v);
  const int       lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
  const int       hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// sum_a c[a] v(lane a of this quad), a ascending (the x line's contraction)
__device__ __forceinline__ double quad_contract(const double c[4], double v) {
  double acc = 0;
  acc += c[0] * dpp_d<0x00>(v);
  acc += c[1] * dpp_d<0x55>(v);
  acc += c[2] * dpp_d<0xAA>(v);
  acc += c[3] * dpp_d<0xFF>(v);
  return acc;
}

// sum_n c[n] v(lane rotated by 4 n within this row): the y line's contraction with coefficients already permuted to the
// rotation order (see RowRot)
__device__ __forceinline__ double row_contract(const double c[4], double v) {
  double acc = 0;
  acc += c[0] * v;
  acc += c[1] * dpp_d<0x124>(v);
  acc += c[2] * dpp_d<0x128>(v);
  acc += c[3] * dpp_d<0x12C>(v);
  return acc;
}

// The four values of this lane's z line (the lanes with the same qx and qy), z ascending: a 32-lane swap gives every lane
// the values of lanes L & 31 and (L & 31) + 32, and a 16-lane swap of each gives lanes with bit 4 cleared and set
// (v_permlane32_swap, v_permlane16_swap; both operands the same register)
__device__ __forceinline__ void z_line(double v, double z[4]) {
  const long long b = __double_as_longlong(v);
  unsigned        h[2] = {(unsigned)b, (unsigned)(b >> 32)};
  unsigned        zz[2][4];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const auto a  = __builtin_amdgcn_permlane32_swap(h[k], h[k], false, false);
    const auto lo = __builtin_amdgcn_permlane16_swap(a[0], a[0], false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(a[1], a[1], false, false);
    zz[k][0] = lo[0], zz[k][1] = lo[1], zz[k][2] = hi[0], zz[k][3] = hi[1];
  }
#pragma unroll
  for (int i = 0; i < 4; i++) z[i] = __longlong_as_double(((long long)zz[1][i] << 32) | (long long)zz[0][i]);
}

// sum_z c[z] v(lane of this z line at z), z ascending
__device__ __forceinline__ double z_contract(const double c[4], double v) {
  double z[4];
  z_line(v, z);
  double acc = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) acc += c[i] * z[i];
  return acc;
}

// The y index of the lane each row rotation reads, found by rotating the lanes' own y index with the same DPP controls
// (independent of the rotation's direction)
struct RowRot {
  int y[4];
  __device__ explicit RowRot(int qy) {
    y[0] = qy;
    y[1] = __builtin_amdgcn_mov_dpp(qy, 0x124, 0xF, 0xF, true);
    y[2] = __builtin_amdgcn_mov_dpp(qy, 0x128, 0xF, 0xF, true);
    y[3] = __builtin_amdgcn_mov_dpp(qy, 0x12C, 0xF, 0xF, true);
  }
};

// Diagonal-assembly contraction of the ten slot-pair weights W[pair][k][q] (k < NC components; pairs 0..3 the diagonal
// pairs of {value, d/dX0, d/dX1, d/dX2}, then (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)) with their 1D factor triples
// (Fx, Fy, Fz) drawn from BB = B.B, GB = G.B and GG = G.G:
//   0 (BB,BB,BB)  1 (GG,BB,BB)  2 (BB,GG,BB)  3 (BB,BB,GG)  4 (GB,BB,BB)
//   5 (BB,GB,BB)  6 (BB,BB,GB)  7 (GB,GB,BB)  8 (GB,BB,GB)  9 (BB,GB,GB).
// r[k] += sum_pair sum_q Fz Fy Fx W[pair][k](q) for node t.  Sum factorization with grouping: after the z pass the
// pairs that share (Fx, Fy) are summed (10 z passes -> 6 arrays), after the y pass those that share Fx (6 y passes
// -> 3 arrays), so 19 passes replace 30.  W is overwritten.  Barriers are block-wide; W complete before the call.
template <int P, int Q, int NC>
struct ContractPairs {
  using D = Dims<P, Q>;
  int t;

  __device__ void zacc(const double (*src)[D::MAX3], const double *Fz, double (&acc)[NC]) const {
    if (t < D::PQQ) {
      const int ix = t % Q, iy = (t / Q) % Q, c = t / (Q * Q);
#pragma unroll
      for (int k = 0; k < NC; k++)
#pragma unroll
        for (int z = 0; z < Q; z++) acc[k] += Fz[z * P + c] * src[k][(z * Q + iy) * Q + ix];
    }
  }
  __device__ void yacc(const double (*src)[D::MAX3], const double *Fy, double (&acc)[NC]) const {
    if (t < D::PPQ) {
      const int ix = t % Q, b = (t / Q) % P, c = t / (Q * P);
#pragma unroll
      for (int k = 0; k < NC; k++)
#pragma unroll
        for (int y = 0; y < Q; y++) acc[k] += Fy[y * P + b] * src[k][(c * Q + y) * Q + ix];
    }
  }
  __device__ void xacc(const double (*src)[D::MAX3], const double *Fx, double (&acc)[NC]) const {
    if (t < D::P3) {
      const int a = t % P, cb = t / P;
#pragma unroll
      for (int k = 0; k < NC; k++)
#pragma unroll
        for (int x = 0; x < Q; x++) acc[k] += Fx[x * P + a] * src[k][cb * Q + x];
    }
  }
  // every thread has read the group's inputs before any thread overwrites the first of them
  __device__ void put(double (*dst)[D::MAX3], const double (&acc)[NC], int n) const {
    __syncthreads();
    if (t < n)
#pragma unroll
      for (int k = 0; k < NC; k++) dst[k][t] = acc[k];
  }

  __device__ void add(double (*W)[NC][D::MAX3], const double *BB, const double *GB, const double *GG, double (&r)[NC]) const {
    if constexpr (P == 4 && Q == 4) {  // one wave per element: the y and x passes across lanes, on registers
      add_cross_lane(W, BB, GB, GG, r);
      return;
    }
    double a[NC];
    __syncthreads();
    // z pass, grouped by (Fx, Fy); each group's sum goes to its first pair's slot (read by no other group)
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = 0;
    zacc(W[0], BB, a), zacc(W[6], GB, a), zacc(W[3], GG, a);
    put(W[0], a, D::PQQ);
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = 0;
    zacc(W[4], BB, a), zacc(W[8], GB, a);
    put(W[4], a, D::PQQ);
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = 0;
    zacc(W[5], BB, a), zacc(W[9], GB, a);
    put(W[5], a, D::PQQ);
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = 0;
    zacc(W[1], BB, a);
    put(W[1], a, D::PQQ);
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = 0;
    zacc(W[2], BB, a);
    put(W[2], a, D::PQQ);
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = 0;
    zacc(W[7], BB, a);
    put(W[7], a, D::PQQ);
    __syncthreads();
    // y pass, grouped by Fx: (BB,BB) (BB,GB) (BB,GG) in slots 0, 5, 2; (GB,BB) (GB,GB) in 4, 7; (GG,BB) in 1
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = 0;
    yacc(W[0], BB, a), yacc(W[5], GB, a), yacc(W[2], GG, a);
    put(W[0], a, D::PPQ);
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = 0;
    yacc(W[4], BB, a), yacc(W[7], GB, a);
    put(W[4], a, D::PPQ);
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = 0;
    yacc(W[1], BB, a);
    put(W[1], a, D::PPQ);
    __syncthreads();
    // x pass
    xacc(W[0], BB, r), xacc(W[4], GB, r), xacc(W[1], GG, r);
  }

  // The 4 x 4 x 4 element (lane = qx + 4 qy + 16 qz): the z passes read W from LDS into registers, the six (Fx, Fy)
  // group sums stay in registers, and the y and x passes run across lanes (DPP row rotations and quad broadcasts, as in
  // Stage's transpose).  All 64 lanes must be active.
  __device__ void add_cross_lane(double (*W)[NC][D::MAX3], const double *BB, const double *GB, const double *GG, double (&r)[NC]) const {
    static_assert(P == 4 && Q == 4, "cross-lane passes need the 4 x 4 x 4 element");
    const int qx = t & 3, qy = (t >> 2) & 3;
    double    g[6][NC];
#pragma unroll
    for (int m = 0; m < 6; m++)
#pragma unroll
      for (int k = 0; k < NC; k++) g[m][k] = 0;
    __syncthreads();  // W complete
    zacc(W[0], BB, g[0]), zacc(W[6], GB, g[0]), zacc(W[3], GG, g[0]);  // (BB, BB)
    zacc(W[4], BB, g[1]), zacc(W[8], GB, g[1]);                        // (GB, BB)
    zacc(W[5], BB, g[2]), zacc(W[9], GB, g[2]);                        // (BB, GB)
    zacc(W[1], BB, g[3]);                                              // (GG, BB)
    zacc(W[2], BB, g[4]);                                              // (BB, GG)
    zacc(W[7], BB, g[5]);                                              // (GB, GB)
    yx_groups<NC>(g, BB, GB, GG, qx, qy, r);
  }

  // The y and x passes of the six z-contracted (Fx, Fy) groups g[m][k] of the 4 x 4 x 4 element (lane layout kept:
  // qx, qy, node z), across lanes: r[k] += sum of the groups' y then x contractions
  template <int K>
  __device__ static void yx_groups(const double (&g)[6][K], const double *BB, const double *GB, const double *GG, int qx, int qy,
                                   double (&r)[K]) {
    const RowRot rr(qy);
    double       yBB[4], yGB[4], yGG[4], xBB[4], xGB[4], xGG[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      yBB[i] = BB[rr.y[i] * 4 + qy], yGB[i] = GB[rr.y[i] * 4 + qy], yGG[i] = GG[rr.y[i] * 4 + qy];
      xBB[i] = BB[i * 4 + qx], xGB[i] = GB[i * 4 + qx], xGG[i] = GG[i * 4 + qx];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double h0 = row_contract(yBB, g[0][k]) + row_contract(yGB, g[2][k]) + row_contract(yGG, g[4][k]);  // Fx = BB
      const double h1 = row_contract(yBB, g[1][k]) + row_contract(yGB, g[5][k]);                                // Fx = GB
      const double h2 = row_contract(yBB, g[3][k]);                                                             // Fx = GG
      r[k] += quad_contract(xBB, h0) + quad_contract(xGB, h1) + quad_contract(xGG, h2);
    }
  }
};

// ZX: the z passes of the 4 x 4 x 4 element also across lanes (v_permlane swaps) instead of through LDS; measured
// faster for the FP32 Jacobian and slower for the FP64 operators
template <int P, int Q, bool ZX = false>
struct Stage {
  using D = Dims<P, Q>;
  static constexpr int Q3 = D::Q3;
  const double *B;   // [Q][P]
  const double *Dc;  // [Q][Q]
  double (*X)[Q3];
  int t, qx, qy, qz;

  __device__ Stage(const double *B_, const double *Dc_, const double *, double (*X_)[Q3], int t_) : B(B_), Dc(Dc_), X(X_), t(t_) {
    qx = t % Q;
    qy = (t / Q) % Q;
    qz = t / (Q * Q);
  }

  // interp() followed by grad(dst, gdst)
  template <int NC>
  __device__ void interp_grad(const double (&u)[NC], int s0, int s1, int dst, int gdst) {
    interp<NC>(u, s0, s1, dst);
    grad<NC>(dst, gdst);
  }

  // Interpolate NC components from node values u[c] (threads t < P^3) to this thread's quadrature point, leaving the
  // results in slots dst..dst+NC-1.  Slots s0..s0+NC-1 and s1..s1+NC-1 are scratch; dst may equal s1 but not s0.
  // Ends without a barrier: only this thread's own dst slots are written by the last pass.
  template <int NC>
  __device__ void interp(const double (&u)[NC], int s0, int s1, int dst) {
    if constexpr (P == 4 && Q == 4 && ZX) {  // all three passes across the lanes of the element's wave
      const RowRot rr(qy);
      double       bx[4], by[4], bz[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        bx[i] = B[qx * 4 + i];
        by[i] = B[qy * 4 + rr.y[i]];
        bz[i] = B[qz * 4 + i];
      }
#pragma unroll
      for (int c = 0; c < NC; c++) X[dst + c][t] = z_contract(bz, row_contract(by, quad_contract(bx, u[c])));
      return;
    } else if constexpr (P == 4 && Q == 4) {  // x and y passes across lanes, z through LDS (slot s0)
      const RowRot rr(qy);
      double       bx[4], by[4], bz[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        bx[i] = B[qx * 4 + i];
        by[i] = B[qy * 4 + rr.y[i]];
        bz[i] = B[qz * 4 + i];
      }
#pragma unroll
      for (int c = 0; c < NC; c++) X[s0 + c][t] = row_contract(by, quad_contract(bx, u[c]));
      __syncthreads();
      double r[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) acc += bz[i] * X[s0 + c][(i * 4 + qy) * 4 + qx];
        r[c] = acc;
      }
#pragma unroll
      for (int c = 0; c < NC; c++) X[dst + c][t] = r[c];
      return;
    }
    if (t < D::P3) {
#pragma unroll
      for (int c = 0; c < NC; c++) X[s0 + c][t] = u[c];
    }
    __syncthreads();
    if (t < D::PPQ) {  // x: s1[k][j][ix] = sum_a B[ix][a] u[k][j][a]
      const int ix = t % Q, kj = t / Q;
      double    b[P];
#pragma unroll
      for (int i = 0; i < P; i++) b[i] = B[ix * P + i];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < P; i++) acc += b[i] * X[s0 + c][kj * P + i];
        X[s1 + c][t] = acc;
      }
    }
    __syncthreads();
    if (t < D::PQQ) {  // y: s0[k][iy][ix] = sum_b B[iy][b] s1[k][b][ix]
      const int ix = t % Q, iy = (t / Q) % Q, k = t / (Q * Q);
      double    b[P];
#pragma unroll
      for (int i = 0; i < P; i++) b[i] = B[iy * P + i];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < P; i++) acc += b[i] * X[s1 + c][(k * P + i) * Q + ix];
        X[s0 + c][t] = acc;
      }
    }
    __syncthreads();
    {  // z: value at (qx, qy, qz)
      double b[P];
#pragma unroll
      for (int i = 0; i < P; i++) b[i] = B[qz * P + i];
      double r[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < P; i++) acc += b[i] * X[s0 + c][(i * Q + qy) * Q + qx];
        r[c] = acc;
      }
#pragma unroll
      for (int c = 0; c < NC; c++) X[dst + c][t] = r[c];
    }
  }

  // Collocated reference gradient of the NC quadrature-point fields in slots src.. (written before the last barrier
  // or by this thread): slot gdst + NC*d + c receives d/dX_d of component c.  Starts with a barrier.
  template <int NC>
  __device__ void grad(int src, int gdst) {
    __syncthreads();
    if constexpr (P == 4 && Q == 4 && ZX) {  // derivatives across the lanes of the element's wave from this lane's value
      const RowRot rr(qy);
      double       dx[4], dy[4], dz[4];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        dx[m] = Dc[qx * 4 + m];
        dy[m] = Dc[qy * 4 + rr.y[m]];
        dz[m] = Dc[qz * 4 + m];
      }
      double g[3][NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const double v = X[src + c][t];
        g[0][c]        = quad_contract(dx, v);
        g[1][c]        = row_contract(dy, v);
        g[2][c]        = z_contract(dz, v);
      }
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int c = 0; c < NC; c++) X[gdst + NC * d + c][t] = g[d][c];
      return;
    } else if constexpr (P == 4 && Q == 4) {  // x and y derivatives across lanes from this lane's value, z through LDS
      const RowRot rr(qy);
      double       dx[4], dy[4], dz[4];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        dx[m] = Dc[qx * 4 + m];
        dy[m] = Dc[qy * 4 + rr.y[m]];
        dz[m] = Dc[qz * 4 + m];
      }
      double g[3][NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const double v  = X[src + c][t];
        double       gz = 0;
#pragma unroll
        for (int m = 0; m < 4; m++) gz += dz[m] * X[src + c][(m * 4 + qy) * 4 + qx];
        g[0][c] = quad_contract(dx, v);
        g[1][c] = row_contract(dy, v);
        g[2][c] = gz;
      }
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int c = 0; c < NC; c++) X[gdst + NC * d + c][t] = g[d][c];
      return;
    }
    double dx[Q], dy[Q], dz[Q];
#pragma unroll
    for (int m = 0; m < Q; m++) {
      dx[m] = Dc[qx * Q + m];
      dy[m] = Dc[qy * Q + m];
      dz[m] = Dc[qz * Q + m];
    }
    double g[3][NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      double gx = 0, gy = 0, gz = 0;
#pragma unroll
      for (int m = 0; m < Q; m++) {
        gx += dx[m] * X[src + c][(qz * Q + qy) * Q + m];
        gy += dy[m] * X[src + c][(qz * Q + m) * Q + qx];
        gz += dz[m] * X[src + c][(m * Q + qy) * Q + qx];
      }
      g[0][c] = gx;
      g[1][c] = gy;
      g[2][c] = gz;
    }
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int c = 0; c < NC; c++) X[gdst + NC * d + c][t] = g[d][c];
  }

  // Transpose of interp (+ collocated gradient) for NC components: the quadrature-point values are in slots v0..
  // (HAS_V) and the reference-gradient test weights in slots g0 + NC*d + c (HAS_G), all written before the call.
  // Returns in r[c] (threads t < P^3) sum_q B3(q,n) v_c(q) + sum_q sum_d G3_d(q,n) g_{d,c}(q) for node n = t.
  // Uses slots v0.. and g0.. as scratch (needs g0 + 2*NC <= slot count when !HAS_G).  Starts and ends with a barrier.
  template <int NC, bool HAS_V, bool HAS_G>
  __device__ void transpose(int v0, int g0, double (&r)[NC]) {
    __syncthreads();
    if constexpr (P == 4 && Q == 4 && ZX) {  // all passes across the lanes of the element's wave
      const RowRot rr(qy);
      double       bxT[4], byT[4], bzT[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        bxT[i] = B[i * 4 + qx];
        byT[i] = B[rr.y[i] * 4 + qy];
        bzT[i] = B[i * 4 + qz];
      }
      double dxT[4], dyT[4], dzT[4];
      if (HAS_G) {
#pragma unroll
        for (int m = 0; m < 4; m++) {
          dxT[m] = Dc[m * 4 + qx];
          dyT[m] = Dc[rr.y[m] * 4 + qy];
          dzT[m] = Dc[m * 4 + qz];
        }
      }
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double w = HAS_V ? X[v0 + c][t] : 0.0;
        if (HAS_G) {
          w += z_contract(dzT, X[g0 + 2 * NC + c][t]);
          w += quad_contract(dxT, X[g0 + c][t]);
          w += row_contract(dyT, X[g0 + NC + c][t]);
        }
        r[c] = quad_contract(bxT, row_contract(byT, z_contract(bzT, w)));
      }
      __syncthreads();
      return;
    } else if constexpr (P == 4 && Q == 4) {  // z passes through LDS (slot v0 as scratch), x and y passes across lanes
      const RowRot rr(qy);
      double       w[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) w[c] = HAS_V ? X[v0 + c][t] : 0.0;
      if (HAS_G) {
        double dxT[4], dyT[4], dz[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
          dxT[m] = Dc[m * 4 + qx];
          dyT[m] = Dc[rr.y[m] * 4 + qy];
          dz[m]  = Dc[m * 4 + qz];
        }
#pragma unroll
        for (int c = 0; c < NC; c++) {
          double acc = w[c];
#pragma unroll
          for (int m = 0; m < 4; m++) acc += dz[m] * X[g0 + 2 * NC + c][(m * 4 + qy) * 4 + qx];
          acc += quad_contract(dxT, X[g0 + c][t]);
          acc += row_contract(dyT, X[g0 + NC + c][t]);
          w[c] = acc;
        }
        __syncthreads();
      }
#pragma unroll
      for (int c = 0; c < NC; c++) X[v0 + c][t] = w[c];
      __syncthreads();
      double bz[4], byT[4], bxT[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        bz[i]  = B[i * 4 + qz];
        byT[i] = B[rr.y[i] * 4 + qy];
        bxT[i] = B[i * 4 + qx];
      }
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = 0;
#pragma unroll
        for (int z = 0; z < 4; z++) acc += bz[z] * X[v0 + c][(z * 4 + qy) * 4 + qx];
        r[c] = quad_contract(bxT, row_contract(byT, acc));
      }
      __syncthreads();
      return;
    }
    double w[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) w[c] = HAS_V ? X[v0 + c][t] : 0.0;
    if (HAS_G) {
      double dx[Q], dy[Q], dz[Q];
#pragma unroll
      for (int m = 0; m < Q; m++) {
        dx[m] = Dc[m * Q + qx];
        dy[m] = Dc[m * Q + qy];
        dz[m] = Dc[m * Q + qz];
      }
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = w[c];
#pragma unroll
        for (int m = 0; m < Q; m++) {
          acc += dx[m] * X[g0 + c][(qz * Q + qy) * Q + m];
          acc += dy[m] * X[g0 + NC + c][(qz * Q + m) * Q + qx];
          acc += dz[m] * X[g0 + 2 * NC + c][(m * Q + qy) * Q + qx];
        }
        w[c] = acc;
      }
      __syncthreads();
    }
    const int sa = v0, sb = g0, sc = g0 + NC;
#pragma unroll
    for (int c = 0; c < NC; c++) X[sa + c][t] = w[c];
    __syncthreads();
    if (t < D::PQQ) {  // z^T: sb[c][iy][ix] = sum_qz B[qz][c] w[qz][iy][ix]
      const int ix = t % Q, iy = (t / Q) % Q, cz = t / (Q * Q);
      double    b[Q];
#pragma unroll
      for (int z = 0; z < Q; z++) b[z] = B[z * P + cz];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = 0;
#pragma unroll
        for (int z = 0; z < Q; z++) acc += b[z] * X[sa + c][(z * Q + iy) * Q + ix];
        X[sb + c][t] = acc;
      }
    }
    __syncthreads();
    if (t < D::PPQ) {  // y^T: sc[c][b][ix] = sum_qy B[qy][b] sb[c][qy][ix]
      const int ix = t % Q, bb = (t / Q) % P, cz = t / (Q * P);
      double    b[Q];
#pragma unroll
      for (int y = 0; y < Q; y++) b[y] = B[y * P + bb];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = 0;
#pragma unroll
        for (int y = 0; y < Q; y++) acc += b[y] * X[sb + c][(cz * Q + y) * Q + ix];
        X[sc + c][t] = acc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; c++) r[c] = 0;
    if (t < D::P3) {  // x^T: r[c][b][a] = sum_qx B[qx][a] sc[c][b][qx]
      const int a = t % P, cb = t / P;
      double    b[Q];
#pragma unroll
      for (int x = 0; x < Q; x++) b[x] = B[x * P + a];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = 0;
#pragma unroll
        for (int x = 0; x < Q; x++) acc += b[x] * X[sc + c][cb * Q + x];
        r[c] = acc;
      }
    }
    __syncthreads();
  }
};

// FP64 matrix-core passes for the 4 x 4 x 4 element.  v_mfma_f64_4x4x4_4b on one 64-lane wave: lane L = 4 b + 16 r + c
// (block b = (L >> 2) & 3); operand A holds A[i][k] at (c = i, r = k), operand B holds B[k][j] at (r = k, c = j), and the
// result D[i][j] + C lands at (r = i, c = j): the sum runs over r.  In the element layout L = qx + 4 qy + 16 qz (c = qx,
// b = qy, r = qz):
//   mf(M, u, acc) with the coefficients as A: w(qz') = sum_qz M[qz'][qz] u(qz) + acc, in the same layout (a z pass);
//   mf(u, M, acc) with the data as A: sum over qz with M[qz][j], the result's r taking the old qx and its c the new j:
//                 x and z trade places, so two such passes contract z and then x and return to the element layout.
// A 4 x 4 table T (row-major) is held per lane as cr(T) = T[c][r] or rc(T) = T[r][c].  The y passes (across the four
// blocks) stay DPP row rotations.  No pass reads LDS or needs a barrier; X only parks values across the point body.
__device__ __forceinline__ double mfma44(double a, double b, double c) { return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0); }

struct StageMF {
  static constexpr int Q3 = 64;
  const double *B;   // [4][4] interpolation B[q][n]
  const double *Dc;  // [4][4] collocated derivative
  const double *G;   // [4][4] derivative of the interpolation G[q][n]
  double (*X)[Q3];
  int t, qy;

  __device__ StageMF(const double *B_, const double *Dc_, const double *G_, double (*X_)[Q3], int t_) : B(B_), Dc(Dc_), G(G_), X(X_), t(t_) {
    qy = (t >> 2) & 3;
  }
  __device__ double cr(const double *T) const { return T[(t & 3) * 4 + (t >> 4)]; }
  __device__ double rc(const double *T) const { return T[(t >> 4) * 4 + (t & 3)]; }

  // Node values u[c] (lane = node a + 4 b + 16 c) -> quadrature-point values in slots dst..
  template <int NC>
  __device__ void interp(const double (&u)[NC], int, int, int dst) {
    const RowRot rr(qy);
    double       by[4];
#pragma unroll
    for (int i = 0; i < 4; i++) by[i] = B[qy * 4 + rr.y[i]];
    const double bcr = cr(B);
#pragma unroll
    for (int c = 0; c < NC; c++) X[dst + c][t] = mfma44(mfma44(row_contract(by, u[c]), bcr, 0.0), bcr, 0.0);
  }

  // Values (slots dst..) and reference gradients (slots gdst + NC d + c): d/dX0 from the z-contracted intermediate with
  // G (G_x B_y B_z), d/dX1 and d/dX2 collocated from the values (Dc_y, Dc_z)
  template <int NC>
  __device__ void interp_grad(const double (&u)[NC], int, int, int dst, int gdst) {
    const RowRot rr(qy);
    double       by[4], dy[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      by[i] = B[qy * 4 + rr.y[i]];
      dy[i] = Dc[qy * 4 + rr.y[i]];
    }
    const double bcr = cr(B), gcr = cr(G), dcr = cr(Dc);
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double s   = mfma44(row_contract(by, u[c]), bcr, 0.0);  // y, then z (x and z traded)
      const double val = mfma44(s, bcr, 0.0);                       // x: element layout
      X[dst + c][t]             = val;
      X[gdst + c][t]            = mfma44(s, gcr, 0.0);
      X[gdst + NC + c][t]       = row_contract(dy, val);
      X[gdst + 2 * NC + c][t]   = mfma44(dcr, val, 0.0);
    }
  }

  // Transpose: node contributions r[c] = B3^T v_c + sum_d G3_d^T g_{d,c} from slots v0.. and g0 + NC d + c, in the
  // collocated form w = v + sum_d Dc_d^T g_d, then B^T along y (DPP), z and x (matrix passes)
  template <int NC, bool HAS_V, bool HAS_G>
  __device__ void transpose(int v0, int g0, double (&r)[NC]) {
    const RowRot rr(qy);
    double       byT[4], dyT[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      byT[i] = B[rr.y[i] * 4 + qy];
      dyT[i] = Dc[rr.y[i] * 4 + qy];
    }
    const double brc = rc(B), drc = rc(Dc), eye = ((t >> 4) == (t & 3)) ? 1.0 : 0.0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      double w = HAS_V ? X[v0 + c][t] : 0.0;
      if (HAS_G) {
        w                = mfma44(drc, X[g0 + 2 * NC + c][t], w);         // + Dc_z^T g_z
        w               += row_contract(dyT, X[g0 + NC + c][t]);          // + Dc_y^T g_y
        const double gxf = mfma44(X[g0 + c][t], eye, 0.0);                // g_x with x and z traded
        w                = mfma44(mfma44(drc, gxf, 0.0), eye, w);         // + Dc_x^T g_x, traded back
      }
      r[c] = mfma44(mfma44(row_contract(byT, w), brc, 0.0), brc, 0.0);  // B^T along y, then z and x
    }
  }
};

}  // namespace basis3
