// Vector kernels of the time integrator and its Krylov solvers (GMRES with point-Jacobi, Newton, explicit RK).
//
// Reductions are deterministic: a grid of fixed size (a function of the length only) reduces contiguous grid-stride
// slices per block in a fixed tree order into partials[slot][block] (slot-major, so that the second pass reads each
// slot's partials contiguously), and a second kernel sums each slot over the blocks in a fixed order.  The same inputs
// give bitwise identical results from run to run.
//
// Multi-dot: the dots of w with up to kMaxDots basis vectors are taken in groups of 8 (w is read once per group).
// Projection update: w -= sum_j h_j V_j with the h_j read from device memory (the reduced dots), fused with the
// partial sums of ||w||^2, so that one Arnoldi step needs a single host read of (h_0..h_k, ||w||^2), which the final
// reduction kernel stores into mapped host memory (Publish).
#include <hip/hip_runtime.h>

#include "krylov.hpp"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double block_sum(double v, double *s) {
  s[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  const double r = s[0];
  __syncthreads();
  return r;
}

// partial[j0 + j][block] = sum over this block's slice of w[i] V[(j0 + j) ldv + i], j < nj <= 8
__global__ void __launch_bounds__(kThreads) k_mdot8(const double *__restrict__ w, const double *__restrict__ V, long long ldv, int j0, int nj,
                                                    long long n, double *__restrict__ partial) {
  __shared__ double s[kThreads];
  double            acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const long long   stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const double wi = w[i];
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (j < nj) acc[j] += wi * V[(long long)(j0 + j) * ldv + i];
  }
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (j >= nj) break;
    const double r = block_sum(acc[j], s);
    if (threadIdx.x == 0) partial[(long long)(j0 + j) * kMaxReductionBlocks + blockIdx.x] = r;
  }
}

// partial[slot][block] = sum over this block's slice of x[i] y[i]
__global__ void __launch_bounds__(kThreads) k_dot(const double *__restrict__ x, const double *__restrict__ y, long long n, int slot,
                                                  double *__restrict__ partial) {
  __shared__ double s[kThreads];
  double            acc    = 0;
  const long long   stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) acc += x[i] * y[i];
  const double r = block_sum(acc, s);
  if (threadIdx.x == 0) partial[(long long)slot * kMaxReductionBlocks + blockIdx.x] = r;
}

// w[i] -= sum_{j < k} h[j] V[j ldv + i] (in order j = 0, 1, ...); partial[slot][block] = sum of the new w[i]^2
__global__ void __launch_bounds__(kThreads) k_project_out(double *__restrict__ w, const double *__restrict__ V, long long ldv, int k,
                                                          const double *__restrict__ h, long long n, int slot, double *__restrict__ partial) {
  __shared__ double s[kThreads];
  __shared__ double c[kMaxDots];
  if ((int)threadIdx.x < k) c[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  double          ss     = 0;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    double acc = w[i];
    for (int j = 0; j < k; j++) acc -= c[j] * V[(long long)j * ldv + i];
    w[i] = acc;
    ss += acc * acc;
  }
  const double r = block_sum(ss, s);
  if (threadIdx.x == 0) partial[(long long)slot * kMaxReductionBlocks + blockIdx.x] = r;
}

// out[slot] = sum over blocks b (in order) of partial[slot][b], for slots [s0, s0 + gridDim.x).  A single-block launch
// with pub.h set then stores out[0 .. pub.count) into mapped host memory and pub.seq into *pub.flag (Publish).
__global__ void __launch_bounds__(kThreads) k_reduce_partials(const double *__restrict__ partial, int nblocks, int s0, double *__restrict__ out,
                                                              Publish pub) {
  __shared__ double s[kThreads];
  const int         slot = s0 + blockIdx.x;
  double            acc  = 0;
  for (int b = threadIdx.x; b < nblocks; b += kThreads) acc += partial[(long long)slot * kMaxReductionBlocks + b];
  const double r = block_sum(acc, s);
  if (threadIdx.x == 0) out[slot] = r;
  if (pub.h) {
    const int i = threadIdx.x;
    if (i < pub.count) pub.h[i] = i == slot ? r : out[i];  // the other slots come from earlier launches
    __threadfence_system();  // each thread's value reaches host memory before the sequence number
    __syncthreads();
    if (i == 0) __hip_atomic_store(pub.flag, pub.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// y = y / sqrt(*hsq) unless *hsq is zero (a Krylov vector normalized by its squared norm in device memory)
__global__ void __launch_bounds__(kThreads) k_normalize(double *__restrict__ y, const double *__restrict__ hsq, long long n) {
  const double h = *hsq;
  if (h == 0) return;
  const double    a      = 1.0 / sqrt(h);
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) y[i] = a * y[i];
}

// y[i] = a y[i] + sum_{j < m} c_j X_j[i]
__global__ void __launch_bounds__(kThreads) k_lincomb(double *__restrict__ y, double a, LinComb lc, long long n) {
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    double acc = a == 0 ? 0.0 : a * y[i];
    for (int j = 0; j < lc.m; j++) acc += lc.c[j] * lc.x[j][i];
    y[i] = acc;
  }
}

// y[i] = a y[i] + sum_{j < k} c_j V[j ldv + i] with basis-vector coefficients
__global__ void __launch_bounds__(kThreads) k_basis_comb(double *__restrict__ y, const double *__restrict__ V, long long ldv, Coefs cf, long long n) {
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    double acc = y[i];
    for (int j = 0; j < cf.k; j++) acc += cf.c[j] * V[(long long)j * ldv + i];
    y[i] = acc;
  }
}

// z[i] = d[i] (x[i] - y[i]) (y may be null: z = d x)
__global__ void __launch_bounds__(kThreads) k_dscale_diff(double *__restrict__ z, const double *__restrict__ d, const double *__restrict__ x,
                                                          const double *__restrict__ y, long long n) {
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) z[i] = d[i] * (y ? x[i] - y[i] : x[i]);
}

// d[i] = 1 / d[i], with 1 where d[i] == 0 (PCJACOBI's zero-diagonal rule)
__global__ void __launch_bounds__(kThreads) k_jacobi_invert(double *__restrict__ d, long long n) {
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) d[i] = d[i] != 0 ? 1.0 / d[i] : 1.0;
}

__global__ void __launch_bounds__(kThreads) k_set(double *__restrict__ x, double v, long long n) {
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) x[i] = v;
}

// partial[slot][block] = number of non-finite entries in the block's slice (as a double)
__global__ void __launch_bounds__(kThreads) k_count_nonfinite(const double *__restrict__ x, long long n, int slot, double *__restrict__ partial) {
  __shared__ double s[kThreads];
  double            cnt    = 0;
  const long long   stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) cnt += isfinite(x[i]) ? 0.0 : 1.0;
  const double r = block_sum(cnt, s);
  if (threadIdx.x == 0) partial[(long long)slot * kMaxReductionBlocks + blockIdx.x] = r;
}

// In-place inverse of each node's 5x5 block (row-major [c_out][c_in]) by Gauss-Jordan elimination with partial
// pivoting on [A | I]; every index is a compile-time constant so the blocks stay in registers.
// partial[slot][block] counts the singular blocks of the block's slice.
__global__ void __launch_bounds__(kThreads) k_pb_invert(double *__restrict__ pb, long long nn, int slot, double *__restrict__ partial) {
  __shared__ double s[kThreads];
  double            bad    = 0;
  const long long   stride = (long long)gridDim.x * kThreads;
  for (long long node = (long long)blockIdx.x * kThreads + threadIdx.x; node < nn; node += stride) {
    double A[5][5], X[5][5];
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
      for (int j = 0; j < 5; j++) {
        A[i][j] = pb[node * 25 + i * 5 + j];
        X[i][j] = i == j ? 1.0 : 0.0;
      }
    bool singular = false;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      int    piv  = k;
      double best = fabs(A[k][k]);
#pragma unroll
      for (int i = k + 1; i < 5; i++)
        if (fabs(A[i][k]) > best) {
          best = fabs(A[i][k]);
          piv  = i;
        }
#pragma unroll
      for (int i = k + 1; i < 5; i++)
        if (i == piv)
#pragma unroll
          for (int j = 0; j < 5; j++) {
            double ta = A[k][j], tx = X[k][j];
            A[k][j] = A[i][j], X[k][j] = X[i][j];
            A[i][j] = ta, X[i][j] = tx;
          }
      if (A[k][k] == 0) singular = true;
      const double inv = A[k][k] != 0 ? 1.0 / A[k][k] : 0.0;
#pragma unroll
      for (int j = 0; j < 5; j++) {
        A[k][j] *= inv;
        X[k][j] *= inv;
      }
#pragma unroll
      for (int i = 0; i < 5; i++) {
        if (i == k) continue;
        const double f = A[i][k];
#pragma unroll
        for (int j = 0; j < 5; j++) {
          A[i][j] -= f * A[k][j];
          X[i][j] -= f * X[k][j];
        }
      }
    }
    if (singular) bad += 1.0;
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
      for (int j = 0; j < 5; j++) pb[node * 25 + i * 5 + j] = X[i][j];
  }
  const double r = block_sum(bad, s);
  if (threadIdx.x == 0) partial[(long long)slot * kMaxReductionBlocks + blockIdx.x] = r;
}

// z_n = Binv_n (x_n - y_n) per node n (y may be null)
__global__ void __launch_bounds__(kThreads) k_pb_apply(double *__restrict__ z, const double *__restrict__ pbinv, const double *__restrict__ x,
                                                       const double *__restrict__ y, long long nn) {
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long node = (long long)blockIdx.x * kThreads + threadIdx.x; node < nn; node += stride) {
    double r[5];
#pragma unroll
    for (int j = 0; j < 5; j++) r[j] = y ? x[node * 5 + j] - y[node * 5 + j] : x[node * 5 + j];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      double acc = 0;
#pragma unroll
      for (int j = 0; j < 5; j++) acc += pbinv[node * 25 + i * 5 + j] * r[j];
      z[node * 5 + i] = acc;
    }
  }
}

int elementwise_grid(long long n) {
  long long g = (n + kThreads - 1) / kThreads;
  if (g > 4096) g = 4096;
  return g < 1 ? 1 : (int)g;
}

}  // namespace

int krylov_reduction_blocks(long long n) {
  long long g = (n + kThreads - 1) / kThreads;
  if (g > kMaxReductionBlocks) g = kMaxReductionBlocks;
  return g < 1 ? 1 : (int)g;
}

hipError_t launch_mdot(const double *w, const double *V, long long ldv, int k, long long n, double *partial, double *out, hipStream_t s) {
  if (k < 1 || k > kMaxDots) return hipErrorInvalidValue;
  const int nb = krylov_reduction_blocks(n);
  for (int j0 = 0; j0 < k; j0 += 8) {
    const int nj = k - j0 < 8 ? k - j0 : 8;
    hipLaunchKernelGGL(k_mdot8, dim3(nb), dim3(kThreads), 0, s, w, V, ldv, j0, nj, n, partial);
  }
  hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(kThreads), 0, s, partial, nb, 0, out, Publish{});
  return hipGetLastError();
}

hipError_t launch_dot(const double *x, const double *y, long long n, int slot, double *partial, double *out, hipStream_t s, const Publish &pub) {
  if (slot < 0 || slot >= kMaxDots || pub.count > kThreads) return hipErrorInvalidValue;
  const int nb = krylov_reduction_blocks(n);
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kThreads), 0, s, x, y, n, slot, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kThreads), 0, s, partial, nb, slot, out, pub);
  return hipGetLastError();
}

hipError_t launch_project_out(double *w, const double *V, long long ldv, int k, const double *h, long long n, int slot, double *partial, double *out,
                              hipStream_t s, const Publish &pub) {
  if (k < 0 || k > kMaxDots || slot < 0 || slot >= kMaxDots || pub.count > kThreads) return hipErrorInvalidValue;
  const int nb = krylov_reduction_blocks(n);
  hipLaunchKernelGGL(k_project_out, dim3(nb), dim3(kThreads), 0, s, w, V, ldv, k, h, n, slot, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kThreads), 0, s, partial, nb, slot, out, pub);
  return hipGetLastError();
}

hipError_t launch_normalize(double *y, const double *hsq, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_normalize, dim3(elementwise_grid(n)), dim3(kThreads), 0, s, y, hsq, n);
  return hipGetLastError();
}

hipError_t launch_lincomb(double *y, double a, const LinComb &lc, long long n, hipStream_t s) {
  if (lc.m < 0 || lc.m > kMaxLinComb) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_lincomb, dim3(elementwise_grid(n)), dim3(kThreads), 0, s, y, a, lc, n);
  return hipGetLastError();
}

hipError_t launch_basis_comb(double *y, const double *V, long long ldv, const Coefs &cf, long long n, hipStream_t s) {
  if (cf.k < 0 || cf.k > kMaxDots) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_basis_comb, dim3(elementwise_grid(n)), dim3(kThreads), 0, s, y, V, ldv, cf, n);
  return hipGetLastError();
}

hipError_t launch_dscale_diff(double *z, const double *d, const double *x, const double *y, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_dscale_diff, dim3(elementwise_grid(n)), dim3(kThreads), 0, s, z, d, x, y, n);
  return hipGetLastError();
}

hipError_t launch_jacobi_invert(double *d, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_jacobi_invert, dim3(elementwise_grid(n)), dim3(kThreads), 0, s, d, n);
  return hipGetLastError();
}

hipError_t launch_set(double *x, double v, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_set, dim3(elementwise_grid(n)), dim3(kThreads), 0, s, x, v, n);
  return hipGetLastError();
}

hipError_t launch_count_nonfinite(const double *x, long long n, int slot, double *partial, double *out, hipStream_t s, const Publish &pub) {
  if (slot < 0 || slot >= kMaxDots || pub.count > kThreads) return hipErrorInvalidValue;
  const int nb = krylov_reduction_blocks(n);
  hipLaunchKernelGGL(k_count_nonfinite, dim3(nb), dim3(kThreads), 0, s, x, n, slot, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kThreads), 0, s, partial, nb, slot, out, pub);
  return hipGetLastError();
}

hipError_t launch_pb_invert(double *pb, long long nnodes, int slot, double *partial, double *out, hipStream_t s, const Publish &pub) {
  if (slot < 0 || slot >= kMaxDots || pub.count > kThreads) return hipErrorInvalidValue;
  const int nb = krylov_reduction_blocks(nnodes);
  hipLaunchKernelGGL(k_pb_invert, dim3(nb), dim3(kThreads), 0, s, pb, nnodes, slot, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kThreads), 0, s, partial, nb, slot, out, pub);
  return hipGetLastError();
}

hipError_t launch_pb_apply(double *z, const double *pbinv, const double *x, const double *y, long long nnodes, hipStream_t s) {
  hipLaunchKernelGGL(k_pb_apply, dim3(elementwise_grid(nnodes)), dim3(kThreads), 0, s, z, pbinv, x, y, nnodes);
  return hipGetLastError();
}
