// Vector and restriction kernels of the MI355X backend (CeedVector arithmetic, norms, multiplicity, restriction apply).
// All are bandwidth-bound streaming kernels with grid-stride loops.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace {

constexpr int kBlock = 256;

inline int grid_for(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

__global__ void k_fill(double *x, double v, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) x[i] = v;
}
__global__ void k_axpby(double *y, double a, double b, const double *x, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) y[i] = a * x[i] + b * y[i];
}
__global__ void k_scale(double *x, double a, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) x[i] *= a;
}
__global__ void k_pmult(double *w, const double *x, const double *y, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) w[i] = x[i] * y[i];
}
__global__ void k_recip(double *x, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    if (fabs(x[i]) > 1e-300) x[i] = 1.0 / x[i];
}

// Per-block partial: 0 -> sum|x|, 1 -> sum x^2, 2 -> max|x|; npartial blocks, results in partial[block]
__global__ void k_norm(const double *x, long long n, int type, double *partial) {
  __shared__ double red[kBlock];
  double            acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double v = x[i];
    if (type == 0) acc += fabs(v);
    else if (type == 1) acc += v * v;
    else acc = fmax(acc, fabs(v));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = (type == 2) ? fmax(red[threadIdx.x], red[threadIdx.x + s]) : red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_multiplicity(const int *off, long long n, int ncomp, int cs, double *mult) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    for (int c = 0; c < ncomp; c++) atomicAdd(&mult[off[i] + c * cs], 1.0);
}

// E-vector layout [elem][comp][node] (libCEED's offset-restriction E-vector ordering)
__global__ void k_rstr_n(const int *off, int nelem, int esize, int ncomp, int cs, const double *u, double *e) {
  const long long n = (long long)nelem * esize;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long el = i / esize, node = i % esize;
    for (int c = 0; c < ncomp; c++) e[(el * ncomp + c) * esize + node] = u[off[i] + c * cs];
  }
}
__global__ void k_rstr_t(const int *off, int nelem, int esize, int ncomp, int cs, const double *e, double *u) {
  const long long n = (long long)nelem * esize;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long el = i / esize, node = i % esize;
    for (int c = 0; c < ncomp; c++) atomicAdd(&u[off[i] + c * cs], e[(el * ncomp + c) * esize + node]);
  }
}

// Deterministic second scatter pass: one thread per (L-vector node, component) item.  Row k's contributions are the
// E-vector slots t_rowptr[k] .. t_rowptr[k+1]-1 (node-sorted E-vector) or the slots index[] of that range
// (element-ordered E-vector); they are summed in ascending order.  Each thread handles NPT items a grid stride apart
// and issues all their row-pointer loads, then all their contribution loads (up to 8 per row, the most a hexahedral
// mesh node has), so a thread waits on two memory round trips for NPT items.
template <bool INDEXED, int NPT>
__global__ void __launch_bounds__(256) k_node_gather(int nrows, int ncomp, int cs, const int *__restrict__ rowptr, const int *__restrict__ index,
                                                     const int *__restrict__ toff, const double *__restrict__ ev, double *__restrict__ v,
                                                     bool overwrite) {
  const long long n      = (long long)nrows * ncomp;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += stride * NPT) {
    int row[NPT], c[NPT], b[NPT], m[NPT];
#pragma unroll
    for (int j = 0; j < NPT; j++) {
      const long long i = i0 + j * stride;
      row[j]            = i < n ? (int)(i / ncomp) : 0;
      c[j]              = i < n ? (int)(i % ncomp) : 0;
      b[j]              = rowptr[row[j]];
      m[j]              = i < n ? rowptr[row[j] + 1] - b[j] : 0;
    }
    double acc[NPT];
#pragma unroll
    for (int j = 0; j < NPT; j++) acc[j] = 0;
    bool more = false;
#pragma unroll
    for (int j = 0; j < NPT; j++) more |= m[j] > 8;
    if (!more) {  // every row has at most 8 contributions: one batch of loads
      double x[NPT][8];
#pragma unroll
      for (int j = 0; j < NPT; j++)
#pragma unroll
        for (int k = 0; k < 8; k++) {
          x[j][k] = 0;
          if (k < m[j]) x[j][k] = ev[(INDEXED ? (long long)index[b[j] + k] : (long long)(b[j] + k)) * ncomp + c[j]];
        }
#pragma unroll
      for (int j = 0; j < NPT; j++)
#pragma unroll
        for (int k = 0; k < 8; k++)
          if (k < m[j]) acc[j] += x[j][k];
    } else {
#pragma unroll
      for (int j = 0; j < NPT; j++)
        for (int k = 0; k < m[j]; k++) acc[j] += ev[(INDEXED ? (long long)index[b[j] + k] : (long long)(b[j] + k)) * ncomp + c[j]];
    }
#pragma unroll
    for (int j = 0; j < NPT; j++) {
      if (i0 + j * stride < n) {
        double *dst = &v[toff[row[j]] + (long long)c[j] * cs];
        *dst        = overwrite ? acc[j] : *dst + acc[j];
      }
    }
  }
}

// The same sum for the element-ordered E-vector, one thread per L-vector node: each contribution is the node's NC
// consecutive values in its element's block (40 contiguous bytes for 5 components), read whole; the sums run over the
// slots in ascending order, as in k_node_gather (identical results).  Measured at 32^3 p=3: 50 us against 49-57 us for
// one thread per (node, component) in node or in element order.
// The rows are split into 8 contiguous ranges, one per XCD (block b runs on XCD b % 8; the grid is a multiple of 8):
// the nodes of one range belong to a contiguous range of elements, whose E-vector lines are then fetched into one XCD's
// L2 instead of into several.
// E is the E-vector's type: double, or float for the FP32 Jacobian's contributions (summed in FP64 here).
template <int NC, typename E>
__global__ void __launch_bounds__(512) k_node_gather_rows(int nrows, int cs, const int *__restrict__ rowptr, const int *__restrict__ index,
                                                          const int *__restrict__ toff, const E *__restrict__ ev, double *__restrict__ v,
                                                          bool overwrite) {
  const long long per = ((long long)nrows + 7) / 8;
  const long long x = blockIdx.x % 8, j = blockIdx.x / 8, nb = gridDim.x / 8;
  const long long lo = x * per, hi = (x + 1) * per < nrows ? (x + 1) * per : nrows;
  for (long long row = lo + j * blockDim.x + threadIdx.x; row < hi; row += nb * blockDim.x) {
    const int b = rowptr[row], m = rowptr[row + 1] - b;
    double    acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0;
    if (m <= 8) {  // at most 8 contributions (a hexahedral mesh node): all slot indices, then all values
      int slot[8];
#pragma unroll
      for (int k = 0; k < 8; k++) slot[k] = k < m ? index[b + k] : 0;
      double x[8][NC];
#pragma unroll
      for (int k = 0; k < 8; k++)
#pragma unroll
        for (int c = 0; c < NC; c++) x[k][c] = k < m ? (double)ev[(long long)slot[k] * NC + c] : 0.0;
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k < m)
#pragma unroll
          for (int c = 0; c < NC; c++) acc[c] += x[k][c];
    } else {
      for (int k = 0; k < m; k++) {
        const long long sl = index[b + k];
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c] += (double)ev[sl * NC + c];
      }
    }
    const long long o = toff[row];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      double *dst = &v[o + (long long)c * cs];
      *dst        = overwrite ? acc[c] : *dst + acc[c];
    }
  }
}

// The point-block form of the element-ordered gather: the E-vector holds a 5 x 5 block [c_out][c_in] per element node
// (25 consecutive values), the output a block per node at 25 * (offset / 5) (the L-vector offsets of an interleaved
// 5-component field are 5 * node).  One thread per (node, c_out) row of the block sums its 5 values over the node's
// contributions in ascending slot order and adds them to the output.
__global__ void __launch_bounds__(256) k_node_gather_blocks(int nrows, const int *__restrict__ rowptr, const int *__restrict__ index,
                                                            const int *__restrict__ toff, const double *__restrict__ ev, double *__restrict__ v) {
  const long long n = (long long)nrows * 5;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int row = (int)(i / 5), co = (int)(i % 5);
    const int b = rowptr[row], m = rowptr[row + 1] - b;
    double    acc[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < m; k++) {
      const double *src = ev + (long long)index[b + k] * 25 + co * 5;
#pragma unroll
      for (int c = 0; c < 5; c++) acc[c] += src[c];
    }
    double *dst = v + (long long)toff[row] * 5 + co * 5;
#pragma unroll
    for (int c = 0; c < 5; c++) dst[c] += acc[c];
  }
}

// Per-block partial sums of x . y over a fixed grid (deterministic for a given n and npartial)
__global__ void k_dot(const double *x, const double *y, long long n, double *partial) {
  __shared__ double red[kBlock];
  double            acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) acc += x[i] * y[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

}  // namespace

hipError_t launch_node_gather(int nrows, int ncomp, int comp_stride, const int *t_rowptr, const int *t_index, const int *t_offset,
                              const double *ev, double *v, bool overwrite, hipStream_t s) {
  constexpr int   NPT = 2;  // measured: 2 items per thread 31.8 us against 39.1 us (1) and 39.0 us (4) at 884,736 rows
  const long long n   = (long long)nrows * ncomp;
  if (n <= 0) return hipSuccess;
  if (t_index && ncomp == 5) {  // element-ordered E-vector: one thread per node, 8 row ranges (one per XCD)
    // 512-thread blocks: 49.4 us against 50.6 us (256) and 50.3 us (128) at 884,736 rows, two A/B pairs
    constexpr int kRowsBlock = 512;
    const int     grid       = (int)(((nrows + kRowsBlock - 1) / kRowsBlock + 7) / 8 * 8);
    hipLaunchKernelGGL((k_node_gather_rows<5, double>), dim3(grid), dim3(kRowsBlock), 0, s, nrows, comp_stride, t_rowptr, t_index, t_offset, ev, v,
                       overwrite);
    return hipGetLastError();
  }
  const int grid = grid_for((n + NPT - 1) / NPT);
  if (t_index)
    hipLaunchKernelGGL((k_node_gather<true, NPT>), dim3(grid), dim3(kBlock), 0, s, nrows, ncomp, comp_stride, t_rowptr, t_index, t_offset, ev, v,
                       overwrite);
  else
    hipLaunchKernelGGL((k_node_gather<false, NPT>), dim3(grid), dim3(kBlock), 0, s, nrows, ncomp, comp_stride, t_rowptr, t_index, t_offset, ev,
                       v, overwrite);
  return hipGetLastError();
}

hipError_t launch_node_gather_rows_f32(int nrows, int comp_stride, const int *t_rowptr, const int *t_index, const int *t_offset, const float *ev,
                                      double *v, bool overwrite, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  const int grid = (grid_for(nrows) + 7) / 8 * 8;
  hipLaunchKernelGGL((k_node_gather_rows<5, float>), dim3(grid), dim3(kBlock), 0, s, nrows, comp_stride, t_rowptr, t_index, t_offset, ev, v, overwrite);
  return hipGetLastError();
}

hipError_t launch_node_gather_blocks(int nrows, const int *t_rowptr, const int *t_index, const int *t_offset, const double *ev, double *v,
                                     hipStream_t s) {
  const long long n = (long long)nrows * 5;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_node_gather_blocks, dim3(grid_for(n)), dim3(kBlock), 0, s, nrows, t_rowptr, t_index, t_offset, ev, v);
  return hipGetLastError();
}

hipError_t launch_fill(double *x, double value, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, x, value, n);
  return hipGetLastError();
}
hipError_t launch_axpby(double *y, double alpha, double beta, const double *x, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kBlock), 0, s, y, alpha, beta, x, n);
  return hipGetLastError();
}
hipError_t launch_scale(double *x, double alpha, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kBlock), 0, s, x, alpha, n);
  return hipGetLastError();
}
hipError_t launch_pointwise_mult(double *w, const double *x, const double *y, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pmult, dim3(grid_for(n)), dim3(kBlock), 0, s, w, x, y, n);
  return hipGetLastError();
}
hipError_t launch_reciprocal(double *x, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_recip, dim3(grid_for(n)), dim3(kBlock), 0, s, x, n);
  return hipGetLastError();
}
hipError_t launch_dot(const double *x, const double *y, long long n, double *d_partial, int npartial, hipStream_t s) {
  hipLaunchKernelGGL(k_dot, dim3(npartial), dim3(kBlock), 0, s, x, y, n, d_partial);
  return hipGetLastError();
}
hipError_t launch_norm(const double *x, long long n, int type, double *d_partial, int npartial, hipStream_t s) {
  hipLaunchKernelGGL(k_norm, dim3(npartial), dim3(kBlock), 0, s, x, n, type, d_partial);
  return hipGetLastError();
}
hipError_t launch_multiplicity(const int *offsets, long long n_offsets, int ncomp, int comp_stride, double *mult, hipStream_t s) {
  if (n_offsets <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_multiplicity, dim3(grid_for(n_offsets)), dim3(kBlock), 0, s, offsets, n_offsets, ncomp, comp_stride, mult);
  return hipGetLastError();
}
hipError_t launch_restriction_notranspose(const int *offsets, int nelem, int elemsize, int ncomp, int comp_stride, const double *u, double *e,
                                          hipStream_t s) {
  const long long n = (long long)nelem * elemsize;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rstr_n, dim3(grid_for(n)), dim3(kBlock), 0, s, offsets, nelem, elemsize, ncomp, comp_stride, u, e);
  return hipGetLastError();
}
hipError_t launch_restriction_transpose(const int *offsets, int nelem, int elemsize, int ncomp, int comp_stride, const double *e, double *u,
                                        hipStream_t s) {
  const long long n = (long long)nelem * elemsize;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rstr_t, dim3(grid_for(n)), dim3(kBlock), 0, s, offsets, nelem, elemsize, ncomp, comp_stride, e, u);
  return hipGetLastError();
}
