// Device vector operations of the time integrator (kernels_krylov.hip).  All lengths are in doubles; every launch is
// asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>

constexpr int kMaxDots            = 32;    // GMRES restart + 1 must not exceed this
constexpr int kMaxReductionBlocks = 1024;  // partial buffers hold kMaxDots x kMaxReductionBlocks doubles, slot-major
constexpr int kMaxLinComb         = 8;

struct LinComb {  // sum_j c[j] x[j], passed by value
  int           m;
  double        c[kMaxLinComb];
  const double *x[kMaxLinComb];
};

// Host read of reduction results without a copy: the final (single-block) reduction kernel of a launch stores
// out[0 .. count) into h and then seq into *flag with system-scope release ordering.  h and flag are device addresses
// of mapped, coherent host memory; h == nullptr publishes nothing.
struct Publish {
  double             *h     = nullptr;
  unsigned long long *flag  = nullptr;
  unsigned long long  seq   = 0;
  int                 count = 0;
};

struct Coefs {  // basis coefficients, passed by value
  int    k;
  double c[kMaxDots];
};

int krylov_reduction_blocks(long long n);
// out[j] = <w, V_j>, j < k (V_j = V + j ldv)
hipError_t launch_mdot(const double *w, const double *V, long long ldv, int k, long long n, double *partial, double *out, hipStream_t s);
// out[slot] = <x, y>
hipError_t launch_dot(const double *x, const double *y, long long n, int slot, double *partial, double *out, hipStream_t s,
                      const Publish &pub = Publish{});
// w -= sum_{j<k} h[j] V_j (h in device memory), then out[slot] = <w, w>
hipError_t launch_project_out(double *w, const double *V, long long ldv, int k, const double *h, long long n, int slot, double *partial, double *out,
                              hipStream_t s, const Publish &pub = Publish{});
// y /= sqrt(*hsq) unless *hsq == 0 (hsq in device memory)
hipError_t launch_normalize(double *y, const double *hsq, long long n, hipStream_t s);
// y = a y + sum_j lc.c[j] lc.x[j]  (a == 0 does not read y)
hipError_t launch_lincomb(double *y, double a, const LinComb &lc, long long n, hipStream_t s);
// y += sum_{j<cf.k} cf.c[j] V_j
hipError_t launch_basis_comb(double *y, const double *V, long long ldv, const Coefs &cf, long long n, hipStream_t s);
// z = d (x - y), or z = d x when y is null
hipError_t launch_dscale_diff(double *z, const double *d, const double *x, const double *y, long long n, hipStream_t s);
// d = 1/d with 1 for zero entries
hipError_t launch_jacobi_invert(double *d, long long n, hipStream_t s);
hipError_t launch_set(double *x, double v, long long n, hipStream_t s);
// out[slot] = number of NaN/Inf entries of x
hipError_t launch_count_nonfinite(const double *x, long long n, int slot, double *partial, double *out, hipStream_t s,
                                  const Publish &pub = Publish{});
// In-place inverse of each node's 5x5 block [c_out][c_in]; out[slot] = number of singular blocks
hipError_t launch_pb_invert(double *pb, long long nnodes, int slot, double *partial, double *out, hipStream_t s,
                            const Publish &pub = Publish{});
// z_n = Binv_n (x_n - y_n) for each node n (y may be null; 5 interleaved components per node)
hipError_t launch_pb_apply(double *z, const double *pbinv, const double *x, const double *y, long long nnodes, hipStream_t s);
