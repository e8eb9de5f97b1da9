// Launch interfaces of the SGS model training kernels (kernels_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "honee_mi355x.h"

// Layer sizes (inputs, hidden neurons, outputs), the leaky-ReLU slope and the neurons of a second hidden layer (0: one
// hidden layer).  Parameter layout (floats): W1 [nh][ni], b1 [nh], then with one hidden layer W2 [no][nh], b2 [no];
// with two, W2 [nh2][nh], b2 [nh2], W3 [no][nh2], b3 [no].
struct MlpDims {
  int   ni, nh, no;
  float alpha;
  int   nh2;
};

__host__ __device__ constexpr int mlp_nparams(const MlpDims &d) {
  return d.nh * d.ni + d.nh + (d.nh2 > 0 ? d.nh2 * d.nh + d.nh2 + d.no * d.nh2 : d.no * d.nh) + d.no;
}

// Adam step constants; bc1 = 1 - beta1^t and bc2 = 1 - beta2^t are the bias corrections of step t
struct MlpAdam {
  float lr, beta1, beta2, eps, bc1, bc2;
};

bool       mlp_dims_supported(const MlpDims &d);
// Samples per block of launch_mlp_grad (the number of per-block partials is ceil(count / this))
int        mlp_grad_block(const MlpDims &d);
// Batch gradient of the mean squared error over samples idx[first .. first+count) (idx null: first .. first+count);
// inv_norm = 1 / (count * no).  Writes nblocks per-block partials [nblocks][nparams] and loss partials.
hipError_t launch_mlp_grad(const MlpDims &d, const float *theta, const float *X, const float *T, const int *idx, int first, int count, float inv_norm,
                           float *partial, float *loss_partial, int *nblocks, hipStream_t s);
hipError_t launch_mlp_reduce_adam(int np, int nblocks, const float *partial, const float *loss_partial, float *grad, float *loss, float *theta, float *m,
                                  float *v, const MlpAdam &adam, bool apply, hipStream_t s);
hipError_t launch_mlp_predict(const MlpDims &d, const float *theta, const float *X, int n, float *Y, hipStream_t s);
hipError_t launch_sgs_model_samples(long long n, const double *train, const double *filtered, const double *G, const double *A, const CeedMiGas &gas,
                                    float *X, double *S, hipStream_t s);
