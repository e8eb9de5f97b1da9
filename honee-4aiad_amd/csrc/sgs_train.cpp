// Host side of the SGS model training (FP32 MLP, Adam) on MI355X: parameters, Adam moments and scratch live in
// device memory; every batch runs k_mlp_grad then k_mlp_reduce_adam on the object's stream.
#include "sgs_train.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "kernels_mlp.hpp"

namespace {

thread_local std::string g_err;

int fail(const char *msg, hipError_t e = hipSuccess) {
  g_err = msg;
  if (e != hipSuccess) g_err += std::string(": ") + hipGetErrorString(e);
  return 1;
}

#define HIPCHK(expr)                              \
  do {                                            \
    hipError_t e_ = (expr);                       \
    if (e_ != hipSuccess) return fail(#expr, e_); \
  } while (0)

// splitmix64: a deterministic stream of 64-bit values
struct SplitMix {
  unsigned long long s;
  unsigned long long next() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z                    = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z                    = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

struct HoneeMiMlp_ {
  MlpDims     d;
  int         np;
  hipStream_t s = nullptr;
  float      *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *loss = nullptr;
  float      *partial = nullptr, *loss_partial = nullptr;
  int         partial_blocks = 0;
  long long   step           = 0;
  // staging of host inputs
  float      *X = nullptr, *T = nullptr;
  long long   staged = 0;
  int        *idx    = nullptr;
  long long   idx_len = 0;
};

static int ensure_partials(HoneeMiMlp p, int nblocks) {
  if (nblocks <= p->partial_blocks) return 0;
  if (p->partial) HIPCHK(hipFree(p->partial));
  if (p->loss_partial) HIPCHK(hipFree(p->loss_partial));
  HIPCHK(hipMalloc((void **)&p->partial, sizeof(float) * (size_t)nblocks * p->np));
  HIPCHK(hipMalloc((void **)&p->loss_partial, sizeof(float) * nblocks));
  p->partial_blocks = nblocks;
  return 0;
}

// Device views of X and T (copied once per call when they are host arrays)
static int stage(HoneeMiMlp p, int n, const float *X, const float *T, int mem_type, const float **dX, const float **dT) {
  if (mem_type == 1) {
    *dX = X;
    *dT = T;
    return 0;
  }
  if (n > p->staged) {
    if (p->X) HIPCHK(hipFree(p->X));
    if (p->T) HIPCHK(hipFree(p->T));
    HIPCHK(hipMalloc((void **)&p->X, sizeof(float) * (size_t)n * p->d.ni));
    HIPCHK(hipMalloc((void **)&p->T, sizeof(float) * (size_t)n * p->d.no));
    p->staged = n;
  }
  HIPCHK(hipMemcpyAsync(p->X, X, sizeof(float) * (size_t)n * p->d.ni, hipMemcpyHostToDevice, p->s));
  if (T) HIPCHK(hipMemcpyAsync(p->T, T, sizeof(float) * (size_t)n * p->d.no, hipMemcpyHostToDevice, p->s));
  *dX = p->X;
  *dT = p->T;
  return 0;
}

// Gradient (and optionally an Adam step) on samples idx[first .. first+count) of dX/dT
static int batch(HoneeMiMlp p, const float *dX, const float *dT, const int *idx, int first, int count, const HoneeMiMlpTrainOptions *o, bool adam,
                 float *loss_dst) {
  const int tpb = mlp_grad_block(p->d), nb = (count + tpb - 1) / tpb;
  if (ensure_partials(p, nb)) return 1;
  int nblocks = 0;
  HIPCHK(launch_mlp_grad(p->d, p->theta, dX, dT, idx, first, count, 1.0f / ((float)count * p->d.no), p->partial, p->loss_partial, &nblocks, p->s));
  MlpAdam a{};
  if (adam) {
    p->step++;
    a = MlpAdam{o->lr, o->beta1, o->beta2, o->eps, (float)(1.0 - pow((double)o->beta1, (double)p->step)),
                (float)(1.0 - pow((double)o->beta2, (double)p->step))};
  }
  HIPCHK(launch_mlp_reduce_adam(p->np, nblocks, p->partial, p->loss_partial, p->grad, loss_dst, p->theta, p->m, p->v, a, adam, p->s));
  return 0;
}

extern "C" {

const char *HoneeMiMlpLastError(void) { return g_err.c_str(); }

int HoneeMiMlpCreate(int ni, int nh, int no, float alpha, unsigned long long seed, HoneeMiMlp *mlp) {
  return HoneeMiMlpCreate2(ni, nh, 0, no, alpha, seed, mlp);
}

int HoneeMiMlpCreate2(int ni, int nh, int nh2, int no, float alpha, unsigned long long seed, HoneeMiMlp *mlp) {
  const MlpDims d{ni, nh, no, alpha, nh2};
  if (!mlp_dims_supported(d)) return fail("HoneeMiMlpCreate: only the 6-20-6 and 6-20-20-6 networks have training kernels");
  auto p = new HoneeMiMlp_();
  p->d   = d;
  p->np  = mlp_nparams(d);
  *mlp   = p;
  HIPCHK(hipStreamCreate(&p->s));
  HIPCHK(hipMalloc((void **)&p->theta, sizeof(float) * p->np));
  HIPCHK(hipMalloc((void **)&p->m, sizeof(float) * p->np));
  HIPCHK(hipMalloc((void **)&p->v, sizeof(float) * p->np));
  HIPCHK(hipMalloc((void **)&p->grad, sizeof(float) * p->np));
  HIPCHK(hipMalloc((void **)&p->loss, sizeof(float)));
  std::vector<float> theta(p->np, 0.f);
  SplitMix           rng{seed};
  // Glorot-uniform weights layer by layer, biases zero
  const int    nl = nh2 > 0 ? nh2 : no;
  const double r1 = sqrt(6.0 / (ni + nh)), r2 = sqrt(6.0 / (nh + nl));
  for (int i = 0; i < nh * ni; i++) theta[i] = (float)((2 * rng.uniform() - 1) * r1);
  for (int i = 0; i < nl * nh; i++) theta[nh * ni + nh + i] = (float)((2 * rng.uniform() - 1) * r2);
  if (nh2 > 0) {
    const double r3 = sqrt(6.0 / (nh2 + no));
    const int    o3 = nh * ni + nh + nh2 * nh + nh2;
    for (int i = 0; i < no * nh2; i++) theta[o3 + i] = (float)((2 * rng.uniform() - 1) * r3);
  }
  return HoneeMiMlpSetParams(p, theta.data());
}

int HoneeMiMlpNumParams(HoneeMiMlp p) { return p->np; }

int HoneeMiMlpSetParams(HoneeMiMlp p, const float *theta) {
  HIPCHK(hipMemcpyAsync(p->theta, theta, sizeof(float) * p->np, hipMemcpyHostToDevice, p->s));
  HIPCHK(hipMemsetAsync(p->m, 0, sizeof(float) * p->np, p->s));
  HIPCHK(hipMemsetAsync(p->v, 0, sizeof(float) * p->np, p->s));
  HIPCHK(hipStreamSynchronize(p->s));
  p->step = 0;
  return 0;
}

int HoneeMiMlpGetParams(HoneeMiMlp p, float *theta) {
  HIPCHK(hipMemcpyAsync(theta, p->theta, sizeof(float) * p->np, hipMemcpyDeviceToHost, p->s));
  HIPCHK(hipStreamSynchronize(p->s));
  return 0;
}

int HoneeMiMlpLossGrad(HoneeMiMlp p, int n, const float *X, const float *T, int mem_type, float *loss, float *grad) {
  const float *dX, *dT;
  if (n <= 0) return fail("HoneeMiMlpLossGrad: no samples");
  if (stage(p, n, X, T, mem_type, &dX, &dT)) return 1;
  if (batch(p, dX, dT, nullptr, 0, n, nullptr, false, p->loss)) return 1;
  HIPCHK(hipMemcpyAsync(grad, p->grad, sizeof(float) * p->np, hipMemcpyDeviceToHost, p->s));
  HIPCHK(hipMemcpyAsync(loss, p->loss, sizeof(float), hipMemcpyDeviceToHost, p->s));
  HIPCHK(hipStreamSynchronize(p->s));
  return 0;
}

int HoneeMiMlpAdamStep(HoneeMiMlp p, int n, const float *X, const float *T, int mem_type, const HoneeMiMlpTrainOptions *o, float *loss) {
  const float *dX, *dT;
  if (n <= 0) return fail("HoneeMiMlpAdamStep: no samples");
  if (stage(p, n, X, T, mem_type, &dX, &dT)) return 1;
  if (batch(p, dX, dT, nullptr, 0, n, o, true, p->loss)) return 1;
  HIPCHK(hipMemcpyAsync(loss, p->loss, sizeof(float), hipMemcpyDeviceToHost, p->s));
  HIPCHK(hipStreamSynchronize(p->s));
  return 0;
}

int HoneeMiMlpTrain(HoneeMiMlp p, int n, const float *X, const float *T, int mem_type, const HoneeMiMlpTrainOptions *o, float *epoch_loss) {
  const float *dX, *dT;
  if (n <= 0 || o->batch <= 0 || o->epochs <= 0) return fail("HoneeMiMlpTrain: empty training set or schedule");
  if (stage(p, n, X, T, mem_type, &dX, &dT)) return 1;
  if (n > p->idx_len) {
    if (p->idx) HIPCHK(hipFree(p->idx));
    HIPCHK(hipMalloc((void **)&p->idx, sizeof(int) * (size_t)n));
    p->idx_len = n;
  }
  const int nbatch = (n + o->batch - 1) / o->batch;
  float    *d_bl   = nullptr;
  HIPCHK(hipMalloc((void **)&d_bl, sizeof(float) * (size_t)nbatch * o->epochs));
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  SplitMix rng{o->seed};
  int      ierr = 0;
  for (int ep = 0; ep < o->epochs && !ierr; ep++) {
    for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[rng.next() % (unsigned long long)(i + 1)]);  // Fisher-Yates
    if (hipMemcpyAsync(p->idx, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, p->s) != hipSuccess) ierr = fail("index upload");
    // the host permutation buffer is rewritten next epoch: wait for the copy
    if (!ierr && hipStreamSynchronize(p->s) != hipSuccess) ierr = fail("stream synchronize");
    for (int b = 0; b < nbatch && !ierr; b++) {
      const int first = b * o->batch, count = std::min(o->batch, n - first);
      ierr = batch(p, dX, dT, p->idx, first, count, o, true, d_bl + (size_t)ep * nbatch + b);
    }
  }
  std::vector<float> bl((size_t)nbatch * o->epochs);
  if (!ierr && hipMemcpyAsync(bl.data(), d_bl, sizeof(float) * bl.size(), hipMemcpyDeviceToHost, p->s) != hipSuccess) ierr = fail("loss copy");
  if (!ierr && hipStreamSynchronize(p->s) != hipSuccess) ierr = fail("stream synchronize");
  (void)hipFree(d_bl);
  if (ierr) return ierr;
  for (int ep = 0; ep < o->epochs; ep++) {
    double acc = 0;
    for (int b = 0; b < nbatch; b++) acc += bl[(size_t)ep * nbatch + b];
    epoch_loss[ep] = (float)(acc / nbatch);
  }
  return 0;
}

int HoneeMiMlpPredict(HoneeMiMlp p, int n, const float *X, int mem_type, float *Y) {
  if (n <= 0) return 0;
  const float *dX, *dT;
  if (stage(p, n, X, nullptr, mem_type, &dX, &dT)) return 1;
  float *dY = Y;
  if (mem_type != 1) HIPCHK(hipMalloc((void **)&dY, sizeof(float) * (size_t)n * p->d.no));
  HIPCHK(launch_mlp_predict(p->d, p->theta, dX, n, dY, p->s));
  if (mem_type != 1) {
    HIPCHK(hipMemcpyAsync(Y, dY, sizeof(float) * (size_t)n * p->d.no, hipMemcpyDeviceToHost, p->s));
    HIPCHK(hipStreamSynchronize(p->s));
    HIPCHK(hipFree(dY));
  } else {
    HIPCHK(hipStreamSynchronize(p->s));
  }
  return 0;
}

int HoneeMiMlpDestroy(HoneeMiMlp *mlp) {
  HoneeMiMlp p = *mlp;
  if (!p) return 0;
  if (p->s) (void)hipStreamSynchronize(p->s);
  for (void *a : {(void *)p->theta, (void *)p->m, (void *)p->v, (void *)p->grad, (void *)p->loss, (void *)p->partial, (void *)p->loss_partial,
                  (void *)p->X, (void *)p->T, (void *)p->idx})
    if (a) (void)hipFree(a);
  if (p->s) (void)hipStreamDestroy(p->s);
  delete p;
  *mlp = nullptr;
  return 0;
}

int HoneeMiSgsModelSamples(int n, int mem_type, const double *train, const double *filtered, const double *grad_velo, const double *aniso,
                           const CeedMiGas *gas, float *X, double *S) {
  if (n <= 0) return 0;
  const size_t sizes[6] = {12, 11, 9, 7, 6, 6};
  if (mem_type == 1) {
    HIPCHK(launch_sgs_model_samples(n, train, filtered, grad_velo, aniso, *gas, X, S, nullptr));
    HIPCHK(hipStreamSynchronize(nullptr));
    return 0;
  }
  void       *d[6] = {nullptr};
  const void *h[4] = {train, filtered, grad_velo, aniso};
  for (int i = 0; i < 6; i++) HIPCHK(hipMalloc(&d[i], (i == 4 ? sizeof(float) : sizeof(double)) * sizes[i] * (size_t)n));
  for (int i = 0; i < 4; i++) HIPCHK(hipMemcpy(d[i], h[i], sizeof(double) * sizes[i] * (size_t)n, hipMemcpyHostToDevice));
  HIPCHK(launch_sgs_model_samples(n, (double *)d[0], (double *)d[1], (double *)d[2], (double *)d[3], *gas, (float *)d[4], (double *)d[5], nullptr));
  HIPCHK(hipMemcpy(X, d[4], sizeof(float) * 6 * (size_t)n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(S, d[5], sizeof(double) * 6 * (size_t)n, hipMemcpyDeviceToHost));
  for (int i = 0; i < 6; i++) HIPCHK(hipFree(d[i]));
  return 0;
}

}  // extern "C"
