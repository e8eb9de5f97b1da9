// FP32 training kernels of the subgrid-stress MLP (leaky ReLU), the network that the fused FP64 closure evaluates
// (qfunctions/sgs_dd_model.h:37-53): y = W2 lrelu(W1 x + b1) + b2 (the reference's one hidden layer) or, with a second
// hidden layer, y = W3 lrelu(W2 lrelu(W1 x + b1) + b2) + b3; loss = mean over samples and outputs of (y - t)^2.
//
// k_mlp_grad: one thread per sample runs the forward and backward pass in registers and stages the sample's input,
// activation and the two error vectors in LDS; the block's parameter-gradient partials are then summed over its
// samples in sample order (every gradient entry is a length-256 dot product of two LDS columns), so the result does
// not depend on scheduling.  k_mlp_reduce_adam sums the per-block partials in block order and applies Adam.
// Parameter layout (floats): see MlpDims.  The layer sizes are compile-time constants (6-20-6 and 6-20-20-6) so every
// activation stays in registers.
#include <hip/hip_runtime.h>

#include "kernels_mlp.hpp"
#include "ns_physics.hpp"
#include "sgs_ops.hpp"

namespace {

constexpr int kTpb = 256;

template <int NI, int NH, int NO>
__global__ void __launch_bounds__(kTpb) k_mlp_grad(float alpha, const float *__restrict__ theta, const float *__restrict__ X, const float *__restrict__ T,
                                                   const int *__restrict__ idx, int first, int count, float inv_norm, float *__restrict__ partial,
                                                   float *__restrict__ loss_partial) {
  constexpr int    ni = NI, nh = NH, no = NO, np = NH * NI + NH + NO * NH + NO;
  __shared__ float sP[np];
  __shared__ float sX[NI][kTpb], sA[NH][kTpb], sDH[NH][kTpb], sDY[NO][kTpb];
  __shared__ float sLoss[kTpb];
  const MlpDims    d{NI, NH, NO, alpha};
  for (int i = threadIdx.x; i < np; i += blockDim.x) sP[i] = theta[i];
  __syncthreads();
  const float *W1 = sP, *b1 = sP + nh * ni, *W2 = b1 + nh, *b2 = W2 + no * nh;
  const int    s  = blockIdx.x * kTpb + threadIdx.x;
  const bool   ok = s < count;
  const int    smp = ok ? (idx ? idx[first + s] : first + s) : 0;
  float        x[NI], h[NH], dy[NO], loss = 0;
  for (int i = 0; i < ni; i++) x[i] = ok ? X[(long long)smp * ni + i] : 0.f;
  for (int j = 0; j < nh; j++) {
    float acc = b1[j];
    for (int i = 0; i < ni; i++) acc += W1[j * ni + i] * x[i];
    h[j] = acc;
  }
  for (int k = 0; k < no; k++) {
    float acc = b2[k];
    for (int j = 0; j < nh; j++) acc += W2[k * nh + j] * (h[j] < 0 ? d.alpha * h[j] : h[j]);
    const float e = ok ? acc - T[(long long)smp * no + k] : 0.f;
    loss += e * e;
    dy[k] = 2.f * e * inv_norm;
  }
  for (int i = 0; i < ni; i++) sX[i][threadIdx.x] = x[i];
  for (int j = 0; j < nh; j++) {
    float da = 0;
    for (int k = 0; k < no; k++) da += W2[k * nh + j] * dy[k];
    sA[j][threadIdx.x]  = ok ? (h[j] < 0 ? d.alpha * h[j] : h[j]) : 0.f;
    sDH[j][threadIdx.x] = da * (h[j] < 0 ? d.alpha : 1.f);
  }
  for (int k = 0; k < no; k++) sDY[k][threadIdx.x] = dy[k];
  sLoss[threadIdx.x] = loss * inv_norm;
  __syncthreads();
  // gradient entries of this block, each a fixed-order sum over the block's samples
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    const float *u, *v = nullptr;
    if (p < nh * ni) {
      u = sDH[p / ni], v = sX[p % ni];
    } else if (p < nh * ni + nh) {
      u = sDH[p - nh * ni];
    } else if (p < nh * ni + nh + no * nh) {
      const int q = p - nh * ni - nh;
      u = sDY[q / nh], v = sA[q % nh];
    } else {
      u = sDY[p - nh * ni - nh - no * nh];
    }
    float acc = 0;
    if (v)
      for (int t = 0; t < kTpb; t++) acc += u[t] * v[t];
    else
      for (int t = 0; t < kTpb; t++) acc += u[t];
    partial[(long long)blockIdx.x * np + p] = acc;
  }
  for (int w = kTpb / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sLoss[threadIdx.x] += sLoss[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_partial[blockIdx.x] = sLoss[0];
}

// The same for two hidden layers: y = W3 lrelu(W2 lrelu(W1 x + b1) + b2) + b3 (6-20-20-6).  Six LDS columns per sample
// (input, both activations, both hidden errors, output error), 128 samples per block.
constexpr int kTpb2 = 128;

template <int NI, int NH, int NH2, int NO>
__global__ void __launch_bounds__(kTpb2) k_mlp2_grad(float alpha, const float *__restrict__ theta, const float *__restrict__ X, const float *__restrict__ T,
                                                     const int *__restrict__ idx, int first, int count, float inv_norm, float *__restrict__ partial,
                                                     float *__restrict__ loss_partial) {
  constexpr int    np = NH * NI + NH + NH2 * NH + NH2 + NO * NH2 + NO;
  constexpr int    o1 = NH * NI, o2 = o1 + NH, o3 = o2 + NH2 * NH, o4 = o3 + NH2, o5 = o4 + NO * NH2;
  __shared__ float sP[np];
  __shared__ float sX[NI][kTpb2], sA1[NH][kTpb2], sD1[NH][kTpb2], sA2[NH2][kTpb2], sD2[NH2][kTpb2], sDY[NO][kTpb2];
  __shared__ float sLoss[kTpb2];
  for (int i = threadIdx.x; i < np; i += blockDim.x) sP[i] = theta[i];
  __syncthreads();
  const float *W1 = sP, *b1 = sP + o1, *W2 = sP + o2, *b2 = sP + o3, *W3 = sP + o4, *b3 = sP + o5;
  const int    s   = blockIdx.x * kTpb2 + threadIdx.x;
  const bool   ok  = s < count;
  const int    smp = ok ? (idx ? idx[first + s] : first + s) : 0;
  float        x[NI], h1[NH], a1[NH], h2[NH2], a2[NH2], dy[NO], loss = 0;
  for (int i = 0; i < NI; i++) x[i] = ok ? X[(long long)smp * NI + i] : 0.f;
  for (int j = 0; j < NH; j++) {
    float acc = b1[j];
    for (int i = 0; i < NI; i++) acc += W1[j * NI + i] * x[i];
    h1[j] = acc;
    a1[j] = acc < 0 ? alpha * acc : acc;
  }
  for (int k = 0; k < NH2; k++) {
    float acc = b2[k];
    for (int j = 0; j < NH; j++) acc += W2[k * NH + j] * a1[j];
    h2[k] = acc;
    a2[k] = acc < 0 ? alpha * acc : acc;
  }
  for (int o = 0; o < NO; o++) {
    float acc = b3[o];
    for (int k = 0; k < NH2; k++) acc += W3[o * NH2 + k] * a2[k];
    const float e = ok ? acc - T[(long long)smp * NO + o] : 0.f;
    loss += e * e;
    dy[o] = 2.f * e * inv_norm;
  }
  float d2[NH2];
  for (int k = 0; k < NH2; k++) {
    float da = 0;
    for (int o = 0; o < NO; o++) da += W3[o * NH2 + k] * dy[o];
    d2[k] = da * (h2[k] < 0 ? alpha : 1.f);
  }
  for (int j = 0; j < NH; j++) {
    float da = 0;
    for (int k = 0; k < NH2; k++) da += W2[k * NH + j] * d2[k];
    sD1[j][threadIdx.x] = da * (h1[j] < 0 ? alpha : 1.f);
    sA1[j][threadIdx.x] = ok ? a1[j] : 0.f;
  }
  for (int i = 0; i < NI; i++) sX[i][threadIdx.x] = x[i];
  for (int k = 0; k < NH2; k++) {
    sA2[k][threadIdx.x] = ok ? a2[k] : 0.f;
    sD2[k][threadIdx.x] = d2[k];
  }
  for (int o = 0; o < NO; o++) sDY[o][threadIdx.x] = dy[o];
  sLoss[threadIdx.x] = loss * inv_norm;
  __syncthreads();
  // gradient entries of this block, each a fixed-order sum over the block's samples
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    const float *u, *v = nullptr;
    if (p < o1) {
      u = sD1[p / NI], v = sX[p % NI];
    } else if (p < o2) {
      u = sD1[p - o1];
    } else if (p < o3) {
      u = sD2[(p - o2) / NH], v = sA1[(p - o2) % NH];
    } else if (p < o4) {
      u = sD2[p - o3];
    } else if (p < o5) {
      u = sDY[(p - o4) / NH2], v = sA2[(p - o4) % NH2];
    } else {
      u = sDY[p - o5];
    }
    float acc = 0;
    if (v)
      for (int t = 0; t < kTpb2; t++) acc += u[t] * v[t];
    else
      for (int t = 0; t < kTpb2; t++) acc += u[t];
    partial[(long long)blockIdx.x * np + p] = acc;
  }
  for (int w = kTpb2 / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sLoss[threadIdx.x] += sLoss[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_partial[blockIdx.x] = sLoss[0];
}

template <int NI, int NH, int NH2, int NO>
__global__ void k_mlp2_predict(float alpha, const float *__restrict__ theta, const float *__restrict__ X, int n, float *__restrict__ Y) {
  constexpr int    np = NH * NI + NH + NH2 * NH + NH2 + NO * NH2 + NO;
  constexpr int    o1 = NH * NI, o2 = o1 + NH, o3 = o2 + NH2 * NH, o4 = o3 + NH2, o5 = o4 + NO * NH2;
  __shared__ float sP[np];
  for (int i = threadIdx.x; i < np; i += blockDim.x) sP[i] = theta[i];
  __syncthreads();
  const float *W1 = sP, *b1 = sP + o1, *W2 = sP + o2, *b2 = sP + o3, *W3 = sP + o4, *b3 = sP + o5;
  const int    s  = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[NI], a1[NH], a2[NH2];
  for (int i = 0; i < NI; i++) x[i] = X[(long long)s * NI + i];
  for (int j = 0; j < NH; j++) {
    float acc = b1[j];
    for (int i = 0; i < NI; i++) acc += W1[j * NI + i] * x[i];
    a1[j] = acc < 0 ? alpha * acc : acc;
  }
  for (int k = 0; k < NH2; k++) {
    float acc = b2[k];
    for (int j = 0; j < NH; j++) acc += W2[k * NH + j] * a1[j];
    a2[k] = acc < 0 ? alpha * acc : acc;
  }
  for (int o = 0; o < NO; o++) {
    float acc = b3[o];
    for (int k = 0; k < NH2; k++) acc += W3[o * NH2 + k] * a2[k];
    Y[(long long)s * NO + o] = acc;
  }
}

// Sum of the block partials in block order; Adam (bias-corrected moments) when adam is set, else the gradient only
__global__ void k_mlp_reduce_adam(int np, int nblocks, const float *__restrict__ partial, const float *__restrict__ loss_partial, float *__restrict__ grad,
                                  float *__restrict__ loss, float *theta, float *m, float *v, MlpAdam adam, int apply) {
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    float g = 0;
    for (int b = 0; b < nblocks; b++) g += partial[(long long)b * np + p];
    grad[p] = g;
    if (apply) {
      const float mp = adam.beta1 * m[p] + (1.f - adam.beta1) * g;
      const float vp = adam.beta2 * v[p] + (1.f - adam.beta2) * g * g;
      m[p]           = mp;
      v[p]           = vp;
      const float mh = mp / adam.bc1, vh = vp / adam.bc2;
      theta[p] -= adam.lr * mh / (sqrtf(vh) + adam.eps);
    }
  }
  if (threadIdx.x == 0) {
    float l = 0;
    for (int b = 0; b < nblocks; b++) l += loss_partial[b];
    *loss = l;
  }
}

template <int NI, int NH, int NO>
__global__ void k_mlp_predict(float alpha, const float *__restrict__ theta, const float *__restrict__ X, int n, float *__restrict__ Y) {
  constexpr int    ni = NI, nh = NH, no = NO, np = NH * NI + NH + NO * NH + NO;
  __shared__ float sP[np];
  const MlpDims    d{NI, NH, NO, alpha};
  for (int i = threadIdx.x; i < np; i += blockDim.x) sP[i] = theta[i];
  __syncthreads();
  const float *W1 = sP, *b1 = sP + nh * ni, *W2 = b1 + nh, *b2 = W2 + no * nh;
  const int    s  = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[NI], a[NH];
  for (int i = 0; i < ni; i++) x[i] = X[(long long)s * ni + i];
  for (int j = 0; j < nh; j++) {
    float acc = b1[j];
    for (int i = 0; i < ni; i++) acc += W1[j * ni + i] * x[i];
    a[j] = acc < 0 ? d.alpha * acc : acc;
  }
  for (int k = 0; k < no; k++) {
    float acc = b2[k];
    for (int j = 0; j < nh; j++) acc += W2[k * nh + j] * a[j];
    Y[(long long)s * no + k] = acc;
  }
}

// Model-space samples from the nodal training fields: the closure inputs of the filtered state (ComputeSgsDDInputs,
// qfunctions/sgs_dd_utils.h:58-90) and the filtered SGS stress tau rotated into the input's eigenframe and scaled as
// ComputeSgsDDOutputs undoes it (qfunctions/sgs_dd_utils.h:92-127): S = E tau E^T / (delta^2 |grad u|^2), stored in
// the model's output order S00 S11 S22 S01 S02 S12.
__global__ void k_sgs_model_samples(long long n, const double *__restrict__ train, const double *__restrict__ filtered, const double *__restrict__ G,
                                    const double *__restrict__ A, CeedMiGas gas, float *__restrict__ X, double *__restrict__ S) {
  const auto g = ns::Gas<double>::from(gas);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double        *yf = filtered + i * 11;
    double               yi[5], gu[3][3], kmA[6], E[3][3], in[6], gmag;
#pragma unroll
    for (int c = 0; c < 5; c++) yi[c] = yf[c];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int k = 0; k < 3; k++) gu[j][k] = G[i * 9 + j * 3 + k];
#pragma unroll
    for (int c = 0; c < 6; c++) kmA[c] = A[i * 7 + c];
    const double         delta = A[i * 7 + 6];
    const ns::St<double> s     = ns::state_from_q<ns::VAR_PRIM>(g, yi);
    sgs::sgs_inputs(gu, kmA, delta, g.mu / s.U[0], E, in, gmag);
    const double *t       = train + i * 12 + 6;  // XX YY ZZ YZ XZ XY
    const double  tau[3][3] = {{t[0], t[5], t[4]}, {t[5], t[1], t[3]}, {t[4], t[3], t[2]}};
    double        Et[3][3], R[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) Et[a][b] = E[a][0] * tau[0][b] + E[a][1] * tau[1][b] + E[a][2] * tau[2][b];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) R[a][b] = Et[a][0] * E[b][0] + Et[a][1] * E[b][1] + Et[a][2] * E[b][2];
    const double sc = 1.0 / (delta * delta * (gmag * gmag));
#pragma unroll
    for (int c = 0; c < 6; c++) X[i * 6 + c] = (float)in[c];
    double *o = S + i * 6;
    o[0]      = R[0][0] * sc;
    o[1]      = R[1][1] * sc;
    o[2]      = R[2][2] * sc;
    o[3]      = R[0][1] * sc;
    o[4]      = R[0][2] * sc;
    o[5]      = R[1][2] * sc;
  }
}

}  // namespace

// 6-20-6 (the reference's closure) and 6-20-20-6
bool mlp_dims_supported(const MlpDims &d) { return d.ni == 6 && d.nh == 20 && d.no == 6 && (d.nh2 == 0 || d.nh2 == 20); }

int mlp_grad_block(const MlpDims &d) { return d.nh2 > 0 ? kTpb2 : kTpb; }

hipError_t launch_mlp_grad(const MlpDims &d, const float *theta, const float *X, const float *T, const int *idx, int first, int count, float inv_norm,
                           float *partial, float *loss_partial, int *nblocks, hipStream_t s) {
  const int tpb = mlp_grad_block(d);
  const int nb  = (count + tpb - 1) / tpb;
  *nblocks      = nb;
  if (!mlp_dims_supported(d)) return hipErrorInvalidValue;
  if (nb <= 0) return hipSuccess;
  if (d.nh2 > 0)
    hipLaunchKernelGGL((k_mlp2_grad<6, 20, 20, 6>), dim3(nb), dim3(tpb), 0, s, d.alpha, theta, X, T, idx, first, count, inv_norm, partial, loss_partial);
  else
    hipLaunchKernelGGL((k_mlp_grad<6, 20, 6>), dim3(nb), dim3(tpb), 0, s, d.alpha, theta, X, T, idx, first, count, inv_norm, partial, loss_partial);
  return hipGetLastError();
}

hipError_t launch_mlp_reduce_adam(int np, int nblocks, const float *partial, const float *loss_partial, float *grad, float *loss, float *theta, float *m,
                                  float *v, const MlpAdam &adam, bool apply, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(1), dim3(512), 0, s, np, nblocks, partial, loss_partial, grad, loss, theta, m, v, adam, apply ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_mlp_predict(const MlpDims &d, const float *theta, const float *X, int n, float *Y, hipStream_t s) {
  if (!mlp_dims_supported(d)) return hipErrorInvalidValue;
  if (n <= 0) return hipSuccess;
  if (d.nh2 > 0) hipLaunchKernelGGL((k_mlp2_predict<6, 20, 20, 6>), dim3((n + kTpb - 1) / kTpb), dim3(kTpb), 0, s, d.alpha, theta, X, n, Y);
  else hipLaunchKernelGGL((k_mlp_predict<6, 20, 6>), dim3((n + kTpb - 1) / kTpb), dim3(kTpb), 0, s, d.alpha, theta, X, n, Y);
  return hipGetLastError();
}

hipError_t launch_sgs_model_samples(long long n, const double *train, const double *filtered, const double *G, const double *A, const CeedMiGas &gas,
                                    float *X, double *S, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
  hipLaunchKernelGGL(k_sgs_model_samples, dim3(grid), dim3(256), 0, s, n, train, filtered, G, A, gas, X, S);
  return hipGetLastError();
}
