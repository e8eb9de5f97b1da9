// Training of the data-driven SGS model on the GPU (FP32): the 6-20-6 (or 6-20-20-6) leaky-ReLU MLP that the fused
// FP64 closure evaluates (problems/sgs_dd_model.c, qfunctions/sgs_dd_model.h:37-53), trained by mean-squared error with
// Adam on model-space samples built from the filtered training fields (src/smartsim/sgs_dd_training.c).  The reference
// trains its network outside the solver; these entry points have no counterpart in it.
#pragma once

#include "honee_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct HoneeMiMlp_ *HoneeMiMlp;

typedef struct {
  float              lr, beta1, beta2, eps;  // Adam
  int                batch, epochs;
  unsigned long long seed;  // sample order of each epoch
} HoneeMiMlpTrainOptions;

// Layer sizes must be 6, 20, 6.  Weights Glorot-uniform from seed, biases zero.  Parameter vector layout (FP32):
// W1 [nh][ni], b1 [nh], W2 [no][nh], b2 [no].
int HoneeMiMlpCreate(int ni, int nh, int no, float alpha, unsigned long long seed, HoneeMiMlp *mlp);
// With a second hidden layer of nh2 neurons (0: as HoneeMiMlpCreate); sizes 6, 20, 20, 6.  Layout: W1 [nh][ni], b1 [nh],
// W2 [nh2][nh], b2 [nh2], W3 [no][nh2], b3 [no].
int HoneeMiMlpCreate2(int ni, int nh, int nh2, int no, float alpha, unsigned long long seed, HoneeMiMlp *mlp);
int HoneeMiMlpNumParams(HoneeMiMlp mlp);
int HoneeMiMlpSetParams(HoneeMiMlp mlp, const float *theta);  // host arrays; resets the Adam state
int HoneeMiMlpGetParams(HoneeMiMlp mlp, float *theta);
// Mean squared error and its gradient over n samples X [n][ni], T [n][no] (mem_type 0 host, 1 device arrays)
int HoneeMiMlpLossGrad(HoneeMiMlp mlp, int n, const float *X, const float *T, int mem_type, float *loss, float *grad);
// One Adam step on the batch of all n samples; loss before the step
int HoneeMiMlpAdamStep(HoneeMiMlp mlp, int n, const float *X, const float *T, int mem_type, const HoneeMiMlpTrainOptions *o, float *loss);
// Mini-batch Adam over epochs (sample order shuffled per epoch from o->seed); epoch_loss [epochs] (mean batch loss)
int HoneeMiMlpTrain(HoneeMiMlp mlp, int n, const float *X, const float *T, int mem_type, const HoneeMiMlpTrainOptions *o, float *epoch_loss);
int HoneeMiMlpPredict(HoneeMiMlp mlp, int n, const float *X, int mem_type, float *Y);
int HoneeMiMlpDestroy(HoneeMiMlp *mlp);

// Model-space samples of n nodes from the training fields (host or device arrays): train [n][12] (model inputs,
// filtered SGS stress XX YY ZZ YZ XZ XY), filtered [n][11], grad_velo [n][9], aniso [n][7] -> X [n][6] (FP32 model
// inputs) and S [n][6] (FP64 stress in the inputs' eigenframe divided by delta^2 |grad u|^2, order S00 S11 S22 S01
// S02 S12: the values the closure's de-normalization maps the network outputs to).
int HoneeMiSgsModelSamples(int n, int mem_type, const double *train, const double *filtered, const double *grad_velo, const double *aniso,
                           const CeedMiGas *gas, float *X, double *S);

const char *HoneeMiMlpLastError(void);

#ifdef __cplusplus
}
#endif
