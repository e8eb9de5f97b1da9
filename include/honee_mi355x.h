// MI355X backend extensions and the QFunction context layouts the backend interprets.
//
// The fused kernels read HONEE's QFunction context structs by value, so their binary layout is part of the interface:
// each struct below has exactly the member order and types of the reference struct it mirrors (cited), and the
// backend checks the registered context size against sizeof() before binding an operator.
#pragma once

#include "ceed.h"

#ifdef __cplusplus
extern "C" {
#endif

// struct NewtonianIdealGasContext_ (qfunctions/newtonian_types.h:20-47, stabilization_types.h:32-38)
typedef struct {
  double lambda, mu, k, cv, cp;  // NewtonianIGProperties
} CeedMiGas;

typedef struct {
  double Ctau_t, Ctau_v, Ctau_C, Ctau_M, Ctau_E;  // TauDiagCoefficients
} CeedMiTauCoeffs;

typedef struct {
  CeedMiGas       gas;
  double          g[3];
  double          time;
  double          ijacobian_time_shift;
  bool            is_implicit;
  int             state_var;      // StateVariable: 0 conservative, 1 primitive, 2 entropy
  int             stabilization;  // StabilizationType: 0 none, 1 SU, 2 SUPG
  bool            idl_enable;
  double          idl_pressure;
  double          idl_amplitude;
  double          idl_start;
  double          idl_length;
  CeedMiTauCoeffs tau_coeffs;
  double          dt;
  int             divFdiff_method;  // DivDiffFluxProjectionMethod: only NONE (0) is supported by the fused kernels
} CeedMiNewtonianContext;

// State (qfunctions/newtonian_state.h:13-22): conservative then primitive
typedef struct {
  double U[5];
  double Y[5];
} CeedMiState;

// struct FreestreamContext_ (qfunctions/bc_freestream_type.h:9-12)
typedef struct {
  CeedMiNewtonianContext newt_ctx;
  CeedMiState            S_infty;
} CeedMiFreestreamContext;

// struct OutflowContext_ (qfunctions/bc_freestream_type.h:14-21)
typedef struct {
  CeedMiNewtonianContext newt_ctx;
  double                 recirc;             // amount of recirculation allowed in the exterior state [0, 1]
  double                 softplus_velocity;  // width of the softplus velocity regularization
  double                 pressure;
  double                 temperature;
} CeedMiOutflowContext;

// struct TaylorGreenContext_ (qfunctions/taylorgreen.h:13-20)
typedef struct {
  double                 reference[5];  // StatePrimitive: pressure, velocity[3], temperature
  CeedMiNewtonianContext newt_ctx;
  double                 lx, ly, lz;
  double                 u[3];
} CeedMiTaylorGreenContext;

// struct SgsDDContext_ (qfunctions/sgs_dd_model.h:15-31): the data-driven SGS model, its parameters in data[] at the
// given offsets (bias1[num_neurons], bias2[num_outputs], weight1[num_neurons][num_inputs],
// weight2[num_outputs][num_neurons], out_scaling[num_outputs][2]); total_bytes is the size of the whole allocation.
// Extension (the reference evaluates one hidden layer): with num_layers = 2, a second hidden layer of num_neurons
// follows out_scaling in data[]: its biases [num_neurons], then its weights [num_neurons (out)][num_neurons (in)];
// weight2 then maps the second hidden layer to the outputs.
typedef struct {
  CeedInt                num_inputs, num_outputs;
  CeedInt                num_layers;
  CeedInt                num_neurons;
  double                 alpha;
  CeedMiNewtonianContext newt_ctx;
  struct {
    size_t bias1, bias2;
    size_t weight1, weight2;
    size_t out_scaling;
  } offsets;
  size_t total_bytes;
  double data[1];
} CeedMiSgsDDContext;

// struct DifferentialFilterContext_ (qfunctions/differential_filter.h:16-25); damping_function is
// enum DifferentialFilterDampingFunction: 0 none, 1 van Driest, 2 MMS
typedef struct {
  bool                   grid_based_width;
  double                 width_scaling[3];
  double                 kernel_scaling;
  double                 friction_length;
  int                    damping_function;
  double                 damping_constant;
  CeedMiNewtonianContext newt_ctx;
} CeedMiDiffFilterContext;

// struct SGS_DD_TrainingContext_ (qfunctions/sgs_dd_training.h:17-20)
typedef struct {
  CeedMiNewtonianContext newt_ctx;
} CeedMiSgsTrainingContext;

// ---- Backend extensions (no counterpart in libCEED; used by this repository's driver, tests and benchmark) ----

// HIP stream the backend launches on (a hipStream_t).
int CeedMiGetStream(Ceed ceed, void **hip_stream);

// Kernel timing: when enabled, every fused kernel launch of an operator apply is bracketed by HIP events on the
// backend stream.  CeedMiGetKernelTimes returns, per kernel name (comma-separated in names), the number of launches
// and the summed event time in milliseconds since the last reset.
int CeedMiSetKernelTiming(Ceed ceed, bool enable);
int CeedMiResetKernelTimes(Ceed ceed);
int CeedMiGetKernelTimes(Ceed ceed, int max_kernels, char *names, size_t names_len, int *launches, double *total_ms, int *num_kernels);
// Region timing on the backend stream, independent of per-kernel timing: the work launched between Begin and End is
// bracketed by two events (no events between its kernels) and reported by CeedMiGetKernelTimes under name, one
// launch per region.  For steady-state durations of back-to-back operator applies.
int CeedMiTimeRegionBegin(Ceed ceed);
int CeedMiTimeRegionEnd(Ceed ceed, const char *name);

// Device pointer of a vector's FP32 shadow array (the jac_data mirror written in jacobian_precision=fp32 mode).
int CeedMiVectorGetFp32ShadowRead(CeedVector vec, const float **d_array, CeedSize *length);

// Dot product x . y of two vectors of equal length, reduced on the device over a fixed grid and summed on the host in
// a fixed order (the same result for the same inputs on every call).
int CeedMiVectorDot(CeedVector x, CeedVector y, double *result);

// Synchronize the backend stream.
int CeedMiSynchronize(Ceed ceed);

// CeedOperatorApply that allocates nothing.  It has no counterpart in libCEED or HONEE: it is for callers that cannot
// allocate device memory, such as code under HIP graph capture or a tight serving loop.  The operator must have been
// applied once with CeedOperatorApply (which binds it and uploads its offsets and scatter maps), and in and out must
// hold device arrays.  The two-pass scatter's E-vector then lives in a fixed device array of 96 MiB
// (12 Mi doubles: elements x nodes per element x 5 must fit, about 39,000 p=3 elements); larger operators, and any
// step that would allocate, return an error instead.  Every call on one GPU shares that array and it keeps the last
// call's contents: calls that overlap on one GPU (other streams or threads) race.  Kernel timing is skipped.
int CeedMiOperatorApplyNoAlloc(CeedOperator op, CeedVector in, CeedVector out);

#ifdef __cplusplus
}
#endif
