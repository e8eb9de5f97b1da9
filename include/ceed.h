// Drop-in libCEED public C API (the subset HONEE calls), implemented by the MI355X backend library libceed.so.
//
// Resource strings: "/gpu/hip/mi355x" (fused HIP kernels for gfx950), with optional ":option=value" suffixes
//   :device_id=<n>              HIP device ordinal (default: current device)
//   :jacobian_precision=fp32    run Newtonian IJacobian operators in FP32 (FP32 jac_data mirror, FP32 Tau and flux
//                               Jacobians); FP64 is the default
//   :scatter=twopass|atomic     scatter-add strategy of volume operators: twopass (the default) writes element-node
//                               contributions and sums them per node in a fixed order (deterministic); atomic adds
//                               them with FP64 atomics
//   :qdata=stored|recompute     stored (the default): volume kernels read the Setup operator's quadrature data;
//                               recompute: the IFunction (P = Q = 4, 5) and FP64 IJacobian (P = Q = 4) kernels rebuild
//                               it from the element vertex coordinates while the qdata vector holds the Setup
//                               operator's unmodified output (bitwise-identical results)
// A CeedInit failure stores its message without printing; CeedGetErrorMessage(NULL, &msg) returns it.
// The resource name sits under /gpu/hip so HONEE selects device vectors (examples/navierstokes.c:95-103).
//
// Every entry point returns an int error code (CEED_ERROR_SUCCESS == 0) and never aborts; the message of the last
// error is available from CeedGetErrorMessage, as HONEE's PetscCallCeed expects (include/petsc-ceed.h:25-36).
// Names, argument order and meaning follow libCEED so that HONEE's call sites compile unchanged; the reference call
// site each group replaces is cited beside it (paths relative to the HONEE source tree).
#pragma once

#include <stdarg.h>
#include <stdio.h>

#include "ceed/types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CEED_VERSION_MAJOR 0
#define CEED_VERSION_MINOR 12
#define CEED_VERSION_PATCH 0
#define CEED_VERSION_RELEASE false

typedef struct Ceed_private                 *Ceed;
typedef struct CeedVector_private           *CeedVector;
typedef struct CeedElemRestriction_private  *CeedElemRestriction;
typedef struct CeedBasis_private            *CeedBasis;
typedef struct CeedQFunction_private        *CeedQFunction;
typedef struct CeedQFunctionContext_private *CeedQFunctionContext;
typedef struct CeedOperator_private         *CeedOperator;
typedef struct CeedOperatorField_private    *CeedOperatorField;
typedef struct CeedContextFieldLabel_private *CeedContextFieldLabel;
typedef struct CeedRequest_private          *CeedRequest;

typedef enum { CEED_MEM_HOST, CEED_MEM_DEVICE } CeedMemType;
// Names of the memory types, indexed by CeedMemType (src/misc.c:330)
extern const char *const CeedMemTypes[];
typedef enum { CEED_COPY_VALUES, CEED_USE_POINTER, CEED_OWN_POINTER } CeedCopyMode;
typedef enum { CEED_NORM_1, CEED_NORM_2, CEED_NORM_MAX } CeedNormType;
typedef enum {
  CEED_EVAL_NONE   = 0,
  CEED_EVAL_INTERP = 1,
  CEED_EVAL_GRAD   = 2,
  CEED_EVAL_DIV    = 4,
  CEED_EVAL_CURL   = 8,
  CEED_EVAL_WEIGHT = 16,
} CeedEvalMode;
typedef enum {
  CEED_GAUSS         = 0,
  CEED_GAUSS_LOBATTO = 1,
} CeedQuadMode;
#define CEED_TOPOLOGY_DIM(topo) ((topo) >> 16)
typedef enum {
  CEED_TOPOLOGY_LINE  = 1 << 16 | 0,
  CEED_TOPOLOGY_TRIANGLE = 2 << 16 | 1,
  CEED_TOPOLOGY_QUAD  = 2 << 16 | 2,
  CEED_TOPOLOGY_TET   = 3 << 16 | 3,
  CEED_TOPOLOGY_PYRAMID = 3 << 16 | 4,
  CEED_TOPOLOGY_PRISM = 3 << 16 | 5,
  CEED_TOPOLOGY_HEX   = 3 << 16 | 6,
} CeedElemTopology;
typedef enum {
  CEED_CONTEXT_FIELD_DOUBLE = 1,
  CEED_CONTEXT_FIELD_INT32  = 2,
  CEED_CONTEXT_FIELD_BOOL   = 3,
} CeedContextFieldType;

typedef int (*CeedQFunctionUser)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out);
typedef int (*CeedQFunctionContextDataDestroyUser)(void *data);
typedef int (*CeedErrorHandler)(Ceed, const char *, int, const char *, int, const char *, va_list *);

// Sentinel objects (libCEED include/ceed/ceed.h)
extern const CeedVector          CEED_VECTOR_ACTIVE;
extern const CeedVector          CEED_VECTOR_NONE;
extern const CeedBasis           CEED_BASIS_NONE;
extern const CeedElemRestriction CEED_ELEMRESTRICTION_NONE;
extern const CeedQFunction       CEED_QFUNCTION_NONE;
extern const CeedInt             CEED_STRIDES_BACKEND[3];
extern CeedRequest *const        CEED_REQUEST_IMMEDIATE;
extern CeedRequest *const        CEED_REQUEST_ORDERED;

// ---- Library / errors: examples/navierstokes.c:64-66, include/petsc-ceed.h:25-36, src/honee_versioning.c
int  CeedInit(const char *resource, Ceed *ceed);
int  CeedReferenceCopy(Ceed ceed, Ceed *ceed_copy);
int  CeedGetResource(Ceed ceed, const char **resource);
int  CeedGetPreferredMemType(Ceed ceed, CeedMemType *type);
int  CeedSetStream(Ceed ceed, void *handle);
int  CeedSetErrorHandler(Ceed ceed, CeedErrorHandler handler);
int  CeedGetErrorMessage(Ceed ceed, const char **err_msg);
int  CeedResetErrorMessage(Ceed ceed, const char **err_msg);
int  CeedGetVersion(int *major, int *minor, int *patch, bool *release);
int  CeedGetGitVersion(const char **git_version);
int  CeedDestroy(Ceed *ceed);
int  CeedErrorImpl(Ceed ceed, const char *filename, int lineno, const char *func, int ecode, const char *format, ...);
int  CeedErrorReturn(Ceed ceed, const char *filename, int lineno, const char *func, int ecode, const char *format, va_list *args);
int  CeedErrorStore(Ceed ceed, const char *filename, int lineno, const char *func, int ecode, const char *format, va_list *args);
int  CeedErrorAbort(Ceed ceed, const char *filename, int lineno, const char *func, int ecode, const char *format, va_list *args);
#define CeedError(ceed, ecode, ...) CeedErrorImpl((ceed), __FILE__, __LINE__, __func__, (ecode), __VA_ARGS__)

// ---- Vectors: include/petsc-ceed-utils.h:127-160, src/qdata.c:36-47, src/setuplibceed.c:131-135
int  CeedVectorCreate(Ceed ceed, CeedSize len, CeedVector *vec);
int  CeedVectorReferenceCopy(CeedVector vec, CeedVector *vec_copy);
int  CeedVectorCopy(CeedVector vec, CeedVector vec_copy);
int  CeedVectorSetArray(CeedVector vec, CeedMemType mem_type, CeedCopyMode copy_mode, CeedScalar *array);
int  CeedVectorSetValue(CeedVector vec, CeedScalar value);
int  CeedVectorSyncArray(CeedVector vec, CeedMemType mem_type);
int  CeedVectorTakeArray(CeedVector vec, CeedMemType mem_type, CeedScalar **array);
int  CeedVectorGetArray(CeedVector vec, CeedMemType mem_type, CeedScalar **array);
int  CeedVectorGetArrayRead(CeedVector vec, CeedMemType mem_type, const CeedScalar **array);
int  CeedVectorGetArrayWrite(CeedVector vec, CeedMemType mem_type, CeedScalar **array);
int  CeedVectorRestoreArray(CeedVector vec, CeedScalar **array);
int  CeedVectorRestoreArrayRead(CeedVector vec, const CeedScalar **array);
int  CeedVectorNorm(CeedVector vec, CeedNormType type, CeedScalar *norm);
int  CeedVectorScale(CeedVector x, CeedScalar alpha);
int  CeedVectorAXPY(CeedVector y, CeedScalar alpha, CeedVector x);
int  CeedVectorAXPBY(CeedVector y, CeedScalar alpha, CeedScalar beta, CeedVector x);
int  CeedVectorPointwiseMult(CeedVector w, CeedVector x, CeedVector y);
int  CeedVectorReciprocal(CeedVector vec);
int  CeedVectorView(CeedVector vec, const char *fp_fmt, FILE *stream);
int  CeedVectorGetCeed(CeedVector vec, Ceed *ceed);
Ceed CeedVectorReturnCeed(CeedVector vec);
int  CeedVectorGetLength(CeedVector vec, CeedSize *length);
int  CeedVectorDestroy(CeedVector *vec);

// ---- Element restrictions: src/dm-utils.c:108-112 (offsets), src/dm-utils.c:235 (strided qdata), src/misc.c:29
int  CeedElemRestrictionCreate(Ceed ceed, CeedInt num_elem, CeedInt elem_size, CeedInt num_comp, CeedInt comp_stride, CeedSize l_size,
                               CeedMemType mem_type, CeedCopyMode copy_mode, const CeedInt *offsets, CeedElemRestriction *rstr);
int  CeedElemRestrictionCreateStrided(Ceed ceed, CeedInt num_elem, CeedInt elem_size, CeedInt num_comp, CeedSize l_size, const CeedInt strides[3],
                                      CeedElemRestriction *rstr);
int  CeedElemRestrictionReferenceCopy(CeedElemRestriction rstr, CeedElemRestriction *rstr_copy);
int  CeedElemRestrictionCreateVector(CeedElemRestriction rstr, CeedVector *l_vec, CeedVector *e_vec);
int  CeedElemRestrictionApply(CeedElemRestriction rstr, CeedTransposeMode t_mode, CeedVector u, CeedVector ru, CeedRequest *request);
int  CeedElemRestrictionGetCeed(CeedElemRestriction rstr, Ceed *ceed);
Ceed CeedElemRestrictionReturnCeed(CeedElemRestriction rstr);
int  CeedElemRestrictionGetNumElements(CeedElemRestriction rstr, CeedInt *num_elem);
int  CeedElemRestrictionGetElementSize(CeedElemRestriction rstr, CeedInt *elem_size);
int  CeedElemRestrictionGetLVectorSize(CeedElemRestriction rstr, CeedSize *l_size);
int  CeedElemRestrictionGetEVectorSize(CeedElemRestriction rstr, CeedSize *e_size);
int  CeedElemRestrictionGetNumComponents(CeedElemRestriction rstr, CeedInt *num_comp);
int  CeedElemRestrictionGetMultiplicity(CeedElemRestriction rstr, CeedVector mult);
int  CeedElemRestrictionDestroy(CeedElemRestriction *rstr);

// ---- Bases: src/dm-utils.c:844-850 (tensor H1), src/dm-utils.c:831 (non-tensor), src/setuplibceed.c:165 (projection)
int  CeedBasisCreateTensorH1(Ceed ceed, CeedInt dim, CeedInt num_comp, CeedInt P_1d, CeedInt Q_1d, const CeedScalar *interp_1d,
                             const CeedScalar *grad_1d, const CeedScalar *q_ref_1d, const CeedScalar *q_weight_1d, CeedBasis *basis);
int  CeedBasisCreateH1(Ceed ceed, CeedElemTopology topo, CeedInt num_comp, CeedInt num_nodes, CeedInt num_qpts, const CeedScalar *interp,
                       const CeedScalar *grad, const CeedScalar *q_ref, const CeedScalar *q_weights, CeedBasis *basis);
int  CeedBasisCreateProjection(CeedBasis basis_from, CeedBasis basis_to, CeedBasis *basis_project);
int  CeedBasisReferenceCopy(CeedBasis basis, CeedBasis *basis_copy);
int  CeedBasisGetCeed(CeedBasis basis, Ceed *ceed);
Ceed CeedBasisReturnCeed(CeedBasis basis);
int  CeedBasisGetDimension(CeedBasis basis, CeedInt *dim);
int  CeedBasisGetNumComponents(CeedBasis basis, CeedInt *num_comp);
int  CeedBasisGetNumNodes(CeedBasis basis, CeedInt *P);
int  CeedBasisGetNumNodes1D(CeedBasis basis, CeedInt *P_1d);
int  CeedBasisGetNumQuadraturePoints(CeedBasis basis, CeedInt *Q);
int  CeedBasisGetNumQuadraturePoints1D(CeedBasis basis, CeedInt *Q_1d);
int  CeedBasisDestroy(CeedBasis *basis);

// ---- QFunctions: src/setuplibceed.c:212-226, 248-258; src/qdata.c:154-169
int  CeedQFunctionCreateInterior(Ceed ceed, CeedInt vec_length, CeedQFunctionUser f, const char *source, CeedQFunction *qf);
int  CeedQFunctionCreateIdentity(Ceed ceed, CeedInt size, CeedEvalMode in_mode, CeedEvalMode out_mode, CeedQFunction *qf);
int  CeedQFunctionReferenceCopy(CeedQFunction qf, CeedQFunction *qf_copy);
int  CeedQFunctionAddInput(CeedQFunction qf, const char *field_name, CeedInt size, CeedEvalMode eval_mode);
int  CeedQFunctionAddOutput(CeedQFunction qf, const char *field_name, CeedInt size, CeedEvalMode eval_mode);
int  CeedQFunctionSetContext(CeedQFunction qf, CeedQFunctionContext ctx);
int  CeedQFunctionGetContext(CeedQFunction qf, CeedQFunctionContext *ctx);
int  CeedQFunctionSetUserFlopsEstimate(CeedQFunction qf, CeedSize flops);
int  CeedQFunctionDestroy(CeedQFunction *qf);

// ---- QFunction contexts: problems/newtonian.c:731-753
int  CeedQFunctionContextCreate(Ceed ceed, CeedQFunctionContext *ctx);
int  CeedQFunctionContextReference(CeedQFunctionContext ctx);
int  CeedQFunctionContextReferenceCopy(CeedQFunctionContext ctx, CeedQFunctionContext *ctx_copy);
int  CeedQFunctionContextSetData(CeedQFunctionContext ctx, CeedMemType mem_type, CeedCopyMode copy_mode, size_t size, void *data);
int  CeedQFunctionContextSetDataDestroy(CeedQFunctionContext ctx, CeedMemType f_mem_type, CeedQFunctionContextDataDestroyUser f);
int  CeedQFunctionContextGetData(CeedQFunctionContext ctx, CeedMemType mem_type, void *data);
int  CeedQFunctionContextGetDataRead(CeedQFunctionContext ctx, CeedMemType mem_type, void *data);
int  CeedQFunctionContextRestoreData(CeedQFunctionContext ctx, void *data);
int  CeedQFunctionContextRestoreDataRead(CeedQFunctionContext ctx, void *data);
int  CeedQFunctionContextGetContextSize(CeedQFunctionContext ctx, size_t *ctx_size);
int  CeedQFunctionContextRegisterDouble(CeedQFunctionContext ctx, const char *field_name, size_t field_offset, size_t num_values,
                                        const char *field_description);
int  CeedQFunctionContextRegisterInt32(CeedQFunctionContext ctx, const char *field_name, size_t field_offset, size_t num_values,
                                       const char *field_description);
int  CeedQFunctionContextRegisterBoolean(CeedQFunctionContext ctx, const char *field_name, size_t field_offset, size_t num_values,
                                         const char *field_description);
int  CeedQFunctionContextDestroy(CeedQFunctionContext *ctx);

// ---- Operators: src/setuplibceed.c:228-318, src/setupts.c:91-109, src/mat-ceed.c:1603-1694, src/petsc_ops.c:218-244
int  CeedOperatorCreate(Ceed ceed, CeedQFunction qf, CeedQFunction dqf, CeedQFunction dqfT, CeedOperator *op);
int  CeedOperatorCreateComposite(Ceed ceed, CeedOperator *op);
int  CeedOperatorCompositeAddSub(CeedOperator composite_op, CeedOperator sub_op);
int  CeedOperatorCompositeGetNumSub(CeedOperator op, CeedInt *num_sub);
int  CeedOperatorCompositeGetSubList(CeedOperator op, CeedOperator **sub_ops);
int  CeedOperatorIsComposite(CeedOperator op, bool *is_composite);
int  CeedOperatorReferenceCopy(CeedOperator op, CeedOperator *op_copy);
int  CeedOperatorSetField(CeedOperator op, const char *field_name, CeedElemRestriction rstr, CeedBasis basis, CeedVector vec);
int  CeedOperatorGetFields(CeedOperator op, CeedInt *num_input_fields, CeedOperatorField **input_fields, CeedInt *num_output_fields,
                           CeedOperatorField **output_fields);
int  CeedOperatorGetFieldByName(CeedOperator op, const char *field_name, CeedOperatorField *op_field);
int  CeedOperatorFieldGetName(CeedOperatorField op_field, const char **field_name);
int  CeedOperatorFieldGetElemRestriction(CeedOperatorField op_field, CeedElemRestriction *rstr);
int  CeedOperatorFieldGetBasis(CeedOperatorField op_field, CeedBasis *basis);
int  CeedOperatorFieldGetVector(CeedOperatorField op_field, CeedVector *vec);
int  CeedOperatorFieldGetData(CeedOperatorField op_field, const char **field_name, CeedElemRestriction *rstr, CeedBasis *basis, CeedVector *vec);
int  CeedOperatorSetName(CeedOperator op, const char *name);
int  CeedOperatorGetName(CeedOperator op, const char **name);
int  CeedOperatorGetContext(CeedOperator op, CeedQFunctionContext *ctx);
int  CeedOperatorGetContextFieldLabel(CeedOperator op, const char *field_name, CeedContextFieldLabel *field_label);
int  CeedOperatorSetContextDouble(CeedOperator op, CeedContextFieldLabel field_label, double *values);
int  CeedOperatorGetContextDoubleRead(CeedOperator op, CeedContextFieldLabel field_label, size_t *num_values, const double **values);
int  CeedOperatorRestoreContextDoubleRead(CeedOperator op, CeedContextFieldLabel field_label, const double **values);
int  CeedOperatorGetActiveVectorLengths(CeedOperator op, CeedSize *input_size, CeedSize *output_size);
int  CeedOperatorApply(CeedOperator op, CeedVector in, CeedVector out, CeedRequest *request);
int  CeedOperatorApplyAdd(CeedOperator op, CeedVector in, CeedVector out, CeedRequest *request);
int  CeedOperatorLinearAssembleDiagonal(CeedOperator op, CeedVector assembled, CeedRequest *request);
int  CeedOperatorLinearAssembleAddDiagonal(CeedOperator op, CeedVector assembled, CeedRequest *request);
int  CeedOperatorLinearAssembleAddPointBlockDiagonal(CeedOperator op, CeedVector assembled, CeedRequest *request);
int  CeedOperatorLinearAssemblePointBlockDiagonal(CeedOperator op, CeedVector assembled, CeedRequest *request);
int  CeedOperatorLinearAssemblePointBlockDiagonalSymbolic(CeedOperator op, CeedSize *num_entries, CeedInt **rows, CeedInt **cols);
int  CeedOperatorLinearAssembleSymbolic(CeedOperator op, CeedSize *num_entries, CeedInt **rows, CeedInt **cols);
int  CeedOperatorLinearAssemble(CeedOperator op, CeedVector values);
int  CeedOperatorLinearAssembleGetNumEntries(CeedOperator op, CeedSize *num_entries);
int  CeedOperatorGetFlopsEstimate(CeedOperator op, CeedSize *flops);
int  CeedOperatorSetQFunctionAssemblyDataUpdateNeeded(CeedOperator op, bool needs_data_update);
int  CeedOperatorView(CeedOperator op, FILE *stream);
int  CeedOperatorViewTerse(CeedOperator op, FILE *stream);
int  CeedOperatorGetNumViewTabs(CeedOperator op, CeedInt *num_tabs);
int  CeedOperatorSetNumViewTabs(CeedOperator op, CeedInt num_tabs);
int  CeedOperatorGetCeed(CeedOperator op, Ceed *ceed);
Ceed CeedOperatorReturnCeed(CeedOperator op);
int  CeedOperatorDestroy(CeedOperator *op);

int  CeedFree(void *p);

#ifdef __cplusplus
}
#endif
