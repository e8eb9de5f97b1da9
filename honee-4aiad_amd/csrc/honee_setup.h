// PETSc-free Newtonian problem driver over the libCEED API (mirrors HONEE's SetupLibceed wiring); used by the tests,
// the benchmark and the time integrator of this package.
#pragma once

#include <hip/hip_runtime_api.h>

#include "ceed.h"
#include "honee_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

// Mesh and 1D tabulation supplied in place of DMPlex/PetscFE.
typedef struct {
  int           nelem, P, Q, nnodes, nxnodes;
  const int    *offsets;   // [nelem][P^3] node index of each element node (x fastest)
  const int    *xoffsets;  // [nelem][8] coordinate node index
  const double *xcoord;    // [nxnodes][3]
  const double *B, *D;     // [Q][P]
  const double *Bx, *Dx;   // [Q][2]
  const double *qref, *qweight;  // [Q]
} HoneeMiMesh;

typedef struct HoneeMiProblem_ *HoneeMiProblem;

int          HoneeMiProblemCreate(const char *resource, const HoneeMiMesh *mesh, const CeedMiNewtonianContext *ctx, int state_var,
                                  HoneeMiProblem *problem);
int          HoneeMiProblemDestroy(HoneeMiProblem *problem);
Ceed         HoneeMiProblemGetCeed(HoneeMiProblem p);
CeedOperator HoneeMiProblemGetOperator(HoneeMiProblem p, int which);  // 0: IFunction composite, 1: IJacobian composite
int          HoneeMiApplyIFunction(HoneeMiProblem p, int mem_type, double time, double dt, const double *q, const double *q_dot, double *g);
int          HoneeMiApplyIJacobian(HoneeMiProblem p, int mem_type, double shift, const double *dq, double *v);
int          HoneeMiIJacobianDiagonal(HoneeMiProblem p, int mem_type, double shift, double *diag);
int          HoneeMiGetQData(HoneeMiProblem p, double *out);
int          HoneeMiGetJacData(HoneeMiProblem p, double *out);
int          HoneeMiGetJacData32(HoneeMiProblem p, float *out);
const char  *HoneeMiProblemError(HoneeMiProblem p);
const char  *HoneeMiLastError(void);

#ifdef __cplusplus
}
#endif
