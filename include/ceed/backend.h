// libCEED backend-header entry points that HONEE calls, implemented by the MI355X library libceed.so.
//
// HONEE includes <ceed/backend.h> in src/mat-ceed.c, problems/newtonian.c and problems/channel.c (among others) and
// calls the functions below; names, argument order and meaning follow libCEED's backend header so those call sites
// compile unchanged.  Everything in ceed.h is available through this header as well.
#pragma once

#include "../ceed.h"

#ifdef __cplusplus
extern "C" {
#endif

// Any libCEED object (Ceed, CeedVector, CeedElemRestriction, CeedBasis, CeedQFunction, CeedQFunctionContext,
// CeedOperator) converts to a CeedObject by a pointer cast: every object of this library starts with the Ceed it
// belongs to (a Ceed starts with itself).
typedef struct CeedObject_private *CeedObject;

// The Ceed an object belongs to, without taking a reference (src/misc.c:198, src/honee-wall-distance.c:31)
Ceed CeedObjectReturnCeed(CeedObject obj);

// Assembly data of a non-composite operator: its active input and output bases.  The data belongs to the operator
// and lives until the operator is destroyed; the caller does not destroy it (src/mat-ceed.c:478-479,494-495).
typedef struct CeedOperatorAssemblyData_private *CeedOperatorAssemblyData;

int CeedOperatorGetOperatorAssemblyData(CeedOperator op, CeedOperatorAssemblyData *data);
// Active bases of the operator's input and output fields (each distinct basis once, in field order).  Any output
// pointer may be NULL.  The assembled (non-tensor) basis matrices are not formed by this library: requesting them
// (assembled_bases_in/out non-NULL) returns CEED_ERROR_UNSUPPORTED.  HONEE passes NULL for them.
int CeedOperatorAssemblyDataGetBases(CeedOperatorAssemblyData data, CeedInt *num_active_bases_in, CeedBasis **active_bases_in,
                                     const CeedScalar ***assembled_bases_in, CeedInt *num_active_bases_out, CeedBasis **active_bases_out,
                                     const CeedScalar ***assembled_bases_out);

#ifdef __cplusplus
}
#endif
