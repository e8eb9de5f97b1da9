// Scalar/QFunction definitions needed by QFunction source headers (libCEED include/ceed/types.h).
//
// HONEE's qfunctions/*.h include <ceed/types.h>; this header lets them compile unchanged against this library.
// Names and values follow libCEED's public header so the ABI matches.
#pragma once

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

typedef int32_t CeedInt;
typedef ptrdiff_t CeedSize;
// printf conversions of CeedInt and CeedSize (src/petsc_ops.c:69, src/misc.c:398, src/mat-ceed.c:1497)
#define CeedInt_FMT "d"
#define CeedSize_FMT "td"
typedef int8_t  CeedInt8;
typedef double  CeedScalar;
#define CEED_SCALAR_TYPE CEED_SCALAR_FP64
#define CEED_EPSILON 2.220446049250313e-16

#ifndef CEED_QFUNCTION_ATTR
#define CEED_QFUNCTION_ATTR
#endif
#ifndef CEED_QFUNCTION_HELPER_ATTR
#define CEED_QFUNCTION_HELPER_ATTR inline
#endif
#define CEED_QFUNCTION_HELPER CEED_QFUNCTION_HELPER_ATTR static

// A QFunction body plus its "<source file>:<name>" location string, which the MI355X backend uses to identify it.
#ifndef CEED_QFUNCTION
#define CEED_QFUNCTION(name)                                  \
  static const char name##_loc[] = __FILE__ ":" #name;        \
  CEED_QFUNCTION_ATTR static int name
#endif

#ifndef CeedPragmaSIMD
#define CeedPragmaSIMD
#endif
#define CEED_Q_VLA Q

typedef enum {
  CEED_ERROR_SUCCESS      = 0,
  CEED_ERROR_MINOR        = 1,
  CEED_ERROR_DIMENSION    = 2,
  CEED_ERROR_INCOMPLETE   = 3,
  CEED_ERROR_INCOMPATIBLE = 4,
  CEED_ERROR_ACCESS       = 5,
  CEED_ERROR_MAJOR        = -1,
  CEED_ERROR_BACKEND      = -2,
  CEED_ERROR_UNSUPPORTED  = -3,
} CeedErrorType;

typedef enum {
  CEED_NOTRANSPOSE,
  CEED_TRANSPOSE,
} CeedTransposeMode;

typedef enum {
  CEED_SCALAR_FP32,
  CEED_SCALAR_FP64,
} CeedScalarType;
