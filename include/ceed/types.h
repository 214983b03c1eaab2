/*
 * ceed/types.h -- scalar types, enums and QFunction source macros of the
 * libCEED C API, as used by the /gpu/hip/mi355 backend and by user QFunction
 * sources.  Names, signatures and enum values are the reference's
 * (CEED/libCEED include/ceed/types.h) so that user code and QFunction headers
 * compile unchanged; the text is this project's own.
 *
 * This header is also included by the runtime (hipRTC) compilation of
 * QFunction sources, where CEED_RUNNING_JIT_PASS is defined and the system
 * headers are not available.
 */
#ifndef CEED_QFUNCTION_DEFS_H
#define CEED_QFUNCTION_DEFS_H

#ifndef CEED_RUNNING_JIT_PASS
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#endif

/* ---- QFunction source macros ------------------------------------------- */
#ifndef CEED_QFUNCTION_ATTR
#if defined(__GNUC__) || defined(__clang__)
#define CEED_QFUNCTION_ATTR __attribute__((flatten))
#else
#define CEED_QFUNCTION_ATTR
#endif
#endif

#ifndef CEED_QFUNCTION_HELPER_ATTR
#if defined(__GNUC__) || defined(__clang__)
#define CEED_QFUNCTION_HELPER_ATTR CEED_QFUNCTION_ATTR __attribute__((always_inline))
#else
#define CEED_QFUNCTION_HELPER_ATTR CEED_QFUNCTION_ATTR
#endif
#endif

/* Defines the QFunction `name` and a string `name_loc` = "<file>:<name>" that
 * locates its source for runtime compilation. */
#ifndef CEED_QFUNCTION
#define CEED_QFUNCTION(name)                                        \
  static const char              name##_loc[] = __FILE__ ":" #name; \
  CEED_QFUNCTION_ATTR static int name
#endif

#ifndef CEED_QFUNCTION_HELPER
#define CEED_QFUNCTION_HELPER CEED_QFUNCTION_HELPER_ATTR static inline
#endif

/* Length of the quadrature-point axis in VLA casts of QFunction arguments. */
#ifndef CEED_Q_VLA
#define CEED_Q_VLA Q
#endif

#ifndef CeedPragmaSIMD
#if defined(__GNUC__) && !defined(__clang__) && __GNUC__ >= 5
#define CeedPragmaSIMD _Pragma("GCC ivdep")
#elif defined(__clang__)
#define CeedPragmaSIMD _Pragma("clang loop vectorize(enable)")
#else
#define CeedPragmaSIMD
#endif
#endif

/* ---- scalar types ------------------------------------------------------ */
typedef int32_t     CeedInt;
#define CeedInt_FMT "d"
typedef ptrdiff_t   CeedSize;
#define CeedSize_FMT "td"
typedef signed char CeedInt8;
#define CeedInt8_FMT "d"

typedef enum {
  CEED_SCALAR_FP32,
  CEED_SCALAR_FP64
} CeedScalarType;

/* CeedScalar is FP64 in this build; the FP32 operator is a backend mode. */
#define CEED_SCALAR_IS_FP64
#define CEED_SCALAR_TYPE CEED_SCALAR_FP64
typedef double CeedScalar;
#define CEED_EPSILON 1e-16

/* ---- enums (values are part of the ABI) -------------------------------- */
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
  CEED_MEM_HOST,
  CEED_MEM_DEVICE,
} CeedMemType;

typedef enum {
  CEED_COPY_VALUES,
  CEED_USE_POINTER,
  CEED_OWN_POINTER,
} CeedCopyMode;

typedef enum {
  CEED_NORM_1,
  CEED_NORM_2,
  CEED_NORM_MAX,
} CeedNormType;

typedef enum {
  CEED_NOTRANSPOSE,
  CEED_TRANSPOSE
} CeedTransposeMode;

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

typedef enum {
  CEED_TOPOLOGY_LINE     = 1 << 16 | 0,
  CEED_TOPOLOGY_TRIANGLE = 2 << 16 | 1,
  CEED_TOPOLOGY_QUAD     = 2 << 16 | 2,
  CEED_TOPOLOGY_TET      = 3 << 16 | 3,
  CEED_TOPOLOGY_PYRAMID  = 3 << 16 | 4,
  CEED_TOPOLOGY_PRISM    = 3 << 16 | 5,
  CEED_TOPOLOGY_HEX      = 3 << 16 | 6,
} CeedElemTopology;

typedef enum {
  CEED_CONTEXT_FIELD_DOUBLE = 1,
  CEED_CONTEXT_FIELD_INT32  = 2,
  CEED_CONTEXT_FIELD_BOOL   = 3,
} CeedContextFieldType;

#endif /* CEED_QFUNCTION_DEFS_H */
