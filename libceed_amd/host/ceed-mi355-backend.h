/*
 * Private declarations of the /gpu/hip/mi355 backend: per-object backend data
 * and the helpers shared by its object and operator implementations.
 */
#ifndef CEED_MI355_BACKEND_H
#define CEED_MI355_BACKEND_H

#include <ceed.h>
#include <ceed/backend.h>
#include <ceed-mi355.h>

typedef struct {
  CeedMi355Restriction handle;
  CeedRestrictionType  type;
  CeedInt             *offsets, *offsets_owned; /* host offsets (standard restrictions) */
} RestrictionData;

typedef struct {
  CeedMi355Basis handle;
} BasisData;

typedef struct {
  CeedMi355QFunction kernel; /* compiled on first apply */
} QFunctionData;

typedef struct {
  void    *data, *owned;   /* the context bytes (host memory), and the allocation the backend frees */
  void    *device;         /* device copy of the bytes */
  size_t   device_size;
  uint64_t device_state;   /* context state the copy was taken at */
  bool     has_device_copy;
} ContextData;

/* Turn a kernel-library status into a CeedError carrying its message. */
CEED_INTERN int CeedMi355Check(Ceed ceed, int ierr);
/* Device array of a vector, synced to the device; write marks the host copy stale. */
CEED_INTERN int CeedMi355VectorDevice(CeedVector vec, bool write, CeedScalar **array);
CEED_INTERN int CeedMi355QFunctionKernel(CeedQFunction qf, CeedMi355QFunction *kernel);
CEED_INTERN int CeedMi355QFunctionContextDevice(CeedQFunction qf, void **ctx_device);
CEED_INTERN int CeedOperatorCreate_Mi355(CeedOperator op);

#endif /* CEED_MI355_BACKEND_H */
