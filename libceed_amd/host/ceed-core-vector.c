/*
 * CeedVector interface: argument checks, the access lock (state counter odd
 * while a write access is out, reader count for reads) and dispatch to the
 * backend slots, following interface/ceed-vector.c (:527-547, :630-637).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ceed-core.h"

int CeedVectorCreate(Ceed ceed, CeedSize len, CeedVector *vec) {
  CeedCheck(len >= 0, ceed, CEED_ERROR_DIMENSION, "a vector length cannot be negative");
  CeedCheck(ceed->VectorCreate, ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorCreate");
  CeedCall(CeedCalloc(1, vec));
  CeedCall(CeedReferenceCopy(ceed, &(*vec)->ceed));
  (*vec)->ref_count = 1;
  (*vec)->length    = len;
  CeedCall(ceed->VectorCreate(len, *vec));
  return CEED_ERROR_SUCCESS;
}

int CeedVectorReferenceCopy(CeedVector vec, CeedVector *vec_copy) {
  if (vec && !CeedCoreIsSpecialVector(vec)) vec->ref_count++;
  if (*vec_copy && !CeedCoreIsSpecialVector(*vec_copy)) CeedCall(CeedVectorDestroy(vec_copy));
  *vec_copy = vec;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorGetCeed(CeedVector vec, Ceed *ceed) {
  *ceed = NULL;
  return CeedReferenceCopy(vec->ceed, ceed);
}

int CeedVectorGetState(CeedVector vec, uint64_t *state) {
  *state = vec->state;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorGetData(CeedVector vec, void *data) {
  *(void **)data = vec->data;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorSetData(CeedVector vec, void *data) {
  vec->data = data;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorGetLength(CeedVector vec, CeedSize *length) {
  *length = vec->length;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorHasValidArray(CeedVector vec, bool *has_valid_array) {
  CeedCheck(vec->HasValidArray, vec->ceed, CEED_ERROR_UNSUPPORTED, "backend does not report whether a CeedVector holds data");
  if (vec->length == 0) {
    *has_valid_array = true;
    return CEED_ERROR_SUCCESS;
  }
  return vec->HasValidArray(vec, has_valid_array);
}

int CeedVectorHasBorrowedArrayOfType(CeedVector vec, CeedMemType mem_type, bool *has_borrowed_array_of_type) {
  CeedCheck(vec->HasBorrowedArrayOfType, vec->ceed, CEED_ERROR_UNSUPPORTED, "backend does not report borrowed CeedVector arrays");
  return vec->HasBorrowedArrayOfType(vec, mem_type, has_borrowed_array_of_type);
}

/* reads need data: set by CeedVectorSetArray, CeedVectorSetValue or a write access */
static int check_valid(CeedVector vec, const char *what) {
  bool valid;
  CeedCall(CeedVectorHasValidArray(vec, &valid));
  CeedCheck(valid, vec->ceed, CEED_ERROR_BACKEND, "%s: the CeedVector holds no data yet (set it with CeedVectorSetValue or CeedVectorSetArray)", what);
  return CEED_ERROR_SUCCESS;
}

static int check_unlocked(CeedVector vec) {
  CeedCheck(vec->state % 2 == 0, vec->ceed, CEED_ERROR_ACCESS, "CeedVector access lock in use: a write access is out");
  CeedCheck(vec->num_readers == 0, vec->ceed, CEED_ERROR_ACCESS, "CeedVector access lock in use: a read access is out");
  return CEED_ERROR_SUCCESS;
}

int CeedVectorSetArray(CeedVector vec, CeedMemType mem_type, CeedCopyMode copy_mode, CeedScalar *array) {
  CeedCheck(vec->SetArray, vec->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorSetArray");
  CeedCall(check_unlocked(vec));
  if (vec->length > 0) CeedCall(vec->SetArray(vec, mem_type, copy_mode, array));
  vec->state += 2;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorSetValue(CeedVector vec, CeedScalar value) {
  CeedCall(check_unlocked(vec));
  if (vec->SetValue) {
    if (vec->length > 0) CeedCall(vec->SetValue(vec, value));
    vec->state += 2;
  } else {
    CeedScalar *a;
    CeedCall(CeedVectorGetArrayWrite(vec, CEED_MEM_HOST, &a));
    for (CeedSize i = 0; i < vec->length; i++) a[i] = value;
    CeedCall(CeedVectorRestoreArray(vec, &a));
  }
  return CEED_ERROR_SUCCESS;
}

int CeedVectorSyncArray(CeedVector vec, CeedMemType mem_type) {
  CeedCheck(vec->state % 2 == 0, vec->ceed, CEED_ERROR_ACCESS, "cannot sync a vector with a write access out");
  if (vec->SyncArray && vec->length > 0) CeedCall(vec->SyncArray(vec, mem_type));
  return CEED_ERROR_SUCCESS;
}

int CeedVectorTakeArray(CeedVector vec, CeedMemType mem_type, CeedScalar **array) {
  CeedCall(check_unlocked(vec));
  CeedCheck(vec->TakeArray, vec->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorTakeArray");
  *array = NULL;
  if (vec->length > 0) {
    bool borrowed;
    CeedCall(CeedVectorHasBorrowedArrayOfType(vec, mem_type, &borrowed));
    CeedCheck(borrowed, vec->ceed, CEED_ERROR_BACKEND, "CeedVectorTakeArray: no array of this memory type was lent with CEED_USE_POINTER");
    CeedCall(check_valid(vec, "CeedVectorTakeArray"));
    CeedCall(vec->TakeArray(vec, mem_type, array));
  }
  return CEED_ERROR_SUCCESS;
}

int CeedVectorGetArray(CeedVector vec, CeedMemType mem_type, CeedScalar **array) {
  CeedCheck(vec->GetArray, vec->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorGetArray");
  CeedCall(check_unlocked(vec));
  *array = NULL;
  if (vec->length > 0) {
    CeedCall(check_valid(vec, "CeedVectorGetArray"));
    CeedCall(vec->GetArray(vec, mem_type, array));
  }
  vec->state++;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorGetArrayRead(CeedVector vec, CeedMemType mem_type, const CeedScalar **array) {
  CeedCheck(vec->GetArrayRead, vec->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorGetArrayRead");
  CeedCheck(vec->state % 2 == 0, vec->ceed, CEED_ERROR_ACCESS, "cannot grant read access while a write access is out");
  *array = NULL;
  if (vec->length > 0) {
    CeedCall(check_valid(vec, "CeedVectorGetArrayRead"));
    CeedCall(vec->GetArrayRead(vec, mem_type, array));
  }
  vec->num_readers++;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorGetArrayWrite(CeedVector vec, CeedMemType mem_type, CeedScalar **array) {
  CeedCheck(vec->GetArrayWrite, vec->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorGetArrayWrite");
  CeedCall(check_unlocked(vec));
  *array = NULL;
  if (vec->length > 0) CeedCall(vec->GetArrayWrite(vec, mem_type, array));
  vec->state++;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorRestoreArray(CeedVector vec, CeedScalar **array) {
  CeedCheck(vec->state % 2 == 1, vec->ceed, CEED_ERROR_ACCESS, "cannot restore a CeedVector array that was not granted");
  if (vec->RestoreArray && vec->length > 0) CeedCall(vec->RestoreArray(vec));
  *array = NULL;
  vec->state++;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorRestoreArrayRead(CeedVector vec, const CeedScalar **array) {
  CeedCheck(vec->num_readers > 0, vec->ceed, CEED_ERROR_ACCESS, "cannot restore a read access that was not granted");
  vec->num_readers--;
  if (vec->RestoreArrayRead && vec->length > 0 && vec->num_readers == 0) CeedCall(vec->RestoreArrayRead(vec));
  *array = NULL;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorCopy(CeedVector vec, CeedVector vec_copy) {
  CeedCheck(vec->length == vec_copy->length, vec->ceed, CEED_ERROR_INCOMPATIBLE, "vector lengths differ");
  CeedMemType       mt;
  const CeedScalar *a;
  CeedScalar       *b;
  CeedCall(CeedGetPreferredMemType(vec->ceed, &mt));
  CeedCall(CeedVectorGetArrayRead(vec, mt, &a));
  CeedCall(CeedVectorSetArray(vec_copy, mt, CEED_COPY_VALUES, (CeedScalar *)a));
  CeedCall(CeedVectorRestoreArrayRead(vec, &a));
  (void)b;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorNorm(CeedVector vec, CeedNormType type, CeedScalar *norm) {
  CeedCheck(vec->state % 2 == 0, vec->ceed, CEED_ERROR_ACCESS, "cannot compute a norm while a write access is out");
  *norm = 0.;
  if (vec->length == 0) return CEED_ERROR_SUCCESS;
  CeedCall(check_valid(vec, "CeedVectorNorm"));
  if (vec->Norm) return vec->Norm(vec, type, norm);
  const CeedScalar *a;
  CeedCall(CeedVectorGetArrayRead(vec, CEED_MEM_HOST, &a));
  for (CeedSize i = 0; i < vec->length; i++) {
    if (type == CEED_NORM_1) *norm += fabs(a[i]);
    else if (type == CEED_NORM_2) *norm += a[i] * a[i];
    else *norm = fmax(*norm, fabs(a[i]));
  }
  if (type == CEED_NORM_2) *norm = sqrt(*norm);
  CeedCall(CeedVectorRestoreArrayRead(vec, &a));
  return CEED_ERROR_SUCCESS;
}

int CeedVectorScale(CeedVector x, CeedScalar alpha) {
  CeedCall(check_unlocked(x));
  CeedCheck(x->Scale, x->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorScale");
  if (x->length > 0) {
    CeedCall(check_valid(x, "CeedVectorScale"));
    CeedCall(x->Scale(x, alpha));
  }
  x->state += 2;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorAXPY(CeedVector y, CeedScalar alpha, CeedVector x) {
  CeedCheck(y->length == x->length, y->ceed, CEED_ERROR_UNSUPPORTED, "cannot add vectors of different lengths (%td, %td)", x->length, y->length);
  CeedCheck(x != y, y->ceed, CEED_ERROR_UNSUPPORTED, "cannot use same vector for x and y in CeedVectorAXPY");
  CeedCall(check_unlocked(y));
  CeedCheck(y->AXPY, y->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorAXPY");
  if (y->length > 0) {
    CeedCall(check_valid(x, "CeedVectorAXPY (x)"));
    CeedCall(check_valid(y, "CeedVectorAXPY (y)"));
    CeedCall(y->AXPY(y, alpha, x));
  }
  y->state += 2;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorAXPBY(CeedVector y, CeedScalar alpha, CeedScalar beta, CeedVector x) {
  CeedCheck(y->length == x->length, y->ceed, CEED_ERROR_UNSUPPORTED, "cannot add vectors of different lengths (%td, %td)", x->length, y->length);
  CeedCheck(x != y, y->ceed, CEED_ERROR_UNSUPPORTED, "cannot use same vector for x and y in CeedVectorAXPBY");
  CeedCall(check_unlocked(y));
  CeedCheck(y->AXPBY, y->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorAXPBY");
  if (y->length > 0) {
    CeedCall(check_valid(x, "CeedVectorAXPBY (x)"));
    CeedCall(check_valid(y, "CeedVectorAXPBY (y)"));
    CeedCall(y->AXPBY(y, alpha, beta, x));
  }
  y->state += 2;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorPointwiseMult(CeedVector w, CeedVector x, CeedVector y) {
  CeedCheck(w->length == x->length && w->length == y->length, w->ceed, CEED_ERROR_UNSUPPORTED, "vector lengths differ");
  CeedCall(check_unlocked(w));
  CeedCheck(w->PointwiseMult, w->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorPointwiseMult");
  if (w->length > 0) {
    CeedCall(check_valid(x, "CeedVectorPointwiseMult (x)"));
    CeedCall(check_valid(y, "CeedVectorPointwiseMult (y)"));
    CeedCall(w->PointwiseMult(w, x, y));
  }
  w->state += 2;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorReciprocal(CeedVector vec) {
  CeedCall(check_unlocked(vec));
  CeedCheck(vec->Reciprocal, vec->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support VectorReciprocal");
  if (vec->length > 0) {
    CeedCall(check_valid(vec, "CeedVectorReciprocal"));
    CeedCall(vec->Reciprocal(vec));
  }
  vec->state += 2;
  return CEED_ERROR_SUCCESS;
}

int CeedVectorView(CeedVector vec, const char *fp_fmt, FILE *stream) {
  const CeedScalar *a;
  char              fmt[64];
  snprintf(fmt, sizeof(fmt), "%s\n", fp_fmt ? fp_fmt : "\t%g");
  fprintf(stream, "CeedVector length %td\n", vec->length);
  CeedCall(CeedVectorGetArrayRead(vec, CEED_MEM_HOST, &a));
  for (CeedSize i = 0; i < vec->length; i++) fprintf(stream, fmt, a[i]);
  CeedCall(CeedVectorRestoreArrayRead(vec, &a));
  return CEED_ERROR_SUCCESS;
}

int CeedVectorDestroy(CeedVector *vec) {
  if (!*vec || CeedCoreIsSpecialVector(*vec) || --(*vec)->ref_count > 0) {
    *vec = NULL;
    return CEED_ERROR_SUCCESS;
  }
  CeedCheck((*vec)->state % 2 == 0, (*vec)->ceed, CEED_ERROR_ACCESS, "cannot destroy a CeedVector with a write access out");
  CeedCheck((*vec)->num_readers == 0, (*vec)->ceed, CEED_ERROR_ACCESS, "cannot destroy a CeedVector with a read access out");
  if ((*vec)->Destroy) CeedCall((*vec)->Destroy(*vec));
  CeedCall(CeedDestroy(&(*vec)->ceed));
  free(*vec);
  *vec = NULL;
  return CEED_ERROR_SUCCESS;
}
