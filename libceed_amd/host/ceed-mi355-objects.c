/*
 * /gpu/hip/mi355 backend: registration, init, and the CeedVector,
 * CeedElemRestriction, CeedBasis, CeedQFunction(Context) implementations.
 * Written against the libCEED backend API only (ceed/backend.h); every device
 * operation goes through the C ABI of ceed-mi355-kernels.h.
 *
 * Vectors keep a host and a device copy with validity flags, as the
 * reference HIP backend does (backends/hip-ref/ceed-hip-ref-vector.c:41-95):
 * a read access syncs the requested side, a write access invalidates the other.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <ceed.h>
#include <ceed/backend.h>

#include "ceed-mi355-backend.h"

/* ------------------------------------------------------------------------ */
/* device helpers                                                            */
/* ------------------------------------------------------------------------ */
int CeedMi355Check(Ceed ceed, int ierr) {
  if (ierr == CEED_MI355_SUCCESS) return CEED_ERROR_SUCCESS;
  return CeedError(ceed, CEED_ERROR_BACKEND, "MI355X kernel library: %s", CeedMi355GetLastError());
}

static int dev_alloc(Ceed ceed, size_t bytes, void *ptr) { return CeedMi355Check(ceed, CeedMi355Malloc(bytes, (void **)ptr)); }

static int copy_sync(Ceed ceed, void *dst, const void *src, size_t bytes) {
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355Memcpy(dst, src, bytes, NULL)));
  return CeedMi355Check(ceed, CeedMi355StreamSynchronize(NULL));
}

/* ------------------------------------------------------------------------ */
/* vector                                                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
  CeedScalar *h_array, *h_owned; /* current host array (owned or borrowed) */
  CeedScalar *d_array, *d_owned; /* current device array */
  bool        h_valid, d_valid;
} VectorData;

static int vec_length_bytes(CeedVector vec, size_t *bytes) {
  CeedSize n;
  CeedCallBackend(CeedVectorGetLength(vec, &n));
  *bytes = sizeof(CeedScalar) * (size_t)n;
  return CEED_ERROR_SUCCESS;
}

static int vec_ensure(CeedVector vec, VectorData *d, CeedMemType mt) {
  Ceed   ceed;
  size_t bytes;
  CeedCallBackend(CeedVectorGetCeed(vec, &ceed));
  CeedCallBackend(vec_length_bytes(vec, &bytes));
  if (mt == CEED_MEM_HOST && !d->h_array) {
    d->h_owned = calloc(1, bytes);
    CeedCheck(d->h_owned, ceed, CEED_ERROR_MAJOR, "host allocation failed");
    d->h_array = d->h_owned;
  }
  if (mt == CEED_MEM_DEVICE && !d->d_array) {
    CeedCallBackend(dev_alloc(ceed, bytes, &d->d_owned));
    d->d_array = d->d_owned;
  }
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* make the `mt` copy valid */
static int vec_sync(CeedVector vec, VectorData *d, CeedMemType mt) {
  Ceed   ceed;
  size_t bytes;
  CeedCallBackend(vec_ensure(vec, d, mt));
  if ((mt == CEED_MEM_HOST && d->h_valid) || (mt == CEED_MEM_DEVICE && d->d_valid)) return CEED_ERROR_SUCCESS;
  CeedCallBackend(CeedVectorGetCeed(vec, &ceed));
  CeedCallBackend(vec_length_bytes(vec, &bytes));
  if (mt == CEED_MEM_HOST) {
    if (d->d_valid) CeedCallBackend(copy_sync(ceed, d->h_array, d->d_array, bytes));
    else memset(d->h_array, 0, bytes);
    d->h_valid = true;
  } else {
    if (d->h_valid) CeedCallBackend(copy_sync(ceed, d->d_array, d->h_array, bytes));
    else CeedCallBackend(CeedMi355Check(ceed, CeedMi355VectorOp(CEED_MI355_VEC_SET, CEED_MI355_FP64, (int64_t)(bytes / 8), 0.0, 0.0, NULL, NULL,
                                                                  d->d_array, NULL)));
    d->d_valid = true;
  }
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int VectorSetArray(CeedVector vec, CeedMemType mt, CeedCopyMode cm, CeedScalar *array) {
  VectorData *d;
  Ceed        ceed;
  size_t      bytes;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  CeedCallBackend(CeedVectorGetCeed(vec, &ceed));
  CeedCallBackend(vec_length_bytes(vec, &bytes));
  if (mt == CEED_MEM_HOST) {
    if (cm == CEED_COPY_VALUES) {
      if (!d->h_owned) d->h_owned = calloc(1, bytes);
      d->h_array = d->h_owned;
      if (array) memcpy(d->h_array, array, bytes);
    } else {
      if (d->h_owned && d->h_owned != array) free(d->h_owned);
      d->h_owned = cm == CEED_OWN_POINTER ? array : NULL;
      d->h_array = array;
    }
    d->h_valid = array != NULL;
    d->d_valid = false;
  } else {
    if (cm == CEED_COPY_VALUES) {
      if (!d->d_owned) CeedCallBackend(dev_alloc(ceed, bytes, &d->d_owned));
      d->d_array = d->d_owned;
      if (array) CeedCallBackend(copy_sync(ceed, d->d_array, array, bytes));
    } else {
      if (d->d_owned && d->d_owned != array) CeedMi355Free(d->d_owned);
      d->d_owned = cm == CEED_OWN_POINTER ? array : NULL;
      d->d_array = array;
    }
    d->d_valid = array != NULL;
    d->h_valid = false;
  }
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int VectorSyncArray(CeedVector vec, CeedMemType mt) {
  VectorData *d;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  return vec_sync(vec, d, mt);
}

static int VectorTakeArray(CeedVector vec, CeedMemType mt, CeedScalar **array) {
  VectorData *d;
  Ceed        ceed;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  CeedCallBackend(CeedVectorGetCeed(vec, &ceed));
  CeedCallBackend(vec_sync(vec, d, mt));
  if (mt == CEED_MEM_HOST) {
    CeedCheck(d->h_owned != d->h_array, ceed, CEED_ERROR_BACKEND, "CeedVectorTakeArray needs an array set with CEED_USE_POINTER");
    *array     = d->h_array;
    d->h_array = NULL;
    d->h_valid = false;
  } else {
    CeedCheck(d->d_owned != d->d_array, ceed, CEED_ERROR_BACKEND, "CeedVectorTakeArray needs an array set with CEED_USE_POINTER");
    *array     = d->d_array;
    d->d_array = NULL;
    d->d_valid = false;
  }
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int VectorGetArrayRead(CeedVector vec, CeedMemType mt, const CeedScalar **array) {
  VectorData *d;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  CeedCallBackend(vec_sync(vec, d, mt));
  *array = mt == CEED_MEM_HOST ? d->h_array : d->d_array;
  return CEED_ERROR_SUCCESS;
}

static int VectorGetArray(CeedVector vec, CeedMemType mt, CeedScalar **array) {
  VectorData *d;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  CeedCallBackend(vec_sync(vec, d, mt));
  if (mt == CEED_MEM_HOST) d->d_valid = false;
  else d->h_valid = false;
  *array = mt == CEED_MEM_HOST ? d->h_array : d->d_array;
  return CEED_ERROR_SUCCESS;
}

static int VectorGetArrayWrite(CeedVector vec, CeedMemType mt, CeedScalar **array) {
  VectorData *d;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  CeedCallBackend(vec_ensure(vec, d, mt));
  if (mt == CEED_MEM_HOST) d->h_valid = true, d->d_valid = false;
  else d->d_valid = true, d->h_valid = false;
  *array = mt == CEED_MEM_HOST ? d->h_array : d->d_array;
  return CEED_ERROR_SUCCESS;
}

static int VectorRestoreArray(CeedVector vec) {
  (void)vec;
  return CEED_ERROR_SUCCESS;
}

/* valid data on either side (interface/ceed-vector.c:89 asks before every read) */
static int VectorHasValidArray(CeedVector vec, bool *has_valid_array) {
  VectorData *d;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  *has_valid_array = d->h_valid || d->d_valid;
  return CEED_ERROR_SUCCESS;
}

/* an array of the given memory type set with CEED_USE_POINTER (CeedVectorTakeArray's precondition) */
static int VectorHasBorrowedArrayOfType(CeedVector vec, CeedMemType mt, bool *has_borrowed) {
  VectorData *d;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  if (mt == CEED_MEM_HOST) *has_borrowed = d->h_array && d->h_array != d->h_owned;
  else *has_borrowed = d->d_array && d->d_array != d->d_owned;
  return CEED_ERROR_SUCCESS;
}

int CeedMi355VectorDevice(CeedVector vec, bool write, CeedScalar **array) {
  VectorData *d;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  CeedCallBackend(vec_sync(vec, d, CEED_MEM_DEVICE));
  if (write) d->h_valid = false;
  *array = d->d_array;
  return CEED_ERROR_SUCCESS;
}

static int VectorSetValue(CeedVector vec, CeedScalar value) {
  VectorData *d;
  Ceed        ceed;
  CeedSize    n;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  CeedCallBackend(CeedVectorGetCeed(vec, &ceed));
  CeedCallBackend(CeedVectorGetLength(vec, &n));
  CeedCallBackend(vec_ensure(vec, d, CEED_MEM_DEVICE));
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355VectorOp(CEED_MI355_VEC_SET, CEED_MI355_FP64, n, value, 0, NULL, NULL, d->d_array, NULL)));
  d->d_valid = true;
  d->h_valid = false;
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* y := op(x, y) on the device */
static int vector_binary(CeedVector y, int op, double a, double b, CeedVector x, CeedVector z) {
  Ceed        ceed;
  CeedSize    n;
  CeedScalar *dy, *dx = NULL, *dz = NULL;
  CeedCallBackend(CeedVectorGetCeed(y, &ceed));
  CeedCallBackend(CeedVectorGetLength(y, &n));
  if (x) CeedCallBackend(CeedMi355VectorDevice(x, false, &dx));
  if (z) CeedCallBackend(CeedMi355VectorDevice(z, false, &dz));
  CeedCallBackend(CeedMi355VectorDevice(y, true, &dy));
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355VectorOp(op, CEED_MI355_FP64, n, a, b, dx, dz, dy, NULL)));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int VectorScale(CeedVector x, CeedScalar alpha) { return vector_binary(x, CEED_MI355_VEC_SCALE, alpha, 0, NULL, NULL); }
static int VectorAXPY(CeedVector y, CeedScalar alpha, CeedVector x) { return vector_binary(y, CEED_MI355_VEC_AXPY, alpha, 0, x, NULL); }
static int VectorAXPBY(CeedVector y, CeedScalar alpha, CeedScalar beta, CeedVector x) {
  return vector_binary(y, CEED_MI355_VEC_AXPBY, alpha, beta, x, NULL);
}
static int VectorPointwiseMult(CeedVector w, CeedVector x, CeedVector y) { return vector_binary(w, CEED_MI355_VEC_PMULT, 0, 0, x, y); }
static int VectorReciprocal(CeedVector x) { return vector_binary(x, CEED_MI355_VEC_RECIPROCAL, CEED_EPSILON, 0, NULL, NULL); }

static int VectorNorm(CeedVector vec, CeedNormType type, CeedScalar *norm) {
  Ceed        ceed;
  CeedSize    n;
  CeedScalar *dx;
  double     *d_result;
  CeedCallBackend(CeedVectorGetCeed(vec, &ceed));
  CeedCallBackend(CeedVectorGetLength(vec, &n));
  CeedCallBackend(CeedMi355VectorDevice(vec, false, &dx));
  CeedCallBackend(dev_alloc(ceed, sizeof(double), &d_result));
  const int op = type == CEED_NORM_1 ? CEED_MI355_RED_ABS_SUM : (type == CEED_NORM_2 ? CEED_MI355_RED_DOT : CEED_MI355_RED_MAX_ABS);
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355VectorReduce(op, CEED_MI355_FP64, n, dx, dx, d_result, NULL)));
  CeedCallBackend(copy_sync(ceed, norm, d_result, sizeof(double)));
  CeedMi355Free(d_result);
  if (type == CEED_NORM_2) *norm = sqrt(*norm);
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int VectorDestroy(CeedVector vec) {
  VectorData *d;
  CeedCallBackend(CeedVectorGetData(vec, &d));
  free(d->h_owned);
  CeedMi355Free(d->d_owned);
  free(d);
  return CEED_ERROR_SUCCESS;
}

static int VectorCreate(CeedSize n, CeedVector vec) {
  Ceed        ceed;
  VectorData *d = calloc(1, sizeof(VectorData));
  (void)n;
  CeedCallBackend(CeedVectorGetCeed(vec, &ceed));
  CeedCallBackend(CeedVectorSetData(vec, d));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "HasValidArray", VectorHasValidArray));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "HasBorrowedArrayOfType", VectorHasBorrowedArrayOfType));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "SetArray", VectorSetArray));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "TakeArray", VectorTakeArray));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "SyncArray", VectorSyncArray));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "SetValue", VectorSetValue));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "GetArray", VectorGetArray));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "GetArrayRead", VectorGetArrayRead));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "GetArrayWrite", VectorGetArrayWrite));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "RestoreArray", VectorRestoreArray));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "RestoreArrayRead", VectorRestoreArray));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "Norm", VectorNorm));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "Scale", VectorScale));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "AXPY", VectorAXPY));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "AXPBY", VectorAXPBY));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "PointwiseMult", VectorPointwiseMult));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "Reciprocal", VectorReciprocal));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Vector", vec, "Destroy", VectorDestroy));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* element restriction                                                       */
/* ------------------------------------------------------------------------ */
static int RestrictionApply(CeedElemRestriction rstr, CeedTransposeMode t_mode, CeedVector u, CeedVector v, CeedRequest *request) {
  RestrictionData *d;
  Ceed             ceed;
  CeedScalar      *du, *dv;
  (void)request;
  CeedCallBackend(CeedElemRestrictionGetData(rstr, &d));
  CeedCallBackend(CeedElemRestrictionGetCeed(rstr, &ceed));
  CeedCallBackend(CeedMi355VectorDevice(u, false, &du));
  if (t_mode == CEED_NOTRANSPOSE) {
    CeedCallBackend(CeedVectorGetArrayWrite(v, CEED_MEM_DEVICE, &dv));
  } else {
    CeedCallBackend(CeedVectorGetArray(v, CEED_MEM_DEVICE, &dv));
  }
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355RestrictionApply(d->handle, t_mode, du, dv, NULL)));
  CeedCallBackend(CeedVectorRestoreArray(v, &dv));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* the offsets as created, on the host (the device copy lives inside the kernel
 * library's restriction and is not handed out) */
static int RestrictionGetOffsets(CeedElemRestriction rstr, CeedMemType mt, const CeedInt **offsets) {
  RestrictionData *d;
  Ceed             ceed;
  CeedCallBackend(CeedElemRestrictionGetData(rstr, &d));
  CeedCallBackend(CeedElemRestrictionGetCeed(rstr, &ceed));
  CeedCheck(d->offsets, ceed, CEED_ERROR_BACKEND, "a strided restriction has no offsets");
  CeedCheck(mt == CEED_MEM_HOST, ceed, CEED_ERROR_UNSUPPORTED, "/gpu/hip/mi355 provides restriction offsets in host memory only");
  *offsets = d->offsets;
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int RestrictionDestroy(CeedElemRestriction rstr) {
  RestrictionData *d;
  CeedCallBackend(CeedElemRestrictionGetData(rstr, &d));
  CeedMi355RestrictionDestroy(&d->handle);
  free(d->offsets_owned);
  free(d);
  return CEED_ERROR_SUCCESS;
}

static int ElemRestrictionCreate(CeedMemType mt, CeedCopyMode cm, const CeedInt *offsets, const bool *orients, const CeedInt8 *curl_orients,
                                 CeedElemRestriction rstr) {
  Ceed                ceed;
  CeedInt             num_elem, elem_size, num_comp, comp_stride = 1;
  CeedSize            l_size;
  CeedRestrictionType type;
  CeedCallBackend(CeedElemRestrictionGetCeed(rstr, &ceed));
  CeedCheck(mt == CEED_MEM_HOST, ceed, CEED_ERROR_BACKEND, "offsets must be host memory");
  CeedCheck(!orients && !curl_orients, ceed, CEED_ERROR_UNSUPPORTED, "oriented restrictions are not supported by /gpu/hip/mi355");
  CeedCallBackend(CeedElemRestrictionGetType(rstr, &type));
  CeedCallBackend(CeedElemRestrictionGetNumElements(rstr, &num_elem));
  CeedCallBackend(CeedElemRestrictionGetElementSize(rstr, &elem_size));
  CeedCallBackend(CeedElemRestrictionGetNumComponents(rstr, &num_comp));
  CeedCallBackend(CeedElemRestrictionGetLVectorSize(rstr, &l_size));
  RestrictionData *d = calloc(1, sizeof(RestrictionData));
  d->type            = type;
  if (type == CEED_RESTRICTION_STRIDED) {
    CeedInt strides[3];
    bool    backend_strides;
    CeedCallBackend(CeedElemRestrictionHasBackendStrides(rstr, &backend_strides));
    if (backend_strides) { /* CEED_STRIDES_BACKEND: this backend's E-vector layout, [elem][comp][node] */
      strides[0] = 1;
      strides[1] = elem_size;
      strides[2] = elem_size * num_comp;
    } else {
      CeedCallBackend(CeedElemRestrictionGetStrides(rstr, strides));
    }
    CeedCallBackend(CeedMi355Check(ceed, CeedMi355RestrictionCreateStrided(num_elem, elem_size, num_comp, l_size, strides, &d->handle)));
  } else {
    CeedCheck(type == CEED_RESTRICTION_STANDARD, ceed, CEED_ERROR_UNSUPPORTED, "restriction type not supported by /gpu/hip/mi355");
    CeedCallBackend(CeedElemRestrictionGetCompStride(rstr, &comp_stride));
    CeedCallBackend(CeedMi355Check(ceed, CeedMi355RestrictionCreateOffset(num_elem, elem_size, num_comp, comp_stride, l_size, offsets, &d->handle)));
    /* host offsets for CeedElemRestrictionGetOffsets: the caller's array when
     * it is handed over (OWN) or lent (USE), else a copy */
    const size_t n = (size_t)num_elem * elem_size;
    if (cm == CEED_COPY_VALUES) {
      d->offsets_owned = malloc(sizeof(CeedInt) * (n ? n : 1));
      CeedCheck(d->offsets_owned, ceed, CEED_ERROR_MAJOR, "host allocation failed");
      if (n) memcpy(d->offsets_owned, offsets, sizeof(CeedInt) * n);
      d->offsets = d->offsets_owned;
    } else {
      d->offsets       = (CeedInt *)offsets;
      d->offsets_owned = cm == CEED_OWN_POINTER ? (CeedInt *)offsets : NULL;
    }
  }
  {
    CeedInt layout[3] = {1, elem_size, elem_size * num_comp}; /* E-vectors are [elem][comp][node] */
    CeedCallBackend(CeedElemRestrictionSetELayout(rstr, layout));
  }
  CeedCallBackend(CeedElemRestrictionSetData(rstr, d));
  CeedCallBackend(CeedSetBackendFunction(ceed, "ElemRestriction", rstr, "Apply", RestrictionApply));
  CeedCallBackend(CeedSetBackendFunction(ceed, "ElemRestriction", rstr, "GetOffsets", RestrictionGetOffsets));
  CeedCallBackend(CeedSetBackendFunction(ceed, "ElemRestriction", rstr, "Destroy", RestrictionDestroy));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* basis                                                                     */
/* ------------------------------------------------------------------------ */
static int basis_apply(CeedBasis basis, bool add, CeedInt num_elem, CeedTransposeMode t_mode, CeedEvalMode eval_mode, CeedVector u, CeedVector v) {
  BasisData  *d;
  Ceed        ceed;
  CeedScalar *du = NULL, *dv;
  CeedCallBackend(CeedBasisGetData(basis, &d));
  CeedCallBackend(CeedBasisGetCeed(basis, &ceed));
  if (u != CEED_VECTOR_NONE) CeedCallBackend(CeedMi355VectorDevice(u, false, &du));
  if (add) CeedCallBackend(CeedVectorGetArray(v, CEED_MEM_DEVICE, &dv));
  else CeedCallBackend(CeedVectorGetArrayWrite(v, CEED_MEM_DEVICE, &dv));
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355BasisApply(d->handle, num_elem, t_mode, eval_mode, add, du, dv, NULL)));
  CeedCallBackend(CeedVectorRestoreArray(v, &dv));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int BasisApply(CeedBasis basis, CeedInt num_elem, CeedTransposeMode t_mode, CeedEvalMode eval_mode, CeedVector u, CeedVector v) {
  return basis_apply(basis, false, num_elem, t_mode, eval_mode, u, v);
}

static int BasisApplyAdd(CeedBasis basis, CeedInt num_elem, CeedTransposeMode t_mode, CeedEvalMode eval_mode, CeedVector u, CeedVector v) {
  return basis_apply(basis, true, num_elem, t_mode, eval_mode, u, v);
}

static int BasisDestroy(CeedBasis basis) {
  BasisData *d;
  CeedCallBackend(CeedBasisGetData(basis, &d));
  CeedMi355BasisDestroy(&d->handle);
  free(d);
  return CEED_ERROR_SUCCESS;
}

static int BasisCreateTensorH1(CeedInt dim, CeedInt P_1d, CeedInt Q_1d, const CeedScalar *interp_1d, const CeedScalar *grad_1d,
                               const CeedScalar *q_ref_1d, const CeedScalar *q_weight_1d, CeedBasis basis) {
  Ceed    ceed;
  CeedInt num_comp;
  (void)q_ref_1d;
  CeedCallBackend(CeedBasisGetCeed(basis, &ceed));
  CeedCallBackend(CeedBasisGetNumComponents(basis, &num_comp));
  BasisData *d = calloc(1, sizeof(BasisData));
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355BasisCreateTensorH1(dim, num_comp, P_1d, Q_1d, interp_1d, grad_1d, q_weight_1d, &d->handle)));
  CeedCallBackend(CeedBasisSetData(basis, d));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Basis", basis, "Apply", BasisApply));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Basis", basis, "ApplyAdd", BasisApplyAdd));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Basis", basis, "Destroy", BasisDestroy));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* QFunction and context                                                     */
/* ------------------------------------------------------------------------ */
int CeedMi355QFunctionKernel(CeedQFunction qf, CeedMi355QFunction *kernel) {
  QFunctionData *d;
  Ceed           ceed;
  CeedCallBackend(CeedQFunctionGetData(qf, &d));
  if (!d->kernel) {
    const char        *path, *name;
    CeedInt            num_in, num_out;
    CeedQFunctionField *in, *out;
    int32_t            in_sizes[CEED_FIELD_MAX], out_sizes[CEED_FIELD_MAX];
    CeedCallBackend(CeedQFunctionGetCeed(qf, &ceed));
    CeedCallBackend(CeedQFunctionGetSourcePath(qf, &path));
    CeedCallBackend(CeedQFunctionGetKernelName(qf, &name));
    CeedCallBackend(CeedQFunctionGetFields(qf, &num_in, &in, &num_out, &out));
    for (CeedInt i = 0; i < num_in; i++) CeedCallBackend(CeedQFunctionFieldGetSize(in[i], &in_sizes[i]));
    for (CeedInt i = 0; i < num_out; i++) CeedCallBackend(CeedQFunctionFieldGetSize(out[i], &out_sizes[i]));
    CeedCallBackend(CeedMi355Check(ceed, CeedMi355QFunctionCompile(path, name, num_in, in_sizes, num_out, out_sizes, &d->kernel)));
    CeedCallBackend(CeedDestroy(&ceed));
  }
  *kernel = d->kernel;
  return CEED_ERROR_SUCCESS;
}

/* device copy of the QFunction's context data, refreshed when the context changes */
int CeedMi355QFunctionContextDevice(CeedQFunction qf, void **ctx_device) {
  CeedQFunctionContext ctx;
  *ctx_device = NULL;
  CeedCallBackend(CeedQFunctionGetContext(qf, &ctx));
  if (!ctx) return CEED_ERROR_SUCCESS;
  ContextData *d;
  Ceed         ceed;
  uint64_t     state;
  size_t       size;
  CeedCallBackend(CeedQFunctionContextGetCeed(ctx, &ceed));
  CeedCallBackend(CeedQFunctionContextGetBackendData(ctx, &d));
  CeedCallBackend(CeedQFunctionContextGetState(ctx, &state));
  CeedCallBackend(CeedQFunctionContextGetContextSize(ctx, &size));
  if ((!d->has_device_copy || state != d->device_state) && size > 0 && d->data) {
    void *host;
    if (!d->device || d->device_size < size) {
      CeedMi355Free(d->device);
      CeedCallBackend(dev_alloc(ceed, size, &d->device));
      d->device_size = size;
    }
    CeedCallBackend(CeedQFunctionContextGetDataRead(ctx, CEED_MEM_HOST, &host));
    CeedCallBackend(copy_sync(ceed, d->device, host, size));
    CeedCallBackend(CeedQFunctionContextRestoreDataRead(ctx, &host));
    d->device_state    = state;
    d->has_device_copy = true;
  }
  *ctx_device = d->device;
  CeedCallBackend(CeedDestroy(&ceed));
  CeedCallBackend(CeedQFunctionContextDestroy(&ctx));
  return CEED_ERROR_SUCCESS;
}

static int QFunctionApply(CeedQFunction qf, CeedInt Q, CeedVector *u, CeedVector *v) {
  Ceed               ceed;
  CeedMi355QFunction kernel;
  CeedInt            num_in, num_out;
  const double      *in[CEED_FIELD_MAX];
  double            *out[CEED_FIELD_MAX];
  void              *ctx;
  CeedCallBackend(CeedQFunctionGetCeed(qf, &ceed));
  CeedCallBackend(CeedQFunctionGetNumArgs(qf, &num_in, &num_out));
  CeedCallBackend(CeedMi355QFunctionKernel(qf, &kernel));
  CeedCallBackend(CeedMi355QFunctionContextDevice(qf, &ctx));
  for (CeedInt i = 0; i < num_in; i++) CeedCallBackend(CeedMi355VectorDevice(u[i], false, (CeedScalar **)&in[i]));
  for (CeedInt i = 0; i < num_out; i++) CeedCallBackend(CeedVectorGetArrayWrite(v[i], CEED_MEM_DEVICE, &out[i]));
  /* a single "element" of Q points: field f, component k of point q at k*Q + q */
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355QFunctionApply(kernel, 1, Q, ctx, in, out, NULL)));
  for (CeedInt i = 0; i < num_out; i++) CeedCallBackend(CeedVectorRestoreArray(v[i], &out[i]));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int QFunctionDestroy(CeedQFunction qf) {
  QFunctionData *d;
  CeedCallBackend(CeedQFunctionGetData(qf, &d));
  CeedMi355QFunctionDestroy(&d->kernel);
  free(d);
  return CEED_ERROR_SUCCESS;
}

static int QFunctionCreate(CeedQFunction qf) {
  Ceed ceed;
  CeedCallBackend(CeedQFunctionGetCeed(qf, &ceed));
  QFunctionData *d = calloc(1, sizeof(QFunctionData));
  CeedCallBackend(CeedQFunctionSetData(qf, d));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunction", qf, "Apply", QFunctionApply));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunction", qf, "Destroy", QFunctionDestroy));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* The context bytes live in host memory (QFunctions read them through a
 * device copy refreshed on state change, CeedMi355QFunctionContextDevice);
 * the slots follow the backend contract of interface/ceed-qfunctioncontext.c. */
static int context_data(CeedQFunctionContext ctx, CeedMemType mt, ContextData **d) {
  CeedCallBackend(CeedQFunctionContextGetBackendData(ctx, d));
  if (mt != CEED_MEM_HOST) {
    Ceed ceed;
    CeedCallBackend(CeedQFunctionContextGetCeed(ctx, &ceed));
    return CeedError(ceed, CEED_ERROR_UNSUPPORTED, "/gpu/hip/mi355 keeps QFunction context data in host memory");
  }
  return CEED_ERROR_SUCCESS;
}

static int ContextHasValidData(CeedQFunctionContext ctx, bool *has_valid_data) {
  ContextData *d;
  CeedCallBackend(CeedQFunctionContextGetBackendData(ctx, &d));
  *has_valid_data = d->data != NULL;
  return CEED_ERROR_SUCCESS;
}

static int ContextHasBorrowedDataOfType(CeedQFunctionContext ctx, CeedMemType mt, bool *has_borrowed) {
  ContextData *d;
  CeedCallBackend(CeedQFunctionContextGetBackendData(ctx, &d));
  *has_borrowed = mt == CEED_MEM_HOST && d->data && d->data != d->owned;
  return CEED_ERROR_SUCCESS;
}

static int ContextSetData(CeedQFunctionContext ctx, CeedMemType mt, CeedCopyMode cm, void *data) {
  ContextData *d;
  size_t       size;
  CeedCallBackend(context_data(ctx, mt, &d));
  CeedCallBackend(CeedQFunctionContextGetContextSize(ctx, &size));
  free(d->owned);
  d->owned = NULL;
  if (cm == CEED_COPY_VALUES) {
    d->owned = malloc(size ? size : 1);
    if (!d->owned) {
      Ceed ceed;
      CeedCallBackend(CeedQFunctionContextGetCeed(ctx, &ceed));
      return CeedError(ceed, CEED_ERROR_MAJOR, "host allocation failed");
    }
    if (size) memcpy(d->owned, data, size);
    d->data = d->owned;
  } else {
    d->data  = data;
    d->owned = cm == CEED_OWN_POINTER ? data : NULL;
  }
  d->has_device_copy = false;
  return CEED_ERROR_SUCCESS;
}

static int ContextTakeData(CeedQFunctionContext ctx, CeedMemType mt, void *data) {
  ContextData *d;
  CeedCallBackend(context_data(ctx, mt, &d));
  *(void **)data     = d->data;
  d->data            = NULL;
  d->has_device_copy = false;
  return CEED_ERROR_SUCCESS;
}

static int ContextGetData(CeedQFunctionContext ctx, CeedMemType mt, void *data) {
  ContextData *d;
  CeedCallBackend(context_data(ctx, mt, &d));
  *(void **)data = d->data;
  return CEED_ERROR_SUCCESS;
}

static int ContextRestoreData(CeedQFunctionContext ctx) {
  (void)ctx;
  return CEED_ERROR_SUCCESS;
}

static int QFunctionContextDestroy(CeedQFunctionContext ctx) {
  ContextData *d;
  CeedCallBackend(CeedQFunctionContextGetBackendData(ctx, &d));
  if (d) {
    CeedMi355Free(d->device);
    free(d->owned);
    free(d);
  }
  return CEED_ERROR_SUCCESS;
}

static int QFunctionContextCreate(CeedQFunctionContext ctx) {
  Ceed         ceed;
  ContextData *d = calloc(1, sizeof(ContextData));
  CeedCallBackend(CeedQFunctionContextGetCeed(ctx, &ceed));
  CeedCheck(d, ceed, CEED_ERROR_MAJOR, "host allocation failed");
  CeedCallBackend(CeedQFunctionContextSetBackendData(ctx, d));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunctionContext", ctx, "HasValidData", ContextHasValidData));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunctionContext", ctx, "HasBorrowedDataOfType", ContextHasBorrowedDataOfType));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunctionContext", ctx, "SetData", ContextSetData));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunctionContext", ctx, "TakeData", ContextTakeData));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunctionContext", ctx, "GetData", ContextGetData));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunctionContext", ctx, "GetDataRead", ContextGetData));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunctionContext", ctx, "RestoreData", ContextRestoreData));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunctionContext", ctx, "RestoreDataRead", ContextRestoreData));
  CeedCallBackend(CeedSetBackendFunction(ceed, "QFunctionContext", ctx, "Destroy", QFunctionContextDestroy));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* registration                                                              */
/* ------------------------------------------------------------------------ */
static int GetPreferredMemType(CeedMemType *mt) {
  *mt = CEED_MEM_DEVICE;
  return CEED_ERROR_SUCCESS;
}

static int CeedDestroy_Mi355(Ceed ceed) {
  void *data;
  CeedCallBackend(CeedGetData(ceed, &data));
  free(data);
  return CEED_ERROR_SUCCESS;
}

int CeedInit_Mi355(const char *resource, Ceed ceed) {
  int         device_id = 0;
  const char *opt       = strstr(resource, ":device_id=");
  CeedCheck(!strncmp(resource, "/gpu/hip/mi355", 14) || !strcmp(resource, "/gpu/hip") || !strncmp(resource, "/gpu/hip:", 9) ||
                !strcmp(resource, "/gpu") || !strncmp(resource, "/gpu:", 5),
            ceed, CEED_ERROR_BACKEND, "Mi355 backend cannot use resource: %s", resource);
  if (opt) device_id = atoi(opt + 11);
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355DeviceInit(device_id)));
  int *data = calloc(1, sizeof(int));
  *data     = device_id;
  CeedCallBackend(CeedSetData(ceed, data));
  {
    /* Every default E^T mode gives the same bits on every run: colours, the
     * transpose map and bricks sum in a fixed order, and the multiplicity
     * split's atomic adds are two-term sums, which do not depend on the order.
     * CEED_MI355_SCATTER=atomic switches plans to order-dependent atomics. */
    const char *scatter = getenv("CEED_MI355_SCATTER");
    bool        deterministic = true;
    if (scatter) {
      CeedCheck(!strcmp(scatter, "color") || !strcmp(scatter, "tmap") || !strcmp(scatter, "split") || !strcmp(scatter, "atomic") ||
                    !strcmp(scatter, "brick"),
                ceed, CEED_ERROR_BACKEND, "CEED_MI355_SCATTER=%s: expected color, tmap, split, atomic or brick", scatter);
      deterministic = strcmp(scatter, "atomic") != 0;
    }
    CeedCallBackend(CeedSetDeterministic(ceed, deterministic));
  }
  CeedCallBackend(CeedSetBackendFunction(ceed, "Ceed", ceed, "GetPreferredMemType", GetPreferredMemType));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Ceed", ceed, "Destroy", CeedDestroy_Mi355));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Ceed", ceed, "VectorCreate", VectorCreate));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Ceed", ceed, "ElemRestrictionCreate", ElemRestrictionCreate));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Ceed", ceed, "BasisCreateTensorH1", BasisCreateTensorH1));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Ceed", ceed, "QFunctionCreate", QFunctionCreate));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Ceed", ceed, "QFunctionContextCreate", QFunctionContextCreate));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Ceed", ceed, "OperatorCreate", CeedOperatorCreate_Mi355));
  return CEED_ERROR_SUCCESS;
}

int CeedRegister_Mi355(void) { return CeedRegister("/gpu/hip/mi355", CeedInit_Mi355, 20); }
