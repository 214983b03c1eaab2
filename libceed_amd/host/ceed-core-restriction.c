/*
 * CeedElemRestriction interface: creation (standard offsets or strides),
 * size checks and dispatch, following interface/ceed-elemrestriction.c
 * (create :657-680, strided :824, apply :1314-1336).  The backend keeps the
 * offsets (the caller's array under its copy mode) and hands them out through
 * its GetOffsets slot (:269-299).
 */
#include <stdlib.h>
#include <string.h>

#include "ceed-core.h"

static int rstr_alloc(Ceed ceed, CeedInt num_elem, CeedInt elem_size, CeedInt num_comp, CeedSize l_size, CeedElemRestriction *rstr) {
  CeedCheck(num_elem >= 0, ceed, CEED_ERROR_DIMENSION, "CeedElemRestriction must have at least 0 elements");
  CeedCheck(elem_size > 0, ceed, CEED_ERROR_DIMENSION, "element size must be at least 1");
  CeedCheck(num_comp > 0, ceed, CEED_ERROR_DIMENSION, "a restriction needs one or more components");
  CeedCheck(ceed->ElemRestrictionCreate, ceed, CEED_ERROR_UNSUPPORTED, "backend does not support ElemRestrictionCreate");
  CeedCall(CeedCalloc(1, rstr));
  CeedCall(CeedReferenceCopy(ceed, &(*rstr)->ceed));
  (*rstr)->ref_count   = 1;
  (*rstr)->num_elem    = num_elem;
  (*rstr)->elem_size   = elem_size;
  (*rstr)->num_comp    = num_comp;
  (*rstr)->l_size      = l_size;
  (*rstr)->e_layout[0] = 1;
  (*rstr)->e_layout[1] = elem_size;
  (*rstr)->e_layout[2] = elem_size * num_comp;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionCreate(Ceed ceed, CeedInt num_elem, CeedInt elem_size, CeedInt num_comp, CeedInt comp_stride, CeedSize l_size,
                              CeedMemType mem_type, CeedCopyMode copy_mode, const CeedInt *offsets, CeedElemRestriction *rstr) {
  CeedCheck(num_comp == 1 || comp_stride > 0, ceed, CEED_ERROR_DIMENSION, "components of a multi-component restriction need a positive stride");
  CeedCheck(mem_type == CEED_MEM_HOST, ceed, CEED_ERROR_UNSUPPORTED, "offsets must be host memory");
  CeedCall(rstr_alloc(ceed, num_elem, elem_size, num_comp, l_size, rstr));
  CeedElemRestriction r = *rstr;
  r->type               = CEED_RESTRICTION_STANDARD;
  r->comp_stride        = comp_stride;
  const size_t n        = (size_t)num_elem * elem_size;
  for (size_t i = 0; i < n; i++) {
    const CeedSize max_l = (CeedSize)offsets[i] + (CeedSize)(num_comp - 1) * comp_stride;
    CeedCheck(offsets[i] >= 0 && max_l < l_size, ceed, CEED_ERROR_DIMENSION, "restriction offset %d out of range for L-vector of size %td",
              offsets[i], l_size);
  }
  CeedCall(ceed->ElemRestrictionCreate(mem_type, copy_mode, offsets, NULL, NULL, r));
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionCreateStrided(Ceed ceed, CeedInt num_elem, CeedInt elem_size, CeedInt num_comp, CeedSize l_size, const CeedInt strides[3],
                                     CeedElemRestriction *rstr) {
  CeedCall(rstr_alloc(ceed, num_elem, elem_size, num_comp, l_size, rstr));
  CeedElemRestriction r = *rstr;
  r->type               = CEED_RESTRICTION_STRIDED;
  r->backend_strides    = strides == CEED_STRIDES_BACKEND;
  if (r->backend_strides) {
    /* this backend's choice: [elem][comp][node], the layout of its E-vectors */
    r->strides[0] = 1;
    r->strides[1] = elem_size;
    r->strides[2] = elem_size * num_comp;
  } else {
    memcpy(r->strides, strides, sizeof(r->strides));
  }
  CeedCall(ceed->ElemRestrictionCreate(CEED_MEM_HOST, CEED_COPY_VALUES, NULL, NULL, NULL, r));
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionReferenceCopy(CeedElemRestriction rstr, CeedElemRestriction *rstr_copy) {
  if (rstr && rstr != CEED_ELEMRESTRICTION_NONE) rstr->ref_count++;
  if (*rstr_copy && *rstr_copy != CEED_ELEMRESTRICTION_NONE) CeedCall(CeedElemRestrictionDestroy(rstr_copy));
  *rstr_copy = rstr;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetCeed(CeedElemRestriction rstr, Ceed *ceed) {
  *ceed = NULL;
  return CeedReferenceCopy(rstr->ceed, ceed);
}

int CeedElemRestrictionGetType(CeedElemRestriction rstr, CeedRestrictionType *rstr_type) {
  *rstr_type = rstr->type;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionIsStrided(CeedElemRestriction rstr, bool *is_strided) {
  *is_strided = rstr->type == CEED_RESTRICTION_STRIDED;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetStrides(CeedElemRestriction rstr, CeedInt strides[3]) {
  CeedCheck(rstr->type == CEED_RESTRICTION_STRIDED, rstr->ceed, CEED_ERROR_MINOR, "restriction is not strided");
  memcpy(strides, rstr->strides, sizeof(rstr->strides));
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionHasBackendStrides(CeedElemRestriction rstr, bool *has_backend_strides) {
  *has_backend_strides = rstr->type == CEED_RESTRICTION_STRIDED && rstr->backend_strides;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetOffsets(CeedElemRestriction rstr, CeedMemType mem_type, const CeedInt **offsets) {
  CeedCheck(rstr->type == CEED_RESTRICTION_STANDARD, rstr->ceed, CEED_ERROR_MINOR, "restriction has no offsets");
  CeedCheck(rstr->GetOffsets, rstr->ceed, CEED_ERROR_UNSUPPORTED, "backend does not hand out restriction offsets");
  CeedCall(rstr->GetOffsets(rstr, mem_type, offsets));
  rstr->num_readers++;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionRestoreOffsets(CeedElemRestriction rstr, const CeedInt **offsets) {
  CeedCheck(rstr->num_readers > 0, rstr->ceed, CEED_ERROR_ACCESS, "no offsets array of this restriction is out");
  rstr->num_readers--;
  *offsets = NULL;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetELayout(CeedElemRestriction rstr, CeedInt layout[3]) {
  memcpy(layout, rstr->e_layout, sizeof(rstr->e_layout));
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionSetELayout(CeedElemRestriction rstr, CeedInt layout[3]) {
  memcpy(rstr->e_layout, layout, sizeof(rstr->e_layout));
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetData(CeedElemRestriction rstr, void *data) {
  *(void **)data = rstr->data;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionSetData(CeedElemRestriction rstr, void *data) {
  rstr->data = data;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionCreateVector(CeedElemRestriction rstr, CeedVector *lvec, CeedVector *evec) {
  if (lvec) CeedCall(CeedVectorCreate(rstr->ceed, rstr->l_size, lvec));
  if (evec) CeedCall(CeedVectorCreate(rstr->ceed, (CeedSize)rstr->num_elem * rstr->elem_size * rstr->num_comp, evec));
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionApply(CeedElemRestriction rstr, CeedTransposeMode t_mode, CeedVector u, CeedVector ru, CeedRequest *request) {
  const CeedSize e_size = (CeedSize)rstr->num_elem * rstr->elem_size * rstr->num_comp;
  const CeedSize in_min = t_mode == CEED_NOTRANSPOSE ? rstr->l_size : e_size, out_min = t_mode == CEED_NOTRANSPOSE ? e_size : rstr->l_size;
  CeedCheck(u->length >= in_min, rstr->ceed, CEED_ERROR_DIMENSION, "input vector size %td not compatible with element restriction (%td)", u->length,
            in_min);
  CeedCheck(ru->length >= out_min, rstr->ceed, CEED_ERROR_DIMENSION, "output vector size %td not compatible with element restriction (%td)",
            ru->length, out_min);
  if (rstr->num_elem > 0) CeedCall(rstr->Apply(rstr, t_mode, u, ru, request));
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetCompStride(CeedElemRestriction rstr, CeedInt *comp_stride) {
  *comp_stride = rstr->comp_stride;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetNumElements(CeedElemRestriction rstr, CeedInt *num_elem) {
  *num_elem = rstr->num_elem;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetElementSize(CeedElemRestriction rstr, CeedInt *elem_size) {
  *elem_size = rstr->elem_size;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetLVectorSize(CeedElemRestriction rstr, CeedSize *l_size) {
  *l_size = rstr->l_size;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetEVectorSize(CeedElemRestriction rstr, CeedSize *e_size) {
  *e_size = (CeedSize)rstr->num_elem * rstr->elem_size * rstr->num_comp;
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionGetNumComponents(CeedElemRestriction rstr, CeedInt *num_comp) {
  *num_comp = rstr->num_comp;
  return CEED_ERROR_SUCCESS;
}

/* mult = E^T 1: number of element nodes mapping to each L entry (t209). */
int CeedElemRestrictionGetMultiplicity(CeedElemRestriction rstr, CeedVector mult) {
  CeedVector e_vec;
  CeedCall(CeedElemRestrictionCreateVector(rstr, NULL, &e_vec));
  CeedCall(CeedVectorSetValue(e_vec, 1.0));
  CeedCall(CeedVectorSetValue(mult, 0.0));
  CeedCall(CeedElemRestrictionApply(rstr, CEED_TRANSPOSE, e_vec, mult, CEED_REQUEST_IMMEDIATE));
  CeedCall(CeedVectorDestroy(&e_vec));
  return CEED_ERROR_SUCCESS;
}

int CeedElemRestrictionDestroy(CeedElemRestriction *rstr) {
  if (!*rstr || *rstr == CEED_ELEMRESTRICTION_NONE || --(*rstr)->ref_count > 0) {
    *rstr = NULL;
    return CEED_ERROR_SUCCESS;
  }
  if ((*rstr)->Destroy) CeedCall((*rstr)->Destroy(*rstr));
  CeedCall(CeedDestroy(&(*rstr)->ceed));
  free(*rstr);
  *rstr = NULL;
  return CEED_ERROR_SUCCESS;
}
