/*
 * CeedOperator interface: creation, field binding with the consistency
 * checks of interface/ceed-operator.c:931-1040 (element counts, quadrature
 * point counts, vector lengths), readiness (:1450-1500) and apply:
 * CeedOperatorApply uses the backend's Apply when it has one, otherwise zeroes
 * the active output and calls ApplyAdd (:2271-2294); ApplyAdd (:2313-2340).
 */
#include <stdlib.h>
#include <string.h>

#include "ceed-core.h"

int CeedOperatorCreate(Ceed ceed, CeedQFunction qf, CeedQFunction dqf, CeedQFunction dqfT, CeedOperator *op) {
  (void)dqf, (void)dqfT;
  CeedCheck(qf && qf != CEED_QFUNCTION_NONE, ceed, CEED_ERROR_MINOR, "operator must have a valid QFunction");
  CeedCheck(ceed->OperatorCreate, ceed, CEED_ERROR_UNSUPPORTED, "backend does not support OperatorCreate");
  CeedCall(CeedCalloc(1, op));
  CeedOperator o = *op;
  CeedCall(CeedReferenceCopy(ceed, &o->ceed));
  o->ref_count = 1;
  CeedCall(CeedQFunctionReferenceCopy(qf, &o->qf));
  for (int i = 0; i < CEED_CORE_MAX_FIELDS; i++) {
    o->in_ptr[i]  = &o->in[i];
    o->out_ptr[i] = &o->out[i];
  }
  CeedCall(ceed->OperatorCreate(o));
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorReferenceCopy(CeedOperator op, CeedOperator *op_copy) {
  if (op) op->ref_count++;
  if (*op_copy) CeedCall(CeedOperatorDestroy(op_copy));
  *op_copy = op;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorSetField(CeedOperator op, const char *field_name, CeedElemRestriction rstr, CeedBasis basis, CeedVector vec) {
  CeedQFunction     qf = op->qf;
  CeedQFunctionField qf_field = NULL;
  struct CeedOperatorField_private *slot = NULL;
  CeedCheck(!op->ready, op->ceed, CEED_ERROR_MAJOR, "operator cannot be changed after it is applied");
  for (CeedInt i = 0; i < qf->num_in; i++)
    if (!strcmp(qf->in[i].name, field_name)) qf_field = &qf->in[i], slot = &op->in[i];
  for (CeedInt i = 0; i < qf->num_out; i++)
    if (!strcmp(qf->out[i].name, field_name)) qf_field = &qf->out[i], slot = &op->out[i];
  CeedCheck(qf_field, op->ceed, CEED_ERROR_INCOMPLETE, "QFunction has no field named %s", field_name);
  CeedCheck(!slot->name, op->ceed, CEED_ERROR_INCOMPLETE, "field %s already set", field_name);
  CeedCheck(rstr, op->ceed, CEED_ERROR_INCOMPATIBLE, "restriction for field %s must be non-NULL", field_name);
  CeedCheck(basis, op->ceed, CEED_ERROR_INCOMPATIBLE, "basis for field %s must be non-NULL", field_name);
  CeedCheck(vec, op->ceed, CEED_ERROR_INCOMPATIBLE, "vector for field %s must be non-NULL", field_name);
  const CeedEvalMode mode = qf_field->eval_mode;
  CeedCheck((rstr == CEED_ELEMRESTRICTION_NONE) == (mode == CEED_EVAL_WEIGHT), op->ceed, CEED_ERROR_INCOMPATIBLE,
            "CEED_ELEMRESTRICTION_NONE goes with CEED_EVAL_WEIGHT only (field %s)", field_name);
  CeedCheck((basis == CEED_BASIS_NONE) == (mode == CEED_EVAL_NONE), op->ceed, CEED_ERROR_INCOMPATIBLE,
            "CEED_BASIS_NONE goes with CEED_EVAL_NONE only (field %s)", field_name);
  if (rstr != CEED_ELEMRESTRICTION_NONE) {
    CeedCheck(!op->has_num_elem || rstr->num_elem == op->num_elem, op->ceed, CEED_ERROR_DIMENSION,
              "restriction of field %s has %d elements, operator has %d", field_name, rstr->num_elem, op->num_elem);
    op->num_elem     = rstr->num_elem;
    op->has_num_elem = true;
  }
  CeedInt num_qpts = -1;
  if (basis != CEED_BASIS_NONE) {
    CeedCall(CeedBasisGetNumQuadraturePoints(basis, &num_qpts));
    CeedInt expect = basis->num_comp;
    if (mode == CEED_EVAL_GRAD) expect *= basis->dim;
    if (mode == CEED_EVAL_WEIGHT) expect = 1;
    CeedCheck(qf_field->size == expect, op->ceed, CEED_ERROR_DIMENSION, "field %s: QFunction size %d does not match basis (%d)", field_name,
              qf_field->size, expect);
    if (rstr != CEED_ELEMRESTRICTION_NONE)
      CeedCheck(rstr->num_comp == basis->num_comp, op->ceed, CEED_ERROR_DIMENSION, "field %s: restriction and basis components differ",
                field_name);
  } else {
    num_qpts = rstr->elem_size;
    CeedCheck(qf_field->size == rstr->num_comp, op->ceed, CEED_ERROR_DIMENSION, "field %s: QFunction size %d vs restriction components %d",
              field_name, qf_field->size, rstr->num_comp);
  }
  CeedCheck(!op->has_num_qpts || num_qpts == op->num_qpts, op->ceed, CEED_ERROR_DIMENSION, "field %s has %d quadrature points, operator has %d",
            field_name, num_qpts, op->num_qpts);
  op->num_qpts     = num_qpts;
  op->has_num_qpts = true;
  if (vec != CEED_VECTOR_ACTIVE && vec != CEED_VECTOR_NONE && rstr != CEED_ELEMRESTRICTION_NONE)
    CeedCheck(vec->length >= rstr->l_size, op->ceed, CEED_ERROR_DIMENSION, "vector of field %s is shorter than its L-vector", field_name);
  CeedCall(CeedStringAllocCopy(field_name, &slot->name));
  CeedCall(CeedElemRestrictionReferenceCopy(rstr, &slot->rstr));
  CeedCall(CeedBasisReferenceCopy(basis, &slot->basis));
  CeedCall(CeedVectorReferenceCopy(vec, &slot->vec));
  op->num_fields_set++;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorCheckReady(CeedOperator op) {
  if (op->ready) return CEED_ERROR_SUCCESS;
  CeedCheck(op->num_fields_set == op->qf->num_in + op->qf->num_out, op->ceed, CEED_ERROR_INCOMPLETE, "not all operator fields set");
  CeedCheck(op->has_num_elem, op->ceed, CEED_ERROR_INCOMPLETE, "at least one restriction required");
  op->ready            = true;
  op->qf->is_immutable = true;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorGetCeed(CeedOperator op, Ceed *ceed) {
  *ceed = NULL;
  return CeedReferenceCopy(op->ceed, ceed);
}

int CeedOperatorGetQFunction(CeedOperator op, CeedQFunction *qf) {
  *qf = NULL;
  return CeedQFunctionReferenceCopy(op->qf, qf);
}

int CeedOperatorGetFields(CeedOperator op, CeedInt *num_input_fields, CeedOperatorField **input_fields, CeedInt *num_output_fields,
                          CeedOperatorField **output_fields) {
  if (num_input_fields) *num_input_fields = op->qf->num_in;
  if (input_fields) *input_fields = op->in_ptr;
  if (num_output_fields) *num_output_fields = op->qf->num_out;
  if (output_fields) *output_fields = op->out_ptr;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorFieldGetElemRestriction(CeedOperatorField f, CeedElemRestriction *rstr) {
  *rstr = NULL;
  return CeedElemRestrictionReferenceCopy(f->rstr, rstr);
}

int CeedOperatorFieldGetBasis(CeedOperatorField f, CeedBasis *basis) {
  *basis = NULL;
  return CeedBasisReferenceCopy(f->basis, basis);
}

int CeedOperatorFieldGetVector(CeedOperatorField f, CeedVector *vec) {
  *vec = NULL;
  return CeedVectorReferenceCopy(f->vec, vec);
}

int CeedOperatorIsComposite(CeedOperator op, bool *is_composite) {
  (void)op;
  *is_composite = false;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorGetData(CeedOperator op, void *data) {
  *(void **)data = op->data;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorSetData(CeedOperator op, void *data) {
  op->data = data;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorGetNumElements(CeedOperator op, CeedInt *num_elem) {
  *num_elem = op->num_elem;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorGetNumQuadraturePoints(CeedOperator op, CeedInt *num_qpts) {
  *num_qpts = op->num_qpts;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorApply(CeedOperator op, CeedVector in, CeedVector out, CeedRequest *request) {
  CeedCall(CeedOperatorCheckReady(op));
  if (op->Apply) return op->Apply(op, in, out, request);
  if (out != CEED_VECTOR_NONE) CeedCall(CeedVectorSetValue(out, 0.0));
  return CeedOperatorApplyAdd(op, in, out, request);
}

int CeedOperatorApplyAdd(CeedOperator op, CeedVector in, CeedVector out, CeedRequest *request) {
  CeedCall(CeedOperatorCheckReady(op));
  CeedCheck(op->ApplyAdd, op->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support OperatorApplyAdd");
  if (op->num_elem == 0) return CEED_ERROR_SUCCESS;
  return op->ApplyAdd(op, in, out, request);
}

int CeedOperatorDestroy(CeedOperator *op) {
  if (!*op || --(*op)->ref_count > 0) {
    *op = NULL;
    return CEED_ERROR_SUCCESS;
  }
  if ((*op)->Destroy) CeedCall((*op)->Destroy(*op));
  for (int i = 0; i < CEED_CORE_MAX_FIELDS; i++) {
    struct CeedOperatorField_private *f[2] = {&(*op)->in[i], &(*op)->out[i]};
    for (int k = 0; k < 2; k++) {
      if (!f[k]->name) continue;
      free(f[k]->name);
      CeedCall(CeedElemRestrictionDestroy(&f[k]->rstr));
      CeedCall(CeedBasisDestroy(&f[k]->basis));
      CeedCall(CeedVectorDestroy(&f[k]->vec));
    }
  }
  CeedCall(CeedQFunctionDestroy(&(*op)->qf));
  CeedCall(CeedDestroy(&(*op)->ceed));
  free(*op);
  *op = NULL;
  return CEED_ERROR_SUCCESS;
}
