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
  CeedCheck(!op->is_composite, op->ceed, CEED_ERROR_INCOMPATIBLE, "cannot add a field to a composite operator");
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

/* Composite operators: the sum of their sub-operators (interface/ceed-operator.c:863-1440). */
int CeedOperatorCreateComposite(Ceed ceed, CeedOperator *op) {
  CeedCall(CeedCalloc(1, op));
  CeedOperator o = *op;
  CeedCall(CeedReferenceCopy(ceed, &o->ceed));
  o->ref_count    = 1;
  o->is_composite = true;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorCompositeAddSub(CeedOperator composite_op, CeedOperator sub_op) {
  CeedCheck(composite_op->is_composite, composite_op->ceed, CEED_ERROR_MINOR, "this call needs a composite CeedOperator");
  CeedCheck(!sub_op->is_composite, composite_op->ceed, CEED_ERROR_UNSUPPORTED, "nested composite operators are not supported");
  CeedCheck(!composite_op->ready, composite_op->ceed, CEED_ERROR_MAJOR, "composite operator cannot be changed after it is applied");
  if (composite_op->num_sub > 0 && sub_op->has_num_elem) {
    CeedSize in0, out0, in1, out1;
    CeedCall(CeedOperatorGetActiveVectorLengths(composite_op->sub[0], &in0, &out0));
    CeedCall(CeedOperatorGetActiveVectorLengths(sub_op, &in1, &out1));
    CeedCheck((in0 < 0 || in1 < 0 || in0 == in1) && (out0 < 0 || out1 < 0 || out0 == out1), composite_op->ceed, CEED_ERROR_MAJOR,
              "sub-operators must have compatible dimensions");
  }
  CeedCall(CeedRealloc(composite_op->num_sub + 1, &composite_op->sub));
  composite_op->sub[composite_op->num_sub] = NULL;
  CeedCall(CeedOperatorReferenceCopy(sub_op, &composite_op->sub[composite_op->num_sub]));
  composite_op->num_sub++;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorCompositeGetNumSub(CeedOperator op, CeedInt *num_suboperators) {
  CeedCheck(op->is_composite, op->ceed, CEED_ERROR_MINOR, "this call needs a composite CeedOperator");
  *num_suboperators = op->num_sub;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorCompositeGetSubList(CeedOperator op, CeedOperator **sub_operators) {
  CeedCheck(op->is_composite, op->ceed, CEED_ERROR_MINOR, "this call needs a composite CeedOperator");
  *sub_operators = op->sub;
  return CEED_ERROR_SUCCESS;
}

/* L-vector lengths of the active input and output (-1 if there is none) (interface/ceed-operator.c:1513). */
int CeedOperatorGetActiveVectorLengths(CeedOperator op, CeedSize *input_size, CeedSize *output_size) {
  *input_size = *output_size = -1;
  if (op->is_composite) {
    for (CeedInt i = 0; i < op->num_sub; i++) {
      CeedSize a, b;
      CeedCall(CeedOperatorGetActiveVectorLengths(op->sub[i], &a, &b));
      if (a >= 0) *input_size = a;
      if (b >= 0) *output_size = b;
    }
    return CEED_ERROR_SUCCESS;
  }
  for (CeedInt i = 0; op->qf && i < op->qf->num_in; i++)
    if (op->in[i].vec == CEED_VECTOR_ACTIVE) *input_size = op->in[i].rstr->l_size;
  for (CeedInt i = 0; op->qf && i < op->qf->num_out; i++)
    if (op->out[i].vec == CEED_VECTOR_ACTIVE) *output_size = op->out[i].rstr->l_size;
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorCheckReady(CeedOperator op) {
  if (op->ready) return CEED_ERROR_SUCCESS;
  if (op->is_composite) {
    for (CeedInt i = 0; i < op->num_sub; i++) CeedCall(CeedOperatorCheckReady(op->sub[i]));
    op->ready = true;
    return CEED_ERROR_SUCCESS;
  }
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
  CeedCheck(!op->is_composite, op->ceed, CEED_ERROR_MINOR, "a composite operator has no fields");
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
  *is_composite = op->is_composite;
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
  if (!op->is_composite && op->Apply) return op->Apply(op, in, out, request);
  if (out != CEED_VECTOR_NONE) CeedCall(CeedVectorSetValue(out, 0.0));
  return CeedOperatorApplyAdd(op, in, out, request);
}

int CeedOperatorApplyAdd(CeedOperator op, CeedVector in, CeedVector out, CeedRequest *request) {
  CeedCall(CeedOperatorCheckReady(op));
  if (op->is_composite) {  /* sum of the sub-operators, in order (interface/ceed-operator.c:2327-2333) */
    for (CeedInt i = 0; i < op->num_sub; i++) CeedCall(CeedOperatorApplyAdd(op->sub[i], in, out, request));
    return CEED_ERROR_SUCCESS;
  }
  CeedCheck(op->ApplyAdd, op->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support OperatorApplyAdd");
  if (op->num_elem == 0) return CEED_ERROR_SUCCESS;
  return op->ApplyAdd(op, in, out, request);
}

/* Diagonal assembly (interface/ceed-preconditioning.c:2395-2480): the backend's
 * Diagonal slot, else zero then AddDiagonal; a composite sums its sub-operators. */
int CeedOperatorLinearAssembleAddDiagonal(CeedOperator op, CeedVector assembled, CeedRequest *request) {
  CeedSize in, out;
  CeedCall(CeedOperatorCheckReady(op));
  CeedCall(CeedOperatorGetActiveVectorLengths(op, &in, &out));
  CeedCheck(in == out, op->ceed, CEED_ERROR_DIMENSION, "a diagonal needs an operator whose input and output spaces agree");
  if (op->is_composite) {
    for (CeedInt i = 0; i < op->num_sub; i++) CeedCall(CeedOperatorLinearAssembleAddDiagonal(op->sub[i], assembled, request));
    return CEED_ERROR_SUCCESS;
  }
  if (op->num_elem == 0) return CEED_ERROR_SUCCESS;
  CeedCheck(op->LinearAssembleAddDiagonal, op->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support LinearAssembleAddDiagonal for this operator");
  return op->LinearAssembleAddDiagonal(op, assembled, request);
}

int CeedOperatorLinearAssembleDiagonal(CeedOperator op, CeedVector assembled, CeedRequest *request) {
  CeedSize in, out;
  CeedCall(CeedOperatorCheckReady(op));
  CeedCall(CeedOperatorGetActiveVectorLengths(op, &in, &out));
  CeedCheck(in == out, op->ceed, CEED_ERROR_DIMENSION, "a diagonal needs an operator whose input and output spaces agree");
  if (!op->is_composite && op->num_elem == 0) return CEED_ERROR_SUCCESS;
  if (!op->is_composite && op->LinearAssembleDiagonal) return op->LinearAssembleDiagonal(op, assembled, request);
  CeedCall(CeedVectorSetValue(assembled, 0.0));
  return CeedOperatorLinearAssembleAddDiagonal(op, assembled, request);
}

int CeedOperatorDestroy(CeedOperator *op) {
  if (!*op || --(*op)->ref_count > 0) {
    *op = NULL;
    return CEED_ERROR_SUCCESS;
  }
  if ((*op)->Destroy) CeedCall((*op)->Destroy(*op));
  for (CeedInt i = 0; i < (*op)->num_sub; i++) CeedCall(CeedOperatorDestroy(&(*op)->sub[i]));
  free((*op)->sub);
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
  if ((*op)->qf) CeedCall(CeedQFunctionDestroy(&(*op)->qf));
  CeedCall(CeedDestroy(&(*op)->ceed));
  free(*op);
  *op = NULL;
  return CEED_ERROR_SUCCESS;
}
