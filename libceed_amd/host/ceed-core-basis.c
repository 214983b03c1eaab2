/*
 * CeedBasis interface for tensor-product H1 bases: Lagrange construction
 * (GLL nodes, Gauss or Gauss-Lobatto points; interface/ceed-basis.c:1617-1690),
 * getters, the text view whose format tests/output/t300-basis.out records,
 * and apply with the size checks of CeedBasisApplyCheckDims (:350-380).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ceed-core.h"
#include "ceed-mi355.h"

/* every size of a tensor basis counts something that must exist */
static int tensor_sizes_valid(Ceed ceed, CeedInt dim, CeedInt num_comp, CeedInt P_1d, CeedInt Q_1d) {
  CeedCheck(dim > 0 && num_comp > 0 && P_1d > 0 && Q_1d > 0, ceed, CEED_ERROR_DIMENSION,
            "tensor basis sizes must all be positive: dim=%d, %d components, P_1d=%d, Q_1d=%d", dim, num_comp, P_1d, Q_1d);
  return CEED_ERROR_SUCCESS;
}

int CeedBasisCreateTensorH1(Ceed ceed, CeedInt dim, CeedInt num_comp, CeedInt P_1d, CeedInt Q_1d, const CeedScalar *interp_1d,
                            const CeedScalar *grad_1d, const CeedScalar *q_ref_1d, const CeedScalar *q_weight_1d, CeedBasis *basis) {
  CeedCall(tensor_sizes_valid(ceed, dim, num_comp, P_1d, Q_1d));
  CeedCheck(ceed->BasisCreateTensorH1, ceed, CEED_ERROR_UNSUPPORTED, "backend does not support BasisCreateTensorH1");
  CeedCall(CeedCalloc(1, basis));
  CeedBasis b = *basis;
  CeedCall(CeedReferenceCopy(ceed, &b->ceed));
  b->ref_count = 1;
  b->dim       = dim;
  b->num_comp  = num_comp;
  b->P_1d      = P_1d;
  b->Q_1d      = Q_1d;
  CeedCall(CeedMalloc(Q_1d * P_1d, &b->interp_1d));
  CeedCall(CeedMalloc(Q_1d * P_1d, &b->grad_1d));
  CeedCall(CeedMalloc(Q_1d, &b->q_ref_1d));
  CeedCall(CeedMalloc(Q_1d, &b->q_weight_1d));
  memcpy(b->interp_1d, interp_1d, sizeof(CeedScalar) * Q_1d * P_1d);
  memcpy(b->grad_1d, grad_1d, sizeof(CeedScalar) * Q_1d * P_1d);
  if (q_ref_1d) memcpy(b->q_ref_1d, q_ref_1d, sizeof(CeedScalar) * Q_1d);
  if (q_weight_1d) memcpy(b->q_weight_1d, q_weight_1d, sizeof(CeedScalar) * Q_1d);
  CeedCall(ceed->BasisCreateTensorH1(dim, P_1d, Q_1d, b->interp_1d, b->grad_1d, b->q_ref_1d, b->q_weight_1d, b));
  return CEED_ERROR_SUCCESS;
}

int CeedBasisCreateTensorH1Lagrange(Ceed ceed, CeedInt dim, CeedInt num_comp, CeedInt P, CeedInt Q, CeedQuadMode quad_mode, CeedBasis *basis) {
  CeedCall(tensor_sizes_valid(ceed, dim, num_comp, P, Q));
  CeedScalar *interp, *grad, *q_ref, *q_weight;
  CeedCall(CeedCalloc(P * Q, &interp));
  CeedCall(CeedCalloc(P * Q, &grad));
  CeedCall(CeedCalloc(Q, &q_ref));
  CeedCall(CeedCalloc(Q, &q_weight));
  int ierr = CeedMi355LagrangeBasis1D(P, Q, quad_mode == CEED_GAUSS_LOBATTO ? 1 : 0, interp, grad, q_ref, q_weight);
  if (!ierr) ierr = CeedBasisCreateTensorH1(ceed, dim, num_comp, P, Q, interp, grad, q_ref, q_weight, basis);
  else ierr = CeedError(ceed, CEED_ERROR_DIMENSION, "cannot build a Lagrange basis with P=%d Q=%d", P, Q);
  free(interp);
  free(grad);
  free(q_ref);
  free(q_weight);
  return ierr;
}

int CeedBasisReferenceCopy(CeedBasis basis, CeedBasis *basis_copy) {
  if (basis && basis != CEED_BASIS_NONE) basis->ref_count++;
  if (*basis_copy && *basis_copy != CEED_BASIS_NONE) CeedCall(CeedBasisDestroy(basis_copy));
  *basis_copy = basis;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetCeed(CeedBasis basis, Ceed *ceed) {
  *ceed = NULL;
  return CeedReferenceCopy(basis->ceed, ceed);
}

int CeedBasisIsTensor(CeedBasis basis, bool *is_tensor) {
  (void)basis;
  *is_tensor = true;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetData(CeedBasis basis, void *data) {
  *(void **)data = basis->data;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisSetData(CeedBasis basis, void *data) {
  basis->data = data;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetDimension(CeedBasis basis, CeedInt *dim) {
  *dim = basis->dim;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetNumComponents(CeedBasis basis, CeedInt *num_comp) {
  *num_comp = basis->num_comp;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetNumNodes(CeedBasis basis, CeedInt *P) {
  *P = CeedIntPow(basis->P_1d, basis->dim);
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetNumNodes1D(CeedBasis basis, CeedInt *P_1d) {
  *P_1d = basis->P_1d;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetNumQuadraturePoints(CeedBasis basis, CeedInt *Q) {
  *Q = CeedIntPow(basis->Q_1d, basis->dim);
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetNumQuadraturePoints1D(CeedBasis basis, CeedInt *Q_1d) {
  *Q_1d = basis->Q_1d;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetQRef(CeedBasis basis, const CeedScalar **q_ref) {
  *q_ref = basis->q_ref_1d;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetQWeights(CeedBasis basis, const CeedScalar **q_weights) {
  *q_weights = basis->q_weight_1d;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetInterp1D(CeedBasis basis, const CeedScalar **interp_1d) {
  *interp_1d = basis->interp_1d;
  return CEED_ERROR_SUCCESS;
}

int CeedBasisGetGrad1D(CeedBasis basis, const CeedScalar **grad_1d) {
  *grad_1d = basis->grad_1d;
  return CEED_ERROR_SUCCESS;
}

/* one row per line; entries below 1e-14 print as 0, as the reference's views do */
static void view_scalars(FILE *stream, const char *name, CeedInt m, CeedInt n, const CeedScalar *a) {
  if (m > 1) {
    fprintf(stream, "  %s:\n", name);
  } else {
    char label[12];
    snprintf(label, 11, "%s:", name);
    fprintf(stream, "  %-10s", label);
  }
  for (CeedInt i = 0; i < m; i++) {
    if (m > 1) fprintf(stream, "    [%d]", i);
    for (CeedInt j = 0; j < n; j++) fprintf(stream, "\t% 12.8f", fabs(a[i * n + j]) > 1e-14 ? a[i * n + j] : 0.0);
    fputs("\n", stream);
  }
}

int CeedBasisView(CeedBasis basis, FILE *stream) {
  static const char *topo[] = {"", "line", "quadrilateral", "hexahedron"};
  fprintf(stream, "CeedBasis in a H^1 space on a %s element\n", basis->dim <= 3 ? topo[basis->dim] : "tensor");
  fprintf(stream, "  P: %d\n  Q: %d\n  dimension: %d\n  field components: %d\n", basis->P_1d, basis->Q_1d, basis->dim, basis->num_comp);
  view_scalars(stream, "qref1d", 1, basis->Q_1d, basis->q_ref_1d);
  view_scalars(stream, "qweight1d", 1, basis->Q_1d, basis->q_weight_1d);
  view_scalars(stream, "interp1d", basis->Q_1d, basis->P_1d, basis->interp_1d);
  view_scalars(stream, "grad1d", basis->Q_1d, basis->P_1d, basis->grad_1d);
  return CEED_ERROR_SUCCESS;
}

static int basis_apply_check(CeedBasis basis, CeedInt num_elem, CeedTransposeMode t_mode, CeedEvalMode eval_mode, CeedVector u, CeedVector v) {
  CeedInt        nn, nq;
  CeedCall(CeedBasisGetNumNodes(basis, &nn));
  CeedCall(CeedBasisGetNumQuadraturePoints(basis, &nq));
  const CeedSize nodes = (CeedSize)num_elem * basis->num_comp * nn;
  CeedSize       qpts  = 0;
  switch (eval_mode) {
    case CEED_EVAL_INTERP: qpts = (CeedSize)num_elem * basis->num_comp * nq; break;
    case CEED_EVAL_GRAD: qpts = (CeedSize)num_elem * basis->num_comp * nq * basis->dim; break;
    case CEED_EVAL_WEIGHT: qpts = (CeedSize)num_elem * nq; break;
    default: return CeedError(basis->ceed, CEED_ERROR_UNSUPPORTED, "evaluation mode %d not supported by a tensor H1 basis", eval_mode);
  }
  CeedCheck(eval_mode != CEED_EVAL_WEIGHT || t_mode == CEED_NOTRANSPOSE, basis->ceed, CEED_ERROR_BACKEND,
            "quadrature weights have no transpose (CEED_EVAL_WEIGHT with CEED_TRANSPOSE)");
  const CeedSize u_need = t_mode == CEED_TRANSPOSE ? qpts : nodes, v_need = t_mode == CEED_TRANSPOSE ? nodes : qpts;
  if (eval_mode != CEED_EVAL_WEIGHT)
    CeedCheck(u != CEED_VECTOR_NONE && u->length >= u_need, basis->ceed, CEED_ERROR_DIMENSION, "input vector too short for basis apply");
  CeedCheck(v->length >= v_need, basis->ceed, CEED_ERROR_DIMENSION, "output vector too short for basis apply");
  return CEED_ERROR_SUCCESS;
}

int CeedBasisApply(CeedBasis basis, CeedInt num_elem, CeedTransposeMode t_mode, CeedEvalMode eval_mode, CeedVector u, CeedVector v) {
  CeedCall(basis_apply_check(basis, num_elem, t_mode, eval_mode, u, v));
  CeedCheck(basis->Apply, basis->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support BasisApply");
  return basis->Apply(basis, num_elem, t_mode, eval_mode, u, v);
}

int CeedBasisApplyAdd(CeedBasis basis, CeedInt num_elem, CeedTransposeMode t_mode, CeedEvalMode eval_mode, CeedVector u, CeedVector v) {
  CeedCall(basis_apply_check(basis, num_elem, t_mode, eval_mode, u, v));
  CeedCheck(eval_mode != CEED_EVAL_WEIGHT, basis->ceed, CEED_ERROR_UNSUPPORTED, "CEED_EVAL_WEIGHT has no ApplyAdd");
  CeedCheck(basis->ApplyAdd, basis->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support BasisApplyAdd");
  return basis->ApplyAdd(basis, num_elem, t_mode, eval_mode, u, v);
}

int CeedBasisDestroy(CeedBasis *basis) {
  if (!*basis || *basis == CEED_BASIS_NONE || --(*basis)->ref_count > 0) {
    *basis = NULL;
    return CEED_ERROR_SUCCESS;
  }
  if ((*basis)->Destroy) CeedCall((*basis)->Destroy(*basis));
  free((*basis)->interp_1d);
  free((*basis)->grad_1d);
  free((*basis)->q_ref_1d);
  free((*basis)->q_weight_1d);
  CeedCall(CeedDestroy(&(*basis)->ceed));
  free(*basis);
  *basis = NULL;
  return CEED_ERROR_SUCCESS;
}
