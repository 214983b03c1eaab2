/*
 * CeedOperator for /gpu/hip/mi355.
 *
 * Fused path: an operator made of one gallery pointwise operator
 * (MassApply, Vector3MassApply, Poisson3DApply, Vector3Poisson3DApply), an
 * active field with a standard restriction and a 3D tensor basis used for both
 * input and output, and passive geometric factors (EVAL_NONE), runs as one
 * fused kernel: v (+)= E^T B^T D B E u, whose E^T is completed without
 * atomics by an ordered gather over the transpose offset map (see the
 * CeedMi355OperatorCreateHex plan).  This is the path of the reference's
 * fused GPU backends (backends/hip-gen/ceed-hip-gen-operator.c:52-258),
 * rebuilt for gfx950.
 *
 * Fused path, run-time compiled: an operator whose pointwise function is a
 * user (or any non-inlined) QFunction, with scalar active fields on one
 * standard restriction and one 3D tensor basis (inputs INTERP and/or GRAD,
 * outputs INTERP and/or GRAD) and at most one passive EVAL_NONE input, runs
 * as one fused kernel too: the QFunction source is compiled with hipRTC into
 * the line-layout kernel (CeedMi355OperatorCreateHexJit), the role of the
 * reference's hip-gen backend (backends/hip-gen/ceed-hip-gen-operator-build.cpp).
 * CEED_MI355_FUSE_QFUNCTIONS=0 keeps such operators on the generic path.
 *
 * Generic path (any other operator): the staged apply of the reference
 * backends (backends/hip-ref/ceed-hip-ref-operator.c:258-638 and
 * backends/ref/ceed-ref-operator.c:145-500): E-vectors for all elements,
 * basis apply, the run-time compiled QFunction over every point, transposed
 * basis, and the ordered E^T.  Passive inputs are restricted again only when
 * their vector's state changes.
 */
#include <stdlib.h>
#include <string.h>

#include "ceed-mi355-backend.h"

typedef struct {
  bool     is_set_up, fused, jit;
  /* fused path */
  CeedMi355Operator hex;
  CeedSize          l_size;
  uint64_t          qdata_state;
  bool              qdata_loaded;
  CeedInt           qdata_field; /* input field of the geometric factors, -1: none */
  /* generic path */
  CeedInt  num_in, num_out, num_elem, Q;
  double  *e_in[CEED_FIELD_MAX], *q_in[CEED_FIELD_MAX];
  double  *e_out[CEED_FIELD_MAX], *q_out[CEED_FIELD_MAX];
  bool     q_in_is_e[CEED_FIELD_MAX], q_out_is_e[CEED_FIELD_MAX];
  uint64_t in_state[CEED_FIELD_MAX];
  bool     in_valid[CEED_FIELD_MAX];
} OperatorData;

/* E-vector of qdata in [elem][comp][qpt]: the field's L-vector itself when its
 * restriction has the backend strides, else a temporary restricted copy. */
static int qdata_evector(Ceed ceed, CeedElemRestriction rstr, CeedVector vec, const double **e, double **temp) {
  bool        backend;
  CeedScalar *dl;
  *temp = NULL;
  CeedCallBackend(CeedMi355VectorDevice(vec, false, &dl));
  CeedCallBackend(CeedElemRestrictionHasBackendStrides(rstr, &backend));
  if (backend) {
    *e = dl;
    return CEED_ERROR_SUCCESS;
  }
  RestrictionData *rd;
  CeedSize         e_size;
  CeedCallBackend(CeedElemRestrictionGetData(rstr, &rd));
  CeedCallBackend(CeedElemRestrictionGetEVectorSize(rstr, &e_size));
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355Malloc(sizeof(double) * (size_t)e_size, (void **)temp)));
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355RestrictionApply(rd->handle, CEED_NOTRANSPOSE, dl, *temp, NULL)));
  *e = *temp;
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* fused path                                                                */
/* ------------------------------------------------------------------------ */
static bool ends_with(const char *s, const char *suffix) {
  const size_t ls = s ? strlen(s) : 0, lx = strlen(suffix);
  return ls >= lx && !strcmp(s + ls - lx, suffix);
}

/* Gallery pointwise operators the fused kernel inlines, recognised by kernel
 * name and source file: this library's gallery header, or the reference
 * gallery's files when the backend is built into a libCEED library
 * (include/ceed/jit-source/gallery/ceed-{massapply,vectormassapply,
 * poisson3dapply,vectorpoisson3dapply}.h). */
static int fused_kind(const char *path, const char *name, int *qf_kind, CeedInt *num_comp) {
  *qf_kind = -1;
  if (!path || !name) return 0;
  const bool own = ends_with(path, "ceed-mi355-gallery.h");
  if (!own && !ends_with(path, "ceed-massapply.h") && !ends_with(path, "ceed-vectormassapply.h") && !ends_with(path, "ceed-poisson3dapply.h") &&
      !ends_with(path, "ceed-vectorpoisson3dapply.h"))
    return 0;
  if (!strcmp(name, "MassApply")) *qf_kind = CEED_MI355_QF_MASS, *num_comp = 1;
  else if (!strcmp(name, "Vector3MassApply")) *qf_kind = CEED_MI355_QF_MASS, *num_comp = 3;
  else if (!strcmp(name, "Poisson3DApply")) *qf_kind = CEED_MI355_QF_DIFFUSION, *num_comp = 1;
  else if (!strcmp(name, "Vector3Poisson3DApply")) *qf_kind = CEED_MI355_QF_DIFFUSION, *num_comp = 3;
  return 0;
}

/* Build the fused plan when the operator has the supported shape; leaves
 * d->fused false otherwise. */
static int try_fused(CeedOperator op, OperatorData *d) {
  Ceed                ceed;
  CeedQFunction       qf;
  CeedInt             num_in, num_out, num_comp = 0, dim, P, Q, rstr_comp, comp_stride, elem_size;
  CeedOperatorField  *in, *out;
  CeedQFunctionContext ctx;
  const char         *path, *name;
  int                 kind;
  CeedElemRestriction ru = NULL, rv = NULL, rq = NULL;
  CeedBasis           bu = NULL, bv = NULL;
  CeedVector          vu = NULL, vv = NULL, vq = NULL;
  CeedRestrictionType type;
  int                 ok = 0;

  CeedCallBackend(CeedOperatorGetCeed(op, &ceed));
  CeedCallBackend(CeedOperatorGetQFunction(op, &qf));
  CeedCallBackend(CeedQFunctionGetSourcePath(qf, &path));
  CeedCallBackend(CeedQFunctionGetKernelName(qf, &name));
  CeedCallBackend(CeedQFunctionGetContext(qf, &ctx));
  CeedCallBackend(fused_kind(path, name, &kind, &num_comp));
  CeedCallBackend(CeedOperatorGetFields(op, &num_in, &in, &num_out, &out));
  if (kind < 0 || ctx || num_in != 2 || num_out != 1) goto done;
  CeedCallBackend(CeedOperatorFieldGetElemRestriction(in[0], &ru));
  CeedCallBackend(CeedOperatorFieldGetBasis(in[0], &bu));
  CeedCallBackend(CeedOperatorFieldGetVector(in[0], &vu));
  CeedCallBackend(CeedOperatorFieldGetElemRestriction(in[1], &rq));
  CeedCallBackend(CeedOperatorFieldGetVector(in[1], &vq));
  CeedCallBackend(CeedOperatorFieldGetElemRestriction(out[0], &rv));
  CeedCallBackend(CeedOperatorFieldGetBasis(out[0], &bv));
  CeedCallBackend(CeedOperatorFieldGetVector(out[0], &vv));
  if (vu != CEED_VECTOR_ACTIVE || vv != CEED_VECTOR_ACTIVE || ru != rv || bu != bv) goto done;
  if (vq == CEED_VECTOR_ACTIVE || vq == CEED_VECTOR_NONE) goto done;
  CeedCallBackend(CeedElemRestrictionGetType(ru, &type));
  if (type != CEED_RESTRICTION_STANDARD) goto done;
  CeedCallBackend(CeedBasisGetDimension(bu, &dim));
  CeedCallBackend(CeedBasisGetNumNodes1D(bu, &P));
  CeedCallBackend(CeedBasisGetNumQuadraturePoints1D(bu, &Q));
  CeedCallBackend(CeedElemRestrictionGetNumComponents(ru, &rstr_comp));
  if (dim != 3 || rstr_comp != num_comp) goto done;
  {
    CeedInt q_comp, q_elem;
    CeedCallBackend(CeedElemRestrictionGetNumComponents(rq, &q_comp));
    CeedCallBackend(CeedElemRestrictionGetElementSize(rq, &q_elem));
    if (q_comp != (kind == CEED_MI355_QF_MASS ? 1 : 6) || q_elem != Q * Q * Q) goto done;
  }
  {
    const CeedScalar *interp, *grad;
    const CeedInt    *offsets;
    CeedInt           num_elem;
    CeedCallBackend(CeedBasisGetInterp1D(bu, &interp));
    CeedCallBackend(CeedBasisGetGrad1D(bu, &grad));
    CeedCallBackend(CeedElemRestrictionGetNumElements(ru, &num_elem));
    CeedCallBackend(CeedElemRestrictionGetElementSize(ru, &elem_size));
    CeedCallBackend(CeedElemRestrictionGetCompStride(ru, &comp_stride));
    CeedCallBackend(CeedElemRestrictionGetLVectorSize(ru, &d->l_size));
    CeedCallBackend(CeedElemRestrictionGetOffsets(ru, CEED_MEM_HOST, &offsets));
    /* shapes the fused kernels are instantiated for; anything else takes the generic path */
    int ierr = CEED_MI355_ERROR_UNSUPPORTED;
    if (elem_size == P * P * P && P >= 2 && P <= 9 && (Q == P || Q == P + 1))
      ierr = CeedMi355OperatorCreateHex(kind, CEED_MI355_FP64, num_comp, num_elem, P, Q, interp, grad, offsets, comp_stride, d->l_size, 0,
                                        &d->hex);
    CeedCallBackend(CeedElemRestrictionRestoreOffsets(ru, &offsets));
    if (ierr != CEED_MI355_ERROR_UNSUPPORTED) {
      CeedCallBackend(CeedMi355Check(ceed, ierr));
      ok = 1;
    }
  }
done:
  d->fused       = ok;
  d->qdata_field = 1;
  CeedCallBackend(CeedElemRestrictionDestroy(&ru));
  CeedCallBackend(CeedElemRestrictionDestroy(&rv));
  CeedCallBackend(CeedElemRestrictionDestroy(&rq));
  CeedCallBackend(CeedBasisDestroy(&bu));
  CeedCallBackend(CeedBasisDestroy(&bv));
  CeedCallBackend(CeedVectorDestroy(&vu));
  CeedCallBackend(CeedVectorDestroy(&vv));
  CeedCallBackend(CeedVectorDestroy(&vq));
  CeedCallBackend(CeedQFunctionContextDestroy(&ctx));
  CeedCallBackend(CeedQFunctionDestroy(&qf));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* Fused plan with a run-time compiled QFunction, when the operator has the
 * supported shape; leaves d->fused false otherwise. */
static int try_fused_jit(CeedOperator op, OperatorData *d) {
  Ceed                 ceed;
  CeedQFunction        qf;
  CeedInt              num_in, num_out, q_comp = 0, qdata_field = -1, P = 0, Q = 0;
  CeedOperatorField   *in, *out;
  CeedQFunctionField  *qf_in, *qf_out;
  const char          *path, *name;
  CeedElemRestriction  ru = NULL;
  CeedBasis            bu = NULL;
  int32_t              in_kinds[3], out_kinds[2];
  int                  ok = 1;
  const char          *env = getenv("CEED_MI355_FUSE_QFUNCTIONS");

  d->fused = false;
  if (env && !strcmp(env, "0")) return CEED_ERROR_SUCCESS;
  CeedCallBackend(CeedOperatorGetCeed(op, &ceed));
  CeedCallBackend(CeedOperatorGetQFunction(op, &qf));
  CeedCallBackend(CeedQFunctionGetSourcePath(qf, &path));
  CeedCallBackend(CeedQFunctionGetKernelName(qf, &name));
  CeedCallBackend(CeedOperatorGetFields(op, &num_in, &in, &num_out, &out));
  CeedCallBackend(CeedQFunctionGetFields(qf, NULL, &qf_in, NULL, &qf_out));
  if (!path || !name || num_in < 1 || num_in > 3 || num_out < 1 || num_out > 2) ok = 0;
  for (CeedInt f = 0; ok && f < num_in + num_out; f++) {
    const bool          is_in = f < num_in;
    const CeedInt       i     = is_in ? f : f - num_in;
    CeedElemRestriction rstr;
    CeedBasis           basis;
    CeedVector          vec;
    CeedEvalMode        mode;
    CeedInt             size;
    CeedCallBackend(CeedOperatorFieldGetElemRestriction(is_in ? in[i] : out[i], &rstr));
    CeedCallBackend(CeedOperatorFieldGetBasis(is_in ? in[i] : out[i], &basis));
    CeedCallBackend(CeedOperatorFieldGetVector(is_in ? in[i] : out[i], &vec));
    CeedCallBackend(CeedQFunctionFieldGetEvalMode(is_in ? qf_in[i] : qf_out[i], &mode));
    CeedCallBackend(CeedQFunctionFieldGetSize(is_in ? qf_in[i] : qf_out[i], &size));
    if (vec == CEED_VECTOR_ACTIVE) {
      /* all active fields: the same restriction and basis, scalar */
      if ((mode != CEED_EVAL_INTERP && mode != CEED_EVAL_GRAD) || (ru && (rstr != ru || basis != bu)) || size != (mode == CEED_EVAL_GRAD ? 3 : 1))
        ok = 0;
      if (ok && !ru) {
        CeedCallBackend(CeedElemRestrictionReferenceCopy(rstr, &ru));
        CeedCallBackend(CeedBasisReferenceCopy(basis, &bu));
      }
      const int32_t kind = mode == CEED_EVAL_GRAD ? CEED_MI355_FIELD_GRAD : CEED_MI355_FIELD_INTERP;
      if (is_in) in_kinds[i] = kind;
      else out_kinds[i] = kind;
    } else if (is_in && vec != CEED_VECTOR_NONE && mode == CEED_EVAL_NONE && qdata_field < 0 && rstr != CEED_ELEMRESTRICTION_NONE) {
      in_kinds[i] = CEED_MI355_FIELD_PASSIVE;
      qdata_field = i;
      q_comp      = size;
    } else {
      ok = 0;
    }
    CeedCallBackend(CeedVectorDestroy(&vec));
    CeedCallBackend(CeedBasisDestroy(&basis));
    CeedCallBackend(CeedElemRestrictionDestroy(&rstr));
  }
  if (ok && ru) {
    CeedRestrictionType type;
    CeedInt             dim, nc, rc, elem_size, num_elem;
    CeedCallBackend(CeedElemRestrictionGetType(ru, &type));
    CeedCallBackend(CeedBasisGetDimension(bu, &dim));
    CeedCallBackend(CeedBasisGetNumComponents(bu, &nc));
    CeedCallBackend(CeedElemRestrictionGetNumComponents(ru, &rc));
    CeedCallBackend(CeedBasisGetNumNodes1D(bu, &P));
    CeedCallBackend(CeedBasisGetNumQuadraturePoints1D(bu, &Q));
    CeedCallBackend(CeedElemRestrictionGetElementSize(ru, &elem_size));
    CeedCallBackend(CeedElemRestrictionGetNumElements(ru, &num_elem));
    ok = type == CEED_RESTRICTION_STANDARD && dim == 3 && nc == 1 && rc == 1 && elem_size == P * P * P && P >= 2 && P <= 9 && Q >= P && Q <= 10;
    if (ok && qdata_field >= 0) {  /* the passive field: Q^3 points per element, its size per point */
      CeedElemRestriction rq;
      CeedInt             q_elem, q_nc, q_ne;
      CeedCallBackend(CeedOperatorFieldGetElemRestriction(in[qdata_field], &rq));
      CeedCallBackend(CeedElemRestrictionGetElementSize(rq, &q_elem));
      CeedCallBackend(CeedElemRestrictionGetNumComponents(rq, &q_nc));
      CeedCallBackend(CeedElemRestrictionGetNumElements(rq, &q_ne));
      ok = q_elem == Q * Q * Q && q_nc == q_comp && q_ne == num_elem;
      CeedCallBackend(CeedElemRestrictionDestroy(&rq));
    }
    if (ok) {
      const CeedScalar *interp, *grad;
      const CeedInt    *offsets;
      CeedCallBackend(CeedBasisGetInterp1D(bu, &interp));
      CeedCallBackend(CeedBasisGetGrad1D(bu, &grad));
      CeedCallBackend(CeedElemRestrictionGetLVectorSize(ru, &d->l_size));
      CeedCallBackend(CeedElemRestrictionGetOffsets(ru, CEED_MEM_HOST, &offsets));
      int ierr = CeedMi355OperatorCreateHexJit(path, name, num_in, in_kinds, num_out, out_kinds, q_comp, num_elem, P, Q, interp, grad, offsets,
                                               d->l_size, 0, &d->hex);
      CeedCallBackend(CeedElemRestrictionRestoreOffsets(ru, &offsets));
      if (ierr == CEED_MI355_ERROR_UNSUPPORTED) ok = 0;
      else CeedCallBackend(CeedMi355Check(ceed, ierr));
    }
  } else {
    ok = 0;
  }
  d->fused       = ok;
  d->jit         = ok;
  d->qdata_field = qdata_field;
  CeedCallBackend(CeedElemRestrictionDestroy(&ru));
  CeedCallBackend(CeedBasisDestroy(&bu));
  CeedCallBackend(CeedQFunctionDestroy(&qf));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* reload the geometric factors into the plan when their vector changed, and
 * (run-time compiled plans) hand over the current device context */
static int fused_sync_qdata(CeedOperator op, OperatorData *d) {
  Ceed                ceed;
  CeedOperatorField  *in;
  CeedElemRestriction rq;
  CeedVector          vq;
  uint64_t            state;
  CeedCallBackend(CeedOperatorGetCeed(op, &ceed));
  if (d->jit) {
    CeedQFunction qf;
    void         *ctx;
    CeedCallBackend(CeedOperatorGetQFunction(op, &qf));
    CeedCallBackend(CeedMi355QFunctionContextDevice(qf, &ctx));
    CeedCallBackend(CeedMi355Check(ceed, CeedMi355OperatorSetContext(d->hex, ctx)));
    CeedCallBackend(CeedQFunctionDestroy(&qf));
  }
  if (d->qdata_field < 0) return CeedDestroy(&ceed);
  CeedCallBackend(CeedOperatorGetFields(op, NULL, &in, NULL, NULL));
  CeedCallBackend(CeedOperatorFieldGetElemRestriction(in[d->qdata_field], &rq));
  CeedCallBackend(CeedOperatorFieldGetVector(in[d->qdata_field], &vq));
  CeedCallBackend(CeedVectorGetState(vq, &state));
  if (!d->qdata_loaded || state != d->qdata_state) {
    const double *e;
    double       *temp;
    CeedCallBackend(qdata_evector(ceed, rq, vq, &e, &temp));
    CeedCallBackend(CeedMi355Check(ceed, CeedMi355OperatorSetQData(d->hex, e, NULL)));
    if (temp) CeedMi355Free(temp);
    d->qdata_state  = state;
    d->qdata_loaded = true;
  }
  CeedCallBackend(CeedVectorDestroy(&vq));
  CeedCallBackend(CeedElemRestrictionDestroy(&rq));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* generic path                                                              */
/* ------------------------------------------------------------------------ */
static int field_size(CeedBasis basis, CeedElemRestriction rstr, CeedEvalMode mode, CeedInt *per_point) {
  CeedInt nc = 1, dim = 1;
  if (basis != CEED_BASIS_NONE) {
    CeedCallBackend(CeedBasisGetNumComponents(basis, &nc));
    CeedCallBackend(CeedBasisGetDimension(basis, &dim));
  } else if (rstr != CEED_ELEMRESTRICTION_NONE) {
    CeedCallBackend(CeedElemRestrictionGetNumComponents(rstr, &nc));
  }
  *per_point = mode == CEED_EVAL_GRAD ? nc * dim : (mode == CEED_EVAL_WEIGHT ? 1 : nc);
  return CEED_ERROR_SUCCESS;
}

static int generic_setup(CeedOperator op, OperatorData *d) {
  Ceed               ceed;
  CeedQFunction      qf;
  CeedInt            num_in, num_out;
  CeedOperatorField *in, *out;
  CeedQFunctionField *qf_in, *qf_out;
  CeedCallBackend(CeedOperatorGetCeed(op, &ceed));
  CeedCallBackend(CeedOperatorGetQFunction(op, &qf));
  CeedCallBackend(CeedOperatorGetFields(op, &num_in, &in, &num_out, &out));
  CeedCallBackend(CeedQFunctionGetFields(qf, NULL, &qf_in, NULL, &qf_out));
  CeedCallBackend(CeedOperatorGetNumElements(op, &d->num_elem));
  CeedCallBackend(CeedOperatorGetNumQuadraturePoints(op, &d->Q));
  d->num_in  = num_in;
  d->num_out = num_out;
  const size_t points = (size_t)d->num_elem * (size_t)d->Q;
  for (CeedInt f = 0; f < num_in + num_out; f++) {
    const bool          is_in = f < num_in;
    const CeedInt       i     = is_in ? f : f - num_in;
    CeedElemRestriction rstr;
    CeedBasis           basis;
    CeedEvalMode        mode;
    CeedInt             per_point = 0;
    CeedCallBackend(CeedOperatorFieldGetElemRestriction(is_in ? in[i] : out[i], &rstr));
    CeedCallBackend(CeedOperatorFieldGetBasis(is_in ? in[i] : out[i], &basis));
    CeedCallBackend(CeedQFunctionFieldGetEvalMode(is_in ? qf_in[i] : qf_out[i], &mode));
    CeedCallBackend(field_size(basis, rstr, mode, &per_point));
    double **e = is_in ? &d->e_in[i] : &d->e_out[i], **q = is_in ? &d->q_in[i] : &d->q_out[i];
    bool    *same = is_in ? &d->q_in_is_e[i] : &d->q_out_is_e[i];
    if (rstr != CEED_ELEMRESTRICTION_NONE) {
      CeedSize e_size;
      CeedCallBackend(CeedElemRestrictionGetEVectorSize(rstr, &e_size));
      CeedCallBackend(CeedMi355Check(ceed, CeedMi355Malloc(sizeof(double) * (size_t)(e_size > 0 ? e_size : 1), (void **)e)));
    }
    if (mode == CEED_EVAL_NONE) {
      *q    = *e;
      *same = true;
    } else {
      CeedCallBackend(CeedMi355Check(ceed, CeedMi355Malloc(sizeof(double) * (points * (size_t)per_point + 1), (void **)q)));
    }
    if (is_in && mode == CEED_EVAL_WEIGHT) {
      BasisData *bd;
      CeedCallBackend(CeedBasisGetData(basis, &bd));
      CeedCallBackend(CeedMi355Check(ceed, CeedMi355BasisApply(bd->handle, d->num_elem, CEED_NOTRANSPOSE, CEED_EVAL_WEIGHT, 0, NULL, *q, NULL)));
    }
    CeedCallBackend(CeedElemRestrictionDestroy(&rstr));
    CeedCallBackend(CeedBasisDestroy(&basis));
  }
  CeedCallBackend(CeedQFunctionDestroy(&qf));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int generic_apply_add(CeedOperator op, OperatorData *d, CeedVector in_vec, CeedVector out_vec) {
  Ceed               ceed;
  CeedQFunction      qf;
  CeedOperatorField *in, *out;
  CeedQFunctionField *qf_in, *qf_out;
  CeedMi355QFunction kernel;
  void              *ctx;
  CeedCallBackend(CeedOperatorGetCeed(op, &ceed));
  CeedCallBackend(CeedOperatorGetQFunction(op, &qf));
  CeedCallBackend(CeedOperatorGetFields(op, NULL, &in, NULL, &out));
  CeedCallBackend(CeedQFunctionGetFields(qf, NULL, &qf_in, NULL, &qf_out));
  CeedCallBackend(CeedMi355QFunctionKernel(qf, &kernel));
  CeedCallBackend(CeedMi355QFunctionContextDevice(qf, &ctx));

  /* inputs: E then B */
  for (CeedInt i = 0; i < d->num_in; i++) {
    CeedElemRestriction rstr;
    CeedBasis           basis;
    CeedVector          vec;
    CeedEvalMode        mode;
    CeedCallBackend(CeedQFunctionFieldGetEvalMode(qf_in[i], &mode));
    if (mode == CEED_EVAL_WEIGHT) continue; /* filled once at set-up */
    CeedCallBackend(CeedOperatorFieldGetElemRestriction(in[i], &rstr));
    CeedCallBackend(CeedOperatorFieldGetBasis(in[i], &basis));
    CeedCallBackend(CeedOperatorFieldGetVector(in[i], &vec));
    const bool active = vec == CEED_VECTOR_ACTIVE;
    CeedVector l_vec  = active ? in_vec : vec;
    bool       fresh  = true;
    if (!active) {
      uint64_t state;
      CeedCallBackend(CeedVectorGetState(l_vec, &state));
      fresh          = !d->in_valid[i] || state != d->in_state[i];
      d->in_state[i] = state;
      d->in_valid[i] = true;
    }
    if (fresh) {
      RestrictionData *rd;
      CeedScalar      *dl;
      CeedCallBackend(CeedElemRestrictionGetData(rstr, &rd));
      CeedCallBackend(CeedMi355VectorDevice(l_vec, false, &dl));
      CeedCallBackend(CeedMi355Check(ceed, CeedMi355RestrictionApply(rd->handle, CEED_NOTRANSPOSE, dl, d->e_in[i], NULL)));
      if (mode != CEED_EVAL_NONE) {
        BasisData *bd;
        CeedCallBackend(CeedBasisGetData(basis, &bd));
        CeedCallBackend(
            CeedMi355Check(ceed, CeedMi355BasisApply(bd->handle, d->num_elem, CEED_NOTRANSPOSE, mode, 0, d->e_in[i], d->q_in[i], NULL)));
      }
    }
    CeedCallBackend(CeedVectorDestroy(&vec));
    CeedCallBackend(CeedBasisDestroy(&basis));
    CeedCallBackend(CeedElemRestrictionDestroy(&rstr));
  }

  /* D at every quadrature point */
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355QFunctionApply(kernel, d->num_elem, d->Q, ctx, (const double *const *)d->q_in, d->q_out, NULL)));

  /* outputs: B^T then E^T (summed into the L-vector) */
  for (CeedInt i = 0; i < d->num_out; i++) {
    CeedElemRestriction rstr;
    CeedBasis           basis;
    CeedVector          vec;
    CeedEvalMode        mode;
    RestrictionData    *rd;
    CeedScalar         *dl;
    CeedCallBackend(CeedQFunctionFieldGetEvalMode(qf_out[i], &mode));
    CeedCallBackend(CeedOperatorFieldGetElemRestriction(out[i], &rstr));
    CeedCallBackend(CeedOperatorFieldGetBasis(out[i], &basis));
    CeedCallBackend(CeedOperatorFieldGetVector(out[i], &vec));
    if (mode != CEED_EVAL_NONE) {
      BasisData *bd;
      CeedCallBackend(CeedBasisGetData(basis, &bd));
      CeedCallBackend(CeedMi355Check(ceed, CeedMi355BasisApply(bd->handle, d->num_elem, CEED_TRANSPOSE, mode, 0, d->q_out[i], d->e_out[i], NULL)));
    }
    CeedVector l_vec = vec == CEED_VECTOR_ACTIVE ? out_vec : vec;
    if (l_vec != CEED_VECTOR_NONE) {
      CeedCallBackend(CeedElemRestrictionGetData(rstr, &rd));
      CeedCallBackend(CeedVectorGetArray(l_vec, CEED_MEM_DEVICE, &dl));
      CeedCallBackend(CeedMi355Check(ceed, CeedMi355RestrictionApply(rd->handle, CEED_TRANSPOSE, d->e_out[i], dl, NULL)));
      CeedCallBackend(CeedVectorRestoreArray(l_vec, &dl));
    }
    CeedCallBackend(CeedVectorDestroy(&vec));
    CeedCallBackend(CeedBasisDestroy(&basis));
    CeedCallBackend(CeedElemRestrictionDestroy(&rstr));
  }
  CeedCallBackend(CeedQFunctionDestroy(&qf));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* slots                                                                     */
/* ------------------------------------------------------------------------ */
static int setup(CeedOperator op, OperatorData *d) {
  if (d->is_set_up) return CEED_ERROR_SUCCESS;
  CeedCallBackend(try_fused(op, d));
  if (!d->fused) CeedCallBackend(try_fused_jit(op, d));
  if (!d->fused) CeedCallBackend(generic_setup(op, d));
  d->is_set_up = true;
  return CEED_ERROR_SUCCESS;
}

static int OperatorApplyAdd(CeedOperator op, CeedVector in, CeedVector out, CeedRequest *request) {
  OperatorData *d;
  Ceed          ceed;
  (void)request;
  CeedCallBackend(CeedOperatorGetData(op, &d));
  CeedCallBackend(setup(op, d));
  if (!d->fused) return generic_apply_add(op, d, in, out);
  CeedScalar *du, *dv;
  CeedCallBackend(CeedOperatorGetCeed(op, &ceed));
  CeedCallBackend(fused_sync_qdata(op, d));
  CeedCallBackend(CeedMi355VectorDevice(in, false, &du));
  CeedCallBackend(CeedVectorGetArray(out, CEED_MEM_DEVICE, &dv));
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355OperatorApplyAdd(d->hex, du, dv, NULL)));
  CeedCallBackend(CeedVectorRestoreArray(out, &dv));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int OperatorApply(CeedOperator op, CeedVector in, CeedVector out, CeedRequest *request) {
  OperatorData *d;
  Ceed          ceed;
  CeedSize      out_len;
  CeedCallBackend(CeedOperatorGetData(op, &d));
  CeedCallBackend(setup(op, d));
  if (out != CEED_VECTOR_NONE) CeedCallBackend(CeedVectorGetLength(out, &out_len));
  /* the fused plan overwrites its whole L-vector; a longer output (or the
   * generic path) is zeroed first and accumulated into */
  if (!d->fused || out == CEED_VECTOR_NONE || out_len != d->l_size) {
    if (out != CEED_VECTOR_NONE) CeedCallBackend(CeedVectorSetValue(out, 0.0));
    return OperatorApplyAdd(op, in, out, request);
  }
  CeedScalar *du, *dv;
  CeedCallBackend(CeedOperatorGetCeed(op, &ceed));
  CeedCallBackend(fused_sync_qdata(op, d));
  CeedCallBackend(CeedMi355VectorDevice(in, false, &du));
  CeedCallBackend(CeedVectorGetArrayWrite(out, CEED_MEM_DEVICE, &dv));
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355OperatorApply(d->hex, du, dv, NULL)));
  CeedCallBackend(CeedVectorRestoreArray(out, &dv));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

/* Diagonal of the operator (Jacobi preconditioning): the fused plan's
 * diagonal kernel; the staged path has no diagonal assembly. */
static int assemble_diagonal(CeedOperator op, CeedVector assembled, bool add) {
  OperatorData *d;
  Ceed          ceed;
  CeedScalar   *da;
  CeedCallBackend(CeedOperatorGetData(op, &d));
  CeedCallBackend(setup(op, d));
  CeedCallBackend(CeedOperatorGetCeed(op, &ceed));
  CeedCheck(d->fused && !d->jit, ceed, CEED_ERROR_UNSUPPORTED, "diagonal assembly needs a gallery mass or diffusion operator on 3D tensor hexes");
  CeedCallBackend(CeedVectorGetArray(assembled, CEED_MEM_DEVICE, &da));
  CeedCallBackend(CeedMi355Check(ceed, CeedMi355OperatorAssembleDiagonal(d->hex, da, add, NULL)));
  CeedCallBackend(CeedVectorRestoreArray(assembled, &da));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}

static int OperatorLinearAssembleDiagonal(CeedOperator op, CeedVector assembled, CeedRequest *request) {
  (void)request;
  return assemble_diagonal(op, assembled, false);
}

static int OperatorLinearAssembleAddDiagonal(CeedOperator op, CeedVector assembled, CeedRequest *request) {
  (void)request;
  return assemble_diagonal(op, assembled, true);
}

static int OperatorDestroy(CeedOperator op) {
  OperatorData *d;
  CeedCallBackend(CeedOperatorGetData(op, &d));
  if (d->hex) CeedMi355OperatorDestroy(&d->hex);
  for (CeedInt i = 0; i < CEED_FIELD_MAX; i++) {
    if (d->q_in[i] && !d->q_in_is_e[i]) CeedMi355Free(d->q_in[i]);
    if (d->e_in[i]) CeedMi355Free(d->e_in[i]);
    if (d->q_out[i] && !d->q_out_is_e[i]) CeedMi355Free(d->q_out[i]);
    if (d->e_out[i]) CeedMi355Free(d->e_out[i]);
  }
  free(d);
  return CEED_ERROR_SUCCESS;
}

int CeedOperatorCreate_Mi355(CeedOperator op) {
  Ceed ceed;
  CeedCallBackend(CeedOperatorGetCeed(op, &ceed));
  OperatorData *d = calloc(1, sizeof(OperatorData));
  CeedCheck(d, ceed, CEED_ERROR_MAJOR, "host allocation failed");
  CeedCallBackend(CeedOperatorSetData(op, d));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Operator", op, "Apply", OperatorApply));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Operator", op, "ApplyAdd", OperatorApplyAdd));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Operator", op, "LinearAssembleDiagonal", OperatorLinearAssembleDiagonal));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Operator", op, "LinearAssembleAddDiagonal", OperatorLinearAssembleAddDiagonal));
  CeedCallBackend(CeedSetBackendFunction(ceed, "Operator", op, "Destroy", OperatorDestroy));
  CeedCallBackend(CeedDestroy(&ceed));
  return CEED_ERROR_SUCCESS;
}
