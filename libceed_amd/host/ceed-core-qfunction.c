/*
 * CeedQFunction and CeedQFunctionContext interfaces, and the gallery of
 * QFunctions available by name, following interface/ceed-qfunction.c
 * (create :642-676, by name :689, apply :1117-1127) and
 * interface/ceed-qfunctioncontext.c.  A QFunction records its user function,
 * the "file:function" location of its source (compiled at run time for the
 * GPU) and its fields.
 */
#include <stdlib.h>
#include <string.h>

#include "ceed-core.h"
#include "ceed/jit-source/gallery/ceed-mi355-gallery.h"

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */
int CeedQFunctionContextCreate(Ceed ceed, CeedQFunctionContext *ctx) {
  CeedCall(CeedCalloc(1, ctx));
  CeedCall(CeedReferenceCopy(ceed, &(*ctx)->ceed));
  (*ctx)->ref_count = 1;
  if (ceed->QFunctionContextCreate) CeedCall(ceed->QFunctionContextCreate(*ctx));
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextReferenceCopy(CeedQFunctionContext ctx, CeedQFunctionContext *ctx_copy) {
  if (ctx) ctx->ref_count++;
  if (*ctx_copy) CeedCall(CeedQFunctionContextDestroy(ctx_copy));
  *ctx_copy = ctx;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextGetCeed(CeedQFunctionContext ctx, Ceed *ceed) {
  *ceed = NULL;
  return CeedReferenceCopy(ctx->ceed, ceed);
}

/* Data access: the backend stores the bytes (slots SetData, GetData, ...);
 * the core keeps the size, the state (odd while a write access is out) and
 * the reader count, as interface/ceed-qfunctioncontext.c:643-820 does. */
int CeedQFunctionContextHasValidData(CeedQFunctionContext ctx, bool *has_valid_data) {
  CeedCheck(ctx->HasValidData, ctx->ceed, CEED_ERROR_UNSUPPORTED, "backend does not report whether a context holds data");
  return ctx->HasValidData(ctx, has_valid_data);
}

int CeedQFunctionContextHasBorrowedDataOfType(CeedQFunctionContext ctx, CeedMemType mem_type, bool *has_borrowed_data_of_type) {
  CeedCheck(ctx->HasBorrowedDataOfType, ctx->ceed, CEED_ERROR_UNSUPPORTED, "backend does not report borrowed context data");
  return ctx->HasBorrowedDataOfType(ctx, mem_type, has_borrowed_data_of_type);
}

static int ctx_check_valid(CeedQFunctionContext ctx, const char *what) {
  bool valid;
  CeedCall(CeedQFunctionContextHasValidData(ctx, &valid));
  CeedCheck(valid, ctx->ceed, CEED_ERROR_BACKEND, "%s: the context holds no data yet (set it with CeedQFunctionContextSetData)", what);
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextSetData(CeedQFunctionContext ctx, CeedMemType mem_type, CeedCopyMode copy_mode, size_t size, void *data) {
  CeedCheck(ctx->SetData, ctx->ceed, CEED_ERROR_UNSUPPORTED, "backend does not store context data");
  CeedCheck(ctx->state % 2 == 0, ctx->ceed, CEED_ERROR_ACCESS, "context data access already granted");
  ctx->size = size;
  CeedCall(ctx->SetData(ctx, mem_type, copy_mode, data));
  ctx->state += 2;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextTakeData(CeedQFunctionContext ctx, CeedMemType mem_type, void *data) {
  bool  borrowed;
  void *taken = NULL;
  CeedCall(ctx_check_valid(ctx, "CeedQFunctionContextTakeData"));
  CeedCheck(ctx->TakeData, ctx->ceed, CEED_ERROR_UNSUPPORTED, "backend does not hand context data back");
  CeedCheck(ctx->state % 2 == 0, ctx->ceed, CEED_ERROR_ACCESS, "context data access already granted");
  CeedCall(CeedQFunctionContextHasBorrowedDataOfType(ctx, mem_type, &borrowed));
  CeedCheck(borrowed, ctx->ceed, CEED_ERROR_BACKEND, "CeedQFunctionContextTakeData: no data of this memory type was lent with CEED_USE_POINTER");
  CeedCall(ctx->TakeData(ctx, mem_type, &taken));
  if (data) *(void **)data = taken;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextGetData(CeedQFunctionContext ctx, CeedMemType mem_type, void *data) {
  CeedCheck(ctx->GetData, ctx->ceed, CEED_ERROR_UNSUPPORTED, "backend does not hand out context data");
  CeedCheck(ctx->state % 2 == 0, ctx->ceed, CEED_ERROR_ACCESS, "context data access already granted");
  CeedCheck(ctx->num_readers == 0, ctx->ceed, CEED_ERROR_ACCESS, "context data is being read");
  CeedCall(ctx_check_valid(ctx, "CeedQFunctionContextGetData"));
  CeedCall(ctx->GetData(ctx, mem_type, data));
  ctx->state++;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextGetDataRead(CeedQFunctionContext ctx, CeedMemType mem_type, void *data) {
  CeedCheck(ctx->GetDataRead, ctx->ceed, CEED_ERROR_UNSUPPORTED, "backend does not hand out context data");
  CeedCheck(ctx->state % 2 == 0, ctx->ceed, CEED_ERROR_ACCESS, "context data access already granted");
  CeedCall(ctx_check_valid(ctx, "CeedQFunctionContextGetDataRead"));
  CeedCall(ctx->GetDataRead(ctx, mem_type, data));
  ctx->num_readers++;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextRestoreData(CeedQFunctionContext ctx, void *data) {
  CeedCheck(ctx->state % 2 == 1, ctx->ceed, CEED_ERROR_ACCESS, "context data access was not granted");
  if (ctx->RestoreData) CeedCall(ctx->RestoreData(ctx));
  *(void **)data = NULL;
  ctx->state++;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextRestoreDataRead(CeedQFunctionContext ctx, void *data) {
  CeedCheck(ctx->num_readers > 0, ctx->ceed, CEED_ERROR_ACCESS, "context read access was not granted");
  ctx->num_readers--;
  if (ctx->num_readers == 0 && ctx->RestoreDataRead) CeedCall(ctx->RestoreDataRead(ctx));
  *(void **)data = NULL;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextGetContextSize(CeedQFunctionContext ctx, size_t *ctx_size) {
  *ctx_size = ctx->size;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextGetState(CeedQFunctionContext ctx, uint64_t *state) {
  *state = ctx->state;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextGetBackendData(CeedQFunctionContext ctx, void *data) {
  *(void **)data = ctx->backend_data;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextSetBackendData(CeedQFunctionContext ctx, void *data) {
  ctx->backend_data = data;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionContextDestroy(CeedQFunctionContext *ctx) {
  if (!*ctx || --(*ctx)->ref_count > 0) {
    *ctx = NULL;
    return CEED_ERROR_SUCCESS;
  }
  if ((*ctx)->Destroy) CeedCall((*ctx)->Destroy(*ctx));
  CeedCall(CeedDestroy(&(*ctx)->ceed));
  free(*ctx);
  *ctx = NULL;
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* QFunction                                                                 */
/* ------------------------------------------------------------------------ */
int CeedQFunctionCreateInterior(Ceed ceed, CeedInt vec_length, CeedQFunctionUser f, const char *source, CeedQFunction *qf) {
  CeedCheck(source, ceed, CEED_ERROR_INCOMPLETE, "missing QFunction source");
  const char *colon = strrchr(source, ':');
  CeedCheck(colon && colon != source && colon[1], ceed, CEED_ERROR_INCOMPLETE, "QFunction source must be \"file:function\", got %s", source);
  CeedCheck(ceed->QFunctionCreate, ceed, CEED_ERROR_UNSUPPORTED, "backend does not support QFunctionCreate");
  CeedCall(CeedCalloc(1, qf));
  CeedQFunction q = *qf;
  CeedCall(CeedReferenceCopy(ceed, &q->ceed));
  q->ref_count  = 1;
  q->vec_length = vec_length;
  q->function   = f;
  CeedCall(CeedCallocArray((size_t)(colon - source) + 1, 1, &q->source_path));
  memcpy(q->source_path, source, (size_t)(colon - source));
  CeedCall(CeedStringAllocCopy(colon + 1, &q->kernel_name));
  for (int i = 0; i < CEED_CORE_MAX_FIELDS; i++) {
    q->in_ptr[i]  = &q->in[i];
    q->out_ptr[i] = &q->out[i];
  }
  CeedCall(ceed->QFunctionCreate(q));
  return CEED_ERROR_SUCCESS;
}

typedef struct {
  const char       *name;
  CeedQFunctionUser f;
  int               num_in, num_out;
  struct {
    const char  *name;
    CeedInt      size;
    CeedEvalMode mode;
  } fields[3];
  CeedSize flops;
} GalleryEntry;

/* name, function, inputs, outputs, fields (inputs then outputs), flops per point */
static const GalleryEntry gallery[] = {
    {"Mass1DBuild", Mass1DBuild, 2, 1, {{"dx", 1, CEED_EVAL_GRAD}, {"weights", 1, CEED_EVAL_WEIGHT}, {"qdata", 1, CEED_EVAL_NONE}}, 1},
    {"Mass2DBuild", Mass2DBuild, 2, 1, {{"dx", 4, CEED_EVAL_GRAD}, {"weights", 1, CEED_EVAL_WEIGHT}, {"qdata", 1, CEED_EVAL_NONE}}, 4},
    {"Mass3DBuild", Mass3DBuild, 2, 1, {{"dx", 9, CEED_EVAL_GRAD}, {"weights", 1, CEED_EVAL_WEIGHT}, {"qdata", 1, CEED_EVAL_NONE}}, 15},
    {"MassApply", MassApply, 2, 1, {{"u", 1, CEED_EVAL_INTERP}, {"qdata", 1, CEED_EVAL_NONE}, {"v", 1, CEED_EVAL_INTERP}}, 1},
    {"Vector3MassApply", Vector3MassApply, 2, 1, {{"u", 3, CEED_EVAL_INTERP}, {"qdata", 1, CEED_EVAL_NONE}, {"v", 3, CEED_EVAL_INTERP}}, 3},
    {"Poisson1DBuild", Poisson1DBuild, 2, 1, {{"dx", 1, CEED_EVAL_GRAD}, {"weights", 1, CEED_EVAL_WEIGHT}, {"qdata", 1, CEED_EVAL_NONE}}, 1},
    {"Poisson2DBuild", Poisson2DBuild, 2, 1, {{"dx", 4, CEED_EVAL_GRAD}, {"weights", 1, CEED_EVAL_WEIGHT}, {"qdata", 3, CEED_EVAL_NONE}}, 12},
    {"Poisson3DBuild", Poisson3DBuild, 2, 1, {{"dx", 9, CEED_EVAL_GRAD}, {"weights", 1, CEED_EVAL_WEIGHT}, {"qdata", 6, CEED_EVAL_NONE}}, 69},
    {"Poisson1DApply", Poisson1DApply, 2, 1, {{"du", 1, CEED_EVAL_GRAD}, {"qdata", 1, CEED_EVAL_NONE}, {"dv", 1, CEED_EVAL_GRAD}}, 1},
    {"Poisson2DApply", Poisson2DApply, 2, 1, {{"du", 2, CEED_EVAL_GRAD}, {"qdata", 3, CEED_EVAL_NONE}, {"dv", 2, CEED_EVAL_GRAD}}, 6},
    {"Poisson3DApply", Poisson3DApply, 2, 1, {{"du", 3, CEED_EVAL_GRAD}, {"qdata", 6, CEED_EVAL_NONE}, {"dv", 3, CEED_EVAL_GRAD}}, 15},
    {"Vector3Poisson1DApply", Vector3Poisson1DApply, 2, 1, {{"du", 3, CEED_EVAL_GRAD}, {"qdata", 1, CEED_EVAL_NONE}, {"dv", 3, CEED_EVAL_GRAD}}, 3},
    {"Vector3Poisson2DApply", Vector3Poisson2DApply, 2, 1, {{"du", 6, CEED_EVAL_GRAD}, {"qdata", 3, CEED_EVAL_NONE}, {"dv", 6, CEED_EVAL_GRAD}}, 18},
    {"Vector3Poisson3DApply", Vector3Poisson3DApply, 2, 1, {{"du", 9, CEED_EVAL_GRAD}, {"qdata", 6, CEED_EVAL_NONE}, {"dv", 9, CEED_EVAL_GRAD}}, 45},
};

int CeedQFunctionCreateInteriorByName(Ceed ceed, const char *name, CeedQFunction *qf) {
  const GalleryEntry *g = NULL;
  for (size_t i = 0; i < sizeof(gallery) / sizeof(gallery[0]); i++)
    if (!strcmp(gallery[i].name, name)) g = &gallery[i];
  CeedCheck(g, ceed, CEED_ERROR_UNSUPPORTED, "no gallery QFunction named %s", name);
  const char *root;
  char        source[4096];
  CeedCall(CeedCoreJitRoot(&root));
  snprintf(source, sizeof(source), "%s/ceed/jit-source/gallery/ceed-mi355-gallery.h:%s", root, name);
  CeedCall(CeedQFunctionCreateInterior(ceed, 1, g->f, source, qf));
  for (int i = 0; i < g->num_in; i++) CeedCall(CeedQFunctionAddInput(*qf, g->fields[i].name, g->fields[i].size, g->fields[i].mode));
  for (int i = 0; i < g->num_out; i++)
    CeedCall(CeedQFunctionAddOutput(*qf, g->fields[g->num_in + i].name, g->fields[g->num_in + i].size, g->fields[g->num_in + i].mode));
  (*qf)->user_flops = g->flops;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionReferenceCopy(CeedQFunction qf, CeedQFunction *qf_copy) {
  if (qf && qf != CEED_QFUNCTION_NONE) qf->ref_count++;
  if (*qf_copy && *qf_copy != CEED_QFUNCTION_NONE) CeedCall(CeedQFunctionDestroy(qf_copy));
  *qf_copy = qf;
  return CEED_ERROR_SUCCESS;
}

static int add_field(CeedQFunction qf, bool input, const char *field_name, CeedInt size, CeedEvalMode eval_mode) {
  CeedCheck(!qf->is_immutable, qf->ceed, CEED_ERROR_MAJOR, "fields cannot be added to a QFunction an operator already uses");
  CeedCheck(eval_mode != CEED_EVAL_WEIGHT || size == 1, qf->ceed, CEED_ERROR_DIMENSION, "a quadrature-weight field holds one value per point");
  CeedCheck(input || eval_mode != CEED_EVAL_WEIGHT, qf->ceed, CEED_ERROR_DIMENSION, "quadrature weights are an input only");
  CeedInt                           *n = input ? &qf->num_in : &qf->num_out;
  struct CeedQFunctionField_private *f = input ? qf->in : qf->out;
  for (CeedInt i = 0; i < qf->num_in; i++)
    CeedCheck(strcmp(qf->in[i].name, field_name), qf->ceed, CEED_ERROR_MINOR, "QFunction field names must be unique: %s", field_name);
  for (CeedInt i = 0; i < qf->num_out; i++)
    CeedCheck(strcmp(qf->out[i].name, field_name), qf->ceed, CEED_ERROR_MINOR, "QFunction field names must be unique: %s", field_name);
  CeedCheck(*n < CEED_CORE_MAX_FIELDS, qf->ceed, CEED_ERROR_DIMENSION, "too many QFunction fields");
  CeedCall(CeedStringAllocCopy(field_name, &f[*n].name));
  f[*n].size      = size;
  f[*n].eval_mode = eval_mode;
  (*n)++;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionAddInput(CeedQFunction qf, const char *field_name, CeedInt size, CeedEvalMode eval_mode) {
  return add_field(qf, true, field_name, size, eval_mode);
}

int CeedQFunctionAddOutput(CeedQFunction qf, const char *field_name, CeedInt size, CeedEvalMode eval_mode) {
  return add_field(qf, false, field_name, size, eval_mode);
}

int CeedQFunctionSetContext(CeedQFunction qf, CeedQFunctionContext ctx) { return CeedQFunctionContextReferenceCopy(ctx, &qf->ctx); }

int CeedQFunctionSetUserFlopsEstimate(CeedQFunction qf, CeedSize flops) {
  qf->user_flops = flops;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionGetCeed(CeedQFunction qf, Ceed *ceed) {
  *ceed = NULL;
  return CeedReferenceCopy(qf->ceed, ceed);
}

int CeedQFunctionGetFields(CeedQFunction qf, CeedInt *num_input_fields, CeedQFunctionField **input_fields, CeedInt *num_output_fields,
                           CeedQFunctionField **output_fields) {
  if (num_input_fields) *num_input_fields = qf->num_in;
  if (input_fields) *input_fields = qf->in_ptr;
  if (num_output_fields) *num_output_fields = qf->num_out;
  if (output_fields) *output_fields = qf->out_ptr;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionGetNumArgs(CeedQFunction qf, CeedInt *num_input_fields, CeedInt *num_output_fields) {
  if (num_input_fields) *num_input_fields = qf->num_in;
  if (num_output_fields) *num_output_fields = qf->num_out;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionFieldGetName(CeedQFunctionField f, const char **field_name) {
  *field_name = f->name;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionFieldGetSize(CeedQFunctionField f, CeedInt *size) {
  *size = f->size;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionFieldGetEvalMode(CeedQFunctionField f, CeedEvalMode *eval_mode) {
  *eval_mode = f->eval_mode;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionGetSourcePath(CeedQFunction qf, const char **source_path) {
  *source_path = qf->source_path;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionGetKernelName(CeedQFunction qf, const char **kernel_name) {
  *kernel_name = qf->kernel_name;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionGetContext(CeedQFunction qf, CeedQFunctionContext *ctx) {
  *ctx = NULL;
  if (qf->ctx) CeedCall(CeedQFunctionContextReferenceCopy(qf->ctx, ctx));
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionGetData(CeedQFunction qf, void *data) {
  *(void **)data = qf->data;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionSetData(CeedQFunction qf, void *data) {
  qf->data = data;
  return CEED_ERROR_SUCCESS;
}

int CeedQFunctionApply(CeedQFunction qf, CeedInt Q, CeedVector *u, CeedVector *v) {
  CeedCheck(qf->Apply, qf->ceed, CEED_ERROR_UNSUPPORTED, "backend does not support QFunctionApply");
  CeedCheck(Q % qf->vec_length == 0, qf->ceed, CEED_ERROR_DIMENSION, "number of points %d must be a multiple of %d", Q, qf->vec_length);
  qf->is_immutable = true;
  return qf->Apply(qf, Q, u, v);
}

int CeedQFunctionDestroy(CeedQFunction *qf) {
  if (!*qf || *qf == CEED_QFUNCTION_NONE || --(*qf)->ref_count > 0) {
    *qf = NULL;
    return CEED_ERROR_SUCCESS;
  }
  if ((*qf)->Destroy) CeedCall((*qf)->Destroy(*qf));
  for (CeedInt i = 0; i < (*qf)->num_in; i++) free((*qf)->in[i].name);
  for (CeedInt i = 0; i < (*qf)->num_out; i++) free((*qf)->out[i].name);
  free((*qf)->source_path);
  free((*qf)->kernel_name);
  CeedCall(CeedQFunctionContextDestroy(&(*qf)->ctx));
  CeedCall(CeedDestroy(&(*qf)->ceed));
  free(*qf);
  *qf = NULL;
  return CEED_ERROR_SUCCESS;
}
