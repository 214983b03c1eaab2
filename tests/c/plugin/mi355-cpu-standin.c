/*
 * CPU stand-in for the device ABI of include/ceed-mi355-kernels.h, for
 * testing the /gpu/hip/mi355 backend sources inside the reference's own
 * libCEED on a machine without a GPU.  "Device" memory is host memory; the
 * arithmetic is the oracle's (oracle/ceed_oracle.c, the restatement of
 * /cpu/self/ref/serial); user QFunctions are compiled from their source file
 * with the host C compiler and loaded with dlopen, where the real library
 * compiles them with hipRTC.  Run-time fused kernels (CeedMi355OperatorCreateHexJit)
 * report CEED_MI355_ERROR_UNSUPPORTED, so operators with user QFunctions take
 * the backend's staged path, which the stand-in covers call by call.
 *
 * STANDIN_CEED_INCLUDE names the directory holding ceed/types.h for the
 * QFunction compiles; STANDIN_LOG (optional) receives one line per entry
 * point called, so that a test can check which paths the backend took.
 */
#define _DEFAULT_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "ceed-mi355-kernels.h"
#include "ceed_oracle.h"

static char g_error[1024];

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
  return code;
}

static void trace(const char *name) {
  const char *path = getenv("STANDIN_LOG");
  if (!path) return;
  FILE *f = fopen(path, "a");
  if (f) {
    fprintf(f, "%s\n", name);
    fclose(f);
  }
}

const char *CeedMi355GetLastError(void) { return g_error; }
int         CeedMi355DeviceInit(int device_id) {
  trace("DeviceInit");
  return device_id == 0 ? CEED_MI355_SUCCESS : fail(CEED_MI355_ERROR_BACKEND, "stand-in has device 0 only");
}
int CeedMi355Malloc(size_t bytes, void **ptr) {
  *ptr = calloc(1, bytes ? bytes : 1);
  return *ptr ? CEED_MI355_SUCCESS : fail(CEED_MI355_ERROR_MAJOR, "allocation of %zu bytes failed", bytes);
}
int CeedMi355Free(void *ptr) {
  free(ptr);
  return CEED_MI355_SUCCESS;
}
int CeedMi355Memcpy(void *dst, const void *src, size_t bytes, void *stream) {
  (void)stream;
  if (bytes) memmove(dst, src, bytes);
  return CEED_MI355_SUCCESS;
}
int CeedMi355Memset(void *dst, int value, size_t bytes, void *stream) {
  (void)stream;
  memset(dst, value, bytes);
  return CEED_MI355_SUCCESS;
}
int CeedMi355StreamSynchronize(void *stream) {
  (void)stream;
  return CEED_MI355_SUCCESS;
}

/* ---- restriction ------------------------------------------------------- */
struct CeedMi355Restriction_private {
  int      strided, num_elem, elem_size, num_comp, comp_stride;
  int32_t *offsets, strides[3];
};

int CeedMi355RestrictionCreateOffset(int32_t num_elem, int32_t elem_size, int32_t num_comp, int32_t comp_stride, int64_t l_size,
                                     const int32_t *offsets_host, CeedMi355Restriction *rstr) {
  (void)l_size;
  trace("RestrictionCreateOffset");
  CeedMi355Restriction r = calloc(1, sizeof(*r));
  r->num_elem = num_elem, r->elem_size = elem_size, r->num_comp = num_comp, r->comp_stride = comp_stride;
  r->offsets = malloc(sizeof(int32_t) * ((size_t)num_elem * elem_size + 1));
  memcpy(r->offsets, offsets_host, sizeof(int32_t) * (size_t)num_elem * elem_size);
  *rstr = r;
  return CEED_MI355_SUCCESS;
}

int CeedMi355RestrictionCreateStrided(int32_t num_elem, int32_t elem_size, int32_t num_comp, int64_t l_size, const int32_t *strides_host,
                                      CeedMi355Restriction *rstr) {
  (void)l_size;
  trace("RestrictionCreateStrided");
  CeedMi355Restriction r = calloc(1, sizeof(*r));
  r->strided = 1, r->num_elem = num_elem, r->elem_size = elem_size, r->num_comp = num_comp;
  memcpy(r->strides, strides_host, sizeof(r->strides));
  *rstr = r;
  return CEED_MI355_SUCCESS;
}

int CeedMi355RestrictionApply(CeedMi355Restriction r, int t_mode, const double *u, double *v, void *stream) {
  (void)stream;
  trace(t_mode ? "RestrictionApply(T)" : "RestrictionApply");
  if (r->strided) orc_restriction_strided(r->num_elem, r->elem_size, r->num_comp, r->strides, t_mode, u, v);
  else orc_restriction_offset(r->num_elem, r->elem_size, r->num_comp, r->comp_stride, r->offsets, t_mode, u, v);
  return CEED_MI355_SUCCESS;
}

int CeedMi355RestrictionDestroy(CeedMi355Restriction *r) {
  if (*r) free((*r)->offsets);
  free(*r);
  *r = NULL;
  return CEED_MI355_SUCCESS;
}

/* ---- basis ------------------------------------------------------------- */
struct CeedMi355Basis_private {
  int     dim, num_comp, P, Q;
  double *interp, *grad, *q_weight;
};

int CeedMi355BasisCreateTensorH1(int32_t dim, int32_t num_comp, int32_t P_1d, int32_t Q_1d, const double *interp_1d_host,
                                 const double *grad_1d_host, const double *q_weight_1d_host, CeedMi355Basis *basis) {
  trace("BasisCreateTensorH1");
  CeedMi355Basis b = calloc(1, sizeof(*b));
  b->dim = dim, b->num_comp = num_comp, b->P = P_1d, b->Q = Q_1d;
  b->interp   = malloc(sizeof(double) * P_1d * Q_1d);
  b->grad     = malloc(sizeof(double) * P_1d * Q_1d);
  b->q_weight = malloc(sizeof(double) * Q_1d);
  memcpy(b->interp, interp_1d_host, sizeof(double) * P_1d * Q_1d);
  memcpy(b->grad, grad_1d_host, sizeof(double) * P_1d * Q_1d);
  memcpy(b->q_weight, q_weight_1d_host, sizeof(double) * Q_1d);
  *basis = b;
  return CEED_MI355_SUCCESS;
}

static int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

int CeedMi355BasisApply(CeedMi355Basis b, int32_t num_elem, int t_mode, int eval_mode, int add, const double *u, double *v, void *stream) {
  (void)stream;
  trace("BasisApply");
  const int    nn = ipow(b->P, b->dim), nq = ipow(b->Q, b->dim);
  const size_t per_node = (size_t)b->num_comp * nn;
  const size_t per_qpt  = eval_mode == ORC_EVAL_WEIGHT ? (size_t)nq : (size_t)b->num_comp * nq * (eval_mode == ORC_EVAL_GRAD ? b->dim : 1);
  for (int e = 0; e < num_elem; e++) {
    const double *ue = u ? u + (t_mode ? per_qpt : per_node) * e : NULL;
    double       *ve = v + (t_mode ? per_node : per_qpt) * e;
    if (orc_basis_apply(b->dim, b->num_comp, b->P, b->Q, b->interp, b->grad, b->q_weight, 1, t_mode, eval_mode, add, ue, ve))
      return fail(CEED_MI355_ERROR_BACKEND, "basis apply failed");
  }
  return CEED_MI355_SUCCESS;
}

int CeedMi355BasisDestroy(CeedMi355Basis *b) {
  if (*b) free((*b)->interp), free((*b)->grad), free((*b)->q_weight);
  free(*b);
  *b = NULL;
  return CEED_MI355_SUCCESS;
}

/* ---- vectors ----------------------------------------------------------- */
int CeedMi355VectorOp(int op, int scalar_type, int64_t n, double alpha, double beta, const void *xv, const void *yv, void *zv, void *stream) {
  (void)stream;
  trace("VectorOp");
  if (scalar_type != CEED_MI355_FP64) return fail(CEED_MI355_ERROR_UNSUPPORTED, "stand-in vectors are FP64");
  const double *x = xv, *y = yv;
  double       *z = zv;
  for (int64_t i = 0; i < n; i++) {
    switch (op) {
      case CEED_MI355_VEC_SET: z[i] = alpha; break;
      case CEED_MI355_VEC_SCALE: z[i] *= alpha; break;
      case CEED_MI355_VEC_AXPY: z[i] += alpha * x[i]; break;
      case CEED_MI355_VEC_AXPBY: z[i] = alpha * x[i] + beta * z[i]; break;
      case CEED_MI355_VEC_PMULT: z[i] = x[i] * y[i]; break;
      case CEED_MI355_VEC_RECIPROCAL:
        if (fabs(z[i]) > alpha) z[i] = 1.0 / z[i];
        break;
      default: return fail(CEED_MI355_ERROR_UNSUPPORTED, "vector op %d", op);
    }
  }
  return CEED_MI355_SUCCESS;
}

int CeedMi355VectorReduce(int op, int scalar_type, int64_t n, const void *xv, const void *yv, double *result, void *stream) {
  (void)stream;
  trace("VectorReduce");
  if (scalar_type != CEED_MI355_FP64) return fail(CEED_MI355_ERROR_UNSUPPORTED, "stand-in vectors are FP64");
  const double *x = xv, *y = yv;
  double        r = 0;
  for (int64_t i = 0; i < n; i++) {
    if (op == CEED_MI355_RED_DOT) r += x[i] * y[i];
    else if (op == CEED_MI355_RED_ABS_SUM) r += fabs(x[i]);
    else r = fmax(r, fabs(x[i]));
  }
  *result = r;
  return CEED_MI355_SUCCESS;
}

/* ---- QFunctions: compiled with the host compiler, loaded with dlopen ---- */
typedef int (*standin_qf)(void *ctx, int Q, const double *const *in, double *const *out);
struct CeedMi355QFunction_private {
  void      *handle;
  standin_qf f;
  int        num_in, num_out, in_sizes[16], out_sizes[16];
  char       so_path[256];
};

int CeedMi355QFunctionCompile(const char *source_path, const char *function_name, int32_t num_in, const int32_t *in_sizes, int32_t num_out,
                              const int32_t *out_sizes, CeedMi355QFunction *qf) {
  trace("QFunctionCompile");
  const char *inc = getenv("STANDIN_CEED_INCLUDE");
  char        dir[] = "/tmp/mi355-standin-XXXXXX", src[512], cmd[2048];
  if (!inc) return fail(CEED_MI355_ERROR_BACKEND, "STANDIN_CEED_INCLUDE is not set");
  if (!mkdtemp(dir)) return fail(CEED_MI355_ERROR_BACKEND, "mkdtemp failed");
  CeedMi355QFunction q = calloc(1, sizeof(*q));
  snprintf(src, sizeof(src), "%s/qf.c", dir);
  snprintf(q->so_path, sizeof(q->so_path), "%s/qf.so", dir);
  FILE *f = fopen(src, "w");
  if (!f) return fail(CEED_MI355_ERROR_BACKEND, "cannot write %s", src);
  fprintf(f, "#include <ceed/types.h>\n#include \"%s\"\n", source_path);
  fprintf(f, "int standin_entry(void *ctx, CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) { return %s(ctx, Q, in, out); }\n",
          function_name);
  fclose(f);
  snprintf(cmd, sizeof(cmd), "cc -O1 -shared -fPIC -std=c11 -I%s -o %s %s -lm", inc, q->so_path, src);
  if (system(cmd) != 0) return fail(CEED_MI355_ERROR_BACKEND, "QFunction compile failed: %s", cmd);
  q->handle = dlopen(q->so_path, RTLD_NOW | RTLD_LOCAL);
  if (!q->handle) return fail(CEED_MI355_ERROR_BACKEND, "dlopen: %s", dlerror());
  q->f = (standin_qf)dlsym(q->handle, "standin_entry");
  if (!q->f) return fail(CEED_MI355_ERROR_BACKEND, "dlsym: %s", dlerror());
  q->num_in = num_in, q->num_out = num_out;
  for (int i = 0; i < num_in; i++) q->in_sizes[i] = in_sizes[i];
  for (int i = 0; i < num_out; i++) q->out_sizes[i] = out_sizes[i];
  *qf = q;
  return CEED_MI355_SUCCESS;
}

/* field i, component k of point q of element e at in[i][(e*size_i + k)*Q + q]:
 * the QFunction sees one element's Q points at a time */
int CeedMi355QFunctionApply(CeedMi355QFunction qf, int64_t num_elem, int32_t Q, void *ctx, const double *const *in_host, double *const *out_host,
                            void *stream) {
  (void)stream;
  trace("QFunctionApply");
  const double *in[16];
  double       *out[16];
  for (int64_t e = 0; e < num_elem; e++) {
    for (int i = 0; i < qf->num_in; i++) in[i] = in_host[i] + (size_t)e * qf->in_sizes[i] * Q;
    for (int i = 0; i < qf->num_out; i++) out[i] = out_host[i] + (size_t)e * qf->out_sizes[i] * Q;
    if (qf->f(ctx, Q, in, out)) return fail(CEED_MI355_ERROR_BACKEND, "QFunction returned an error");
  }
  return CEED_MI355_SUCCESS;
}

int CeedMi355QFunctionDestroy(CeedMi355QFunction *qf) {
  if (*qf && (*qf)->handle) dlclose((*qf)->handle);
  if (*qf) unlink((*qf)->so_path);
  free(*qf);
  *qf = NULL;
  return CEED_MI355_SUCCESS;
}

/* ---- fused hex plans (gallery QFunctions), by the oracle's element loop -- */
struct CeedMi355Operator_private {
  int      qf, num_comp, num_elem, P, Q, comp_stride, q_comp;
  int64_t  l_size;
  int32_t *offsets;
  double  *interp, *grad, *q_weight, *q_data;
};

int CeedMi355OperatorCreateHex(int qf, int scalar_type, int32_t num_comp, int32_t num_elem, int32_t P_1d, int32_t Q_1d, const double *interp_1d_host,
                               const double *grad_1d_host, const int32_t *offsets_host, int32_t comp_stride, int64_t l_size, int flags,
                               CeedMi355Operator *op) {
  (void)flags;
  trace("OperatorCreateHex");
  if (scalar_type != CEED_MI355_FP64 || (qf != CEED_MI355_QF_MASS && qf != CEED_MI355_QF_DIFFUSION))
    return fail(CEED_MI355_ERROR_UNSUPPORTED, "stand-in plans: FP64 mass or diffusion");
  CeedMi355Operator o = calloc(1, sizeof(*o));
  o->qf = qf, o->num_comp = num_comp, o->num_elem = num_elem, o->P = P_1d, o->Q = Q_1d, o->comp_stride = comp_stride, o->l_size = l_size;
  o->q_comp = qf == CEED_MI355_QF_MASS ? 1 : 6;
  const size_t es = (size_t)P_1d * P_1d * P_1d, q3 = (size_t)Q_1d * Q_1d * Q_1d;
  o->offsets = malloc(sizeof(int32_t) * (num_elem * es + 1));
  memcpy(o->offsets, offsets_host, sizeof(int32_t) * num_elem * es);
  o->interp   = malloc(sizeof(double) * P_1d * Q_1d);
  o->grad     = malloc(sizeof(double) * P_1d * Q_1d);
  o->q_weight = calloc(Q_1d, sizeof(double));
  memcpy(o->interp, interp_1d_host, sizeof(double) * P_1d * Q_1d);
  memcpy(o->grad, grad_1d_host, sizeof(double) * P_1d * Q_1d);
  o->q_data = calloc((size_t)num_elem * o->q_comp * q3 + 1, sizeof(double));
  *op = o;
  return CEED_MI355_SUCCESS;
}

int CeedMi355OperatorCreateHexJit(const char *source_path, const char *function_name, int32_t num_in, const int32_t *in_kinds_host,
                                  int32_t num_out, const int32_t *out_kinds_host, int32_t q_comp, int32_t num_elem, int32_t P_1d, int32_t Q_1d,
                                  const double *interp_1d_host, const double *grad_1d_host, const int32_t *offsets_host, int64_t l_size, int flags,
                                  CeedMi355Operator *op) {
  (void)source_path, (void)function_name, (void)num_in, (void)in_kinds_host, (void)num_out, (void)out_kinds_host, (void)q_comp;
  (void)num_elem, (void)P_1d, (void)Q_1d, (void)interp_1d_host, (void)grad_1d_host, (void)offsets_host, (void)l_size, (void)flags;
  trace("OperatorCreateHexJit");
  *op = NULL;
  return fail(CEED_MI355_ERROR_UNSUPPORTED, "the stand-in has no run-time fused kernels");
}

int CeedMi355OperatorSetContext(CeedMi355Operator op, void *ctx_device) {
  (void)op, (void)ctx_device;
  return CEED_MI355_SUCCESS;
}

int CeedMi355OperatorSetQData(CeedMi355Operator op, const void *q_data, void *stream) {
  (void)stream;
  trace("OperatorSetQData");
  const size_t q3 = (size_t)op->Q * op->Q * op->Q;
  memcpy(op->q_data, q_data, sizeof(double) * op->num_elem * op->q_comp * q3);
  return CEED_MI355_SUCCESS;
}

int CeedMi355OperatorApplyAdd(CeedMi355Operator op, const void *u, void *v, void *stream) {
  (void)stream;
  trace("OperatorApplyAdd");
  const int     P = op->P, Q = op->Q, nc = op->num_comp, grad = op->qf == CEED_MI355_QF_DIFFUSION;
  const int32_t qstr[3] = {1, Q * Q * Q, op->q_comp * Q * Q * Q};
  orc_field     in[2], out[1];
  memset(in, 0, sizeof(in));
  memset(out, 0, sizeof(out));
  in[0].eval_mode = grad ? ORC_EVAL_GRAD : ORC_EVAL_INTERP;
  in[0].size = nc * (grad ? 3 : 1), in[0].num_comp = nc, in[0].elem_size = P * P * P, in[0].comp_stride = op->comp_stride;
  in[0].offsets = op->offsets, in[0].dim = 3, in[0].P = P, in[0].Q = Q, in[0].interp = op->interp, in[0].grad = op->grad;
  in[0].q_weight = op->q_weight, in[0].vec = (double *)u;
  in[1].eval_mode = ORC_EVAL_NONE, in[1].size = op->q_comp, in[1].num_comp = op->q_comp, in[1].elem_size = Q * Q * Q;
  memcpy(in[1].strides, qstr, sizeof(qstr));
  in[1].vec = op->q_data;
  out[0]     = in[0];
  out[0].vec = (double *)v;
  orc_qfunction f = op->qf == CEED_MI355_QF_MASS ? (nc == 3 ? orc_qf_vector3_mass_apply : orc_qf_mass_apply)
                                                 : (nc == 3 ? orc_qf_vector3_poisson3d_apply : orc_qf_poisson3d_apply);
  if (orc_operator_apply_add(op->num_elem, Q * Q * Q, 2, in, 1, out, f, NULL)) return fail(CEED_MI355_ERROR_BACKEND, "oracle apply failed");
  return CEED_MI355_SUCCESS;
}

int CeedMi355OperatorApply(CeedMi355Operator op, const void *u, void *v, void *stream) {
  trace("OperatorApply");
  memset(v, 0, sizeof(double) * op->l_size);
  return CeedMi355OperatorApplyAdd(op, u, v, stream);
}

int CeedMi355OperatorAssembleDiagonal(CeedMi355Operator op, void *diag, int add, void *stream) {
  (void)op, (void)diag, (void)add, (void)stream;
  return fail(CEED_MI355_ERROR_UNSUPPORTED, "no diagonal in the stand-in");
}

int CeedMi355OperatorDestroy(CeedMi355Operator *op) {
  if (*op) free((*op)->offsets), free((*op)->interp), free((*op)->grad), free((*op)->q_weight), free((*op)->q_data);
  free(*op);
  *op = NULL;
  return CEED_MI355_SUCCESS;
}
