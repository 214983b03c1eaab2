/*
 * Host core of the libCEED API: the Ceed context, the backend registry
 * (CeedRegister / CeedInit resource matching), backend slot installation
 * (CeedSetBackendFunction), error handlers and memory helpers.
 *
 * Behaviour follows the reference interface (interface/ceed.c): a resource
 * selects the registered backend whose prefix covers the resource stem (the
 * part before ':'), lowest priority winning ties (:1175-1262); the error
 * handler comes from CEED_ERROR_HANDLER = abort (default) | exit | store
 * (:1269-1277) and "return" for errors without a Ceed (:1743-1750).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ceed-core.h"
#include "ceed-mi355.h"

/* ------------------------------------------------------------------------ */
/* special arguments                                                         */
/* ------------------------------------------------------------------------ */
static CeedRequest                       request_immediate, request_ordered;
CeedRequest *const                       CEED_REQUEST_IMMEDIATE = &request_immediate;
CeedRequest *const                       CEED_REQUEST_ORDERED   = &request_ordered;
static struct CeedVector_private         vector_active, vector_none;
const CeedVector                         CEED_VECTOR_ACTIVE = &vector_active;
const CeedVector                         CEED_VECTOR_NONE   = &vector_none;
static struct CeedBasis_private          basis_none;
const CeedBasis                          CEED_BASIS_NONE = &basis_none;
static struct CeedElemRestriction_private rstr_none;
const CeedElemRestriction                CEED_ELEMRESTRICTION_NONE = &rstr_none;
static struct CeedQFunction_private      qfunction_none;
const CeedQFunction                      CEED_QFUNCTION_NONE     = &qfunction_none;
const CeedInt                            CEED_STRIDES_BACKEND[3] = {0};

int CeedCoreIsSpecialVector(CeedVector v) { return v == CEED_VECTOR_ACTIVE || v == CEED_VECTOR_NONE; }

int CeedRequestWait(CeedRequest *req) {
  (void)req;  /* every request completes immediately */
  return CEED_ERROR_SUCCESS;
}

int CeedGetVersion(int *major, int *minor, int *patch, bool *release) {
  if (major) *major = CEED_VERSION_MAJOR;
  if (minor) *minor = CEED_VERSION_MINOR;
  if (patch) *patch = CEED_VERSION_PATCH;
  if (release) *release = CEED_VERSION_RELEASE;
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* memory                                                                    */
/* ------------------------------------------------------------------------ */
int CeedMallocArray(size_t n, size_t unit, void *p) {
  *(void **)p = n ? malloc(n * unit) : NULL;
  if (n && !*(void **)p) return CeedError(NULL, CEED_ERROR_MAJOR, "malloc of %zu x %zu bytes failed", n, unit);
  return CEED_ERROR_SUCCESS;
}

int CeedCallocArray(size_t n, size_t unit, void *p) {
  *(void **)p = n ? calloc(n, unit) : NULL;
  if (n && !*(void **)p) return CeedError(NULL, CEED_ERROR_MAJOR, "calloc of %zu x %zu bytes failed", n, unit);
  return CEED_ERROR_SUCCESS;
}

int CeedReallocArray(size_t n, size_t unit, void *p) {
  void *q = realloc(*(void **)p, n * unit);
  if (n && !q) return CeedError(NULL, CEED_ERROR_MAJOR, "realloc of %zu x %zu bytes failed", n, unit);
  *(void **)p = q;
  return CEED_ERROR_SUCCESS;
}

int CeedStringAllocCopy(const char *source, char **copy) {
  const size_t len = strlen(source);
  CeedCall(CeedCallocArray(len + 1, 1, copy));
  memcpy(*copy, source, len);
  return CEED_ERROR_SUCCESS;
}

int CeedFree(void *p) {
  free(*(void **)p);
  *(void **)p = NULL;
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* errors                                                                    */
/* ------------------------------------------------------------------------ */
static void format_error(char *buf, size_t len, const char *filename, int line_no, const char *func, const char *format, va_list *args) {
  int n = snprintf(buf, len, "%s:%d in %s(): ", filename, line_no, func);
  if (n < 0) n = 0;
  if ((size_t)n < len) vsnprintf(buf + n, len - (size_t)n, format, *args);
}

int CeedErrorReturn(Ceed ceed, const char *filename, int line_no, const char *func, int err_code, const char *format, va_list *args) {
  (void)ceed, (void)filename, (void)line_no, (void)func, (void)format, (void)args;
  return err_code;
}

int CeedErrorStore(Ceed ceed, const char *filename, int line_no, const char *func, int err_code, const char *format, va_list *args) {
  if (ceed) format_error(ceed->err_msg, sizeof(ceed->err_msg), filename, line_no, func, format, args);
  return err_code;
}

int CeedErrorAbort(Ceed ceed, const char *filename, int line_no, const char *func, int err_code, const char *format, va_list *args) {
  char buf[CEED_MAX_RESOURCE_LEN];
  (void)ceed;
  format_error(buf, sizeof(buf), filename, line_no, func, format, args);
  fprintf(stderr, "%s\n", buf);
  abort();
  return err_code;
}

int CeedErrorExit(Ceed ceed, const char *filename, int line_no, const char *func, int err_code, const char *format, va_list *args) {
  char buf[CEED_MAX_RESOURCE_LEN];
  (void)ceed;
  format_error(buf, sizeof(buf), filename, line_no, func, format, args);
  fprintf(stderr, "%s\n", buf);
  exit(err_code);
  return err_code;
}

int CeedErrorImpl(Ceed ceed, const char *filename, int lineno, const char *func, int ecode, const char *format, ...) {
  va_list args;
  int     ret;
  va_start(args, format);
  if (ceed) {
    ret = ceed->Error(ceed, filename, lineno, func, ecode, format, &args);
  } else {
    const char *h = getenv("CEED_ERROR_HANDLER");
    if (h && !strcmp(h, "return")) ret = CeedErrorReturn(NULL, filename, lineno, func, ecode, format, &args);
    else if (h && (!strcmp(h, "store") || !strcmp(h, "exit"))) {
      char buf[CEED_MAX_RESOURCE_LEN];
      format_error(buf, sizeof(buf), filename, lineno, func, format, &args);
      fprintf(stderr, "%s\n", buf);
      ret = ecode;
      if (!strcmp(h, "exit")) exit(ecode);
    } else ret = CeedErrorAbort(NULL, filename, lineno, func, ecode, format, &args);
  }
  va_end(args);
  return ret;
}

int CeedSetErrorHandler(Ceed ceed, CeedErrorHandler handler) {
  ceed->Error = handler;
  return CEED_ERROR_SUCCESS;
}

int CeedGetErrorMessage(Ceed ceed, const char **err_msg) {
  *err_msg = ceed->err_msg;
  return CEED_ERROR_SUCCESS;
}

int CeedResetErrorMessage(Ceed ceed, const char **err_msg) {
  *err_msg = NULL;
  snprintf(ceed->err_msg, sizeof(ceed->err_msg), "%s", "(no message recorded)");
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* registry                                                                  */
/* ------------------------------------------------------------------------ */
#define MAX_BACKENDS 32
static struct {
  char     prefix[CEED_MAX_RESOURCE_LEN];
  int    (*init)(const char *, Ceed);
  unsigned priority;
} backends[MAX_BACKENDS];
static size_t num_backends;
static bool   builtin_registered;

int CeedRegister(const char *prefix, int (*init)(const char *, Ceed), unsigned int priority) {
  if (num_backends >= MAX_BACKENDS) return CeedError(NULL, CEED_ERROR_MAJOR, "too many backends registered");
  for (size_t i = 0; i < num_backends; i++)
    if (!strcmp(backends[i].prefix, prefix)) return CEED_ERROR_SUCCESS;  /* idempotent */
  snprintf(backends[num_backends].prefix, CEED_MAX_RESOURCE_LEN, "%s", prefix);
  backends[num_backends].init     = init;
  backends[num_backends].priority = priority;
  num_backends++;
  return CEED_ERROR_SUCCESS;
}

static int register_builtin(void) {
  if (builtin_registered) return CEED_ERROR_SUCCESS;
  builtin_registered = true;
  return CeedRegister_Mi355();
}

int CeedRegistryGetList(size_t *n, char ***const resources, CeedInt **priorities) {
  CeedCall(register_builtin());
  *n = num_backends;
  if (resources) {
    CeedCall(CeedCalloc(num_backends, resources));
    for (size_t i = 0; i < num_backends; i++) (*resources)[i] = backends[i].prefix;
  }
  if (priorities) {
    CeedCall(CeedCalloc(num_backends, priorities));
    for (size_t i = 0; i < num_backends; i++) (*priorities)[i] = (CeedInt)backends[i].priority;
  }
  return CEED_ERROR_SUCCESS;
}

int CeedInit(const char *resource, Ceed *ceed) {
  *ceed = NULL;
  if (!resource) return CeedError(NULL, CEED_ERROR_MAJOR, "no resource provided");
  CeedCall(register_builtin());
  const size_t stem = strcspn(resource, ":");
  int          match = -1;
  for (size_t i = 0; i < num_backends; i++) {
    size_t n = 0;
    while (n < stem && backends[i].prefix[n] && backends[i].prefix[n] == resource[n]) n++;
    if (n != stem) continue;
    if (match < 0 || backends[i].priority < backends[match].priority) match = (int)i;
  }
  if (match < 0) {
    char known[CEED_MAX_RESOURCE_LEN] = "";
    for (size_t i = 0; i < num_backends; i++) {
      strncat(known, " ", sizeof(known) - strlen(known) - 1);
      strncat(known, backends[i].prefix, sizeof(known) - strlen(known) - 1);
    }
    return CeedError(NULL, CEED_ERROR_MAJOR, "No suitable backend for resource %s; registered:%s", resource, known);
  }
  Ceed c;
  CeedCall(CeedCalloc(1, &c));
  c->ref_count = 1;
  const char *h = getenv("CEED_ERROR_HANDLER");
  if (h && !strcmp(h, "exit")) c->Error = CeedErrorExit;
  else if (h && !strcmp(h, "store")) c->Error = CeedErrorStore;
  else if (h && !strcmp(h, "return")) c->Error = CeedErrorReturn;
  else c->Error = CeedErrorAbort;
  snprintf(c->err_msg, sizeof(c->err_msg), "%s", "(no message recorded)");
  CeedCall(CeedStringAllocCopy(resource, &c->resource));
  /* a failing backend init is stored (abort and exit act at once) and
   * printed, since the context is released before the caller sees it */
  CeedErrorHandler chosen = c->Error;
  if (chosen == CeedErrorReturn) c->Error = CeedErrorStore;
  int ierr = backends[match].init(resource, c);
  c->Error = chosen;
  if (ierr) {
    fprintf(stderr, "CeedInit(%s): %s\n", resource, c->err_msg);
    free(c->resource);
    free(c);
    return ierr;
  }
  *ceed = c;
  return CEED_ERROR_SUCCESS;
}

int CeedReferenceCopy(Ceed ceed, Ceed *ceed_copy) {
  if (ceed) ceed->ref_count++;
  if (*ceed_copy) CeedCall(CeedDestroy(ceed_copy));
  *ceed_copy = ceed;
  return CEED_ERROR_SUCCESS;
}

int CeedGetResource(Ceed ceed, const char **resource) {
  *resource = ceed->resource;
  return CEED_ERROR_SUCCESS;
}

int CeedIsDeterministic(Ceed ceed, bool *is_deterministic) {
  *is_deterministic = ceed->is_deterministic;
  return CEED_ERROR_SUCCESS;
}

int CeedSetDeterministic(Ceed ceed, bool is_deterministic) {
  ceed->is_deterministic = is_deterministic;
  return CEED_ERROR_SUCCESS;
}

int CeedGetPreferredMemType(Ceed ceed, CeedMemType *type) {
  if (ceed->GetPreferredMemType) return ceed->GetPreferredMemType(type);
  *type = CEED_MEM_HOST;
  return CEED_ERROR_SUCCESS;
}

int CeedGetData(Ceed ceed, void *data) {
  *(void **)data = ceed->data;
  return CEED_ERROR_SUCCESS;
}

int CeedSetData(Ceed ceed, void *data) {
  ceed->data = data;
  return CEED_ERROR_SUCCESS;
}

int CeedGetParent(Ceed ceed, Ceed *parent) {
  *parent = NULL;
  return CeedReferenceCopy(ceed, parent);
}

int CeedView(Ceed ceed, FILE *stream) {
  CeedMemType mem_type;
  CeedCall(CeedGetPreferredMemType(ceed, &mem_type));
  fprintf(stream, "Ceed\n  Ceed Resource: %s\n  Preferred MemType: %s\n", ceed->resource, mem_type == CEED_MEM_HOST ? "host" : "device");
  return CEED_ERROR_SUCCESS;
}

int CeedDestroy(Ceed *ceed) {
  if (!*ceed || --(*ceed)->ref_count > 0) {
    *ceed = NULL;
    return CEED_ERROR_SUCCESS;
  }
  if ((*ceed)->Destroy) CeedCall((*ceed)->Destroy(*ceed));
  free((*ceed)->resource);
  free(*ceed);
  *ceed = NULL;
  return CEED_ERROR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* backend slots: (object type, function name) -> field of the object        */
/* ------------------------------------------------------------------------ */
typedef struct {
  const char *type, *name;
  size_t      offset;
} SlotEntry;

#define SLOT(T, S, F) {#T, #F, offsetof(struct S, F)}
static const SlotEntry slots[] = {
    SLOT(Ceed, Ceed_private, GetPreferredMemType),
    SLOT(Ceed, Ceed_private, Destroy),
    SLOT(Ceed, Ceed_private, VectorCreate),
    SLOT(Ceed, Ceed_private, ElemRestrictionCreate),
    SLOT(Ceed, Ceed_private, BasisCreateTensorH1),
    SLOT(Ceed, Ceed_private, QFunctionCreate),
    SLOT(Ceed, Ceed_private, QFunctionContextCreate),
    SLOT(Ceed, Ceed_private, OperatorCreate),
    SLOT(Vector, CeedVector_private, HasValidArray),
    SLOT(Vector, CeedVector_private, HasBorrowedArrayOfType),
    SLOT(Vector, CeedVector_private, SetArray),
    SLOT(Vector, CeedVector_private, TakeArray),
    SLOT(Vector, CeedVector_private, SyncArray),
    SLOT(Vector, CeedVector_private, SetValue),
    SLOT(Vector, CeedVector_private, GetArray),
    SLOT(Vector, CeedVector_private, GetArrayRead),
    SLOT(Vector, CeedVector_private, GetArrayWrite),
    SLOT(Vector, CeedVector_private, RestoreArray),
    SLOT(Vector, CeedVector_private, RestoreArrayRead),
    SLOT(Vector, CeedVector_private, Norm),
    SLOT(Vector, CeedVector_private, Scale),
    SLOT(Vector, CeedVector_private, AXPY),
    SLOT(Vector, CeedVector_private, AXPBY),
    SLOT(Vector, CeedVector_private, PointwiseMult),
    SLOT(Vector, CeedVector_private, Reciprocal),
    SLOT(Vector, CeedVector_private, Destroy),
    SLOT(ElemRestriction, CeedElemRestriction_private, Apply),
    SLOT(ElemRestriction, CeedElemRestriction_private, GetOffsets),
    SLOT(ElemRestriction, CeedElemRestriction_private, Destroy),
    SLOT(Basis, CeedBasis_private, Apply),
    SLOT(Basis, CeedBasis_private, ApplyAdd),
    SLOT(Basis, CeedBasis_private, Destroy),
    SLOT(QFunction, CeedQFunction_private, Apply),
    SLOT(QFunction, CeedQFunction_private, Destroy),
    SLOT(QFunctionContext, CeedQFunctionContext_private, HasValidData),
    SLOT(QFunctionContext, CeedQFunctionContext_private, HasBorrowedDataOfType),
    SLOT(QFunctionContext, CeedQFunctionContext_private, SetData),
    SLOT(QFunctionContext, CeedQFunctionContext_private, TakeData),
    SLOT(QFunctionContext, CeedQFunctionContext_private, GetData),
    SLOT(QFunctionContext, CeedQFunctionContext_private, GetDataRead),
    SLOT(QFunctionContext, CeedQFunctionContext_private, RestoreData),
    SLOT(QFunctionContext, CeedQFunctionContext_private, RestoreDataRead),
    SLOT(QFunctionContext, CeedQFunctionContext_private, Destroy),
    SLOT(Operator, CeedOperator_private, Apply),
    SLOT(Operator, CeedOperator_private, ApplyAdd),
    SLOT(Operator, CeedOperator_private, LinearAssembleDiagonal),
    SLOT(Operator, CeedOperator_private, LinearAssembleAddDiagonal),
    SLOT(Operator, CeedOperator_private, Destroy),
};
#undef SLOT

int CeedSetBackendFunctionImpl(Ceed ceed, const char *type, void *object, const char *func_name, void (*f)(void)) {
  for (size_t i = 0; i < sizeof(slots) / sizeof(slots[0]); i++) {
    if (!strcmp(slots[i].type, type) && !strcmp(slots[i].name, func_name)) {
      *(void (**)(void))((char *)object + slots[i].offset) = f;
      return CEED_ERROR_SUCCESS;
    }
  }
  /* slots this core has no use for (assembly, composite, ...) are accepted and ignored */
  (void)ceed;
  return CEED_ERROR_SUCCESS;
}

/* Directory holding include/ (for gallery QFunction sources): two levels
 * above this shared library (libceed_amd/lib/ -> repository root). */
int CeedCoreJitRoot(const char **root) {
  static char buf[4096];
  if (!buf[0]) {
    Dl_info info;
    if (!dladdr((void *)&CeedCoreJitRoot, &info) || !info.dli_fname) return CeedError(NULL, CEED_ERROR_MAJOR, "cannot locate the library");
    snprintf(buf, sizeof(buf), "%s", info.dli_fname);
    for (int k = 0; k < 3; k++) {
      char *s = strrchr(buf, '/');
      if (s) *s = '\0';
    }
    strncat(buf, "/include", sizeof(buf) - strlen(buf) - 1);
  }
  *root = buf;
  return CEED_ERROR_SUCCESS;
}
