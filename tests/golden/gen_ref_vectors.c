/*
 * Golden-vector driver: runs the reference library built by oracle/Makefile
 * (oracle/_ref/libceed.so, /cpu/self/ref/serial) on inputs written by
 * make_golden.py and writes the outputs as raw little-endian arrays.
 * Built and run only in the build container, where the reference exists.
 *
 *   gen basis  P Q qmode                         out
 *   gen bapply dim ncomp P Q tmode emode in      out     (one element)
 *   gen rstr   E es nc cs lsize offsets u v0     out_e out_v
 *   gen op     kind ncomp E P Q nverts N prefix  (prefix.{off,moff,coords,u} -> prefix.{qdata,v})
 */
#include <ceed.h>
#include <ceed/backend.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void *read_file(const char *path, size_t *len) {
  FILE *f = fopen(path, "rb");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", path);
    exit(2);
  }
  fseek(f, 0, SEEK_END);
  *len = (size_t)ftell(f);
  fseek(f, 0, SEEK_SET);
  void *buf = malloc(*len + 8);
  if (fread(buf, 1, *len, f) != *len) exit(3);
  fclose(f);
  return buf;
}

static void write_file(const char *path, const void *buf, size_t len) {
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(buf, 1, len, f) != len) exit(4);
  fclose(f);
}

static void write_vec(const char *path, CeedVector v) {
  const CeedScalar *a;
  CeedSize          n;
  CeedVectorGetLength(v, &n);
  CeedVectorGetArrayRead(v, CEED_MEM_HOST, &a);
  write_file(path, a, sizeof(CeedScalar) * (size_t)n);
  CeedVectorRestoreArrayRead(v, &a);
}

static char *cat(const char *a, const char *b) {
  static char buf[4][1024];
  static int  k = 0;
  k = (k + 1) % 4;
  snprintf(buf[k], sizeof(buf[k]), "%s%s", a, b);
  return buf[k];
}

int main(int argc, char **argv) {
  Ceed ceed;
  if (argc < 2) return 1;
  CeedInit("/cpu/self/ref/serial", &ceed);
  if (!strcmp(argv[1], "basis")) {
    int       P = atoi(argv[2]), Q = atoi(argv[3]), qmode = atoi(argv[4]);
    CeedBasis b;
    CeedBasisCreateTensorH1Lagrange(ceed, 1, 1, P, Q, qmode, &b);
    const CeedScalar *interp, *grad, *qref, *qw;
    CeedBasisGetInterp1D(b, &interp);
    CeedBasisGetGrad1D(b, &grad);
    CeedBasisGetQRef(b, &qref);
    CeedBasisGetQWeights(b, &qw);
    size_t      n   = (size_t)(2 * P * Q + 2 * Q + Q * Q);
    CeedScalar *out = calloc(n, sizeof(CeedScalar));
    memcpy(out, interp, sizeof(CeedScalar) * P * Q);
    memcpy(out + P * Q, grad, sizeof(CeedScalar) * P * Q);
    memcpy(out + 2 * P * Q, qref, sizeof(CeedScalar) * Q);
    memcpy(out + 2 * P * Q + Q, qw, sizeof(CeedScalar) * Q);
    if (Q >= P) CeedBasisGetCollocatedGrad(b, out + 2 * P * Q + 2 * Q);
    write_file(argv[5], out, sizeof(CeedScalar) * n);
    CeedBasisDestroy(&b);
  } else if (!strcmp(argv[1], "bapply")) {
    int       dim = atoi(argv[2]), nc = atoi(argv[3]), P = atoi(argv[4]), Q = atoi(argv[5]), tm = atoi(argv[6]), em = atoi(argv[7]);
    CeedBasis b;
    CeedBasisCreateTensorH1Lagrange(ceed, dim, nc, P, Q, CEED_GAUSS, &b);
    int nn = 1, nq = 1;
    for (int d = 0; d < dim; d++) nn *= P, nq *= Q;
    int        qsize = em == CEED_EVAL_GRAD ? dim * nc * nq : (em == CEED_EVAL_WEIGHT ? nq : nc * nq);
    int        in_n = tm == CEED_NOTRANSPOSE ? nc * nn : qsize, out_n = tm == CEED_NOTRANSPOSE ? qsize : nc * nn;
    CeedVector U = CEED_VECTOR_NONE, V;
    CeedVectorCreate(ceed, out_n, &V);
    CeedVectorSetValue(V, 0.0);
    if (em != CEED_EVAL_WEIGHT) {
      size_t      len;
      CeedScalar *in = read_file(argv[8], &len);
      CeedVectorCreate(ceed, in_n, &U);
      CeedVectorSetArray(U, CEED_MEM_HOST, CEED_COPY_VALUES, in);
      free(in);
    }
    CeedBasisApply(b, 1, tm, em, U, V);
    write_vec(argv[9], V);
    if (U != CEED_VECTOR_NONE) CeedVectorDestroy(&U);
    CeedVectorDestroy(&V);
    CeedBasisDestroy(&b);
  } else if (!strcmp(argv[1], "rstr")) {
    int                 E = atoi(argv[2]), es = atoi(argv[3]), nc = atoi(argv[4]), cs = atoi(argv[5]), lsize = atoi(argv[6]);
    size_t              len;
    CeedInt            *off = read_file(argv[7], &len);
    CeedScalar         *u = read_file(argv[8], &len), *v0 = read_file(argv[9], &len);
    CeedElemRestriction r;
    CeedVector          L, Ev, L2;
    CeedElemRestrictionCreate(ceed, E, es, nc, cs, lsize, CEED_MEM_HOST, CEED_COPY_VALUES, off, &r);
    CeedVectorCreate(ceed, lsize, &L);
    CeedVectorCreate(ceed, lsize, &L2);
    CeedVectorCreate(ceed, (CeedSize)E * es * nc, &Ev);
    CeedVectorSetArray(L, CEED_MEM_HOST, CEED_COPY_VALUES, u);
    CeedVectorSetArray(L2, CEED_MEM_HOST, CEED_COPY_VALUES, v0);
    CeedElemRestrictionApply(r, CEED_NOTRANSPOSE, L, Ev, CEED_REQUEST_IMMEDIATE);
    write_vec(argv[10], Ev);
    CeedElemRestrictionApply(r, CEED_TRANSPOSE, Ev, L2, CEED_REQUEST_IMMEDIATE);
    write_vec(argv[11], L2);
    CeedVectorDestroy(&L);
    CeedVectorDestroy(&L2);
    CeedVectorDestroy(&Ev);
    CeedElemRestrictionDestroy(&r);
    free(off);
    free(u);
    free(v0);
  } else if (!strcmp(argv[1], "op")) {
    int         kind = atoi(argv[2]), nc = atoi(argv[3]), E = atoi(argv[4]), P = atoi(argv[5]), Q = atoi(argv[6]), nv = atoi(argv[7]);
    int         N = atoi(argv[8]);
    const char *pre = argv[9];
    size_t      len;
    CeedInt    *off = read_file(cat(pre, ".off"), &len), *moff = read_file(cat(pre, ".moff"), &len);
    CeedScalar *coords = read_file(cat(pre, ".coords"), &len), *u = read_file(cat(pre, ".u"), &len);
    const int   qc = kind == 0 ? 1 : 6, nq = Q * Q * Q;
    CeedBasis   mb, sb;
    CeedElemRestriction mr, sr, qr;
    CeedQFunction       qfb, qfa;
    CeedOperator        opb, opa;
    CeedVector          X, U, V, QD;
    CeedBasisCreateTensorH1Lagrange(ceed, 3, 3, 2, Q, CEED_GAUSS, &mb);
    CeedBasisCreateTensorH1Lagrange(ceed, 3, nc, P, Q, CEED_GAUSS, &sb);
    CeedElemRestrictionCreate(ceed, E, 8, 3, nv, 3 * nv, CEED_MEM_HOST, CEED_COPY_VALUES, moff, &mr);
    CeedElemRestrictionCreate(ceed, E, P * P * P, nc, N, nc * N, CEED_MEM_HOST, CEED_COPY_VALUES, off, &sr);
    CeedElemRestrictionCreateStrided(ceed, E, nq, qc, (CeedSize)qc * E * nq, CEED_STRIDES_BACKEND, &qr);
    CeedVectorCreate(ceed, 3 * nv, &X);
    CeedVectorSetArray(X, CEED_MEM_HOST, CEED_COPY_VALUES, coords);
    CeedVectorCreate(ceed, (CeedSize)qc * E * nq, &QD);
    CeedQFunctionCreateInteriorByName(ceed, kind == 0 ? "Mass3DBuild" : "Poisson3DBuild", &qfb);
    CeedOperatorCreate(ceed, qfb, CEED_QFUNCTION_NONE, CEED_QFUNCTION_NONE, &opb);
    CeedOperatorSetField(opb, "dx", mr, mb, CEED_VECTOR_ACTIVE);
    CeedOperatorSetField(opb, "weights", CEED_ELEMRESTRICTION_NONE, mb, CEED_VECTOR_NONE);
    CeedOperatorSetField(opb, "qdata", qr, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE);
    CeedOperatorApply(opb, X, QD, CEED_REQUEST_IMMEDIATE);
    write_vec(cat(pre, ".qdata"), QD);
    const char *name = kind == 0 ? (nc == 1 ? "MassApply" : "Vector3MassApply") : (nc == 1 ? "Poisson3DApply" : "Vector3Poisson3DApply");
    CeedQFunctionCreateInteriorByName(ceed, name, &qfa);
    CeedOperatorCreate(ceed, qfa, CEED_QFUNCTION_NONE, CEED_QFUNCTION_NONE, &opa);
    CeedOperatorSetField(opa, kind == 0 ? "u" : "du", sr, sb, CEED_VECTOR_ACTIVE);
    CeedOperatorSetField(opa, "qdata", qr, CEED_BASIS_NONE, QD);
    CeedOperatorSetField(opa, kind == 0 ? "v" : "dv", sr, sb, CEED_VECTOR_ACTIVE);
    CeedVectorCreate(ceed, (CeedSize)nc * N, &U);
    CeedVectorCreate(ceed, (CeedSize)nc * N, &V);
    CeedVectorSetArray(U, CEED_MEM_HOST, CEED_COPY_VALUES, u);
    CeedOperatorApply(opa, U, V, CEED_REQUEST_IMMEDIATE);
    write_vec(cat(pre, ".v"), V);
    CeedVectorDestroy(&X);
    CeedVectorDestroy(&U);
    CeedVectorDestroy(&V);
    CeedVectorDestroy(&QD);
    CeedOperatorDestroy(&opb);
    CeedOperatorDestroy(&opa);
    CeedQFunctionDestroy(&qfb);
    CeedQFunctionDestroy(&qfa);
    CeedElemRestrictionDestroy(&mr);
    CeedElemRestrictionDestroy(&sr);
    CeedElemRestrictionDestroy(&qr);
    CeedBasisDestroy(&mb);
    CeedBasisDestroy(&sb);
    free(off);
    free(moff);
    free(coords);
    free(u);
  } else {
    return 1;
  }
  CeedDestroy(&ceed);
  return 0;
}
