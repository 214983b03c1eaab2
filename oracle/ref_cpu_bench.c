/*
 * CPU baseline: times CeedOperatorApply of the 3D Poisson operator on a
 * reference CPU backend (default /cpu/self/avx/blocked), built from the
 * reference sources by oracle/Makefile (oracle/_ref/perf).  Used only by
 * bench.py's cpu_baseline leg.  One process = one core; bench.py runs one
 * process per core on its own mesh and sums the rates.
 *
 *   ref_cpu_bench <resource> <p> <n> <seconds>
 * prints: {"dofs": N, "applies": K, "seconds": T}
 *
 * Mesh: n^3 elements on the unit cube, lexicographic numbering as in
 * examples/ceed/ex1-volume.c; geometry is affine (the apply cost does not
 * depend on the coordinate values).
 */
#define _POSIX_C_SOURCE 199309L
#include <ceed.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

static double now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void lex_offsets(int n, int p, CeedInt *off) {
  const int P = p + 1, M = n * p + 1;
  for (int ez = 0; ez < n; ez++)
    for (int ey = 0; ey < n; ey++)
      for (int ex = 0; ex < n; ex++) {
        const int e = ex + n * (ey + n * ez);
        for (int k = 0; k < P; k++)
          for (int j = 0; j < P; j++)
            for (int i = 0; i < P; i++)
              off[(size_t)e * P * P * P + i + P * (j + P * k)] = (ex * p + i) + M * ((ey * p + j) + M * (ez * p + k));
      }
}

int main(int argc, char **argv) {
  if (argc < 5) return 1;
  const char *resource = argv[1];
  const int   p = atoi(argv[2]), n = atoi(argv[3]), P = p + 1, Q = p + 2, E = n * n * n;
  const double secs = atof(argv[4]);
  const int    M = n * p + 1, N = M * M * M, nv = (n + 1) * (n + 1) * (n + 1), nq = Q * Q * Q;
  Ceed         ceed;
  if (CeedInit(resource, &ceed)) return 2;
  CeedInt    *off = malloc(sizeof(CeedInt) * (size_t)E * P * P * P), *moff = malloc(sizeof(CeedInt) * (size_t)E * 8);
  CeedScalar *xyz = malloc(sizeof(CeedScalar) * 3 * (size_t)nv);
  lex_offsets(n, p, off);
  lex_offsets(n, 1, moff);
  for (int k = 0; k <= n; k++)
    for (int j = 0; j <= n; j++)
      for (int i = 0; i <= n; i++) {
        const int v = i + (n + 1) * (j + (n + 1) * k);
        xyz[v] = (double)i / n;
        xyz[nv + v] = (double)j / n;
        xyz[2 * nv + v] = (double)k / n;
      }
  CeedBasis           mb, sb;
  CeedElemRestriction mr, sr, qr;
  CeedQFunction       qfb, qfa;
  CeedOperator        opb, opa;
  CeedVector          X, U, V, QD;
  CeedBasisCreateTensorH1Lagrange(ceed, 3, 3, 2, Q, CEED_GAUSS, &mb);
  CeedBasisCreateTensorH1Lagrange(ceed, 3, 1, P, Q, CEED_GAUSS, &sb);
  CeedElemRestrictionCreate(ceed, E, 8, 3, nv, 3 * nv, CEED_MEM_HOST, CEED_COPY_VALUES, moff, &mr);
  CeedElemRestrictionCreate(ceed, E, P * P * P, 1, 1, N, CEED_MEM_HOST, CEED_COPY_VALUES, off, &sr);
  CeedElemRestrictionCreateStrided(ceed, E, nq, 6, (CeedSize)6 * E * nq, CEED_STRIDES_BACKEND, &qr);
  CeedVectorCreate(ceed, 3 * nv, &X);
  CeedVectorSetArray(X, CEED_MEM_HOST, CEED_COPY_VALUES, xyz);
  CeedVectorCreate(ceed, (CeedSize)6 * E * nq, &QD);
  CeedQFunctionCreateInteriorByName(ceed, "Poisson3DBuild", &qfb);
  CeedOperatorCreate(ceed, qfb, CEED_QFUNCTION_NONE, CEED_QFUNCTION_NONE, &opb);
  CeedOperatorSetField(opb, "dx", mr, mb, CEED_VECTOR_ACTIVE);
  CeedOperatorSetField(opb, "weights", CEED_ELEMRESTRICTION_NONE, mb, CEED_VECTOR_NONE);
  CeedOperatorSetField(opb, "qdata", qr, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE);
  CeedOperatorApply(opb, X, QD, CEED_REQUEST_IMMEDIATE);
  CeedQFunctionCreateInteriorByName(ceed, "Poisson3DApply", &qfa);
  CeedOperatorCreate(ceed, qfa, CEED_QFUNCTION_NONE, CEED_QFUNCTION_NONE, &opa);
  CeedOperatorSetField(opa, "du", sr, sb, CEED_VECTOR_ACTIVE);
  CeedOperatorSetField(opa, "qdata", qr, CEED_BASIS_NONE, QD);
  CeedOperatorSetField(opa, "dv", sr, sb, CEED_VECTOR_ACTIVE);
  CeedVectorCreate(ceed, N, &U);
  CeedVectorCreate(ceed, N, &V);
  CeedVectorSetValue(U, 1.0);
  CeedOperatorApply(opa, U, V, CEED_REQUEST_IMMEDIATE); /* warm-up */
  int    k  = 0;
  double t0 = now(), t = 0;
  do {
    CeedOperatorApply(opa, U, V, CEED_REQUEST_IMMEDIATE);
    k++;
    t = now() - t0;
  } while (t < secs);
  printf("{\"dofs\": %d, \"applies\": %d, \"seconds\": %.6f}\n", N, k, t);
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
  CeedDestroy(&ceed);
  free(off);
  free(moff);
  free(xyz);
  return 0;
}
