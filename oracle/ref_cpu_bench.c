/*
 * CPU baseline: times CeedOperatorApply of the 3D Poisson operator on a
 * reference CPU backend (default /cpu/self/avx/blocked), built from the
 * reference sources by oracle/Makefile (oracle/_ref/perf).  Used only by
 * bench.py's cpu_baseline leg.  One process = one core; bench.py runs one
 * process per core on its own mesh and sums the rates.
 *
 *   ref_cpu_bench <resource> <p> <nx> <ny> <nz> <seconds>
 * prints: {"dofs": N, "applies": K, "seconds": T}
 *
 * Mesh: nx x ny x nz elements (a slab of a larger structured mesh: bench.py
 * gives each process an equal share of the GPU workload's elements),
 * lexicographic numbering as in examples/ceed/ex1-volume.c; geometry is
 * affine (the apply cost does not depend on the coordinate values).
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

static void lex_offsets(const int n[3], int p, CeedInt *off) {
  const int P = p + 1, Mx = n[0] * p + 1, My = n[1] * p + 1;
  for (int ez = 0; ez < n[2]; ez++)
    for (int ey = 0; ey < n[1]; ey++)
      for (int ex = 0; ex < n[0]; ex++) {
        const int e = ex + n[0] * (ey + n[1] * ez);
        for (int k = 0; k < P; k++)
          for (int j = 0; j < P; j++)
            for (int i = 0; i < P; i++)
              off[(size_t)e * P * P * P + i + P * (j + P * k)] = (ex * p + i) + Mx * ((ey * p + j) + My * (ez * p + k));
      }
}

int main(int argc, char **argv) {
  if (argc < 7) return 1;
  const char *resource = argv[1];
  const int   p = atoi(argv[2]), n[3] = {atoi(argv[3]), atoi(argv[4]), atoi(argv[5])}, P = p + 1, Q = p + 2, E = n[0] * n[1] * n[2];
  const double secs = atof(argv[6]);
  const int    N = (n[0] * p + 1) * (n[1] * p + 1) * (n[2] * p + 1), nv = (n[0] + 1) * (n[1] + 1) * (n[2] + 1), nq = Q * Q * Q;
  Ceed         ceed;
  if (CeedInit(resource, &ceed)) return 2;
  CeedInt    *off = malloc(sizeof(CeedInt) * (size_t)E * P * P * P), *moff = malloc(sizeof(CeedInt) * (size_t)E * 8);
  CeedScalar *xyz = malloc(sizeof(CeedScalar) * 3 * (size_t)nv);
  lex_offsets(n, p, off);
  lex_offsets(n, 1, moff);
  for (int k = 0; k <= n[2]; k++)
    for (int j = 0; j <= n[1]; j++)
      for (int i = 0; i <= n[0]; i++) {
        const int v = i + (n[0] + 1) * (j + (n[1] + 1) * k);
        xyz[v]          = (double)i / n[0];
        xyz[nv + v]     = (double)j / n[1];
        xyz[2 * nv + v] = (double)k / n[2];
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
