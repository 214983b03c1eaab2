/* bp-apply: throughput of CeedOperatorApply through the libCEED C API (the
 * drop-in path, library defaults, no tuning): the gallery Poisson3DBuild sets
 * up q_data on the perturbed trilinear hex mesh of the benchmark problems
 * (n^3 elements on the unit cube, interior vertices moved by up to 10% of the
 * element size, seeded; trilinear geometry) and the gallery Poisson3DApply
 * operator (scalar, or the 3-component Vector3Poisson3DApply) of degree p,
 * Q = p + 2, is applied `reps` times.
 *
 * Timing: one loop of `reps` applies between two synchronisations, each a
 * CeedVectorNorm of v (a device reduction whose 8-byte result the host waits
 * for), so no L-vector crosses PCIe inside the timed region.
 * Check: the operator is pure Neumann, so A 1 = 0; the largest |(A 1)_i| is
 * printed ("ones_max_abs").
 * With a sixth argument "user", the scalar operator's pointwise function is
 * the user QFunction of ex-hex instead of the gallery Poisson3DApply: the
 * backend compiles it at run time into its fused kernel, or, with
 * CEED_MI355_FUSE_QFUNCTIONS=0, into the staged path.
 * Usage: bp-apply [resource] [p] [n] [reps] [ncomp] [user]; prints one JSON line. */
#define _POSIX_C_SOURCE 199309L
#include <ceed.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ex-hex.h"

static double now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* seconds for k applies, from a synchronised start to a synchronised end */
static double run(CeedOperator op, CeedVector u, CeedVector v, int k) {
  CeedScalar norm;
  CeedVectorNorm(v, CEED_NORM_MAX, &norm);
  const double t0 = now();
  for (int i = 0; i < k; i++) CeedOperatorApply(op, u, v, CEED_REQUEST_IMMEDIATE);
  CeedVectorNorm(v, CEED_NORM_MAX, &norm);
  return now() - t0;
}

static double lcg(unsigned *s) {
  *s = *s * 1664525u + 1013904223u;
  return *s / 4294967296.0;
}

int main(int argc, char **argv) {
  const char   *resource = argc > 1 ? argv[1] : "/gpu/hip/mi355";
  const CeedInt p = argc > 2 ? atoi(argv[2]) : 4, n = argc > 3 ? atoi(argv[3]) : 50, reps = argc > 4 ? atoi(argv[4]) : 50;
  const CeedInt nc   = argc > 5 ? atoi(argv[5]) : 1;
  const int     user = argc > 6 && !strcmp(argv[6], "user") && nc == 1;
  const CeedInt P = p + 1, Q = p + 2, E = n * n * n, nd = n * p + 1, N = nd * nd * nd;
  Ceed          ceed;
  if (CeedInit(resource, &ceed)) return 2;

  CeedInt *off = malloc(sizeof(CeedInt) * (size_t)E * P * P * P);
  for (CeedInt ez = 0; ez < n; ez++)
    for (CeedInt ey = 0; ey < n; ey++)
      for (CeedInt ex = 0; ex < n; ex++) {
        const CeedInt e = (ez * n + ey) * n + ex;
        for (CeedInt k = 0; k < P; k++)
          for (CeedInt j = 0; j < P; j++)
            for (CeedInt i = 0; i < P; i++)
              off[(((size_t)e * P + k) * P + j) * P + i] = ((ez * p + k) * nd + ey * p + j) * nd + ex * p + i;
      }
  /* trilinear geometry: vertices of the n^3 grid, interior ones perturbed */
  const CeedInt nv1 = n + 1, NV = nv1 * nv1 * nv1;
  CeedInt      *voff = malloc(sizeof(CeedInt) * 8 * (size_t)E);
  CeedScalar   *xv   = malloc(sizeof(CeedScalar) * 3 * (size_t)NV);
  unsigned      seed = 2024u;
  for (CeedInt k = 0; k < nv1; k++)
    for (CeedInt j = 0; j < nv1; j++)
      for (CeedInt i = 0; i < nv1; i++) {
        const CeedInt idx[3] = {i, j, k}, vtx = (k * nv1 + j) * nv1 + i;
        const int     interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
        for (int d = 0; d < 3; d++) xv[(size_t)d * NV + vtx] = (idx[d] + (interior ? 0.2 * lcg(&seed) - 0.1 : 0.0)) / n;
      }
  for (CeedInt ez = 0; ez < n; ez++)
    for (CeedInt ey = 0; ey < n; ey++)
      for (CeedInt ex = 0; ex < n; ex++)
        for (CeedInt c = 0; c < 8; c++)
          voff[8 * (size_t)((ez * n + ey) * n + ex) + c] = ((ez + c / 4) * nv1 + ey + (c / 2) % 2) * nv1 + ex + c % 2;

  CeedElemRestriction r_u, r_x, r_qd;
  CeedBasis           b_u, b_x;
  CeedVector          x, qd, u, v;
  CeedElemRestrictionCreate(ceed, E, P * P * P, nc, N, nc * N, CEED_MEM_HOST, CEED_USE_POINTER, off, &r_u);
  CeedElemRestrictionCreate(ceed, E, 8, 3, NV, 3 * NV, CEED_MEM_HOST, CEED_USE_POINTER, voff, &r_x);
  CeedElemRestrictionCreateStrided(ceed, E, Q * Q * Q, 6, 6 * E * Q * Q * Q, CEED_STRIDES_BACKEND, &r_qd);
  CeedBasisCreateTensorH1Lagrange(ceed, 3, nc, P, Q, CEED_GAUSS, &b_u);
  CeedBasisCreateTensorH1Lagrange(ceed, 3, 3, 2, Q, CEED_GAUSS, &b_x);
  CeedVectorCreate(ceed, 3 * NV, &x);
  CeedVectorSetArray(x, CEED_MEM_HOST, CEED_USE_POINTER, xv);
  CeedVectorCreate(ceed, 6 * E * Q * Q * Q, &qd);
  CeedVectorCreate(ceed, nc * N, &u);
  CeedVectorCreate(ceed, nc * N, &v);

  CeedQFunction qf_build, qf_apply;
  CeedOperator  op_build, op_apply;
  CeedQFunctionCreateInteriorByName(ceed, "Poisson3DBuild", &qf_build);
  CeedOperatorCreate(ceed, qf_build, CEED_QFUNCTION_NONE, CEED_QFUNCTION_NONE, &op_build);
  CeedOperatorSetField(op_build, "dx", r_x, b_x, CEED_VECTOR_ACTIVE);
  CeedOperatorSetField(op_build, "weights", CEED_ELEMRESTRICTION_NONE, b_x, CEED_VECTOR_NONE);
  CeedOperatorSetField(op_build, "qdata", r_qd, CEED_BASIS_NONE, CEED_VECTOR_ACTIVE);
  CeedOperatorApply(op_build, x, qd, CEED_REQUEST_IMMEDIATE);

  if (user) {
    CeedQFunctionCreateInterior(ceed, 1, ex_diffusion, ex_diffusion_loc, &qf_apply);
    CeedQFunctionAddInput(qf_apply, "du", 3, CEED_EVAL_GRAD);
    CeedQFunctionAddInput(qf_apply, "qdata", 6, CEED_EVAL_NONE);
    CeedQFunctionAddOutput(qf_apply, "dv", 3, CEED_EVAL_GRAD);
  } else {
    CeedQFunctionCreateInteriorByName(ceed, nc == 3 ? "Vector3Poisson3DApply" : "Poisson3DApply", &qf_apply);
  }
  CeedOperatorCreate(ceed, qf_apply, CEED_QFUNCTION_NONE, CEED_QFUNCTION_NONE, &op_apply);
  CeedOperatorSetField(op_apply, "du", r_u, b_u, CEED_VECTOR_ACTIVE);
  CeedOperatorSetField(op_apply, "qdata", r_qd, CEED_BASIS_NONE, qd);
  CeedOperatorSetField(op_apply, "dv", r_u, b_u, CEED_VECTOR_ACTIVE);

  {
    CeedScalar *a;
    CeedVectorGetArrayWrite(u, CEED_MEM_HOST, &a);
    unsigned s = 12345u;
    for (CeedInt i = 0; i < nc * N; i++) {
      s = s * 1664525u + 1013904223u;
      a[i] = 2.0 * (s / 4294967296.0) - 1.0;
    }
    CeedVectorRestoreArray(u, &a);
  }
  CeedVectorSetValue(v, 0.0);
  run(op_apply, u, v, 3 + reps / 2); /* warm-up: plan set-up, first transfers, the GPU at its working clocks */
  const double ms = 1e3 * run(op_apply, u, v, reps) / reps;
  CeedScalar   ones_max;
  CeedVectorSetValue(u, 1.0);
  CeedOperatorApply(op_apply, u, v, CEED_REQUEST_IMMEDIATE);
  CeedVectorNorm(v, CEED_NORM_MAX, &ones_max);
  const char *fuse = getenv("CEED_MI355_FUSE_QFUNCTIONS");
  printf("{\"resource\": \"%s\", \"p\": %d, \"n\": %d, \"num_comp\": %d, \"qfunction\": \"%s\", \"mesh\": \"perturbed trilinear\", "
         "\"dofs\": %lld, \"reps\": %d, \"ms_per_apply\": %.5f, \"gdofs\": %.4f, \"ones_max_abs\": %.3e}\n",
         resource, (int)p, (int)n, (int)nc, !user ? "gallery" : (fuse && !strcmp(fuse, "0")) ? "user (staged)" : "user (run-time fused)",
         (long long)nc * N, (int)reps, ms, (double)nc * N / (ms * 1e-3) / 1e9, ones_max);

  CeedVectorDestroy(&x);
  CeedVectorDestroy(&qd);
  CeedVectorDestroy(&u);
  CeedVectorDestroy(&v);
  CeedOperatorDestroy(&op_build);
  CeedOperatorDestroy(&op_apply);
  CeedQFunctionDestroy(&qf_build);
  CeedQFunctionDestroy(&qf_apply);
  CeedBasisDestroy(&b_u);
  CeedBasisDestroy(&b_x);
  CeedElemRestrictionDestroy(&r_u);
  CeedElemRestrictionDestroy(&r_x);
  CeedElemRestrictionDestroy(&r_qd);
  free(off);
  free(voff);
  free(xv);
  return CeedDestroy(&ceed);
}
