/* bp-apply: throughput of CeedOperatorApply through the libCEED C API (the
 * drop-in path, library defaults, no tuning): the gallery Poisson3DBuild sets
 * up q_data and the gallery Poisson3DApply operator (scalar, or the
 * 3-component Vector3Poisson3DApply) is applied `reps` times on an n^3 hex
 * mesh of degree p, Q = p + 2, on the affine map x -> A x.
 *
 * Timing: CeedOperatorApply is asynchronous on the device; reading v on the
 * host waits for it.  Two runs of 5 and 5 + reps applies, each closed by a
 * host read, give the time per apply as their difference over reps.
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

static const double Amap[3][3] = {{1.0, 0.5, 0.0}, {0.0, 2.0, 0.25}, {0.0, 0.0, 3.0}};

static double now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

static double run(CeedOperator op, CeedVector u, CeedVector v, int k, CeedInt n_out) {
  const double t0 = now();
  for (int i = 0; i < k; i++) CeedOperatorApply(op, u, v, CEED_REQUEST_IMMEDIATE);
  const CeedScalar *a;
  CeedVectorGetArrayRead(v, CEED_MEM_HOST, &a);
  volatile double sink = a[n_out - 1];
  (void)sink;
  CeedVectorRestoreArrayRead(v, &a);
  return now() - t0;
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
  CeedScalar *xq = malloc(sizeof(CeedScalar) * 3 * (size_t)N), *gll = malloc(sizeof(CeedScalar) * P);
  CeedLobattoQuadrature(P, gll, NULL);
  for (CeedInt k = 0; k < nd; k++)
    for (CeedInt j = 0; j < nd; j++)
      for (CeedInt i = 0; i < nd; i++) {
        const CeedInt idx[3] = {i, j, k};
        double        r[3];
        for (int d = 0; d < 3; d++) r[d] = ((idx[d] / p) + (idx[d] % p == 0 ? 0.0 : (gll[idx[d] % p] + 1.0) / 2.0)) / n;
        for (int d = 0; d < 3; d++) xq[(size_t)d * N + (k * nd + j) * nd + i] = Amap[d][0] * r[0] + Amap[d][1] * r[1] + Amap[d][2] * r[2];
      }

  CeedElemRestriction r_u, r_x, r_qd;
  CeedBasis           b_u, b_x;
  CeedVector          x, qd, u, v;
  CeedElemRestrictionCreate(ceed, E, P * P * P, nc, N, nc * N, CEED_MEM_HOST, CEED_USE_POINTER, off, &r_u);
  CeedElemRestrictionCreate(ceed, E, P * P * P, 3, N, 3 * N, CEED_MEM_HOST, CEED_USE_POINTER, off, &r_x);
  CeedElemRestrictionCreateStrided(ceed, E, Q * Q * Q, 6, 6 * E * Q * Q * Q, CEED_STRIDES_BACKEND, &r_qd);
  CeedBasisCreateTensorH1Lagrange(ceed, 3, nc, P, Q, CEED_GAUSS, &b_u);
  CeedBasisCreateTensorH1Lagrange(ceed, 3, 3, P, Q, CEED_GAUSS, &b_x);
  CeedVectorCreate(ceed, 3 * N, &x);
  CeedVectorSetArray(x, CEED_MEM_HOST, CEED_USE_POINTER, xq);
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
  run(op_apply, u, v, 3, nc * N); /* warm-up: plan set-up, first transfers */
  const double t_short = run(op_apply, u, v, 5, nc * N);
  const double t_long  = run(op_apply, u, v, 5 + reps, nc * N);
  const double ms      = 1e3 * (t_long - t_short) / reps;
  printf("{\"resource\": \"%s\", \"p\": %d, \"n\": %d, \"num_comp\": %d, \"qfunction\": \"%s\", \"dofs\": %lld, \"reps\": %d, "
         "\"ms_per_apply\": %.5f, \"gdofs\": %.4f}\n",
         resource, (int)p, (int)n, (int)nc, !user ? "gallery" : (getenv("CEED_MI355_FUSE_QFUNCTIONS") && !strcmp(getenv("CEED_MI355_FUSE_QFUNCTIONS"), "0")) ? "user (staged)" : "user (run-time fused)", (long long)nc * N, (int)reps, ms,
         (double)nc * N / (ms * 1e-3) / 1e9);

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
  free(xq);
  free(gll);
  return CeedDestroy(&ceed);
}
