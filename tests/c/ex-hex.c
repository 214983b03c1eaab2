/* ex-hex: a libCEED program of the kind examples/ceed/ex1-volume.c and
 * ex2-surface.c are, run on /gpu/hip/mi355 through the C API.
 *
 * Mesh: n^3 hexes of degree p on the unit cube mapped by an affine map
 * x -> A x (det A = 6).  Checks:
 *   volume   1^T M 1                  = det A          (gallery Mass3DBuild + MassApply, fused path)
 *   energy   x1^T K x1                = det A |A^-T e1|^2 ... computed exactly below
 *   the gallery Poisson3DApply (fused) and a user QFunction (staged path, run-time compiled) agree to 1e-12.
 * Usage: ex-hex [resource] [p] [n]; prints one line, exit status 0 on success. */
#include <ceed.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "ex-hex.h"

static const double Amap[3][3] = {{1.0, 0.5, 0.0}, {0.0, 2.0, 0.25}, {0.0, 0.0, 3.0}};

/* lexicographic H1 offsets of an n^3 mesh of degree p (P = p + 1 nodes per direction) */
static CeedInt *hex_offsets(CeedInt n, CeedInt p) {
  const CeedInt P = p + 1, nd = n * p + 1;
  CeedInt      *off = malloc(sizeof(CeedInt) * n * n * n * P * P * P);
  for (CeedInt ez = 0; ez < n; ez++)
    for (CeedInt ey = 0; ey < n; ey++)
      for (CeedInt ex = 0; ex < n; ex++) {
        const CeedInt e = (ez * n + ey) * n + ex;
        for (CeedInt k = 0; k < P; k++)
          for (CeedInt j = 0; j < P; j++)
            for (CeedInt i = 0; i < P; i++)
              off[((e * P + k) * P + j) * P + i] = ((ez * p + k) * nd + ey * p + j) * nd + ex * p + i;
      }
  return off;
}

int main(int argc, char **argv) {
  const char *resource = argc > 1 ? argv[1] : "/gpu/hip/mi355";
  const CeedInt p = argc > 2 ? atoi(argv[2]) : 4, n = argc > 3 ? atoi(argv[3]) : 4;
  const CeedInt P = p + 1, Q = p + 2, E = n * n * n, nd = n * p + 1, N = nd * nd * nd;
  Ceed          ceed;
  if (CeedInit(resource, &ceed)) return 2;

  /* geometry: degree-p nodes carry the affine coordinates (exact for an affine map) */
  CeedScalar *xq = malloc(sizeof(CeedScalar) * 3 * N), *gll = malloc(sizeof(CeedScalar) * P);
  CeedLobattoQuadrature(P, gll, NULL);
  for (CeedInt k = 0; k < nd; k++)
    for (CeedInt j = 0; j < nd; j++)
      for (CeedInt i = 0; i < nd; i++) {
        const CeedInt idx[3] = {i, j, k};
        double        r[3];
        for (int d = 0; d < 3; d++) r[d] = ((idx[d] / p) + (idx[d] % p == 0 ? 0.0 : (gll[idx[d] % p] + 1.0) / 2.0)) / n;
        for (int d = 0; d < 3; d++) xq[d * N + (k * nd + j) * nd + i] = Amap[d][0] * r[0] + Amap[d][1] * r[1] + Amap[d][2] * r[2];
      }
  CeedInt            *off = hex_offsets(n, p);
  CeedElemRestriction r_u, r_x, r_qm, r_qd;
  CeedBasis           b_u, b_x;
  CeedVector          x, qm, qd, u, v, w;
  CeedElemRestrictionCreate(ceed, E, P * P * P, 1, N, N, CEED_MEM_HOST, CEED_USE_POINTER, off, &r_u);
  CeedElemRestrictionCreate(ceed, E, P * P * P, 3, N, 3 * N, CEED_MEM_HOST, CEED_USE_POINTER, off, &r_x);
  CeedElemRestrictionCreateStrided(ceed, E, Q * Q * Q, 1, E * Q * Q * Q, CEED_STRIDES_BACKEND, &r_qm);
  CeedElemRestrictionCreateStrided(ceed, E, Q * Q * Q, 6, 6 * E * Q * Q * Q, CEED_STRIDES_BACKEND, &r_qd);
  CeedBasisCreateTensorH1Lagrange(ceed, 3, 1, P, Q, CEED_GAUSS, &b_u);
  CeedBasisCreateTensorH1Lagrange(ceed, 3, 3, P, Q, CEED_GAUSS, &b_x);
  CeedVectorCreate(ceed, 3 * N, &x);
  CeedVectorSetArray(x, CEED_MEM_HOST, CEED_USE_POINTER, xq);
  CeedVectorCreate(ceed, E * Q * Q * Q, &qm);
  CeedVectorCreate(ceed, 6 * E * Q * Q * Q, &qd);
  CeedVectorCreate(ceed, N, &u);
  CeedVectorCreate(ceed, N, &v);
  CeedVectorCreate(ceed, N, &w);

  /* q_data set-up operators (gallery) */
  const char         *build[2] = {"Mass3DBuild", "Poisson3DBuild"};
  CeedElemRestriction r_q[2]   = {r_qm, r_qd};
  CeedVector          q[2]     = {qm, qd};
  for (int b = 0; b < 2; b++) {
    CeedQFunction qf;
    CeedOperator  op;
    CeedQFunctionCreateInteriorByName(ceed, build[b], &qf);
    CeedOperatorCreate(ceed, qf, CEED_QFUNCTION_NONE, CEED_QFUNCTION_NONE, &op);
    CeedOperatorSetField(op, "dx", r_x, b_x, CEED_VECTOR_ACTIVE);
    CeedOperatorSetField(op, "weights", CEED_ELEMRESTRICTION_NONE, b_x, CEED_VECTOR_NONE);
    CeedOperatorSetField(op, "qdata", r_q[b], CEED_BASIS_NONE, CEED_VECTOR_ACTIVE);
    CeedOperatorApply(op, x, q[b], CEED_REQUEST_IMMEDIATE);
    CeedOperatorDestroy(&op);
    CeedQFunctionDestroy(&qf);
  }

  /* mass: volume */
  CeedQFunction qf_mass, qf_diff, qf_user;
  CeedOperator  op_mass, op_diff, op_user;
  CeedQFunctionCreateInteriorByName(ceed, "MassApply", &qf_mass);
  CeedOperatorCreate(ceed, qf_mass, CEED_QFUNCTION_NONE, CEED_QFUNCTION_NONE, &op_mass);
  CeedOperatorSetField(op_mass, "u", r_u, b_u, CEED_VECTOR_ACTIVE);
  CeedOperatorSetField(op_mass, "qdata", r_qm, CEED_BASIS_NONE, qm);
  CeedOperatorSetField(op_mass, "v", r_u, b_u, CEED_VECTOR_ACTIVE);
  CeedVectorSetValue(u, 1.0);
  CeedOperatorApply(op_mass, u, v, CEED_REQUEST_IMMEDIATE);
  double volume = 0;
  {
    const CeedScalar *a;
    CeedVectorGetArrayRead(v, CEED_MEM_HOST, &a);
    for (CeedInt i = 0; i < N; i++) volume += a[i];
    CeedVectorRestoreArrayRead(v, &a);
  }

  /* diffusion of u = x_1 (first physical coordinate): gallery (fused) and user QFunction (staged) */
  CeedQFunctionCreateInteriorByName(ceed, "Poisson3DApply", &qf_diff);
  CeedQFunctionCreateInterior(ceed, 1, ex_diffusion, ex_diffusion_loc, &qf_user);
  CeedQFunctionAddInput(qf_user, "du", 3, CEED_EVAL_GRAD);
  CeedQFunctionAddInput(qf_user, "qdata", 6, CEED_EVAL_NONE);
  CeedQFunctionAddOutput(qf_user, "dv", 3, CEED_EVAL_GRAD);
  CeedOperator  *ops[2] = {&op_diff, &op_user};
  CeedQFunction  qfs[2] = {qf_diff, qf_user};
  for (int k = 0; k < 2; k++) {
    CeedOperatorCreate(ceed, qfs[k], CEED_QFUNCTION_NONE, CEED_QFUNCTION_NONE, ops[k]);
    CeedOperatorSetField(*ops[k], "du", r_u, b_u, CEED_VECTOR_ACTIVE);
    CeedOperatorSetField(*ops[k], "qdata", r_qd, CEED_BASIS_NONE, qd);
    CeedOperatorSetField(*ops[k], "dv", r_u, b_u, CEED_VECTOR_ACTIVE);
  }
  CeedVectorSetArray(u, CEED_MEM_HOST, CEED_COPY_VALUES, xq);  /* x_1 at the nodes */
  CeedOperatorApply(op_diff, u, v, CEED_REQUEST_IMMEDIATE);
  CeedOperatorApply(op_user, u, w, CEED_REQUEST_IMMEDIATE);
  double energy = 0, diff = 0, norm = 0;
  {
    const CeedScalar *a, *b;
    CeedVectorGetArrayRead(v, CEED_MEM_HOST, &a);
    CeedVectorGetArrayRead(w, CEED_MEM_HOST, &b);
    for (CeedInt i = 0; i < N; i++) {
      energy += xq[i] * a[i];
      diff += (a[i] - b[i]) * (a[i] - b[i]);
      norm += a[i] * a[i];
    }
    CeedVectorRestoreArrayRead(v, &a);
    CeedVectorRestoreArrayRead(w, &b);
  }
  /* x_1^T K x_1 = int |grad x_1|^2 = |Omega| (grad of a physical coordinate is a unit vector) */
  const double det = Amap[0][0] * Amap[1][1] * Amap[2][2];
  const int    ok  = fabs(volume - det) < 1e-12 * det && fabs(energy - det) < 1e-11 * det && sqrt(diff / norm) < 1e-12;
  printf("ex-hex %s p=%d n=%d: volume %.15f energy %.15f (exact %.1f) fused/staged rel diff %.2e %s\n", resource, (int)p, (int)n, volume, energy,
         det, sqrt(diff / norm), ok ? "ok" : "FAILED");

  CeedVectorDestroy(&x);
  CeedVectorDestroy(&qm);
  CeedVectorDestroy(&qd);
  CeedVectorDestroy(&u);
  CeedVectorDestroy(&v);
  CeedVectorDestroy(&w);
  CeedOperatorDestroy(&op_mass);
  CeedOperatorDestroy(&op_diff);
  CeedOperatorDestroy(&op_user);
  CeedQFunctionDestroy(&qf_mass);
  CeedQFunctionDestroy(&qf_diff);
  CeedQFunctionDestroy(&qf_user);
  CeedElemRestrictionDestroy(&r_u);
  CeedElemRestrictionDestroy(&r_x);
  CeedElemRestrictionDestroy(&r_qm);
  CeedElemRestrictionDestroy(&r_qd);
  CeedBasisDestroy(&b_u);
  CeedBasisDestroy(&b_x);
  CeedDestroy(&ceed);
  free(off);
  free(xq);
  free(gll);
  return ok ? 0 : 1;
}
