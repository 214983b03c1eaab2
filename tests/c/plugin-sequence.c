/*
 * Drives the libCEED calls of examples/ceed/ex1-volume.c and ex2-surface.c,
 * and the vector, restriction-offset, QFunction-context and QFunctionApply
 * entry points behind them, on /gpu/hip/mi355 inside whichever libCEED it is
 * linked with (the reference's own, with the backend registered as a plugin).
 * Every call must succeed; values are checked against closed forms.
 * Prints "ok" and exits 0, or names the failing step and exits 1.
 */
#include <ceed.h>
#include <ceed/backend.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#define CHECK(call)                                                   \
  do {                                                                \
    int ierr_ = (call);                                               \
    if (ierr_) {                                                      \
      printf("FAIL: %s returned %d (line %d)\n", #call, ierr_, __LINE__); \
      return 1;                                                       \
    }                                                                 \
  } while (0)
#define EXPECT(cond)                                             \
  do {                                                           \
    if (!(cond)) {                                               \
      printf("FAIL: %s (line %d)\n", #cond, __LINE__);           \
      return 1;                                                  \
    }                                                            \
  } while (0)

typedef struct {
  double scale;
  int    id;
} Ctx;

int main(int argc, char **argv) {
  const char *resource = argc > 1 ? argv[1] : "/gpu/hip/mi355";
  Ceed        ceed;
  CHECK(CeedInit(resource, &ceed));
  bool det;
  CHECK(CeedIsDeterministic(ceed, &det));
  EXPECT(det);

  /* ---- vectors: validity, reads, norms, BLAS-1, borrowed arrays ---- */
  {
    CeedVector  x, y;
    CeedScalar  a[10], b[10], *taken;
    const CeedScalar *r;
    bool        valid, borrowed;
    for (int i = 0; i < 10; i++) a[i] = i - 4.5, b[i] = 2.0;
    CHECK(CeedVectorCreate(ceed, 10, &x));
    CHECK(CeedVectorCreate(ceed, 10, &y));
    CHECK(CeedVectorHasValidArray(x, &valid));
    EXPECT(!valid);
    CHECK(CeedVectorSetArray(x, CEED_MEM_HOST, CEED_COPY_VALUES, a));
    CHECK(CeedVectorHasValidArray(x, &valid));
    EXPECT(valid);
    CHECK(CeedVectorGetArrayRead(x, CEED_MEM_HOST, &r));
    for (int i = 0; i < 10; i++) EXPECT(r[i] == a[i]);
    CHECK(CeedVectorRestoreArrayRead(x, &r));
    CeedScalar n1, n2, nm;
    CHECK(CeedVectorNorm(x, CEED_NORM_1, &n1));
    CHECK(CeedVectorNorm(x, CEED_NORM_2, &n2));
    CHECK(CeedVectorNorm(x, CEED_NORM_MAX, &nm));
    EXPECT(fabs(n1 - 25.0) < 1e-13 && fabs(n2 - sqrt(82.5)) < 1e-13 && nm == 4.5);
    CHECK(CeedVectorSetValue(y, 3.0));
    CHECK(CeedVectorAXPY(y, 2.0, x));   /* y = 3 + 2a */
    CHECK(CeedVectorAXPBY(y, 1.0, 0.5, x)); /* y = 1.5 + 2a */
    CHECK(CeedVectorScale(y, 2.0));      /* y = 3 + 4a */
    CHECK(CeedVectorPointwiseMult(y, y, x));
    CHECK(CeedVectorGetArrayRead(y, CEED_MEM_HOST, &r));
    for (int i = 0; i < 10; i++) EXPECT(fabs(r[i] - (3 + 4 * a[i]) * a[i]) < 1e-12);
    CHECK(CeedVectorRestoreArrayRead(y, &r));
    CHECK(CeedVectorSetArray(y, CEED_MEM_HOST, CEED_USE_POINTER, b));
    CHECK(CeedVectorHasBorrowedArrayOfType(y, CEED_MEM_HOST, &borrowed));
    EXPECT(borrowed);
    CHECK(CeedVectorReciprocal(y));
    CHECK(CeedVectorTakeArray(y, CEED_MEM_HOST, &taken));
    EXPECT(taken == b && b[0] == 0.5);
    CHECK(CeedVectorDestroy(&x));
    CHECK(CeedVectorDestroy(&y));
  }

  /* ---- QFunction context: the backend stores the data ---- */
  {
    CeedQFunctionContext ctx;
    Ctx                  c = {2.5, 7}, lent = {1.0, 1}, *p;
    const Ctx           *rp;
    bool                 valid;
    CHECK(CeedQFunctionContextCreate(ceed, &ctx));
    CHECK(CeedQFunctionContextSetData(ctx, CEED_MEM_HOST, CEED_COPY_VALUES, sizeof(c), &c));
    CHECK(CeedQFunctionContextGetDataRead(ctx, CEED_MEM_HOST, &rp));
    EXPECT(rp != &c && rp->scale == 2.5 && rp->id == 7);
    CHECK(CeedQFunctionContextRestoreDataRead(ctx, &rp));
    CHECK(CeedQFunctionContextGetData(ctx, CEED_MEM_HOST, &p));
    p->id = 8;
    CHECK(CeedQFunctionContextRestoreData(ctx, &p));
    CHECK(CeedQFunctionContextGetDataRead(ctx, CEED_MEM_HOST, &rp));
    EXPECT(rp->id == 8);
    CHECK(CeedQFunctionContextRestoreDataRead(ctx, &rp));
    CHECK(CeedQFunctionContextSetData(ctx, CEED_MEM_HOST, CEED_USE_POINTER, sizeof(lent), &lent));
    CHECK(CeedQFunctionContextTakeData(ctx, CEED_MEM_HOST, &p));
    EXPECT(p == &lent);
    CHECK(CeedQFunctionContextHasValidData(ctx, &valid));
    EXPECT(!valid);
    CHECK(CeedQFunctionContextDestroy(&ctx));
  }

  /* ---- restriction offsets and apply ---- */
  {
    CeedElemRestriction rstr;
    const CeedInt      *off;
    CeedInt             ind[8] = {0, 1, 1, 2, 2, 3, 3, 4};
    CeedVector          l, e;
    CeedScalar          lv[5] = {1, 2, 3, 4, 5};
    const CeedScalar   *r;
    CHECK(CeedElemRestrictionCreate(ceed, 4, 2, 1, 1, 5, CEED_MEM_HOST, CEED_COPY_VALUES, ind, &rstr));
    CHECK(CeedElemRestrictionGetOffsets(rstr, CEED_MEM_HOST, &off));
    for (int i = 0; i < 8; i++) EXPECT(off[i] == ind[i]);
    CHECK(CeedElemRestrictionRestoreOffsets(rstr, &off));
    CHECK(CeedVectorCreate(ceed, 5, &l));
    CHECK(CeedVectorCreate(ceed, 8, &e));
    CHECK(CeedVectorSetArray(l, CEED_MEM_HOST, CEED_COPY_VALUES, lv));
    CHECK(CeedElemRestrictionApply(rstr, CEED_NOTRANSPOSE, l, e, CEED_REQUEST_IMMEDIATE));
    CHECK(CeedVectorGetArrayRead(e, CEED_MEM_HOST, &r));
    for (int i = 0; i < 8; i++) EXPECT(r[i] == lv[ind[i]]);
    CHECK(CeedVectorRestoreArrayRead(e, &r));
    CHECK(CeedVectorSetValue(l, 0.0));
    CHECK(CeedElemRestrictionApply(rstr, CEED_TRANSPOSE, e, l, CEED_REQUEST_IMMEDIATE));
    CHECK(CeedVectorGetArrayRead(l, CEED_MEM_HOST, &r));
    EXPECT(r[0] == 1 && r[1] == 4 && r[2] == 6 && r[3] == 8 && r[4] == 5);
    CHECK(CeedVectorRestoreArrayRead(l, &r));
    CHECK(CeedVectorDestroy(&l));
    CHECK(CeedVectorDestroy(&e));
    CHECK(CeedElemRestrictionDestroy(&rstr));
  }

  /* ---- CeedQFunctionApply on gallery QFunctions (tests/t4xx pattern) ---- */
  {
    const CeedInt Q = 8;
    CeedQFunction qf;
    CeedVector    in[2], out[1];
    CeedScalar    u[8], w[8], ug[3 * 8], qd[6 * 8];
    const CeedScalar *r;
    for (int i = 0; i < Q; i++) u[i] = 1 + i, w[i] = 0.5 * i;
    CHECK(CeedQFunctionCreateInteriorByName(ceed, "MassApply", &qf));
    CHECK(CeedVectorCreate(ceed, Q, &in[0]));
    CHECK(CeedVectorCreate(ceed, Q, &in[1]));
    CHECK(CeedVectorCreate(ceed, Q, &out[0]));
    CHECK(CeedVectorSetArray(in[0], CEED_MEM_HOST, CEED_COPY_VALUES, u));
    CHECK(CeedVectorSetArray(in[1], CEED_MEM_HOST, CEED_COPY_VALUES, w));
    CHECK(CeedQFunctionApply(qf, Q, in, out));
    CHECK(CeedVectorGetArrayRead(out[0], CEED_MEM_HOST, &r));
    for (int i = 0; i < Q; i++) EXPECT(r[i] == u[i] * w[i]);
    CHECK(CeedVectorRestoreArrayRead(out[0], &r));
    for (int k = 0; k < 3; k++) CHECK(CeedVectorDestroy(k < 2 ? &in[k] : &out[0]));
    CHECK(CeedQFunctionDestroy(&qf));
    /* Poisson3DApply: vg = S ug, S in Voigt order {xx, yy, zz, yz, xz, xy} */
    for (int i = 0; i < Q; i++) {
      for (int d = 0; d < 3; d++) ug[d * Q + i] = d + 1 + 0.1 * i;
      for (int k = 0; k < 6; k++) qd[k * Q + i] = k < 3 ? 2.0 + k : 0.25 * (k - 2);
    }
    CHECK(CeedQFunctionCreateInteriorByName(ceed, "Poisson3DApply", &qf));
    CHECK(CeedVectorCreate(ceed, 3 * Q, &in[0]));
    CHECK(CeedVectorCreate(ceed, 6 * Q, &in[1]));
    CHECK(CeedVectorCreate(ceed, 3 * Q, &out[0]));
    CHECK(CeedVectorSetArray(in[0], CEED_MEM_HOST, CEED_COPY_VALUES, ug));
    CHECK(CeedVectorSetArray(in[1], CEED_MEM_HOST, CEED_COPY_VALUES, qd));
    CHECK(CeedQFunctionApply(qf, Q, in, out));
    CHECK(CeedVectorGetArrayRead(out[0], CEED_MEM_HOST, &r));
    for (int i = 0; i < Q; i++) {
      const double gx = ug[i], gy = ug[Q + i], gz = ug[2 * Q + i];
      const double sxx = qd[i], syy = qd[Q + i], szz = qd[2 * Q + i], syz = qd[3 * Q + i], sxz = qd[4 * Q + i], sxy = qd[5 * Q + i];
      EXPECT(fabs(r[i] - (sxx * gx + sxy * gy + sxz * gz)) < 1e-13);
      EXPECT(fabs(r[Q + i] - (sxy * gx + syy * gy + syz * gz)) < 1e-13);
      EXPECT(fabs(r[2 * Q + i] - (sxz * gx + syz * gy + szz * gz)) < 1e-13);
    }
    CHECK(CeedVectorRestoreArrayRead(out[0], &r));
    for (int k = 0; k < 3; k++) CHECK(CeedVectorDestroy(k < 2 ? &in[k] : &out[0]));
    CHECK(CeedQFunctionDestroy(&qf));
  }
  CHECK(CeedDestroy(&ceed));
  printf("ok\n");
  return 0;
}
