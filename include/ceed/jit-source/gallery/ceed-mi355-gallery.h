/*
 * Gallery QFunctions of the /gpu/hip/mi355 backend, by name: the pointwise
 * operators CeedQFunctionCreateInteriorByName provides.  The same text is
 * compiled into the host library (function pointers) and, at run time, by
 * hipRTC for the GPU.  Names, field sizes and the data layouts (geometric
 * factors in Voigt order) are those of the reference gallery
 * (include/ceed/jit-source/gallery/ceed-*.h) so operators built from them are
 * interchangeable; the pointwise algebra is restated here.
 *
 * Conventions: J[d][c][i] = dx_c/dX_d at point i (GRAD of the coordinates);
 * 3D Voigt order {xx, yy, zz, yz, xz, xy}, 2D {xx, yy, xy}.
 */
#ifndef CEED_MI355_GALLERY_H
#define CEED_MI355_GALLERY_H

#include <ceed/types.h>

/* ---- mass ---------------------------------------------------------------- */
CEED_QFUNCTION(Mass1DBuild)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const CeedScalar *J = in[0], *w = in[1];
  for (CeedInt i = 0; i < Q; i++) out[0][i] = J[i] * w[i];
  return 0;
}

CEED_QFUNCTION(Mass2DBuild)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const CeedScalar *J = in[0], *w = in[1];
  for (CeedInt i = 0; i < Q; i++) out[0][i] = (J[0 * Q + i] * J[3 * Q + i] - J[1 * Q + i] * J[2 * Q + i]) * w[i];
  return 0;
}

CEED_QFUNCTION(Mass3DBuild)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const CeedScalar *J = in[0], *w = in[1];
  for (CeedInt i = 0; i < Q; i++) {
    const CeedScalar a = J[0 * Q + i], b = J[1 * Q + i], c = J[2 * Q + i];  /* J[0][*] */
    const CeedScalar d = J[3 * Q + i], e = J[4 * Q + i], f = J[5 * Q + i];  /* J[1][*] */
    const CeedScalar g = J[6 * Q + i], h = J[7 * Q + i], k = J[8 * Q + i];  /* J[2][*] */
    out[0][i] = (a * (e * k - f * h) - b * (d * k - f * g) + c * (d * h - e * g)) * w[i];
  }
  return 0;
}

CEED_QFUNCTION(MassApply)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++) out[0][i] = in[0][i] * in[1][i];
  return 0;
}

CEED_QFUNCTION(Vector3MassApply)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++)
    for (CeedInt c = 0; c < 3; c++) out[0][c * Q + i] = in[0][c * Q + i] * in[1][i];
  return 0;
}

/* ---- Poisson --------------------------------------------------------------- */
CEED_QFUNCTION(Poisson1DBuild)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++) out[0][i] = in[1][i] / in[0][i];
  return 0;
}

CEED_QFUNCTION(Poisson2DBuild)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const CeedScalar *J = in[0], *w = in[1];
  for (CeedInt i = 0; i < Q; i++) {
    const CeedScalar j00 = J[0 * Q + i], j10 = J[1 * Q + i], j01 = J[2 * Q + i], j11 = J[3 * Q + i];
    const CeedScalar qw  = w[i] / (j00 * j11 - j10 * j01);
    out[0][0 * Q + i]    = qw * (j01 * j01 + j11 * j11);
    out[0][1 * Q + i]    = qw * (j00 * j00 + j10 * j10);
    out[0][2 * Q + i]    = -qw * (j00 * j01 + j10 * j11);
  }
  return 0;
}

/* (w / det J) adj(J) adj(J)^T, symmetric part in Voigt order */
CEED_QFUNCTION(Poisson3DBuild)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const CeedScalar *J = in[0], *w = in[1];
  for (CeedInt i = 0; i < Q; i++) {
    CeedScalar A[3][3];
    for (CeedInt j = 0; j < 3; j++)
      for (CeedInt k = 0; k < 3; k++) {
        const CeedInt k1 = (k + 1) % 3, k2 = (k + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        A[k][j] = J[(k1 * 3 + j1) * Q + i] * J[(k2 * 3 + j2) * Q + i] - J[(k2 * 3 + j1) * Q + i] * J[(k1 * 3 + j2) * Q + i];
      }
    const CeedScalar qw = w[i] / (J[0 * Q + i] * A[0][0] + J[1 * Q + i] * A[0][1] + J[2 * Q + i] * A[0][2]);
    out[0][0 * Q + i]   = qw * (A[0][0] * A[0][0] + A[0][1] * A[0][1] + A[0][2] * A[0][2]);
    out[0][1 * Q + i]   = qw * (A[1][0] * A[1][0] + A[1][1] * A[1][1] + A[1][2] * A[1][2]);
    out[0][2 * Q + i]   = qw * (A[2][0] * A[2][0] + A[2][1] * A[2][1] + A[2][2] * A[2][2]);
    out[0][3 * Q + i]   = qw * (A[1][0] * A[2][0] + A[1][1] * A[2][1] + A[1][2] * A[2][2]);
    out[0][4 * Q + i]   = qw * (A[0][0] * A[2][0] + A[0][1] * A[2][1] + A[0][2] * A[2][2]);
    out[0][5 * Q + i]   = qw * (A[0][0] * A[1][0] + A[0][1] * A[1][1] + A[0][2] * A[1][2]);
  }
  return 0;
}

CEED_QFUNCTION(Poisson1DApply)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++) out[0][i] = in[0][i] * in[1][i];
  return 0;
}

/* vg_j = sum_d ug_d S_dj for each of nc components; S from 2D Voigt {xx, yy, xy} */
CEED_QFUNCTION_HELPER void mi355_poisson2d_point(CeedInt Q, CeedInt i, CeedInt nc, const CeedScalar *ug, const CeedScalar *qd, CeedScalar *vg) {
  const CeedScalar s00 = qd[0 * Q + i], s11 = qd[1 * Q + i], s01 = qd[2 * Q + i];
  for (CeedInt c = 0; c < nc; c++) {
    const CeedScalar u0 = ug[(0 * nc + c) * Q + i], u1 = ug[(1 * nc + c) * Q + i];
    vg[(0 * nc + c) * Q + i] = u0 * s00 + u1 * s01;
    vg[(1 * nc + c) * Q + i] = u0 * s01 + u1 * s11;
  }
}

/* vg_j = sum_d ug_d S_dj for each of nc components; S from 3D Voigt {xx, yy, zz, yz, xz, xy} */
CEED_QFUNCTION_HELPER void mi355_poisson3d_point(CeedInt Q, CeedInt i, CeedInt nc, const CeedScalar *ug, const CeedScalar *qd, CeedScalar *vg) {
  const CeedScalar S[3][3] = {
      {qd[0 * Q + i], qd[5 * Q + i], qd[4 * Q + i]},
      {qd[5 * Q + i], qd[1 * Q + i], qd[3 * Q + i]},
      {qd[4 * Q + i], qd[3 * Q + i], qd[2 * Q + i]},
  };
  for (CeedInt c = 0; c < nc; c++) {
    const CeedScalar u0 = ug[(0 * nc + c) * Q + i], u1 = ug[(1 * nc + c) * Q + i], u2 = ug[(2 * nc + c) * Q + i];
    for (CeedInt j = 0; j < 3; j++) vg[(j * nc + c) * Q + i] = u0 * S[0][j] + u1 * S[1][j] + u2 * S[2][j];
  }
}

CEED_QFUNCTION(Poisson2DApply)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++) mi355_poisson2d_point(Q, i, 1, in[0], in[1], out[0]);
  return 0;
}

CEED_QFUNCTION(Poisson3DApply)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++) mi355_poisson3d_point(Q, i, 1, in[0], in[1], out[0]);
  return 0;
}

CEED_QFUNCTION(Vector3Poisson1DApply)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++)
    for (CeedInt c = 0; c < 3; c++) out[0][c * Q + i] = in[0][c * Q + i] * in[1][i];
  return 0;
}

CEED_QFUNCTION(Vector3Poisson2DApply)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++) mi355_poisson2d_point(Q, i, 3, in[0], in[1], out[0]);
  return 0;
}

CEED_QFUNCTION(Vector3Poisson3DApply)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++) mi355_poisson3d_point(Q, i, 3, in[0], in[1], out[0]);
  return 0;
}

#endif /* CEED_MI355_GALLERY_H */
