/*
 * User QFunctions for the API tests: they are not gallery functions, so
 * operators built from them are compiled at run time, either into the fused
 * line-layout kernel (scalar active fields, at most one passive field) or as
 * the pointwise stage between the standalone restriction and basis kernels.
 */
#include <ceed/types.h>

struct ScaleContext {
  double scale;
};

/* qd = scale * w * det J  (context-dependent Mass3DBuild) */
CEED_QFUNCTION(scaled_mass_build)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const struct ScaleContext *c = (const struct ScaleContext *)ctx;
  const CeedScalar          *J = in[0], *w = in[1];
  for (CeedInt i = 0; i < Q; i++) {
    const CeedScalar m00 = J[0 * Q + i], m01 = J[1 * Q + i], m02 = J[2 * Q + i];
    const CeedScalar m10 = J[3 * Q + i], m11 = J[4 * Q + i], m12 = J[5 * Q + i];
    const CeedScalar m20 = J[6 * Q + i], m21 = J[7 * Q + i], m22 = J[8 * Q + i];
    const CeedScalar det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
    out[0][i]            = c->scale * det * w[i];
  }
  return 0;
}

/* v = qd * u */
CEED_QFUNCTION(user_mass)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  for (CeedInt i = 0; i < Q; i++) out[0][i] = in[1][i] * in[0][i];
  return 0;
}

/* dv_j = sum_d du_d S_dj with S in Voigt order {xx, yy, zz, yz, xz, xy} */
CEED_QFUNCTION(user_diffusion)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const CeedScalar *du = in[0], *qd = in[1];
  CeedScalar       *dv = out[0];
  for (CeedInt i = 0; i < Q; i++) {
    const CeedScalar g0 = du[0 * Q + i], g1 = du[1 * Q + i], g2 = du[2 * Q + i];
    dv[0 * Q + i]       = g0 * qd[0 * Q + i] + g1 * qd[5 * Q + i] + g2 * qd[4 * Q + i];
    dv[1 * Q + i]       = g0 * qd[5 * Q + i] + g1 * qd[1 * Q + i] + g2 * qd[3 * Q + i];
    dv[2 * Q + i]       = g0 * qd[4 * Q + i] + g1 * qd[3 * Q + i] + g2 * qd[2 * Q + i];
  }
  return 0;
}

/* mass plus diffusion in one pointwise operator: two active inputs and two
 * active outputs of different evaluation modes */
CEED_QFUNCTION(user_mass_diffusion)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const CeedScalar *u = in[0], *du = in[1], *qm = in[2], *qd = in[3];
  for (CeedInt i = 0; i < Q; i++) {
    const CeedScalar g0 = du[0 * Q + i], g1 = du[1 * Q + i], g2 = du[2 * Q + i];
    out[0][i]           = qm[i] * u[i];
    out[1][0 * Q + i]   = g0 * qd[0 * Q + i] + g1 * qd[5 * Q + i] + g2 * qd[4 * Q + i];
    out[1][1 * Q + i]   = g0 * qd[5 * Q + i] + g1 * qd[1 * Q + i] + g2 * qd[3 * Q + i];
    out[1][2 * Q + i]   = g0 * qd[4 * Q + i] + g1 * qd[3 * Q + i] + g2 * qd[2 * Q + i];
  }
  return 0;
}

/* v = scale * qd * u, scale from the context */
CEED_QFUNCTION(scaled_mass)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const struct ScaleContext *c = (const struct ScaleContext *)ctx;
  for (CeedInt i = 0; i < Q; i++) out[0][i] = c->scale * in[1][i] * in[0][i];
  return 0;
}

/* mass plus diffusion with one passive field of 7 values per point: the mass
 * factor, then the 6 diffusion factors in Voigt order */
CEED_QFUNCTION(user_mass_diffusion7)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const CeedScalar *u = in[0], *du = in[1], *q = in[2], *qd = in[2] + Q;
  for (CeedInt i = 0; i < Q; i++) {
    const CeedScalar g0 = du[0 * Q + i], g1 = du[1 * Q + i], g2 = du[2 * Q + i];
    out[0][i]           = q[i] * u[i];
    out[1][0 * Q + i]   = g0 * qd[0 * Q + i] + g1 * qd[5 * Q + i] + g2 * qd[4 * Q + i];
    out[1][1 * Q + i]   = g0 * qd[5 * Q + i] + g1 * qd[1 * Q + i] + g2 * qd[3 * Q + i];
    out[1][2 * Q + i]   = g0 * qd[4 * Q + i] + g1 * qd[3 * Q + i] + g2 * qd[2 * Q + i];
  }
  return 0;
}
