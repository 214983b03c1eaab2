/* Pointwise operators of the ex-hex example: affine-mesh diffusion applied
 * through a user QFunction (run-time compiled on the GPU). */
#include <ceed/types.h>

/* dv = S du with S in Voigt order {xx, yy, zz, yz, xz, xy} */
CEED_QFUNCTION(ex_diffusion)(void *ctx, const CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out) {
  const CeedScalar *du = in[0], *q = in[1];
  for (CeedInt i = 0; i < Q; i++) {
    const CeedScalar a = du[i], b = du[Q + i], c = du[2 * Q + i];
    out[0][i]         = q[i] * a + q[5 * Q + i] * b + q[4 * Q + i] * c;
    out[0][Q + i]     = q[5 * Q + i] * a + q[Q + i] * b + q[3 * Q + i] * c;
    out[0][2 * Q + i] = q[4 * Q + i] * a + q[3 * Q + i] * b + q[2 * Q + i] * c;
  }
  return 0;
}
