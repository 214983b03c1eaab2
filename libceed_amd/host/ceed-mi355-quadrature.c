/*
 * 1D quadrature and Lagrange basis matrices for the tensor H1 bases of the
 * /gpu/hip/mi355 backend (host code, FP64).
 *
 * CeedGaussQuadrature / CeedLobattoQuadrature keep the reference's public
 * signatures (include/ceed/ceed.h) and its Newton iteration on Legendre
 * polynomials (interface/ceed-basis.c:2529-2636); the Lagrange matrices use
 * Fornberg's recursion as interface/ceed-basis.c:1645-1666 does, so basis
 * matrices agree with the reference to the last bit when both are compiled
 * without FMA contraction.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ceed-mi355.h"

#define MI355_EPSILON 1e-16

/* P_n(x) by the three-term recurrence; *pm1 receives P_{n-1}. */
static double legendre_pn(int n, double x, double *pm1) {
  double a = 1.0, b = x, c = 0.0;
  for (int j = 2; j <= n; j++) {
    c = (((double)(2 * j - 1)) * x * b - ((double)(j - 1)) * a) / ((double)(j));
    a = b;
    b = c;
  }
  *pm1 = a;
  return c;
}

int CeedGaussQuadrature(CeedInt Q, CeedScalar *q_ref_1d, CeedScalar *q_weight_1d) {
  const double pi = 4.0 * atan(1.0);
  if (Q < 1) return CEED_MI355_ERROR_DIMENSION;
  for (int i = 0; i <= Q / 2; i++) {
    double x  = cos(pi * (double)(2 * i + 1) / ((double)(2 * Q)));
    double pm = 0, pq = legendre_pn(Q, x, &pm);
    double dp = (x * pq - pm) * (double)Q / (x * x - 1.0);
    x -= pq / dp;
    for (int it = 0; it < 100 && fabs(pq) > 10 * MI355_EPSILON; it++) {
      pq = legendre_pn(Q, x, &pm);
      dp = (x * pq - pm) * (double)Q / (x * x - 1.0);
      x -= pq / dp;
    }
    const double w = 2.0 / ((1.0 - x * x) * dp * dp);
    q_weight_1d[i] = q_weight_1d[Q - 1 - i] = w;
    q_ref_1d[i]                             = -x;
    q_ref_1d[Q - 1 - i]                     = x;
  }
  return 0;
}

int CeedLobattoQuadrature(CeedInt Q, CeedScalar *q_ref_1d, CeedScalar *q_weight_1d) {
  const double pi = 4.0 * atan(1.0);
  if (Q < 2) return CEED_MI355_ERROR_DIMENSION;
  const double wq = 2.0 / ((double)(Q * (Q - 1)));
  if (q_weight_1d) q_weight_1d[0] = q_weight_1d[Q - 1] = wq;
  q_ref_1d[0]     = -1.0;
  q_ref_1d[Q - 1] = 1.0;
  for (int i = 1; i <= (Q - 1) / 2; i++) {
    double x  = cos(pi * (double)(i) / (double)(Q - 1));
    double pm = 0, pq = legendre_pn(Q - 1, x, &pm);
    double d1 = (x * pq - pm) * (double)Q / (x * x - 1.0);
    double d2 = (2 * x * d1 - (double)(Q * (Q - 1)) * pq) / (1.0 - x * x);
    x -= d1 / d2;
    for (int it = 0; it < 100 && fabs(d1) > 10 * MI355_EPSILON; it++) {
      pq = legendre_pn(Q - 1, x, &pm);
      d1 = (x * pq - pm) * (double)Q / (x * x - 1.0);
      d2 = (2 * x * d1 - (double)(Q * (Q - 1)) * pq) / (1.0 - x * x);
      x -= d1 / d2;
    }
    const double w = 2.0 / (((double)(Q * (Q - 1))) * pq * pq);
    if (q_weight_1d) q_weight_1d[i] = q_weight_1d[Q - 1 - i] = w;
    q_ref_1d[i]         = -x;
    q_ref_1d[Q - 1 - i] = x;
  }
  return 0;
}

int CeedMi355LagrangeBasis1D(int32_t P, int32_t Q, int32_t quad_mode, double *interp_1d, double *grad_1d, double *q_ref_1d,
                             double *q_weight_1d) {
  if (P < 2 || Q < 1) return CEED_MI355_ERROR_DIMENSION;
  double *x = (double *)calloc((size_t)P, sizeof(double));
  int     ierr = CeedLobattoQuadrature(P, x, NULL);
  if (!ierr) ierr = quad_mode == 0 ? CeedGaussQuadrature(Q, q_ref_1d, q_weight_1d) : CeedLobattoQuadrature(Q, q_ref_1d, q_weight_1d);
  if (ierr) {
    free(x);
    return ierr;
  }
  memset(interp_1d, 0, sizeof(double) * P * Q);
  memset(grad_1d, 0, sizeof(double) * P * Q);
  /* Fornberg (1998): weights of the Lagrange interpolant and its derivative at q_ref[i] */
  for (int i = 0; i < Q; i++) {
    double *L = interp_1d + i * P, *dL = grad_1d + i * P;
    double  prod_prev = 1.0, dist = x[0] - q_ref_1d[i];
    L[0] = 1.0;
    for (int j = 1; j < P; j++) {
      double prod = 1.0, dist_prev = dist;
      dist = x[j] - q_ref_1d[i];
      for (int k = 0; k < j; k++) {
        const double h = x[j] - x[k];
        prod *= h;
        if (k == j - 1) {
          dL[j] = prod_prev * (L[k] - dist_prev * dL[k]) / prod;
          L[j]  = -prod_prev * dist_prev * L[k] / prod;
        }
        dL[k] = (dist * dL[k] - L[k]) / h;
        L[k]  = dist * L[k] / h;
      }
      prod_prev = prod;
    }
  }
  free(x);
  return 0;
}
