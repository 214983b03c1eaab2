/*
 * ceed_oracle.c -- CPU oracle (test infrastructure only; never the product).
 *
 * Restates the reference CPU backend /cpu/self/ref/serial for the 3D tensor
 * hex operator path.  Arithmetic is kept in the reference's order so that,
 * compiled with -ffp-contract=off, results match a reference build of the
 * same sources bit for bit (checked in tests/test_oracle.py against
 * oracle/_ref and against the committed fixtures in tests/golden/).
 */
#include "ceed_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EPSILON 1e-16 /* CEED_EPSILON for FP64, include/ceed/ceed-f64.h:20 */

static int ipow(int b, int e) {
  int r = 1;
  for (int i = 0; i < e; i++) r *= b;
  return r;
}

/* ======================================================================== */
/* Quadrature                                                               */
/* ======================================================================== */

/* Legendre P_n(x) by the three-term recurrence; returns P_n, stores P_{n-1}. */
static double legendre(int n, double x, double *p_prev) {
  double p0 = 1.0, p1 = x, p2 = 0.0;
  for (int j = 2; j <= n; j++) {
    p2 = (((double)(2 * j - 1)) * x * p1 - ((double)(j - 1)) * p0) / ((double)(j));
    p0 = p1;
    p1 = p2;
  }
  *p_prev = p0;
  return p2;
}

/* Gauss-Legendre nodes/weights: Newton on P_Q from the Chebyshev guess.
 * Restates interface/ceed-basis.c:2529-2573 (CeedGaussQuadrature). */
int orc_gauss_quadrature(int Q, double *q_ref, double *q_weight) {
  const double pi = 4.0 * atan(1.0);
  for (int i = 0; i <= Q / 2; i++) {
    double x = cos(pi * (double)(2 * i + 1) / ((double)(2 * Q)));
    double pm1, p = legendre(Q, x, &pm1);
    double dp = (x * p - pm1) * (double)Q / (x * x - 1.0);
    x = x - p / dp;
    for (int k = 0; k < 100 && fabs(p) > 10 * ORC_EPSILON; k++) {
      p  = legendre(Q, x, &pm1);
      dp = (x * p - pm1) * (double)Q / (x * x - 1.0);
      x  = x - p / dp;
    }
    const double w = 2.0 / ((1.0 - x * x) * dp * dp);
    q_weight[i] = w;
    q_weight[Q - 1 - i] = w;
    q_ref[i] = -x;
    q_ref[Q - 1 - i] = x;
  }
  return 0;
}

/* Gauss-Lobatto nodes/weights: Newton on P'_{Q-1}.
 * Restates interface/ceed-basis.c:2581-2636 (CeedLobattoQuadrature). */
int orc_lobatto_quadrature(int Q, double *q_ref, double *q_weight) {
  const double pi = 4.0 * atan(1.0);
  if (Q < 2) return -1;
  double w = 2.0 / ((double)(Q * (Q - 1)));
  if (q_weight) {
    q_weight[0] = w;
    q_weight[Q - 1] = w;
  }
  q_ref[0] = -1.0;
  q_ref[Q - 1] = 1.0;
  for (int i = 1; i <= (Q - 1) / 2; i++) {
    double x = cos(pi * (double)(i) / (double)(Q - 1));
    double pm1, p = legendre(Q - 1, x, &pm1); /* interior nodes exist only for Q >= 3 */
    double dp  = (x * p - pm1) * (double)Q / (x * x - 1.0);
    double d2p = (2 * x * dp - (double)(Q * (Q - 1)) * p) / (1.0 - x * x);
    x = x - dp / d2p;
    for (int k = 0; k < 100 && fabs(dp) > 10 * ORC_EPSILON; k++) {
      p = legendre(Q - 1, x, &pm1);
      dp  = (x * p - pm1) * (double)Q / (x * x - 1.0);
      d2p = (2 * x * dp - (double)(Q * (Q - 1)) * p) / (1.0 - x * x);
      x   = x - dp / d2p;
    }
    w = 2.0 / (((double)(Q * (Q - 1))) * p * p);
    if (q_weight) {
      q_weight[i] = w;
      q_weight[Q - 1 - i] = w;
    }
    q_ref[i] = -x;
    q_ref[Q - 1 - i] = x;
  }
  return 0;
}

/* Lagrange interpolation/derivative matrices (Q x P, row-major) on GLL nodes
 * by Fornberg's recursion (Fornberg 1998, "Calculation of weights in finite
 * difference formulas").  Restates interface/ceed-basis.c:1617-1690. */
int orc_lagrange_1d(int P, int Q, int quad_mode, double *interp, double *grad, double *q_ref, double *q_weight) {
  double *nodes = (double *)calloc((size_t)P, sizeof(double));
  int     ierr  = 0;
  memset(interp, 0, sizeof(double) * P * Q);
  memset(grad, 0, sizeof(double) * P * Q);
  if (P < 2 || orc_lobatto_quadrature(P, nodes, NULL)) {
    free(nodes);
    return -1;
  }
  ierr = quad_mode == ORC_GAUSS ? orc_gauss_quadrature(Q, q_ref, q_weight) : orc_lobatto_quadrature(Q, q_ref, q_weight);
  if (ierr) {
    free(nodes);
    return ierr;
  }
  for (int i = 0; i < Q; i++) {
    double *b = &interp[i * P], *g = &grad[i * P];
    double  c1 = 1.0, c3 = nodes[0] - q_ref[i];
    b[0] = 1.0;
    for (int j = 1; j < P; j++) {
      double c2 = 1.0, c4 = c3;
      c3 = nodes[j] - q_ref[i];
      for (int k = 0; k < j; k++) {
        const double dx = nodes[j] - nodes[k];
        c2 *= dx;
        if (k == j - 1) {
          g[j] = c1 * (b[k] - c4 * g[k]) / c2;
          b[j] = -c1 * c4 * b[k] / c2;
        }
        g[k] = (c3 * g[k] - b[k]) / dx;
        b[k] = c3 * b[k] / dx;
      }
      c1 = c2;
    }
  }
  free(nodes);
  return 0;
}

/* Householder reflection A <- (I - b v v^T) A on an m x n block with strides.
 * Restates interface/ceed-basis.c:104-113. */
static void householder_reflect(double *A, const double *v, double b, int m, int n, int row, int col) {
  for (int j = 0; j < n; j++) {
    double w = A[j * col];
    for (int i = 1; i < m; i++) w += v[i] * A[i * row + j * col];
    A[j * col] -= b * w;
    for (int i = 1; i < m; i++) A[i * row + j * col] -= b * w * v[i];
  }
}

/* Householder QR of an m x n row-major matrix in place.
 * Restates interface/ceed-basis.c:1188-1220 (CeedQRFactorization). */
static int qr_factor(double *mat, double *tau, int m, int n) {
  double *v = (double *)calloc((size_t)m, sizeof(double));
  if (n > m) {
    free(v);
    return -1;
  }
  for (int i = 0; i < n; i++) {
    if (i >= m - 1) {
      tau[i] = 0.;
      break;
    }
    double sigma = 0.0;
    v[i] = mat[i + n * i];
    for (int j = i + 1; j < m; j++) {
      v[j] = mat[i + n * j];
      sigma += v[j] * v[j];
    }
    const double nrm = sqrt(v[i] * v[i] + sigma);
    const double Rii = -copysign(nrm, v[i]);
    v[i] -= Rii;
    tau[i] = 2 * v[i] * v[i] / (v[i] * v[i] + sigma);
    for (int j = i + 1; j < m; j++) v[j] /= v[i];
    householder_reflect(&mat[i * n + i + 1], &v[i], tau[i], m - i, n - i - 1, n, 1);
    mat[i + n * i] = Rii;
    for (int j = i + 1; j < m; j++) mat[i + n * j] = v[j];
  }
  free(v);
  return 0;
}

/* Apply Q^T of a factored QR to A (m x n).  Restates ceed-basis.c:1240-1252. */
static void householder_apply_qt(double *A, const double *Qf, const double *tau, int m, int n, int k, int row, int col) {
  double *v = (double *)calloc((size_t)m, sizeof(double));
  for (int i = 0; i < k; i++) {
    for (int j = i + 1; j < m; j++) v[j] = Qf[j * k + i];
    householder_reflect(&A[i * row], &v[i], tau[i], m - i, n, row, col);
  }
  free(v);
}

/* Pseudoinverse of a full-column-rank m x n matrix, R^-1 Q^T.
 * Restates interface/ceed-basis.c:1268-1300 (CeedMatrixPseudoinverse). */
int orc_pseudoinverse(const double *mat, int m, int n, double *pinv) {
  double *tau = (double *)calloc((size_t)m, sizeof(double));
  double *I   = (double *)calloc((size_t)m * m, sizeof(double));
  double *A   = (double *)malloc(sizeof(double) * m * n);
  memcpy(A, mat, sizeof(double) * m * n);
  int ierr = qr_factor(A, tau, m, n);
  if (!ierr) {
    for (int i = 0; i < m; i++) I[i * m + i] = 1.0;
    householder_apply_qt(I, A, tau, m, m, n, m, 1);
    for (int j = 0; j < m; j++) {
      pinv[j + m * (n - 1)] = I[j + m * (n - 1)] / A[n * n - 1];
      for (int i = n - 2; i >= 0; i--) {
        pinv[j + m * i] = I[j + m * i];
        for (int kk = i + 1; kk < n; kk++) pinv[j + m * i] -= A[kk + n * i] * pinv[j + m * kk];
        pinv[j + m * i] /= A[i + n * i];
      }
    }
  }
  free(tau);
  free(I);
  free(A);
  return ierr;
}

/* Collocated derivative D_Q = G * pinv(B) (Q x Q).
 * Restates interface/ceed-basis.c:750-771 and the product at :1163-1173. */
int orc_collocated_grad(int P, int Q, const double *interp, const double *grad, double *collo) {
  double *pinv = (double *)calloc((size_t)P * Q, sizeof(double));
  int     ierr = orc_pseudoinverse(interp, Q, P, pinv);
  if (!ierr) {
    for (int i = 0; i < Q; i++) {
      for (int j = 0; j < Q; j++) {
        double sum = 0;
        for (int k = 0; k < P; k++) sum += grad[k + i * P] * pinv[j + k * Q];
        collo[j + i * Q] = sum;
      }
    }
  }
  free(pinv);
  return ierr;
}

/* Square basis whose interpolation is the identity.  Restates ceed-basis.c:840-854. */
int orc_is_collocated(int P, int Q, const double *interp) {
  if (P != Q) return 0;
  int c = 1;
  for (int i = 0; i < P; i++) {
    c = c && (fabs(interp[i + P * i] - 1.0) < 10 * ORC_EPSILON);
    for (int j = 0; j < Q; j++)
      if (j != i) c = c && (fabs(interp[j + P * i]) < 10 * ORC_EPSILON);
  }
  return c;
}

/* ======================================================================== */
/* Tensor contraction and basis apply                                      */
/* ======================================================================== */

/* v[a][j][c] (+)= sum_b t[j][b] u[a][b][c]   (t is J x B; transposed: B x J).
 * Restates backends/ref/ceed-ref-tensor.c:16-37. */
void orc_tensor_contract(int A, int B, int C, int J, const double *t, int t_mode, int add, const double *u, double *v) {
  const int s0 = t_mode == ORC_TRANSPOSE ? 1 : B, s1 = t_mode == ORC_TRANSPOSE ? J : 1;
  if (!add) memset(v, 0, sizeof(double) * (size_t)A * J * C);
  for (int a = 0; a < A; a++)
    for (int b = 0; b < B; b++)
      for (int j = 0; j < J; j++) {
        const double tq = t[j * s0 + b * s1];
        for (int c = 0; c < C; c++) v[(a * J + j) * C + c] += tq * u[(a * B + b) * C + c];
      }
}

/* Tensor H1 basis apply on num_elem elements, element index fastest
 * (u: [comp][node][elem], v: [dim][comp][qpt][elem] for GRAD).
 * Restates backends/ref/ceed-ref-basis.c:19-201 (tensor branch). */
int orc_basis_apply(int dim, int num_comp, int P1, int Q1, const double *interp, const double *grad, const double *q_weight,
                    int num_elem, int t_mode, int eval_mode, int apply_add, const double *u, double *v) {
  const int add = apply_add || t_mode == ORC_TRANSPOSE;
  const int num_nodes = ipow(P1, dim), num_qpts = ipow(Q1, dim);
  const int collocated = orc_is_collocated(P1, Q1, interp);
  const int mx = P1 > Q1 ? P1 : Q1;
  const size_t tmp_len = (size_t)num_elem * num_comp * Q1 * ipow(mx, dim - 1) + (size_t)num_elem * num_comp * ipow(mx, dim);

  if (t_mode == ORC_TRANSPOSE && !apply_add) {
    size_t len = (size_t)num_elem * num_comp * num_nodes;
    memset(v, 0, sizeof(double) * len);
  }
  double *tmp0 = (double *)calloc(tmp_len, sizeof(double));
  double *tmp1 = (double *)calloc(tmp_len, sizeof(double));
  double *ival = (double *)calloc(tmp_len, sizeof(double));
  double *tmp[2] = {tmp0, tmp1};
  int ierr = 0;

  switch (eval_mode) {
    case ORC_EVAL_INTERP: {
      if (collocated) {
        memcpy(v, u, sizeof(double) * (size_t)num_elem * num_comp * num_nodes);
        break;
      }
      int P = P1, Q = Q1;
      if (t_mode == ORC_TRANSPOSE) P = Q1, Q = P1;
      int pre = num_comp * ipow(P, dim - 1), post = num_elem;
      for (int d = 0; d < dim; d++) {
        orc_tensor_contract(pre, P, post, Q, interp, t_mode, add && d == dim - 1, d == 0 ? u : tmp[d % 2],
                            d == dim - 1 ? v : tmp[(d + 1) % 2]);
        pre /= P;
        post *= Q;
      }
    } break;
    case ORC_EVAL_GRAD: {
      int P = P1, Q = Q1;
      if (t_mode == ORC_TRANSPOSE) P = Q1, Q = Q1;
      int pre = num_comp * ipow(P, dim - 1), post = num_elem;
      if (Q1 >= P1 && !collocated) {
        double *collo = (double *)malloc(sizeof(double) * Q1 * Q1);
        ierr = orc_collocated_grad(P1, Q1, interp, grad, collo);
        const int nt = t_mode == ORC_NOTRANSPOSE;
        /* interpolate to qpts (NOTRANSPOSE) / collocated grad^T (TRANSPOSE) */
        for (int d = 0; d < dim; d++) {
          orc_tensor_contract(pre, P, post, Q, nt ? interp : collo, t_mode, !nt && d > 0,
                              nt ? (d == 0 ? u : tmp[d % 2]) : &u[(size_t)d * num_qpts * num_comp * num_elem],
                              nt ? (d == dim - 1 ? ival : tmp[(d + 1) % 2]) : ival);
          pre /= P;
          post *= Q;
        }
        /* collocated grad (NOTRANSPOSE) / interp^T to nodes (TRANSPOSE) */
        P = Q1, Q = Q1;
        if (!nt) P = Q1, Q = P1;
        pre = num_comp * ipow(P, dim - 1), post = num_elem;
        for (int d = 0; d < dim; d++) {
          orc_tensor_contract(pre, P, post, Q, nt ? collo : interp, t_mode, (nt && apply_add) || (!nt && d == dim - 1),
                              nt ? ival : (d == 0 ? ival : tmp[d % 2]),
                              nt ? &v[(size_t)d * num_qpts * num_comp * num_elem] : (d == dim - 1 ? v : tmp[(d + 1) % 2]));
          pre /= P;
          post *= Q;
        }
        free(collo);
      } else if (collocated) {
        for (int d = 0; d < dim; d++) {
          orc_tensor_contract(pre, P, post, Q, grad, t_mode, add && d > 0,
                              t_mode == ORC_NOTRANSPOSE ? u : &u[(size_t)d * num_comp * num_qpts * num_elem],
                              t_mode == ORC_TRANSPOSE ? v : &v[(size_t)d * num_comp * num_qpts * num_elem]);
          pre /= P;
          post *= Q;
        }
      } else { /* P > Q: dim^2 contractions */
        if (t_mode == ORC_TRANSPOSE) P = Q1, Q = P1;
        for (int p = 0; p < dim; p++) {
          int pr = num_comp * ipow(P, dim - 1), po = num_elem;
          for (int d = 0; d < dim; d++) {
            orc_tensor_contract(pr, P, po, Q, p == d ? grad : interp, t_mode, add && d == dim - 1,
                                d == 0 ? (t_mode == ORC_NOTRANSPOSE ? u : &u[(size_t)p * num_comp * num_qpts * num_elem]) : tmp[d % 2],
                                d == dim - 1 ? (t_mode == ORC_TRANSPOSE ? v : &v[(size_t)p * num_comp * num_qpts * num_elem])
                                             : tmp[(d + 1) % 2]);
            pr /= P;
            po *= Q;
          }
        }
      }
    } break;
    case ORC_EVAL_WEIGHT: {
      if (t_mode != ORC_NOTRANSPOSE) {
        ierr = -1;
        break;
      }
      const int Q = Q1;
      for (int d = 0; d < dim; d++) {
        const int pre = ipow(Q, dim - d - 1), post = ipow(Q, d);
        for (int i = 0; i < pre; i++)
          for (int j = 0; j < Q; j++)
            for (int k = 0; k < post; k++) {
              const size_t idx = ((size_t)(i * Q + j) * post + k) * num_elem;
              const double w   = q_weight[j] * (d == 0 ? 1 : v[idx]);
              for (int e = 0; e < num_elem; e++) v[idx + e] = w;
            }
      }
    } break;
    default:
      ierr = -1;
  }
  free(tmp0);
  free(tmp1);
  free(ival);
  return ierr;
}

/* ======================================================================== */
/* Element restriction (block size 1, /cpu/self/ref/serial)                 */
/* ======================================================================== */

/* Offset restriction.  E-vector layout [elem][comp][node].
 * NOTRANSPOSE restates ceed-ref-restriction.c:59-75;
 * TRANSPOSE (sum into v, ascending elem, comp, node) restates :220-242. */
void orc_restriction_offset(int num_elem, int elem_size, int num_comp, int comp_stride, const int32_t *offsets, int t_mode,
                            const double *u, double *v) {
  for (size_t e = 0; e < (size_t)num_elem; e++)
    for (size_t k = 0; k < (size_t)num_comp; k++)
      for (size_t i = 0; i < (size_t)elem_size; i++) {
        const size_t ev = elem_size * (k + e * num_comp) + i;
        const size_t lv = (size_t)offsets[i + e * elem_size] + k * comp_stride;
        if (t_mode == ORC_NOTRANSPOSE) v[ev] = u[lv];
        else v[lv] += u[ev];
      }
}

/* Strided restriction.  Restates ceed-ref-restriction.c:19-57 and :181-218. */
void orc_restriction_strided(int num_elem, int elem_size, int num_comp, const int32_t *strides, int t_mode, const double *u,
                             double *v) {
  for (size_t e = 0; e < (size_t)num_elem; e++)
    for (size_t k = 0; k < (size_t)num_comp; k++)
      for (size_t n = 0; n < (size_t)elem_size; n++) {
        const size_t ev = e * elem_size * num_comp + k * elem_size + n;
        const size_t lv = n * strides[0] + k * strides[1] + e * strides[2];
        if (t_mode == ORC_NOTRANSPOSE) v[ev] = u[lv];
        else v[lv] += u[ev];
      }
}

/* ======================================================================== */
/* Gallery QFunctions                                                       */
/* ======================================================================== */

/* q = w det(J).  Restates include/ceed/jit-source/gallery/ceed-mass3dbuild.h. */
int orc_qf_mass3d_build(void *ctx, int Q, const double *const *in, double *const *out) {
  const double *J = in[0], *w = in[1];
  double       *qd = out[0];
#define JJ(a, b) J[((a) * 3 + (b)) * Q + i]
  for (int i = 0; i < Q; i++) {
    qd[i] = (JJ(0, 0) * (JJ(1, 1) * JJ(2, 2) - JJ(1, 2) * JJ(2, 1)) - JJ(0, 1) * (JJ(1, 0) * JJ(2, 2) - JJ(1, 2) * JJ(2, 0)) +
             JJ(0, 2) * (JJ(1, 0) * JJ(2, 1) - JJ(1, 1) * JJ(2, 0))) *
            w[i];
  }
  return 0;
}

/* v = u q.  Restates gallery/ceed-massapply.h. */
int orc_qf_mass_apply(void *ctx, int Q, const double *const *in, double *const *out) {
  for (int i = 0; i < Q; i++) out[0][i] = in[0][i] * in[1][i];
  return 0;
}

/* Voigt {xx, yy, zz, yz, xz, xy} of (w / det J) adj(J) adj(J)^T.
 * Restates gallery/ceed-poisson3dbuild.h. */
int orc_qf_poisson3d_build(void *ctx, int Q, const double *const *in, double *const *out) {
  const double *J = in[0], *w = in[1];
  double       *qd = out[0];
  for (int i = 0; i < Q; i++) {
    double A[3][3];
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        A[k][j] = JJ((k + 1) % 3, (j + 1) % 3) * JJ((k + 2) % 3, (j + 2) % 3) - JJ((k + 2) % 3, (j + 1) % 3) * JJ((k + 1) % 3, (j + 2) % 3);
    const double qw = w[i] / (JJ(0, 0) * A[0][0] + JJ(0, 1) * A[0][1] + JJ(0, 2) * A[0][2]);
    qd[0 * Q + i] = qw * (A[0][0] * A[0][0] + A[0][1] * A[0][1] + A[0][2] * A[0][2]);
    qd[1 * Q + i] = qw * (A[1][0] * A[1][0] + A[1][1] * A[1][1] + A[1][2] * A[1][2]);
    qd[2 * Q + i] = qw * (A[2][0] * A[2][0] + A[2][1] * A[2][1] + A[2][2] * A[2][2]);
    qd[3 * Q + i] = qw * (A[1][0] * A[2][0] + A[1][1] * A[2][1] + A[1][2] * A[2][2]);
    qd[4 * Q + i] = qw * (A[0][0] * A[2][0] + A[0][1] * A[2][1] + A[0][2] * A[2][2]);
    qd[5 * Q + i] = qw * (A[0][0] * A[1][0] + A[0][1] * A[1][1] + A[0][2] * A[1][2]);
  }
#undef JJ
  return 0;
}

/* vg_j = sum_i ug_i S_ij, S from Voigt {0:xx 1:yy 2:zz 3:yz 4:xz 5:xy}.
 * Restates gallery/ceed-poisson3dapply.h. */
int orc_qf_poisson3d_apply(void *ctx, int Q, const double *const *in, double *const *out) {
  const double *ug = in[0], *qd = in[1];
  double       *vg = out[0];
  for (int i = 0; i < Q; i++) {
    const double S[3][3] = {
        {qd[0 * Q + i], qd[5 * Q + i], qd[4 * Q + i]},
        {qd[5 * Q + i], qd[1 * Q + i], qd[3 * Q + i]},
        {qd[4 * Q + i], qd[3 * Q + i], qd[2 * Q + i]},
    };
    for (int j = 0; j < 3; j++) vg[j * Q + i] = (ug[0 * Q + i] * S[0][j] + ug[1 * Q + i] * S[1][j] + ug[2 * Q + i] * S[2][j]);
  }
  return 0;
}

/* Restates gallery/ceed-vectormassapply.h (3 components, shared q). */
int orc_qf_vector3_mass_apply(void *ctx, int Q, const double *const *in, double *const *out) {
  for (int i = 0; i < Q; i++)
    for (int c = 0; c < 3; c++) out[0][c * Q + i] = in[0][c * Q + i] * in[1][i];
  return 0;
}

/* Restates gallery/ceed-vectorpoisson3dapply.h: ug[dir][comp][Q]. */
int orc_qf_vector3_poisson3d_apply(void *ctx, int Q, const double *const *in, double *const *out) {
  const double *ug = in[0], *qd = in[1];
  double       *vg = out[0];
  for (int i = 0; i < Q; i++) {
    const double S[3][3] = {
        {qd[0 * Q + i], qd[5 * Q + i], qd[4 * Q + i]},
        {qd[5 * Q + i], qd[1 * Q + i], qd[3 * Q + i]},
        {qd[4 * Q + i], qd[3 * Q + i], qd[2 * Q + i]},
    };
    for (int j = 0; j < 3; j++)
      for (int c = 0; c < 3; c++)
        vg[(j * 3 + c) * Q + i] = (ug[(0 * 3 + c) * Q + i] * S[0][j] + ug[(1 * 3 + c) * Q + i] * S[1][j] + ug[(2 * 3 + c) * Q + i] * S[2][j]);
  }
  return 0;
}

/* Variable-coefficient diffusion: the Poisson3DApply algebra scaled by kappa
 * (in[2], one value per point).  No gallery counterpart exists in the
 * reference; this is the pointwise D of -div(kappa grad u). */
int orc_qf_diffusion_kappa_apply(void *ctx, int Q, const double *const *in, double *const *out) {
  const double *ug = in[0], *qd = in[1], *kappa = in[2];
  double       *vg = out[0];
  for (int i = 0; i < Q; i++) {
    const double S[3][3] = {
        {qd[0 * Q + i], qd[5 * Q + i], qd[4 * Q + i]},
        {qd[5 * Q + i], qd[1 * Q + i], qd[3 * Q + i]},
        {qd[4 * Q + i], qd[3 * Q + i], qd[2 * Q + i]},
    };
    for (int j = 0; j < 3; j++)
      vg[j * Q + i] = kappa[i] * (ug[0 * Q + i] * S[0][j] + ug[1 * Q + i] * S[1][j] + ug[2 * Q + i] * S[2][j]);
  }
  return 0;
}

/* ======================================================================== */
/* Operator: element loop with full E-vectors                              */
/* ======================================================================== */

static size_t evec_len(const orc_field *f, int num_elem) { return (size_t)num_elem * f->elem_size * f->num_comp; }

static void restrict_field(const orc_field *f, int num_elem, int t_mode, const double *u, double *v) {
  if (f->offsets) orc_restriction_offset(num_elem, f->elem_size, f->num_comp, f->comp_stride, f->offsets, t_mode, u, v);
  else orc_restriction_strided(num_elem, f->elem_size, f->num_comp, f->strides, t_mode, u, v);
}

/* v += A u with ref/serial semantics: input restrictions to full E-vectors,
 * per element basis -> QFunction -> transpose basis, then output transpose
 * restrictions.  Restates backends/ref/ceed-ref-operator.c:381-470 (and the
 * per-field setup at :19-142; duplicate-restriction sharing is not needed
 * for the single-active-field operators used here). */
int orc_operator_apply_add(int num_elem, int num_qpts, int num_in, const orc_field *in, int num_out, const orc_field *out,
                           orc_qfunction qf, void *ctx) {
  double *e_full[32] = {0}, *q_in[16] = {0}, *q_out[16] = {0};
  int     ierr = 0;
  if (num_in > 16 || num_out > 16) return -1;

  for (int i = 0; i < num_in; i++) {
    const orc_field *f = &in[i];
    q_in[i] = (double *)calloc((size_t)num_qpts * f->size, sizeof(double));
    if (f->eval_mode == ORC_EVAL_WEIGHT) {
      ierr |= orc_basis_apply(f->dim, 1, f->P, f->Q, f->interp, f->grad, f->q_weight, 1, ORC_NOTRANSPOSE, ORC_EVAL_WEIGHT, 0, NULL, q_in[i]);
      continue;
    }
    e_full[i] = (double *)calloc(evec_len(f, num_elem), sizeof(double));
    restrict_field(f, num_elem, ORC_NOTRANSPOSE, f->vec, e_full[i]);
  }
  for (int i = 0; i < num_out; i++) {
    const orc_field *f = &out[i];
    q_out[i] = (double *)calloc((size_t)num_qpts * f->size, sizeof(double));
    e_full[16 + i] = (double *)calloc(evec_len(f, num_elem), sizeof(double));
  }

  for (int e = 0; e < num_elem; e++) {
    for (int i = 0; i < num_in; i++) {
      const orc_field *f = &in[i];
      switch (f->eval_mode) {
        case ORC_EVAL_NONE:
          memcpy(q_in[i], &e_full[i][(size_t)e * num_qpts * f->size], sizeof(double) * num_qpts * f->size);
          break;
        case ORC_EVAL_INTERP:
        case ORC_EVAL_GRAD:
          ierr |= orc_basis_apply(f->dim, f->num_comp, f->P, f->Q, f->interp, f->grad, f->q_weight, 1, ORC_NOTRANSPOSE, f->eval_mode, 0,
                                  &e_full[i][(size_t)e * f->elem_size * f->num_comp], q_in[i]);
          break;
        default:
          break;
      }
    }
    ierr |= qf(ctx, num_qpts, (const double *const *)q_in, q_out);
    for (int i = 0; i < num_out; i++) {
      const orc_field *f = &out[i];
      double          *ee = &e_full[16 + i][(size_t)e * f->elem_size * f->num_comp];
      if (f->eval_mode == ORC_EVAL_NONE) {
        memcpy(ee, q_out[i], sizeof(double) * num_qpts * f->size);
      } else {
        ierr |= orc_basis_apply(f->dim, f->num_comp, f->P, f->Q, f->interp, f->grad, f->q_weight, 1, ORC_TRANSPOSE, f->eval_mode, 0, q_out[i],
                                ee);
      }
    }
  }
  for (int i = 0; i < num_out; i++) restrict_field(&out[i], num_elem, ORC_TRANSPOSE, e_full[16 + i], out[i].vec);

  for (int i = 0; i < 16; i++) {
    free(q_in[i]);
    free(q_out[i]);
  }
  for (int i = 0; i < 32; i++) free(e_full[i]);
  return ierr;
}

/* ======================================================================== */
/* Hex-mesh wrappers (mirror examples/ceed/ex1-volume.c:160-248 set-up)     */
/* ======================================================================== */

typedef struct {
  double interp[16 * 16], grad[16 * 16], qref[16], qweight[16];
} basis1d;

static int make_basis(int P, int Q, int quad_mode, basis1d *b) {
  if (P > 16 || Q > 16) return -1;
  return orc_lagrange_1d(P, Q, quad_mode, b->interp, b->grad, b->qref, b->qweight);
}

int orc_hex_build_qdata(int kind, int num_elem, int P_mesh, int Q, int quad_mode, const int32_t *mesh_offsets, int mesh_comp_stride,
                        const double *coords, double *q_data) {
  basis1d mb;
  if (make_basis(P_mesh, Q, quad_mode, &mb)) return -1;
  const int nq = Q * Q * Q, qcomp = kind == 0 ? 1 : 6;
  orc_field in[2] = {0}, out[1] = {0};
  in[0] = (orc_field){.eval_mode = ORC_EVAL_GRAD, .size = 9, .num_comp = 3, .elem_size = P_mesh * P_mesh * P_mesh,
                      .comp_stride = mesh_comp_stride, .offsets = mesh_offsets, .dim = 3, .P = P_mesh, .Q = Q,
                      .interp = mb.interp, .grad = mb.grad, .q_weight = mb.qweight, .vec = (double *)coords};
  in[1] = (orc_field){.eval_mode = ORC_EVAL_WEIGHT, .size = 1, .dim = 3, .P = P_mesh, .Q = Q, .interp = mb.interp, .grad = mb.grad,
                      .q_weight = mb.qweight};
  out[0] = (orc_field){.eval_mode = ORC_EVAL_NONE, .size = qcomp, .num_comp = qcomp, .elem_size = nq, .offsets = NULL,
                       .strides = {1, nq, nq * qcomp}, .vec = q_data};
  memset(q_data, 0, sizeof(double) * (size_t)num_elem * nq * qcomp);
  return orc_operator_apply_add(num_elem, nq, 2, in, 1, out, kind == 0 ? orc_qf_mass3d_build : orc_qf_poisson3d_build, NULL);
}

int orc_hex_apply(int kind, int num_comp, int num_elem, int P, int Q, int quad_mode, const int32_t *offsets, int comp_stride,
                  const double *q_data, const double *u, double *v) {
  basis1d sb;
  if (make_basis(P, Q, quad_mode, &sb)) return -1;
  const int    nq = Q * Q * Q, qcomp = kind == 0 ? 1 : 6, emode = kind == 0 ? ORC_EVAL_INTERP : ORC_EVAL_GRAD;
  const int    fsize = kind == 0 ? num_comp : 3 * num_comp;
  orc_qfunction qf;
  if (kind == 0) qf = num_comp == 1 ? orc_qf_mass_apply : orc_qf_vector3_mass_apply;
  else qf = num_comp == 1 ? orc_qf_poisson3d_apply : orc_qf_vector3_poisson3d_apply;
  if (num_comp != 1 && num_comp != 3) return -1;
  orc_field act = {.eval_mode = emode, .size = fsize, .num_comp = num_comp, .elem_size = P * P * P, .comp_stride = comp_stride,
                   .offsets = offsets, .dim = 3, .P = P, .Q = Q, .interp = sb.interp, .grad = sb.grad, .q_weight = sb.qweight};
  orc_field in[2], out[1];
  in[0] = act;
  in[0].vec = (double *)u;
  in[1] = (orc_field){.eval_mode = ORC_EVAL_NONE, .size = qcomp, .num_comp = qcomp, .elem_size = nq, .offsets = NULL,
                      .strides = {1, nq, nq * qcomp}, .vec = (double *)q_data};
  out[0] = act;
  out[0].vec = v;
  return orc_operator_apply_add(num_elem, nq, 2, in, 1, out, qf, NULL);
}

int orc_hex_apply_kappa(int num_elem, int P, int Q, int quad_mode, const int32_t *offsets, const double *q_data, const double *kappa,
                        const double *u, double *v) {
  basis1d sb;
  if (make_basis(P, Q, quad_mode, &sb)) return -1;
  const int nq = Q * Q * Q;
  orc_field act = {.eval_mode = ORC_EVAL_GRAD, .size = 3, .num_comp = 1, .elem_size = P * P * P, .comp_stride = 1, .offsets = offsets,
                   .dim = 3, .P = P, .Q = Q, .interp = sb.interp, .grad = sb.grad, .q_weight = sb.qweight};
  orc_field in[3], out[1];
  in[0] = act;
  in[0].vec = (double *)u;
  in[1] = (orc_field){.eval_mode = ORC_EVAL_NONE, .size = 6, .num_comp = 6, .elem_size = nq, .strides = {1, nq, nq * 6},
                      .vec = (double *)q_data};
  in[2] = (orc_field){.eval_mode = ORC_EVAL_NONE, .size = 1, .num_comp = 1, .elem_size = nq, .strides = {1, nq, nq},
                      .vec = (double *)kappa};
  out[0] = act;
  out[0].vec = v;
  return orc_operator_apply_add(num_elem, nq, 3, in, 1, out, orc_qf_diffusion_kappa_apply, NULL);
}

int orc_hex_kappa_gradient(int num_elem, int P, int Q, int quad_mode, const int32_t *offsets, const double *q_data, const double *u,
                           const double *lam, double *dkappa) {
  basis1d sb;
  if (make_basis(P, Q, quad_mode, &sb)) return -1;
  const int nq = Q * Q * Q, es = P * P * P;
  double   *eu = (double *)malloc(sizeof(double) * es), *el = (double *)malloc(sizeof(double) * es);
  double   *gu = (double *)malloc(sizeof(double) * 3 * nq), *gl = (double *)malloc(sizeof(double) * 3 * nq);
  int       ierr = 0;
  for (int e = 0; e < num_elem; e++) {
    for (int i = 0; i < es; i++) {
      eu[i] = u[offsets[(size_t)e * es + i]];
      el[i] = lam[offsets[(size_t)e * es + i]];
    }
    ierr |= orc_basis_apply(3, 1, P, Q, sb.interp, sb.grad, sb.qweight, 1, ORC_NOTRANSPOSE, ORC_EVAL_GRAD, 0, eu, gu);
    ierr |= orc_basis_apply(3, 1, P, Q, sb.interp, sb.grad, sb.qweight, 1, ORC_NOTRANSPOSE, ORC_EVAL_GRAD, 0, el, gl);
    const double *qd = &q_data[(size_t)e * 6 * nq];
    for (int i = 0; i < nq; i++) {
      const double S[3][3] = {
          {qd[0 * nq + i], qd[5 * nq + i], qd[4 * nq + i]},
          {qd[5 * nq + i], qd[1 * nq + i], qd[3 * nq + i]},
          {qd[4 * nq + i], qd[3 * nq + i], qd[2 * nq + i]},
      };
      double s = 0;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) s += gu[a * nq + i] * S[a][b] * gl[b * nq + i];
      dkappa[(size_t)e * nq + i] = -s;
    }
  }
  free(eu);
  free(el);
  free(gu);
  free(gl);
  return ierr;
}
