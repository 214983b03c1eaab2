// Standalone tensor-product H1 basis apply (CeedBasisApply drop-in), dim 1..3.
//
// One workgroup per element, one component at a time: the element's data is
// staged in LDS and contracted one dimension at a time.  The contraction and
// the choice of path (collocated gradient when Q >= P, direct when the basis is
// collocated, dim^2 contractions when P > Q) follow the reference CPU backend
// (backends/ref/ceed-ref-basis.c:19-201), so results agree to rounding.
//
// Layouts (element-major, this backend's choice):
//   nodes : [elem][comp][P^dim]
//   INTERP: [elem][comp][Q^dim]        GRAD: [elem][dim][comp][Q^dim]
//   WEIGHT: [elem][Q^dim]
// The fused operator kernels do not use this path.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "mi355_common.h"

struct CeedMi355Basis_private {
  int32_t dim, num_comp, P, Q;
  bool    collocated;   // P == Q and interp is the identity
  bool    collo_grad;   // Q >= P and not collocated: collocated-gradient path
  double *d_mats;       // interp (Q*P) | grad (Q*P) | collo (Q*Q) | q_weight (Q)
};

namespace {

constexpr int kMaxT = 10;  // max(P, Q) supported by the LDS sizing

// out[a][j][c] (+)= sum_b M(j, b) in[a][b][c];  M(j, b) = m[j*B + b], or m[b*J + j] when transposed.
__device__ void lds_contract(int A, int B, int C, int J, const double *m, bool transpose, bool add, const double *in, double *out) {
  const int n = A * J * C;
  for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
    const int c = idx % C, j = (idx / C) % J, a = idx / (C * J);
    double    s = add ? out[idx] : 0.0;
    for (int b = 0; b < B; b++) s += (transpose ? m[b * J + j] : m[j * B + b]) * in[(a * B + b) * C + c];
    out[idx] = s;
  }
  __syncthreads();
}

__device__ int ipow_d(int b, int e) {
  int r = 1;
  for (int i = 0; i < e; i++) r *= b;
  return r;
}

__global__ void __launch_bounds__(256) basis_tensor_kernel(int num_elem, int dim, int num_comp, int P1, int Q1, int collocated, int collo_grad,
                                                           int t_mode, int eval_mode, int add, const double *__restrict__ mats,
                                                           const double *__restrict__ u, double *__restrict__ v) {
  extern __shared__ double lds[];
  const int T = P1 > Q1 ? P1 : Q1, Tn = ipow_d(T, dim);
  double   *s_interp = lds, *s_grad = s_interp + Q1 * P1, *s_collo = s_grad + Q1 * P1;
  double   *buf0 = s_collo + Q1 * Q1, *buf1 = buf0 + Tn, *buf2 = buf1 + Tn, *acc = buf2 + Tn;
  for (int i = threadIdx.x; i < 2 * Q1 * P1 + Q1 * Q1; i += blockDim.x) lds[i] = mats[i];
  __syncthreads();
  const int  n_nodes = ipow_d(P1, dim), n_qpts = ipow_d(Q1, dim);
  const bool tr = t_mode == 1;

  for (int e = blockIdx.x; e < num_elem; e += gridDim.x) {
    for (int comp = 0; comp < num_comp; comp++) {
      // ---- load input for this component
      const int n_in = tr ? n_qpts : n_nodes;
      const int n_dir_in = (tr && eval_mode == 2) ? dim : 1;
      if (eval_mode == 2 && tr) {
        // gradient transpose input: dim slices, handled inside the GRAD branch
      } else {
        const double *src = &u[((size_t)e * num_comp + comp) * n_in];
        for (int i = threadIdx.x; i < n_in; i += blockDim.x) buf0[i] = src[i];
      }
      __syncthreads();
      const int n_out = tr ? n_nodes : n_qpts;
      double   *result = nullptr;  // LDS buffer holding the output of this component
      if (eval_mode == 1) {
        if (collocated) {
          result = buf0;
        } else {
          int P = tr ? Q1 : P1, Q = tr ? P1 : Q1, pre = ipow_d(P, dim - 1), post = 1;
          double *in = buf0, *out = buf1;
          for (int d = 0; d < dim; d++) {
            lds_contract(pre, P, post, Q, s_interp, tr, false, in, out);
            pre /= P;
            post *= Q;
            double *t = in;
            in = out;
            out = (t == buf0) ? buf2 : t;
          }
          result = in;
        }
        const size_t base = ((size_t)e * num_comp + comp) * n_out;
        for (int i = threadIdx.x; i < n_out; i += blockDim.x) v[base + i] = (add ? v[base + i] : 0.0) + result[i];
        __syncthreads();
      } else if (eval_mode == 2) {
        const size_t qstride = (size_t)num_comp * n_qpts;  // between directions inside one element
        const size_t ebase   = (size_t)e * dim * qstride;
        if (!tr) {
          // gradient to quadrature points: write one direction at a time
          if (collo_grad) {
            int     pre = ipow_d(P1, dim - 1), post = 1;
            double *in = buf0, *out = buf1;
            for (int d = 0; d < dim; d++) {
              lds_contract(pre, P1, post, Q1, s_interp, false, false, in, out);
              pre /= P1;
              post *= Q1;
              double *t = in;
              in = out;
              out = (t == buf0) ? buf2 : t;
            }
            // `in` holds the interpolated values; copy to acc so buffers can be reused
            for (int i = threadIdx.x; i < n_qpts; i += blockDim.x) acc[i] = in[i];
            __syncthreads();
            for (int d = 0; d < dim; d++) {
              const int pre_d = ipow_d(Q1, dim - 1 - d), post_d = ipow_d(Q1, d);
              lds_contract(pre_d, Q1, post_d, Q1, s_collo, false, false, acc, buf1);
              const size_t base = ebase + d * qstride + (size_t)comp * n_qpts;
              for (int i = threadIdx.x; i < n_qpts; i += blockDim.x) v[base + i] = (add ? v[base + i] : 0.0) + buf1[i];
              __syncthreads();
            }
          } else {
            // collocated basis (grad directly) or P > Q (dim^2 contractions)
            for (int p = 0; p < dim; p++) {
              int     pre = ipow_d(P1, dim - 1), post = 1;
              double *in = buf0, *out = buf1;
              for (int d = 0; d < dim; d++) {
                if (collocated && d != p) {
                  pre /= P1;
                  post *= Q1;
                  continue;
                }
                lds_contract(pre, P1, post, Q1, d == p ? s_grad : s_interp, false, false, in, out);
                pre /= P1;
                post *= Q1;
                double *t = in;
                in = out;
                out = (t == buf0) ? buf2 : t;
              }
              const size_t base = ebase + p * qstride + (size_t)comp * n_qpts;
              for (int i = threadIdx.x; i < n_qpts; i += blockDim.x) v[base + i] = (add ? v[base + i] : 0.0) + in[i];
              __syncthreads();
            }
          }
        } else {
          // transpose: sum over directions into nodes
          for (int i = threadIdx.x; i < n_nodes; i += blockDim.x) acc[i] = 0.0;
          __syncthreads();
          if (collo_grad) {
            // acc_q = sum_d D^T_d u_d  (at quadrature points), then interp^T
            double *qacc = buf2;
            for (int i = threadIdx.x; i < n_qpts; i += blockDim.x) qacc[i] = 0.0;
            __syncthreads();
            for (int d = 0; d < dim; d++) {
              const double *src = &u[ebase + d * qstride + (size_t)comp * n_qpts];
              for (int i = threadIdx.x; i < n_qpts; i += blockDim.x) buf0[i] = src[i];
              __syncthreads();
              const int pre_d = ipow_d(Q1, dim - 1 - d), post_d = ipow_d(Q1, d);
              lds_contract(pre_d, Q1, post_d, Q1, s_collo, true, true, buf0, qacc);
            }
            int     pre = ipow_d(Q1, dim - 1), post = 1;
            double *in = qacc, *out = buf0;
            for (int d = 0; d < dim; d++) {
              lds_contract(pre, Q1, post, P1, s_interp, true, false, in, out);
              pre /= Q1;
              post *= P1;
              double *t = in;
              in = out;
              out = t;
            }
            for (int i = threadIdx.x; i < n_nodes; i += blockDim.x) acc[i] = in[i];
            __syncthreads();
          } else {
            for (int p = 0; p < dim; p++) {
              const double *src = &u[ebase + p * qstride + (size_t)comp * n_qpts];
              for (int i = threadIdx.x; i < n_qpts; i += blockDim.x) buf0[i] = src[i];
              __syncthreads();
              int     pre = ipow_d(Q1, dim - 1), post = 1;
              double *in = buf0, *out = buf1;
              for (int d = 0; d < dim; d++) {
                if (collocated && d != p) {
                  pre /= Q1;
                  post *= P1;
                  continue;
                }
                lds_contract(pre, Q1, post, P1, d == p ? s_grad : s_interp, true, false, in, out);
                pre /= Q1;
                post *= P1;
                double *t = in;
                in = out;
                out = (t == buf0) ? buf2 : t;
              }
              for (int i = threadIdx.x; i < n_nodes; i += blockDim.x) acc[i] += in[i];
              __syncthreads();
            }
          }
          const size_t base = ((size_t)e * num_comp + comp) * n_nodes;
          for (int i = threadIdx.x; i < n_nodes; i += blockDim.x) v[base + i] = (add ? v[base + i] : 0.0) + acc[i];
          __syncthreads();
        }
      }
      (void)n_dir_in;
    }
  }
}

__global__ void __launch_bounds__(256) basis_weight_kernel(int num_elem, int dim, int Q1, const double *__restrict__ q_weight,
                                                           double *__restrict__ v) {
  const int     nq = ipow_d(Q1, dim);
  const int64_t n  = (int64_t)num_elem * nq;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    int    q = (int)(idx % nq);
    // same product order as the reference: w[x], then * w[y], then * w[z]
    double w = q_weight[q % Q1];
    q /= Q1;
    for (int d = 1; d < dim; d++) {
      w = q_weight[q % Q1] * w;
      q /= Q1;
    }
    v[idx] = w;
  }
}

// Host restatement of the collocated derivative D = G pinv(B) (Householder QR),
// same steps as interface/ceed-basis.c:750-771 and :1188-1300.
void householder_reflect(double *A, const double *v, double b, int m, int n, int row, int col) {
  for (int j = 0; j < n; j++) {
    double w = A[j * col];
    for (int i = 1; i < m; i++) w += v[i] * A[i * row + j * col];
    A[j * col] -= b * w;
    for (int i = 1; i < m; i++) A[i * row + j * col] -= b * w * v[i];
  }
}

void collocated_grad_host(int P, int Q, const double *interp, const double *grad, double *collo) {
  const int           m = Q, n = P;
  std::vector<double> A(interp, interp + m * n), tau(m, 0.0), v(m, 0.0), I(m * m, 0.0), pinv(m * n, 0.0);
  for (int i = 0; i < n; i++) {
    if (i >= m - 1) {
      tau[i] = 0.;
      break;
    }
    double sigma = 0.0;
    v[i] = A[i + n * i];
    for (int j = i + 1; j < m; j++) {
      v[j] = A[i + n * j];
      sigma += v[j] * v[j];
    }
    const double nrm = sqrt(v[i] * v[i] + sigma), Rii = -copysign(nrm, v[i]);
    v[i] -= Rii;
    tau[i] = 2 * v[i] * v[i] / (v[i] * v[i] + sigma);
    for (int j = i + 1; j < m; j++) v[j] /= v[i];
    householder_reflect(&A[i * n + i + 1], &v[i], tau[i], m - i, n - i - 1, n, 1);
    A[i + n * i] = Rii;
    for (int j = i + 1; j < m; j++) A[i + n * j] = v[j];
  }
  for (int i = 0; i < m; i++) I[i * m + i] = 1.0;
  for (int i = 0; i < n; i++) {
    for (int j = i + 1; j < m; j++) v[j] = A[j * n + i];
    householder_reflect(&I[i * m], &v[i], tau[i], m - i, m, m, 1);
  }
  for (int j = 0; j < m; j++) {
    pinv[j + m * (n - 1)] = I[j + m * (n - 1)] / A[n * n - 1];
    for (int i = n - 2; i >= 0; i--) {
      pinv[j + m * i] = I[j + m * i];
      for (int k = i + 1; k < n; k++) pinv[j + m * i] -= A[k + n * i] * pinv[j + m * k];
      pinv[j + m * i] /= A[i + n * i];
    }
  }
  for (int i = 0; i < Q; i++)
    for (int j = 0; j < Q; j++) {
      double s = 0;
      for (int k = 0; k < P; k++) s += grad[k + i * P] * pinv[j + k * Q];
      collo[j + i * Q] = s;
    }
}

}  // namespace

namespace mi355 {
void collocated_grad(int P, int Q, const double *interp, const double *grad, double *collo) {
  collocated_grad_host(P, Q, interp, grad, collo);
}
}  // namespace mi355

extern "C" int CeedMi355BasisCreateTensorH1(int32_t dim, int32_t num_comp, int32_t P_1d, int32_t Q_1d, const double *interp_1d_host,
                                            const double *grad_1d_host, const double *q_weight_1d_host, CeedMi355Basis *basis) {
  *basis = nullptr;
  MI355_CHECK(dim >= 1 && dim <= 3, CEED_MI355_ERROR_UNSUPPORTED, "tensor basis dim %d not supported (1..3)", dim);
  MI355_CHECK(P_1d >= 1 && Q_1d >= 1 && P_1d <= kMaxT && Q_1d <= kMaxT, CEED_MI355_ERROR_UNSUPPORTED,
              "P_1d=%d Q_1d=%d outside 1..%d", P_1d, Q_1d, kMaxT);
  MI355_CHECK(num_comp >= 1, CEED_MI355_ERROR_DIMENSION, "num_comp must be positive");
  CeedMi355Basis b = (CeedMi355Basis)calloc(1, sizeof(*b));
  b->dim = dim;
  b->num_comp = num_comp;
  b->P = P_1d;
  b->Q = Q_1d;
  bool coll = P_1d == Q_1d;
  for (int i = 0; i < P_1d && coll; i++)
    for (int j = 0; j < Q_1d; j++) {
      const double x = interp_1d_host[j + P_1d * i];
      if ((i == j && fabs(x - 1.0) >= 1e-15) || (i != j && fabs(x) >= 1e-15)) coll = false;
    }
  b->collocated = coll;
  b->collo_grad = Q_1d >= P_1d && !coll;
  const int           QP = Q_1d * P_1d, QQ = Q_1d * Q_1d;
  std::vector<double> mats(2 * QP + QQ + Q_1d, 0.0);
  memcpy(&mats[0], interp_1d_host, sizeof(double) * QP);
  memcpy(&mats[QP], grad_1d_host, sizeof(double) * QP);
  if (b->collo_grad) collocated_grad_host(P_1d, Q_1d, interp_1d_host, grad_1d_host, &mats[2 * QP]);
  if (q_weight_1d_host) memcpy(&mats[2 * QP + QQ], q_weight_1d_host, sizeof(double) * Q_1d);
  MI355_HIP(hipMalloc(&b->d_mats, sizeof(double) * mats.size()));
  MI355_HIP(hipMemcpy(b->d_mats, mats.data(), sizeof(double) * mats.size(), hipMemcpyHostToDevice));
  *basis = b;
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355BasisApply(CeedMi355Basis b, int32_t num_elem, int t_mode, int eval_mode, int add, const double *u, double *v,
                                   void *stream) {
  MI355_CHECK(b, CEED_MI355_ERROR_MAJOR, "null basis");
  hipStream_t s = (hipStream_t)stream;
  if (num_elem == 0) return CEED_MI355_SUCCESS;
  if (eval_mode == 16) {
    MI355_CHECK(t_mode == 0, CEED_MI355_ERROR_BACKEND, "CEED_EVAL_WEIGHT incompatible with CEED_TRANSPOSE");
    int nq = 1;
    for (int d = 0; d < b->dim; d++) nq *= b->Q;
    basis_weight_kernel<<<mi355_blocks((int64_t)num_elem * nq, 256), 256, 0, s>>>(num_elem, b->dim, b->Q,
                                                                                  b->d_mats + 2 * b->Q * b->P + b->Q * b->Q, v);
    MI355_LAUNCH_CHECK();
    return CEED_MI355_SUCCESS;
  }
  MI355_CHECK(eval_mode == 1 || eval_mode == 2, CEED_MI355_ERROR_BACKEND, "eval mode %d not supported by tensor basis", eval_mode);
  MI355_CHECK(eval_mode != 2 || b->collo_grad || b->collocated || b->P > b->Q, CEED_MI355_ERROR_BACKEND, "unsupported gradient path");
  int T = b->P > b->Q ? b->P : b->Q, Tn = 1;
  for (int d = 0; d < b->dim; d++) Tn *= T;
  const size_t lds = sizeof(double) * (2 * b->Q * b->P + b->Q * b->Q + 4 * (size_t)Tn);
  const int    grid = num_elem < 65536 ? num_elem : 65536;
  basis_tensor_kernel<<<grid, 256, lds, s>>>(num_elem, b->dim, b->num_comp, b->P, b->Q, b->collocated, b->collo_grad, t_mode, eval_mode, add,
                                             b->d_mats, u, v);
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355BasisDestroy(CeedMi355Basis *basis) {
  if (!basis || !*basis) return CEED_MI355_SUCCESS;
  (void)hipFree((*basis)->d_mats);
  free(*basis);
  *basis = nullptr;
  return CEED_MI355_SUCCESS;
}
