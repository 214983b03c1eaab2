// Diagonal of a fused hex plan's operator, assembled on the GPU: the input of
// Jacobi preconditioning (the reference's CeedOperatorLinearAssembleDiagonal,
// interface/ceed-preconditioning.c:2395, used by examples/petsc/src/matops.c:54).
//
// Element diagonal of node i = (a, b, c) (x fastest):
//   mass:       sum_q (B_a(qx) B_b(qy) B_c(qz))^2 qd(q)
//   diffusion:  sum_q g_i(q)^T S(q) g_i(q),  g_i = (G_a B_b B_c, B_a G_b B_c, B_a B_b G_c)
//               (times kappa(q) for the kappa-weighted operator),
// with G = D B, D the plan's collocated derivative; every component of a
// vector operator has the same diagonal.  The element diagonals are summed
// into the L-vector by the plan's E^T mode (transpose map, colours, atomics).
// One workgroup per element; the element's geometric factors sit in LDS.
#include "hex_plan.hpp"

namespace mi355 {
namespace {

template <int P, int Q, typename T>
__global__ void __launch_bounds__(128) hex_diag_kernel(int64_t elem_begin, int64_t elem_end, int scatter, int qf, int num_comp,
                                                       const int32_t *__restrict__ offsets, int32_t comp_stride, const T *__restrict__ q_data,
                                                       int q_comp, const T *__restrict__ mats, T *__restrict__ v, T *__restrict__ evec,
                                                       const int32_t *__restrict__ tpos) {
  constexpr int P3 = P * P * P, Q3 = Q * Q * Q;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *sB  = reinterpret_cast<T *>(lds_raw);  // Q x P
  T *sG  = sB + Q * P;                       // Q x P, G = D B
  T *sQd = sG + Q * P;                       // q_comp x Q^3
  const int64_t e = elem_begin + xcd_chunk(blockIdx.x, gridDim.x);
  if (e >= elem_end) return;  // uniform per workgroup
  const T *D = mats + Q * P;
  for (int i = threadIdx.x; i < Q * P; i += blockDim.x) {
    sB[i]            = mats[i];
    const int qx = i / P, a = i % P;
    T         g  = 0;
    for (int c = 0; c < Q; c++) g += D[qx * Q + c] * mats[c * P + a];
    sG[i] = g;
  }
  for (int i = threadIdx.x; i < q_comp * Q3; i += blockDim.x) sQd[i] = q_data[e * (int64_t)q_comp * Q3 + i];
  __syncthreads();
  for (int i = threadIdx.x; i < P3; i += blockDim.x) {
    const int a = i % P, b = (i / P) % P, c = i / (P * P);
    T         acc = 0;
    for (int qz = 0; qz < Q; qz++) {
      const T bz = sB[qz * P + c], gz = sG[qz * P + c];
      for (int qy = 0; qy < Q; qy++) {
        const T by = sB[qy * P + b], gy = sG[qy * P + b];
        for (int qx = 0; qx < Q; qx++) {
          const T   bx = sB[qx * P + a], gx = sG[qx * P + a];
          const int q  = (qz * Q + qy) * Q + qx;
          if (qf == kQfMass) {
            const T phi = bx * by * bz;
            acc += phi * phi * sQd[q];
          } else {
            const T d0 = gx * by * bz, d1 = bx * gy * bz, d2 = bx * by * gz;
            const T s00 = sQd[q], s11 = sQd[Q3 + q], s22 = sQd[2 * Q3 + q], s12 = sQd[3 * Q3 + q], s02 = sQd[4 * Q3 + q], s01 = sQd[5 * Q3 + q];
            T       t   = d0 * (s00 * d0 + s01 * d1 + s02 * d2) + d1 * (s01 * d0 + s11 * d1 + s12 * d2) + d2 * (s02 * d0 + s12 * d1 + s22 * d2);
            if (qf == kQfDiffusionKappa) t *= sQd[6 * Q3 + q];
            acc += t;
          }
        }
      }
    }
    const int32_t o = offsets[e * P3 + i];
    for (int k = 0; k < num_comp; k++) {
      if (scatter == kScatterElem) {  // element-major E-vector (brick plans)
        evec[(e * num_comp + k) * (int64_t)P3 + i] = acc;
        continue;
      }
      if (scatter == kScatterEvec || (scatter == kScatterSplit && (o & 1))) {  // L-entry-sorted E-vector
        evec[tpos[(e * num_comp + k) * (int64_t)P3 + i]] = acc;
        continue;
      }
      T *dst = &v[(o >> 1) + (int64_t)k * comp_stride];
      if (scatter == kScatterSplit) {  // at most two contributions: onto the zeroed target
        unsafeAtomicAdd(dst, acc);
        continue;
      }
      if (scatter == kScatterAtomic) unsafeAtomicAdd(dst, acc);
      else if (scatter == kScatterOverwrite && (o & 1)) *dst = acc;
      else *dst += acc;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) diag_zero(int64_t n, T *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = T(0);
}

template <int P, int Q, typename T>
int launch_diag(const CeedMi355Operator_private *op, void *v, int add, hipStream_t s) {
  const size_t lds  = sizeof(T) * (size_t)(2 * Q * P + op->q_comp * Q * Q * Q);
  const bool   tmap = op->scatter == 0;
  const bool   split = op->scatter == CEED_MI355_PLAN_SPLIT;
  const bool   brick = op->scatter == CEED_MI355_PLAN_BRICK;
  MI355_CHECK(op->scatter != CEED_MI355_PLAN_SEGMENT, CEED_MI355_ERROR_UNSUPPORTED,
              "diagonal assembly does not take segment plans (elements of a colour share segment faces)");
  if (brick) {  // element diagonals into the element-major E-vector, then the brick E^T
    if (!add && !op->covers_all) {
      diag_zero<T><<<mi355_blocks(op->l_size, 256, 4096), 256, 0, s>>>(op->l_size, (T *)v);
      MI355_LAUNCH_CHECK();
    }
    hex_diag_kernel<P, Q, T><<<(unsigned)op->num_elem, 128, lds, s>>>(0, op->num_elem, kScatterElem, op->qf, op->num_comp, op->d_offsets,
                                                                     op->comp_stride, (const T *)op->d_qdata, op->q_comp,
                                                                     (const T *)op->d_mats, (T *)v, (T *)op->d_evec, op->d_tpos);
    MI355_LAUNCH_CHECK();
    return brick_evec(op, op->d_evec, v, !add, s);
  }
  // the split mode's atomic adds land on a zeroed target: v, or with add the scratch vector
  T *target = split && add ? (T *)op->d_scratch : (T *)v;
  if ((!add && !tmap && !op->covers_all) || (split && add)) {
    diag_zero<T><<<mi355_blocks(op->l_size, 256, 4096), 256, 0, s>>>(op->l_size, target);
    MI355_LAUNCH_CHECK();
  }
  for (size_t c = 0; c + 1 < op->color_start.size(); c++) {
    const int64_t b = op->color_start[c], e = op->color_start[c + 1];
    if (e <= b) continue;
    const int scatter = tmap ? kScatterEvec
                        : split ? kScatterSplit
                                : (op->scatter == CEED_MI355_PLAN_ATOMIC ? kScatterAtomic : (add ? kScatterAdd : kScatterOverwrite));
    hex_diag_kernel<P, Q, T><<<(unsigned)(e - b), 128, lds, s>>>(b, e, scatter, op->qf, op->num_comp, op->d_offsets, op->comp_stride,
                                                               (const T *)op->d_qdata, op->q_comp, (const T *)op->d_mats, target,
                                                               (T *)op->d_evec, op->d_tpos);
  }
  MI355_LAUNCH_CHECK();
  if (tmap) return transpose_gather(op, v, !add, s);
  if (split) return split_finish(op, v, !add, s);
  return CEED_MI355_SUCCESS;
}

using DiagFn = int (*)(const CeedMi355Operator_private *, void *, int, hipStream_t);

template <typename T>
DiagFn pick_diag(int P, int Q) {
#define MI355_DIAG_CASE(PP, QQ) \
  if (P == PP && Q == QQ) return launch_diag<PP, QQ, T>;
  MI355_DIAG_CASE(2, 2) MI355_DIAG_CASE(2, 3) MI355_DIAG_CASE(3, 3) MI355_DIAG_CASE(3, 4) MI355_DIAG_CASE(4, 4) MI355_DIAG_CASE(4, 5)
  MI355_DIAG_CASE(5, 5) MI355_DIAG_CASE(5, 6) MI355_DIAG_CASE(6, 6) MI355_DIAG_CASE(6, 7) MI355_DIAG_CASE(7, 7) MI355_DIAG_CASE(7, 8)
  MI355_DIAG_CASE(8, 8) MI355_DIAG_CASE(8, 9) MI355_DIAG_CASE(9, 9) MI355_DIAG_CASE(9, 10)
#undef MI355_DIAG_CASE
  return nullptr;
}

}  // namespace
}  // namespace mi355

extern "C" int CeedMi355OperatorAssembleDiagonal(CeedMi355Operator op, void *diag, int add, void *stream) {
  MI355_CHECK(op, CEED_MI355_ERROR_MAJOR, "null operator");
  MI355_CHECK(op->qf == mi355::kQfMass || op->qf == mi355::kQfDiffusion || op->qf == mi355::kQfDiffusionKappa, CEED_MI355_ERROR_UNSUPPORTED,
              "diagonal assembly needs a mass or diffusion plan");
  if (op->num_elem == 0) return CEED_MI355_SUCCESS;
  if (op->scatter == CEED_MI355_PLAN_BRICK && !op->d_evec)  // element-major E-vector of the brick plan's diagonal, made once
    MI355_HIP(hipMalloc(&op->d_evec, (size_t)op->scalar_bytes * ((size_t)op->num_elem * op->num_comp * op->P * op->P * op->P + 1)));
  auto fn = op->scalar_type == CEED_MI355_FP64 ? mi355::pick_diag<double>(op->P, op->Q) : mi355::pick_diag<float>(op->P, op->Q);
  MI355_CHECK(fn, CEED_MI355_ERROR_UNSUPPORTED, "no diagonal kernel for P=%d Q=%d", op->P, op->Q);
  return fn(op, diag, add, (hipStream_t)stream);
}
