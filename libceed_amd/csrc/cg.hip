// Conjugate gradients on a fused hex operator plan, resident on the GPU.
//
// libCEED has no solver; its benchmark problems run PETSc's KSPCG around
// CeedOperatorApply (examples/petsc/bps.c:210-239, natural norm, rtol 1e-10).
// This is the same unpreconditioned CG (Hestenes-Stiefel recurrences), with
// every scalar kept in device memory so an iteration never waits for the
// host: alpha and beta are formed by the reduction kernels and read by the
// update kernels.  A segment of iterations can be captured once in a hipGraph
// and replayed; the host only reads the residual between segments.
//
// Mixed precision: the inner iteration runs on an FP32 plan with FP32
// vectors (dot products accumulated in FP64); every k iterations the FP32
// solution increment is added into the FP64 solution and the residual is
// replaced by the true FP64 residual b - A64 x (a "reliable update":
// Sleijpen and van der Vorst, Computing 56 (1996) 141-163), so the search
// directions carry on and the attainable accuracy is FP64's.
//
// Jacobi preconditioning (optional, CeedMi355CGSetJacobi): z = D^-1 r with D
// the operator's assembled diagonal (CeedMi355OperatorAssembleDiagonal, the
// input the reference's PETSc examples build their Jacobi preconditioner
// from: examples/petsc/src/matops.c:54); then rho = r . z drives alpha and
// beta while r . r drives the stopping test.
//
// Per iteration (n = L-vector length): the operator apply, then
//   k_dot:       p . Ap                                        (reads 2n)
//   k_update_r:  r -= alpha Ap [z = D^-1 r], partial r.r [r.z]  (reads 2n [3n], writes n [2n])
//   k_update_px: x += alpha p, p = z + beta p  (z = r unpreconditioned)  (reads 3n, writes 2n)
// plus single-block reductions that finish the dot products and form
// alpha and beta.
#include <math.h>
#include <string.h>

#include "hex_plan.hpp"
#include "mi355_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kBlocks  = 1024;  // fixed reduction grid: bitwise-deterministic sums

// device scalar slots: kRho = r . z (= r . r unpreconditioned) drives alpha and beta, kRR = r . r the stopping test
enum { kRR = 0, kPAP, kAlpha, kBeta, kRRTrue, kBB, kRho, kNumScal = 8 };

__device__ double block_sum(double v) {
  __shared__ double s[kThreads];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  return s[0];
}

template <typename A, typename B>
__global__ void __launch_bounds__(kThreads) k_dot(int64_t n, const A *__restrict__ a, const B *__restrict__ b, double *__restrict__ partial) {
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) acc += (double)a[i] * (double)b[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Finish a dot product into scal[slot]; mode 1: alpha = Rho / PAP; mode 2: beta = new / Rho, Rho = new
// (and scal[slot] = new, slot kRR when unpreconditioned, kRho otherwise).
__global__ void __launch_bounds__(kThreads) k_finish(int nb, const double *__restrict__ partial, double *__restrict__ scal, int slot, int mode) {
  double acc = 0;
  for (int i = threadIdx.x; i < nb; i += kThreads) acc += partial[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) {
    if (mode == 1) {
      scal[kPAP]   = s;
      scal[kAlpha] = s != 0.0 ? scal[kRho] / s : 0.0;
    } else if (mode == 2) {
      scal[kBeta] = scal[kRho] != 0.0 ? s / scal[kRho] : 0.0;
      scal[kRho]  = s;
      scal[slot]  = s;
    } else {
      scal[slot] = s;
    }
  }
}

// r -= alpha Ap, and the partial sums of r . r
template <typename T>
__global__ void __launch_bounds__(kThreads) k_update_r(int64_t n, const double *__restrict__ scal, T *__restrict__ r, const T *__restrict__ Ap,
                                                       double *__restrict__ partial) {
  const T alpha = (T)scal[kAlpha];
  double  acc   = 0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const T ri = r[i] - alpha * Ap[i];
    r[i]       = ri;
    acc += (double)ri * (double)ri;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// r -= alpha Ap, z = invd r, and the partial sums of r . r and r . z
template <typename T>
__global__ void __launch_bounds__(kThreads) k_update_r_jacobi(int64_t n, const double *__restrict__ scal, T *__restrict__ r, const T *__restrict__ Ap,
                                                              const T *__restrict__ invd, T *__restrict__ z, double *__restrict__ partial_rr,
                                                              double *__restrict__ partial_rz) {
  const T alpha = (T)scal[kAlpha];
  double  rr = 0, rz = 0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const T ri = r[i] - alpha * Ap[i];
    const T zi = invd[i] * ri;
    r[i]       = ri;
    z[i]       = zi;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)zi;
  }
  const double s = block_sum(rr);
  if (threadIdx.x == 0) partial_rr[blockIdx.x] = s;
  __syncthreads();  // block_sum's shared buffer is reused
  const double t = block_sum(rz);
  if (threadIdx.x == 0) partial_rz[blockIdx.x] = t;
}

// z = invd r
template <typename T>
__global__ void __launch_bounds__(kThreads) k_scale(int64_t n, const T *__restrict__ invd, const T *__restrict__ r, T *__restrict__ z) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) z[i] = invd[i] * r[i];
}

// invd = 1 / diag (in the inner scalar type)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_invert(int64_t n, const double *__restrict__ diag, T *__restrict__ invd) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) invd[i] = (T)(1.0 / diag[i]);
}

// x += alpha p;  p = z + beta p  (z = r unpreconditioned)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_update_px(int64_t n, const double *__restrict__ scal, T *__restrict__ x, T *__restrict__ p,
                                                        const T *__restrict__ r) {
  const T alpha = (T)scal[kAlpha], beta = (T)scal[kBeta];
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const T pi = p[i];
    x[i] += alpha * pi;
    p[i] = r[i] + beta * pi;
  }
}

// r = b - Ap (FP64), partial r . r
__global__ void __launch_bounds__(kThreads) k_residual(int64_t n, const double *__restrict__ b, const double *__restrict__ Ap, double *__restrict__ r,
                                                       double *__restrict__ partial) {
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double ri = b[i] - Ap[i];
    r[i]            = ri;
    acc += ri * ri;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// start of a search: p = r (and the FP32 copies of a replaced residual)
template <typename S, typename D>
__global__ void __launch_bounds__(kThreads) k_copy2(int64_t n, const S *__restrict__ src, D *__restrict__ a, D *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const D v = (D)src[i];
    a[i]      = v;
    if (b) b[i] = v;
  }
}

// x64 += x32; x32 = 0
__global__ void __launch_bounds__(kThreads) k_accumulate(int64_t n, double *__restrict__ x, float *__restrict__ dx) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    x[i] += (double)dx[i];
    dx[i] = 0.0f;
  }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_fill(int64_t n, T v, T *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) x[i] = v;
}

}  // namespace

struct CeedMi355CG_private {
  CeedMi355Operator op = nullptr, op64 = nullptr;  // inner plan; FP64 plan for mixed precision
  bool              mixed = false;
  int64_t           n     = 0;
  hipStream_t       stream = nullptr;
  hipEvent_t        ev     = nullptr;
  double           *scal = nullptr, *h_scal = nullptr, *partial = nullptr, *partial2 = nullptr;
  void             *r = nullptr, *p = nullptr, *Ap = nullptr, *dx = nullptr;  // inner scalar type
  bool              jacobi = false;
  void             *invd = nullptr, *z = nullptr;  // Jacobi: 1 / diag and z = invd r, inner scalar type
  double           *r64 = nullptr, *Ap64 = nullptr;                            // mixed only
  hipGraphExec_t    graph = nullptr;
  int               graph_len = 0;
  const void       *graph_b = nullptr;
  void             *graph_x = nullptr;
};

namespace {

int nblocks(int64_t n) { return mi355_blocks(n, kThreads, kBlocks); }

// one inner iteration on cg->op with vectors of type T; x is the inner solution accumulator
template <typename T>
int iteration(CeedMi355CG cg, T *x) {
  hipStream_t s  = cg->stream;
  const int   nb = nblocks(cg->n);
  T *r = (T *)cg->r, *p = (T *)cg->p, *Ap = (T *)cg->Ap, *z = cg->jacobi ? (T *)cg->z : r;
  int ierr = CeedMi355OperatorApply(cg->op, p, Ap, s);
  if (ierr) return ierr;
  k_dot<T, T><<<nb, kThreads, 0, s>>>(cg->n, p, Ap, cg->partial);
  k_finish<<<1, kThreads, 0, s>>>(nb, cg->partial, cg->scal, kPAP, 1);
  if (cg->jacobi) {
    k_update_r_jacobi<T><<<nb, kThreads, 0, s>>>(cg->n, cg->scal, r, Ap, (const T *)cg->invd, z, cg->partial, cg->partial2);
    k_finish<<<1, kThreads, 0, s>>>(nb, cg->partial2, cg->scal, kRho, 2);
    k_finish<<<1, kThreads, 0, s>>>(nb, cg->partial, cg->scal, kRR, 0);
  } else {
    k_update_r<T><<<nb, kThreads, 0, s>>>(cg->n, cg->scal, r, Ap, cg->partial);
    k_finish<<<1, kThreads, 0, s>>>(nb, cg->partial, cg->scal, kRR, 2);
  }
  k_update_px<T><<<nb, kThreads, 0, s>>>(cg->n, cg->scal, x, p, z);
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

// FP64 true residual r64 = b - A64 x into scal[kRRTrue]
int true_residual(CeedMi355CG cg, const double *b, const double *x, double *r64, double *Ap64) {
  hipStream_t s   = cg->stream;
  const int   nb  = nblocks(cg->n);
  int         err = CeedMi355OperatorApply(cg->mixed ? cg->op64 : cg->op, x, Ap64, s);
  if (err) return err;
  k_residual<<<nb, kThreads, 0, s>>>(cg->n, b, Ap64, r64, cg->partial);
  k_finish<<<1, kThreads, 0, s>>>(nb, cg->partial, cg->scal, kRRTrue, 0);
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

// With r in place: [z = invd r], optionally p = z, then r . r into kRR and r . z into kRho.
template <typename T>
int start_search(CeedMi355CG cg, bool set_p) {
  hipStream_t s  = cg->stream;
  const int   nb = nblocks(cg->n);
  const T    *r  = (const T *)cg->r;
  const T    *z  = r;
  if (cg->jacobi) {
    k_scale<T><<<nb, kThreads, 0, s>>>(cg->n, (const T *)cg->invd, r, (T *)cg->z);
    z = (const T *)cg->z;
  }
  if (set_p) k_copy2<T, T><<<nb, kThreads, 0, s>>>(cg->n, z, (T *)cg->p, nullptr);
  k_dot<T, T><<<nb, kThreads, 0, s>>>(cg->n, r, r, cg->partial);
  k_finish<<<1, kThreads, 0, s>>>(nb, cg->partial, cg->scal, kRR, 0);
  if (cg->jacobi) k_dot<T, T><<<nb, kThreads, 0, s>>>(cg->n, r, z, cg->partial);
  k_finish<<<1, kThreads, 0, s>>>(nb, cg->partial, cg->scal, kRho, 0);
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

// mixed: fold the FP32 increment into x, replace the residual by the FP64 one, restart nothing
int reliable_update(CeedMi355CG cg, const double *b, double *x) {
  hipStream_t s  = cg->stream;
  const int   nb = nblocks(cg->n);
  k_accumulate<<<nb, kThreads, 0, s>>>(cg->n, x, (float *)cg->dx);
  MI355_LAUNCH_CHECK();
  int err = true_residual(cg, b, x, cg->r64, cg->Ap64);
  if (err) return err;
  k_copy2<double, float><<<nb, kThreads, 0, s>>>(cg->n, cg->r64, (float *)cg->r, nullptr);
  return start_search<float>(cg, false);
}

// `len` iterations (plus, for mixed precision, the reliable update that ends them)
int segment(CeedMi355CG cg, const double *b, double *x, int len, bool update) {
  for (int i = 0; i < len; i++) {
    const int err = cg->mixed ? iteration<float>(cg, (float *)cg->dx) : iteration<double>(cg, x);
    if (err) return err;
  }
  if (cg->mixed && update) return reliable_update(cg, b, x);
  return CEED_MI355_SUCCESS;
}

int run_segment(CeedMi355CG cg, const double *b, double *x, int len, bool update, bool use_graph) {
  if (!use_graph) return segment(cg, b, x, len, update);
  if (!cg->graph || cg->graph_len != len || cg->graph_b != b || cg->graph_x != x) {
    if (cg->graph) (void)hipGraphExecDestroy(cg->graph);
    cg->graph = nullptr;
    hipGraph_t g;
    MI355_HIP(hipStreamBeginCapture(cg->stream, hipStreamCaptureModeThreadLocal));
    const int err = segment(cg, b, x, len, update);
    hipError_t e  = hipStreamEndCapture(cg->stream, &g);
    if (err) return err;
    MI355_HIP(e);
    e = hipGraphInstantiate(&cg->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MI355_HIP(e);
    cg->graph_len = len;
    cg->graph_b   = b;
    cg->graph_x   = x;
  }
  MI355_HIP(hipGraphLaunch(cg->graph, cg->stream));
  return CEED_MI355_SUCCESS;
}

int read_scalars(CeedMi355CG cg) {
  MI355_HIP(hipMemcpyAsync(cg->h_scal, cg->scal, sizeof(double) * kNumScal, hipMemcpyDeviceToHost, cg->stream));
  MI355_HIP(hipStreamSynchronize(cg->stream));
  return CEED_MI355_SUCCESS;
}

}  // namespace

extern "C" int CeedMi355CGCreate(CeedMi355Operator op, CeedMi355Operator op64, CeedMi355CG *cg_out) {
  *cg_out = nullptr;
  MI355_CHECK(op, CEED_MI355_ERROR_MAJOR, "null operator");
  const bool mixed = op64 != nullptr;
  MI355_CHECK(!mixed || (op->scalar_type == CEED_MI355_FP32 && op64->scalar_type == CEED_MI355_FP64), CEED_MI355_ERROR_BACKEND,
              "mixed-precision CG needs an FP32 inner plan and an FP64 outer plan");
  MI355_CHECK(mixed || op->scalar_type == CEED_MI355_FP64, CEED_MI355_ERROR_BACKEND, "plain CG needs an FP64 plan");
  MI355_CHECK(!mixed || op64->l_size == op->l_size, CEED_MI355_ERROR_DIMENSION, "FP32 and FP64 plans have different L-vector sizes");
  CeedMi355CG cg = new CeedMi355CG_private();
  cg->op         = op;
  cg->op64       = op64;
  cg->mixed      = mixed;
  cg->n          = op->l_size;
  const size_t s = mixed ? sizeof(float) : sizeof(double), n = (size_t)cg->n + 1;
  bool         ok = hipStreamCreateWithFlags(&cg->stream, hipStreamNonBlocking) == hipSuccess && hipEventCreateWithFlags(&cg->ev, hipEventDisableTiming) == hipSuccess &&
            hipMalloc(&cg->scal, sizeof(double) * kNumScal) == hipSuccess &&
            hipHostMalloc(&cg->h_scal, sizeof(double) * kNumScal, hipHostMallocDefault) == hipSuccess &&
            hipMalloc(&cg->partial, sizeof(double) * kBlocks) == hipSuccess && hipMalloc(&cg->partial2, sizeof(double) * kBlocks) == hipSuccess &&
            hipMalloc(&cg->r, s * n) == hipSuccess &&
            hipMalloc(&cg->p, s * n) == hipSuccess && hipMalloc(&cg->Ap, s * n) == hipSuccess;
  if (ok && mixed)
    ok = hipMalloc(&cg->dx, sizeof(float) * n) == hipSuccess && hipMalloc(&cg->r64, sizeof(double) * n) == hipSuccess &&
         hipMalloc(&cg->Ap64, sizeof(double) * n) == hipSuccess;
  if (!ok) {
    CeedMi355CGDestroy(&cg);
    MI355_CHECK(false, CEED_MI355_ERROR_BACKEND, "CG work-space allocation failed (%lld entries)", (long long)n);
  }
  *cg_out = cg;
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355CGSetJacobi(CeedMi355CG cg, const double *diag, void *stream) {
  MI355_CHECK(cg, CEED_MI355_ERROR_MAJOR, "null CG");
  if (cg->graph) (void)hipGraphExecDestroy(cg->graph);  // captured with the other recurrences
  cg->graph = nullptr;
  if (!diag) {
    cg->jacobi = false;
    return CEED_MI355_SUCCESS;
  }
  const size_t s = cg->mixed ? sizeof(float) : sizeof(double), n = (size_t)cg->n + 1;
  if (!cg->invd) {
    MI355_HIP(hipMalloc(&cg->invd, s * n));
    MI355_HIP(hipMalloc(&cg->z, s * n));
  }
  MI355_HIP(hipEventRecord(cg->ev, (hipStream_t)stream));
  MI355_HIP(hipStreamWaitEvent(cg->stream, cg->ev, 0));
  const int nb = nblocks(cg->n);
  if (cg->mixed) k_invert<float><<<nb, kThreads, 0, cg->stream>>>(cg->n, diag, (float *)cg->invd);
  else k_invert<double><<<nb, kThreads, 0, cg->stream>>>(cg->n, diag, (double *)cg->invd);
  MI355_LAUNCH_CHECK();
  MI355_HIP(hipEventRecord(cg->ev, cg->stream));
  MI355_HIP(hipStreamWaitEvent((hipStream_t)stream, cg->ev, 0));
  cg->jacobi = true;
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355CGSolve(CeedMi355CG cg, const double *b, double *x, int32_t max_iter, double rtol, int32_t check_every, int use_graph,
                                void *stream, CeedMi355CGStats *stats) {
  MI355_CHECK(cg && stats, CEED_MI355_ERROR_MAJOR, "null CG or stats");
  MI355_CHECK(check_every >= 1, CEED_MI355_ERROR_BACKEND, "check_every must be >= 1");
  memset(stats, 0, sizeof(*stats));
  hipStream_t s  = cg->stream;
  const int   nb = nblocks(cg->n);
  // order after the caller's work on its stream
  MI355_HIP(hipEventRecord(cg->ev, (hipStream_t)stream));
  MI355_HIP(hipStreamWaitEvent(s, cg->ev, 0));

  k_fill<double><<<1, kThreads, 0, s>>>(kNumScal, 0.0, cg->scal);
  k_dot<double, double><<<nb, kThreads, 0, s>>>(cg->n, b, b, cg->partial);
  k_finish<<<1, kThreads, 0, s>>>(nb, cg->partial, cg->scal, kBB, 0);
  MI355_LAUNCH_CHECK();
  int err;
  if (cg->mixed) {  // r64 = b - A x, r = (float) r64, p = [D^-1] r, dx = 0
    if ((err = true_residual(cg, b, x, cg->r64, cg->Ap64))) return err;
    k_copy2<double, float><<<nb, kThreads, 0, s>>>(cg->n, cg->r64, (float *)cg->r, nullptr);
    k_fill<float><<<nb, kThreads, 0, s>>>(cg->n, 0.0f, (float *)cg->dx);
    err = start_search<float>(cg, true);
  } else {  // r = b - A x, p = [D^-1] r
    if ((err = true_residual(cg, b, x, (double *)cg->r, (double *)cg->Ap))) return err;
    err = start_search<double>(cg, true);
  }
  if (err) return err;
  MI355_LAUNCH_CHECK();
  if ((err = read_scalars(cg))) return err;
  const double bb   = cg->h_scal[kBB];
  const double stop = rtol * rtol * bb;
  double       rr   = cg->h_scal[kRR];
  int          it   = 0;
  while (it < max_iter && rr > stop && bb > 0) {
    const int len = check_every < max_iter - it ? check_every : max_iter - it;
    if ((err = run_segment(cg, b, x, len, true, use_graph && len == check_every))) return err;
    it += len;
    if ((err = read_scalars(cg))) return err;
    rr = cg->mixed ? cg->h_scal[kRRTrue] : cg->h_scal[kRR];
    if (cg->mixed) stats->replacements++;
  }
  // report the true FP64 residual of the returned x
  if (cg->mixed) {
    if ((err = true_residual(cg, b, x, cg->r64, cg->Ap64))) return err;
  } else {
    if ((err = true_residual(cg, b, x, (double *)cg->p, (double *)cg->Ap))) return err;
  }
  if ((err = read_scalars(cg))) return err;
  stats->iterations   = it;
  stats->rel_residual = bb > 0 ? sqrt(cg->h_scal[kRRTrue] / bb) : 0.0;
  stats->rel_residual_recursive = bb > 0 ? sqrt(rr / bb) : 0.0;
  // the caller's stream waits for the solve
  MI355_HIP(hipEventRecord(cg->ev, s));
  MI355_HIP(hipStreamWaitEvent((hipStream_t)stream, cg->ev, 0));
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355CGDestroy(CeedMi355CG *cg) {
  if (!cg || !*cg) return CEED_MI355_SUCCESS;
  CeedMi355CG c = *cg;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  (void)hipFree(c->scal);
  (void)hipHostFree(c->h_scal);
  (void)hipFree(c->partial);
  (void)hipFree(c->partial2);
  (void)hipFree(c->invd);
  (void)hipFree(c->z);
  (void)hipFree(c->r);
  (void)hipFree(c->p);
  (void)hipFree(c->Ap);
  (void)hipFree(c->dx);
  (void)hipFree(c->r64);
  (void)hipFree(c->Ap64);
  if (c->ev) (void)hipEventDestroy(c->ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  *cg = nullptr;
  return CEED_MI355_SUCCESS;
}
