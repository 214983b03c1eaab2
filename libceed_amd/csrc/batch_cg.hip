// Batched CG over many independent systems A_s x_s = b_s (block-diagonal).
//
// The systems' L-vectors are concatenated into one vector; system s owns the
// entries [sys_begin[s], sys_begin[s+1]).  The operator is a list of fused
// hex plans, each covering a contiguous range of the vector (in practice one
// plan per (pointwise operator, P, Q) group holding the elements of every
// system of that group, with offsets relative to the range start), so one CG
// iteration is one fused launch per group plus five vector kernels over the
// whole batch, whatever the number of systems.  Per-system scalars (r.r,
// p.Ap, alpha, beta) live in device memory and come from segmented,
// fixed-order reductions: the vector is cut into chunks that never straddle
// two systems, a block reduces a chunk, and one thread per system adds its
// chunks in order.  A system whose residual has reached its tolerance gets
// alpha = beta = 0 and stays frozen while the others go on.  A segment of
// iterations is captured once into a hipGraph and replayed.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <utility>
#include <vector>

#include "hex_plan.hpp"
#include "mi355_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kChunk   = 8192;  // entries per reduction chunk

enum { kRR = 0, kPAP, kAlpha, kBeta, kBB, kTol2, kRRTrue, kNumSlots };

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

// one block per chunk: partial[c] = sum over the chunk of a*b (chunks of frozen systems: skipped)
__global__ void __launch_bounds__(kThreads) k_chunk_dot(const int64_t *__restrict__ chunk_begin, const int32_t *__restrict__ chunk_sys,
                                                        const int32_t *__restrict__ active, const double *__restrict__ a,
                                                        const double *__restrict__ b, double *__restrict__ partial) {
  const int64_t c = blockIdx.x, lo = chunk_begin[c], hi = chunk_begin[c + 1];
  if (active && !active[chunk_sys[c]]) return;  // uniform per block; the finish kernel ignores frozen systems
  double        acc = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) acc += a[i] * b[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[c] = s;
}

// r -= alpha_s Ap, and the chunk partials of r.r
__global__ void __launch_bounds__(kThreads) k_chunk_update_r(const int64_t *__restrict__ chunk_begin, const int32_t *__restrict__ chunk_sys,
                                                             const double *__restrict__ scal, double *__restrict__ r, const double *__restrict__ Ap,
                                                             const int32_t *__restrict__ active, double *__restrict__ partial) {
  const int64_t c = blockIdx.x, lo = chunk_begin[c], hi = chunk_begin[c + 1];
  if (!active[chunk_sys[c]]) return;
  const double alpha = scal[(int64_t)chunk_sys[c] * kNumSlots + kAlpha];
  double        acc   = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) {
    const double ri = r[i] - alpha * Ap[i];
    r[i]            = ri;
    acc += ri * ri;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[c] = s;
}

// x += alpha_s p; p = r + beta_s p
__global__ void __launch_bounds__(kThreads) k_chunk_update_px(const int64_t *__restrict__ chunk_begin, const int32_t *__restrict__ chunk_sys,
                                                              const double *__restrict__ scal, double *__restrict__ x, double *__restrict__ p,
                                                              const double *__restrict__ r, const int32_t *__restrict__ was_active) {
  const int64_t c = blockIdx.x, lo = chunk_begin[c], hi = chunk_begin[c + 1];
  if (!was_active[chunk_sys[c]]) return;
  const double *sc = scal + (int64_t)chunk_sys[c] * kNumSlots;
  const double  alpha = sc[kAlpha], beta = sc[kBeta];
  for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) {
    const double pi = p[i];
    x[i] += alpha * pi;
    p[i] = r[i] + beta * pi;
  }
}

// r = b - Ap; p = r; chunk partials of r.r
__global__ void __launch_bounds__(kThreads) k_chunk_residual(const int64_t *__restrict__ chunk_begin, const double *__restrict__ b,
                                                             const double *__restrict__ Ap, double *__restrict__ r, double *__restrict__ p,
                                                             double *__restrict__ partial) {
  const int64_t c = blockIdx.x, lo = chunk_begin[c], hi = chunk_begin[c + 1];
  double        acc = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) {
    const double ri = b[i] - Ap[i];
    r[i]            = ri;
    if (p) p[i] = ri;
    acc += ri * ri;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[c] = s;
}

// per system: the ordered sum of its chunk partials, then the scalar update
//   mode 0: scal[slot] = sum;  mode 1: p.Ap -> alpha;  mode 2: r.r -> beta, RR, iteration count;
//   mode 3: b.b -> BB and the stopping threshold rtol^2 b.b
//   modes 0 and 2 also refresh active[s] = (r.r > threshold); prev[s] keeps the flag the
//   iteration started with (the x/p update of a system that just converged still runs once)
__global__ void __launch_bounds__(kThreads) k_sys_finish(int32_t num_sys, const int32_t *__restrict__ sys_chunk, const double *__restrict__ partial,
                                                         double *__restrict__ scal, int32_t *__restrict__ iters, int32_t *__restrict__ active_flag,
                                                         int32_t *__restrict__ prev_flag, int slot, int mode, double rtol2) {
  for (int32_t s = blockIdx.x * kThreads + threadIdx.x; s < num_sys; s += gridDim.x * kThreads) {
    double *sc = scal + (int64_t)s * kNumSlots;
    const bool active = sc[kRR] > sc[kTol2];
    double     acc    = 0;
    if (mode != 1 && mode != 2 || active)
      for (int32_t c = sys_chunk[s]; c < sys_chunk[s + 1]; c++) acc += partial[c];
    if (mode == 1) {
      sc[kPAP]   = acc;
      sc[kAlpha] = active && acc != 0.0 ? sc[kRR] / acc : 0.0;
    } else if (mode == 2) {
      prev_flag[s] = active;
      if (active) {
        sc[kBeta] = sc[kRR] != 0.0 ? acc / sc[kRR] : 0.0;
        sc[kRR]   = acc;
        iters[s]++;
      } else {
        sc[kBeta] = 0.0;
      }
      active_flag[s] = sc[kRR] > sc[kTol2];
    } else if (mode == 3) {
      sc[kBB]   = acc;
      sc[kTol2] = rtol2 * acc;
    } else {
      sc[slot] = acc;
      if (slot == kRR) active_flag[s] = prev_flag[s] = acc > sc[kTol2];
    }
  }
}

__global__ void __launch_bounds__(kThreads) k_zero(int64_t n, double *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) x[i] = 0.0;
}
__global__ void __launch_bounds__(kThreads) k_zero_i(int64_t n, int32_t *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) x[i] = 0;
}

}  // namespace

struct CeedMi355BatchCG_private {
  std::vector<CeedMi355Operator> ops;
  std::vector<int64_t>           op_begin;
  int32_t                        num_sys = 0, num_chunks = 0;
  int64_t                        n       = 0;
  int64_t                       *d_chunk_begin = nullptr;
  int32_t                       *d_chunk_sys = nullptr, *d_sys_chunk = nullptr, *d_iters = nullptr, *h_iters = nullptr;
  int32_t                       *d_active = nullptr, *d_prev = nullptr;            // per system
  std::vector<int32_t *>         d_elem_sys;                                       // per plan: system of each packed element
  std::vector<int64_t *>         d_op_chunk_begin;                                 // per plan: its L range in chunks of one system
  std::vector<int32_t *>         d_op_chunk_sys;                                   //   (relative begins + end) and their systems
  std::vector<int64_t>           op_num_chunks;
  double                        *d_scal = nullptr, *h_scal = nullptr, *d_partial = nullptr;
  double                        *r = nullptr, *p = nullptr, *Ap = nullptr;
  hipStream_t                    stream = nullptr;
  hipEvent_t                     ev     = nullptr;
  hipGraphExec_t                 graph  = nullptr;
  int                            graph_len = 0;
  const void                    *graph_b = nullptr;
  void                          *graph_x = nullptr;
};

namespace {

// A p for every plan; with skip, the elements of frozen systems are left out
// (their stale entries of Ap are never used: a frozen system's alpha is 0)
int apply_all(CeedMi355BatchCG cg, const double *u, double *v, bool skip = false) {
  for (size_t i = 0; i < cg->ops.size(); i++) {
    CeedMi355Operator op = cg->ops[i];
    op->d_elem_sys           = skip ? cg->d_elem_sys[i] : nullptr;
    op->d_sys_active         = skip ? cg->d_active : nullptr;
    op->d_gather_chunk_begin = skip ? cg->d_op_chunk_begin[i] : nullptr;
    op->d_gather_chunk_sys   = skip ? cg->d_op_chunk_sys[i] : nullptr;
    op->gather_num_chunks    = skip ? cg->op_num_chunks[i] : 0;
    const int err            = CeedMi355OperatorApply(op, u + cg->op_begin[i], v + cg->op_begin[i], cg->stream);
    op->d_elem_sys           = nullptr;
    op->d_sys_active         = nullptr;
    op->d_gather_chunk_begin = nullptr;
    op->d_gather_chunk_sys   = nullptr;
    op->gather_num_chunks    = 0;
    if (err) return err;
  }
  return CEED_MI355_SUCCESS;
}

int finish(CeedMi355BatchCG cg, int slot, int mode, double rtol2 = 0) {
  k_sys_finish<<<mi355_blocks(cg->num_sys, kThreads, 1024), kThreads, 0, cg->stream>>>(cg->num_sys, cg->d_sys_chunk, cg->d_partial, cg->d_scal,
                                                                                      cg->d_iters, cg->d_active, cg->d_prev, slot, mode, rtol2);
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

int iteration(CeedMi355BatchCG cg, double *x) {
  hipStream_t s  = cg->stream;
  const int   nc = cg->num_chunks;
  int         err;
  if ((err = apply_all(cg, cg->p, cg->Ap, true))) return err;
  k_chunk_dot<<<nc, kThreads, 0, s>>>(cg->d_chunk_begin, cg->d_chunk_sys, cg->d_active, cg->p, cg->Ap, cg->d_partial);
  if ((err = finish(cg, kPAP, 1))) return err;
  k_chunk_update_r<<<nc, kThreads, 0, s>>>(cg->d_chunk_begin, cg->d_chunk_sys, cg->d_scal, cg->r, cg->Ap, cg->d_active, cg->d_partial);
  if ((err = finish(cg, kRR, 2))) return err;
  k_chunk_update_px<<<nc, kThreads, 0, s>>>(cg->d_chunk_begin, cg->d_chunk_sys, cg->d_scal, x, cg->p, cg->r, cg->d_prev);
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

int segment(CeedMi355BatchCG cg, double *x, int len) {
  for (int i = 0; i < len; i++) {
    const int err = iteration(cg, x);
    if (err) return err;
  }
  return CEED_MI355_SUCCESS;
}

int run_segment(CeedMi355BatchCG cg, const double *b, double *x, int len, bool use_graph) {
  if (!use_graph) return segment(cg, x, len);
  if (!cg->graph || cg->graph_len != len || cg->graph_b != b || cg->graph_x != x) {
    if (cg->graph) (void)hipGraphExecDestroy(cg->graph);
    cg->graph = nullptr;
    hipGraph_t g;
    MI355_HIP(hipStreamBeginCapture(cg->stream, hipStreamCaptureModeThreadLocal));
    const int  err = segment(cg, x, len);
    hipError_t e   = hipStreamEndCapture(cg->stream, &g);
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

int read_back(CeedMi355BatchCG cg) {
  MI355_HIP(hipMemcpyAsync(cg->h_scal, cg->d_scal, sizeof(double) * kNumSlots * cg->num_sys, hipMemcpyDeviceToHost, cg->stream));
  MI355_HIP(hipMemcpyAsync(cg->h_iters, cg->d_iters, sizeof(int32_t) * cg->num_sys, hipMemcpyDeviceToHost, cg->stream));
  MI355_HIP(hipStreamSynchronize(cg->stream));
  return CEED_MI355_SUCCESS;
}

}  // namespace

extern "C" int CeedMi355BatchCGCreate(int32_t num_ops, const CeedMi355Operator *ops_host, const int64_t *op_begin_host, int32_t num_systems,
                                      const int64_t *sys_begin_host, CeedMi355BatchCG *cg_out) {
  *cg_out = nullptr;
  MI355_CHECK(num_ops > 0 && num_systems > 0, CEED_MI355_ERROR_DIMENSION, "batched CG needs operators and systems");
  const int64_t n = sys_begin_host[num_systems];
  MI355_CHECK(sys_begin_host[0] == 0 && n > 0, CEED_MI355_ERROR_DIMENSION, "system ranges must start at 0 and be non-empty");
  for (int32_t s = 0; s < num_systems; s++)
    MI355_CHECK(sys_begin_host[s + 1] > sys_begin_host[s], CEED_MI355_ERROR_DIMENSION, "system %d has an empty range", s);
  for (int32_t i = 0; i < num_ops; i++) {
    MI355_CHECK(ops_host[i] && ops_host[i]->scalar_type == CEED_MI355_FP64, CEED_MI355_ERROR_BACKEND, "operator %d is not an FP64 plan", i);
    MI355_CHECK(op_begin_host[i] >= 0 && op_begin_host[i] + ops_host[i]->l_size <= n, CEED_MI355_ERROR_DIMENSION,
                "operator %d range [%lld, %lld) outside the batch vector of %lld entries", i, (long long)op_begin_host[i],
                (long long)(op_begin_host[i] + ops_host[i]->l_size), (long long)n);
  }
  {  // the operator ranges tile [0, n)
    std::vector<std::pair<int64_t, int64_t>> rg;
    for (int32_t i = 0; i < num_ops; i++) rg.emplace_back(op_begin_host[i], op_begin_host[i] + ops_host[i]->l_size);
    std::sort(rg.begin(), rg.end());
    int64_t at = 0;
    for (auto &r : rg) {
      MI355_CHECK(r.first == at, CEED_MI355_ERROR_DIMENSION, "operator ranges must tile the batch vector (gap or overlap at %lld)", (long long)at);
      at = r.second;
    }
    MI355_CHECK(at == n, CEED_MI355_ERROR_DIMENSION, "operator ranges end at %lld, the batch vector has %lld entries", (long long)at, (long long)n);
  }
  // reduction chunks, none straddling two systems
  std::vector<int64_t> chunk_begin;
  std::vector<int32_t> chunk_sys, sys_chunk(num_systems + 1);
  for (int32_t s = 0; s < num_systems; s++) {
    sys_chunk[s] = (int32_t)chunk_sys.size();
    for (int64_t i = sys_begin_host[s]; i < sys_begin_host[s + 1]; i += kChunk) {
      chunk_begin.push_back(i);
      chunk_sys.push_back(s);
    }
  }
  sys_chunk[num_systems] = (int32_t)chunk_sys.size();
  chunk_begin.push_back(n);

  CeedMi355BatchCG cg = new CeedMi355BatchCG_private();
  cg->ops.assign(ops_host, ops_host + num_ops);
  cg->op_begin.assign(op_begin_host, op_begin_host + num_ops);
  cg->num_sys    = num_systems;
  cg->num_chunks = (int32_t)chunk_sys.size();
  cg->n          = n;
  bool ok = hipStreamCreateWithFlags(&cg->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&cg->ev, hipEventDisableTiming) == hipSuccess &&
            hipMalloc(&cg->d_chunk_begin, sizeof(int64_t) * chunk_begin.size()) == hipSuccess &&
            hipMalloc(&cg->d_chunk_sys, sizeof(int32_t) * chunk_sys.size()) == hipSuccess &&
            hipMalloc(&cg->d_sys_chunk, sizeof(int32_t) * sys_chunk.size()) == hipSuccess &&
            hipMalloc(&cg->d_iters, sizeof(int32_t) * num_systems) == hipSuccess &&
            hipHostMalloc(&cg->h_iters, sizeof(int32_t) * num_systems, hipHostMallocDefault) == hipSuccess &&
            hipMalloc(&cg->d_scal, sizeof(double) * kNumSlots * num_systems) == hipSuccess &&
            hipHostMalloc(&cg->h_scal, sizeof(double) * kNumSlots * num_systems, hipHostMallocDefault) == hipSuccess &&
            hipMalloc(&cg->d_partial, sizeof(double) * chunk_sys.size()) == hipSuccess &&
            hipMalloc(&cg->r, sizeof(double) * n) == hipSuccess && hipMalloc(&cg->p, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&cg->Ap, sizeof(double) * n) == hipSuccess;
  ok = ok && hipMalloc(&cg->d_active, sizeof(int32_t) * num_systems) == hipSuccess &&
       hipMalloc(&cg->d_prev, sizeof(int32_t) * num_systems) == hipSuccess;
  // system of every packed element of every plan (from the element's first node)
  for (int32_t i = 0; ok && i < num_ops; i++) {
    CeedMi355Operator    op = ops_host[i];
    const int64_t        es = (int64_t)op->P * op->P * op->P;
    std::vector<int32_t> off((size_t)(op->num_elem * es)), sys((size_t)op->num_elem + 1, 0);
    ok = op->num_elem == 0 || hipMemcpy(off.data(), op->d_offsets, sizeof(int32_t) * off.size(), hipMemcpyDeviceToHost) == hipSuccess;
    for (int64_t e = 0; ok && e < op->num_elem; e++) {
      const int64_t node = op_begin_host[i] + (off[(size_t)(e * es)] >> 1);
      sys[(size_t)e]     = (int32_t)(std::upper_bound(sys_begin_host, sys_begin_host + num_systems + 1, node) - sys_begin_host - 1);
    }
    int32_t *d = nullptr;
    ok = ok && hipMalloc(&d, sizeof(int32_t) * sys.size()) == hipSuccess &&
         hipMemcpy(d, sys.data(), sizeof(int32_t) * sys.size(), hipMemcpyHostToDevice) == hipSuccess;
    cg->d_elem_sys.push_back(d);
    // the plan's range [b, e) cut at the reduction chunks' bounds (a system may span plans)
    const int64_t        b = op_begin_host[i], e = b + op->l_size;
    std::vector<int64_t> cb;
    std::vector<int32_t> cs;
    for (size_t c = 0; c + 1 < chunk_begin.size(); c++) {
      const int64_t lo = std::max(chunk_begin[c], b), hi = std::min(chunk_begin[c + 1], e);
      if (lo >= hi) continue;
      cb.push_back(lo - b);
      cs.push_back(chunk_sys[c]);
    }
    cb.push_back(e - b);
    int64_t *dcb = nullptr;
    int32_t *dcs = nullptr;
    ok = ok && hipMalloc(&dcb, sizeof(int64_t) * cb.size()) == hipSuccess && hipMalloc(&dcs, sizeof(int32_t) * (cs.size() + 1)) == hipSuccess &&
         hipMemcpy(dcb, cb.data(), sizeof(int64_t) * cb.size(), hipMemcpyHostToDevice) == hipSuccess &&
         (cs.empty() || hipMemcpy(dcs, cs.data(), sizeof(int32_t) * cs.size(), hipMemcpyHostToDevice) == hipSuccess);
    cg->d_op_chunk_begin.push_back(dcb);
    cg->d_op_chunk_sys.push_back(dcs);
    cg->op_num_chunks.push_back((int64_t)cs.size());
  }
  ok = ok && hipMemcpy(cg->d_chunk_begin, chunk_begin.data(), sizeof(int64_t) * chunk_begin.size(), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(cg->d_chunk_sys, chunk_sys.data(), sizeof(int32_t) * chunk_sys.size(), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(cg->d_sys_chunk, sys_chunk.data(), sizeof(int32_t) * sys_chunk.size(), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    CeedMi355BatchCGDestroy(&cg);
    MI355_CHECK(false, CEED_MI355_ERROR_BACKEND, "batched CG set-up failed (%lld entries)", (long long)n);
  }
  *cg_out = cg;
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355BatchCGSolve(CeedMi355BatchCG cg, const double *b, double *x, int32_t max_iter, double rtol, int32_t check_every,
                                     int use_graph, void *stream, int32_t *iterations_host, double *rel_residual_host, int32_t *converged_at_check_host,
                                     int32_t *num_checks, double *converged_ms_host) {
  MI355_CHECK(cg && check_every >= 1, CEED_MI355_ERROR_BACKEND, "bad batched CG arguments");
  const auto t_start = std::chrono::steady_clock::now();
  hipStream_t s  = cg->stream;
  const int   nc = cg->num_chunks;
  int         err;
  MI355_HIP(hipEventRecord(cg->ev, (hipStream_t)stream));
  MI355_HIP(hipStreamWaitEvent(s, cg->ev, 0));
  k_zero<<<mi355_blocks((int64_t)kNumSlots * cg->num_sys, kThreads, 1024), kThreads, 0, s>>>((int64_t)kNumSlots * cg->num_sys, cg->d_scal);
  k_zero_i<<<mi355_blocks(cg->num_sys, kThreads, 1024), kThreads, 0, s>>>(cg->num_sys, cg->d_iters);
  k_chunk_dot<<<nc, kThreads, 0, s>>>(cg->d_chunk_begin, cg->d_chunk_sys, nullptr, b, b, cg->d_partial);
  MI355_LAUNCH_CHECK();
  if ((err = finish(cg, kBB, 3, rtol * rtol))) return err;
  if ((err = apply_all(cg, x, cg->Ap))) return err;
  k_chunk_residual<<<nc, kThreads, 0, s>>>(cg->d_chunk_begin, b, cg->Ap, cg->r, cg->p, cg->d_partial);
  MI355_LAUNCH_CHECK();
  if ((err = finish(cg, kRR, 0))) return err;
  if ((err = read_back(cg))) return err;
  for (int32_t i = 0; i < cg->num_sys; i++) {
    converged_at_check_host[i] = -1;
    if (converged_ms_host) converged_ms_host[i] = NAN;
  }
  int it = 0, checks = 0;
  for (;;) {
    int          active = 0;
    const double now_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    for (int32_t i = 0; i < cg->num_sys; i++) {
      const double *sc = cg->h_scal + (int64_t)i * kNumSlots;
      if (sc[kRR] > sc[kTol2]) active++;
      else if (converged_at_check_host[i] < 0) {
        converged_at_check_host[i] = checks;
        if (converged_ms_host) converged_ms_host[i] = now_ms;  // read_back has waited for this check's iterations
      }
    }
    if (!active || it >= max_iter) break;
    const int len = check_every < max_iter - it ? check_every : max_iter - it;
    if ((err = run_segment(cg, b, x, len, use_graph && len == check_every))) return err;
    it += len;
    checks++;
    if ((err = read_back(cg))) return err;
  }
  // true residuals of the returned solutions
  if ((err = apply_all(cg, x, cg->Ap))) return err;
  k_chunk_residual<<<nc, kThreads, 0, s>>>(cg->d_chunk_begin, b, cg->Ap, cg->r, nullptr, cg->d_partial);
  MI355_LAUNCH_CHECK();
  if ((err = finish(cg, kRRTrue, 0))) return err;
  if ((err = read_back(cg))) return err;
  for (int32_t i = 0; i < cg->num_sys; i++) {
    const double *sc     = cg->h_scal + (int64_t)i * kNumSlots;
    iterations_host[i]   = cg->h_iters[i];
    rel_residual_host[i] = sc[kBB] > 0 ? sqrt(sc[kRRTrue] / sc[kBB]) : 0.0;
  }
  *num_checks = checks;
  MI355_HIP(hipEventRecord(cg->ev, s));
  MI355_HIP(hipStreamWaitEvent((hipStream_t)stream, cg->ev, 0));
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355BatchCGDestroy(CeedMi355BatchCG *cg) {
  if (!cg || !*cg) return CEED_MI355_SUCCESS;
  CeedMi355BatchCG c = *cg;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  (void)hipFree(c->d_chunk_begin);
  (void)hipFree(c->d_chunk_sys);
  (void)hipFree(c->d_sys_chunk);
  (void)hipFree(c->d_iters);
  (void)hipHostFree(c->h_iters);
  (void)hipFree(c->d_scal);
  (void)hipHostFree(c->h_scal);
  (void)hipFree(c->d_partial);
  (void)hipFree(c->r);
  (void)hipFree(c->p);
  (void)hipFree(c->Ap);
  (void)hipFree(c->d_active);
  (void)hipFree(c->d_prev);
  for (int32_t *d : c->d_elem_sys) (void)hipFree(d);
  for (int64_t *d : c->d_op_chunk_begin) (void)hipFree(d);
  for (int32_t *d : c->d_op_chunk_sys) (void)hipFree(d);
  if (c->ev) (void)hipEventDestroy(c->ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  *cg = nullptr;
  return CEED_MI355_SUCCESS;
}
