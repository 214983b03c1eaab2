// L-vector kernels for the backend's CeedVector and the CG solvers:
// set / scale / axpy / axpby / pointwise multiply / reciprocal, and
// deterministic reductions (fixed grid, fixed-order tree: the same inputs give
// the same bits on every run).  FP64 and FP32 variants.
#include "mi355_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kReduceBlocks = 1024;

template <typename T>
__global__ void __launch_bounds__(kThreads) k_set(int64_t n, T a, T *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = a;
}
template <typename T>
__global__ void __launch_bounds__(kThreads) k_scale(int64_t n, T a, T *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] *= a;
}
template <typename T>
__global__ void __launch_bounds__(kThreads) k_axpby(int64_t n, T a, T b, const T *__restrict__ x, T *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = a * x[i] + b * y[i];
}
template <typename T>
__global__ void __launch_bounds__(kThreads) k_axpy(int64_t n, T a, const T *__restrict__ x, T *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] += a * x[i];
}
template <typename T>
__global__ void __launch_bounds__(kThreads) k_pmult(int64_t n, T *__restrict__ w, const T *__restrict__ x, const T *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) w[i] = x[i] * y[i];
}
// 1/x where |x| > eps (CeedVectorReciprocal leaves (near-)zero entries alone)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_recip(int64_t n, T eps, T *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (fabs((double)x[i]) > (double)eps) x[i] = T(1) / x[i];
}
template <typename TS, typename TD>
__global__ void __launch_bounds__(kThreads) k_convert(int64_t n, const TS *__restrict__ x, TD *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = (TD)x[i];
}

enum { kSum = 0, kAbsSum = 1, kMaxAbs = 2 };

__device__ inline double combine(int op, double a, double b) { return op == kMaxAbs ? fmax(a, b) : a + b; }

// Fixed-order tree over the block; returns the block value in thread 0.
__device__ double block_reduce(int op, double v) {
  __shared__ double s[kThreads];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = combine(op, s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  return s[0];
}

// stage 1: partial[b] over a fixed grid-stride range; terms x*y (kSum), |x| (kAbsSum), |x| (kMaxAbs)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_reduce1(int op, int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                      double *__restrict__ partial) {
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double xi = (double)x[i];
    const double t  = op == kSum ? xi * (double)y[i] : fabs(xi);
    acc             = combine(op, acc, t);
  }
  const double b = block_reduce(op, acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

__global__ void __launch_bounds__(kThreads) k_reduce2(int op, int nb, const double *__restrict__ partial, double *__restrict__ out) {
  double acc = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc = combine(op, acc, partial[i]);
  const double b = block_reduce(op, acc);
  if (threadIdx.x == 0) *out = b;
}

// Per-device scratch for block partials (allocated on first use, kept for the process).
double *partials(int &ierr) {
  static double *buf[64] = {nullptr};
  int            dev     = 0;
  ierr                   = hipGetDevice(&dev) == hipSuccess ? 0 : CEED_MI355_ERROR_BACKEND;
  if (ierr || dev < 0 || dev >= 64) return nullptr;
  if (!buf[dev] && hipMalloc(&buf[dev], sizeof(double) * (kReduceBlocks + 8)) != hipSuccess) ierr = CEED_MI355_ERROR_BACKEND;
  return buf[dev];
}

template <typename T>
int reduce(int op, int64_t n, const T *x, const T *y, double *result_device, double *partial, hipStream_t s) {
  const int nb = mi355_blocks(n, kThreads, kReduceBlocks);
  k_reduce1<T><<<nb, kThreads, 0, s>>>(op, n, x, y, partial);
  k_reduce2<<<1, kThreads, 0, s>>>(op, nb, partial, result_device);
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

// The partial sums of CeedMi355VectorReduceNoAlloc: one fixed array in device
// memory, which every GPU the library runs on holds (its own copy, at its own
// address) until the process ends.  Host code reaches it only through the
// per-device address hipGetSymbolAddress returns: the array's name is a
// host-side stand-in, not a device address.  Every call on one GPU shares its
// copy, so calls that overlap on one GPU race.
constexpr int kScratchPartials = kReduceBlocks;
static_assert(kScratchPartials >= kReduceBlocks, "one partial per block of the reduction grid");
__device__ __align__(256) double scratch[kScratchPartials];
double *scratch_ptr[64];  // per device id: the address of that GPU's copy, nullptr until set up

template <typename T>
int run_vector_op(int op, int64_t n, double a, double b, const void *x, const void *y, void *z, hipStream_t s) {
  const int nb = mi355_blocks(n, kThreads, 8192);
  switch (op) {
    case 0: k_set<T><<<nb, kThreads, 0, s>>>(n, (T)a, (T *)z); break;
    case 1: k_scale<T><<<nb, kThreads, 0, s>>>(n, (T)a, (T *)z); break;
    case 2: k_axpy<T><<<nb, kThreads, 0, s>>>(n, (T)a, (const T *)x, (T *)z); break;
    case 3: k_axpby<T><<<nb, kThreads, 0, s>>>(n, (T)a, (T)b, (const T *)x, (T *)z); break;
    case 4: k_pmult<T><<<nb, kThreads, 0, s>>>(n, (T *)z, (const T *)x, (const T *)y); break;
    case 5: k_recip<T><<<nb, kThreads, 0, s>>>(n, (T)a, (T *)z); break;
    default: MI355_CHECK(false, CEED_MI355_ERROR_BACKEND, "unknown vector op %d", op);
  }
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

}  // namespace

// The current device's copy of the no-allocation scratch array: its address,
// looked up once per device (hipGetSymbolAddress allocates nothing).
int mi355::reduce_scratch_setup(double **ptr) {
  int dev = 0;
  MI355_HIP(hipGetDevice(&dev));
  MI355_CHECK(dev >= 0 && dev < 64, CEED_MI355_ERROR_BACKEND, "device id %d outside the scratch table", dev);
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    MI355_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MI355_CHECK(p != nullptr, CEED_MI355_ERROR_BACKEND, "no device address for the reduction scratch array on device %d", dev);
    scratch_ptr[dev] = (double *)p;
  }
  if (ptr) *ptr = scratch_ptr[dev];
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355VectorOp(int op, int scalar_type, int64_t n, double alpha, double beta, const void *x, const void *y, void *z,
                                 void *stream) {
  if (n <= 0) return CEED_MI355_SUCCESS;
  hipStream_t s = (hipStream_t)stream;
  return scalar_type == CEED_MI355_FP64 ? run_vector_op<double>(op, n, alpha, beta, x, y, z, s)
                                        : run_vector_op<float>(op, n, alpha, beta, x, y, z, s);
}

extern "C" int CeedMi355VectorReduce(int op, int scalar_type, int64_t n, const void *x, const void *y, double *result_device, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  MI355_CHECK(op >= 0 && op <= 2, CEED_MI355_ERROR_BACKEND, "unknown reduction %d", op);
  if (n <= 0) {  // fill kernel, not a memset node: stays valid under graph capture
    k_set<double><<<1, kThreads, 0, s>>>(1, 0.0, result_device);
    MI355_LAUNCH_CHECK();
    return CEED_MI355_SUCCESS;
  }
  int     ierr = 0;
  double *p    = partials(ierr);
  MI355_CHECK(!ierr, CEED_MI355_ERROR_BACKEND, "reduction scratch allocation failed");
  return scalar_type == CEED_MI355_FP64 ? reduce<double>(op, n, (const double *)x, (const double *)y, result_device, p, s)
                                        : reduce<float>(op, n, (const float *)x, (const float *)y, result_device, p, s);
}

extern "C" int CeedMi355VectorReduceNoAlloc(int op, int scalar_type, int64_t n, const void *x, const void *y, double *result_device,
                                            void *stream) {
  hipStream_t s = (hipStream_t)stream;
  MI355_CHECK(op >= 0 && op <= 2, CEED_MI355_ERROR_BACKEND, "unknown reduction %d", op);
  MI355_CHECK(scalar_type == CEED_MI355_FP64 || scalar_type == CEED_MI355_FP32, CEED_MI355_ERROR_BACKEND, "unknown scalar type %d",
              scalar_type);
  if (n <= 0) {
    k_set<double><<<1, kThreads, 0, s>>>(1, 0.0, result_device);
    MI355_LAUNCH_CHECK();
    return CEED_MI355_SUCCESS;
  }
  // the partials take one slot per block of the reduction grid, which is
  // capped at kReduceBlocks blocks: every n fits the fixed array
  const int nb = mi355_blocks(n, kThreads, kReduceBlocks);
  MI355_CHECK(nb <= kScratchPartials, CEED_MI355_ERROR_BACKEND, "reduction of %lld entries needs %d partial slots, the scratch array holds %d",
              (long long)n, nb, kScratchPartials);
  double   *p    = nullptr;
  const int ierr = mi355::reduce_scratch_setup(&p);  // first call on this device only; later calls read the table
  if (ierr) return ierr;
  return scalar_type == CEED_MI355_FP64 ? reduce<double>(op, n, (const double *)x, (const double *)y, result_device, p, s)
                                        : reduce<float>(op, n, (const float *)x, (const float *)y, result_device, p, s);
}

extern "C" int CeedMi355VectorConvert(int64_t n, int from_type, const void *x, int to_type, void *y, void *stream) {
  if (n <= 0) return CEED_MI355_SUCCESS;
  hipStream_t s  = (hipStream_t)stream;
  const int   nb = mi355_blocks(n, kThreads, 8192);
  if (from_type == CEED_MI355_FP64 && to_type == CEED_MI355_FP32) k_convert<double, float><<<nb, kThreads, 0, s>>>(n, (const double *)x, (float *)y);
  else if (from_type == CEED_MI355_FP32 && to_type == CEED_MI355_FP64)
    k_convert<float, double><<<nb, kThreads, 0, s>>>(n, (const float *)x, (double *)y);
  else if (from_type == to_type) MI355_HIP(hipMemcpyAsync(y, x, n * (from_type == CEED_MI355_FP64 ? 8 : 4), hipMemcpyDeviceToDevice, s));
  else MI355_CHECK(false, CEED_MI355_ERROR_BACKEND, "bad conversion");
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}
