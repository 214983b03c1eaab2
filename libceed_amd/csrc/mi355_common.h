// Shared helpers for the MI355X (gfx950) kernels: error reporting and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/ceed-mi355-kernels.h"

namespace mi355 {

// Last error message of the calling host thread (read by CeedMi355GetLastError).
void set_error(const char *fmt, ...);

// The current device's address of the fixed scratch array of
// CeedMi355VectorReduceNoAlloc (vector.hip), looked up on the first call for
// a device and kept; ptr may be nullptr (set-up only).
int reduce_scratch_setup(double **ptr);

}  // namespace mi355

#define MI355_CHECK(cond, code, ...)    \
  do {                                  \
    if (!(cond)) {                      \
      mi355::set_error(__VA_ARGS__);    \
      return code;                      \
    }                                   \
  } while (0)

#define MI355_HIP(call)                                                                                         \
  do {                                                                                                          \
    hipError_t e_ = (call);                                                                                     \
    if (e_ != hipSuccess) {                                                                                     \
      mi355::set_error("HIP error %s (%d) at %s:%d: %s", hipGetErrorName(e_), (int)e_, __FILE__, __LINE__, #call); \
      return CEED_MI355_ERROR_BACKEND;                                                                          \
    }                                                                                                           \
  } while (0)

#define MI355_LAUNCH_CHECK() MI355_HIP(hipGetLastError())

static inline int mi355_blocks(int64_t n, int threads, int64_t cap = 1 << 20) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}
