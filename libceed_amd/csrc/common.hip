// Device selection, memory helpers and the per-thread error message.
#include <stdarg.h>
#include <string.h>

#include "mi355_common.h"

namespace mi355 {

static thread_local char last_error[1024];

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(last_error, sizeof(last_error), fmt, ap);
  va_end(ap);
}

}  // namespace mi355

extern "C" const char *CeedMi355GetLastError(void) { return mi355::last_error; }

extern "C" int CeedMi355DeviceInit(int device_id) {
  int count = 0;
  MI355_HIP(hipGetDeviceCount(&count));
  MI355_CHECK(count > 0, CEED_MI355_ERROR_BACKEND, "no HIP device visible");
  MI355_CHECK(device_id >= 0 && device_id < count, CEED_MI355_ERROR_BACKEND, "device_id %d out of range [0, %d)", device_id, count);
  MI355_HIP(hipSetDevice(device_id));
  hipDeviceProp_t prop;
  MI355_HIP(hipGetDeviceProperties(&prop, device_id));
  MI355_CHECK(strncmp(prop.gcnArchName, "gfx950", 6) == 0, CEED_MI355_ERROR_UNSUPPORTED,
              "device %d is %s; this library is built for gfx950 (MI355X) only", device_id, prop.gcnArchName);
  // the device address of this GPU's no-allocation reduction scratch (vector.hip)
  return mi355::reduce_scratch_setup(nullptr);
}

extern "C" int CeedMi355Malloc(size_t bytes, void **ptr) {
  *ptr = nullptr;
  if (bytes == 0) return CEED_MI355_SUCCESS;
  MI355_HIP(hipMalloc(ptr, bytes));
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355Free(void *ptr) {
  if (ptr) MI355_HIP(hipFree(ptr));
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355Memcpy(void *dst, const void *src, size_t bytes, void *stream) {
  if (bytes == 0) return CEED_MI355_SUCCESS;
  MI355_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream));
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355Memset(void *dst, int value, size_t bytes, void *stream) {
  if (bytes == 0) return CEED_MI355_SUCCESS;
  MI355_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355StreamSynchronize(void *stream) {
  MI355_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CEED_MI355_SUCCESS;
}
