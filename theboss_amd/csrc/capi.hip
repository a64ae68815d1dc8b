// C ABI plumbing shared by every entry point: version, per-thread error message,
// HIP error mapping. No C++ exception crosses the boundary.
#include <stdarg.h>
#include <stdio.h>

#include "tb_common.hpp"

namespace {
thread_local char g_last_error[512] = "";
}

namespace tbh {

int32_t fail(int32_t code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

int32_t check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return TB_OK;
  return fail(TB_EHIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

}  // namespace tbh

extern "C" int32_t tb_version(void) { return 100; /* 0.1.0 */ }

extern "C" const char *tb_last_error(void) { return g_last_error; }

extern "C" int32_t tb_device_count(void) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return tbh::check_hip(e, "hipGetDeviceCount");
  return n;
}
