// C ABI plumbing shared by every entry point: version, per-thread error message,
// HIP error mapping. No C++ exception crosses the boundary.
#include <stdarg.h>
#include <stdio.h>

#include <atomic>

#include "tb_common.hpp"

namespace {
thread_local char g_last_error[512] = "";
// Tuning table, zero (built-in plan) until set; static storage is zero-initialised.
std::atomic<int32_t> g_tune[TB_TUNE_GRAD + 1][TB_TUNE_MAX_SIZE + 1][TB_TUNE_BATCH_BUCKETS];
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

// Bucket ceil(log2(batch)): 1 | 2 | 3-4 | 5-8 | ... | above 2^(buckets-2).
int tune_bucket(int64_t batch) {
  int b = 0;
  while (b < TB_TUNE_BATCH_BUCKETS - 1 && ((int64_t)1 << b) < batch) ++b;
  return b;
}

int32_t tuned(int kernel, int size, int64_t batch) {
  if (kernel < TB_TUNE_SWEEP || kernel > TB_TUNE_GRAD || size < 0 || size > TB_TUNE_MAX_SIZE)
    return 0;
  return g_tune[kernel][size][tune_bucket(batch)].load(std::memory_order_relaxed);
}

}  // namespace tbh

namespace {
int32_t tune_args(int32_t kernel, int32_t size) {
  if (kernel < TB_TUNE_SWEEP || kernel > TB_TUNE_GRAD)
    return tbh::fail(TB_EINVAL, "tune: unknown kernel %d", kernel);
  if (size < 0 || size > TB_TUNE_MAX_SIZE)
    return tbh::fail(TB_ERANGE, "tune: size %d outside 0..%d", size, TB_TUNE_MAX_SIZE);
  return TB_OK;
}
}  // namespace

extern "C" int32_t tb_tune_set(int32_t kernel, int32_t size, int64_t batch, int32_t chunk_bits) {
  const int32_t rc = tune_args(kernel, size);
  if (rc) return rc;
  if (chunk_bits < 0 || chunk_bits > 16)
    return tbh::fail(TB_EINVAL, "tune: chunk_bits %d outside 0..16", chunk_bits);
  if (batch <= 0) {  // every bucket
    for (int b = 0; b < TB_TUNE_BATCH_BUCKETS; ++b)
      g_tune[kernel][size][b].store(chunk_bits, std::memory_order_relaxed);
  } else {
    g_tune[kernel][size][tbh::tune_bucket(batch)].store(chunk_bits, std::memory_order_relaxed);
  }
  return TB_OK;
}

extern "C" int32_t tb_tune_get(int32_t kernel, int32_t size, int64_t batch) {
  const int32_t rc = tune_args(kernel, size);
  if (rc) return rc;
  if (batch < 1) return tbh::fail(TB_EINVAL, "tune: batch %lld < 1", (long long)batch);
  return g_tune[kernel][size][tbh::tune_bucket(batch)].load(std::memory_order_relaxed);
}

extern "C" int32_t tb_version(void) { return 100; /* 0.1.0 */ }

extern "C" const char *tb_last_error(void) { return g_last_error; }

extern "C" int32_t tb_device_count(void) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return tbh::check_hip(e, "hipGetDeviceCount");
  return n;
}
