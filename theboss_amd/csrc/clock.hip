// Shader clock probe: one wave reads the shader cycle counter (s_memtime) and the
// constant-rate real-time counter (s_memrealtime, tb_realtime_hz) at its start, sleeps
// until the real-time counter has advanced by `ticks` (or an iteration cap is reached,
// so the wave always ends), and stores both deltas. Launched on a second stream beside a
// long kernel, cycles / ticks * tb_realtime_hz() is the average clock of a CU while that
// kernel runs: the figure that separates instruction count from clock in a VALU-bound
// kernel's fraction of peak.
#include "tb_common.hpp"

namespace tb {

__global__ __launch_bounds__(64) void clock_probe_kernel(uint64_t *__restrict__ out,
                                                         uint64_t ticks, int max_iters) {
  const uint64_t c0 = __builtin_amdgcn_s_memtime();
  const uint64_t r0 = __builtin_amdgcn_s_memrealtime();
  uint64_t c = c0, r = r0;
  for (int it = 0; it < max_iters && r - r0 < ticks; ++it) {
    __builtin_amdgcn_s_sleep(16);
    c = __builtin_amdgcn_s_memtime();
    r = __builtin_amdgcn_s_memrealtime();
  }
  if (threadIdx.x == 0) {
    out[0] = c - c0;
    out[1] = r - r0;
  }
}

}  // namespace tb

extern "C" int64_t tb_realtime_hz(void) {
  int dev = 0, khz = 0;
  int32_t rc = tbh::check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  rc = tbh::check_hip(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev),
                      "hipDeviceGetAttribute");
  if (rc) return rc;
  return (int64_t)khz * 1000;
}

extern "C" int32_t tb_clock_probe(uint64_t *out, int64_t ticks, void *stream) {
  if (!out) return tbh::fail(TB_EINVAL, "clock_probe: out is NULL");
  if (ticks < 1 || ticks > ((int64_t)1 << 32))
    return tbh::fail(TB_EINVAL, "clock_probe: ticks=%lld outside 1..2^32", (long long)ticks);
  // One iteration sleeps ~1024 clocks (~0.4 us at 2.4 GHz, 40 ticks of a 100 MHz counter):
  // the cap is reached only if the real-time counter runs far slower than it can.
  const int64_t cap = ticks / 2 + 1024;
  hipLaunchKernelGGL(tb::clock_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, out,
                     (uint64_t)ticks, (int)(cap > (1 << 30) ? (1 << 30) : cap));
  return tbh::check_hip(hipGetLastError(), "clock_probe_kernel");
}
