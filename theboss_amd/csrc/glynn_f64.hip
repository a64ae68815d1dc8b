// FP64 Glynn: instantiates the launch templates for n = 2..TB_GLYNN_MAX_N and holds the
// C entry points.
#include <mutex>

#include "glynn_impl.hpp"

namespace tbgl {
int32_t glynn_dispatch_f64(int n, const tb::cpx<double> *A, int64_t mstride, int batch,
                           tb::cd *out, tb::cd *partial, int *handled, int wpm_cap, int c_pref,
                           hipStream_t s) {
  return dispatch_glynn<double>(n, A, mstride, (const tb::cd *)A, mstride, batch, out, partial,
                                handled, wpm_cap, c_pref, s);
}
}  // namespace tbgl

namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The "handled" flags of the paired walk (one int per matrix) ahead of the partials.
size_t tbgl_flag_bytes(int batch) { return align256((size_t)batch * sizeof(int)); }

// Partials per matrix the workspace holds: the largest over every plan a launch of n can
// make, i.e. every lower bound on C (the static bits, 2..4, and the paired walks' static
// bits plus one) and every residency the plan may count (2, 3 or 4 waves per SIMD).
int tbgl_partial_wpm(int n, int batch, int c_pref) {
  int wpm = 0;
  for (int u : {2, 3, 4, tbgl::glynn_u_max(n)})
    for (int slots : {2, 3, 4}) wpm = std::max(wpm, tbgl::glynn_plan(n, batch, c_pref, u, slots).wpm);
  return wpm;
}

size_t tbgl_partial_bytes(int n, int batch, int c_pref) {
  if (n < tbgl::kWideMinN) return 0;
  return tbgl_flag_bytes(batch) +
         align256((size_t)batch * tbgl_partial_wpm(n, batch, c_pref) * sizeof(tb::cd));
}

// The precision a call runs in: FP32 column sums only up to TB_FP32_COLSUM_MAX_N unless
// the caller asks for them at any n (error studies); larger n keep FP64 end to end.
bool runs_fp32(int n, int precision) {
  return precision == TB_PREC_FP32_COLSUM_ANY_N ||
         (precision == TB_PREC_FP32_COLSUM && n <= TB_FP32_COLSUM_MAX_N);
}

size_t glynn_workspace_bytes(int n, int batch, int precision, int c_pref) {
  if (n <= 1 || batch <= 0 || n > TB_GLYNN_MAX_N) return 0;
  size_t bytes = tbgl_partial_bytes(n, batch, c_pref);
  if (runs_fp32(n, precision)) bytes += align256((size_t)batch * n * n * sizeof(tb::cpx<float>));
  return bytes;
}
}  // namespace

extern "C" size_t tb_glynn_workspace_size(int32_t n, int32_t batch, int32_t precision) {
  return glynn_workspace_bytes(n, batch, precision, tbh::tuned(TB_TUNE_GLYNN, n, batch));
}

extern "C" int32_t tb_glynn_batch(const tb_complex128 *A, int32_t n, int32_t batch,
                                  int64_t matrix_stride, int32_t precision, tb_complex128 *out,
                                  void *workspace, size_t workspace_bytes, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (batch < 0) return tbh::fail(TB_EINVAL, "glynn: batch=%d < 0", batch);
  if (batch == 0) return TB_OK;
  if (!out) return tbh::fail(TB_EINVAL, "glynn: out is NULL");
  if (n < 0) return tbh::fail(TB_EINVAL, "glynn: n=%d < 0", n);
  if (n > TB_GLYNN_MAX_N)
    return tbh::fail(TB_ERANGE, "glynn: n=%d exceeds TB_GLYNN_MAX_N=%d", n, TB_GLYNN_MAX_N);
  if (precision != TB_PREC_FP64 && precision != TB_PREC_FP32_COLSUM &&
      precision != TB_PREC_FP32_COLSUM_ANY_N)
    return tbh::fail(TB_EINVAL, "glynn: unknown precision %d", precision);
  tb::cd *o = (tb::cd *)out;
  if (n == 0) {
    hipLaunchKernelGGL(tb::fill_one_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, o, batch);
    return tbh::check_hip(hipGetLastError(), "fill_one_kernel");
  }
  if (!A) return tbh::fail(TB_EINVAL, "glynn: A is NULL");
  if (matrix_stride < (int64_t)n * n)
    return tbh::fail(TB_EINVAL, "glynn: matrix_stride=%lld < n*n=%d", (long long)matrix_stride,
                     n * n);
  const int c_pref = tbh::tuned(TB_TUNE_GLYNN, n, batch);  // read once: size check and launch agree
  const size_t need = glynn_workspace_bytes(n, batch, precision, c_pref);
  if (need > 0 && (workspace == nullptr || workspace_bytes < need))
    return tbh::fail(TB_EWORKSPACE, "glynn: workspace %zu bytes < required %zu", workspace_bytes,
                     need);
  if (n == 1) {
    hipLaunchKernelGGL(tb::copy_scalar_kernel, dim3((batch + 255) / 256), dim3(256), 0, s,
                       (const tb::cd *)A, matrix_stride, batch, o);
    return tbh::check_hip(hipGetLastError(), "copy_scalar_kernel");
  }
  char *ws = (char *)workspace;
  int *handled = (int *)ws;
  tb::cd *partial = ws ? (tb::cd *)(ws + tbgl_flag_bytes(batch)) : nullptr;
  const int wpm_cap = n >= tbgl::kWideMinN ? tbgl_partial_wpm(n, batch, c_pref) : 0;
  if (!runs_fp32(n, precision))
    return tbgl::glynn_dispatch_f64(n, (const tb::cpx<double> *)A, matrix_stride, batch, o,
                                    partial, handled, wpm_cap, c_pref, s);
  // FP32 column sums: convert once into the tail of the workspace.
  float *A32 = (float *)(ws + tbgl_partial_bytes(n, batch, c_pref));
  const int64_t total = (int64_t)batch * n * n;
  hipLaunchKernelGGL(tb::to_fp32_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     (const tb::cd *)A, matrix_stride, n, batch, A32);
  int32_t rc = tbh::check_hip(hipGetLastError(), "to_fp32_kernel");
  if (rc) return rc;
  return tbgl::glynn_dispatch_f32(n, (const tb::cpx<float> *)A32, (int64_t)n * n,
                                  (const tb::cd *)A, matrix_stride, batch, o, partial, handled,
                                  wpm_cap, c_pref, s);
}

// ------------------------------------------------------------------------------------
// tb_glynn_batch_noalloc: the same launches with their scratch memory (per-wave partial
// sums, the FP32 copy of the matrices) in a static device array instead of a caller's
// workspace, for callers that cannot allocate. `scratch` is the array's host-side name
// only; its address on a device comes from hipGetSymbolAddress, taken once per device
// with that device current, and that address is what the launches receive.
// ------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[TB_GLYNN_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_rc[kMaxDevices];
std::once_flag g_scratch_once[kMaxDevices];

// Address of `scratch` on the current device (resolved on the first call on it).
int32_t scratch_on_current_device(void **ptr) {
  int dev = 0;
  int32_t rc = tbh::check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices)
    return tbh::fail(TB_ERANGE, "glynn_noalloc: device %d beyond %d", dev, kMaxDevices - 1);
  std::call_once(g_scratch_once[dev], [dev]() {
    g_scratch_rc[dev] = hipGetSymbolAddress(&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
  });
  rc = tbh::check_hip(g_scratch_rc[dev], "hipGetSymbolAddress(scratch)");
  if (rc) return rc;
  *ptr = g_scratch_ptr[dev];
  return TB_OK;
}
}  // namespace

extern "C" int32_t tb_glynn_batch_noalloc(const tb_complex128 *A, int32_t n, int32_t batch,
                                          int64_t matrix_stride, int32_t precision,
                                          tb_complex128 *out, void *stream) {
  if (batch < 0) return tbh::fail(TB_EINVAL, "glynn_noalloc: batch=%d < 0", batch);
  if (n > TB_GLYNN_MAX_N)
    return tbh::fail(TB_ERANGE, "glynn_noalloc: n=%d exceeds TB_GLYNN_MAX_N=%d", n, TB_GLYNN_MAX_N);
  const size_t need =
      n < 0 ? 0 : glynn_workspace_bytes(n, batch, precision, tbh::tuned(TB_TUNE_GLYNN, n, batch));
  if (need > (size_t)TB_GLYNN_NOALLOC_SCRATCH_BYTES)
    return tbh::fail(TB_ERANGE,
                     "glynn_noalloc: n=%d batch=%d needs %zu bytes of scratch, more than the "
                     "static %u (split the batch or use tb_glynn_batch with a workspace)",
                     n, batch, need, (unsigned)TB_GLYNN_NOALLOC_SCRATCH_BYTES);
  void *ws = nullptr;
  if (need > 0) {
    const int32_t rc = scratch_on_current_device(&ws);
    if (rc) return rc;
  }
  // The workspace entry point re-validates every argument and re-checks the size against
  // the whole array, so a tuning entry changed in between cannot overrun it.
  return tb_glynn_batch(A, n, batch, matrix_stride, precision, out, ws,
                        ws ? (size_t)TB_GLYNN_NOALLOC_SCRATCH_BYTES : 0, stream);
}
