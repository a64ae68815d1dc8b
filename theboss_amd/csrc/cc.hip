// Entry points built on the binary C&C sweep: the submatrices calculator and the
// permanent gradient. The mixed-radix fallback lives in guan.hip.
#include <algorithm>

#include "tb_common.hpp"

namespace tbs {
size_t sweep_partial_bytes(int k, int items, int c_override);
int32_t run_sweep(int k, const tb::cd *Vt, int ldv, int64_t item_stride, const int *item_map,
                  const int *item_count, int items, int c_override, tb::cd *partial,
                  tb::cd *minors, hipStream_t s);
}  // namespace tbs
namespace tbgr {
bool grad_supported(int n);
size_t grad_partial_bytes(int n, int batch, int c_pref);
int32_t run_grad(int n, const tb::cd *A, int64_t mstride, int batch, int c_pref, tb::cd *partial,
                 tb::cd *grad, hipStream_t s);
}  // namespace tbgr
namespace tbg {
int32_t run_guan(const tb::cd *U, int ldu, const int *out_modes, const int *out_mult, int out_max,
                 const int *in_modes, const int *in_mult, int in_max, int batch, int formula,
                 int mode, tb::cd *out, hipStream_t s, int n_in = 0);
}  // namespace tbg

namespace tb {

// Vt[b][nu][j] = U[out row j][in particle nu]: particles are the input slots with
// multiplicity 1 in slot order, output rows expand each slot's multiplicity.
__global__ void cc_gather_kernel(const cd *__restrict__ U, int ldu, const int *__restrict__ out_modes,
                                 const int *__restrict__ out_mult, int out_max,
                                 const int *__restrict__ in_modes, const int *__restrict__ in_mult,
                                 int in_max, int batch, int k, cd *__restrict__ Vt) {
  const int R = k - 1;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)batch * k * R) return;
  const int b = (int)(idx / ((int64_t)k * R));
  const int rem = (int)(idx - (int64_t)b * k * R);
  const int nu = rem / R, j = rem - nu * R;
  int col = -1, cnt = 0;
  for (int d = 0; d < in_max; ++d) {
    if (in_mult[(size_t)b * in_max + d] > 0) {
      if (cnt == nu) {
        col = in_modes[(size_t)b * in_max + d];
        break;
      }
      ++cnt;
    }
  }
  int row = -1, acc = 0;
  for (int q = 0; q < out_max; ++q) {
    acc += out_mult[(size_t)b * out_max + q];
    if (j < acc) {
      row = out_modes[(size_t)b * out_max + q];
      break;
    }
  }
  cd v = {0.0, 0.0};
  if (row >= 0 && col >= 0) v = U[(size_t)row * ldu + col];
  Vt[idx] = v;
}

// minors[b][slot] = compact[b][nu(slot)] for occupied slots, 0 elsewhere.
__global__ void cc_scatter_kernel(const cd *__restrict__ compact, const int *__restrict__ in_mult,
                                  int in_max, int batch, int k, cd *__restrict__ minors) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= batch * in_max) return;
  const int b = idx / in_max, d = idx - b * in_max;
  cd v = {0.0, 0.0};
  if (in_mult[idx] > 0) {
    int nu = 0;
    for (int e = 0; e < d; ++e)
      if (in_mult[(size_t)b * in_max + e] > 0) ++nu;
    if (nu < k) v = compact[(size_t)b * k + nu];
  }
  minors[idx] = v;
}

// k == 1: bs_submatrices_permanent_calculator_base.py:156-158 returns s itself.
__global__ void cc_single_kernel(const int *__restrict__ in_mult, int n, cd *__restrict__ minors) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) minors[idx] = {(double)in_mult[idx], 0.0};
}

// Gradient items (b, r): Vt[nu][j'] = A[b][row j' (skipping r)][nu].
__global__ void grad_gather_kernel(const cd *__restrict__ A, int n, int batch, int64_t mstride,
                                   cd *__restrict__ Vt) {
  const int R = n - 1;
  const int64_t per_item = (int64_t)n * R;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)batch * n * per_item) return;
  const int64_t item = idx / per_item;
  const int rem = (int)(idx - item * per_item);
  const int nu = rem / R, jp = rem - nu * R;
  const int b = (int)(item / n), r = (int)(item - (int64_t)b * n);
  const int row = jp < r ? jp : jp + 1;
  Vt[idx] = A[(size_t)b * mstride + (size_t)row * n + nu];
}

// perm[b] = sum_nu A[b][0][nu] * G[b][0][nu] (Laplace expansion along row 0).
__global__ void grad_perm_kernel(const cd *__restrict__ A, int n, int batch, int64_t mstride,
                                 const cd *__restrict__ G, cd *__restrict__ perm) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  double re = 0.0, im = 0.0;
  for (int nu = 0; nu < n; ++nu) {
    const cd a = A[(size_t)b * mstride + nu];
    const cd g = G[(size_t)b * n * n + nu];
    re += a.re * g.re - a.im * g.im;
    im += a.re * g.im + a.im * g.re;
  }
  perm[b] = {re, im};
}

__global__ void grad_small_kernel(const cd *__restrict__ A, int n, int batch, int64_t mstride,
                                  cd *__restrict__ perm, cd *__restrict__ G) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  if (n == 0) {
    perm[b] = {1.0, 0.0};
  } else {  // n == 1
    perm[b] = A[(size_t)b * mstride];
    G[b] = {1.0, 0.0};
  }
}

}  // namespace tb

namespace {
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t cc_workspace_bytes(int k, int batch, int c) {
  if (k < 2 || batch <= 0 || k > TB_SWEEP_MAX_K) return 0;
  const size_t vt = align256((size_t)batch * k * (k - 1) * sizeof(tb::cd));
  const size_t compact = align256((size_t)batch * k * sizeof(tb::cd));
  return vt + compact + tbs::sweep_partial_bytes(k, batch, c);
}
}  // namespace

extern "C" size_t tb_cc_workspace_size(int32_t k, int32_t out_max, int32_t in_max, int32_t batch) {
  (void)out_max;
  (void)in_max;
  if (k < 2 || k > TB_SWEEP_MAX_K) return 0;
  return cc_workspace_bytes(k, batch, tbh::tuned(TB_TUNE_SWEEP, k, batch));
}

extern "C" int32_t tb_cc_submatrices_batch(const tb_complex128 *U, int32_t ldu,
                                           const int32_t *out_modes, const int32_t *out_mult,
                                           int32_t out_max, const int32_t *in_modes,
                                           const int32_t *in_mult, int32_t in_max, int32_t batch,
                                           int32_t k, int32_t collision_free,
                                           tb_complex128 *minors, void *workspace,
                                           size_t workspace_bytes, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (batch < 0) return tbh::fail(TB_EINVAL, "cc: batch=%d < 0", batch);
  if (batch == 0) return TB_OK;
  if (!in_mult || !minors) return tbh::fail(TB_EINVAL, "cc: NULL pointer");
  if (in_max < 1) return tbh::fail(TB_EINVAL, "cc: in_max=%d < 1", in_max);
  if (k < 1) return tbh::fail(TB_EINVAL, "cc: k=%d < 1", k);
  tb::cd *mo = (tb::cd *)minors;
  if (k == 1) {
    const int n = batch * in_max;
    hipLaunchKernelGGL(tb::cc_single_kernel, dim3((n + 255) / 256), dim3(256), 0, s, in_mult, n, mo);
    return tbh::check_hip(hipGetLastError(), "cc_single_kernel");
  }
  if (!U || !out_modes || !out_mult || !in_modes) return tbh::fail(TB_EINVAL, "cc: NULL pointer");
  if (out_max < 1 || ldu < 1)
    return tbh::fail(TB_EINVAL, "cc: out_max=%d ldu=%d must be >= 1", out_max, ldu);
  if (!collision_free) {
    if (out_max > TB_GUAN_MAX_ROWS || in_max > TB_GUAN_MAX_COLS)
      return tbh::fail(TB_ERANGE, "cc: out_max=%d > %d or in_max=%d > %d", out_max,
                       TB_GUAN_MAX_ROWS, in_max, TB_GUAN_MAX_COLS);
    return tbg::run_guan((const tb::cd *)U, ldu, out_modes, out_mult, out_max, in_modes, in_mult,
                         in_max, batch, TB_FORMULA_RYSER, 1, mo, s);
  }
  if (k > TB_SWEEP_MAX_K)
    return tbh::fail(TB_ERANGE, "cc: k=%d exceeds TB_SWEEP_MAX_K=%d", k, TB_SWEEP_MAX_K);
  const int c = tbh::tuned(TB_TUNE_SWEEP, k, batch);  // read once: size check and launch agree
  const size_t need = cc_workspace_bytes(k, batch, c);
  if (!workspace || workspace_bytes < need)
    return tbh::fail(TB_EWORKSPACE, "cc: workspace %zu bytes < required %zu", workspace_bytes,
                     need);
  char *ws = (char *)workspace;
  tb::cd *Vt = (tb::cd *)ws;
  ws += align256((size_t)batch * k * (k - 1) * sizeof(tb::cd));
  tb::cd *compact = (tb::cd *)ws;
  ws += align256((size_t)batch * k * sizeof(tb::cd));
  tb::cd *partial = (tb::cd *)ws;
  const int64_t total = (int64_t)batch * k * (k - 1);
  hipLaunchKernelGGL(tb::cc_gather_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     (const tb::cd *)U, ldu, out_modes, out_mult, out_max, in_modes, in_mult,
                     in_max, batch, k, Vt);
  int32_t rc = tbh::check_hip(hipGetLastError(), "cc_gather_kernel");
  if (rc) return rc;
  rc = tbs::run_sweep(k, Vt, k - 1, (int64_t)k * (k - 1), nullptr, nullptr, batch, c, partial,
                      compact, s);
  if (rc) return rc;
  const int n = batch * in_max;
  hipLaunchKernelGGL(tb::cc_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, compact,
                     in_mult, in_max, batch, k, mo);
  return tbh::check_hip(hipGetLastError(), "cc_scatter_kernel");
}

namespace {
// c: the tuning entry of the path that n takes (single-sweep gradient over the batch, or
// n sweeps per matrix over batch * n sweep items).
int grad_tuned(int n, int batch) {
  if (tbgr::grad_supported(n)) return tbh::tuned(TB_TUNE_GRAD, n, batch);
  return tbh::tuned(TB_TUNE_SWEEP, n, (int64_t)batch * n);
}

size_t grad_workspace_bytes(int n, int batch, int c) {
  if (n < 2 || batch <= 0 || n > TB_SWEEP_MAX_K) return 0;
  if (tbgr::grad_supported(n)) return tbgr::grad_partial_bytes(n, batch, c);
  const int items = batch * n;
  return align256((size_t)items * n * (n - 1) * sizeof(tb::cd)) +
         tbs::sweep_partial_bytes(n, items, c);
}
}  // namespace

extern "C" size_t tb_perm_grad_workspace_size(int32_t n, int32_t batch) {
  if (n < 2 || n > TB_SWEEP_MAX_K) return 0;
  return grad_workspace_bytes(n, batch, grad_tuned(n, batch));
}

extern "C" int32_t tb_perm_grad_batch(const tb_complex128 *A, int32_t n, int32_t batch,
                                      int64_t matrix_stride, tb_complex128 *perm_out,
                                      tb_complex128 *grad_out, void *workspace,
                                      size_t workspace_bytes, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (batch < 0) return tbh::fail(TB_EINVAL, "grad: batch=%d < 0", batch);
  if (batch == 0) return TB_OK;
  if (n < 0) return tbh::fail(TB_EINVAL, "grad: n=%d < 0", n);
  if (n > TB_SWEEP_MAX_K)
    return tbh::fail(TB_ERANGE, "grad: n=%d exceeds TB_SWEEP_MAX_K=%d", n, TB_SWEEP_MAX_K);
  if (!perm_out || (n > 0 && (!A || !grad_out))) return tbh::fail(TB_EINVAL, "grad: NULL pointer");
  if (n > 0 && matrix_stride < (int64_t)n * n)
    return tbh::fail(TB_EINVAL, "grad: matrix_stride=%lld < n*n", (long long)matrix_stride);
  if (n < 2) {
    hipLaunchKernelGGL(tb::grad_small_kernel, dim3((batch + 255) / 256), dim3(256), 0, s,
                       (const tb::cd *)A, n, batch, matrix_stride, (tb::cd *)perm_out,
                       (tb::cd *)grad_out);
    return tbh::check_hip(hipGetLastError(), "grad_small_kernel");
  }
  const int c = grad_tuned(n, batch);  // read once: size check and launch agree
  const size_t need = grad_workspace_bytes(n, batch, c);
  if (!workspace || workspace_bytes < need)
    return tbh::fail(TB_EWORKSPACE, "grad: workspace %zu bytes < required %zu", workspace_bytes,
                     need);
  if (tbgr::grad_supported(n)) {  // one sweep per group of 4 or 8 columns
    int32_t rc = tbgr::run_grad(n, (const tb::cd *)A, matrix_stride, batch, c,
                                (tb::cd *)workspace, (tb::cd *)grad_out, s);
    if (rc) return rc;
  } else {  // one C&C minors sweep per deleted row
    const int items = batch * n;
    tb::cd *Vt = (tb::cd *)workspace;
    tb::cd *partial =
        (tb::cd *)((char *)workspace + align256((size_t)items * n * (n - 1) * sizeof(tb::cd)));
    const int64_t total = (int64_t)items * n * (n - 1);
    hipLaunchKernelGGL(tb::grad_gather_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       s, (const tb::cd *)A, n, batch, matrix_stride, Vt);
    int32_t rc = tbh::check_hip(hipGetLastError(), "grad_gather_kernel");
    if (rc) return rc;
    rc = tbs::run_sweep(n, Vt, n - 1, (int64_t)n * (n - 1), nullptr, nullptr, items, c, partial,
                        (tb::cd *)grad_out, s);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(tb::grad_perm_kernel, dim3((batch + 255) / 256), dim3(256), 0, s,
                     (const tb::cd *)A, n, batch, matrix_stride, (const tb::cd *)grad_out,
                     (tb::cd *)perm_out);
  return tbh::check_hip(hipGetLastError(), "grad_perm_kernel");
}
