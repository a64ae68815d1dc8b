// Clifford & Clifford submatrices sweep kernels and their host launch (the Gray-walk
// building blocks and the formula are in sweep_impl.hpp).
#include "sweep_impl.hpp"

namespace tb {

// Persistent grid: G one-wave blocks loop over the tasks (slot, w), slot < active
// items, w < wpi. Blocks are dispatched round-robin over the 8 XCDs (each with its own
// L2); XCD x takes the contiguous task range [x T/8, (x+1) T/8) so that the waves of one
// item run on one XCD and its columns are fetched into one L2 only. The active item
// count may come from device memory, so a captured graph launches the same grid for
// any number of active items and idle tasks cost nothing.
template <int K, int U, int R0>
__global__ __launch_bounds__(64) TB_SWEEP_ATTR(K) void sweep_kernel(
    const cd *__restrict__ Vt, int ldv, int64_t item_stride, const int *__restrict__ item_map,
    const int *__restrict__ item_count, int items, int c, int wpi, int iters,
    cd *__restrict__ partial) {
  constexpr int R = K - 1;
  extern __shared__ cd smem[];
  cd *Vs = smem;          // K*R
  cd *Ed = smem + K * R;  // (c - U) * 64 corrector slots
  const int lane = threadIdx.x;
  int nitems = items;
  if (item_count) nitems = min(items, *item_count);
  constexpr int kXcd = 8;
  const int G = gridDim.x, p = blockIdx.x;
  const int x = p % kXcd, bi = p / kXcd;
  const int nblk_x = G / kXcd + (x < G % kXcd ? 1 : 0);
  const int64_t total = (int64_t)nitems * wpi;
  const int64_t per_x = (total + kXcd - 1) / kXcd;
  const int64_t lo = (int64_t)x * per_x;
  const int64_t hi = min(total, lo + per_x);
  for (int64_t task = lo + bi; task < hi; task += nblk_x) {
    const int slot = (int)(task / wpi);
    const int w = (int)(task - (int64_t)slot * wpi);
    const int item = item_map ? item_map[slot] : slot;
    const cd *src = Vt + (size_t)item * item_stride;
    __syncthreads();  // previous task's reads of Vs are done
    for (int e = lane; e < K * R; e += 64) {
      const int nu = e / R, j = e - nu * R;
      Vs[e] = src[(size_t)nu * ldv + j];
    }
    __syncthreads();
    const cd acc = sweep_wave<K, U, R0>(Vs, Ed, lane, c, w, wpi, iters);
    if (lane < K) partial[((size_t)slot * wpi + w) * K + lane] = acc;
  }
}

// Paired variant of sweep_kernel (same grid, tasks and partial layout). Per task the wave
// stages the item, picks the particle whose smallest column entry is largest (relative to
// the item's largest entry, at least 2^-16), swaps it into particle 1, divides the other
// columns by it and runs the paired walk; an item without such a particle takes the plain
// walk. The swap is undone when the minors are written. LDS: Vs | Ed | inv[R] | pmin[K] |
// pmax[K] | Kc, exponent, pick.
template <int K, int U>
__global__ __launch_bounds__(64) TB_SWEEP_PAIR_ATTR(K) void sweep_pair_kernel(
    const cd *__restrict__ Vt, int ldv, int64_t item_stride, const int *__restrict__ item_map,
    const int *__restrict__ item_count, int items, int c, int wpi, int iters,
    cd *__restrict__ partial) {
  constexpr int R = K - 1;
  extern __shared__ cd smem[];
  cd *Vs = smem;                                       // K*R
  cd *Ed = smem + K * R;                               // (c - U) * 64
  cd *inv = Ed + (size_t)(c > U ? c - U : 0) * 64;     // R
  double *pmin = (double *)(inv + R);                  // K
  double *pmax = pmin + K;                             // K
  cd *kc = (cd *)(pmax + K);                           // 1 (2K doubles keep 16-byte alignment)
  int *ke = (int *)(kc + 1);                           // exponent, pick
  const int lane = threadIdx.x;
  int nitems = items;
  if (item_count) nitems = min(items, *item_count);
  constexpr int kXcd = 8;
  const int G = gridDim.x, p = blockIdx.x;
  const int x = p % kXcd, bi = p / kXcd;
  const int nblk_x = G / kXcd + (x < G % kXcd ? 1 : 0);
  const int64_t total = (int64_t)nitems * wpi;
  const int64_t per_x = (total + kXcd - 1) / kXcd;
  const int64_t lo = (int64_t)x * per_x;
  const int64_t hi = min(total, lo + per_x);
  for (int64_t task = lo + bi; task < hi; task += nblk_x) {
    const int slot = (int)(task / wpi);
    const int w = (int)(task - (int64_t)slot * wpi);
    const int item = item_map ? item_map[slot] : slot;
    const cd *src = Vt + (size_t)item * item_stride;
    __syncthreads();  // previous task's reads of LDS are done
    for (int e = lane; e < K * R; e += 64) {
      const int nu = e / R, j = e - nu * R;
      Vs[e] = src[(size_t)nu * ldv + j];
    }
    __syncthreads();
    if (lane < K) {
      double mn = INFINITY, mx = 0.0;
      for (int j = 0; j < R; ++j) {
        const cd v = Vs[lane * R + j];
        const double a = fma(v.re, v.re, v.im * v.im);
        mn = fmin(mn, a);
        mx = fmax(mx, a);
      }
      pmin[lane] = mn;
      pmax[lane] = mx;
    }
    __syncthreads();
    if (lane == 0) {
      int best = 0;
      double mx = 0.0;
      for (int nu = 0; nu < K; ++nu) {
        if (pmin[nu] > pmin[best]) best = nu;
        mx = fmax(mx, pmax[nu]);
      }
      const bool ok = mx > 0.0 && mx <= 1.7976931348623157e308 && pmin[best] >= kPairMinRel2 * mx;
      ke[1] = ok ? best : -1;
    }
    __syncthreads();
    const int pick = ke[1];
    cd acc;
    int dst = lane;  // minor slot this lane's result belongs to
    if (pick < 0) {
      acc = sweep_wave<K, U, 3>(Vs, Ed, lane, c, w, wpi, iters);
    } else {
      if (pick != 1 && lane < R) {
        const cd t = Vs[R + lane];
        Vs[R + lane] = Vs[pick * R + lane];
        Vs[pick * R + lane] = t;
      }
      __syncthreads();
      if (lane < R) inv[lane] = cinv_scaled(Vs[R + lane]);
      if (lane == 63) {
        cd m;
        int e2;
        prod_half_scaled(Vs + R, R, m, e2);
        kc[0] = m;
        ke[0] = e2;
      }
      __syncthreads();
      for (int e = lane; e < K * R; e += 64) {
        const int nu = e / R, j = e - nu * R;
        if (nu != 1) {
          const cd v = cmul(Vs[e], inv[j]);
          Vs[e] = {2.0 * v.re, 2.0 * v.im};
        }
      }
      __syncthreads();
      const cd a = sweep_pair_wave<K, U>(Vs, Ed, lane, c, w, wpi, iters);
      const cd m = kc[0];
      const int e2 = ke[0];
      acc = {ldexp(m.re * a.re - m.im * a.im, e2), ldexp(m.re * a.im + m.im * a.re, e2)};
      if (lane == 1) dst = pick;
      else if (lane == pick) dst = 1;
    }
    if (lane < K) partial[((size_t)slot * wpi + w) * K + dst] = acc;
  }
}

// minors[slot*K + nu] = sum_w partial[(slot*wpi + w)*K + nu], fixed order.
__global__ void sweep_reduce_kernel(const cd *__restrict__ partial, int K, int wpi, int items,
                                    const int *__restrict__ item_count, cd *__restrict__ minors) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= items * K) return;
  const int b = idx / K, nu = idx - b * K;
  if (item_count && b >= *item_count) return;
  double re = 0.0, im = 0.0;
#pragma unroll 8
  for (int w = 0; w < wpi; ++w) {  // unrolled: the loads are in flight together
    const cd v = partial[((size_t)b * wpi + w) * K + nu];
    re += v.re;
    im += v.im;
  }
  minors[idx] = {re, im};
}

}  // namespace tb

// ------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------
namespace tbs {

struct SweepPlan {
  int c, wpi, iters;
};

#ifndef TB_SWEEP_BUILTIN_MAX_C
#define TB_SWEEP_BUILTIN_MAX_C 11
#endif
// Shortest built-in walk per lane between re-seeds (bits): below it the O(k^2) seed and the
// per-chunk wave sums of the k minors outweigh the walk itself.
#ifndef TB_SWEEP_BUILTIN_MIN_C
#define TB_SWEEP_BUILTIN_MIN_C 5
#endif
constexpr int64_t kSweepMaxBlocks = 8192;  // 4 one-wave blocks per SIMD of 256 CUs x 4 SIMDs x 2
// Waves a launch should provide before the built-in plan stops shortening the chunks: 4
// per SIMD (one-wave blocks, 2 resident per SIMD, so two rounds and a short tail).
constexpr int64_t kSweepWaveTarget = 4096;

// Built-in chunk: the longest walk (at most 2^11 Gray steps per lane) that still gives the
// launch kSweepWaveTarget waves, but not below 2^5 steps. A batch of 4096 k = 20 items
// keeps 2^11 (4 waves per item); one k = 20 item gets 2^5 (256 waves) instead of a
// 4-wave launch. The tuner may choose otherwise per (k, batch bucket).
inline SweepPlan sweep_plan(int k, int64_t items, int c_override) {
  SweepPlan p{};
  const int nb = k - 1;                      // Gray bits (particle 0 is fixed)
  const int u = std::min(tb::kSweepU, nb);   // static bits of the kernel for this k
  int c;
  if (c_override > 0) {
    c = c_override;
  } else {
    int want_bits = 0;  // log2(waves per item that bring the launch to the target)
    while (std::max<int64_t>(items, 1) << want_bits < kSweepWaveTarget && want_bits < 30)
      ++want_bits;
    c = std::min(TB_SWEEP_BUILTIN_MAX_C, nb - 6 - want_bits);
    c = std::max(c, std::min(TB_SWEEP_BUILTIN_MIN_C, nb - 6));
  }
  c = std::max(c, u);
  c = std::min(c, std::min(nb, tb::kSweepMaxC));
  p.c = c;
  const int chunk_bits = nb - c;                      // log2(chunks per item)
  const int wave_bits = std::max(0, chunk_bits - 6);  // log2(waves at one chunk per lane)
  const int wpi_bits = std::min(wave_bits, 13);
  p.wpi = 1 << wpi_bits;
  p.iters = 1 << (wave_bits - wpi_bits);
  return p;
}

// LDS per wave: the item's k(k-1) columns plus 64 lanes x (c - U) corrector slots; the
// paired kernel adds the quotients' reciprocals, the particles' extreme magnitudes and the
// column product.
inline size_t sweep_lds_bytes(int k, int c) {
  const int u = std::min(tb::kSweepU, k - 1);
  return sizeof(tb::cd) * ((size_t)k * (k - 1) + (size_t)std::max(c - u, 0) * 64);
}
inline size_t sweep_pair_lds_bytes(int k, int c) {
  return sweep_lds_bytes(k, c) + sizeof(tb::cd) * (size_t)(k - 1) +
         sizeof(double) * (size_t)(2 * k) + sizeof(tb::cd) + 2 * sizeof(int);
}
// Sizes that run the paired sweep (with a chunk holding the static block plus the pair
// bit), at 2 waves per SIMD: up to k = 21 with five particles in broadcast form, from
// k = 22 with the three static ones (196..249 VGPRs, no spills). Five broadcast particles
// spill from k = 22, and at one wave per SIMD the compiler parks broadcast columns in
// AGPRs and reloads them right before their DPP reads, a hazard that
// tests/test_dpp_hazards.py looks for. k = 31 is the largest broadcast-column sweep.
#ifndef TB_SWEEP_PAIR_MAX_K
#define TB_SWEEP_PAIR_MAX_K 31
#endif
template <int K>
constexpr bool sweep_paired() {
  return K >= tb::kSweepPairMinK && K <= TB_SWEEP_PAIR_MAX_K;
}

template <int K>
int32_t launch_sweep_k(const tb::cd *Vt, int ldv, int64_t item_stride, const int *item_map,
                       const int *item_count, int items, const SweepPlan &p, tb::cd *partial,
                       hipStream_t s) {
  // Persistent grid: enough one-wave blocks to fill every SIMD a few times over.
  const int64_t tasks = (int64_t)items * p.wpi;
  const unsigned grid = (unsigned)std::min<int64_t>(tasks, tbs::kSweepMaxBlocks);
  constexpr int R0 = (K <= TB_SWEEP_BCAST_MAX_K) ? 3 : 2;
  constexpr int U = (K - 1 < tb::kSweepU) ? K - 1 : tb::kSweepU;
  if constexpr (sweep_paired<K>()) {
    if (p.c >= U + 1) {
      hipLaunchKernelGGL((tb::sweep_pair_kernel<K, U>), dim3(grid), dim3(64),
                         sweep_pair_lds_bytes(K, p.c), s, Vt, ldv, item_stride, item_map,
                         item_count, items, p.c, p.wpi, p.iters, partial);
      return tbh::check_hip(hipGetLastError(), "sweep_pair_kernel");
    }
  }
  hipLaunchKernelGGL((tb::sweep_kernel<K, U, R0>), dim3(grid), dim3(64), sweep_lds_bytes(K, p.c), s,
                     Vt, ldv, item_stride, item_map, item_count, items, p.c, p.wpi, p.iters,
                     partial);
  return tbh::check_hip(hipGetLastError(), "sweep_kernel");
}

template <int K = 2>
int32_t dispatch_sweep(int k, const tb::cd *Vt, int ldv, int64_t item_stride, const int *item_map,
                       const int *item_count, int items, const SweepPlan &p, tb::cd *partial,
                       hipStream_t s) {
  if constexpr (K > TB_SWEEP_MAX_K) {
    return tbh::fail(TB_ERANGE, "sweep: k=%d exceeds TB_SWEEP_MAX_K=%d", k, TB_SWEEP_MAX_K);
  } else {
    if (k == K) return launch_sweep_k<K>(Vt, ldv, item_stride, item_map, item_count, items, p,
                                         partial, s);
    return dispatch_sweep<K + 1>(k, Vt, ldv, item_stride, item_map, item_count, items, p, partial,
                                 s);
  }
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t sweep_partial_bytes(int k, int items, int c_override) {
  if (k < 2 || items <= 0) return 0;
  const SweepPlan p = sweep_plan(k, items, c_override);
  return align256((size_t)items * p.wpi * k * sizeof(tb::cd));
}

void sweep_plan_of(int k, int items, int c_override, int *c, int *wpi, int *iters) {
  const SweepPlan p = sweep_plan(k, items, c_override);
  *c = p.c;
  *wpi = p.wpi;
  *iters = p.iters;
}

// Partial sums only: partial[(item * wpi + w) * k + nu], w < *wpi. A consumer that sums
// them in w order gets the same minors as run_sweep.
int32_t run_sweep_partial(int k, const tb::cd *Vt, int ldv, int64_t item_stride,
                          const int *item_map, const int *item_count, int items, int c_override,
                          tb::cd *partial, int *wpi, hipStream_t s) {
  const SweepPlan p = sweep_plan(k, items, c_override);
  *wpi = p.wpi;
  return dispatch_sweep(k, Vt, ldv, item_stride, item_map, item_count, items, p, partial, s);
}

// Full sweep: partial sums then the fixed-order reduction into minors[items][k].
int32_t run_sweep(int k, const tb::cd *Vt, int ldv, int64_t item_stride, const int *item_map,
                  const int *item_count, int items, int c_override, tb::cd *partial,
                  tb::cd *minors, hipStream_t s) {
  const SweepPlan p = sweep_plan(k, items, c_override);
  int32_t rc = dispatch_sweep(k, Vt, ldv, item_stride, item_map, item_count, items, p, partial, s);
  if (rc) return rc;
  const int total = items * k;
  hipLaunchKernelGGL(tb::sweep_reduce_kernel, dim3((total + 255) / 256), dim3(256), 0, s, partial,
                     k, p.wpi, items, item_count, minors);
  return tbh::check_hip(hipGetLastError(), "sweep_reduce_kernel");
}

}  // namespace tbs

extern "C" size_t tb_sweep_workspace_size(int32_t k, int32_t batch) {
  if (k > TB_SWEEP_MAX_K) return 0;
  return tbs::sweep_partial_bytes(k, batch, tbh::tuned(TB_TUNE_SWEEP, k, batch));
}

extern "C" int32_t tb_sweep_minors_batch(const tb_complex128 *Vt, int32_t k, int32_t ldv,
                                         int64_t item_stride, int32_t batch,
                                         tb_complex128 *minors, void *workspace,
                                         size_t workspace_bytes, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (batch < 0) return tbh::fail(TB_EINVAL, "sweep: batch=%d < 0", batch);
  if (batch == 0) return TB_OK;
  if (k < 2) return tbh::fail(TB_EINVAL, "sweep: k=%d < 2 (k == 1 has no sweep)", k);
  if (k > TB_SWEEP_MAX_K)
    return tbh::fail(TB_ERANGE, "sweep: k=%d exceeds TB_SWEEP_MAX_K=%d", k, TB_SWEEP_MAX_K);
  if (!Vt || !minors) return tbh::fail(TB_EINVAL, "sweep: NULL pointer");
  if (ldv < k - 1) return tbh::fail(TB_EINVAL, "sweep: ldv=%d < k-1=%d", ldv, k - 1);
  if (item_stride < (int64_t)k * ldv)
    return tbh::fail(TB_EINVAL, "sweep: item_stride=%lld < k*ldv", (long long)item_stride);
  const int c = tbh::tuned(TB_TUNE_SWEEP, k, batch);  // read once: size check and launch agree
  const size_t need = tbs::sweep_partial_bytes(k, batch, c);
  if (!workspace || workspace_bytes < need)
    return tbh::fail(TB_EWORKSPACE, "sweep: workspace %zu bytes < required %zu", workspace_bytes,
                     need);
  return tbs::run_sweep(k, (const tb::cd *)Vt, ldv, item_stride, nullptr, nullptr, batch, c,
                        (tb::cd *)workspace, (tb::cd *)minors, s);
}
