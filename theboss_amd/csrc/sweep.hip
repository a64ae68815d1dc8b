// Binary Clifford & Clifford sweep: all k particle-deleted minors in one Gray walk.
//
// Restates BSCCRyserSubmatricesPermanentCalculator.compute_permanents
// (theboss/boson_sampling_utilities/permanent_calculators/bs_cc_ryser_submatrices_permanent_calculator.py:73-119,
//  bs_submatrices_permanent_calculator_base.py:150-189) for collision-free inputs:
//   minor_nu = (-1)^R sum_{S subset of particles, nu not in S} (-1)^|S| prod_j (sum_{mu in S} V[mu][j])
// over k input particles (Gray bits) and R = k-1 output rows (repeated for output
// multiplicities). With k particles of multiplicity one, every binomial factor of the
// reference's Guan walk is 1 and its Guan code is the reflected binary Gray code.
//
// Work split (same skeleton as the Glynn kernel): aligned chunks of 2^c Gray ranks, one
// per lane; all lanes of a wave belong to one item, so the flipped particle at local
// step l is ctz(l) for the whole wave and its column is a broadcast LDS read.
//
// Minors without a per-step O(k) update: keep the running chunk sum T = sum (-1)^|S| P(S)
// and, per in-chunk bit i, a corrector E_i updated only when bit i flips
// (E_i += 2*sigma_old*T, sigma = +1 while particle i is absent). At chunk end
//   D_i = E_i + sigma_i T,  sum over "i absent" = (T + D_i) / 2.
// Bits above the chunk are constant within it: the chunk adds T to minor_i iff particle i
// is absent. Per Gray step the cost is R complex adds, R-1 complex multiplies and O(1)
// for the minors.
#include <algorithm>

#include "tb_common.hpp"

#ifndef TB_SWEEP_ATTR
#define TB_SWEEP_ATTR __attribute__((amdgpu_waves_per_eu(2)))
#endif
#ifndef TB_STEP_FENCE
#define TB_STEP_FENCE() __builtin_amdgcn_sched_barrier(0)
#endif
#ifndef TB_SWEEP_R0
#define TB_SWEEP_R0 0
#endif

namespace tb {

constexpr int kSweepU = 2;      // Gray bits unrolled per block (static flip rows)
constexpr int kSweepMaxC = 16;  // largest chunk (bits); sizes the LDS corrector store

// h += sg * column (sg = +-1 per lane, or a compile-time sign).
template <int R, int MODE>
__device__ __forceinline__ void sweep_flip(double (&hr)[R], double (&hi)[R], const cd *col,
                                           double sg) {
  constexpr int G = 4;
  constexpr int NG = (R + G - 1) / G;
  cd cur[G], nxt[G];
#pragma unroll
  for (int e = 0; e < G; ++e)
    if (e < R) cur[e] = col[e];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    if (g + 1 < NG) {
#pragma unroll
      for (int e = 0; e < G; ++e)
        if ((g + 1) * G + e < R) nxt[e] = col[(g + 1) * G + e];
    }
#pragma unroll
    for (int e = 0; e < G; ++e) {
      const int j = g * G + e;
      if (j < R) {
        if constexpr (MODE == 1) {
          hr[j] += cur[e].re;
          hi[j] += cur[e].im;
        } else if constexpr (MODE == -1) {
          hr[j] -= cur[e].re;
          hi[j] -= cur[e].im;
        } else {
          hr[j] = fma(sg, cur[e].re, hr[j]);
          hi[j] = fma(sg, cur[e].im, hi[j]);
        }
      }
    }
    TB_STEP_FENCE();
#pragma unroll
    for (int e = 0; e < G; ++e) cur[e] = nxt[e];
  }
}

template <int R>
__device__ __forceinline__ cd sweep_prod(const double (&hr)[R], const double (&hi)[R]) {
  return TreeProd<0, R>::run(hr, hi);
}

// One chunk of 2^c Gray ranks starting at t0 = q << c. Vs: LDS columns [k][R].
// Ed: this lane's LDS corrector slots for bits U..c-1 (stride 64 cd).
// Returns the chunk's contribution to all K minors in mi[] (before the (-1)^R factor).
template <int K, int U>
__device__ __forceinline__ void sweep_chunk(const cd *Vs, cd *Ed, int c, uint64_t q, cd (&mi)[K]) {
  constexpr int R = K - 1;
  const uint64_t t0 = q << c;
  const uint64_t g0 = t0 ^ (t0 >> 1);
  double hr[R], hi[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    hr[j] = 0.0;
    hi[j] = 0.0;
  }
#pragma unroll 1
  for (int i = 0; i < K; ++i) {
    const double b = (double)((g0 >> i) & 1ull);
    const cd *col = Vs + i * R;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      hr[j] = fma(b, col[j].re, hr[j]);
      hi[j] = fma(b, col[j].im, hi[j]);
    }
  }
  double Tr, Ti;
  {
    const cd p = sweep_prod<R>(hr, hi);
    const double s0 = (t0 & 1ull) ? -1.0 : 1.0;
    Tr = s0 * p.re;
    Ti = s0 * p.im;
  }
  double Esr[U > 0 ? U : 1], Esi[U > 0 ? U : 1];
#pragma unroll
  for (int i = 0; i < U; ++i) {
    Esr[i] = 0.0;
    Esi[i] = 0.0;
  }
  for (int i = U; i < c; ++i) Ed[(i - U) * 64] = {0.0, 0.0};
  // t_c: the Gray bit above the chunk is bit 0 of q (0 when the chunk is everything).
  const bool q_odd = (q & 1ull) != 0;
  const int nblocks = 1 << (c - U);
  for (int blk = 0; blk < nblocks; ++blk) {
    if (blk > 0) {
      const int i = U + __builtin_ctz(blk);
      const bool t_next = (i + 1 < c) ? (((blk >> (i + 1 - U)) & 1) != 0) : q_odd;
      const double sg = t_next ? -1.0 : 1.0;  // particle i joins when t_{i+1} == 0
      cd e = Ed[(i - U) * 64];
      e.re = fma(2.0 * sg, Tr, e.re);
      e.im = fma(2.0 * sg, Ti, e.im);
      Ed[(i - U) * 64] = e;
      sweep_flip<R, 0>(hr, hi, Vs + i * R, sg);
      const cd p = sweep_prod<R>(hr, hi);
      Tr += p.re;
      Ti += p.im;
      TB_STEP_FENCE();
    }
#pragma unroll
    for (int j = 1; j < (1 << U); ++j) {
      const int i = __builtin_ctz(j);
      if (i + 1 < U) {
        if ((j >> (i + 1)) & 1) {  // leaves
          Esr[i] -= 2.0 * Tr;
          Esi[i] -= 2.0 * Ti;
          sweep_flip<R, -1>(hr, hi, Vs + i * R, -1.0);
        } else {  // joins
          Esr[i] += 2.0 * Tr;
          Esi[i] += 2.0 * Ti;
          sweep_flip<R, 1>(hr, hi, Vs + i * R, 1.0);
        }
      } else {
        const bool t_next = (U < c) ? ((blk & 1) != 0) : q_odd;
        const double sg = t_next ? -1.0 : 1.0;
        Esr[i] = fma(2.0 * sg, Tr, Esr[i]);
        Esi[i] = fma(2.0 * sg, Ti, Esi[i]);
        sweep_flip<R, 0>(hr, hi, Vs + i * R, sg);
      }
      const cd p = sweep_prod<R>(hr, hi);
      if (j & 1) {
        Tr -= p.re;
        Ti -= p.im;
      } else {
        Tr += p.re;
        Ti += p.im;
      }
      TB_STEP_FENCE();
    }
  }
  const uint64_t tend = t0 + ((uint64_t)1 << c) - 1;
  const uint64_t gend = tend ^ (tend >> 1);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const bool absent = ((gend >> i) & 1ull) == 0;
    if (i < c) {
      double er, ei;
      if (i < U) {
        er = Esr[i < U ? i : 0];
        ei = Esi[i < U ? i : 0];
      } else {
        const cd e = Ed[(i - U) * 64];
        er = e.re;
        ei = e.im;
      }
      const double sig = absent ? 1.0 : -1.0;
      mi[i] = {0.5 * (Tr + er + sig * Tr), 0.5 * (Ti + ei + sig * Ti)};
    } else {
      mi[i] = absent ? cd{Tr, Ti} : cd{0.0, 0.0};
    }
  }
}

// One wave per (item, wave-slot); 64-thread workgroups. Item b's particle columns are
// staged into LDS. Lane q-chunks: q = (w + it*wpi)*64 + lane, it < iters.
// partial[(slot*wpi + w)*K + nu] receives the wave's share of minor nu; item_map
// (optional, device) maps slot -> item whose columns are read.
// item_count (optional, device): only slots < *item_count are processed, so a captured
// graph can run a data-dependent number of items.
template <int K, int U>
__global__ __launch_bounds__(64) TB_SWEEP_ATTR void sweep_kernel(
    const cd *__restrict__ Vt, int ldv, int64_t item_stride, const int *__restrict__ item_map,
    const int *__restrict__ item_count, int items, int c, int wpi, int iters,
    cd *__restrict__ partial) {
  constexpr int R = K - 1;
  extern __shared__ cd smem[];
  cd *Vs = smem;                 // K*R
  cd *Ed = smem + K * R;         // (kSweepMaxC - U) * 64
  const int lane = threadIdx.x;
  const int gw = blockIdx.x;
  const int slot = gw / wpi;
  const int w = gw - slot * wpi;
  int nitems = items;
  if (item_count) nitems = min(items, *item_count);
  if (slot >= nitems) return;
  const int item = item_map ? item_map[slot] : slot;
  const cd *src = Vt + (size_t)item * item_stride;
  for (int e = lane; e < K * R; e += 64) {
    const int nu = e / R, j = e - nu * R;
    Vs[e] = src[(size_t)nu * ldv + j];
  }
  __syncthreads();
  const uint64_t nchunks = (uint64_t)1 << (K - c);
  cd acc = {0.0, 0.0};
  for (int it = 0; it < iters; ++it) {
    const uint64_t q = ((uint64_t)w + (uint64_t)it * wpi) * 64 + lane;
    cd mi[K];
    if (q < nchunks) {
      sweep_chunk<K, U>(Vs, Ed + lane, c, q, mi);
    } else {
#pragma unroll
      for (int i = 0; i < K; ++i) mi[i] = {0.0, 0.0};
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const double re = wave_sum(mi[i].re);
      const double im = wave_sum(mi[i].im);
      if (lane == i) {
        acc.re += re;
        acc.im += im;
      }
    }
  }
  if (lane < K) partial[((size_t)slot * wpi + w) * K + lane] = acc;
}

// minors[slot*K + nu] = (-1)^R * sum_w partial[(slot*wpi + w)*K + nu], fixed order.
__global__ void sweep_reduce_kernel(const cd *__restrict__ partial, int K, int wpi, int items,
                                    const int *__restrict__ item_count, cd *__restrict__ minors) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= items * K) return;
  const int b = idx / K, nu = idx - b * K;
  if (item_count && b >= *item_count) return;
  double re = 0.0, im = 0.0;
  for (int w = 0; w < wpi; ++w) {
    const cd v = partial[((size_t)b * wpi + w) * K + nu];
    re += v.re;
    im += v.im;
  }
  const double sgn = ((K - 1) & 1) ? -1.0 : 1.0;
  minors[idx] = {sgn * re, sgn * im};
}

}  // namespace tb

// ------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------
namespace tbs {

struct SweepPlan {
  int c, wpi, iters;
};

// Default chunk: 2^min(10, k-6) ranks per lane (>= 4), so one wave holds 64 chunks and a
// lane walks at most 1024 Gray steps between re-seeds.
inline SweepPlan sweep_plan(int k, int c_override) {
  SweepPlan p{};
  int c = c_override > 0 ? c_override : std::min(10, k - 6);
  c = std::max(c, tb::kSweepU);
  c = std::min(c, std::min(k, tb::kSweepMaxC));
  p.c = c;
  const int chunk_bits = k - c;                       // log2(chunks per item)
  const int wave_bits = std::max(0, chunk_bits - 6);  // log2(waves at one chunk per lane)
  const int wpi_bits = std::min(wave_bits, 13);
  p.wpi = 1 << wpi_bits;
  p.iters = 1 << (wave_bits - wpi_bits);
  return p;
}

inline size_t sweep_lds_bytes(int k) {
  return sizeof(tb::cd) * ((size_t)k * (k - 1) + (size_t)(tb::kSweepMaxC - tb::kSweepU) * 64);
}

template <int K>
int32_t launch_sweep_k(const tb::cd *Vt, int ldv, int64_t item_stride, const int *item_map,
                       const int *item_count, int items, const SweepPlan &p, tb::cd *partial,
                       hipStream_t s) {
  const unsigned grid = (unsigned)items * (unsigned)p.wpi;
  hipLaunchKernelGGL((tb::sweep_kernel<K, tb::kSweepU>), dim3(grid), dim3(64), sweep_lds_bytes(K), s,
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
  const SweepPlan p = sweep_plan(k, c_override);
  return align256((size_t)items * p.wpi * k * sizeof(tb::cd));
}

// Full sweep: partial sums then the fixed-order reduction into minors[items][k].
int32_t run_sweep(int k, const tb::cd *Vt, int ldv, int64_t item_stride, const int *item_map,
                  const int *item_count, int items, int c_override, tb::cd *partial,
                  tb::cd *minors, hipStream_t s) {
  const SweepPlan p = sweep_plan(k, c_override);
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
  return tbs::sweep_partial_bytes(k, batch, 0);
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
  const size_t need = tb_sweep_workspace_size(k, batch);
  if (!workspace || workspace_bytes < need)
    return tbh::fail(TB_EWORKSPACE, "sweep: workspace %zu bytes < required %zu", workspace_bytes,
                     need);
  return tbs::run_sweep(k, (const tb::cd *)Vt, ldv, item_stride, nullptr, nullptr, batch, 0,
                        (tb::cd *)workspace, (tb::cd *)minors, s);
}
