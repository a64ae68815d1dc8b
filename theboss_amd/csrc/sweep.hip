// Clifford & Clifford submatrices sweep: all k particle-deleted minors in one Gray walk.
//
// Computes what BSCCRyserSubmatricesPermanentCalculator.compute_permanents computes
// (theboss/boson_sampling_utilities/permanent_calculators/bs_cc_ryser_submatrices_permanent_calculator.py:73-119,
//  bs_submatrices_permanent_calculator_base.py:150-189) for collision-free inputs: for k
// input particles (columns V[nu][.]) and R = k-1 output rows (repeated for output
// multiplicities), minor_nu = perm(V without particle nu).
//
// Formula (Glynn-type instead of the reference's Ryser walk): with half sums
// h_j(delta) = (1/2) sum_mu delta_mu V[mu][j] and delta_0 = +1,
//   minor_nu = sum_{delta in {+-1}^k, delta_0 = +1} (prod_mu delta_mu) delta_nu prod_j h_j(delta).
// Expanding prod_j over row->particle maps, the delta-sum keeps exactly the maps that are
// bijections onto the particles other than nu, i.e. perm(V without nu); fixing delta_0
// uses the delta -> -delta symmetry of the summand. Against the reference's Ryser walk
// this halves the Gray steps (2^(k-1) instead of 2^k) at the same per-step cost, and the
// terms cancel far less (all partial sums have the full particle count), which keeps
// k = 20..40 sweeps well inside 1e-10 relative error.
//
// Work split (same skeleton as the Glynn kernel): aligned chunks of 2^c Gray ranks, one
// per lane; all lanes of a wave belong to one item, so the flipped particle at local
// step l is 1 + ctz(l) for the whole wave: its column is wave-uniform, held in
// broadcast-form VGPRs for the most frequent particles and read from LDS otherwise.
//
// All k minors without a per-step O(k) update: keep the running chunk sum
// T = sum (prod delta) prod h and, per in-chunk bit b, a corrector E_b updated only when
// bit b flips (E_b += 2 delta_old T). At chunk end the delta_{b+1}-weighted sum is
// D_b = E_b + delta_b T. Bits above the chunk are constant in it: D_b = delta_b T.
// Per Gray step: R complex adds, R-1 complex multiplies and O(1) for the minors.
#include <algorithm>

#include "flip_common.hpp"

#ifndef TB_SWEEP_ATTR
#define TB_SWEEP_ATTR(K) __attribute__((amdgpu_waves_per_eu((K) <= 26 ? 2 : 1)))
#endif
#ifndef TB_STEP_FENCE
#define TB_STEP_FENCE() __builtin_amdgcn_sched_barrier(0)
#endif
// Where the flipped columns come from (R0 mode of the kernel):
//  3 (k <= TB_SWEEP_BCAST_MAX_K): particles 1..5 in broadcast form in VGPRs (every static
//    flip and 3/4 of the block-start flips, no LDS read); the rest from LDS;
//  2 (larger k, where 5 broadcast columns no longer fit 2 waves per SIMD): the first
//    TB_SWEEP_SROW values of particle 1's column in SGPRs (scalar operands of the adds),
//    everything else from LDS.
// LDS flips load their first column group one step ahead (issued by a step that reads no
// LDS, under its product).
#ifndef TB_SWEEP_BCAST_MAX_K
#define TB_SWEEP_BCAST_MAX_K 31
#endif
#ifndef TB_SWEEP_SROW
#define TB_SWEEP_SROW 20
#endif


namespace tb {

#ifndef TB_SWEEP_U
#define TB_SWEEP_U 3
#endif
constexpr int kSweepU = TB_SWEEP_U;  // Gray bits unrolled per block (static flip rows)
constexpr int kSweepMaxC = 16;  // largest chunk (bits); sizes the LDS corrector store

// h += sg * column (sg = +-1 per lane, or a compile-time sign).
constexpr int kSweepG = 4;  // column values per LDS group

template <int R, int MODE, bool PRE = false>
__device__ __forceinline__ void sweep_flip(double (&hr)[R], double (&hi)[R], const cd *col,
                                           double sg, const cd *pre = nullptr) {
  constexpr int G = kSweepG;
  constexpr int NG = (R + G - 1) / G;
  cd cur[G], nxt[G];
#pragma unroll
  for (int e = 0; e < G; ++e)
    if (e < R) cur[e] = PRE ? pre[e] : col[e];
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

// One chunk of 2^c Gray ranks starting at t0 = q << c. Vs: LDS columns [k][R] (particle
// major).
// Ed: this lane's LDS corrector slots for bits U..c-1 (stride 64 cd).
// Returns the chunk's contribution to all K minors in mi[].
// Register-held column data of one item (R0 mode above).
template <int R, int R0>
struct Col0 {
  static constexpr int S = (R0 == 2) ? (TB_SWEEP_SROW < R ? TB_SWEEP_SROW : R) : 0;
  static constexpr int NU = (R0 == 3) ? (kSweepU + 2 < R ? kSweepU + 2 : R) : 1;
  SRow<S, double> sr;
  Bcast<R> b[NU];  // particles 1..NU
  double one;      // 1.0 in a VGPR (second operand of the fixed-direction flips)
};

template <int R, int MODE, int R0, bool PRE = false>
__device__ __forceinline__ void sweep_flip_col(double (&hr)[R], double (&hi)[R], const cd *Vs,
                                               const Col0<R, R0> &c0, int i, double sg,
                                               const cd *pre = nullptr) {
  if constexpr (R0 == 3) {
    if (i >= 1 && i <= Col0<R, R0>::NU) {  // compile-time at every static flip site
#pragma unroll
      for (int u = 0; u < Col0<R, R0>::NU; ++u)
        if (u == i - 1) flip_bcast<0, R, MODE>(hr, hi, c0.b[u], MODE == 0 ? sg : c0.one);
      return;
    }
  }
  if constexpr (R0 == 2) {
    if (i == 1) {  // particle 1 (compile-time at every static flip site)
      flip_srow<R, Col0<R, R0>::S, MODE, double>(hr, hi, c0.sr,
                                                 (const cpx<double> *)(Vs + R), sg);
      return;
    }
  }
  sweep_flip<R, MODE, PRE>(hr, hi, Vs + i * R, sg, pre);
}

template <int K, int U, int R0>
__device__ __forceinline__ void sweep_chunk(const cd *Vs, const Col0<K - 1, R0> &c1, cd *Ed, int c,
                                            uint64_t q, cd (&mi)[K]) {
  constexpr int R = K - 1;   // output rows
  constexpr int NB = K - 1;  // Gray bits: bit b <-> particle b + 1 (particle 0 fixed at +1)
  const uint64_t t0 = q << c;
  const uint64_t g0 = t0 ^ (t0 >> 1);
  // Half sums h_j = (1/2) sum_nu delta_nu V[nu][j], delta_0 = +1, delta_{b+1} = (-1)^{g_b}.
  double hr[R], hi[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    hr[j] = 0.5 * Vs[j].re;
    hi[j] = 0.5 * Vs[j].im;
  }
#pragma unroll 1
  for (int b = 0; b < NB; ++b) {
    const double d = ((g0 >> b) & 1ull) ? -0.5 : 0.5;
    const cd *col = Vs + (b + 1) * R;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      hr[j] = fma(d, col[j].re, hr[j]);
      hi[j] = fma(d, col[j].im, hi[j]);
    }
  }
  double Tr, Ti;
  {
    const cd p = sweep_prod<R>(hr, hi);
    const double s0 = (t0 & 1ull) ? -1.0 : 1.0;  // prod(delta) = (-1)^popcount(gray(t)) = (-1)^t
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
  // The Gray bit above the chunk is bit 0 of q (0 when the chunk is the whole walk).
  const bool q_odd = (q & 1ull) != 0;
  const int nblocks = 1 << (c - U);
  // Flip of bit i at local step l: the new bit is 1 ^ t_{i+1}. t_{i+1} == 0: delta goes
  // +1 -> -1 (h -= V, corrector += 2T); t_{i+1} == 1: delta goes -1 -> +1 (h += V,
  // corrector -= 2T).
  // Particle 1 (odd local steps) is not read from LDS, so an odd step issues the LDS loads
  // of the next step's first column group; they complete under this step's product.
  constexpr bool kPre = true;
  constexpr bool kStaticLds = R0 != 3;  // static flips of particles >= 2 read LDS
  cd pre[kSweepG];
  auto prefetch = [&](const cd *col) {
#pragma unroll
    for (int e = 0; e < kSweepG; ++e)
      if (e < R) pre[e] = col[e];
  };
  auto static_steps = [&](int blk) {
#pragma unroll
    for (int j = 1; j < (1 << U); ++j) {
      const int i = __builtin_ctz(j);
      if (i + 1 < U) {
        if ((j >> (i + 1)) & 1) {
          Esr[i] -= 2.0 * Tr;
          Esi[i] -= 2.0 * Ti;
          sweep_flip_col<R, 1, R0, kPre>(hr, hi, Vs, c1, i + 1, 1.0, pre);
        } else {
          Esr[i] += 2.0 * Tr;
          Esi[i] += 2.0 * Ti;
          sweep_flip_col<R, -1, R0, kPre>(hr, hi, Vs, c1, i + 1, -1.0, pre);
        }
      } else {
        const bool t_next = (U < c) ? ((blk & 1) != 0) : q_odd;
        const double sg = t_next ? 1.0 : -1.0;
        Esr[i] = fma(-2.0 * sg, Tr, Esr[i]);
        Esi[i] = fma(-2.0 * sg, Ti, Esi[i]);
        sweep_flip_col<R, 0, R0, kPre>(hr, hi, Vs, c1, i + 1, sg, pre);
      }
      if constexpr (kPre) {
        if ((j & 1) && (kStaticLds || j + 1 == (1 << U))) {
          if (j + 1 < (1 << U)) {
            prefetch(Vs + (__builtin_ctz(j + 1) + 1) * R);
          } else {  // the next block's first flip (particle 1's column past the last block)
            const int inext = (blk + 1 < nblocks) ? U + __builtin_ctz(blk + 1) : 0;
            prefetch(Vs + (inext + 1) * R);
          }
        }
      }
      if (j & 1)
        tree_prod_acc<-1, 0, R>(Tr, Ti, hr, hi);
      else
        tree_prod_acc<1, 0, R>(Tr, Ti, hr, hi);
      TB_STEP_FENCE();
    }
  };
  // Block 0 is peeled so the loop body below is branch-free (no phi copies of h).
  static_steps(0);
  for (int blk = 1; blk < nblocks; ++blk) {
    const int i = U + __builtin_ctz(blk);
    const bool t_next = (i + 1 < c) ? (((blk >> (i + 1 - U)) & 1) != 0) : q_odd;
    const double sg = t_next ? 1.0 : -1.0;
    cd e = Ed[(i - U) * 64];
    e.re = fma(-2.0 * sg, Tr, e.re);
    e.im = fma(-2.0 * sg, Ti, e.im);
    Ed[(i - U) * 64] = e;
    if constexpr (R0 == 3) {
      // particles U+1 .. NU (3/4 of the block starts) from the broadcast store
      bool done = false;
#pragma unroll
      for (int u = U; u < Col0<R, R0>::NU; ++u) {
        if (i + 1 == u + 1) {
          flip_bcast<0, R, 0>(hr, hi, c1.b[u], sg);
          done = true;
        }
      }
      if (!done) sweep_flip<R, 0, kPre>(hr, hi, Vs + (i + 1) * R, sg, pre);
    } else {
      sweep_flip<R, 0, kPre>(hr, hi, Vs + (i + 1) * R, sg, pre);
    }
    tree_prod_acc<1, 0, R>(Tr, Ti, hr, hi);
    TB_STEP_FENCE();
    static_steps(blk);
  }
  // minor_0 = T; minor_{b+1} = D_b = E_b + delta_b(end) T for in-chunk bits, delta_b T above.
  const uint64_t tend = t0 + ((uint64_t)1 << c) - 1;
  const uint64_t gend = tend ^ (tend >> 1);
  mi[0] = {Tr, Ti};
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const double d = ((gend >> b) & 1ull) ? -1.0 : 1.0;
    if (b < c) {
      double er, ei;
      if (b < U) {
        er = Esr[b < U ? b : 0];
        ei = Esi[b < U ? b : 0];
      } else {
        const cd e = Ed[(b - U) * 64];
        er = e.re;
        ei = e.im;
      }
      mi[b + 1] = {fma(d, Tr, er), fma(d, Ti, ei)};
    } else {
      mi[b + 1] = {d * Tr, d * Ti};
    }
  }
}

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
  const uint64_t nchunks = (uint64_t)1 << (K - 1 - c);
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
    Col0<R, R0> c1;
    if constexpr (R0 == 2) load_srow(c1.sr, (const cpx<double> *)(Vs + R));
    if constexpr (R0 == 3) {
#pragma unroll
      for (int u = 0; u < Col0<R, R0>::NU; ++u) {
        load_bcast(c1.b[u], (const double *)(Vs + (u + 1) * R), lane);
      }
      c1.one = vgpr_one();
      asm volatile("s_nop 4");  // VALU writes of the store before its first DPP read
    }
    cd acc = {0.0, 0.0};
    for (int it = 0; it < iters; ++it) {
      const uint64_t q = ((uint64_t)w + (uint64_t)it * wpi) * 64 + lane;
      cd mi[K];
      // Lanes past the last chunk walk a valid one and drop it, so every lane stays active
      // (the broadcast flips read other lanes' registers).
      const bool live = q < nchunks;
      sweep_chunk<K, U, R0>(Vs, c1, Ed + lane, c, live ? q : (q & (nchunks - 1)), mi);
      if (!live) {
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
}

// minors[slot*K + nu] = sum_w partial[(slot*wpi + w)*K + nu], fixed order.
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
constexpr int64_t kSweepMaxBlocks = 8192;  // 4 one-wave blocks per SIMD of 256 CUs x 4 SIMDs x 2

// Default chunk: 2^min(11, k-7) ranks per lane (>= 2^U), so an item has at least 64 chunks
// (one wave) and a lane walks at most 2048 Gray steps between re-seeds. The tuner may pick
// longer walks (fewer seedings and reductions) where the batch still fills the GPU.
inline SweepPlan sweep_plan(int k, int c_override) {
  SweepPlan p{};
  const int nb = k - 1;                      // Gray bits (particle 0 is fixed)
  const int u = std::min(tb::kSweepU, nb);   // static bits of the kernel for this k
  int c = c_override > 0 ? c_override : std::min(TB_SWEEP_BUILTIN_MAX_C, nb - 6);
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

// LDS per wave: the item's k(k-1) columns plus 64 lanes x (c - U) corrector slots.
inline size_t sweep_lds_bytes(int k, int c) {
  const int u = std::min(tb::kSweepU, k - 1);
  return sizeof(tb::cd) * ((size_t)k * (k - 1) + (size_t)std::max(c - u, 0) * 64);
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
  const SweepPlan p = sweep_plan(k, c_override);
  return align256((size_t)items * p.wpi * k * sizeof(tb::cd));
}

// Partial sums only: partial[(item * wpi + w) * k + nu], w < *wpi. A consumer that sums
// them in w order gets the same minors as run_sweep.
int32_t run_sweep_partial(int k, const tb::cd *Vt, int ldv, int64_t item_stride,
                          const int *item_map, const int *item_count, int items, int c_override,
                          tb::cd *partial, int *wpi, hipStream_t s) {
  const SweepPlan p = sweep_plan(k, c_override);
  *wpi = p.wpi;
  return dispatch_sweep(k, Vt, ldv, item_stride, item_map, item_count, items, p, partial, s);
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
  return tbs::sweep_partial_bytes(k, batch, tbh::tuned(TB_TUNE_SWEEP, k));
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
  const int c = tbh::tuned(TB_TUNE_SWEEP, k);  // read once: size check and launch agree
  const size_t need = tbs::sweep_partial_bytes(k, batch, c);
  if (!workspace || workspace_bytes < need)
    return tbh::fail(TB_EWORKSPACE, "sweep: workspace %zu bytes < required %zu", workspace_bytes,
                     need);
  return tbs::run_sweep(k, (const tb::cd *)Vt, ldv, item_stride, nullptr, nullptr, batch, c,
                        (tb::cd *)workspace, (tb::cd *)minors, s);
}
