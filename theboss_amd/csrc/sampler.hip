// Batched generalized Clifford & Clifford (GCC-B) sampler on the device.
//
// Restates GeneralizedCliffordsBSimulationStrategy
// (theboss/simulation_strategies/generalized_cliffords_b_simulation_strategy.py:41-110) and
// its uniform-loss variant
// (generalized_cliffords_b_uniform_losses_simulation_strategy.py:50-121) for S samples at
// once. Per sample, step k = 1..l:
//   1. the next input particle joins the current input (the reference pops a uniformly
//      random remaining particle, :102-105; here a per-sample uniformly random order);
//   2. the k minors perm(U[r, s - e_nu]) of the current input s (k particles) and output
//      r (k-1 particles) come from one C&C sweep (:73-77);
//   3. Laplace expansion amp_m = sum_nu U[m][in_nu] * minor_nu over all m output modes,
//      pmf = |amp|^2 / sum (:80-92);
//   4. the output mode is drawn by inverse CDF (numpy.random.choice, :107-110).
// Samples are processed in lock-step per k. Uniform loss draws l ~ Binomial(n, eta) per
// sample by CDF threshold (:67-85); samples are ordered by l (largest first), so the
// active samples at step k are a prefix whose length is read from device memory by the
// persistent sweep and step kernels. Steps 1..12 of every sample run in one launch (one
// wave per sample); each later step is a sweep launch and a step launch. The whole
// sequence is captured once into a hipGraph and replayed; per-call values (seed, eta)
// live in device memory.
//
// Randomness: Philox-4x32-10 (Salmon et al., SC'11) keyed by the seed, counter
// (sample, step, purpose). Callers that need the reference's exact draws pass
// forced orders / uniforms / step counts instead.
#include <stddef.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "sweep_impl.hpp"

namespace tbs {
size_t sweep_partial_bytes(int k, int items, int c_override);
int32_t run_sweep_partial(int k, const tb::cd *Vt, int ldv, int64_t item_stride,
                          const int *item_map, const int *item_count, int items, int c_override,
                          tb::cd *partial, int *wpi, hipStream_t s);
void sweep_plan_of(int k, int items, int c_override, int *c, int *wpi, int *iters);
}  // namespace tbs

namespace tb {

// ---------------------------------------------------------------- Philox-4x32-10
struct u4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = {hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Uniform double in [0,1) with 53 random bits.
__device__ __forceinline__ double philox_uniform(uint64_t seed, uint32_t sample, uint32_t step,
                                                 uint32_t purpose) {
  const u4 r = philox({sample, step, purpose, 0x7462u}, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t bits = ((uint64_t)(r.x >> 5) << 26) | (uint64_t)(r.y >> 6);
  return (double)bits * 0x1.0p-53;
}

enum : uint32_t { kPurposeDraw = 1, kPurposeOrder = 2, kPurposeLoss = 3 };

// Per-call parameters, the caller's buffers included, kept in device memory: a captured
// graph depends only on the shapes and the workspace, and replays for any seed, eta,
// interferometer, output or forced-draw buffers.
struct GccbParams {
  uint64_t seed;
  double eta;
  const cd *U;
  const int *particles;
  const int *forced_order;
  const double *forced_uniforms;
  const int *forced_steps;
  int *out_samples;
  cd *trace;
  int64_t u_stride;  // per-sample interferometers: sample s uses U + s * u_stride
  int ldu;
  int per_sample;    // 1: per-sample U and particle lists (particles + s * n)
  int clamped;       // set by the kernels when an index argument was out of range
};

// Bits of GccbParams::clamped.
enum : int { kClampParticle = 1, kClampOrder = 2 };

__global__ void gccb_set_params_kernel(GccbParams *p, GccbParams v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__device__ __forceinline__ double binom_d(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  if (k > n - k) k = n - k;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
  return rint(r);
}

// Per sample: particle order, number of steps, zeroed output row. One wave per 64
// samples: each lane shuffles its order in LDS (column lane of ord_lds), and the wave
// writes the orders and zeroes the output rows of its samples with coalesced stores.
constexpr int kInitBlock = 64;
__global__ __launch_bounds__(kInitBlock) void gccb_init_kernel(
    const GccbParams *__restrict__ prm, int m, int n, int S, int *__restrict__ order,
    int *__restrict__ steps) {
  __shared__ int ord_lds[TB_SWEEP_MAX_K * kInitBlock];  // ord_lds[i * kInitBlock + lane]
  const int lane = threadIdx.x;
  const int s0 = blockIdx.x * kInitBlock;
  const int rows = min(kInitBlock, S - s0);
  const int s = s0 + lane;
  int *out_samples = prm->out_samples;
  for (int i = lane; i < rows * m; i += kInitBlock) out_samples[(size_t)s0 * m + i] = 0;
  if (s < S) {
    const uint64_t seed = prm->seed;
    const double eta = prm->eta;
    const int *forced_order = prm->forced_order;
    const int *forced_steps = prm->forced_steps;
    int *ord = ord_lds + lane;
    // Per-sample particle lists hold the sample's forced_steps[s] particles first: only
    // that prefix is shuffled, so the padding beyond it never enters a step.
    const int span = (prm->per_sample && forced_steps) ? min(max(forced_steps[s], 0), n) : n;
    if (forced_order) {
      // Entries outside [0, n) would index past the sample's particle rows: clamp them
      // (the call then reports TB_EINVAL through tb_gccb_status).
      bool bad = false;
      for (int i = 0; i < n; ++i) {
        int v = forced_order[(size_t)s * n + i];
        if (v < 0 || v >= n) {
          bad = true;
          v = v < 0 ? 0 : n - 1;
        }
        ord[i * kInitBlock] = v;
      }
      if (bad) atomicOr(const_cast<int *>(&prm->clamped), kClampOrder);
    } else {
      for (int i = 0; i < n; ++i) ord[i * kInitBlock] = i;
      for (int i = span - 1; i >= 1; --i) {  // Fisher-Yates
        const double u = philox_uniform(seed, (uint32_t)s, (uint32_t)i, kPurposeOrder);
        int j = (int)(u * (double)(i + 1));
        j = j > i ? i : j;
        const int t = ord[i * kInitBlock];
        ord[i * kInitBlock] = ord[j * kInitBlock];
        ord[j * kInitBlock] = t;
      }
    }
    int l;
    if (forced_steps) {
      l = forced_steps[s];
    } else if (eta >= 1.0) {
      l = n;
    } else {
      // generalized_cliffords_b_uniform_losses_simulation_strategy.py:56-85
      const double u = philox_uniform(seed, (uint32_t)s, 0u, kPurposeLoss);
      double acc = 0.0;
      l = n;
      for (int k = 0; k <= n; ++k) {
        acc += binom_d(n, k) * pow(eta, (double)k) * pow(1.0 - eta, (double)(n - k));
        if (acc > u) {
          l = k;
          break;
        }
      }
    }
    l = l < 0 ? 0 : (l > n ? n : l);
    steps[s] = l;
  }
  __syncthreads();
  // order[(s0 + r) * n + i] = ord_lds[i * kInitBlock + r]: the block's rows are contiguous.
  for (int e = lane; e < rows * n; e += kInitBlock) {
    const int r = e / n, i = e - r * n;
    order[(size_t)s0 * n + e] = ord_lds[i * kInitBlock + r];
  }
}

// Single workgroup of 16 waves: slots ordered by steps (descending, stable in the sample
// index) and count[v] = number of samples with steps >= v, v = 0..n. Wave w owns the
// contiguous samples [w * seg, (w + 1) * seg): it counts each value there, the offsets
// follow (larger values first, then lower waves), and the wave scatters its samples in
// index order, ranking the lanes of a 64-sample slice that share a value by ballot.
// steps[] holds values in 0..n (the init kernel clamps them).
__global__ __launch_bounds__(1024) void gccb_order_kernel(const int *__restrict__ steps, int S,
                                                          int n, int *__restrict__ slot_map,
                                                          int *__restrict__ count) {
  constexpr int W = 16, V = TB_SWEEP_MAX_K + 1;
  __shared__ int cnt[W][V];
  __shared__ int tot[V];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  for (int i = tid; i < W * V; i += 1024) cnt[i / V][i % V] = 0;
  __syncthreads();
  const int seg = (S + W - 1) / W;
  const int lo = min(S, w * seg), hi = min(S, lo + seg);
  for (int s = lo + lane; s < hi; s += 64) atomicAdd(&cnt[w][steps[s]], 1);
  __syncthreads();
  if (tid <= n) {
    int t = 0;
    for (int ww = 0; ww < W; ++ww) t += cnt[ww][tid];
    tot[tid] = t;
  }
  __syncthreads();
  if (tid <= n) {
    int base = 0;
    for (int v = n; v > tid; --v) base += tot[v];
    for (int ww = 0; ww < W; ++ww) {
      const int c = cnt[ww][tid];
      cnt[ww][tid] = base;
      base += c;
    }
    count[tid] = base;
  }
  __syncthreads();
  for (int s0 = lo; s0 < hi; s0 += 64) {
    const int s = s0 + lane;
    const bool in = s < hi;
    const int v = in ? steps[s] : -1;
    unsigned long long todo = __ballot(in);
    while (todo) {
      const int leader = __ffsll((long long)todo) - 1;
      const int lv = __shfl(v, leader, 64);
      const unsigned long long same = __ballot(in && v == lv);
      const int at = cnt[w][lv];
      if (in && v == lv) slot_map[at + __popcll(same & ((1ull << lane) - 1ull))] = s;
      if (lane == leader) cnt[w][lv] = at + __popcll(same);
      todo &= ~same;
    }
  }
}

// U^T restricted to input particles: UinT[nu][mode] = U[mode][particles[nu]], one copy
// per sample when the samples have their own interferometers (items = S, else 1).
__global__ void gccb_uin_kernel(const GccbParams *__restrict__ prm, int m, int n, int items,
                                cd *__restrict__ UinT) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)items * m * n) return;
  const int s = (int)(idx / ((int64_t)m * n));
  const int rem = (int)(idx - (int64_t)s * m * n);
  const int nu = rem / m, mm = rem - nu * m;
  const cd *U = prm->U + (size_t)s * prm->u_stride;
  int col = prm->particles[(size_t)s * n + nu];
  if (col < 0 || col >= m) {  // not a mode index: clamp, and report through tb_gccb_status
    if (mm == 0) atomicOr(const_cast<int *>(&prm->clamped), kClampParticle);
    col = col < 0 ? 0 : m - 1;
  }
  UinT[idx] = U[(size_t)mm * prm->ldu + col];
}

// Laplace expansion, draw and output of step k for sample s (one wave): the k minors are
// in mi (LDS), rows[nu] is particle nu's UinT row. pm is LDS scratch for m probabilities.
// KC > 0: k == KC at compile time (the particle loop is unrolled, so a lane's loads of one
// mode are in flight together). Returns the chosen output mode (wave-uniform).
template <int KC = 0>
__device__ __forceinline__ int gccb_draw(const GccbParams *__restrict__ prm, int k, int m, int n,
                                         int s, const cd *mi, const int *rows, double *pm,
                                         const cd *__restrict__ UinT, int *__restrict__ outs) {
  if constexpr (KC > 0) k = KC;
  const int lane = threadIdx.x;
  const double *forced_uniforms = prm->forced_uniforms;
  int *out_samples = prm->out_samples;
  cd *trace = prm->trace;
  if (trace && lane < k) trace[((size_t)s * n + (k - 1)) * n + lane] = mi[lane];
  // pmf: amp_m = sum_nu U[m][in_nu] minor_nu over the current k particles. Lanes take
  // modes lane, lane + 64, ... so each UinT row is read with coalesced loads.
  for (int mm = lane; mm < m; mm += 64) {
    double re = 0.0, im = 0.0;
    if constexpr (KC > 0) {
      constexpr int G = KC < 16 ? KC : 16;  // loads in flight per group
#pragma unroll
      for (int nu0 = 0; nu0 < KC; nu0 += G) {
        cd u[G];
#pragma unroll
        for (int g = 0; g < G; ++g)
          if (nu0 + g < KC) u[g] = UinT[(size_t)rows[nu0 + g] * m + mm];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (nu0 + g < KC) {
            const cd v = mi[nu0 + g];
            re += u[g].re * v.re - u[g].im * v.im;
            im += u[g].re * v.im + u[g].im * v.re;
          }
        }
      }
    } else {
      for (int nu = 0; nu < k; ++nu) {
        const cd u = UinT[(size_t)rows[nu] * m + mm];
        const cd v = mi[nu];
        re += u.re * v.re - u.im * v.im;
        im += u.re * v.im + u.im * v.re;
      }
    }
    pm[mm] = re * re + im * im;
  }
  __syncthreads();
  // Inverse CDF over contiguous mode ranges per lane.
  const int per = (m + 63) / 64;
  const int m0 = lane * per;
  const int m1 = min(m, m0 + per);
  double local = 0.0;
  for (int mm = m0; mm < m1; ++mm) local += pm[mm];
  // Inclusive scan of lane sums.
  double incl = local;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  const double total = __shfl(incl, 63, 64);
  const double u = forced_uniforms ? forced_uniforms[(size_t)s * n + (k - 1)]
                                   : philox_uniform(prm->seed, (uint32_t)s, (uint32_t)k, kPurposeDraw);
  const double target = u * total;
  const double excl = incl - local;
  // The lane whose range holds the target picks the first mode with running sum > target.
  int pick = -1;
  if (excl <= target && target < incl) {
    double run = excl;
    for (int mm = m0; mm < m1; ++mm) {
      run += pm[mm];
      if (run > target) {
        pick = mm;
        break;
      }
    }
  }
  // Resolve (first lane wins); rounding at the very top falls back to the last mode with
  // nonzero probability.
  const unsigned long long has = __ballot(pick >= 0);
  int chosen;
  if (has) {
    const int src = __ffsll((long long)has) - 1;
    chosen = __shfl(pick, src, 64);
  } else {
    int last = -1;
    for (int mm = m0; mm < m1; ++mm)
      if (pm[mm] > 0.0) last = mm;
    int best = last;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) best = max(best, __shfl_xor(best, off, 64));
    chosen = best < 0 ? 0 : best;
  }
  if (lane == 0) {
    outs[(size_t)s * n + (k - 1)] = chosen;
    out_samples[(size_t)s * m + chosen] += 1;
  }
  return chosen;
}

// Step-kernel LDS: k minors, m probabilities, n particle rows.
__host__ __device__ constexpr size_t gccb_step_lds(int m, int n) {
  return (size_t)n * sizeof(cd) + (size_t)m * sizeof(double) + (size_t)n * sizeof(int);
}

// One wave per active slot at step k = K: the sweep's wpi partial sums of the k minors
// (added in w order), Laplace over m modes, inverse-CDF draw, output and column
// bookkeeping for the next step.
template <int K>
__global__ __launch_bounds__(64) void gccb_step_kernel(
    const GccbParams *__restrict__ prm, int m, int n, int S, const int *__restrict__ count,
    const int *__restrict__ slot_map, const cd *__restrict__ partial, int wpi,
    const cd *__restrict__ UinT, const int *__restrict__ order, int *__restrict__ outs,
    cd *__restrict__ Vt) {
  constexpr int k = K;
  extern __shared__ char smem[];
  cd *mi = (cd *)smem;                 // k minors
  double *pm = (double *)(mi + n);     // m probabilities
  int *rows = (int *)(pm + m);         // UinT row of each of the k particles
  const int slot = blockIdx.x;
  const int lane = threadIdx.x;
  if (slot >= count[k]) return;
  const int s = slot_map[slot];
  const int *ord = order + (size_t)s * n;
  if (prm->per_sample) UinT += (size_t)s * n * m;
  if (lane < k) {
    cd v{1.0, 0.0};
    if (k > 1) {
      double re = 0.0, im = 0.0;
      const cd *p = partial + (size_t)slot * wpi * k + lane;
#pragma unroll 8
      for (int w = 0; w < wpi; ++w) {  // unrolled: the loads are in flight together
        re += p[(size_t)w * k].re;
        im += p[(size_t)w * k].im;
      }
      v = {re, im};
    }
    mi[lane] = v;
    rows[lane] = ord[lane];
  }
  __syncthreads();
  const int chosen = gccb_draw<K>(prm, k, m, n, s, mi, rows, pm, UinT, outs);
  // New output row j = k-1 for every input particle: Vt[s][nu][k-1] = U[chosen][in_nu].
  for (int nu = lane; nu < n; nu += 64)
    Vt[((size_t)s * n + nu) * n + (k - 1)] = UinT[(size_t)ord[nu] * m + chosen];
}

// Steps 1..kmax of a sample in one wave (kmax <= kSmallK): for small k the separate sweep
// and step launches cost more than their work. Per step the wave sweeps the item's
// columns (kept in LDS; the same chunks, wave sums and w order as sweep_kernel followed by
// gccb_step_kernel, so the minors are bit-identical), then draws as gccb_step_kernel does.
// At the end it writes the sample's first kmax output columns of Vt for the later steps.
#ifndef TB_GCCB_SMALL_K
#define TB_GCCB_SMALL_K 12
#endif
constexpr int kSmallK = TB_GCCB_SMALL_K;
struct SmallPlan {
  int c[kSmallK + 1], wpi[kSmallK + 1], iters[kSmallK + 1];
};

// LDS of the small kernel: Vl [kSmallK][kSmallK] (Vl[nu * kSmallK + j] = V[nu][j]), Vs
// [kSmallK * (kSmallK - 1)], Ed [ed_slots * 64], then the step's minors, probabilities and
// rows (gccb_step_lds), then the chosen mode of each step.
__host__ __device__ constexpr size_t gccb_small_lds(int m, int n, int ed_slots) {
  return sizeof(cd) * ((size_t)kSmallK * kSmallK + (size_t)kSmallK * (kSmallK - 1) +
                       (size_t)ed_slots * 64) +
         gccb_step_lds(m, n) + sizeof(int) * kSmallK;
}

template <int K>
__device__ __forceinline__ void gccb_small_sweep(const cd *Vl, cd *Vs, cd *Ed, cd *mi,
                                                 const SmallPlan &pl) {
  constexpr int R = K - 1;
  constexpr int R0 = (K <= TB_SWEEP_BCAST_MAX_K) ? 3 : 2;
  constexpr int U = (K - 1 < kSweepU) ? K - 1 : kSweepU;
  const int lane = threadIdx.x;
  for (int e = lane; e < K * R; e += 64) {
    const int nu = e / R, j = e - nu * R;
    Vs[e] = Vl[nu * kSmallK + j];
  }
  __syncthreads();
  double re = 0.0, im = 0.0;
  for (int w = 0; w < pl.wpi[K]; ++w) {
    const cd acc = sweep_wave<K, U, R0>(Vs, Ed, lane, pl.c[K], w, pl.wpi[K], pl.iters[K]);
    re += acc.re;
    im += acc.im;
  }
  if (lane < K) mi[lane] = {re, im};
  __syncthreads();
}

template <int K>
__device__ __forceinline__ int gccb_small_steps(const GccbParams *__restrict__ prm, int kmax,
                                                int m, int n, int s, int slot,
                                                const int *__restrict__ count, cd *Vl, cd *Vs,
                                                cd *Ed, cd *mi, double *pm, const int *rows,
                                                int *ch, const cd *__restrict__ UinT,
                                                int *__restrict__ outs, const SmallPlan &pl) {
  if constexpr (K > kSmallK) {
    return K - 1;
  } else {
    if (K > kmax || slot >= count[K]) return K - 1;
    const int lane = threadIdx.x;
    if constexpr (K == 1) {
      if (lane == 0) mi[0] = {1.0, 0.0};
      __syncthreads();
    } else {
      gccb_small_sweep<K>(Vl, Vs, Ed, mi, pl);
    }
    const int chosen = gccb_draw<K>(prm, K, m, n, s, mi, rows, pm, UinT, outs);
    if (lane == 0) ch[K - 1] = chosen;
    for (int nu = lane; nu < kmax; nu += 64)
      Vl[nu * kSmallK + (K - 1)] = UinT[(size_t)rows[nu] * m + chosen];
    __syncthreads();
    return gccb_small_steps<K + 1>(prm, kmax, m, n, s, slot, count, Vl, Vs, Ed, mi, pm, rows, ch,
                                   UinT, outs, pl);
  }
}

__global__ __launch_bounds__(64) void gccb_small_kernel(
    const GccbParams *__restrict__ prm, int kmax, int m, int n, int ed_slots,
    const int *__restrict__ count, const int *__restrict__ slot_map, const cd *__restrict__ UinT,
    const int *__restrict__ order, int *__restrict__ outs, cd *__restrict__ Vt, SmallPlan pl) {
  extern __shared__ char smem[];
  cd *Vl = (cd *)smem;
  cd *Vs = Vl + kSmallK * kSmallK;
  cd *Ed = Vs + kSmallK * (kSmallK - 1);
  cd *mi = Ed + (size_t)ed_slots * 64;
  double *pm = (double *)(mi + n);
  int *rows = (int *)(pm + m);
  int *ch = rows + n;
  const int slot = blockIdx.x;
  const int lane = threadIdx.x;
  if (slot >= count[1]) return;
  const int s = slot_map[slot];
  const int *ord = order + (size_t)s * n;
  if (prm->per_sample) UinT += (size_t)s * n * m;
  for (int nu = lane; nu < n; nu += 64) rows[nu] = ord[nu];
  __syncthreads();
  const int done = gccb_small_steps<1>(prm, kmax, m, n, s, slot, count, Vl, Vs, Ed, mi, pm, rows,
                                       ch, UinT, outs, pl);
  // Vt[s][nu][j] = U[out_j][in_nu] for the done steps, all n particles.
  for (int e = lane; e < n * done; e += 64) {
    const int nu = e / done, j = e - nu * done;
    Vt[((size_t)s * n + nu) * n + j] = UinT[(size_t)rows[nu] * m + ch[j]];
  }
}

// Vt rows follow the sample's particle order: Vt[s][nu][j] uses particle ord[nu]. The
// step kernel writes row entries through ord, and the sweep reads the first k rows.

}  // namespace tb

// ------------------------------------------------------------------------------------
// Host side: workspace layout, launch sequence, graph cache.
// ------------------------------------------------------------------------------------
namespace {

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Layout {
  size_t params, uint, order, steps, slot_map, count, outs, vt, partial, total;
};

Layout layout(const tb_gccb_config &c) {
  const size_t S = (size_t)c.samples, n = (size_t)c.n, m = (size_t)c.m;
  Layout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align256(bytes);
    return at;
  };
  L.params = take(sizeof(tb::GccbParams));
  L.uint = take((c.per_sample ? S : 1) * n * m * sizeof(tb::cd));
  L.order = take(S * n * sizeof(int));
  L.steps = take(S * sizeof(int));
  L.slot_map = take(S * sizeof(int));
  L.count = take((n + 2) * sizeof(int));
  L.outs = take(S * n * sizeof(int));
  L.vt = take(S * n * n * sizeof(tb::cd));
  size_t part = 0;
  for (int k = 2; k <= c.n; ++k) part = std::max(part, tbs::sweep_partial_bytes(k, c.samples, c.chunk_bits[k]));
  L.partial = take(part);
  L.total = off;
  return L;
}

template <int K = 1>
int32_t launch_step(int k, int S, size_t lds, hipStream_t s, const tb::GccbParams *prm, int m,
                    int n, const int *count, const int *slot_map, const tb::cd *partial, int wpi,
                    const tb::cd *UinT, const int *order, int *outs, tb::cd *Vt) {
  if constexpr (K > TB_SWEEP_MAX_K) {
    return tbh::fail(TB_ERANGE, "gccb: step %d exceeds TB_SWEEP_MAX_K=%d", k, TB_SWEEP_MAX_K);
  } else {
    if (k != K)
      return launch_step<K + 1>(k, S, lds, s, prm, m, n, count, slot_map, partial, wpi, UinT,
                                order, outs, Vt);
    hipLaunchKernelGGL(tb::gccb_step_kernel<K>, dim3(S), dim3(64), lds, s, prm, m, n, S, count,
                       slot_map, partial, wpi, UinT, order, outs, Vt);
    return tbh::check_hip(hipGetLastError(), "gccb_step_kernel");
  }
}

int32_t enqueue(const tb_gccb_config &c, char *ws, hipStream_t s) {
  const Layout L = layout(c);
  const int S = c.samples, n = c.n, m = c.m;
  tb::GccbParams *prm = (tb::GccbParams *)(ws + L.params);
  tb::cd *UinT = (tb::cd *)(ws + L.uint);
  int *order = (int *)(ws + L.order);
  int *steps = (int *)(ws + L.steps);
  int *slot_map = (int *)(ws + L.slot_map);
  int *count = (int *)(ws + L.count);
  int *outs = (int *)(ws + L.outs);
  tb::cd *Vt = (tb::cd *)(ws + L.vt);
  tb::cd *partial = (tb::cd *)(ws + L.partial);
  const int uin_items = c.per_sample ? S : 1;
  hipLaunchKernelGGL(tb::gccb_uin_kernel,
                     dim3((unsigned)(((int64_t)uin_items * m * n + 255) / 256)), dim3(256), 0, s,
                     prm, m, n, uin_items, UinT);
  hipLaunchKernelGGL(tb::gccb_init_kernel, dim3((S + tb::kInitBlock - 1) / tb::kInitBlock),
                     dim3(tb::kInitBlock), 0, s, prm, m, n, S, order, steps);
  hipLaunchKernelGGL(tb::gccb_order_kernel, dim3(1), dim3(1024), 0, s, steps, S, n, slot_map,
                     count);
  int32_t rc = tbh::check_hip(hipGetLastError(), "gccb_init");
  if (rc) return rc;
  const size_t step_lds = tb::gccb_step_lds(m, n);
  // Steps 1..kmax in one launch where its LDS fits; the separate sweep and step launches
  // take the rest.
  int k_first = 1;
  {
    const int kmax = std::min(n, tb::kSmallK);
    tb::SmallPlan pl{};
    int ed_slots = 0;
    for (int k = 2; k <= kmax; ++k) {
      tbs::sweep_plan_of(k, S, c.chunk_bits[k], &pl.c[k], &pl.wpi[k], &pl.iters[k]);
      ed_slots = std::max(ed_slots, pl.c[k] - std::min(tb::kSweepU, k - 1));
    }
    const size_t small_lds = tb::gccb_small_lds(m, n, ed_slots);
    if (small_lds <= 64 * 1024) {
      hipLaunchKernelGGL(tb::gccb_small_kernel, dim3(S), dim3(64), small_lds, s, prm, kmax, m, n,
                         ed_slots, count, slot_map, UinT, order, outs, Vt, pl);
      rc = tbh::check_hip(hipGetLastError(), "gccb_small_kernel");
      if (rc) return rc;
      k_first = kmax + 1;
    }
  }
  int wpi = 0;
  for (int k = k_first; k <= n; ++k) {
    if (k >= 2) {
      rc = tbs::run_sweep_partial(k, Vt, n, (int64_t)n * n, slot_map, count + k, S,
                                  c.chunk_bits[k], partial, &wpi, s);
      if (rc) return rc;
    }
    rc = launch_step(k, S, step_lds, s, prm, m, n, count, slot_map, partial, wpi, UinT, order,
                     outs, Vt);
    if (rc) return rc;
  }
  return TB_OK;
}

// Graph cache: one instantiated graph per shape (the caller's buffers are read from the
// parameter block), bounded so that callers with ever-new workspaces cannot grow it
// without limit.
struct Key {
  int device, m, n, S, per_sample;
  std::vector<int> chunk;
  const void *ws;
  bool operator<(const Key &o) const {
    return std::tie(device, m, n, S, per_sample, chunk, ws) <
           std::tie(o.device, o.m, o.n, o.S, o.per_sample, o.chunk, o.ws);
  }
};
constexpr size_t kMaxGraphs = 32;

std::mutex g_mu;
std::map<Key, hipGraphExec_t> g_graphs;
std::map<int, hipStream_t> g_capture_streams;

}  // namespace

// Steps whose chunk bits the caller leaves at 0 take the tuning table's entry.
static tb_gccb_config resolve_chunks(const tb_gccb_config &cfg) {
  tb_gccb_config c = cfg;
  for (int k = 2; k <= std::min(c.n, TB_SWEEP_MAX_K); ++k)
    if (c.chunk_bits[k] == 0) c.chunk_bits[k] = tbh::tuned(TB_TUNE_SWEEP, k, c.samples);
  return c;
}

extern "C" size_t tb_gccb_workspace_size(const tb_gccb_config *cfg) {
  if (!cfg || cfg->m < 1 || cfg->n < 1 || cfg->samples < 1 || cfg->n > TB_SWEEP_MAX_K) return 0;
  return layout(resolve_chunks(*cfg)).total;
}

extern "C" int32_t tb_gccb_sample(const tb_gccb_config *cfg, const tb_complex128 *U, int32_t ldu,
                                  const int32_t *particles, uint64_t seed, double transmissivity,
                                  const int32_t *forced_order, const double *forced_uniforms,
                                  const int32_t *forced_steps, int32_t *out_samples,
                                  tb_complex128 *trace_minors, void *workspace,
                                  size_t workspace_bytes, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!cfg) return tbh::fail(TB_EINVAL, "gccb: cfg is NULL");
  if (cfg->n > TB_SWEEP_MAX_K)
    return tbh::fail(TB_ERANGE, "gccb: n=%d exceeds TB_SWEEP_MAX_K=%d", cfg->n, TB_SWEEP_MAX_K);
  const tb_gccb_config c = resolve_chunks(*cfg);  // read once: layout, graph key and launches
  if (c.m < 1 || c.n < 1 || c.samples < 1)
    return tbh::fail(TB_EINVAL, "gccb: m=%d n=%d samples=%d must be >= 1", c.m, c.n, c.samples);
  if (c.n > TB_SWEEP_MAX_K)
    return tbh::fail(TB_ERANGE, "gccb: n=%d exceeds TB_SWEEP_MAX_K=%d", c.n, TB_SWEEP_MAX_K);
  if (c.samples > (1 << 24)) return tbh::fail(TB_ERANGE, "gccb: samples=%d > 2^24", c.samples);
  for (int k = 0; k <= c.n; ++k)
    if (c.chunk_bits[k] < 0 || c.chunk_bits[k] > 16)
      return tbh::fail(TB_EINVAL, "gccb: chunk_bits[%d]=%d outside 0..16", k, c.chunk_bits[k]);
  if (!U || !particles || !out_samples) return tbh::fail(TB_EINVAL, "gccb: NULL pointer");
  if (ldu < c.m) return tbh::fail(TB_EINVAL, "gccb: ldu=%d < m=%d", ldu, c.m);
  if (c.per_sample != 0 && c.per_sample != 1)
    return tbh::fail(TB_EINVAL, "gccb: per_sample=%d must be 0 or 1", c.per_sample);
  if (c.per_sample && !forced_steps)
    return tbh::fail(TB_EINVAL, "gccb: per-sample particle lists need forced_steps");
  if (c.per_sample && c.u_stride < (int64_t)c.m * ldu)
    return tbh::fail(TB_EINVAL, "gccb: u_stride=%lld < m*ldu", (long long)c.u_stride);
  if (!(transmissivity >= 0.0 && transmissivity <= 1.0))
    return tbh::fail(TB_EINVAL, "gccb: transmissivity %g outside [0,1]", transmissivity);
  const size_t need = layout(c).total;
  if (!workspace || workspace_bytes < need)
    return tbh::fail(TB_EWORKSPACE, "gccb: workspace %zu bytes < required %zu", workspace_bytes,
                     need);
  const size_t step_lds = tb::gccb_step_lds(c.m, c.n);
  if (step_lds > 64 * 1024) return tbh::fail(TB_ERANGE, "gccb: m=%d too large for LDS pmf", c.m);
  char *ws = (char *)workspace;
  const Layout L = layout(c);
  tb::GccbParams prm_v{};
  prm_v.seed = seed;
  prm_v.eta = transmissivity;
  prm_v.U = (const tb::cd *)U;
  prm_v.ldu = ldu;
  prm_v.particles = particles;
  prm_v.forced_order = forced_order;
  prm_v.forced_uniforms = forced_uniforms;
  prm_v.forced_steps = forced_steps;
  prm_v.out_samples = out_samples;
  prm_v.trace = (tb::cd *)trace_minors;
  prm_v.per_sample = c.per_sample ? 1 : 0;
  prm_v.u_stride = c.per_sample ? c.u_stride : 0;
  hipLaunchKernelGGL(tb::gccb_set_params_kernel, dim3(1), dim3(1), 0, s,
                     (tb::GccbParams *)(ws + L.params), prm_v);
  int32_t rc = tbh::check_hip(hipGetLastError(), "gccb_set_params_kernel");
  if (rc) return rc;

  int dev = 0;
  rc = tbh::check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  Key key{dev, c.m, c.n, c.samples, c.per_sample ? 1 : 0,
          std::vector<int>(c.chunk_bits, c.chunk_bits + c.n + 1), workspace};
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_graphs.find(key);
  if (it == g_graphs.end()) {
    size_t on_dev = 0;
    for (auto &kv : g_graphs) on_dev += (kv.first.device == dev);
    if (on_dev >= kMaxGraphs) {  // drop this device's graphs once none can be in flight
      rc = tbh::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
      if (rc) return rc;
      for (auto g = g_graphs.begin(); g != g_graphs.end();) {
        if (g->first.device == dev) {
          (void)hipGraphExecDestroy(g->second);
          g = g_graphs.erase(g);
        } else {
          ++g;
        }
      }
    }
    hipStream_t cap;
    auto cs = g_capture_streams.find(dev);
    if (cs == g_capture_streams.end()) {
      rc = tbh::check_hip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
      if (rc) return rc;
      g_capture_streams[dev] = cap;
    } else {
      cap = cs->second;
    }
    rc = tbh::check_hip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal),
                        "hipStreamBeginCapture");
    if (rc) return rc;
    int32_t erc = enqueue(c, ws, cap);
    hipGraph_t graph = nullptr;
    rc = tbh::check_hip(hipStreamEndCapture(cap, &graph), "hipStreamEndCapture");
    if (erc) {
      if (graph) (void)hipGraphDestroy(graph);
      return erc;
    }
    if (rc) return rc;
    hipGraphExec_t exec = nullptr;
    rc = tbh::check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0),
                        "hipGraphInstantiate");
    (void)hipGraphDestroy(graph);
    if (rc) return rc;
    it = g_graphs.emplace(key, exec).first;
  }
  return tbh::check_hip(hipGraphLaunch(it->second, s), "hipGraphLaunch");
}

extern "C" int32_t tb_gccb_status(const void *workspace, void *stream) {
  if (!workspace) return tbh::fail(TB_EINVAL, "gccb_status: workspace is NULL");
  hipStream_t s = (hipStream_t)stream;
  int clamped = 0;
  const char *prm = (const char *)workspace;  // the parameter block opens the workspace
  int32_t rc = tbh::check_hip(
      hipMemcpyAsync(&clamped, prm + offsetof(tb::GccbParams, clamped), sizeof(int),
                     hipMemcpyDeviceToHost, s),
      "hipMemcpyAsync");
  if (rc) return rc;
  rc = tbh::check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (rc) return rc;
  if (clamped)
    return tbh::fail(TB_EINVAL, "gccb: %s%s out of range in the last call (clamped; its samples "
                     "are not meaningful)",
                     (clamped & tb::kClampParticle) ? "particles[] entries " : "",
                     (clamped & tb::kClampOrder) ? "forced_order[] entries " : "");
  return TB_OK;
}
