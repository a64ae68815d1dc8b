// Histogram of output samples over the reference's outcome order.
//
// Replaces the counting loop of BSSampleBasedDistributionCalculator
// (theboss/distribution_calculators/bs_sample_based_distribution_calculator.py:92-106),
// which scans the outcome list for every sample, by a direct rank: the index of an
// occupation vector in generate_possible_states(n, m, losses)
// (theboss/boson_sampling_utilities/boson_sampling_utilities.py:81-152). That list holds
// one block per particle number l (ascending; only l = n without losses), each block in
// reverse lexicographic order. Within a block, the states ranked before s are those that
// agree with s on modes < i and hold more particles in mode i; with r particles left for
// modes i..m-1 they number
//   sum_{v = s_i + 1}^{r} C(r - v + m - i - 2, m - i - 2) = C(r - s_i - 1 + m - i - 1, m - i - 1)
// (hockey stick), so the rank is a sum of m table lookups. Block l starts at
// sum_{l' < l} C(l' + m - 1, m - 1) = C(l - 1 + m, m).
//
// HBM-bound: each sample is read once (4m bytes, staged through LDS) and adds one to a
// counter (an atomic on the 4-byte counter of its outcome).
#include <algorithm>

#include "tb_common.hpp"

namespace tb {

// Binomial table T[y][r] = C(r + y, y), y = 0..m, r = 0..n, built in LDS by anti-diagonals
// (T[y][r] = T[y-1][r] + T[y][r-1]).
__device__ __forceinline__ uint64_t hist_binom(const uint64_t *T, int n, int y, int r) {
  return T[y * (n + 1) + r];
}

// Samples are staged through LDS in tiles of kHistTileInts ints so the HBM reads are
// coalesced (a lane reading its own row directly would touch one cache line per lane and
// load instruction).
constexpr int kHistTileInts = 6144;

__global__ __launch_bounds__(256) void outcome_histogram_kernel(
    const int *__restrict__ samples, int ld, int S, int m, int n, int losses,
    unsigned *__restrict__ counts, unsigned *__restrict__ outside) {
  extern __shared__ uint64_t T[];
  const int W = n + 1, entries = (m + 1) * W;
  int *tile = (int *)(T + entries);
  for (int d = 0; d <= m + n; ++d) {
    for (int e = threadIdx.x; e < entries; e += blockDim.x) {
      const int y = e / W, r = e - y * W;
      if (y + r != d) continue;
      T[e] = (y == 0 || r == 0) ? 1ull : T[e - W] + T[e - 1];
    }
    __syncthreads();
  }
  const int rows_per_tile = kHistTileInts / m;
  for (int64_t base = (int64_t)blockIdx.x * rows_per_tile; base < S;
       base += (int64_t)gridDim.x * rows_per_tile) {
    const int rows = (int)min((int64_t)rows_per_tile, (int64_t)S - base);
    if (ld == m) {  // contiguous rows: one flat, coalesced copy
      const int *src = samples + base * m;
      for (int e = threadIdx.x; e < rows * m; e += blockDim.x) tile[e] = src[e];
    } else {
      for (int e = threadIdx.x; e < rows * m; e += blockDim.x) {
        const int r = e / m, c = e - r * m;
        tile[e] = samples[(base + r) * ld + c];
      }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
      const int *row = tile + r * m;
      int l = 0;
      bool bad = false;
      for (int i = 0; i < m; ++i) {
        const int v = row[i];
        bad |= (v < 0);
        l += v;
      }
      if (bad || l > n || (!losses && l != n)) {
        if (outside) atomicAdd(outside, 1u);
        continue;
      }
      uint64_t rank = (losses && l > 0) ? hist_binom(T, n, m, l - 1) : 0;  // C(l-1+m, m)
      int rem = l;
      for (int i = 0; i + 1 < m; ++i) {
        const int v = row[i];
        if (rem - v - 1 >= 0) rank += hist_binom(T, n, m - i - 1, rem - v - 1);
        rem -= v;
      }
      atomicAdd(counts + rank, 1u);
    }
    __syncthreads();  // the tile is overwritten next
  }
}

}  // namespace tb

namespace {

// Number of outcomes, or -1 past 2^31 - 1.
int64_t outcome_count(int n, int m, int losses) {
  // C(n + m - 1, m - 1) without losses, C(n + m, m) with (all l <= n), exact in
  // unsigned 128-bit while it stays below the limit.
  const int top = losses ? n + m : n + m - 1;
  const int k = losses ? m : m - 1;
  unsigned __int128 c = 1;
  for (int i = 1; i <= k; ++i) {
    c = c * (unsigned __int128)(top - k + i) / (unsigned __int128)i;
    if (c > (unsigned __int128)0x7fffffff) return -1;
  }
  return (int64_t)c;
}

}  // namespace

extern "C" int64_t tb_outcome_count(int32_t n, int32_t m, int32_t losses) {
  if (n < 0 || m < 1) return tbh::fail(TB_EINVAL, "outcomes: n=%d m=%d", n, m);
  if (m > TB_HIST_MAX_M || n > TB_HIST_MAX_N)
    return tbh::fail(TB_ERANGE, "outcomes: m=%d n=%d beyond %d, %d", m, n, TB_HIST_MAX_M,
                     TB_HIST_MAX_N);
  const int64_t c = outcome_count(n, m, losses ? 1 : 0);
  if (c < 0) return tbh::fail(TB_ERANGE, "outcomes: more than 2^31 - 1 states (n=%d, m=%d)", n, m);
  return c;
}

extern "C" int32_t tb_outcome_histogram(const int32_t *samples, int32_t ld, int32_t samples_n,
                                        int32_t m, int32_t n, int32_t losses, uint32_t *counts,
                                        uint32_t *outside, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (samples_n < 0) return tbh::fail(TB_EINVAL, "histogram: samples=%d < 0", samples_n);
  const int64_t c = tb_outcome_count(n, m, losses);
  if (c < 0) return (int32_t)c;
  if (samples_n == 0) return TB_OK;
  if (!samples || !counts) return tbh::fail(TB_EINVAL, "histogram: NULL pointer");
  if (ld < m) return tbh::fail(TB_EINVAL, "histogram: ld=%d < m=%d", ld, m);
  const size_t lds = (size_t)(m + 1) * (n + 1) * sizeof(uint64_t) + tb::kHistTileInts * sizeof(int);
  const int rows_per_tile = tb::kHistTileInts / m;
  // persistent: the table is built once per block
  const int blocks = (int)std::min<int64_t>(((int64_t)samples_n + rows_per_tile - 1) / rows_per_tile, 1024);
  hipLaunchKernelGGL(tb::outcome_histogram_kernel, dim3(blocks), dim3(256), lds, s, samples, ld,
                     samples_n, m, n, losses ? 1 : 0, counts, outside);
  return tbh::check_hip(hipGetLastError(), "outcome_histogram_kernel");
}
