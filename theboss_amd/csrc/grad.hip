// Permanent gradient d perm(A) / dA[i][j] = perm(A without row i and column j), all n^2
// entries of a batch of n x n complex128 matrices, in one Gray sweep per column group.
//
// The reference has no gradients. Row i of the result is what
// BSCCRyserSubmatricesPermanentCalculator.compute_permanents returns with output rows =
// all rows but i (theboss/boson_sampling_utilities/permanent_calculators/
// bs_cc_ryser_submatrices_permanent_calculator.py:73-119).
//
// Formula. Glynn's identity with delta over the rows (delta_{n-1} = +1) and half column
// sums h_j = (1/2) sum_i delta_i A[i][j] gives perm(A) = 2 sum_delta (prod delta) prod_j h_j.
// perm is linear in A[i][j], and d h_l / d A[i][j] = delta_i [l == j] / 2, so
//   G[i][j] = sum_delta (prod delta) delta_i prod_{l != j} h_l.
// One Gray walk over delta serves every row i and a group of G = 4 or 8 columns j. Per
// step:
//   * one row flip of the column sums (n complex adds);
//   * the G leave-one-out products P_j = prod_{l != j} h_l: a product tree over the
//     other n-G columns, suffix products and a prefix chain over the group
//     (n + 2G - 5 complex multiplies);
//   * the running sums T_j += (prod delta) P_j (G complex adds);
//   * for the flipped bit i, its corrector E_ij += 2 delta_i(old) T_j (G complex FMAs).
// At the end of a walk, sum_t delta_i(t) (prod delta) P_j(t) = E_ij + delta_i(end) T_j.
// 8-column groups halve the sweeps per matrix for 1.3x the work per step; they hold
// twice the correctors, so they unroll 2 static bits where 4-column groups unroll 3-4.
//
// Work split. A lane walks an aligned chunk of 2^c Gray ranks. Bits below U flip inside
// the unrolled static block of 2^U steps and keep per-lane correctors in registers. Bits
// U..c-1 flip at block starts in the same direction in every lane of the wave (bit 0 of
// the chunk index, which decides the top in-chunk bit, comes from the wave index), so
// their correctors are kept per wave: at such a flip the wave sums T_j across its lanes
// and the owner lane of row i adds 2 delta_i(old) sum T_j. Bits above the chunk are
// constant over it and contribute delta_i T_j at the chunk end.
//
// Cross-lane sums use permlane swaps and DPP (no LDS traffic) and reduce 8 values (one
// slot of 4 columns) at once (transposed: each row of 16 lanes ends with one column's
// real and imaginary parts), so the output of a wave lives in lane (r, p) = (row of 16
// lanes, position): per slot q, column 4q + grad_row_col(r) of output rows p and 16 + p.
//
// Columns. The item (matrix, group g) stages its matrix in LDS with the group's columns
// last, so the product structure is static. With 8-column groups and 1..4 columns left
// over, those come from one 4-column group per matrix (columns n-4..n-1, a second launch);
// otherwise the last group starts at n - G and may overlap its predecessor. The reduction
// takes each column from its first group.
#include <algorithm>

#include "flip_common.hpp"

// Gray bits per unrolled static block: 4 while the registers allow it (n <= 19 at 2
// waves per SIMD), 3 above.
#ifndef TB_GRAD_U4_MAX_N
#define TB_GRAD_U4_MAX_N 19
#endif
#ifndef TB_GRAD_SROW
#define TB_GRAD_SROW 20
#endif

#ifndef TB_GRAD_ATTR
#define TB_GRAD_ATTR __attribute__((amdgpu_waves_per_eu(2)))
#endif
// Groups of 8 columns for TB_GRAD_G8_MIN_N <= n <= TB_GRAD_G8_MAX_N (every n of the
// single-sweep kernel by default; builds that narrow the range get 4-column groups, which
// need fewer registers): half the sweeps of 4-column groups at ~1.3x the work per step,
// so 1.1-1.4x faster (n = 16..26), with
// TB_GRAD_G8_U static bits (the 8 columns' correctors take the registers of a third
// bit) and TB_GRAD_G8_B broadcast rows up to n = 24 (none above: fewer spills).
#ifndef TB_GRAD_G8_MIN_N
#define TB_GRAD_G8_MIN_N 14
#endif
#ifndef TB_GRAD_G8_MAX_N
#define TB_GRAD_G8_MAX_N 26
#endif
#ifndef TB_GRAD_G8_U
#define TB_GRAD_G8_U 2
#endif
#ifndef TB_GRAD_G8_B
#define TB_GRAD_G8_B 2
#endif

namespace tb {

constexpr int kGradG = 4;  // columns per slot (one per row of 16 lanes)
// Columns per group: 4 (one slot) or 8 (two slots).
template <int N>
constexpr int grad_g() {
  return (N >= TB_GRAD_G8_MIN_N && N <= TB_GRAD_G8_MAX_N) ? 8 : 4;
}
template <int N, int G>
constexpr int grad_u_of() {
  return G == 8 ? TB_GRAD_G8_U : (N <= TB_GRAD_U4_MAX_N ? 4 : 3);
}
template <int N>
constexpr int grad_u() {
  return grad_u_of<N, grad_g<N>()>();
}
inline int grad_g_of(int n) { return (n >= TB_GRAD_G8_MIN_N && n <= TB_GRAD_G8_MAX_N) ? 8 : 4; }
// 8-column groups with 1..4 columns left over: those go to one 4-column group (a second
// launch) instead of an 8-column group that overlaps its predecessor, which would redo
// 4..7 columns (n = 20: 2 x 274 + 202 flops per Gray step instead of 3 x 274).
#ifndef TB_GRAD_TAIL4
#define TB_GRAD_TAIL4 1
#endif
inline bool grad_tail4(int n) {
  return TB_GRAD_TAIL4 && grad_g_of(n) == 8 && n % 8 != 0 && n % 8 <= 4;
}
constexpr int kGradMinN = 14;
constexpr int kGradMaxN = 26;

// (a, b) -> lanes < 32: a + a[lane ^ 32]; lanes >= 32: b + b[lane ^ 32].
__device__ __forceinline__ double swap32_pair_sum(double a, double b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ua, (uint32_t)ub, false, false);
  const auto hi =
      __builtin_amdgcn_permlane32_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
  return join64(lo[0], hi[0]) + join64(lo[1], hi[1]);
}
// (a, b) -> rows 0, 2: a + a[lane ^ 16]; rows 1, 3: b + b[lane ^ 16].
__device__ __forceinline__ double swap16_pair_sum(double a, double b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)ua, (uint32_t)ub, false, false);
  const auto hi =
      __builtin_amdgcn_permlane16_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
  return join64(lo[0], hi[0]) + join64(lo[1], hi[1]);
}
__device__ __forceinline__ double row16_sum(double v) {
  v += dpp_mov<0x128>(v);  // row_ror:8
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x122>(v);  // row_ror:2
  v += dpp_mov<0x121>(v);  // row_ror:1
  return v;
}

// Column held by lane row r after the transposed sum.
__device__ __forceinline__ int grad_row_col(int r) { return (r == 1) ? 2 : (r == 2) ? 1 : r; }

// Wave totals of x[0..3] (real parts of the 4 columns) and x[4..7] (imaginary parts):
// lane row r receives column grad_row_col(r) in (zr, zi).
__device__ __forceinline__ void wave_sum_cols(const double *xr, const double *xi, double &zr,
                                              double &zi) {
  const double y0 = swap32_pair_sum(xr[0], xr[1]);
  const double y1 = swap32_pair_sum(xr[2], xr[3]);
  const double y2 = swap32_pair_sum(xi[0], xi[1]);
  const double y3 = swap32_pair_sum(xi[2], xi[3]);
  zr = row16_sum(swap16_pair_sum(y0, y1));
  zi = row16_sum(swap16_pair_sum(y2, y3));
}

// Output accumulators of a lane: rows p (r0, i0) and 16 + p (r1, i1) of column
// grad_row_col(r). Named scalars, not an array: a select between array slots would put
// the accumulators in scratch memory.
struct GradAcc {
  double r0, i0, r1, i1;
};

// acc(row i) += coef * z in the lanes that own row i (i wave-uniform).
__device__ __forceinline__ void grad_acc_row(GradAcc &a, int p, int i, double coef, double zr,
                                             double zi) {
  const bool own = (p == (i & 15));
  const double f0 = (own && i < 16) ? coef : 0.0;
  const double f1 = (own && i >= 16) ? coef : 0.0;
  a.r0 = fma(f0, zr, a.r0);
  a.i0 = fma(f0, zi, a.i0);
  a.r1 = fma(f1, zr, a.r1);
  a.i1 = fma(f1, zi, a.i1);
}

// Coefficient of row i in the chunk-end sum of T: in-chunk bits >= U (delta at the
// chunk end), bits above the lane bits (delta constant over the wave), the fixed row.
template <int NB, int U>
__device__ __forceinline__ double grad_end_coef(int i, int c, uint64_t gend, uint64_t g0) {
  if (i >= U && i < c) return ((gend >> i) & 1ull) ? -1.0 : 1.0;
  if (i >= c + 7 && i < NB) return ((g0 >> i) & 1ull) ? -1.0 : 1.0;
  if (i == NB) return 1.0;
  return 0.0;
}

// Leave-one-out products over the group columns N-4..N-1: Q = prod of the other
// columns, prefix chain Q h_{N-4} ... and suffix products h_{k+1} ... h_{N-1} (8 complex
// multiplies besides Q; a two-level tree variant measured slower).
template <int N, int G>
__device__ __forceinline__ void grad_loo(const double (&hr)[N], const double (&hi)[N],
                                         double (&pr)[G], double (&pi)[G]) {
  constexpr int NO = N - G;
  cd q = TreeProd<0, NO>::run(hr, hi);
  cd suf[G];
  suf[G - 2] = {hr[N - 1], hi[N - 1]};
#pragma unroll
  for (int k = G - 3; k >= 0; --k) suf[k] = cmul({hr[NO + k + 1], hi[NO + k + 1]}, suf[k + 1]);
#pragma unroll
  for (int k = 0; k < G; ++k) {
    cd p = (k == G - 1) ? q : cmul(q, suf[k]);
    pr[k] = p.re;
    pi[k] = p.im;
    if (k + 1 < G) q = cmul(q, {hr[NO + k], hi[NO + k]});
  }
}

// Rows 0..B-1 in broadcast form (every static flip when B = U): no LDS read.
template <int N, int B>
struct GradRows {
  Bcast<N> r[B > 0 ? B : 1];
  double one;
};

// T[k] += SIGN * P_k for the 4 leave-one-out products, each product's last multiply fused
// into the accumulation (grad_loo followed by the adds, 6 instructions fewer).
template <int SIGN, int N, int G>
__device__ __forceinline__ void grad_loo_acc(const double (&hr)[N], const double (&hi)[N],
                                             double (&Tr)[G], double (&Ti)[G]) {
  constexpr int NO = N - G;
  cd q = TreeProd<0, NO>::run(hr, hi);
  cd suf[G];
  suf[G - 2] = {hr[N - 1], hi[N - 1]};
#pragma unroll
  for (int k = G - 3; k >= 0; --k) suf[k] = cmul({hr[NO + k + 1], hi[NO + k + 1]}, suf[k + 1]);
#pragma unroll
  for (int k = 0; k < G; ++k) {
    if (k == G - 1) {
      Tr[k] += SIGN * q.re;
      Ti[k] += SIGN * q.im;
    } else {
      cmul_acc<SIGN>(Tr[k], Ti[k], q, suf[k]);
    }
    if (k + 1 < G) q = cmul(q, {hr[NO + k], hi[NO + k]});
  }
}

// The 8 leave-one-out products of the group columns N-8..N-1 by a tree: pair, quad and
// half products of the group, Q (the other columns) times each half, those times the
// other pair of the half, and each product's last multiply fused into T[k] += SIGN * P_k.
// 12 complex multiplies + 8 fused (the prefix/suffix form: 13 + 7 fused + an add) on a
// critical path of 7 multiplies after Q instead of 12, so the two waves of a SIMD have
// more independent work to interleave: 1.3% faster at n = 20 and 1% at n = 22; above
// n = TB_GRAD_LOO_TREE_MAX_N its extra live values spill in the loop and it is slower.
#ifndef TB_GRAD_LOO_TREE_MAX_N
#define TB_GRAD_LOO_TREE_MAX_N 22
#endif
template <int N>
struct GradLooTree {
  cd p01, p23, p45, p67, c01, c23, c45, c67;
};
template <int N>
__device__ __forceinline__ GradLooTree<N> grad_loo_tree(const double (&hr)[N], const double (&hi)[N]) {
  constexpr int NO = N - 8;
  GradLooTree<N> t;
  const cd q = TreeProd<0, NO>::run(hr, hi);
  t.p01 = cmul({hr[NO + 0], hi[NO + 0]}, {hr[NO + 1], hi[NO + 1]});
  t.p23 = cmul({hr[NO + 2], hi[NO + 2]}, {hr[NO + 3], hi[NO + 3]});
  t.p45 = cmul({hr[NO + 4], hi[NO + 4]}, {hr[NO + 5], hi[NO + 5]});
  t.p67 = cmul({hr[NO + 6], hi[NO + 6]}, {hr[NO + 7], hi[NO + 7]});
  const cd a = cmul(q, cmul(t.p45, t.p67)), b = cmul(q, cmul(t.p01, t.p23));
  t.c01 = cmul(a, t.p23);
  t.c23 = cmul(a, t.p01);
  t.c45 = cmul(b, t.p67);
  t.c67 = cmul(b, t.p45);
  return t;
}
template <int SIGN, int N>
__device__ __forceinline__ void grad_loo_acc_tree(const double (&hr)[N], const double (&hi)[N],
                                                  double (&Tr)[8], double (&Ti)[8]) {
  constexpr int NO = N - 8;
  const GradLooTree<N> t = grad_loo_tree<N>(hr, hi);
  cmul_acc<SIGN>(Tr[0], Ti[0], t.c01, {hr[NO + 1], hi[NO + 1]});
  cmul_acc<SIGN>(Tr[1], Ti[1], t.c01, {hr[NO + 0], hi[NO + 0]});
  cmul_acc<SIGN>(Tr[2], Ti[2], t.c23, {hr[NO + 3], hi[NO + 3]});
  cmul_acc<SIGN>(Tr[3], Ti[3], t.c23, {hr[NO + 2], hi[NO + 2]});
  cmul_acc<SIGN>(Tr[4], Ti[4], t.c45, {hr[NO + 5], hi[NO + 5]});
  cmul_acc<SIGN>(Tr[5], Ti[5], t.c45, {hr[NO + 4], hi[NO + 4]});
  cmul_acc<SIGN>(Tr[6], Ti[6], t.c67, {hr[NO + 7], hi[NO + 7]});
  cmul_acc<SIGN>(Tr[7], Ti[7], t.c67, {hr[NO + 6], hi[NO + 6]});
}

// T[k] += SIGN * P_k through the form chosen for N and G.
template <int SIGN, int N, int G>
__device__ __forceinline__ void grad_step_acc(const double (&hr)[N], const double (&hi)[N],
                                              double (&Tr)[G], double (&Ti)[G]) {
  if constexpr (G == 8 && N <= TB_GRAD_LOO_TREE_MAX_N)
    grad_loo_acc_tree<SIGN, N>(hr, hi, Tr, Ti);
  else
    grad_loo_acc<SIGN, N, G>(hr, hi, Tr, Ti);
}

template <int N, int S, int MODE, int B>
__device__ __forceinline__ void grad_flip(double (&hr)[N], double (&hi)[N], const SRow<S, double> &sr,
                                          const GradRows<N, B> &gb, const cpx<double> *M, int i,
                                          double sg) {
  if (i < B) {  // compile-time at the static flip sites
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (u == i) flip_bcast<0, N, MODE>(hr, hi, gb.r[u], MODE == 0 ? sg : gb.one);
    return;
  }
  if (S > 0 && i == 0)
    flip_srow<N, S, MODE, double>(hr, hi, sr, M, sg);
  else
    flip_mem<N, MODE, double>(hr, hi, M + (size_t)i * N, sg);
}

// One wave per workgroup; blockIdx = item * wpi + w, item = matrix * ngroups + group;
// group g holds columns min(col0 + g G, N - G) onwards.
// partial[(item * wpi + w) * N * G + row * G + col]. G = 4 Q columns in Q slots of 4.
template <int N, int U, int G>
__global__ __launch_bounds__(64) TB_GRAD_ATTR void grad_kernel(const cd *__restrict__ A,
                                                               int64_t mstride, int ngroups,
                                                               int col0, int c, int wpi, int iters,
                                                               cd *__restrict__ partial) {
  constexpr int Q = G / kGradG;
  constexpr int NB = N - 1;
  constexpr int NO = N - G;
  // broadcast rows (register budget): with 4 columns all 3 static rows up to n = 22, 2 up
  // to 24, then 1 (the 4-bit blocks of n <= 19 keep row 0 in SGPRs); with 8 columns
  // TB_GRAD_G8_B of the static rows
  constexpr int B = (G == 8) ? (N > 24 ? 0 : TB_GRAD_G8_B < U ? TB_GRAD_G8_B : U)
                             : (U == 3) ? (N <= 22 ? 3 : N <= 24 ? 2 : 1) : 0;
  constexpr int S = B > 0 ? 0 : ((TB_GRAD_SROW < N) ? TB_GRAD_SROW : N);
  __shared__ cpx<double> Ms[N * N];
  const int lane = threadIdx.x;
  const int p = lane & 15;
  const int item = blockIdx.x / wpi;
  const int w = blockIdx.x - item * wpi;
  const int b = item / ngroups, g = item - b * ngroups;
  const int j0 = min(col0 + g * G, N - G);
  const cd *src = A + (size_t)b * mstride;
  for (int e = lane; e < N * N; e += 64) {
    const int i = e / N, jl = e - i * N;
    const int j = (jl >= NO) ? j0 + (jl - NO) : (jl < j0 ? jl : jl + G);
    const cd v = src[(size_t)i * N + j];
    Ms[e] = {v.re, v.im};
  }
  __syncthreads();
  const cpx<double> *M = Ms;
  SRow<S, double> sr;
  if constexpr (S > 0) load_srow(sr, M);
  GradRows<N, B> gb;
  if constexpr (B > 0) {
#pragma unroll
    for (int u = 0; u < B; ++u) load_bcast(gb.r[u], (const double *)(M + (size_t)u * N), lane);
    gb.one = vgpr_one();
    asm volatile("s_nop 4");  // VALU writes of the store before its first DPP read
  }
  GradAcc acc[Q];
#pragma unroll
  for (int qq = 0; qq < Q; ++qq) acc[qq] = {0.0, 0.0, 0.0, 0.0};
  const int nblocks = 1 << (c - U);
  for (int it = 0; it < iters; ++it) {
    const int wt = w + it * wpi;
    const bool q_odd = (wt & 1) != 0;
    const uint64_t q = (((uint64_t)(wt >> 1) * 64 + (uint64_t)lane) << 1) | (q_odd ? 1ull : 0ull);
    const uint64_t t0 = q << c;
    const uint64_t g0 = t0 ^ (t0 >> 1);
    // Half column sums at t0.
    double hr[N], hi[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      hr[j] = 0.5 * M[(size_t)NB * N + j].re;
      hi[j] = 0.5 * M[(size_t)NB * N + j].im;
    }
#pragma unroll 1
    for (int r = 0; r < NB; ++r) {
      const double d = ((g0 >> r) & 1ull) ? -0.5 : 0.5;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        hr[j] = fma(d, M[(size_t)r * N + j].re, hr[j]);
        hi[j] = fma(d, M[(size_t)r * N + j].im, hi[j]);
      }
    }
    double Tr[G], Ti[G];
    grad_loo<N, G>(hr, hi, Tr, Ti);  // prod(delta) = (-1)^t0 = +1 (t0 even)
    double Er[U][G], Ei[U][G];
#pragma unroll
    for (int i = 0; i < U; ++i)
#pragma unroll
      for (int k = 0; k < G; ++k) {
        Er[i][k] = 0.0;
        Ei[i][k] = 0.0;
      }
    // Static steps jj = 1 .. 2^U - 1 of block blk: bit i = ctz(jj) flips; its new value
    // is 1 ^ t_{i+1}. t_{i+1} == 0: delta_i +1 -> -1 (h -= row, E += 2T); else h += row,
    // E -= 2T.
    auto static_steps = [&](int blk) {
#pragma unroll
      for (int jj = 1; jj < (1 << U); ++jj) {
        const int i = __builtin_ctz(jj);
        if (i + 1 < U) {
          if ((jj >> (i + 1)) & 1) {
#pragma unroll
            for (int k = 0; k < G; ++k) {
              Er[i][k] -= 2.0 * Tr[k];
              Ei[i][k] -= 2.0 * Ti[k];
            }
            grad_flip<N, S, 1, B>(hr, hi, sr, gb, M, i, 1.0);
          } else {
#pragma unroll
            for (int k = 0; k < G; ++k) {
              Er[i][k] += 2.0 * Tr[k];
              Ei[i][k] += 2.0 * Ti[k];
            }
            grad_flip<N, S, -1, B>(hr, hi, sr, gb, M, i, -1.0);
          }
        } else {
          const bool t_next = (U < c) ? ((blk & 1) != 0) : q_odd;
          const double sg = t_next ? 1.0 : -1.0;
#pragma unroll
          for (int k = 0; k < G; ++k) {
            Er[i][k] = fma(-2.0 * sg, Tr[k], Er[i][k]);
            Ei[i][k] = fma(-2.0 * sg, Ti[k], Ei[i][k]);
          }
          grad_flip<N, S, 0, B>(hr, hi, sr, gb, M, i, sg);
        }
        if (jj & 1)
          grad_step_acc<-1, N, G>(hr, hi, Tr, Ti);
        else
          grad_step_acc<1, N, G>(hr, hi, Tr, Ti);
        TB_STEP_FENCE();
      }
    };
    static_steps(0);
    for (int blk = 1; blk < nblocks; ++blk) {
      const int i = U + __builtin_ctz(blk);
      const bool t_next = (i + 1 < c) ? (((blk >> (i + 1 - U)) & 1) != 0) : q_odd;
      const double sg = t_next ? 1.0 : -1.0;
#pragma unroll
      for (int qq = 0; qq < Q; ++qq) {  // wave-level corrector of bit i
        double zr, zi;
        wave_sum_cols(Tr + kGradG * qq, Ti + kGradG * qq, zr, zi);
        grad_acc_row(acc[qq], p, i, -2.0 * sg, zr, zi);
      }
      flip_mem<N, 0, double>(hr, hi, M + (size_t)i * N, sg);
      grad_step_acc<1, N, G>(hr, hi, Tr, Ti);
      TB_STEP_FENCE();
      static_steps(blk);
    }
    // End of the chunk.
    const uint64_t tend = t0 + ((uint64_t)1 << c) - 1;
    const uint64_t gend = tend ^ (tend >> 1);
#pragma unroll
    for (int i = 0; i < U; ++i) {  // static rows: E + delta(end) T per lane
      const double d = ((gend >> i) & 1ull) ? -1.0 : 1.0;
      double xr[G], xi[G];
#pragma unroll
      for (int k = 0; k < G; ++k) {
        xr[k] = fma(d, Tr[k], Er[i][k]);
        xi[k] = fma(d, Ti[k], Ei[i][k]);
      }
#pragma unroll
      for (int qq = 0; qq < Q; ++qq) {
        double zr, zi;
        wave_sum_cols(xr + kGradG * qq, xi + kGradG * qq, zr, zi);
        grad_acc_row(acc[qq], p, i, 1.0, zr, zi);
      }
    }
    {  // rows whose delta is wave-uniform at the end: in-chunk bits >= U, bits above the
       // lane bits, and the fixed row N-1
      const double f0 = grad_end_coef<NB, U>(p, c, gend, g0);
      const double f1 = grad_end_coef<NB, U>(16 + p, c, gend, g0);
#pragma unroll
      for (int qq = 0; qq < Q; ++qq) {
        double zr, zi;
        wave_sum_cols(Tr + kGradG * qq, Ti + kGradG * qq, zr, zi);
        acc[qq].r0 = fma(f0, zr, acc[qq].r0);
        acc[qq].i0 = fma(f0, zi, acc[qq].i0);
        acc[qq].r1 = fma(f1, zr, acc[qq].r1);
        acc[qq].i1 = fma(f1, zi, acc[qq].i1);
      }
    }
    // Rows c .. c+6: delta_i depends on the lane (chunk index bits 1..6).
    const int iend = min(c + 7, NB);
#pragma unroll 1
    for (int i = c; i < iend; ++i) {
      const double d = ((g0 >> i) & 1ull) ? -1.0 : 1.0;
      double xr[G], xi[G];
#pragma unroll
      for (int k = 0; k < G; ++k) {
        xr[k] = d * Tr[k];
        xi[k] = d * Ti[k];
      }
#pragma unroll
      for (int qq = 0; qq < Q; ++qq) {
        double zr, zi;
        wave_sum_cols(xr + kGradG * qq, xi + kGradG * qq, zr, zi);
        grad_acc_row(acc[qq], p, i, 1.0, zr, zi);
      }
    }
  }
  cd *dst = partial + ((size_t)item * wpi + w) * N * G;
#pragma unroll
  for (int qq = 0; qq < Q; ++qq) {
    const int col = kGradG * qq + grad_row_col(lane >> 4);
    if (p < N) dst[(size_t)p * G + col] = {acc[qq].r0, acc[qq].i0};
    if (16 + p < N) dst[(size_t)(16 + p) * G + col] = {acc[qq].r1, acc[qq].i1};
  }
}

// grad[b][i][j] = sum_w partial[(b, owner group of j), w][i][j - j0], fixed order. With a
// tail launch (tail = 1), columns from G ngroups on come from its 4-column groups, stored
// after the G-column ones.
__global__ void grad_reduce_kernel(const cd *__restrict__ partial, int n, int G, int ngroups,
                                   int tail, int wpi, int batch, cd *__restrict__ grad) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)batch * n * n) return;
  const int b = (int)(idx / ((int64_t)n * n));
  const int rem = (int)(idx - (int64_t)b * n * n);
  const int i = rem / n, j = rem - i * n;
  const cd *src;
  int stride;
  if (tail && j >= G * ngroups) {
    stride = n * kGradG;
    src = partial + (size_t)batch * ngroups * wpi * n * G + ((size_t)b * wpi) * stride +
          (size_t)i * kGradG + (j - (n - kGradG));
  } else {
    const int g = min(j / G, ngroups - 1);
    const int k = j - min(g * G, n - G);
    stride = n * G;
    src = partial + ((size_t)(b * ngroups + g) * wpi) * n * G + (size_t)i * G + k;
  }
  double re = 0.0, im = 0.0;
#pragma unroll 8
  for (int w = 0; w < wpi; ++w) {  // unrolled: the loads are in flight together
    const cd v = src[(size_t)w * stride];
    re += v.re;
    im += v.im;
  }
  grad[idx] = {re, im};
}

}  // namespace tb

namespace tbgr {

struct GradPlan {
  int c, wpi, iters, ngroups, tail;  // tail: one 4-column group per matrix after ngroups
};

// Chunks of 2^c ranks. Built in: c shrinks (not above 11) until the batch provides
// ~8192 waves, then grows again by up to 2 while the rounds stay full (below); a tuned c
// (c_pref > 0) replaces that choice. Always c >= U and the chunk
// index keeps >= 7 bits (6 lane bits and the wave's bit 0).
inline GradPlan grad_plan(int n, int batch, int c_pref) {
  GradPlan p{};
  const int G = tb::grad_g_of(n);
  p.tail = tb::grad_tail4(n) ? 1 : 0;
  p.ngroups = p.tail ? n / G : (n + G - 1) / G;
  const int nb = n - 1;
  const int64_t items = (int64_t)batch * (p.ngroups + p.tail);
  int c;
  if (c_pref > 0) {
    c = c_pref;
  } else {
    int want_bits = 0;
    while ((items << want_bits) < 8192 && want_bits < 30) ++want_bits;
    c = std::min(nb - 6 - want_bits, 11);
  }
  const int u4 = n <= TB_GRAD_U4_MAX_N ? 4 : 3;
  c = std::max(c, G == 8 ? (p.tail ? std::max(TB_GRAD_G8_U, u4) : TB_GRAD_G8_U) : u4);
  c = std::min(c, nb - 7);
  if (c_pref <= 0) {
    // Up to two sizes larger while the main launch's rounds of 2 waves per SIMD stay as
    // full, scored by the fill of its last round times the share of a chunk left after
    // seeding the column sums (4 n (n - 1) flops against 2^c steps of 8n + 18G - 30):
    // n = 20, 64 matrices: chunks of 2^9 (2048 waves), 5% faster than 2^7 (8192).
    double best = -1.0;
    int best_c = c;
    for (int cc = c; cc <= std::min({c + 2, 11, nb - 7}); ++cc) {
      const double waves = (double)batch * p.ngroups * std::ldexp(1.0, nb - cc - 6);
      const double rounds = waves / (2.0 * 1024.0);
      const double fill = rounds / std::ceil(rounds);
      const double seed = 4.0 * n * (n - 1) / ((8.0 * n + 18.0 * G - 30.0) * std::ldexp(1.0, cc));
      const double score = fill * (1.0 - seed);
      if (score > best + 1e-9) {
        best = score;
        best_c = cc;
      }
    }
    c = best_c;
  }
  p.c = c;
  const int task_bits = nb - c - 6;  // >= 1
  const int wpi_bits = std::min(task_bits, 13);
  p.wpi = 1 << wpi_bits;
  p.iters = 1 << (task_bits - wpi_bits);
  return p;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

bool grad_supported(int n) { return n >= tb::kGradMinN && n <= tb::kGradMaxN; }

size_t grad_partial_bytes(int n, int batch, int c_pref) {
  if (!grad_supported(n) || batch <= 0) return 0;
  const GradPlan p = grad_plan(n, batch, c_pref);
  return align256((size_t)batch * p.wpi * n *
                  (p.ngroups * tb::grad_g_of(n) + p.tail * tb::kGradG) * sizeof(tb::cd));
}

template <int N = tb::kGradMinN>
int32_t dispatch_grad(int n, const tb::cd *A, int64_t mstride, int batch, const GradPlan &p,
                      tb::cd *partial, hipStream_t s) {
  if constexpr (N > tb::kGradMaxN) {
    return tbh::fail(TB_ERANGE, "grad: no single-sweep kernel for n=%d", n);
  } else {
    if (n != N) return dispatch_grad<N + 1>(n, A, mstride, batch, p, partial, s);
    constexpr int G = tb::grad_g<N>();
    const unsigned grid = (unsigned)((int64_t)batch * p.ngroups * p.wpi);
    hipLaunchKernelGGL((tb::grad_kernel<N, tb::grad_u<N>(), G>), dim3(grid), dim3(64), 0, s, A,
                       mstride, p.ngroups, 0, p.c, p.wpi, p.iters, partial);
    int32_t rc = tbh::check_hip(hipGetLastError(), "grad_kernel");
    if (rc || !p.tail) return rc;
    if constexpr (G == 8 && N % 8 != 0 && N % 8 <= 4) {
      // the last N % 8 columns: one 4-column group per matrix (columns N-4..N-1)
      tb::cd *tail_partial = partial + (size_t)batch * p.ngroups * p.wpi * N * G;
      hipLaunchKernelGGL((tb::grad_kernel<N, tb::grad_u_of<N, tb::kGradG>(), tb::kGradG>),
                         dim3((unsigned)((int64_t)batch * p.wpi)), dim3(64), 0, s, A, mstride,
                         1, N - tb::kGradG, p.c, p.wpi, p.iters, tail_partial);
      return tbh::check_hip(hipGetLastError(), "grad_kernel (4-column tail)");
    } else {
      return tbh::fail(TB_EINVAL, "grad: no 4-column tail for n=%d", n);
    }
  }
}

// All n^2 minors of each matrix into grad[b][i][j]; partial holds grad_partial_bytes.
int32_t run_grad(int n, const tb::cd *A, int64_t mstride, int batch, int c_pref, tb::cd *partial,
                 tb::cd *grad, hipStream_t s) {
  const GradPlan p = grad_plan(n, batch, c_pref);
  int32_t rc = dispatch_grad(n, A, mstride, batch, p, partial, s);
  if (rc) return rc;
  const int64_t total = (int64_t)batch * n * n;
  hipLaunchKernelGGL(tb::grad_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     s, partial, n, tb::grad_g_of(n), p.ngroups, p.tail, p.wpi, batch, grad);
  return tbh::check_hip(hipGetLastError(), "grad_reduce_kernel");
}

}  // namespace tbgr
