#pragma once
// Glynn permanents with a reflected Gray code, gfx950 (kernels and launch templates;
// instantiated by glynn_f64.hip and glynn_f32.hip).
//
// Restates GlynnGrayPermanentCalculator.compute_permanent
// (theboss/boson_sampling_utilities/permanent_calculators/glynn_gray_permanent_calculator.py:41-84):
//   perm(M) = 2^-(n-1) * sum_{delta in {+-1}^{n-1} x {+1}} (prod_i delta_i) prod_j sum_i delta_i M[i][j]
// enumerated in Gray order (one delta flip per term).
//
// Work split: the 2^(n-1) Gray ranks are cut into aligned chunks of 2^C ranks. One lane
// owns one chunk: it seeds its column sums at the chunk start in O(n^2) and then walks
// the chunk with O(n) updates per term. Because chunks are aligned, the row flipped at
// local step l is ctz(l) for every lane, so in WIDE mode (all 64 lanes of a wave on one
// matrix) the flipped row is wave-uniform. FP64 keeps the most frequently flipped rows
// (all static ones) in broadcast-form VGPRs (one DPP v_fmac_f64 per value, no LDS read)
// and reads the others from LDS as a broadcast; FP32 holds row 0 in VGPRs.
//
// Half column sums h_j = s_j / 2 are carried so a flip is a plain add/sub of M[i][j];
// perm = 2 * sum_t sign(t) prod_j h_j(t). sign(t) = (-1)^(popcount gray(t)) = (-1)^(t & 1).
#include <algorithm>

#include "flip_common.hpp"

namespace tb {

// FP64 wide kernels keep rows 0..U+1 in broadcast form up to this n, rows 0..U-1 above.
#ifndef TB_GLYNN_BCAST_WIDE_MAX_N
#define TB_GLYNN_BCAST_WIDE_MAX_N 24
#endif

// Occupancy target for the wide kernels: 2 waves per SIMD (<= 256 VGPRs).
#ifndef TB_WIDE_ATTR
#define TB_WIDE_ATTR(N) __attribute__((amdgpu_waves_per_eu((N) <= 40 ? 2 : 1)))
#endif

// Values of row 0 held in SGPRs by the WIDE kernels that keep no row in VGPRs (20
// complex values = 80 SGPRs; more spills SGPRs).
#ifndef TB_GLYNN_SROW
#define TB_GLYNN_SROW 20
#endif

// Rows 0..R0-1 (the most frequently flipped: row i flips on 2^-(i+1) of all steps) are
// held in VGPRs; the rest are read from memory at each flip.
template <int N, int R0, typename F>
struct RegRows {
  F re[R0 > 0 ? R0 : 1][N], im[R0 > 0 ? R0 : 1][N];
};

// Rows 0..B-1 in broadcast form (FP64 only): every static flip and 3/4 of the
// block-start flips read no LDS.
template <int N, int B>
struct BcastRows {
  Bcast<N> r[B > 0 ? B : 1];
  double one;
};

template <int N, int R0, int S, int MODE, typename F, typename P, int B = 0>
__device__ __forceinline__ void flip_row(F (&hr)[N], F (&hi)[N], const RegRows<N, R0, F> &rr,
                                         const SRow<S, F> &sr, P M, int i, F sg,
                                         const BcastRows<N, B> *bs = nullptr) {
  if constexpr (B > 0) {
    if (i < B) {  // compile-time at the static flip sites, uniform at block starts
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (u == i) flip_bcast<0, N, MODE>(hr, hi, bs->r[u], MODE == 0 ? sg : bs->one);
      return;
    }
  }
  if constexpr (S > 0) {
    if (i == 0) {
      flip_srow<N, S, MODE, F>(hr, hi, sr, M, sg);
      return;
    }
  }
  if constexpr (R0 > 0) {
    if (i < R0) {
      // i is a compile-time constant at every static flip site after unrolling.
#pragma unroll
      for (int r = 0; r < R0; ++r)
        if (r == i) flip_reg<N, MODE, F>(hr, hi, rr.re[r], rr.im[r], sg);
      return;
    }
  }
  flip_mem<N, MODE, F>(hr, hi, M + (size_t)i * N, sg);
}

// One lane walks `iters` chunks q = q0 + it*qstride of 2^C Gray ranks each; M is the
// n x n matrix as complex rows. `topbit_fixed` means C == N-1 (the whole enumeration in
// one chunk), so the Gray bit above the chunk is 0.
template <int N, int U, int R0, typename F, typename P, int S = 0, int B = 0>
__device__ __forceinline__ void glynn_lane(P M, int C, uint64_t q0, uint64_t qstride, int iters,
                                           double &acc_re, double &acc_im) {
  constexpr int NB = N - 1;  // Gray bits
  const bool topbit_fixed = (C == NB);
  SRow<S, F> sr;
  if constexpr (S > 0) load_srow(sr, M);
  BcastRows<N, B> bs;
  if constexpr (B > 0) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int u = 0; u < B; ++u) load_bcast(bs.r[u], (const double *)(M + (size_t)u * N), lane);
    bs.one = vgpr_one();
    asm volatile("s_nop 4");  // VALU writes of the store before its first DPP read
  }
  RegRows<N, R0, F> rr;
  if constexpr (R0 > 0) {
#pragma unroll
    for (int r = 0; r < R0; ++r)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const cpx<F> v = M[(size_t)r * N + j];
        rr.re[r][j] = v.re;
        rr.im[r][j] = v.im;
      }
  }
  for (int it = 0; it < iters; ++it) {
    const uint64_t q = q0 + (uint64_t)it * qstride;
    const uint64_t t0 = q << C;
    const uint64_t g0 = t0 ^ (t0 >> 1);
    F hr[N], hi[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const cpx<F> v = M[(size_t)(N - 1) * N + j];
      hr[j] = v.re;
      hi[j] = v.im;
    }
#pragma unroll 1
    for (int i = 0; i < NB; ++i) {
      const F sg = ((g0 >> i) & 1ull) ? F(-1) : F(1);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const cpx<F> v = M[(size_t)i * N + j];
        hr[j] = fma(sg, v.re, hr[j]);
        hi[j] = fma(sg, v.im, hi[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      hr[j] *= F(0.5);
      hi[j] *= F(0.5);
    }
    const bool q_odd = (q & 1ull) != 0;
    double cre, cim;
    {
      cpx<F> p = TreeProdT<F, 0, N>::run(hr, hi);
      cre = (double)p.re;
      cim = (double)p.im;
    }
    const int nblocks = 1 << (C - U);
    // Static steps j = 1 .. 2^U - 1 of block blk (row ctz(j) < U flips).
    auto static_steps = [&](int blk) {
#pragma unroll
      for (int j = 1; j < (1 << U); ++j) {
        const int i = __builtin_ctz(j);
        if (i + 1 < U) {
          // Direction fixed by the local bit i+1 of j.
          if ((j >> (i + 1)) & 1)
            flip_row<N, R0, S, 1, F, P, B>(hr, hi, rr, sr, M, i, F(1), &bs);
          else
            flip_row<N, R0, S, -1, F, P, B>(hr, hi, rr, sr, M, i, F(-1), &bs);
        } else {
          bool add;
          if (U < C)
            add = (blk & 1) != 0;
          else
            add = topbit_fixed ? false : q_odd;
          flip_row<N, R0, S, 0, F, P, B>(hr, hi, rr, sr, M, i, add ? F(1) : F(-1), &bs);
        }
        if constexpr (sizeof(F) == 8) {  // last multiply fused into the accumulation
          if (j & 1)
            tree_prod_acc<-1, 0, N>(cre, cim, (const double *)hr, (const double *)hi);
          else
            tree_prod_acc<1, 0, N>(cre, cim, (const double *)hr, (const double *)hi);
        } else {
          cpx<F> p = TreeProdT<F, 0, N>::run(hr, hi);
          if (j & 1) {
            cre -= (double)p.re;
            cim -= (double)p.im;
          } else {
            cre += (double)p.re;
            cim += (double)p.im;
          }
        }
        TB_STEP_FENCE();
      }
    };
    // Block 0 is peeled so the loop body is branch-free (no phi copies of h).
    static_steps(0);
    for (int blk = 1; blk < nblocks; ++blk) {
      const int i = U + __builtin_ctz(blk);
      bool add;
      if (i + 1 < C)
        add = ((blk >> (i + 1 - U)) & 1) != 0;
      else
        add = topbit_fixed ? false : q_odd;
      if constexpr (B > U) {
        if (i < B)
          flip_row<N, 0, 0, 0, F, P, B>(hr, hi, rr, sr, M, i, add ? F(1) : F(-1), &bs);
        else
          flip_mem<N, 0, F>(hr, hi, M + (size_t)i * N, add ? F(1) : F(-1));
      } else {
        flip_mem<N, 0, F>(hr, hi, M + (size_t)i * N, add ? F(1) : F(-1));
      }
      if constexpr (sizeof(F) == 8) {
        tree_prod_acc<1, 0, N>(cre, cim, (const double *)hr, (const double *)hi);
      } else {
        cpx<F> p = TreeProdT<F, 0, N>::run(hr, hi);
        cre += (double)p.re;
        cim += (double)p.im;
      }
      TB_STEP_FENCE();
      static_steps(blk);
    }
    acc_re += cre;
    acc_im += cim;
  }
}

// WIDE: the 4 waves of a workgroup share one matrix, staged once into LDS; every flip is
// a broadcast LDS read. Grid: batch * (wpm/4) workgroups; wave wq of matrix b covers
// chunks q = (wq + it*wpm)*64 + lane. Strides count complex elements.
// `handled` (may be null): matrices whose partials the paired kernel already wrote.
template <int N, int U, int R0, typename F>
__global__ __launch_bounds__(256) TB_WIDE_ATTR(N) void glynn_wide_kernel(const cpx<F> *__restrict__ A,
                                                         int64_t mstride, int C, int wpm, int iters,
                                                         cd *__restrict__ partial,
                                                         const int *__restrict__ handled) {
  __shared__ cpx<F> Ms[N * N];
  const int lane = threadIdx.x & 63;
  const int wgs_per_matrix = wpm >> 2;
  const int b = blockIdx.x / wgs_per_matrix;
  if (handled != nullptr && handled[b] != 0) return;
  const int wq = (blockIdx.x - b * wgs_per_matrix) * 4 + (threadIdx.x >> 6);
  const cpx<F> *src = A + (size_t)b * mstride;
  for (int e = threadIdx.x; e < N * N; e += 256) Ms[e] = src[e];
  __syncthreads();
  double are = 0.0, aim = 0.0;
  // R0 = 0: row 0's first values in SGPRs; 1: row 0 in VGPRs; 2 (FP64): rows 0..U+1
  // (n <= TB_GLYNN_BCAST_WIDE_MAX_N) or 0..U-1 in broadcast form; 3: rows 0..U-1 (every
  // static flip) in VGPRs.
  constexpr int S = (R0 == 0) ? TB_GLYNN_SROW : 0;
  constexpr int BW = (N <= TB_GLYNN_BCAST_WIDE_MAX_N) ? U + 2 : U;  // rows in broadcast form
  constexpr int B = (R0 == 2) ? (BW < N - 1 ? BW : N - 1) : 0;
  constexpr int RR = (R0 == 2) ? 0 : (R0 == 3 ? U : R0);  // rows in plain VGPRs
  glynn_lane<N, U, RR, F, const cpx<F> *, (S < N ? S : N), B>(
      (const cpx<F> *)Ms, C, (uint64_t)wq * 64 + lane, (uint64_t)wpm * 64, iters, are, aim);
  are = wave_sum(are);
  aim = wave_sum(aim);
  if (lane == 0) partial[(size_t)b * wpm + wq] = {are, aim};
}

// ------------------------------------------------------------------------------------
// Paired Gray walk (FP64 WIDE). Gray codes 2t and 2t+1 differ only in row 0. With row
// N-1 fixed (delta = +1) and u_j = sum_{i >= 1} delta_i M[i][j] / M[0][j], their half
// column sums are h_j = (M[0][j] / 2)(u_j + 1) and (M[0][j] / 2)(u_j - 1), so
//   perm = 2 K sum_t (-1)^t [prod_j (u_j + 1) - prod_j (u_j - 1)],  K = prod_j M[0][j] / 2,
// over the 2^(N-2) Gray ranks t of rows 1..N-2. A flip of row i adds +-2 M[i][j] / M[0][j]
// to every u_j (one update for both codes), and a factor (u_j +- 1) enters a running
// product x as x u_j +- x: four FMAs with the +-1 folded into their addends. A pair of
// codes costs 2N flip instructions and 2 (4N - 2) product instructions (each product two
// chains of N/2 factors, joined by a multiply fused into the accumulation) against
// 2 (6N - 4) for two plain steps. The row whose smallest entry is largest becomes row 0
// (a row permutation leaves the permanent unchanged); a matrix none of whose rows has all
// entries above 2^-16 of its largest entry (zeros included) is left to the plain kernel.
// ------------------------------------------------------------------------------------

// acc += SIGN * (prod (u + 1) - prod (u - 1)).
template <int SIGN, int N>
__device__ __forceinline__ void pair_acc(double &ar, double &ai, const double (&ur)[N],
                                         const double (&ui)[N]) {
  constexpr int H = N / 2;
  const cd ap = fold_chain<1, 0, H>(ur, ui), bp = fold_chain<1, H, N>(ur, ui);
  const cd am = fold_chain<-1, 0, H>(ur, ui), bm = fold_chain<-1, H, N>(ur, ui);
  cmul_acc<SIGN>(ar, ai, ap, bp);
  cmul_acc<-SIGN>(ar, ai, am, bm);
}

// u += sg * row i + 1 of S2 (rows 1..B in broadcast form, the others from LDS).
template <int N, int MODE, int B>
__device__ __forceinline__ void pair_flip(double (&ur)[N], double (&ui)[N], const BcastRows<N, B> &bs,
                                          const cpx<double> *S2, int i, double sg) {
  if constexpr (B > 0) {
    if (i < B) {  // compile-time at the static flip sites, uniform at block starts
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (u == i) flip_bcast<0, N, MODE>(ur, ui, bs.r[u], MODE == 0 ? sg : bs.one);
      return;
    }
  }
  flip_mem<N, MODE, double>(ur, ui, S2 + (size_t)(i + 1) * N, sg);
}

// One lane walks `iters` chunks q = q0 + it*qstride of 2^C super-steps (Gray ranks of rows
// 1..N-2) each; S2 holds 2 M[i][j] / M[0][j] in rows 1..N-1. Same chunk, block and
// direction structure as glynn_lane with row i + 1 at Gray bit i.
template <int N, int U, int B>
__device__ __forceinline__ void glynn_pair_lane(const cpx<double> *S2, int C, uint64_t q0,
                                                uint64_t qstride, int iters, double &acc_re,
                                                double &acc_im) {
  constexpr int NB = N - 2;  // Gray bits
  const bool topbit_fixed = (C == NB);
  BcastRows<N, B> bs;
  if constexpr (B > 0) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int u = 0; u < B; ++u)
      load_bcast(bs.r[u], (const double *)(S2 + (size_t)(u + 1) * N), lane);
    bs.one = vgpr_one();
    asm volatile("s_nop 4");  // VALU writes of the store before its first DPP read
  }
  for (int it = 0; it < iters; ++it) {
    const uint64_t q = q0 + (uint64_t)it * qstride;
    const uint64_t t0 = q << C;
    const uint64_t g0 = t0 ^ (t0 >> 1);
    double ur[N], ui[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const cpx<double> v = S2[(size_t)(N - 1) * N + j];
      ur[j] = v.re;
      ui[j] = v.im;
    }
#pragma unroll 1
    for (int i = 0; i < NB; ++i) {
      const double sg = ((g0 >> i) & 1ull) ? -1.0 : 1.0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const cpx<double> v = S2[(size_t)(i + 1) * N + j];
        ur[j] = fma(sg, v.re, ur[j]);
        ui[j] = fma(sg, v.im, ui[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      ur[j] *= 0.5;
      ui[j] *= 0.5;
    }
    const bool q_odd = (q & 1ull) != 0;
    double cre = 0.0, cim = 0.0;
    pair_acc<1>(cre, cim, ur, ui);
    const int nblocks = 1 << (C - U);
    auto static_steps = [&](int blk) {
#pragma unroll
      for (int j = 1; j < (1 << U); ++j) {
        const int i = __builtin_ctz(j);
        if (i + 1 < U) {
          if ((j >> (i + 1)) & 1)
            pair_flip<N, 1, B>(ur, ui, bs, S2, i, 1.0);
          else
            pair_flip<N, -1, B>(ur, ui, bs, S2, i, -1.0);
        } else {
          bool add;
          if (U < C)
            add = (blk & 1) != 0;
          else
            add = topbit_fixed ? false : q_odd;
          pair_flip<N, 0, B>(ur, ui, bs, S2, i, add ? 1.0 : -1.0);
        }
        if (j & 1)
          pair_acc<-1>(cre, cim, ur, ui);
        else
          pair_acc<1>(cre, cim, ur, ui);
        TB_STEP_FENCE();
      }
    };
    static_steps(0);
    for (int blk = 1; blk < nblocks; ++blk) {
      const int i = U + __builtin_ctz(blk);
      bool add;
      if (i + 1 < C)
        add = ((blk >> (i + 1 - U)) & 1) != 0;
      else
        add = topbit_fixed ? false : q_odd;
      if constexpr (B > U) {
        if (i < B)
          pair_flip<N, 0, B>(ur, ui, bs, S2, i, add ? 1.0 : -1.0);
        else
          flip_mem<N, 0, double>(ur, ui, S2 + (size_t)(i + 1) * N, add ? 1.0 : -1.0);
      } else {
        flip_mem<N, 0, double>(ur, ui, S2 + (size_t)(i + 1) * N, add ? 1.0 : -1.0);
      }
      pair_acc<1>(cre, cim, ur, ui);
      TB_STEP_FENCE();
      static_steps(blk);
    }
    acc_re += cre;
    acc_im += cim;
  }
}

// PAIRED WIDE kernel: grid, chunks and partial layout of glynn_wide_kernel with C Gray
// bits per chunk of codes (C - 1 super-step bits). Writes handled[b] (1: this kernel wrote
// matrix b's partials; 0: left to the plain kernel).
template <int N, int U, int B>
__global__ __launch_bounds__(256) TB_WIDE_ATTR(N) void glynn_pair_kernel(
    const cpx<double> *__restrict__ A, int64_t mstride, int C, int wpm, int iters,
    cd *__restrict__ partial, int *__restrict__ handled) {
  __shared__ cpx<double> Ms[N * N];
  __shared__ cd inv0[N];
  __shared__ double rmin[N], rmax[N];
  __shared__ int pick, Ke;
  __shared__ cd Km;
  const int lane = threadIdx.x & 63;
  const int wgs_per_matrix = wpm >> 2;
  const int b = blockIdx.x / wgs_per_matrix;
  const int wg = blockIdx.x - b * wgs_per_matrix;
  const int wq = wg * 4 + (threadIdx.x >> 6);
  const cpx<double> *src = A + (size_t)b * mstride;
  for (int e = threadIdx.x; e < N * N; e += 256) Ms[e] = src[e];
  __syncthreads();
  if (threadIdx.x < N) {
    double lo = INFINITY, hi = 0.0;
    for (int j = 0; j < N; ++j) {
      const cpx<double> v = Ms[threadIdx.x * N + j];
      const double a = fma(v.re, v.re, v.im * v.im);
      lo = fmin(lo, a);
      hi = fmax(hi, a);
    }
    rmin[threadIdx.x] = lo;
    rmax[threadIdx.x] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = 0;
    double mx = 0.0;
    for (int i = 0; i < N; ++i) {
      if (rmin[i] > rmin[best]) best = i;
      mx = fmax(mx, rmax[i]);
    }
    const bool ok = mx > 0.0 && mx <= 1.7976931348623157e308 && rmin[best] >= kPairMinRel2 * mx;
    pick = ok ? best : -1;
    if (wg == 0) handled[b] = ok ? 1 : 0;
  }
  __syncthreads();
  const int r = pick;
  if (r < 0) return;  // uniform over the workgroup
  if (r != 0 && threadIdx.x < N) {
    const cpx<double> t = Ms[threadIdx.x];
    Ms[threadIdx.x] = Ms[r * N + threadIdx.x];
    Ms[r * N + threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x < N) inv0[threadIdx.x] = cinv_scaled(cd{Ms[threadIdx.x].re, Ms[threadIdx.x].im});
  if (threadIdx.x == 64) {  // K = prod_j M[0][j] / 2 as mantissa x 2^Ke
    cd k;
    int ke;
    prod_half_scaled((const cd *)Ms, N, k, ke);
    Km = k;
    Ke = ke;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < (N - 1) * N; e += 256) {
    const int i = 1 + e / N, j = e - (i - 1) * N;
    const cd v = cmul(cd{Ms[i * N + j].re, Ms[i * N + j].im}, inv0[j]);
    Ms[i * N + j] = {2.0 * v.re, 2.0 * v.im};
  }
  __syncthreads();
  double are = 0.0, aim = 0.0;
  glynn_pair_lane<N, U, B>(Ms, C - 1, (uint64_t)wq * 64 + lane, (uint64_t)wpm * 64, iters, are, aim);
  are = wave_sum(are);
  aim = wave_sum(aim);
  if (lane == 0) {
    const cd k = Km;
    partial[(size_t)b * wpm + wq] = {ldexp(k.re * are - k.im * aim, Ke),
                                     ldexp(k.re * aim + k.im * are, Ke)};
  }
}

// ------------------------------------------------------------------------------------
// Paired walk with FP32 column sums and products (14 <= n <= TB_FP32_COLSUM_MAX_N): the
// walk of glynn_pair_lane in float, the quotients 2 M[i][j] / M[0][j] formed in FP64 and
// rounded once; static row 1 in SGPRs (SR = 1), rows SR + 1..U in plain VGPRs. Each pair's
// difference prod(u + 1) - prod(u - 1) is formed in FP32, summed in FP32 over the 2^U
// pairs of a block and accumulated in FP64 once per block.
// ------------------------------------------------------------------------------------
// Complex values per LDS read group of the paired FP32 kernel's memory flips (two
// groups in flight).
#ifndef TB_GLYNN_F32_PAIR_FLIP_G
#define TB_GLYNN_F32_PAIR_FLIP_G 4
#endif
// One step of the chain x <- x (u + PM), split in its two FMA levels.
template <int PM>
__device__ __forceinline__ void chain_lvl1(float xr, float xi, float u_r, float u_i, float &tr,
                                           float &ti) {
  tr = fmaf(xr, u_r, (float)PM * xr);
  ti = fmaf(xr, u_i, (float)PM * xi);
}
__device__ __forceinline__ void chain_lvl2(float &xr, float &xi, float u_r, float u_i, float tr,
                                           float ti) {
  const float nr = fmaf(-xi, u_i, tr);
  xi = fmaf(xi, u_r, ti);
  xr = nr;
}

// SIGN * (Π(u + 1) - Π(u - 1)) added to the FP32 block sums (ar, ai). Four chains run side by side, Π(u ± 1)
// over columns [0, H) and [H, N): a chain step is two dependent FMA levels, and FP32
// issues fast enough that one chain alone waits on its own results; with the four
// interleaved level by level, each FMA's operand was produced eight instructions earlier.
template <int SIGN, int N>
__device__ __forceinline__ void pair_acc_f32(float &ar, float &ai, const float (&ur)[N],
                                             const float (&ui)[N]) {
  constexpr int H = N / 2;  // the upper half is the longer one for odd N
  float pr = ur[0] + 1.f, pi = ui[0];  // (u + 1), columns [0, H)
  float qr = ur[H] + 1.f, qi = ui[H];  // (u + 1), columns [H, N)
  float mr = ur[0] - 1.f, mi = ui[0];  // (u - 1), columns [0, H)
  float wr = ur[H] - 1.f, wi = ui[H];  // (u - 1), columns [H, N)
#pragma unroll
  for (int k = 1; k < N - H; ++k) {
    float t[8];
    if (k < H) {
      chain_lvl1<1>(pr, pi, ur[k], ui[k], t[0], t[1]);
      chain_lvl1<-1>(mr, mi, ur[k], ui[k], t[4], t[5]);
    }
    chain_lvl1<1>(qr, qi, ur[H + k], ui[H + k], t[2], t[3]);
    chain_lvl1<-1>(wr, wi, ur[H + k], ui[H + k], t[6], t[7]);
    if (k < H) {
      chain_lvl2(pr, pi, ur[k], ui[k], t[0], t[1]);
      chain_lvl2(mr, mi, ur[k], ui[k], t[4], t[5]);
    }
    chain_lvl2(qr, qi, ur[H + k], ui[H + k], t[2], t[3]);
    chain_lvl2(wr, wi, ur[H + k], ui[H + k], t[6], t[7]);
  }
  // (p q) - (m w) with the difference folded into the FMAs
  const float mwr = fmaf(-mi, wi, mr * wr), mwi = fmaf(mi, wr, mr * wi);
  const float dr = fmaf(pr, qr, -fmaf(pi, qi, mwr));
  const float di = fmaf(pr, qi, fmaf(pi, qr, -mwi));
  if constexpr (SIGN > 0) {
    ar += dr;
    ai += di;
  } else {
    ar -= dr;
    ai -= di;
  }
}

// h += v, h -= v or h += sg v for the static row whose first S values sit in SGPRs (s)
// and the rest in VGPRs (rr, ri; entries below S unused).
template <int N, int MODE, int S>
__device__ __forceinline__ void flip_split_f32(float (&hr)[N], float (&hi)[N],
                                               const SRow<S, float> &s, const float (&rr)[N],
                                               const float (&ri)[N], float sg) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const float vr = j < S ? s.re[j < S ? j : 0] : rr[j];
    const float vi = j < S ? s.im[j < S ? j : 0] : ri[j];
    if constexpr (MODE == 1) {
      hr[j] += vr;
      hi[j] += vi;
    } else if constexpr (MODE == -1) {
      hr[j] -= vr;
      hi[j] -= vi;
    } else {
      hr[j] = fmaf(sg, vr, hr[j]);
      hi[j] = fmaf(sg, vi, hi[j]);
    }
  }
}

template <int N, int U, int SR, int S2P>
__device__ __forceinline__ void glynn_pair_lane_f32(const cpx<float> *S2, int C, uint64_t q0,
                                                    uint64_t qstride, int iters, double &acc_re,
                                                    double &acc_im) {
  constexpr int NB = N - 2;  // Gray bits: row i + 1 at bit i
  const bool topbit_fixed = (C == NB);
  // Static row SR (row 1 flips on half of all pairs, row 2 on a quarter) in SGPRs when
  // SR > 0, as the scalar operand of the adds (an FP32 instruction with an SGPR operand
  // does not dual-issue, so the register it frees costs issue rate); the other static
  // rows in VGPRs, except the first S2P values of row 2 beside an SGPR row 1.
  static_assert(SR <= U, "the SGPR row is a static row");
  static_assert(S2P == 0 || (SR == 1 && U >= 2), "row 2 splits only beside an SGPR row 1");
  constexpr int NR = U - (SR ? 1 : 0);  // static rows in VGPRs
  SRow<SR ? N : 0, float> sr;
  if constexpr (SR) load_srow<N>(sr, S2 + (size_t)SR * N);
  SRow<S2P, float> sr2;
  if constexpr (S2P > 0) load_srow<S2P>(sr2, S2 + 2 * N);
  RegRows<N, NR, float> rr;  // rr[r]: static row index r, skipping SR - 1
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (r == 0 && j < S2P) continue;
      const int i = (SR && r >= SR - 1) ? r + 1 : r;
      const cpx<float> v = S2[(size_t)(i + 1) * N + j];
      rr.re[r][j] = v.re;
      rr.im[r][j] = v.im;
    }
  for (int it = 0; it < iters; ++it) {
    const uint64_t q = q0 + (uint64_t)it * qstride;
    const uint64_t t0 = q << C;
    const uint64_t g0 = t0 ^ (t0 >> 1);
    float ur[N], ui[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const cpx<float> v = S2[(size_t)(N - 1) * N + j];
      ur[j] = v.re;
      ui[j] = v.im;
    }
#pragma unroll 1
    for (int i = 0; i < NB; ++i) {
      const float sg = ((g0 >> i) & 1ull) ? -1.0f : 1.0f;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const cpx<float> v = S2[(size_t)(i + 1) * N + j];
        ur[j] = fmaf(sg, v.re, ur[j]);
        ui[j] = fmaf(sg, v.im, ui[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      ur[j] *= 0.5f;
      ui[j] *= 0.5f;
    }
    const bool q_odd = (q & 1ull) != 0;
    // Each block's pair terms (the block-start pair and the 2^U - 1 static ones) are
    // summed in FP32, then added to the FP64 lane sums: one conversion and one FP64 add
    // per block and component.
    float bre = 0.f, bim = 0.f;
    pair_acc_f32<1>(bre, bim, ur, ui);
    const int nblocks = 1 << (C - U);
    // row i + 1: mode +1 / -1 (fixed direction) or 0 (direction sg); every call site
    // passes constants, so the branches fold away after unrolling
    auto do_flip = [&](int i, int mode, float sg) {
      if (SR && i == SR - 1) {
        if constexpr (SR) {
          if (mode == 1)
            flip_reg<N, 1, float>(ur, ui, sr.re, sr.im, sg);
          else if (mode == -1)
            flip_reg<N, -1, float>(ur, ui, sr.re, sr.im, sg);
          else
            flip_reg<N, 0, float>(ur, ui, sr.re, sr.im, sg);
        }
        return;
      }
      if (S2P > 0 && i == 1) {
        if constexpr (S2P > 0) {
          if (mode == 1)
            flip_split_f32<N, 1, S2P>(ur, ui, sr2, rr.re[0], rr.im[0], sg);
          else if (mode == -1)
            flip_split_f32<N, -1, S2P>(ur, ui, sr2, rr.re[0], rr.im[0], sg);
          else
            flip_split_f32<N, 0, S2P>(ur, ui, sr2, rr.re[0], rr.im[0], sg);
        }
        return;
      }
      if (i < U) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
          if (((SR && r >= SR - 1) ? r + 1 : r) == i) {
            if (mode == 1)
              flip_reg<N, 1, float>(ur, ui, rr.re[r], rr.im[r], sg);
            else if (mode == -1)
              flip_reg<N, -1, float>(ur, ui, rr.re[r], rr.im[r], sg);
            else
              flip_reg<N, 0, float>(ur, ui, rr.re[r], rr.im[r], sg);
          }
        return;
      }
      flip_mem<N, 0, float, const cpx<float> *, 0, TB_GLYNN_F32_PAIR_FLIP_G>(ur, ui, S2 + (size_t)(i + 1) * N, sg);
    };
    auto static_steps = [&](int blk) {
#pragma unroll
      for (int j = 1; j < (1 << U); ++j) {
        const int i = __builtin_ctz(j);
        if (i + 1 < U) {
          if ((j >> (i + 1)) & 1)
            do_flip(i, 1, 1.0f);
          else
            do_flip(i, -1, -1.0f);
        } else {
          bool add;
          if (U < C)
            add = (blk & 1) != 0;
          else
            add = topbit_fixed ? false : q_odd;
          do_flip(i, 0, add ? 1.0f : -1.0f);
        }
        if (j & 1)
          pair_acc_f32<-1>(bre, bim, ur, ui);
        else
          pair_acc_f32<1>(bre, bim, ur, ui);
        TB_STEP_FENCE();
      }
      acc_re += (double)bre;
      acc_im += (double)bim;
      bre = 0.f;
      bim = 0.f;
    };
    static_steps(0);
    for (int blk = 1; blk < nblocks; ++blk) {
      const int i = U + __builtin_ctz(blk);
      bool add;
      if (i + 1 < C)
        add = ((blk >> (i + 1 - U)) & 1) != 0;
      else
        add = topbit_fixed ? false : q_odd;
      flip_mem<N, 0, float, const cpx<float> *, 0, TB_GLYNN_F32_PAIR_FLIP_G>(ur, ui, S2 + (size_t)(i + 1) * N, add ? 1.0f : -1.0f);
      pair_acc_f32<1>(bre, bim, ur, ui);
      TB_STEP_FENCE();
      static_steps(blk);
    }
  }
}

// PAIRED WIDE kernel, FP32 column sums: reads the FP64 matrices (the quotients and the
// row-0 product are formed in FP64), writes partials and handled[] like glynn_pair_kernel.
template <int N, int U, int SR, int S2P, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void glynn_pair_kernel_f32(
    const cd *__restrict__ A, int64_t mstride, int C, int wpm, int iters,
    cd *__restrict__ partial, int *__restrict__ handled) {
  __shared__ cd Ms[N * N];
  __shared__ cpx<float> Sf[N * N];
  __shared__ cd inv0[N];
  __shared__ double rmin[N], rmax[N];
  __shared__ int pick, Ke;
  __shared__ cd Km;
  const int lane = threadIdx.x & 63;
  const int wgs_per_matrix = wpm >> 2;
  const int b = blockIdx.x / wgs_per_matrix;
  const int wg = blockIdx.x - b * wgs_per_matrix;
  const int wq = wg * 4 + (threadIdx.x >> 6);
  const cd *src = A + (size_t)b * mstride;
  for (int e = threadIdx.x; e < N * N; e += 256) Ms[e] = src[e];
  __syncthreads();
  if (threadIdx.x < N) {
    double lo = INFINITY, hi = 0.0;
    for (int j = 0; j < N; ++j) {
      const cd v = Ms[threadIdx.x * N + j];
      const double a = fma(v.re, v.re, v.im * v.im);
      lo = fmin(lo, a);
      hi = fmax(hi, a);
    }
    rmin[threadIdx.x] = lo;
    rmax[threadIdx.x] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = 0;
    double mx = 0.0;
    for (int i = 0; i < N; ++i) {
      if (rmin[i] > rmin[best]) best = i;
      mx = fmax(mx, rmax[i]);
    }
    const bool ok = mx > 0.0 && mx <= 1.7976931348623157e308 && rmin[best] >= kPairMinRel2 * mx;
    pick = ok ? best : -1;
    if (wg == 0) handled[b] = ok ? 1 : 0;
  }
  __syncthreads();
  const int r = pick;
  if (r < 0) return;  // uniform over the workgroup
  if (r != 0 && threadIdx.x < N) {
    const cd t = Ms[threadIdx.x];
    Ms[threadIdx.x] = Ms[r * N + threadIdx.x];
    Ms[r * N + threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x < N) inv0[threadIdx.x] = cinv_scaled(Ms[threadIdx.x]);
  if (threadIdx.x == 64) {
    cd k;
    int ke;
    prod_half_scaled(Ms, N, k, ke);
    Km = k;
    Ke = ke;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < (N - 1) * N; e += 256) {
    const int i = 1 + e / N, j = e - (i - 1) * N;
    const cd v = cmul(Ms[i * N + j], inv0[j]);
    Sf[i * N + j] = {(float)(2.0 * v.re), (float)(2.0 * v.im)};
  }
  __syncthreads();
  double are = 0.0, aim = 0.0;
  glynn_pair_lane_f32<N, U, SR, S2P>(Sf, C - 1, (uint64_t)wq * 64 + lane, (uint64_t)wpm * 64, iters, are, aim);
  are = wave_sum(are);
  aim = wave_sum(aim);
  if (lane == 0) {
    const cd k = Km;
    partial[(size_t)b * wpm + wq] = {ldexp(k.re * are - k.im * aim, Ke),
                                     ldexp(k.re * aim + k.im * are, Ke)};
  }
}

// NARROW: one lane per matrix, whole enumeration in the lane (small n), global rows.
template <int N, int U, typename F>
__global__ __launch_bounds__(256) void glynn_narrow_kernel(const cpx<F> *__restrict__ A,
                                                           int64_t mstride, int batch,
                                                           cd *__restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const cpx<F> *M = A + (size_t)b * mstride;
  double are = 0.0, aim = 0.0;
  glynn_lane<N, U, 0, F>(M, N - 1, 0, 0, 1, are, aim);
  out[b] = {2.0 * are, 2.0 * aim};
}

namespace {  // internal linkage: each translation unit keeps its own copy

// Sum the wpm wave partials of each matrix in a fixed order: deterministic results.
__global__ __launch_bounds__(256) void glynn_finish_kernel(const cd *__restrict__ partial, int wpm,
                                                           int batch, cd *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= batch) return;
  double re = 0.0, im = 0.0;
  for (int w = lane; w < wpm; w += 64) {
    re += partial[(size_t)b * wpm + w].re;
    im += partial[(size_t)b * wpm + w].im;
  }
  re = wave_sum(re);
  im = wave_sum(im);
  if (lane == 0) out[b] = {2.0 * re, 2.0 * im};
}

__global__ void fill_one_kernel(cd *out, int batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < batch) out[b] = {1.0, 0.0};
}

__global__ void copy_scalar_kernel(const cd *A, int64_t mstride, int batch, cd *out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < batch) out[b] = A[(size_t)b * mstride];
}

// FP32 copy of the matrices (row store for the FP32 column-sum variant).
__global__ void to_fp32_kernel(const cd *A, int64_t mstride, int n, int batch, float *dst) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)n * n;
  if (idx >= per * batch) return;
  const int64_t b = idx / per, e = idx - b * per;
  const cd v = A[b * mstride + e];
  dst[2 * idx] = (float)v.re;
  dst[2 * idx + 1] = (float)v.im;
}

}  // namespace

}  // namespace tb

// ------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------
namespace tbgl {

// Rows held in VGPRs by the WIDE kernel: row 0 (flipped on half of all steps) when the
// register budget of 2 waves per SIMD allows it -- FP32 always, FP64 up to n = 24.
#ifndef TB_GLYNN_F64_R0_MAX_N
#define TB_GLYNN_F64_R0_MAX_N 24
#endif
// Up to n = 40 the broadcast rows stay in VGPRs; above, the compiler parks them in AGPRs
// and copies them back right before the DPP reads (a VALU write the DPP read must not
// follow within 2 wait states), so larger n keep row 0 in SGPRs instead.
#ifndef TB_GLYNN_F64_BCAST_MAX_N
#define TB_GLYNN_F64_BCAST_MAX_N 40
#endif
// FP32 rows cost half the registers: up to TB_GLYNN_F32_ROWS_MAX_N the FP32 kernels keep
// every static row (0..U-1) in VGPRs, so only block-start flips read LDS.
#ifndef TB_GLYNN_F32_ROWS_MAX_N
#define TB_GLYNN_F32_ROWS_MAX_N 24
#endif
template <int N, typename F>
constexpr int glynn_reg_rows() {
  if (sizeof(F) == 8 && N <= TB_GLYNN_F64_BCAST_MAX_N) return 2;
  if (sizeof(F) == 4) return N <= TB_GLYNN_F32_ROWS_MAX_N ? 3 : 1;
  return N <= TB_GLYNN_F64_R0_MAX_N ? 1 : 0;
}

// Static Gray bits (unrolled block of 2^U steps) of the WIDE kernels: FP64 at n >= 25,
// whose static rows are all in broadcast form, unrolls TB_GLYNN_U_LARGE bits.
#ifndef TB_GLYNN_U_LARGE
#define TB_GLYNN_U_LARGE 3
#endif
// Static Gray bits of the FP32 kernels that keep their static rows in VGPRs: 3 (7/8 of
// the flips from registers; measured 1.05x at n = 20, 1.11x at n = 23 over row 0 alone).
// The plan keeps at least 2^8 chunks per matrix (C <= n - 9) and needs C >= U, so U = 3
// starts at n = 12.
#ifndef TB_GLYNN_F32_U
#define TB_GLYNN_F32_U 3
#endif
// The paired FP32 kernel: 2 static Gray bits, static row SR in SGPRs (row 1 by default)
// and the other in VGPRs. A SIMD issues two FP32 instructions per 4-cycle slot only from
// two different waves, so the kernel wants 4 resident waves (<= 128 VGPRs): row 1 in
// SGPRs and, at n = 19..21, the first values of row 2 too (glynn_f32_pair_s2) give 4 waves
// up to n = 21 and 3 at n = 22, 23. 3 static bits with every row in VGPRs need 217 VGPRs
// (2 waves) at n = 20. SR = 2 (row 2 in SGPRs, so the row-1 flips dual-issue) fits 4
// waves at every n and measured the same time at n = 20 and 5% slower at n = 23: at
// n = 20 the launch runs the clock down to 1.9 GHz whichever rows dual-issue.
#ifndef TB_GLYNN_F32_PAIR_U
#define TB_GLYNN_F32_PAIR_U 2
#endif
#ifndef TB_GLYNN_F32_PAIR_SR
#define TB_GLYNN_F32_PAIR_SR 1
#endif
#ifndef TB_GLYNN_F32_PAIR_WPE
#define TB_GLYNN_F32_PAIR_WPE 3
#endif
template <int N, typename F>
constexpr int glynn_u() {
  if (sizeof(F) == 4)
    return (N <= TB_GLYNN_F32_ROWS_MAX_N && N - 9 >= TB_GLYNN_F32_U) ? TB_GLYNN_F32_U : 2;
  return N >= 25 ? TB_GLYNN_U_LARGE : 2;
}
inline int glynn_u_max(int n) {
  const int u32 = (n <= TB_GLYNN_F32_ROWS_MAX_N && n - 9 >= TB_GLYNN_F32_U) ? TB_GLYNN_F32_U : 2;
  return std::max({2, n >= 25 ? TB_GLYNN_U_LARGE : 2, u32});
}

constexpr int kWideMinN = 11;  // below this a lane owns a whole matrix

// FP64 sizes that run the paired walk first (the plain WIDE kernel then takes the
// matrices it left): from where a chunk holds the static block plus one bit up to the
// largest n that fits 256 VGPRs without spills (n = 40 spills).
#ifndef TB_GLYNN_PAIR_MIN_N
#define TB_GLYNN_PAIR_MIN_N 14
#endif
#ifndef TB_GLYNN_PAIR_MAX_N
#define TB_GLYNN_PAIR_MAX_N 39
#endif
template <int N, typename F>
constexpr bool glynn_paired() {
  if constexpr (sizeof(F) == 8) return N >= TB_GLYNN_PAIR_MIN_N && N <= TB_GLYNN_PAIR_MAX_N;
  // FP32 column sums: where the FP32 kernel is in use (larger n, run only to measure its
  // error, keep the plain FP32 walk)
  return N >= TB_GLYNN_PAIR_MIN_N && N <= TB_FP32_COLSUM_MAX_N;
}
// Static super-step bits of the paired walk and its rows in broadcast form (rows 1..B).
// Static super-step bits: 3 at every paired n (a block of 8 pairs between block-start
// flips; 166..256 VGPRs, no spills). Measured against 2 bits below n = 25: 3.7% faster at
// n = 16 and 3.6-4% at n = 20..24.
template <int N>
constexpr int glynn_pair_u() {
  return 3;
}
template <int N>
constexpr int glynn_pair_b() {
  constexpr int b = (N <= TB_GLYNN_BCAST_WIDE_MAX_N) ? glynn_pair_u<N>() + 2 : glynn_pair_u<N>();
  return b < N - 2 ? b : N - 2;
}
// Lower bound on C of a launch of n: the static bits, and for the paired walk its static
// super-step bits plus the pair bit.
template <int N, typename F>
constexpr int glynn_pair_u_of() {
  return sizeof(F) == 8 ? glynn_pair_u<N>() : TB_GLYNN_F32_PAIR_U;
}
// Values of row 2 the paired FP32 kernel also keeps in SGPRs, so that it fits 4 waves
// per SIMD (<= 128 VGPRs): about 6n + 17 VGPRs without them, 2 fewer per value; up to 11
// values (row 1 takes 2n SGPRs; n = 23 would need 14 and exceed the 102 SGPRs).
#ifndef TB_GLYNN_F32_PAIR_S2_MAX
#define TB_GLYNN_F32_PAIR_S2_MAX 11
#endif
template <int N>
constexpr int glynn_f32_pair_s2() {
  if (TB_GLYNN_F32_PAIR_SR != 1 || TB_GLYNN_F32_PAIR_U < 2) return 0;
  constexpr int over = 6 * N + 17 - 128;
  constexpr int need = over <= 0 ? 0 : (over + 1) / 2 + 1;  // one value of margin
  return (over <= 0 || need > TB_GLYNN_F32_PAIR_S2_MAX) ? 0 : need;
}
// Waves per SIMD the launch is resident at (the plan's round size).
template <int N, typename F>
constexpr int glynn_wave_slots() {
  if constexpr (sizeof(F) == 4 && glynn_paired<N, F>()) {
    if constexpr (TB_GLYNN_F32_PAIR_SR == 1)
      return (6 * N + 17 <= 128 || glynn_f32_pair_s2<N>() > 0) ? 4 : TB_GLYNN_F32_PAIR_WPE;
    if constexpr (TB_GLYNN_F32_PAIR_SR == 2) return 4;
    return TB_GLYNN_F32_PAIR_WPE;
  }
  return 2;
}
template <int N, typename F>
constexpr int glynn_min_c() {
  constexpr int u = glynn_u<N, F>();
  if constexpr (glynn_paired<N, F>())
    return u > glynn_pair_u_of<N, F>() + 1 ? u : glynn_pair_u_of<N, F>() + 1;
  return u;
}
constexpr int kMaxC = 12;      // Gray bits per lane chunk
constexpr int kMaxWpmBits = 13;  // waves per matrix before lanes iterate over chunks

struct GlynnPlan {
  int C, wpm, iters;
};

// WIDE plan: chunks of 2^C ranks. Built in: C shrinks (down to 7) until the batch
// provides at least ~8192 waves (4 rounds of 2 waves on every SIMD), so small batches of
// mid-size matrices still fill the GPU; C <= 12 bounds a lane's walk between re-seeds.
// A tuned C (c_pref > 0) replaces that choice. Always: C >= U and a matrix gets at
// least 4 waves (one workgroup). Re-seeding costs n^2 per 2^C steps.
// Kernels resident at more than 2 waves per SIMD (wave_slots) compare the built-in C with
// C - 1 and C - 2: the waves run in rounds of wave_slots x 1024 SIMDs (256 CUs), and a
// smaller chunk can fill the last round better (16384 waves are 5.33 rounds of 3072 but
// 10.67 rounds at twice the waves), against twice the re-seeding per code.
constexpr int kPlanSimds = 1024;
inline int glynn_plan_clamp(int C, int nb, int u) {
  return std::min(std::max(std::min(C, kMaxC), u), nb - 8);
}
// grow: how many chunk sizes above the ~8192-wave choice a 2-wave kernel may try (the
// FP64 kernels: 2; the FP32 ones keep their chunks, whose length bounds how far their FP32
// column sums drift between re-seeds).
inline GlynnPlan glynn_plan(int n, int batch, int c_pref, int u = 2, int wave_slots = 2,
                            int grow = 0) {
  GlynnPlan p{};
  const int nb = n - 1;
  int C;
  if (c_pref > 0) {
    C = c_pref;
  } else {
    int want_bits = 0;  // log2(waves per matrix needed for ~8192 waves in total)
    while (((int64_t)batch << want_bits) < 8192 && want_bits < 30) ++want_bits;
    C = nb - 6 - want_bits;
    C = std::max(C, std::min(7, nb - 8));
  }
  C = glynn_plan_clamp(C, nb, u);
  if (c_pref <= 0 && (wave_slots > 2 || grow > 0)) {
    // Score chunk sizes near C by how full their last round of wave_slots x kPlanSimds
    // waves is, times the share of a chunk left after seeding its column sums
    // ((n - 2) 2 / (5 2^c) of its instructions). Kernels of 3-4 waves per SIMD try
    // smaller chunks (more waves to fill their rounds); the FP64 kernels (2 waves) try up
    // to `grow` larger ones (less seeding while the rounds stay full: 2-10% faster for
    // batches of up to a few hundred matrices at n = 18..26, never slower in the batches
    // measured). Ties keep the candidate tried first.
    const int step = wave_slots > 2 ? -1 : 1;
    const int count = wave_slots > 2 ? 3 : 1 + grow;
    double best = -1.0;
    int best_c = C;
    for (int k = 0, c = C; k < count; ++k, c += step) {
      if (glynn_plan_clamp(c, nb, u) != c) continue;
      const double waves = (double)batch * std::ldexp(1.0, nb - c - 6);
      const double rounds = waves / ((double)kPlanSimds * wave_slots);
      const double fill = rounds / std::ceil(rounds);
      const double seed = (double)(n - 2) * 2.0 / (5.0 * std::ldexp(1.0, c));
      const double score = fill * (1.0 - seed);
      if (score > best + 1e-9) {
        best = score;
        best_c = c;
      }
    }
    C = best_c;
  }
  p.C = C;
  const int wave_bits = nb - p.C - 6;  // log2(waves needed at one chunk per lane), >= 2
  const int wpm_bits = std::min(wave_bits, kMaxWpmBits);
  p.wpm = 1 << wpm_bits;
  p.iters = 1 << (wave_bits - wpm_bits);
  return p;
}

// A: the matrices in the kernel's precision; A64 / mstride64: the same matrices in FP64
// (the paired FP32 kernel forms its quotients from them).
// wpm_cap: partials per matrix the workspace holds; a plan that needs more fails with
// TB_EWORKSPACE instead of writing past them.
template <int N, typename F>
inline int32_t launch_glynn_n(const tb::cpx<F> *A, int64_t mstride, const tb::cd *A64,
                              int64_t mstride64, int batch, tb::cd *out, tb::cd *partial,
                              int *handled, int wpm_cap, int c_pref, hipStream_t s) {
  if constexpr (N < kWideMinN) {
    constexpr int U = (N - 1) < 2 ? (N - 1) : 2;
    hipLaunchKernelGGL((tb::glynn_narrow_kernel<N, U, F>), dim3((batch + 255) / 256), dim3(256), 0,
                       s, A, mstride, batch, out);
    (void)partial;
    (void)handled;
    (void)wpm_cap;
    (void)c_pref;
    (void)A64;
    (void)mstride64;
    return tbh::check_hip(hipGetLastError(), "glynn_narrow_kernel");
  } else {
    constexpr int U = glynn_u<N, F>();
    constexpr int CMIN = glynn_min_c<N, F>();
    static_assert(CMIN <= N - 9, "the static block must fit the smallest chunk (C <= n - 9)");
    const GlynnPlan p =
        glynn_plan(N, batch, c_pref, CMIN, glynn_wave_slots<N, F>(), sizeof(F) == 8 ? 2 : 0);
    if (p.wpm > wpm_cap)
      return tbh::fail(TB_EWORKSPACE,
                       "glynn: the plan for n=%d batch=%d needs %d partials per matrix, the "
                       "workspace holds %d", N, batch, p.wpm, wpm_cap);
    const unsigned blocks = (unsigned)batch * (unsigned)(p.wpm / 4);
    const int *skip = nullptr;
    if constexpr (glynn_paired<N, F>()) {
      if constexpr (sizeof(F) == 8)
        hipLaunchKernelGGL((tb::glynn_pair_kernel<N, glynn_pair_u<N>(), glynn_pair_b<N>()>),
                           dim3(blocks), dim3(256), 0, s, (const tb::cpx<double> *)A64, mstride64,
                           p.C, p.wpm, p.iters, partial, handled);
      else
        hipLaunchKernelGGL((tb::glynn_pair_kernel_f32<N, TB_GLYNN_F32_PAIR_U, TB_GLYNN_F32_PAIR_SR,
                                                      glynn_f32_pair_s2<N>(),
                                                      glynn_wave_slots<N, F>()>),
                           dim3(blocks), dim3(256),
                           0, s, A64, mstride64, p.C, p.wpm, p.iters, partial, handled);
      int32_t rc = tbh::check_hip(hipGetLastError(), "glynn_pair_kernel");
      if (rc) return rc;
      skip = handled;
    } else {
      (void)handled;
      (void)A64;
      (void)mstride64;
    }
    hipLaunchKernelGGL((tb::glynn_wide_kernel<N, U, glynn_reg_rows<N, F>(), F>), dim3(blocks), dim3(256), 0, s,
                       A, mstride, p.C, p.wpm, p.iters, partial, skip);
    int32_t rc = tbh::check_hip(hipGetLastError(), "glynn_wide_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(tb::glynn_finish_kernel, dim3((batch + 3) / 4), dim3(256), 0, s, partial,
                       p.wpm, batch, out);
    return tbh::check_hip(hipGetLastError(), "glynn_finish_kernel");
  }
}

template <typename F, int N = 2>
inline int32_t dispatch_glynn(int n, const tb::cpx<F> *A, int64_t mstride, const tb::cd *A64,
                              int64_t mstride64, int batch, tb::cd *out, tb::cd *partial,
                              int *handled, int wpm_cap, int c_pref, hipStream_t s) {
  if constexpr (N > TB_GLYNN_MAX_N) {
    return tbh::fail(TB_ERANGE, "glynn: n=%d exceeds TB_GLYNN_MAX_N=%d", n, TB_GLYNN_MAX_N);
  } else {
    if (n == N)
      return launch_glynn_n<N, F>(A, mstride, A64, mstride64, batch, out, partial, handled,
                                  wpm_cap, c_pref, s);
    return dispatch_glynn<F, N + 1>(n, A, mstride, A64, mstride64, batch, out, partial, handled,
                                    wpm_cap, c_pref, s);
  }
}

// `handled`: batch ints of workspace for the paired walks; `wpm_cap`: partials per matrix
// after them. The FP32 dispatch takes the FP32 copy (A) and the FP64 originals (A64).
int32_t glynn_dispatch_f64(int n, const tb::cpx<double> *A, int64_t mstride, int batch,
                           tb::cd *out, tb::cd *partial, int *handled, int wpm_cap, int c_pref,
                           hipStream_t s);
int32_t glynn_dispatch_f32(int n, const tb::cpx<float> *A, int64_t mstride, const tb::cd *A64,
                           int64_t mstride64, int batch, tb::cd *out, tb::cd *partial,
                           int *handled, int wpm_cap, int c_pref, hipStream_t s);

}  // namespace tbgl
