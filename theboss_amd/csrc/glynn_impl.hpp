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
template <int N, int U, int R0, typename F>
__global__ __launch_bounds__(256) TB_WIDE_ATTR(N) void glynn_wide_kernel(const cpx<F> *__restrict__ A,
                                                         int64_t mstride, int C, int wpm, int iters,
                                                         cd *__restrict__ partial) {
  __shared__ cpx<F> Ms[N * N];
  const int lane = threadIdx.x & 63;
  const int wgs_per_matrix = wpm >> 2;
  const int b = blockIdx.x / wgs_per_matrix;
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
inline GlynnPlan glynn_plan(int n, int batch, int c_pref, int u = 2) {
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
  C = std::min(C, kMaxC);
  C = std::max(C, u);
  C = std::min(C, nb - 8);
  p.C = C;
  const int wave_bits = nb - p.C - 6;  // log2(waves needed at one chunk per lane), >= 2
  const int wpm_bits = std::min(wave_bits, kMaxWpmBits);
  p.wpm = 1 << wpm_bits;
  p.iters = 1 << (wave_bits - wpm_bits);
  return p;
}

template <int N, typename F>
inline int32_t launch_glynn_n(const tb::cpx<F> *A, int64_t mstride, int batch, tb::cd *out,
                              tb::cd *partial, int c_pref, hipStream_t s) {
  if constexpr (N < kWideMinN) {
    constexpr int U = (N - 1) < 2 ? (N - 1) : 2;
    hipLaunchKernelGGL((tb::glynn_narrow_kernel<N, U, F>), dim3((batch + 255) / 256), dim3(256), 0,
                       s, A, mstride, batch, out);
    (void)partial;
    (void)c_pref;
    return tbh::check_hip(hipGetLastError(), "glynn_narrow_kernel");
  } else {
    constexpr int U = glynn_u<N, F>();
    static_assert(U <= N - 9, "the static block must fit the smallest chunk (C <= n - 9)");
    const GlynnPlan p = glynn_plan(N, batch, c_pref, U);
    const unsigned blocks = (unsigned)batch * (unsigned)(p.wpm / 4);
    hipLaunchKernelGGL((tb::glynn_wide_kernel<N, U, glynn_reg_rows<N, F>(), F>), dim3(blocks), dim3(256), 0, s,
                       A, mstride, p.C, p.wpm, p.iters, partial);
    int32_t rc = tbh::check_hip(hipGetLastError(), "glynn_wide_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(tb::glynn_finish_kernel, dim3((batch + 3) / 4), dim3(256), 0, s, partial,
                       p.wpm, batch, out);
    return tbh::check_hip(hipGetLastError(), "glynn_finish_kernel");
  }
}

template <typename F, int N = 2>
inline int32_t dispatch_glynn(int n, const tb::cpx<F> *A, int64_t mstride, int batch, tb::cd *out,
                              tb::cd *partial, int c_pref, hipStream_t s) {
  if constexpr (N > TB_GLYNN_MAX_N) {
    return tbh::fail(TB_ERANGE, "glynn: n=%d exceeds TB_GLYNN_MAX_N=%d", n, TB_GLYNN_MAX_N);
  } else {
    if (n == N) return launch_glynn_n<N, F>(A, mstride, batch, out, partial, c_pref, s);
    return dispatch_glynn<F, N + 1>(n, A, mstride, batch, out, partial, c_pref, s);
  }
}

int32_t glynn_dispatch_f64(int n, const tb::cpx<double> *A, int64_t mstride, int batch,
                           tb::cd *out, tb::cd *partial, int c_pref, hipStream_t s);
int32_t glynn_dispatch_f32(int n, const tb::cpx<float> *A, int64_t mstride, int batch,
                           tb::cd *out, tb::cd *partial, int c_pref, hipStream_t s);

}  // namespace tbgl
