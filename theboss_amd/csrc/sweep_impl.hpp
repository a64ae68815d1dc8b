#pragma once
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
// The paired sweep keeps 2 waves per SIMD at every k it runs at: at one wave the compiler
// would park broadcast columns in AGPRs (a DPP hazard), and FP64 issue at one wave is
// latency-bound.
#ifndef TB_SWEEP_PAIR_ATTR
#define TB_SWEEP_PAIR_ATTR(K) __attribute__((amdgpu_waves_per_eu(2)))
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

// One wave's share of an item whose columns are in Vs: chunks q = (w + it * wpi) * 64 + lane
// for it < iters. Lane i returns the sum over those chunks of minor i (in it order).
template <int K, int U, int R0>
__device__ __forceinline__ cd sweep_wave(const cd *Vs, cd *Ed, int lane, int c, int w, int wpi,
                                         int iters) {
  constexpr int R = K - 1;
  const uint64_t nchunks = (uint64_t)1 << (K - 1 - c);
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
  return acc;
}

// ------------------------------------------------------------------------------------
// Paired sweep (kSweepPairMinK <= k <= TB_SWEEP_PAIR_MAX_K). Gray bit 0 is particle 1;
// codes 2t and 2t+1 differ only in it. With the columns of every particle nu != 1 divided
// by particle 1's (S2[nu][j] = 2 V[nu][j] / V[1][j], u_j = (1/2) sum_{nu != 1} delta_nu
// S2[nu][j], delta_0 = +1) the two codes' half sums are (V[1][j]/2)(u_j +- 1), so with
// X = sum_t (-1)^t prod(u + 1), Y = sum_t (-1)^t prod(u - 1) over the Gray ranks t of
// particles 2..k-1 and Kc = prod_j V[1][j] / 2:
//   minor_0 = Kc (X - Y), minor_1 = Kc (X + Y),
//   minor_{b+2} = Kc (E_b + delta_{b+2}(end) (X - Y))
// with the per-bit correctors E_b of the plain sweep kept on T = X - Y. Per pair of codes:
// 2R flip instructions, 2 (4R - 2) for the products and 4 for the corrector, 10R in all
// against 2 (6R - 2) for two plain steps (R = k - 1).
// ------------------------------------------------------------------------------------
constexpr int kSweepPairMinK = 8;

// Particles 2..NU+1 in broadcast form: the static ones and the two most frequent
// block-start ones up to k = TB_SWEEP_PAIR_NU5_MAX_K, the static ones alone above (their
// VGPRs go to the longer column sums; with five the kernel spills from k = 22). Uniform
// loss m = 80, n = 40: 4.8% more samples/s than five up to k = 26 and the plain walk above.
#ifndef TB_SWEEP_PAIR_NU5_MAX_K
#define TB_SWEEP_PAIR_NU5_MAX_K 21
#endif
template <int R>
struct ColP {
  static constexpr int NUW = (R + 1 <= TB_SWEEP_PAIR_NU5_MAX_K) ? kSweepU + 2 : kSweepU;
  static constexpr int NU = (NUW < R - 1) ? NUW : R - 1;
  Bcast<R> b[NU];
  double one;
};

template <int R, int MODE>
__device__ __forceinline__ void sweep_pair_flip(double (&ur)[R], double (&ui)[R], const cd *S2,
                                                const ColP<R> &cp, int b, double sg) {
  if (b < ColP<R>::NU) {  // compile-time at the static flip sites, uniform at block starts
#pragma unroll
    for (int u = 0; u < ColP<R>::NU; ++u)
      if (u == b) flip_bcast<0, R, MODE>(ur, ui, cp.b[u], MODE == 0 ? sg : cp.one);
    return;
  }
  sweep_flip<R, MODE>(ur, ui, S2 + (b + 2) * R, sg);
}

// One chunk of 2^c super-steps (Gray bits of particles 2..k-1) starting at t0 = q << c.
// mi[nu] receives the chunk's sums without the factor Kc.
template <int K, int U>
__device__ __forceinline__ void sweep_pair_chunk(const cd *S2, const ColP<K - 1> &cp, cd *Ed, int c,
                                                 uint64_t q, cd (&mi)[K]) {
  constexpr int R = K - 1;
  constexpr int NB = K - 2;  // Gray bits: bit b <-> particle b + 2
  const uint64_t t0 = q << c;
  const uint64_t g0 = t0 ^ (t0 >> 1);
  double ur[R], ui[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    ur[j] = 0.5 * S2[j].re;
    ui[j] = 0.5 * S2[j].im;
  }
#pragma unroll 1
  for (int b = 0; b < NB; ++b) {
    const double d = ((g0 >> b) & 1ull) ? -0.5 : 0.5;
    const cd *col = S2 + (b + 2) * R;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      ur[j] = fma(d, col[j].re, ur[j]);
      ui[j] = fma(d, col[j].im, ui[j]);
    }
  }
  double Xr = 0.0, Xi = 0.0, Yr = 0.0, Yi = 0.0;
  if (t0 & 1ull)
    pair_acc2<-1>(Xr, Xi, Yr, Yi, ur, ui);
  else
    pair_acc2<1>(Xr, Xi, Yr, Yi, ur, ui);
  double Esr[U], Esi[U];
#pragma unroll
  for (int i = 0; i < U; ++i) {
    Esr[i] = 0.0;
    Esi[i] = 0.0;
  }
  for (int i = U; i < c; ++i) Ed[(i - U) * 64] = {0.0, 0.0};
  const bool q_odd = (q & 1ull) != 0;
  const int nblocks = 1 << (c - U);
  // Flip of bit i at local step l: t_{i+1} == 0: delta goes +1 -> -1 (u -= S2, corrector
  // += 2T); t_{i+1} == 1: delta goes -1 -> +1 (u += S2, corrector -= 2T); T = X - Y.
  auto static_steps = [&](int blk) {
#pragma unroll
    for (int j = 1; j < (1 << U); ++j) {
      const int i = __builtin_ctz(j);
      const double Tr = Xr - Yr, Ti = Xi - Yi;
      if (i + 1 < U) {
        if ((j >> (i + 1)) & 1) {
          Esr[i] = fma(-2.0, Tr, Esr[i]);
          Esi[i] = fma(-2.0, Ti, Esi[i]);
          sweep_pair_flip<R, 1>(ur, ui, S2, cp, i, 1.0);
        } else {
          Esr[i] = fma(2.0, Tr, Esr[i]);
          Esi[i] = fma(2.0, Ti, Esi[i]);
          sweep_pair_flip<R, -1>(ur, ui, S2, cp, i, -1.0);
        }
      } else {
        const bool t_next = (U < c) ? ((blk & 1) != 0) : q_odd;
        const double sg = t_next ? 1.0 : -1.0;
        Esr[i] = fma(-2.0 * sg, Tr, Esr[i]);
        Esi[i] = fma(-2.0 * sg, Ti, Esi[i]);
        sweep_pair_flip<R, 0>(ur, ui, S2, cp, i, sg);
      }
      if (j & 1)
        pair_acc2<-1>(Xr, Xi, Yr, Yi, ur, ui);
      else
        pair_acc2<1>(Xr, Xi, Yr, Yi, ur, ui);
      TB_STEP_FENCE();
    }
  };
  static_steps(0);
  for (int blk = 1; blk < nblocks; ++blk) {
    const int i = U + __builtin_ctz(blk);
    const bool t_next = (i + 1 < c) ? (((blk >> (i + 1 - U)) & 1) != 0) : q_odd;
    const double sg = t_next ? 1.0 : -1.0;
    cd e = Ed[(i - U) * 64];
    e.re = fma(-2.0 * sg, Xr - Yr, e.re);
    e.im = fma(-2.0 * sg, Xi - Yi, e.im);
    Ed[(i - U) * 64] = e;
    if (i < ColP<R>::NU)
      sweep_pair_flip<R, 0>(ur, ui, S2, cp, i, sg);
    else
      sweep_flip<R, 0>(ur, ui, S2 + (i + 2) * R, sg);
    pair_acc2<1>(Xr, Xi, Yr, Yi, ur, ui);
    TB_STEP_FENCE();
    static_steps(blk);
  }
  const double Tr = Xr - Yr, Ti = Xi - Yi;
  const uint64_t tend = t0 + ((uint64_t)1 << c) - 1;
  const uint64_t gend = tend ^ (tend >> 1);
  mi[0] = {Tr, Ti};
  mi[1] = {Xr + Yr, Xi + Yi};
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
      mi[b + 2] = {fma(d, Tr, er), fma(d, Ti, ei)};
    } else {
      mi[b + 2] = {d * Tr, d * Ti};
    }
  }
}

// One wave's share of an item whose scaled columns are in S2, chunk c of codes (c - 1
// super-step bits per lane): the chunk and lane mapping of sweep_wave. Lane i returns
// minor i's sum without Kc.
template <int K, int U>
__device__ __forceinline__ cd sweep_pair_wave(const cd *S2, cd *Ed, int lane, int c, int w,
                                              int wpi, int iters) {
  constexpr int R = K - 1;
  const uint64_t nchunks = (uint64_t)1 << (K - 1 - c);
  ColP<R> cp;
#pragma unroll
  for (int u = 0; u < ColP<R>::NU; ++u) load_bcast(cp.b[u], (const double *)(S2 + (u + 2) * R), lane);
  cp.one = vgpr_one();
  asm volatile("s_nop 4");  // VALU writes of the store before its first DPP read
  cd acc = {0.0, 0.0};
  for (int it = 0; it < iters; ++it) {
    const uint64_t q = ((uint64_t)w + (uint64_t)it * wpi) * 64 + lane;
    cd mi[K];
    const bool live = q < nchunks;
    sweep_pair_chunk<K, U>(S2, cp, Ed + lane, c - 1, live ? q : (q & (nchunks - 1)), mi);
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
  return acc;
}

}  // namespace tb
