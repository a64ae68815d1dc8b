#pragma once
// Gray-walk building blocks shared by the Glynn, sweep and gradient kernels: complex values
// of either precision, balanced product trees, and the flip of one matrix row into the
// half column sums (from LDS/global rows, VGPR rows, SGPR-held values or broadcast-form
// VGPR vectors).
#include "tb_common.hpp"

namespace tb {

// Keeps the scheduler from hoisting the next Gray step's row reads above the current
// step (which otherwise multiplies live registers by the unroll factor and spills).
#ifndef TB_STEP_FENCE
#define TB_STEP_FENCE() __builtin_amdgcn_sched_barrier(0)
#endif

template <typename F>
struct cpx {
  F re, im;
};

template <typename F>
__device__ __forceinline__ cpx<F> cmulT(cpx<F> a, cpx<F> b) {
  cpx<F> r;
  r.re = fma(a.re, b.re, -(a.im * b.im));
  r.im = fma(a.re, b.im, a.im * b.re);
  return r;
}

template <typename F, int LO, int HI>
struct TreeProdT {
  static __device__ __forceinline__ cpx<F> run(const F *re, const F *im) {
    constexpr int MID = (LO + HI) / 2;
    return cmulT<F>(TreeProdT<F, LO, MID>::run(re, im), TreeProdT<F, MID, HI>::run(re, im));
  }
};
template <typename F, int LO>
struct TreeProdT<F, LO, LO + 1> {
  static __device__ __forceinline__ cpx<F> run(const F *re, const F *im) { return {re[LO], im[LO]}; }
};

// h += row (add) or h -= row; `row` is a wave-uniform LDS row (broadcast reads) or a
// per-lane global row (NARROW mode). The row is consumed in groups of G complex values
// with the reads of group g+1 issued before the adds of group g, and a scheduling
// barrier per group, so at most 2*G loaded values are live at once.
#ifndef TB_FLIP_GROUP
#define TB_FLIP_GROUP 4
#endif

// MODE 1: h += row, MODE -1: h -= row, MODE 0: h += sg * row (runtime direction, one
// FMA per component with sg = +-1 held per lane).
// Columns J0..N-1 only (J0 > 0 when the first J0 values of the row are held in SGPRs).
template <int N, int MODE, typename F, typename P, int J0 = 0, int G = TB_FLIP_GROUP>
__device__ __forceinline__ void flip_mem(F (&hr)[N], F (&hi)[N], P row, F sg) {
  constexpr int NG = (N - J0 + G - 1) / G;
  cpx<F> cur[G], nxt[G];
#pragma unroll
  for (int e = 0; e < G; ++e)
    if (J0 + e < N) cur[e] = row[J0 + e];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    if (g + 1 < NG) {
#pragma unroll
      for (int e = 0; e < G; ++e)
        if (J0 + (g + 1) * G + e < N) nxt[e] = row[J0 + (g + 1) * G + e];
    }
#pragma unroll
    for (int e = 0; e < G; ++e) {
      const int j = J0 + g * G + e;
      if (j < N) {
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

template <int N, int MODE, typename F>
__device__ __forceinline__ void flip_reg(F (&hr)[N], F (&hi)[N], const F (&rr)[N], const F (&ri)[N],
                                         F sg) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    if constexpr (MODE == 1) {
      hr[j] += rr[j];
      hi[j] += ri[j];
    } else if constexpr (MODE == -1) {
      hr[j] -= rr[j];
      hi[j] -= ri[j];
    } else {
      hr[j] = fma(sg, rr[j], hr[j]);
      hi[j] = fma(sg, ri[j], hi[j]);
    }
  }
}

// The first S values of row 0 held in SGPRs (wave-uniform): a flip of row 0 reads them
// as scalar operands and only the rest from LDS, which unloads the LDS return path.
template <int S, typename F>
struct SRow {
  F re[S > 0 ? S : 1], im[S > 0 ? S : 1];
};

__device__ __forceinline__ double uniform_value(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float uniform_value(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// Flip of row 0: its first S values from SGPRs, the rest from `row`.
template <int N, int S, int MODE, typename F, typename P>
__device__ __forceinline__ void flip_srow(F (&hr)[N], F (&hi)[N], const SRow<S, F> &sr, P row,
                                          F sg) {
#pragma unroll
  for (int j = 0; j < S; ++j) {
    if constexpr (MODE == 1) {
      hr[j] += sr.re[j];
      hi[j] += sr.im[j];
    } else if constexpr (MODE == -1) {
      hr[j] -= sr.re[j];
      hi[j] -= sr.im[j];
    } else {
      hr[j] = fma(sg, sr.re[j], hr[j]);
      hi[j] = fma(sg, sr.im[j], hi[j]);
    }
  }
  flip_mem<N, MODE, F, P, S>(hr, hi, row, sg);
}

// Load the first S values of a wave-uniform row into SGPRs.
template <int S, typename F, typename P>
__device__ __forceinline__ void load_srow(SRow<S, F> &sr, P row) {
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const cpx<F> v = row[j];
    sr.re[j] = uniform_value(v.re);
    sr.im[j] = uniform_value(v.im);
  }
}

// A row (or column) of N complex values in broadcast form: value t (re/im interleaved)
// sits in lane t mod 16 of every 16-lane row of VGPR pair t / 16, i.e. 2 ceil(2N/16) VGPRs
// per vector. A flip adds it to the half sums with one v_fmac_f64 per value whose first
// operand is taken from that lane by DPP row_newbcast, with no LDS read. Every lane of
// the wave must be active at a flip (DPP reads other lanes' registers).
template <int N>
struct Bcast {
  static constexpr int P = (2 * N + 15) / 16;
  double v[P];
};

// Lane `lane` of the wave loads its slots of the vector `src` (N complex values).
template <int N>
__device__ __forceinline__ void load_bcast(Bcast<N> &b, const double *src, int lane) {
#pragma unroll
  for (int p = 0; p < Bcast<N>::P; ++p) {
    const int t = 16 * p + (lane & 15);
    b.v[p] = t < 2 * N ? src[t] : 0.0;
  }
}

// h += s * v[L] (L = lane within each 16-lane row), and the same with -v[L].
template <int L>
__device__ __forceinline__ void fmac_bcast(double &h, double v, double s) {
  asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(h)
      : "v"(v), "v"(s), "i"(L));
}
template <int L>
__device__ __forceinline__ void fmac_bcast_neg(double &h, double v, double s) {
  asm("v_fmac_f64_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(h)
      : "v"(v), "v"(s), "i"(L));
}

// h += sg * vector (MODE 0), h += vector (MODE 1, s = 1) or h -= vector (MODE -1, s = 1).
template <int T, int N, int MODE>
__device__ __forceinline__ void flip_bcast(double (&hr)[N], double (&hi)[N], const Bcast<N> &b,
                                           double s) {
  if constexpr (T < 2 * N) {
    double &h = (T & 1) ? hi[T / 2] : hr[T / 2];
    if constexpr (MODE == -1)
      fmac_bcast_neg<T % 16>(h, b.v[T / 16], s);
    else
      fmac_bcast<T % 16>(h, b.v[T / 16], s);
    flip_bcast<T + 1, N, MODE>(hr, hi, b, s);
  }
}

// ------------------------------------------------------------------------------------
// Paired walks. Two Gray codes that differ only in one row (particle) r have column sums
// c_j (u_j + 1) and c_j (u_j - 1) once every other row is divided by row r
// (u_j = sum_{i != r} delta_i M[i][j] / M[r][j]), so one update of u serves both codes
// and each factor (u_j +- 1) enters a running product x as x u_j +- x: four FMAs with
// the +-1 folded into their addends.
// ------------------------------------------------------------------------------------

// prod_{j in [LO, HI)} (u_j + PM), PM = +-1: the first factor formed by one add, each
// further one folded into x u_j + PM x.
template <int PM, int LO, int HI, int N>
__device__ __forceinline__ cd fold_chain(const double (&ur)[N], const double (&ui)[N]) {
  double xr = ur[LO] + (double)PM, xi = ui[LO];
#pragma unroll
  for (int j = LO + 1; j < HI; ++j) {
    const double nr = fma(-xi, ui[j], fma(xr, ur[j], (double)PM * xr));
    const double ni = fma(xi, ur[j], fma(xr, ui[j], (double)PM * xi));
    xr = nr;
    xi = ni;
  }
  return {xr, xi};
}

// X += SIGN prod(u + 1), Y += SIGN prod(u - 1); each product is two chains of N/2
// factors joined by a multiply fused into the accumulation.
template <int SIGN, int N>
__device__ __forceinline__ void pair_acc2(double &xr, double &xi, double &yr, double &yi,
                                          const double (&ur)[N], const double (&ui)[N]) {
  constexpr int H = N / 2;
  cmul_acc<SIGN>(xr, xi, fold_chain<1, 0, H>(ur, ui), fold_chain<1, H, N>(ur, ui));
  cmul_acc<SIGN>(yr, yi, fold_chain<-1, 0, H>(ur, ui), fold_chain<-1, H, N>(ur, ui));
}

// 1 / c for c != 0, with c scaled by a power of two so that |c|^2 stays in range.
__device__ __forceinline__ cd cinv_scaled(cd c) {
  const int e = ilogb(fmax(fabs(c.re), fabs(c.im)));
  const double cr = ldexp(c.re, -e), ci = ldexp(c.im, -e);
  const double d = fma(cr, cr, ci * ci);
  return {ldexp(cr / d, -e), ldexp(-ci / d, -e)};
}

// A pair row must have every entry above 2^-16 of the matrix's largest entry (squared
// magnitudes compared): far from zero, its quotients stay in range and well rounded.
constexpr double kPairMinRel2 = 0x1p-32;

// c = prod_j v[j] * 2^-J (J values) as a mantissa and a power of two, renormalised after
// every factor so that no partial product over- or underflows.
__device__ __forceinline__ void prod_half_scaled(const cd *v, int J, cd &mant, int &e2) {
  cd k = {1.0, 0.0};
  int ke = -J;
  for (int j = 0; j < J; ++j) {
    k = cmul(k, v[j]);
    const double a = fmax(fabs(k.re), fabs(k.im));
    if (a > 0.0) {
      const int e = ilogb(a);
      k.re = ldexp(k.re, -e);
      k.im = ldexp(k.im, -e);
      ke += e;
    }
  }
  mant = k;
  e2 = ke;
}

// 1.0 in a VGPR the compiler cannot fold away (second operand of fixed-direction flips).
__device__ __forceinline__ double vgpr_one() {
  double one = 1.0;
  asm volatile("" : "+v"(one));
  return one;
}

}  // namespace tb
