// Multiplicity-aware permanents over Guan codes (reflected mixed-radix Gray codes).
//
// Restates, for a batch of (input, output) occupation pairs on one interferometer:
//  * RyserPermanentCalculator / ChinHuhPermanentCalculator.compute_permanent
//    (theboss/boson_sampling_utilities/permanent_calculators/bs_permanent_calculator_base.py:166-209,
//     ryser_permanent_calculator.py:45-64, chin_huh_permanent_calculator.py:38-59):
//      Ryser:    perm = sum_{r<=s} (-1)^(|s|+|r|) prod_d C(s_d,r_d) prod_j (sum_d r_d U[j][d])^t_j
//      Chin-Huh: perm = 2^-|s| sum_{r<=s} (-1)^|r| prod_d C(s_d,r_d) prod_j (sum_d (s_d-2r_d) U[j][d])^t_j
//  * BSCCRyserSubmatricesPermanentCalculator.compute_permanents for inputs with
//    collisions (bs_submatrices_permanent_calculator_base.py:150-189,
//    bs_cc_ryser_submatrices_permanent_calculator.py:80-119):
//      minor_d = sum_{r<=s, r != 0, r_d < s_d} (-1)^(|t|+|r|) prod C(s,r) (s_d-r_d)/s_d prod_j (...)^t_j
//
// Work split: one wave per item; each lane walks a contiguous range of Guan ranks. A
// lane seeds its Guan state at its first rank from the mixed-radix digits (digit d is
// reflected when the number formed by the higher digits is odd, and its direction is
// the direction of that sweep), then follows the reference's update rule
// (bs_permanent_calculator_base.py:123-149), so every lane visits exactly the codes the
// reference visits at those ranks. The binomial product is kept exact: it is an integer
// below 2^53, updated by exact division and multiplication with table binomials.
#include <algorithm>

#include "tb_common.hpp"

namespace tb {

// C(n, k) for n <= 64 as an exactly representable double (values used here are < 2^53).
__device__ __forceinline__ double binom_exact(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  if (k > n - k) k = n - k;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
  return rint(r);
}

struct GuanItem {
  int D, J, sum_s, sum_t;
};

// mode 0: single permanent (formula 0 Ryser / 1 Chin-Huh) -> out[b]
// mode 1: C&C minors (Ryser) -> out[b*in_max + slot]
template <int JB>
__global__ __launch_bounds__(64) void guan_kernel(const cd *__restrict__ U, int ldu,
                                                  const int *__restrict__ out_modes,
                                                  const int *__restrict__ out_mult, int out_max,
                                                  const int *__restrict__ in_modes,
                                                  const int *__restrict__ in_mult, int in_max,
                                                  int formula, int mode, cd *__restrict__ out) {
  extern __shared__ char smem[];
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  // LDS layout: X[J][D] cd | minors acc [D][64] cd | s[D] t[J] dslot[D] int | r,u [D][64] int8
  cd *X = (cd *)smem;
  cd *Macc = X + out_max * in_max;
  int *s_d = (int *)(Macc + (mode == 1 ? in_max * 64 : 0));
  int *t_j = s_d + in_max;
  int *dslot = t_j + out_max;
  int *hdr = dslot + in_max;  // D, J, sum_s, sum_t
  signed char *rr = (signed char *)(hdr + 4);
  signed char *uu = rr + in_max * 64;

  if (lane == 0) {
    int D = 0, J = 0, ss = 0, st = 0;
    for (int d = 0; d < in_max; ++d) {
      const int m = in_mult[(size_t)b * in_max + d];
      if (m > 0) {
        s_d[D] = m;
        dslot[D] = d;
        ++D;
        ss += m;
      }
    }
    for (int j = 0; j < out_max; ++j) {
      const int m = out_mult[(size_t)b * out_max + j];
      if (m > 0) {
        t_j[J] = m;
        ++J;
        st += m;
      }
    }
    hdr[0] = D;
    hdr[1] = J;
    hdr[2] = ss;
    hdr[3] = st;
  }
  __syncthreads();
  const int D = hdr[0], J = hdr[1], sum_s = hdr[2], sum_t = hdr[3];
  // Gather X[j][d] = U[out_mode_j][in_mode_d] over the compact (nonzero) slots.
  for (int e = lane; e < J * D; e += 64) {
    const int j = e / D, d = e - j * D;
    int jo = -1, cnt = 0;
    for (int q = 0; q < out_max; ++q)
      if (out_mult[(size_t)b * out_max + q] > 0) {
        if (cnt == j) {
          jo = q;
          break;
        }
        ++cnt;
      }
    cd v = {0.0, 0.0};
    if (jo >= 0) {
      const int rowm = out_modes[(size_t)b * out_max + jo];
      const int colm = in_modes[(size_t)b * in_max + dslot[d]];
      v = U[(size_t)rowm * ldu + colm];
    }
    X[e] = v;
  }
  if (mode == 1)
    for (int e = lane; e < D * 64; e += 64) Macc[e] = {0.0, 0.0};
  __syncthreads();

  // Total ranks and this lane's range.
  uint64_t total = 1;
  for (int d = 0; d < D; ++d) total *= (uint64_t)(s_d[d] + 1);
  const uint64_t per = (total + 63) / 64;
  const uint64_t lo = (uint64_t)lane * per;
  const uint64_t hi = min(total, lo + per);
  double acc_re = 0.0, acc_im = 0.0;
  if (lo < hi) {
    // Seed digits/directions at rank lo.
    uint64_t P = 1;
    int sum_r = 0;
    double binom = 1.0;
    for (int d = 0; d < D; ++d) {
      const uint64_t bd = (uint64_t)(s_d[d] + 1);
      const int a = (int)((lo / P) % bd);
      const uint64_t high = lo / (P * bd);
      int g, u;
      if ((high & 1ull) == 0) {
        g = a;
        u = 1;
      } else {
        g = (int)bd - 1 - a;
        u = -1;
      }
      rr[d * 64 + lane] = (signed char)g;
      uu[d * 64 + lane] = (signed char)u;
      sum_r += g;
      binom *= binom_exact(s_d[d], g);
      P *= bd;
    }
    double sr[JB], si[JB];
#pragma unroll
    for (int j = 0; j < JB; ++j) {
      sr[j] = 0.0;
      si[j] = 0.0;
    }
    for (int d = 0; d < D; ++d) {
      const int g = rr[d * 64 + lane];
      const double coef = (formula == 0) ? (double)g : (double)(s_d[d] - 2 * g);
#pragma unroll
      for (int j = 0; j < JB; ++j)
        if (j < J) {
          const cd x = X[j * D + d];
          sr[j] = fma(coef, x.re, sr[j]);
          si[j] = fma(coef, x.im, si[j]);
        }
    }
    // Sign: Ryser perm (-1)^(|s|+|r|); Chin-Huh (-1)^|r| (scale 2^-|s| at the end);
    // C&C minors (-1)^(|t|+|r|).
    int base_par = (mode == 1) ? sum_t : (formula == 0 ? sum_s : 0);
    double sgn = ((base_par + sum_r) & 1) ? -1.0 : 1.0;
    for (uint64_t rank = lo;;) {
      // Term at the current code.
      cd p = {1.0, 0.0};
#pragma unroll
      for (int j = 0; j < JB; ++j)
        if (j < J)
          for (int e = 0; e < t_j[j]; ++e) p = cmul(p, cd{sr[j], si[j]});
      const double w = sgn * binom;
      if (mode == 0) {
        acc_re = fma(w, p.re, acc_re);
        acc_im = fma(w, p.im, acc_im);
      } else if (rank != 0) {
        for (int d = 0; d < D; ++d) {
          const int g = rr[d * 64 + lane];
          if (g < s_d[d]) {
            const double f = w * (double)(s_d[d] - g) / (double)s_d[d];
            cd m = Macc[d * 64 + lane];
            m.re = fma(f, p.re, m.re);
            m.im = fma(f, p.im, m.im);
            Macc[d * 64 + lane] = m;
          }
        }
      }
      if (++rank >= hi) break;
      // Reference Guan update (bs_permanent_calculator_base.py:123-149).
      int d = 0;
      int k = rr[lane] + uu[lane];
      // A movable digit exists because rank < total; the d + 1 < D bound only keeps the
      // scan inside the lane's state.
      while ((k > s_d[d] || k < 0) && d + 1 < D) {
        uu[d * 64 + lane] = (signed char)(-uu[d * 64 + lane]);
        ++d;
        k = rr[d * 64 + lane] + uu[d * 64 + lane];
      }
      const int old = rr[d * 64 + lane];
      rr[d * 64 + lane] = (signed char)k;
      const int dr = k - old;
      binom = binom / binom_exact(s_d[d], old) * binom_exact(s_d[d], k);
      const double coef = (formula == 0) ? (double)dr : -2.0 * (double)dr;
#pragma unroll
      for (int j = 0; j < JB; ++j)
        if (j < J) {
          const cd x = X[j * D + d];
          sr[j] = fma(coef, x.re, sr[j]);
          si[j] = fma(coef, x.im, si[j]);
        }
      sgn = -sgn;
    }
  }
  if (mode == 0) {
    acc_re = wave_sum(acc_re);
    acc_im = wave_sum(acc_im);
    if (lane == 0) {
      const double scale = (formula == 1) ? ldexp(1.0, -sum_s) : 1.0;
      out[b] = {scale * acc_re, scale * acc_im};
    }
  } else {
    __syncthreads();
    for (int q = 0; q < in_max; ++q)
      if (lane == 0) out[(size_t)b * in_max + q] = {0.0, 0.0};
    __syncthreads();
    for (int d = 0; d < D; ++d) {
      const double re = wave_sum(Macc[d * 64 + lane].re);
      const double im = wave_sum(Macc[d * 64 + lane].im);
      if (lane == 0) out[(size_t)b * in_max + dslot[d]] = {re, im};
    }
  }
}

}  // namespace tb

namespace tbg {

size_t guan_lds_bytes(int out_max, int in_max, int mode) {
  size_t bytes = sizeof(tb::cd) * (size_t)out_max * in_max;
  if (mode == 1) bytes += sizeof(tb::cd) * (size_t)in_max * 64;
  bytes += sizeof(int) * ((size_t)in_max * 2 + out_max + 4);
  bytes += (size_t)in_max * 64 * 2;
  return bytes;
}

int32_t run_guan(const tb::cd *U, int ldu, const int *out_modes, const int *out_mult, int out_max,
                 const int *in_modes, const int *in_mult, int in_max, int batch, int formula,
                 int mode, tb::cd *out, hipStream_t s) {
  const size_t lds = guan_lds_bytes(out_max, in_max, mode);
  if (lds > 64 * 1024)
    return tbh::fail(TB_ERANGE, "guan: out_max=%d in_max=%d need %zu bytes of LDS", out_max,
                     in_max, lds);
#define TB_GUAN_LAUNCH(JB)                                                                    \
  hipLaunchKernelGGL((tb::guan_kernel<JB>), dim3(batch), dim3(64), lds, s, U, ldu, out_modes, \
                     out_mult, out_max, in_modes, in_mult, in_max, formula, mode, out)
  if (out_max <= 8)
    TB_GUAN_LAUNCH(8);
  else if (out_max <= 16)
    TB_GUAN_LAUNCH(16);
  else if (out_max <= 32)
    TB_GUAN_LAUNCH(32);
  else
    TB_GUAN_LAUNCH(48);
#undef TB_GUAN_LAUNCH
  return tbh::check_hip(hipGetLastError(), "guan_kernel");
}

}  // namespace tbg

extern "C" int32_t tb_guan_perm_batch(const tb_complex128 *U, int32_t ldu,
                                      const int32_t *out_modes, const int32_t *out_mult,
                                      int32_t out_max, const int32_t *in_modes,
                                      const int32_t *in_mult, int32_t in_max, int32_t batch,
                                      int32_t formula, tb_complex128 *out, void *stream) {
  if (batch < 0) return tbh::fail(TB_EINVAL, "guan: batch=%d < 0", batch);
  if (batch == 0) return TB_OK;
  if (formula != TB_FORMULA_RYSER && formula != TB_FORMULA_CHIN_HUH)
    return tbh::fail(TB_EINVAL, "guan: unknown formula %d", formula);
  if (!U || !out_modes || !out_mult || !in_modes || !in_mult || !out)
    return tbh::fail(TB_EINVAL, "guan: NULL pointer");
  if (out_max < 1 || in_max < 1 || ldu < 1)
    return tbh::fail(TB_EINVAL, "guan: out_max=%d in_max=%d ldu=%d must be >= 1", out_max, in_max,
                     ldu);
  if (out_max > TB_GUAN_MAX_ROWS || in_max > TB_GUAN_MAX_COLS)
    return tbh::fail(TB_ERANGE, "guan: out_max=%d > %d or in_max=%d > %d", out_max,
                     TB_GUAN_MAX_ROWS, in_max, TB_GUAN_MAX_COLS);
  return tbg::run_guan((const tb::cd *)U, ldu, out_modes, out_mult, out_max, in_modes, in_mult,
                       in_max, batch, formula, 0, (tb::cd *)out, (hipStream_t)stream);
}
