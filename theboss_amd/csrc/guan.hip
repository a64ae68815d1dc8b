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
// Work split: an item's Guan ranks are divided into contiguous ranges, one per lane of
// the W lanes that own the item. A lane seeds its Guan state at its first rank from the
// mixed-radix digits (digit d is reflected when the number formed by the higher digits is
// odd, and its direction is the direction of that sweep), then follows the reference's
// update rule (bs_permanent_calculator_base.py:123-149), so every lane visits exactly the
// codes the reference visits at those ranks. The binomial product is kept exact: it is an
// integer below 2^53, updated by exact division and multiplication with table binomials.
//
// Launch shapes. Large batches of small items (GCC-A layers: about 100 codes per item)
// are latency-bound, so a wave takes IPW = 4 consecutive items (8 for all-pairs launches
// of rows with at most 8 slots): the lanes read the items' occupation slots in parallel and compact them
// with ballots, and when every item of the wave has at most 64 codes per lane each item is
// walked by its own W = 64 / IPW lanes; otherwise the wave walks its items one after
// another with all 64 lanes. Small batches keep one item per wave (IPW = 1, W = 64).
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

// A packed wave walks each of its items with its own W = 64 / IPW lanes when every item
// has at most this many Guan codes per lane (1024 codes at 16 lanes, 512 at 8).
constexpr int kGuanPackedCodesPerLane = 64;

// Per-item LDS block: X[J][D] (compact rows x compact columns) for at most out_max x
// in_max, then the compact input multiplicities s, their slots and modes, and the compact
// output multiplicities t and modes.
struct GuanItemLds {
  cd *X;
  int *s, *slot, *cmode, *t, *rmode;
};

__host__ __device__ __forceinline__ size_t guan_item_stride(int out_max, int in_max) {
  const size_t bytes = sizeof(cd) * (size_t)out_max * in_max +
                       sizeof(int) * (size_t)(3 * in_max + 2 * out_max);
  return (bytes + 15) & ~(size_t)15;
}

__device__ __forceinline__ GuanItemLds guan_item_at(char *p, int out_max, int in_max) {
  GuanItemLds it;
  it.X = (cd *)p;
  it.s = (int *)(it.X + (size_t)out_max * in_max);
  it.slot = it.s + in_max;
  it.cmode = it.slot + in_max;
  it.t = it.cmode + in_max;
  it.rmode = it.t + out_max;
  return it;
}

// Sums over aligned groups of W lanes; every lane of a group receives its group's sum.
template <int W>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = W / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
template <int W>
__device__ __forceinline__ int group_sum_i(int v) {
#pragma unroll
  for (int off = W / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Guan state at rank lo: digits rr[d*64 + lane], directions uu[d*64 + lane], their sum and
// the binomial product. One division per digit; R = uint32_t when the item has fewer
// than 2^32 codes.
template <typename R>
__device__ __forceinline__ void guan_seed(R lo, int D, const int *s, signed char *rr,
                                          signed char *uu, int lane, int &sum_r, double &binom) {
  R q = lo;
  for (int d = 0; d < D; ++d) {
    const R bd = (R)(s[d] + 1);
    const R high = q / bd;
    const int a = (int)(q - high * bd);
    int g, u;
    if ((high & 1) == 0) {
      g = a;
      u = 1;
    } else {
      g = (int)bd - 1 - a;
      u = -1;
    }
    rr[d * 64 + lane] = (signed char)g;
    uu[d * 64 + lane] = (signed char)u;
    sum_r += g;
    binom *= binom_exact(s[d], g);
    q = high;
  }
}

// Walks, reduces and stores one item with the W lanes of group l's item (l = lane index
// inside the group). mode 0: single permanent (formula 0 Ryser / 1 Chin-Huh) -> out[b];
// mode 1: C&C minors (Ryser) -> out[b*in_max + slot] for the item's nonzero slots.
template <int JB, int W>
__device__ __forceinline__ void guan_item_run(const GuanItemLds &it, int D, int J, int sum_s,
                                              int sum_t, uint64_t total, int b, int in_max,
                                              int l, int lane, int formula, int mode, cd *Macc,
                                              signed char *rr, signed char *uu,
                                              cd *__restrict__ out) {
  if (mode == 1)
    for (int d = 0; d < D; ++d) Macc[d * 64 + lane] = {0.0, 0.0};
  const uint64_t per = (total + W - 1) / W;
  const uint64_t lo = (uint64_t)l * per;
  const uint64_t hi = min(total, lo + per);
  double acc_re = 0.0, acc_im = 0.0;
  if (lo < hi) {
    int sum_r = 0;
    double binom = 1.0;
    if (total <= 0xffffffffull)
      guan_seed<uint32_t>((uint32_t)lo, D, it.s, rr, uu, lane, sum_r, binom);
    else
      guan_seed<uint64_t>(lo, D, it.s, rr, uu, lane, sum_r, binom);
    double sr[JB], si[JB];
#pragma unroll
    for (int j = 0; j < JB; ++j) {
      sr[j] = 0.0;
      si[j] = 0.0;
    }
    for (int d = 0; d < D; ++d) {
      const int g = rr[d * 64 + lane];
      const double coef = (formula == 0) ? (double)g : (double)(it.s[d] - 2 * g);
#pragma unroll
      for (int j = 0; j < JB; ++j)
        if (j < J) {
          const cd x = it.X[j * D + d];
          sr[j] = fma(coef, x.re, sr[j]);
          si[j] = fma(coef, x.im, si[j]);
        }
    }
    // Sign: Ryser perm (-1)^(|s|+|r|); Chin-Huh (-1)^|r| (scale 2^-|s| at the end);
    // C&C minors (-1)^(|t|+|r|).
    const int base_par = (mode == 1) ? sum_t : (formula == 0 ? sum_s : 0);
    double sgn = ((base_par + sum_r) & 1) ? -1.0 : 1.0;
    for (uint64_t rank = lo;;) {
      // Term at the current code.
      cd p = {1.0, 0.0};
#pragma unroll
      for (int j = 0; j < JB; ++j)
        if (j < J)
          for (int e = 0; e < it.t[j]; ++e) p = cmul(p, cd{sr[j], si[j]});
      const double w = sgn * binom;
      if (mode == 0) {
        acc_re = fma(w, p.re, acc_re);
        acc_im = fma(w, p.im, acc_im);
      } else if (rank != 0) {
        for (int d = 0; d < D; ++d) {
          const int g = rr[d * 64 + lane];
          if (g < it.s[d]) {
            const double f = w * (double)(it.s[d] - g) / (double)it.s[d];
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
      while ((k > it.s[d] || k < 0) && d + 1 < D) {
        uu[d * 64 + lane] = (signed char)(-uu[d * 64 + lane]);
        ++d;
        k = rr[d * 64 + lane] + uu[d * 64 + lane];
      }
      const int old = rr[d * 64 + lane];
      rr[d * 64 + lane] = (signed char)k;
      const int dr = k - old;
      binom = binom / binom_exact(it.s[d], old) * binom_exact(it.s[d], k);
      const double coef = (formula == 0) ? (double)dr : -2.0 * (double)dr;
#pragma unroll
      for (int j = 0; j < JB; ++j)
        if (j < J) {
          const cd x = it.X[j * D + d];
          sr[j] = fma(coef, x.re, sr[j]);
          si[j] = fma(coef, x.im, si[j]);
        }
      sgn = -sgn;
    }
  }
  if (mode == 0) {
    acc_re = group_sum<W>(acc_re);
    acc_im = group_sum<W>(acc_im);
    if (l == 0) {
      const double scale = (formula == 1) ? ldexp(1.0, -sum_s) : 1.0;
      out[b] = {scale * acc_re, scale * acc_im};
    }
  } else {
    for (int d = 0; d < D; ++d) {
      const double re = group_sum<W>(Macc[d * 64 + lane].re);
      const double im = group_sum<W>(Macc[d * 64 + lane].im);
      if (l == 0) out[(size_t)b * in_max + it.slot[d]] = {re, im};
    }
  }
}

// LDS: Macc[in_max][64] cd (mode 1) | IPW item blocks | rr, uu [cap][64] int8.
// Batch form (OUTER false): item b pairs input row b with output row b. All-pairs form
// (OUTER true, mode 0): item b = j * n_in + i pairs output row j (n_out rows) with input
// row i (n_in rows), and the row with fewer Guan codes is enumerated, as the batch callers
// choose per item on the host: perm(U[t, s]) = perm(U^T[s, t]), so a transposed item
// swaps the two rows' roles and reads U with swapped strides. Its slots then hold up to
// max(in_max, out_max) entries on either side.
template <int JB, int IPW, bool OUTER = false>
__global__ __launch_bounds__(64) void guan_kernel(const cd *__restrict__ U, int ldu,
                                                  const int *__restrict__ out_modes,
                                                  const int *__restrict__ out_mult, int out_max,
                                                  const int *__restrict__ in_modes,
                                                  const int *__restrict__ in_mult, int in_max,
                                                  int batch, int n_in, int formula, int mode,
                                                  cd *__restrict__ out) {
  constexpr int W = 64 / IPW;  // lanes per item in a packed wave
  extern __shared__ char smem[];
  const int lane = threadIdx.x;
  const int g = lane / W, l = lane % W;
  const int b0 = blockIdx.x * IPW;
  const int b = b0 + g;
  const bool live = b < batch;
  // Slot capacities of the enumerated (columns) and product (rows) sides: wave-uniform.
  const int cap = OUTER ? max(in_max, out_max) : in_max;
  const int pcap = OUTER ? cap : out_max;
  const int *emode = in_modes, *emult = in_mult, *pmode = out_modes, *pmult = out_mult;
  int ew = in_max, pw = out_max;
  size_t urow = (size_t)ldu, ucol = 1;
  if (live) {
    if constexpr (OUTER) {
      const int j = b / n_in, i = b - j * n_in;
      const int *sm = in_modes + (size_t)i * in_max, *su = in_mult + (size_t)i * in_max;
      const int *tm = out_modes + (size_t)j * out_max, *tu = out_mult + (size_t)j * out_max;
      // Code counts prod (m + 1) in double, as the host computes them (exact below 2^53).
      double cs = 1.0, ct = 1.0;
      for (int q = 0; q < in_max; ++q) cs *= (double)(max(su[q], 0) + 1);
      for (int q = 0; q < out_max; ++q) ct *= (double)(max(tu[q], 0) + 1);
      if (ct < cs) {
        emode = tm;
        emult = tu;
        ew = out_max;
        pmode = sm;
        pmult = su;
        pw = in_max;
        urow = 1;
        ucol = (size_t)ldu;
      } else {
        emode = sm;
        emult = su;
        pmode = tm;
        pmult = tu;
      }
    } else {
      emode = in_modes + (size_t)b * in_max;
      emult = in_mult + (size_t)b * in_max;
      pmode = out_modes + (size_t)b * out_max;
      pmult = out_mult + (size_t)b * out_max;
    }
  }
  cd *Macc = (cd *)smem;
  char *items = smem + (mode == 1 ? sizeof(cd) * (size_t)in_max * 64 : 0);
  const size_t stride = guan_item_stride(pcap, cap);
  signed char *rr = (signed char *)(items + IPW * stride);
  signed char *uu = rr + cap * 64;
  const GuanItemLds it = guan_item_at(items + g * stride, pcap, cap);

  // Compact the nonzero slots of both sides in slot order: the group's lanes read W slots
  // at a time and place each nonzero one by a ballot prefix count. In mode 1 a zero input
  // slot's minor is 0 and is stored here; the nonzero slots are stored after the walk.
  const uint64_t gmask = (W == 64) ? ~0ull : (((1ull << W) - 1ull) << (g * W));
  const uint64_t below = (1ull << lane) - 1ull;
  int D = 0, J = 0, ss = 0, st = 0;
  for (int q0 = 0; q0 < cap; q0 += W) {
    const int q = q0 + l;
    int m = 0, md = 0;
    if (live && q < ew) {
      m = emult[q];
      md = emode[q];
      if (mode == 1 && m <= 0) out[(size_t)b * in_max + q] = {0.0, 0.0};
    }
    const uint64_t bal = __ballot(m > 0) & gmask;
    if (m > 0) {
      const int p = D + __popcll(bal & below);
      it.s[p] = m;
      it.slot[p] = q;
      it.cmode[p] = md;
      ss += m;
    }
    D += __popcll(bal);
  }
  for (int q0 = 0; q0 < pcap; q0 += W) {
    const int q = q0 + l;
    int m = 0, md = 0;
    if (live && q < pw) {
      m = pmult[q];
      md = pmode[q];
    }
    const uint64_t bal = __ballot(m > 0) & gmask;
    if (m > 0) {
      const int p = J + __popcll(bal & below);
      it.t[p] = m;
      it.rmode[p] = md;
      st += m;
    }
    J += __popcll(bal);
  }
  ss = group_sum_i<W>(ss);
  st = group_sum_i<W>(st);
  __syncthreads();
  // Gather X[j][d] = U[row mode j][column mode d] over the compact slots.
  for (int e = l; e < J * D; e += W) {
    const int j = e / D, d = e - j * D;
    it.X[e] = U[(size_t)it.rmode[j] * urow + (size_t)it.cmode[d] * ucol];
  }
  __syncthreads();
  uint64_t total = 1;
  for (int d = 0; d < D; ++d) total *= (uint64_t)(it.s[d] + 1);

  if (IPW == 1 || __all(!live || total <= (uint64_t)kGuanPackedCodesPerLane * W)) {
    if (live)
      guan_item_run<JB, W>(it, D, J, ss, st, total, b, in_max, l, lane, formula, mode, Macc, rr,
                           uu, out);
    return;
  }
  // A large item among the wave's items: walk them one after another with all lanes.
  for (int q = 0; q < IPW; ++q) {
    const int bq = b0 + q;
    if (bq >= batch) break;
    const int Dq = __shfl(D, q * W, 64), Jq = __shfl(J, q * W, 64);
    const int ssq = __shfl(ss, q * W, 64), stq = __shfl(st, q * W, 64);
    const GuanItemLds itq = guan_item_at(items + q * stride, pcap, cap);
    uint64_t totq = 1;
    for (int d = 0; d < Dq; ++d) totq *= (uint64_t)(itq.s[d] + 1);
    guan_item_run<JB, 64>(itq, Dq, Jq, ssq, stq, totq, bq, in_max, lane, lane, formula, mode,
                          Macc, rr, uu, out);
  }
}

}  // namespace tb

namespace tbg {

// Below this batch a launch keeps one item per wave, so a batch of large items still
// spreads over enough waves.
constexpr int kGuanPackMinBatch = 16384;

size_t guan_lds_bytes(int out_max, int in_max, int mode, int ipw) {
  size_t bytes = (mode == 1) ? sizeof(tb::cd) * (size_t)in_max * 64 : 0;
  bytes += (size_t)ipw * tb::guan_item_stride(out_max, in_max);
  bytes += (size_t)in_max * 64 * 2;
  return bytes;
}

// n_in > 0: the all-pairs form over batch = n_out * n_in items (mode 0 only).
int32_t run_guan(const tb::cd *U, int ldu, const int *out_modes, const int *out_mult, int out_max,
                 const int *in_modes, const int *in_mult, int in_max, int batch, int formula,
                 int mode, tb::cd *out, hipStream_t s, int n_in = 0) {
  const bool outer = n_in > 0;
  // slot capacities: an all-pairs item may enumerate either side
  const int ecap = outer ? std::max(in_max, out_max) : in_max;
  const int pcap = outer ? ecap : out_max;
  const size_t lds1 = guan_lds_bytes(pcap, ecap, mode, 1);
  if (lds1 > 64 * 1024)
    return tbh::fail(TB_ERANGE, "guan: out_max=%d in_max=%d need %zu bytes of LDS", out_max,
                     in_max, lds1);
  const size_t lds4 = guan_lds_bytes(pcap, ecap, mode, 4);
  const bool pack = batch >= kGuanPackMinBatch && pcap <= 16 && lds4 <= 32 * 1024;
  // All-pairs launches of rows with at most 8 slots a side (GCC-A layers up to 8 particles,
  // the exact distributions: items of at most 2^8 codes when their multiplicities are 1)
  // put eight items in a wave, 8 lanes each: a lane seeds its sums once per chunk, and at
  // 16 lanes per item of 2^5..2^7 codes the seeds cost as much as the 2..8 codes walked
  // (GCC-A m = 20: 20-40% less kernel time at 3..7 particles, even at 8; at 9, items of 2^9
  // codes, 8 lanes took 239 us against 187 with 16).
  const size_t lds8 = guan_lds_bytes(pcap, ecap, mode, 8);
  const bool pack8 = outer && pack && ecap <= 8 && lds8 <= 32 * 1024;
#define TB_GUAN_LAUNCH(JB, IPW, LDS)                                                          \
  do {                                                                                        \
    if (outer)                                                                                \
      hipLaunchKernelGGL((tb::guan_kernel<JB, IPW, true>), dim3((batch + IPW - 1) / IPW),     \
                         dim3(64), LDS, s, U, ldu, out_modes, out_mult, out_max, in_modes,    \
                         in_mult, in_max, batch, n_in, formula, 0, out);                      \
    else                                                                                      \
      hipLaunchKernelGGL((tb::guan_kernel<JB, IPW, false>), dim3((batch + IPW - 1) / IPW),    \
                         dim3(64), LDS, s, U, ldu, out_modes, out_mult, out_max, in_modes,    \
                         in_mult, in_max, batch, 0, formula, mode, out);                      \
  } while (0)
  if (pack8) {
    if (pcap <= 8)
      TB_GUAN_LAUNCH(8, 8, lds8);
    else
      TB_GUAN_LAUNCH(16, 8, lds8);
  } else if (pack) {
    if (pcap <= 8)
      TB_GUAN_LAUNCH(8, 4, lds4);
    else
      TB_GUAN_LAUNCH(16, 4, lds4);
  } else if (pcap <= 8)
    TB_GUAN_LAUNCH(8, 1, lds1);
  else if (pcap <= 16)
    TB_GUAN_LAUNCH(16, 1, lds1);
  else if (pcap <= 32)
    TB_GUAN_LAUNCH(32, 1, lds1);
  else
    TB_GUAN_LAUNCH(48, 1, lds1);
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

extern "C" int32_t tb_guan_perm_outer(const tb_complex128 *U, int32_t ldu,
                                      const int32_t *out_modes, const int32_t *out_mult,
                                      int32_t out_max, int32_t n_out, const int32_t *in_modes,
                                      const int32_t *in_mult, int32_t in_max, int32_t n_in,
                                      int32_t formula, tb_complex128 *out, void *stream) {
  if (n_out < 0 || n_in < 0) return tbh::fail(TB_EINVAL, "guan_outer: n_out=%d n_in=%d < 0", n_out, n_in);
  if (n_out == 0 || n_in == 0) return TB_OK;
  if ((int64_t)n_out * n_in > INT32_MAX)
    return tbh::fail(TB_ERANGE, "guan_outer: n_out * n_in = %lld pairs > %d",
                     (long long)n_out * n_in, INT32_MAX);
  if (formula != TB_FORMULA_RYSER && formula != TB_FORMULA_CHIN_HUH)
    return tbh::fail(TB_EINVAL, "guan_outer: unknown formula %d", formula);
  if (!U || !out_modes || !out_mult || !in_modes || !in_mult || !out)
    return tbh::fail(TB_EINVAL, "guan_outer: NULL pointer");
  if (out_max < 1 || in_max < 1 || ldu < 1)
    return tbh::fail(TB_EINVAL, "guan_outer: out_max=%d in_max=%d ldu=%d must be >= 1", out_max,
                     in_max, ldu);
  const int w = std::max(out_max, in_max);
  if (w > TB_GUAN_MAX_ROWS || w > TB_GUAN_MAX_COLS)
    return tbh::fail(TB_ERANGE, "guan_outer: out_max=%d, in_max=%d: either side may be "
                     "enumerated, so both must be <= %d", out_max, in_max,
                     std::min(TB_GUAN_MAX_ROWS, TB_GUAN_MAX_COLS));
  return tbg::run_guan((const tb::cd *)U, ldu, out_modes, out_mult, out_max, in_modes, in_mult,
                       in_max, n_out * n_in, formula, 0, (tb::cd *)out, (hipStream_t)stream, n_in);
}
