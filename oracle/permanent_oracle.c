/*
 * CPU oracle for the permanent hot path of theboss.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP kernels: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
 * product library (theboss_amd/libtheboss_amd.so) never links or calls it.
 *
 * Every routine restates one reference routine (cited file:line, paths relative to
 * the reference's repository root). Arithmetic is IEEE FP64 complex (C99 _Complex),
 * with optional double-double accumulation for large Glynn sums.
 *
 * Pinning (which check pins which routine) is listed in oracle/oracle.py.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

#define ORC_MAX_N 64

/* ---------------------------------------------------------------------------------
 * Effective scattering matrix.
 * theboss/boson_sampling_utilities/boson_sampling_utilities.py:595-626
 * M[a][b] = U[out_mode(a)][in_mode(b)], rows repeated t_j times, columns s_i times.
 * Returns the row count (= sum(t)); 0 means "empty" ([]) exactly as the reference,
 * which returns [] when either side is empty (:606-607). *cols receives sum(s).
 * --------------------------------------------------------------------------------- */
int orc_effective_matrix(const cplx *U, int m, const int *s, const int *t, cplx *M,
                         int *cols) {
  int sum_s = 0, sum_t = 0;
  for (int i = 0; i < m; ++i) sum_s += s[i];
  for (int i = 0; i < m; ++i) sum_t += t[i];
  *cols = sum_s;
  if (sum_s == 0 || sum_t == 0) { *cols = 0; return 0; }
  int a = 0;
  for (int j = 0; j < m; ++j) {
    for (int rep = 0; rep < t[j]; ++rep, ++a) {
      int b = 0;
      for (int i = 0; i < m; ++i)
        for (int rr = 0; rr < s[i]; ++rr, ++b) M[(size_t)a * sum_s + b] = U[(size_t)j * m + i];
    }
  }
  return sum_t;
}

/* ---------------------------------------------------------------------------------
 * Classic permanent: recursive row expansion.
 * theboss/boson_sampling_utilities/permanent_calculators/classic_permanent_calculator.py:44-67
 * (walks columns, picks an unused row, multiplies the running product).
 * --------------------------------------------------------------------------------- */
static cplx classic_rec(const cplx *A, int rows, int cols, int column, uint64_t used, cplx prod) {
  if (column == cols) return prod;
  cplx result = 0;
  for (int row = 0; row < rows; ++row) {
    if (used & (1ull << row)) continue;
    result += classic_rec(A, rows, cols, column + 1, used | (1ull << row),
                          prod * A[(size_t)row * cols + column]);
  }
  return result;
}

cplx orc_perm_classic_matrix(const cplx *A, int rows, int cols) {
  return classic_rec(A, rows, cols, 0, 0ull, 1.0);
}

/* classic_permanent_calculator.py:27-41: sum(s)==0 -> 1 if sum(t)==0 else 0. */
cplx orc_perm_classic(const cplx *U, int m, const int *s, const int *t) {
  int sum_s = 0, sum_t = 0;
  for (int i = 0; i < m; ++i) { sum_s += s[i]; sum_t += t[i]; }
  if (sum_s == 0) return sum_t == 0 ? 1.0 : 0.0;
  if (sum_t > ORC_MAX_N || sum_s > ORC_MAX_N) return NAN;
  cplx *M = (cplx *)malloc(sizeof(cplx) * (size_t)sum_s * (size_t)(sum_t ? sum_t : 1));
  int cols = 0;
  int rows = orc_effective_matrix(U, m, s, t, M, &cols);
  cplx r = rows ? orc_perm_classic_matrix(M, rows, cols) : 1.0;
  free(M);
  return r;
}

/* ---------------------------------------------------------------------------------
 * Double-double helpers (Knuth TwoSum) for the Glynn accumulator.
 * --------------------------------------------------------------------------------- */
typedef struct { double hi, lo; } dd;
static inline dd dd_add_d(dd a, double b) {
  double s = a.hi + b;
  double bb = s - a.hi;
  double err = (a.hi - (s - bb)) + (b - bb);
  double lo = a.lo + err;
  double hi = s + lo;
  lo = lo - (hi - s);
  dd r = {hi, lo};
  return r;
}
static inline dd dd_add(dd a, dd b) {
  a = dd_add_d(a, b.hi);
  return dd_add_d(a, b.lo);
}

/* ---------------------------------------------------------------------------------
 * Glynn formula with a reflected Gray code.
 * theboss/boson_sampling_utilities/permanent_calculators/glynn_gray_permanent_calculator.py:41-84
 *   delta = 1, sums_j = sum_i M[i][j], acc = prod(sums)                (:73-84)
 *   for each flip index i (Gray code over n-1 bits, guancodes):        (:57-67)
 *     mult = -mult; delta_i = -delta_i; sums_j += 2 delta_i M[i][j]; acc += mult*prod(sums)
 *   acc /= 2^(n-1)                                                     (:69)
 * The flip order comes from the third-party guancodes package (not vendored); the
 * standard reflected Gray code flips bit ctz(step) at step = 1..2^(n-1)-1, which
 * visits the same 2^(n-1) delta vectors, so the sum is the same up to rounding.
 * mode 0: the reference's arithmetic (serial FP64 accumulation, incremental sums).
 * mode 1: column sums re-seeded from delta every 4096 steps and the alternating sum
 *         accumulated in double-double (FP64 sums and products: ~1e-13 relative at
 *         n = 18, less accurate as n and the cancellation of the sum grow).
 * mode 2: the same walk in 80-bit long double throughout (glynn_ld below).
 * --------------------------------------------------------------------------------- */
static cplx glynn_ld(const cplx *M, int n);

cplx orc_perm_glynn_matrix(const cplx *M, int n, int mode) {
  if (n == 0) return 1.0;
  if (mode == 2) return glynn_ld(M, n);
  int delta[ORC_MAX_N];
  cplx sums[ORC_MAX_N];
  for (int i = 0; i < n; ++i) delta[i] = 1;
  for (int j = 0; j < n; ++j) {
    cplx sj = 0;
    for (int i = 0; i < n; ++i) sj += (double)delta[i] * M[(size_t)i * n + j];
    sums[j] = sj;
  }
  int mult = 1;
  cplx p = 1;
  for (int j = 0; j < n; ++j) p *= sums[j];
  cplx acc = p;
  dd acc_re = {creal(p), 0}, acc_im = {cimag(p), 0};
  uint64_t steps = (n >= 1) ? (1ull << (n - 1)) : 1;
  for (uint64_t step = 1; step < steps; ++step) {
    int i = __builtin_ctzll(step);
    mult = -mult;
    delta[i] = -delta[i];
    if (mode == 1 && (step & 4095ull) == 0) {
      for (int j = 0; j < n; ++j) {
        cplx sj = 0;
        for (int r = 0; r < n; ++r) sj += (double)delta[r] * M[(size_t)r * n + j];
        sums[j] = sj;
      }
    } else {
      for (int j = 0; j < n; ++j) sums[j] += 2.0 * (double)delta[i] * M[(size_t)i * n + j];
    }
    cplx q = 1;
    for (int j = 0; j < n; ++j) q *= sums[j];
    if (mode == 1) {
      acc_re = dd_add_d(acc_re, mult * creal(q));
      acc_im = dd_add_d(acc_im, mult * cimag(q));
    } else {
      acc += (double)mult * q;
    }
  }
  double scale = ldexp(1.0, -(n - 1));
  if (mode == 1) return (acc_re.hi + acc_re.lo) * scale + I * ((acc_im.hi + acc_im.lo) * scale);
  return acc * scale;
}

/* glynn_gray_permanent_calculator.py:48-53: empty effective matrix -> 1. */
cplx orc_perm_glynn(const cplx *U, int m, const int *s, const int *t, int mode) {
  int sum_s = 0, sum_t = 0;
  for (int i = 0; i < m; ++i) { sum_s += s[i]; sum_t += t[i]; }
  if (sum_s == 0 || sum_t == 0) return 1.0;
  if (sum_s != sum_t || sum_s > ORC_MAX_N) return NAN;
  cplx *M = (cplx *)malloc(sizeof(cplx) * (size_t)sum_s * (size_t)sum_t);
  int cols = 0;
  int n = orc_effective_matrix(U, m, s, t, M, &cols);
  cplx r = orc_perm_glynn_matrix(M, n, mode);
  free(M);
  return r;
}

/* ---------------------------------------------------------------------------------
 * Mode 2 (reference grade for n >= 24, where mode 1's FP64 column sums and products
 * limit it to ~1e-13 .. 1e-11 relative): the same Gray walk in 80-bit long double.
 * Column sums are seeded exactly from delta at the start of every chunk of 4096 steps
 * and updated by +-2 M[i][j]; products form a balanced tree; a chunk's alternating sum
 * is compensated (Neumaier) and the chunk sums are added in chunk order, so the result
 * does not depend on the thread count. OpenMP over the chunks of one matrix.
 * --------------------------------------------------------------------------------- */
typedef long double ldbl;

static inline void ld_neumaier(ldbl *s, ldbl *c, ldbl x) {
  const ldbl t = *s + x;
  if (fabsl(*s) >= fabsl(x))
    *c += (*s - t) + x;
  else
    *c += (x - t) + *s;
  *s = t;
}

static inline void ld_tree_prod(const ldbl *re, const ldbl *im, int n, ldbl *pr, ldbl *pi) {
  ldbl tr[ORC_MAX_N], ti[ORC_MAX_N];
  for (int j = 0; j < n; ++j) { tr[j] = re[j]; ti[j] = im[j]; }
  int m = n;
  while (m > 1) {
    const int h = m / 2;
    for (int k = 0; k < h; ++k) {
      const ldbl ar = tr[2 * k], ai = ti[2 * k], br = tr[2 * k + 1], bi = ti[2 * k + 1];
      tr[k] = ar * br - ai * bi;
      ti[k] = ar * bi + ai * br;
    }
    if (m & 1) { tr[h] = tr[m - 1]; ti[h] = ti[m - 1]; }
    m = h + (m & 1);
  }
  *pr = tr[0];
  *pi = ti[0];
}

static cplx glynn_ld(const cplx *M, int n) {
  if (n == 0) return 1.0;
  const int nb = n - 1;
  const int cbits = nb < 12 ? nb : 12;
  const uint64_t nchunks = 1ull << (nb - cbits);
  ldbl *cre = (ldbl *)malloc(sizeof(ldbl) * nchunks), *cim = (ldbl *)malloc(sizeof(ldbl) * nchunks);
  ldbl Mr[ORC_MAX_N * ORC_MAX_N], Mi[ORC_MAX_N * ORC_MAX_N];
  for (int e = 0; e < n * n; ++e) { Mr[e] = creal(M[e]); Mi[e] = cimag(M[e]); }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (uint64_t q = 0; q < nchunks; ++q) {
    const uint64_t t0 = q << cbits;
    const uint64_t g0 = t0 ^ (t0 >> 1);
    int delta[ORC_MAX_N];
    for (int i = 0; i < n; ++i) delta[i] = (i < nb && ((g0 >> i) & 1ull)) ? -1 : 1;
    ldbl sr[ORC_MAX_N], si[ORC_MAX_N];
    for (int j = 0; j < n; ++j) {
      ldbl a = 0, b = 0;
      for (int i = 0; i < n; ++i) { a += delta[i] * Mr[i * n + j]; b += delta[i] * Mi[i * n + j]; }
      sr[j] = a;
      si[j] = b;
    }
    ldbl s_re = 0, c_re = 0, s_im = 0, c_im = 0, pr, pi;
    ld_tree_prod(sr, si, n, &pr, &pi);
    const ldbl sg0 = (t0 & 1ull) ? -1.0L : 1.0L;  /* (-1)^popcount(gray(t)) = (-1)^t */
    ld_neumaier(&s_re, &c_re, sg0 * pr);
    ld_neumaier(&s_im, &c_im, sg0 * pi);
    for (uint64_t l = 1; l < (1ull << cbits); ++l) {
      const uint64_t t = t0 + l;
      const int i = __builtin_ctzll(t);
      delta[i] = -delta[i];
      const ldbl d2 = 2.0L * delta[i];
      for (int j = 0; j < n; ++j) { sr[j] += d2 * Mr[i * n + j]; si[j] += d2 * Mi[i * n + j]; }
      ld_tree_prod(sr, si, n, &pr, &pi);
      const ldbl sg = (t & 1ull) ? -1.0L : 1.0L;
      ld_neumaier(&s_re, &c_re, sg * pr);
      ld_neumaier(&s_im, &c_im, sg * pi);
    }
    cre[q] = s_re + c_re;
    cim[q] = s_im + c_im;
  }
  ldbl s_re = 0, c_re = 0, s_im = 0, c_im = 0;
  for (uint64_t q = 0; q < nchunks; ++q) {
    ld_neumaier(&s_re, &c_re, cre[q]);
    ld_neumaier(&s_im, &c_im, cim[q]);
  }
  free(cre);
  free(cim);
  const ldbl scale = ldexpl(1.0L, -nb);
  return (double)((s_re + c_re) * scale) + I * (double)((s_im + c_im) * scale);
}

/* Batch of square matrices, OpenMP over the batch (CPU baseline leg); mode 2 runs the
 * matrices one after another, each parallel over its chunks. */
void orc_perm_glynn_batch(const cplx *A, int n, int batch, int mode, int threads, cplx *out) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  if (mode == 2) {
    for (int b = 0; b < batch; ++b) out[b] = glynn_ld(A + (size_t)b * n * n, n);
    return;
  }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int b = 0; b < batch; ++b) out[b] = orc_perm_glynn_matrix(A + (size_t)b * n * n, n, mode);
}

/* ---------------------------------------------------------------------------------
 * Guan-code machinery (reflected mixed-radix Gray code over r <= s).
 * theboss/boson_sampling_utilities/permanent_calculators/bs_permanent_calculator_base.py:75-214
 *   _initialize_guan_codes_variables  :118-121   r = 0, u = 1, limits = s
 *   _update_guan_code                 :123-149
 *   _update_binomials_product         :151-164   (float ratio update)
 * --------------------------------------------------------------------------------- */
typedef struct {
  int m;
  const int *limits; /* s */
  int *r, *u;
  int index;         /* index_to_update */
  int last;          /* last_value_at_index */
  double binom;
} guan_state;

static void guan_init(guan_state *g, int m, const int *s) {
  g->m = m;
  g->limits = s;
  g->r = (int *)calloc((size_t)(m ? m : 1), sizeof(int));
  g->u = (int *)malloc(sizeof(int) * (size_t)(m ? m : 1));
  for (int i = 0; i < m; ++i) g->u[i] = 1;
  g->index = 0;
  g->last = 0;
  g->binom = 1.0;
}
static void guan_free(guan_state *g) { free(g->r); free(g->u); }

/* bs_permanent_calculator_base.py:123-149 */
static void guan_update_code(guan_state *g) {
  g->index = 0;
  int k = g->r[0] + g->u[0];
  while (k > g->limits[g->index] || k < 0) {
    g->u[g->index] = -g->u[g->index];
    g->index += 1;
    if (g->index == g->m) return;
    k = g->r[g->index] + g->u[g->index];
  }
  g->last = g->r[g->index];
  g->r[g->index] = k;
}

/* bs_permanent_calculator_base.py:151-164 */
static void guan_update_binom(guan_state *g, const int *s) {
  int i = g->index;
  if (g->r[i] > g->last)
    g->binom *= (double)(s[i] - g->last) / (double)g->r[i];
  else
    g->binom *= (double)g->last / (double)(s[i] - g->r[i]);
}

static cplx cpow_int(cplx z, int e) {
  cplx r = 1;
  for (int k = 0; k < e; ++k) r *= z;
  return r;
}

/* ---------------------------------------------------------------------------------
 * Ryser / Chin-Huh single permanents over Guan codes, with mode multiplicities.
 * bs_permanent_calculator_base.py:166-209 (loop + _update_permanent),
 * ryser_permanent_calculator.py:45-64  (mult = (-1)^sum(s), sums = 0,
 *                                       sums_j += (r_i - last) * U[j][i]),
 * chin_huh_permanent_calculator.py:38-59 (mult = 2^-sum(s), sums_j = sum_i s_i U[j][i],
 *                                         sums_j -= 2 (r_i - last) U[j][i]).
 * formula 0 = Ryser, 1 = Chin-Huh. Returns NAN for the shape failure that raises
 * AttributeError in the reference (bs_permanent_calculator_base.py:61-72,179-180);
 * the caller passes matching lengths, so only m == 0 triggers it here.
 * --------------------------------------------------------------------------------- */
static cplx guan_term(const cplx *sums, const int *cols, int ncols, const int *t, double mult,
                      double binom) {
  /* bs_permanent_calculator_base.py:200-209: reduce(mul, [sums_j^t_j], mult*binom) */
  cplx acc = mult * binom;
  for (int c = 0; c < ncols; ++c) acc *= cpow_int(sums[c], t[cols[c]]);
  return acc;
}

cplx orc_perm_guan(const cplx *U, int m, const int *s, const int *t, int formula) {
  if (m <= 0) return NAN;
  int *cols = (int *)malloc(sizeof(int) * (size_t)m);
  int ncols = 0;
  for (int j = 0; j < m; ++j)
    if (t[j] != 0) cols[ncols++] = j; /* considered columns = nonzero(output_state) */
  cplx *sums = (cplx *)malloc(sizeof(cplx) * (size_t)(ncols ? ncols : 1));
  int sum_s = 0;
  for (int i = 0; i < m; ++i) sum_s += s[i];
  double mult;
  if (formula == 0) {
    mult = (sum_s % 2) ? -1.0 : 1.0;
    for (int c = 0; c < ncols; ++c) sums[c] = 0;
  } else {
    mult = ldexp(1.0, -sum_s);
    for (int c = 0; c < ncols; ++c) {
      cplx acc = 0;
      for (int i = 0; i < m; ++i) acc += (double)s[i] * U[(size_t)cols[c] * m + i];
      sums[c] = acc;
    }
  }
  guan_state g;
  guan_init(&g, m, s);
  cplx perm = guan_term(sums, cols, ncols, t, mult, g.binom);
  while (g.r[m - 1] <= g.limits[m - 1]) {
    guan_update_code(&g);
    if (g.index == m) break;
    mult = -mult;
    int i = g.index;
    int dr = g.r[i] - g.last;
    for (int c = 0; c < ncols; ++c) {
      if (formula == 0)
        sums[c] += (double)dr * U[(size_t)cols[c] * m + i];
      else
        sums[c] -= 2.0 * (double)dr * U[(size_t)cols[c] * m + i];
    }
    guan_update_binom(&g, s);
    perm += guan_term(sums, cols, ncols, t, mult, g.binom);
  }
  guan_free(&g);
  free(sums);
  free(cols);
  return perm;
}

/* ---------------------------------------------------------------------------------
 * Clifford & Clifford submatrices permanents (all input-mode-deleted minors).
 * bs_submatrices_permanent_calculator_base.py:150-189 (shortcut sum(s)==1, loop order:
 *   code, mult flip, binomials, sums, permanents; mult = (-1)^sum(t); no r=0 addend),
 * bs_cc_ryser_submatrices_permanent_calculator.py:73-119 (prod = mult*prod sums^t;
 *   perm_i += prod * binom / (s_i / (s_i - r_i)) for s_i != 0 and r_i != s_i).
 * bs_cc_ch_submatrices_permanent_calculator.py:51-105 (Chin-Huh flavour: mult =
 *   2^-(sum(s)-1), sums_j = sum_i s_i U[j][i], per-i reduced sums sums_j - U[j][i],
 *   initial addend at r = 0).
 * out receives m values (one per input mode; 0 where s_i == 0).
 * formula 0 = Ryser, 1 = Chin-Huh.
 * --------------------------------------------------------------------------------- */
void orc_cc_submatrices(const cplx *U, int m, const int *s, const int *t, int formula, cplx *out) {
  int sum_s = 0;
  for (int i = 0; i < m; ++i) sum_s += s[i];
  if (sum_s == 1) { /* bs_submatrices_permanent_calculator_base.py:156-158 */
    for (int i = 0; i < m; ++i) out[i] = (double)s[i];
    return;
  }
  for (int i = 0; i < m; ++i) out[i] = 0;
  if (m <= 0) return;
  int *cols = (int *)malloc(sizeof(int) * (size_t)m);
  int ncols = 0, sum_t = 0;
  for (int j = 0; j < m; ++j) {
    sum_t += t[j];
    if (t[j] != 0) cols[ncols++] = j;
  }
  cplx *sums = (cplx *)malloc(sizeof(cplx) * (size_t)(ncols ? ncols : 1));
  double mult;
  guan_state g;
  guan_init(&g, m, s);
  if (formula == 0) {
    mult = (sum_t % 2) ? -1.0 : 1.0;
    for (int c = 0; c < ncols; ++c) sums[c] = 0;
  } else {
    mult = ldexp(1.0, -(sum_s - 1));
    for (int c = 0; c < ncols; ++c) {
      cplx acc = 0;
      for (int i = 0; i < m; ++i) acc += (double)s[i] * U[(size_t)cols[c] * m + i];
      sums[c] = acc;
    }
  }
  cplx *reduced = (cplx *)malloc(sizeof(cplx) * (size_t)(ncols ? ncols : 1));
  /* Chin-Huh initial _update_permanents (bs_cc_ch_submatrices_permanent_calculator.py:66) */
  int first = (formula == 1);
  for (;;) {
    if (!first) {
      if (!(g.r[m - 1] <= g.limits[m - 1])) break;
      guan_update_code(&g);
      if (g.index == m) break;
      if (formula == 0) mult = -mult;
      else mult = -mult; /* base class flips the multiplier for both flavours (:170) */
      guan_update_binom(&g, s);
      int i = g.index;
      int dr = g.r[i] - g.last;
      for (int c = 0; c < ncols; ++c) {
        if (formula == 0)
          sums[c] += (double)dr * U[(size_t)cols[c] * m + i];
        else
          sums[c] -= 2.0 * (double)dr * U[(size_t)cols[c] * m + i];
      }
    }
    first = 0;
    if (formula == 0) {
      cplx prod = 1;
      for (int c = 0; c < ncols; ++c) prod *= cpow_int(sums[c], t[cols[c]]);
      prod *= mult;
      for (int i = 0; i < m; ++i) {
        if (s[i] == 0 || g.r[i] == s[i]) continue;
        double ub = g.binom;
        ub /= (double)s[i] / (double)(s[i] - g.r[i]);
        out[i] += prod * ub;
      }
    } else {
      for (int i = 0; i < m; ++i) {
        if (s[i] == 0 || s[i] == g.r[i]) continue;
        for (int c = 0; c < ncols; ++c) reduced[c] = sums[c] - U[(size_t)cols[c] * m + i];
        double ub = g.binom / ((double)s[i] / (double)(s[i] - g.r[i]));
        cplx addend = mult * ub;
        for (int c = 0; c < ncols; ++c) addend *= cpow_int(reduced[c], t[cols[c]]);
        out[i] += addend;
      }
    }
  }
  guan_free(&g);
  free(reduced);
  free(sums);
  free(cols);
}

/* ---------------------------------------------------------------------------------
 * All (n-1)x(n-1) minors of a square matrix: G[i][j] = perm(A without row i, col j)
 * = d perm(A) / d A[i][j]. No reference counterpart (the reference has no gradient);
 * restated as n^2 independent oracle-grade Glynn permanents so that it shares no code
 * path with the single-sweep GPU backward.
 * --------------------------------------------------------------------------------- */
void orc_perm_minors(const cplx *A, int n, int threads, cplx *G) {
  if (n == 1) { G[0] = 1.0; return; }
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int ij = 0; ij < n * n; ++ij) {
    int i = ij / n, j = ij % n;
    cplx sub[ORC_MAX_N * ORC_MAX_N];
    int a = 0;
    for (int r = 0; r < n; ++r) {
      if (r == i) continue;
      int b = 0;
      for (int c = 0; c < n; ++c) {
        if (c == j) continue;
        sub[a * (n - 1) + b] = A[(size_t)r * n + c];
        ++b;
      }
      ++a;
    }
    G[ij] = orc_perm_glynn_matrix(sub, n - 1, 1);
  }
}

/* ---------------------------------------------------------------------------------
 * GCC-B step pmf.
 * theboss/simulation_strategies/generalized_cliffords_b_simulation_strategy.py:69-92
 *   minors = C&C-Ryser submatrices(U, current_input, r_sample)
 *   amp_m  = sum_i current_input_i * minors_i * U[m][i];  pmf = |amp|^2 / sum
 * --------------------------------------------------------------------------------- */
void orc_gccb_pmf(const cplx *U, int m, const int *cur_in, const int *r_sample, double *pmf,
                  cplx *minors_out) {
  cplx *minors = (cplx *)malloc(sizeof(cplx) * (size_t)m);
  orc_cc_submatrices(U, m, cur_in, r_sample, 0, minors);
  double total = 0;
  for (int mm = 0; mm < m; ++mm) {
    cplx amp = 0;
    for (int i = 0; i < m; ++i) {
      if (cur_in[i] == 0) continue;
      amp += (double)cur_in[i] * minors[i] * U[(size_t)mm * m + i];
    }
    double p = creal(amp) * creal(amp) + cimag(amp) * cimag(amp);
    pmf[mm] = p;
    total += p;
  }
  for (int mm = 0; mm < m; ++mm) pmf[mm] /= total;
  if (minors_out) memcpy(minors_out, minors, sizeof(cplx) * (size_t)m);
  free(minors);
}

/* ---------------------------------------------------------------------------------
 * GCC-B sampler, driven by caller-supplied randomness so that runs are reproducible
 * and comparable with the GPU sampler:
 *   order[s*n + k]   the input particle (mode index) added at step k of sample s
 *                    (the reference pops a uniformly random remaining particle,
 *                    generalized_cliffords_b_simulation_strategy.py:102-105);
 *   uniforms[s*n+k]  the uniform in [0,1) that picks the output mode at step k by
 *                    inverse CDF over the normalised pmf (:107-110 uses
 *                    numpy.random.choice, which is an inverse-CDF draw);
 *   steps[s]         how many steps sample s runs (n lossless; l for uniform loss,
 *                    generalized_cliffords_b_uniform_losses_simulation_strategy.py:111-121).
 * out[s*m + mode] receives the output occupation. OpenMP over samples.
 * --------------------------------------------------------------------------------- */
void orc_gccb_sample_batch(const cplx *U, int m, int n, int samples, const int *order,
                           const double *uniforms, const int *steps, int threads, int *out) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int sidx = 0; sidx < samples; ++sidx) {
    int *cur = (int *)calloc((size_t)m, sizeof(int));
    int *r = out + (size_t)sidx * m;
    double *pmf = (double *)malloc(sizeof(double) * (size_t)m);
    for (int i = 0; i < m; ++i) r[i] = 0;
    int nsteps = steps ? steps[sidx] : n;
    for (int k = 0; k < nsteps; ++k) {
      cur[order[(size_t)sidx * n + k]] += 1;
      orc_gccb_pmf(U, m, cur, r, pmf, NULL);
      double u = uniforms[(size_t)sidx * n + k];
      double c = 0;
      int pick = m - 1;
      for (int mm = 0; mm < m; ++mm) {
        c += pmf[mm];
        if (c > u) { pick = mm; break; }
      }
      r[pick] += 1;
    }
    free(pmf);
    free(cur);
  }
}

/* ---------------------------------------------------------------------------------
 * GCC-A layer pmf (the unnormalised pmf of the next particle's output mode).
 * theboss/simulation_strategies/generalized_cliffords_simulation_strategy.py:136-247:
 * the next particle is number k = sum(r) + 1 (:141); the input substates s <= input with
 * sum(s) = k (:94-121, :143-145) get weights w(s) = l! (n - l)! / n! prod_i C(in_i, in_i - s_i)
 * with l = sum(in - s) (:190-205), normalised (:155); for each output r + e_o (:207-224)
 *   pmf[o] = sum_s w(s) |perm(U[r + e_o, s])|^2 / (prod_i s_i! k!)        (:158-167, :226-247)
 * with the permanent by the Ryser restatement above. Substates are enumerated as a
 * mixed-radix counter over the modes (the reference's order differs; the sum is the same up
 * to rounding). Layers: `layers` rows of r (m ints each); pmf: layers x m. OpenMP over layers.
 * --------------------------------------------------------------------------------- */
static double orc_fact(int v) {
  double f = 1.0;
  for (int i = 2; i <= v; ++i) f *= (double)i;
  return f;
}

static double orc_binom(int a, int b) {
  if (b < 0 || b > a) return 0.0;
  return orc_fact(a) / (orc_fact(b) * orc_fact(a - b));
}

void orc_gcca_layer_pmf_batch(const cplx *U, int m, const int *input, const int *r_samples,
                              int layers, int threads, double *pmf) {
  int n = 0;
  for (int i = 0; i < m; ++i) n += input[i];
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int L = 0; L < layers; ++L) {
    const int *r = r_samples + (size_t)L * m;
    double *out = pmf + (size_t)L * m;
    int k = 1;
    for (int i = 0; i < m; ++i) k += r[i];
    int *s = (int *)calloc((size_t)m, sizeof(int));
    int *t = (int *)malloc(sizeof(int) * (size_t)m);
    /* pass 0: sum of the weights; pass 1: weighted probabilities */
    double wsum = 0.0;
    for (int o = 0; o < m; ++o) out[o] = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      for (int i = 0; i < m; ++i) s[i] = 0;
      for (;;) {
        int ks = 0;
        for (int i = 0; i < m; ++i) ks += s[i];
        if (ks == k) {
          const int l = n - k;
          double w = orc_fact(l) * orc_fact(n - l) / orc_fact(n);
          for (int i = 0; i < m; ++i) w *= orc_binom(input[i], input[i] - s[i]);
          if (pass == 0) {
            wsum += w;
          } else {
            double norm = orc_fact(k);
            for (int i = 0; i < m; ++i) norm *= orc_fact(s[i]);
            for (int o = 0; o < m; ++o) {
              memcpy(t, r, sizeof(int) * (size_t)m);
              t[o] += 1;
              const double a = cabs(orc_perm_guan(U, m, s, t, 0));
              out[o] += (w / wsum) * (a * a / norm);
            }
          }
        }
        int i = 0;  /* next substate: mixed-radix increment, digit i in 0..input[i] */
        while (i < m && s[i] == input[i]) s[i++] = 0;
        if (i == m) break;
        s[i] += 1;
      }
    }
    free(t);
    free(s);
  }
}

int orc_omp_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Dump the Guan code sequence (r vectors, row-major, count x m) produced by
 * bs_permanent_calculator_base.py:123-149 starting from r = 0. Returns the number of
 * codes written (including r = 0), at most max_codes. */
int orc_guan_sequence(int m, const int *s, int *codes, int max_codes) {
  guan_state g;
  guan_init(&g, m, s);
  int count = 0;
  if (max_codes > 0) { memcpy(codes, g.r, sizeof(int) * (size_t)m); count = 1; }
  while (count < max_codes) {
    guan_update_code(&g);
    if (g.index == m) break;
    memcpy(codes + (size_t)count * m, g.r, sizeof(int) * (size_t)m);
    ++count;
  }
  guan_free(&g);
  return count;
}
