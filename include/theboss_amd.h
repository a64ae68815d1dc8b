/*
 * theboss_amd: C ABI of the MI355X (gfx950) permanent kernels.
 *
 * Conventions shared by every entry point:
 *   - Array arguments are DEVICE pointers (e.g. torch tensor data_ptr()); complex
 *     values are interleaved double pairs (torch.complex128 / C99 double _Complex).
 *   - Matrices are row-major. The interferometer U is indexed U[output_mode][input_mode]
 *     (theboss convention, ryser_permanent_calculator.py:61-64).
 *   - `stream` is a hipStream_t (0 = the null stream). Calls are asynchronous on it.
 *   - Scratch memory is passed in by the caller (`workspace`, size from the matching
 *     *_workspace_size query); the library never allocates or frees device memory
 *     inside an entry point, so calls can be captured in a hipGraph.
 *   - Return value: TB_OK or a negative TB_E* code; tb_last_error() then holds a
 *     message for the calling thread. Arguments are validated on the host before any
 *     launch. No C++ exception crosses this boundary.
 * Reference paths below are relative to the theboss repository root.
 */
#ifndef THEBOSS_AMD_H
#define THEBOSS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TB_OK 0
#define TB_EINVAL (-1)      /* bad shape / argument */
#define TB_ERANGE (-2)      /* size beyond the compiled maximum */
#define TB_EHIP (-3)        /* HIP runtime error */
#define TB_EWORKSPACE (-4)  /* workspace too small */

#define TB_PREC_FP64 0          /* FP64 column sums, products and accumulation */
#define TB_PREC_FP32_COLSUM 1   /* FP32 column sums and products, FP64 accumulation, for
                                   n <= TB_FP32_COLSUM_MAX_N; larger n run in FP64 */
#define TB_PREC_FP32_COLSUM_ANY_N 2 /* the FP32 column-sum kernel at every n: for measuring
                                       its error; above TB_FP32_COLSUM_MAX_N it misses the
                                       1e-10 relative bar of the FP64 path */
#define TB_FP32_COLSUM_MAX_N 23 /* largest n whose FP32 column sums stay in use */

#define TB_FORMULA_RYSER 0
#define TB_FORMULA_CHIN_HUH 1

#ifndef TB_GLYNN_MAX_N
#define TB_GLYNN_MAX_N 48   /* largest square matrix for tb_glynn_batch */
#endif
#ifndef TB_SWEEP_MAX_K
#define TB_SWEEP_MAX_K 40
#endif
#define TB_SWEEP_MAX_K_DOC   /* largest particle count for the binary C&C sweep */
#define TB_GUAN_MAX_ROWS 48 /* distinct output modes for the mixed-radix kernels */
#define TB_GUAN_MAX_COLS 32 /* distinct input modes for the mixed-radix kernels */
#define TB_HIST_MAX_M 64    /* modes of tb_outcome_histogram */
#define TB_HIST_MAX_N 64    /* particles of tb_outcome_histogram */

typedef struct {
  double re, im;
} tb_complex128;

/* Library version (major*10000 + minor*100 + patch). No reference counterpart. */
int32_t tb_version(void);

/* Message of the last failed call on this thread ("" if none). No reference
 * counterpart: the reference raises Python exceptions
 * (bs_permanent_calculator_base.py:179-180). */
const char *tb_last_error(void);

/* Number of HIP devices visible, or a negative TB_E* code. */
int32_t tb_device_count(void);

/* ---------------------------------------------------------------------------------
 * Tuning table. No reference counterpart: the autotuner (theboss_amd.autotune) times
 * candidate chunk sizes and stores the fastest here. A kernel family takes its chunk
 * size -- log2 of the Gray ranks one lane walks between re-seeds, and with it the
 * waves per permanent -- from this table when a call does not set one:
 *   TB_TUNE_SWEEP  C&C minors sweep, size = k (also the steps of tb_gccb_sample)
 *   TB_TUNE_GLYNN  tb_glynn_batch, size = n
 *   TB_TUNE_GRAD   single-sweep gradient of tb_perm_grad_batch, size = n
 * Entries are kept per batch bucket: the batch of a launch (matrices, sweep items, or
 * samples of tb_gccb_sample) falls in bucket ceil(log2(batch)), i.e. 1 | 2 | 3-4 | 5-8 |
 * ... | above 2^(TB_TUNE_BATCH_BUCKETS - 2), because the best chunk for one permanent is
 * not the best for thousands. tb_tune_set with batch <= 0 writes every bucket of the size.
 * chunk_bits 0 selects the built-in plan (which also depends on the batch); other values
 * are clamped to the kernel's valid range when a call plans its launch. Process-wide and
 * thread-safe; each call reads an entry once, so its workspace check and its launch
 * agree. tb_tune_get returns the entry (>= 0) or a negative TB_E* code.
 * ------------------------------------------------------------------------------- */
#define TB_TUNE_SWEEP 1
#define TB_TUNE_GLYNN 2
#define TB_TUNE_GRAD 3
#define TB_TUNE_MAX_SIZE 48
#define TB_TUNE_BATCH_BUCKETS 16
int32_t tb_tune_set(int32_t kernel, int32_t size, int64_t batch, int32_t chunk_bits);
int32_t tb_tune_get(int32_t kernel, int32_t size, int64_t batch);

/* ---------------------------------------------------------------------------------
 * Glynn permanents of a batch of n x n matrices.
 * Replaces GlynnGrayPermanentCalculator.compute_permanent
 * (theboss/boson_sampling_utilities/permanent_calculators/glynn_gray_permanent_calculator.py:41-71)
 * applied to already-gathered effective matrices (boson_sampling_utilities.py:595-626).
 *   A          batch matrices, matrix b at A + b*matrix_stride (elements), row-major n x n
 *   precision  TB_PREC_FP64, TB_PREC_FP32_COLSUM or TB_PREC_FP32_COLSUM_ANY_N. The
 *              reference computes in complex128 throughout; FP32 column sums lose
 *              accuracy with n (relative error ~1e-5 at n = 24, ~3e-3 at n = 28..32), so
 *              TB_PREC_FP32_COLSUM runs the FP64 kernel for n > TB_FP32_COLSUM_MAX_N.
 *   out        batch permanents
 * n == 0 yields 1 for every matrix (glynn_gray_permanent_calculator.py:52-53).
 * ------------------------------------------------------------------------------- */
size_t tb_glynn_workspace_size(int32_t n, int32_t batch, int32_t precision);
int32_t tb_glynn_batch(const tb_complex128 *A, int32_t n, int32_t batch, int64_t matrix_stride,
                       int32_t precision, tb_complex128 *out, void *workspace,
                       size_t workspace_bytes, void *stream);

/* tb_glynn_batch without a workspace: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The scratch memory of the launches (per-wave partial
 * sums, the FP32 copy of the matrices) is a static device array of
 * TB_GLYNN_NOALLOC_SCRATCH_BYTES on each GPU the library runs on, held until the process
 * ends; a call whose tb_glynn_workspace_size exceeds it returns TB_ERANGE (16 x n=32 needs
 * 2 MiB). Same arguments, results and errors as tb_glynn_batch otherwise. The array is
 * shared by every call on that GPU: calls that overlap in time on one GPU (other streams,
 * other threads) race, so the caller orders them. */
#define TB_GLYNN_NOALLOC_SCRATCH_BYTES (16u << 20)
int32_t tb_glynn_batch_noalloc(const tb_complex128 *A, int32_t n, int32_t batch,
                               int64_t matrix_stride, int32_t precision, tb_complex128 *out,
                               void *stream);

/* ---------------------------------------------------------------------------------
 * Multiplicity-aware permanents over Guan codes (reflected mixed-radix Gray codes).
 * Replaces RyserPermanentCalculator.compute_permanent and
 * ChinHuhPermanentCalculator.compute_permanent
 * (bs_permanent_calculator_base.py:166-209, ryser_permanent_calculator.py:45-64,
 *  chin_huh_permanent_calculator.py:38-59) for a batch of (input, output) states on
 * one interferometer.
 *   Item b uses output modes out_modes[b*out_max + j] with multiplicities
 *   out_mult[b*out_max + j] (t_j; 0 pads) and input modes in_modes[b*in_max + d] with
 *   multiplicities in_mult[b*in_max + d] (s_d; 0 pads). The effective matrix has
 *   rows U[out_mode] repeated t times and columns in_mode repeated s times.
 *   formula    TB_FORMULA_RYSER or TB_FORMULA_CHIN_HUH
 *   out        batch permanents (same value as the reference, which enumerates the
 *              input side: sum over r <= s).
 * ------------------------------------------------------------------------------- */
int32_t tb_guan_perm_batch(const tb_complex128 *U, int32_t ldu, const int32_t *out_modes,
                           const int32_t *out_mult, int32_t out_max, const int32_t *in_modes,
                           const int32_t *in_mult, int32_t in_max, int32_t batch,
                           int32_t formula, tb_complex128 *out, void *stream);

/* All pairs of n_out output rows and n_in input rows in one launch:
 *   out[j*n_in + i] = perm(U[t_j, s_i]), t_j = row j of (out_modes, out_mult) (out_max
 *   slots), s_i = row i of (in_modes, in_mult) (in_max slots), same slot layout as
 *   tb_guan_perm_batch.
 * Replaces the per-layer permanent loop of GeneralizedCliffordsSimulationStrategy
 * (generalized_cliffords_simulation_strategy.py:162-171 over :224-247: every candidate
 * output r + e_o against every substate of the input), which calls compute_permanent
 * once per pair.
 * Each pair enumerates the row with fewer Guan codes, prod (m + 1) (the output row when
 * it has strictly fewer; perm(U[t, s]) = perm(U^T[s, t])), so either side may be
 * enumerated: max(out_max, in_max) <= min(TB_GUAN_MAX_ROWS, TB_GUAN_MAX_COLS), else
 * TB_ERANGE. Every row on both sides holds the same number of particles (square effective
 * matrices). n_out * n_in <= INT32_MAX. */
int32_t tb_guan_perm_outer(const tb_complex128 *U, int32_t ldu, const int32_t *out_modes,
                           const int32_t *out_mult, int32_t out_max, int32_t n_out,
                           const int32_t *in_modes, const int32_t *in_mult, int32_t in_max,
                           int32_t n_in, int32_t formula, tb_complex128 *out, void *stream);

/* ---------------------------------------------------------------------------------
 * Clifford & Clifford submatrices permanents: all input-particle-deleted minors in one
 * sweep. Replaces BSCCRyserSubmatricesPermanentCalculator.compute_permanents
 * (bs_submatrices_permanent_calculator_base.py:150-189,
 *  bs_cc_ryser_submatrices_permanent_calculator.py:73-119).
 * Same item description as tb_guan_perm_batch. minors[b*in_max + d] receives
 * perm(U[t, s - e_{in_modes[b,d]}]) (0 where in_mult is 0). Every item of a call must
 * have the same sum(s) = k and sum(t) = k - 1 (the caller passes k, and
 * collision_free = 1 when every s_d <= 1). k == 1 returns s itself
 * (bs_submatrices_permanent_calculator_base.py:156-158). Collision-free inputs run the
 * binary Gray sweep; others run the mixed-radix kernel.
 * ------------------------------------------------------------------------------- */
size_t tb_cc_workspace_size(int32_t k, int32_t out_max, int32_t in_max, int32_t batch);
int32_t tb_cc_submatrices_batch(const tb_complex128 *U, int32_t ldu, const int32_t *out_modes,
                                const int32_t *out_mult, int32_t out_max,
                                const int32_t *in_modes, const int32_t *in_mult, int32_t in_max,
                                int32_t batch, int32_t k, int32_t collision_free,
                                tb_complex128 *minors, void *workspace, size_t workspace_bytes,
                                void *stream);

/* ---------------------------------------------------------------------------------
 * Binary C&C sweep on pre-gathered column data (the building block shared by the
 * submatrices calculator, the gradient and the batched sampler).
 *   Vt      item b, particle nu, row j at Vt[b*item_stride + nu*ldv + j]:
 *           Vt[nu][j] = U[out_j][in_nu] for the k input particles (Gray bits) and the
 *           k-1 output particles (rows, repeated for multiplicities).
 *   minors  minors[b*k + nu] = perm of the (k-1)x(k-1) matrix without particle nu.
 * No direct reference counterpart (the reference always gathers inside the Python
 * calculator); exposed for callers that keep the gathered data resident.
 * ------------------------------------------------------------------------------- */
size_t tb_sweep_workspace_size(int32_t k, int32_t batch);
int32_t tb_sweep_minors_batch(const tb_complex128 *Vt, int32_t k, int32_t ldv,
                              int64_t item_stride, int32_t batch, tb_complex128 *minors,
                              void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------------------
 * Permanent and its gradient d perm(A) / d A[i][j] = perm(A without row i, column j)
 * for a batch of n x n matrices. No reference counterpart (the reference has no
 * differentiation); row i of the gradient is what the C&C submatrices calculator returns
 * with row i deleted. For 14 <= n <= 26 one Gray sweep per group of 8 columns yields
 * all n rows of those columns (leave-one-out products and per-row correctors); other n
 * run one C&C minors sweep per deleted row. Both recompute instead of storing
 * per-subset sums.
 *   perm_out   batch permanents (Laplace expansion of the gradient along row 0)
 *   grad_out   batch x n x n
 * ------------------------------------------------------------------------------- */
size_t tb_perm_grad_workspace_size(int32_t n, int32_t batch);
int32_t tb_perm_grad_batch(const tb_complex128 *A, int32_t n, int32_t batch,
                           int64_t matrix_stride, tb_complex128 *perm_out,
                           tb_complex128 *grad_out, void *workspace, size_t workspace_bytes,
                           void *stream);

/* ---------------------------------------------------------------------------------
 * Batched generalized Clifford & Clifford (GCC-B) sampler.
 * Replaces GeneralizedCliffordsBSimulationStrategy.simulate / _fill_r_sample
 * (theboss/simulation_strategies/generalized_cliffords_b_simulation_strategy.py:41-110)
 * and its uniform-loss variant
 * (generalized_cliffords_b_uniform_losses_simulation_strategy.py:50-121) for many
 * samples at once: every step k of every sample runs the C&C sweep, the Laplace
 * expansion over all m output modes and the inverse-CDF draw on the device.
 *
 *   U             m x m interferometer (row-major, ldu)
 *   particles     n input particles as mode indices in [0, m) (mode assignment of the
 *                 input state, boson_sampling_utilities.py:60-78)
 *   seed          Philox-4x32-10 key for the device RNG
 *   samples       number of samples S
 *   transmissivity  eta in [0,1]; 1 = lossless GCC-B
 *   forced_order  optional S x n particle orders, each row a permutation of 0..n-1
 *                 (NULL = device RNG)
 *   forced_uniforms optional S x n uniforms for the output draws (NULL = device RNG)
 *   forced_steps  optional S step counts (NULL = n if eta == 1 else binomial draw)
 *   out_samples   S x m int32 output occupations
 *   trace_minors  optional S x n x n complex minors trace (step k row k-1), NULL = off
 * The sampler object owns no device memory: plan with tb_gccb_workspace_size, pass a
 * workspace of that many bytes.
 * The index arrays are read on the device, after the call returns: a particles[] entry
 * outside [0, m) or a forced_order[] entry outside [0, n) is clamped into range (no
 * out-of-bounds read) and recorded in the workspace; tb_gccb_status then reports it.
 * ------------------------------------------------------------------------------- */
typedef struct {
  int32_t m, n, samples;
  int32_t chunk_bits[TB_SWEEP_MAX_K + 1]; /* per step k: Gray bits per lane chunk (0 = auto) */
  int32_t per_sample; /* 1: sample s has its own interferometer U + s*u_stride and particle
                         list particles + s*n, whose first forced_steps[s] entries are its
                         particles (forced_steps required; the rest is padding with valid
                         mode indices). Used by the BOBS approximation strategies. */
  int64_t u_stride;   /* elements between per-sample matrices (>= m*ldu) */
} tb_gccb_config;

size_t tb_gccb_workspace_size(const tb_gccb_config *cfg);
int32_t tb_gccb_sample(const tb_gccb_config *cfg, const tb_complex128 *U, int32_t ldu,
                       const int32_t *particles, uint64_t seed, double transmissivity,
                       const int32_t *forced_order, const double *forced_uniforms,
                       const int32_t *forced_steps, int32_t *out_samples,
                       tb_complex128 *trace_minors, void *workspace, size_t workspace_bytes,
                       void *stream);

/* Waits for `stream`, then returns TB_EINVAL (with a message) if the last tb_gccb_sample
 * on this workspace clamped an out-of-range particles[] or forced_order[] entry, else
 * TB_OK. No reference counterpart: the reference indexes Python lists, which raise. */
int32_t tb_gccb_status(const void *workspace, void *stream);

/* ---------------------------------------------------------------------------------
 * Histogram of output samples over the outcome list generate_possible_states(n, m,
 * losses) (boson_sampling_utilities.py:81-152): one block per particle number l
 * (ascending; only l = n when losses == 0), each in reverse lexicographic order.
 * Replaces the counting loop of BSSampleBasedDistributionCalculator
 * (bs_sample_based_distribution_calculator.py:92-106) with a direct rank per sample.
 *   tb_outcome_count   number of outcomes (the length of counts), or a negative TB_E*
 *                      code (TB_ERANGE past 2^31 - 1 outcomes or TB_HIST_MAX_M/_N)
 *   samples            S x m int32 occupations, row stride ld
 *   counts             uint32 per outcome, incremented (the caller zeroes them)
 *   outside            optional uint32 counter of samples not in the list (other particle
 *                      numbers, negative entries); NULL = not counted
 * ------------------------------------------------------------------------------- */
int64_t tb_outcome_count(int32_t n, int32_t m, int32_t losses);
int32_t tb_outcome_histogram(const int32_t *samples, int32_t ld, int32_t samples_n, int32_t m,
                             int32_t n, int32_t losses, uint32_t *counts, uint32_t *outside,
                             void *stream);

/* ---------------------------------------------------------------------------------
 * Clock probe (measurement only; no reference counterpart). tb_clock_probe launches one
 * wave on `stream` that records, into the device array out[2], the shader cycles and the
 * real-time counter ticks that elapse while it waits for `ticks` real-time ticks (it also
 * stops after a bounded number of waits). Launched on a second stream beside a long
 * kernel, out[0] / out[1] * tb_realtime_hz() is the average shader clock during that
 * kernel. tb_realtime_hz returns the real-time counter frequency of the current device
 * (Hz) or a negative TB_E* code.
 * ------------------------------------------------------------------------------- */
int64_t tb_realtime_hz(void);
int32_t tb_clock_probe(uint64_t *out, int64_t ticks, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* THEBOSS_AMD_H */
