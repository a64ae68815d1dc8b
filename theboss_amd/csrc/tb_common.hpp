// Shared device/host helpers for the gfx950 permanent kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/theboss_amd.h"

namespace tb {

// Interleaved complex128, the layout of torch.complex128 and C99 double _Complex.
struct cd {
  double re, im;
};

// Constant address space: uniform loads through it become scalar (SMEM) loads, so a
// matrix row that every lane of a wave needs is fetched once into SGPRs instead of
// being replicated 64 times through VGPRs.
typedef const __attribute__((address_space(4))) double *kptr;

__device__ __forceinline__ kptr to_k(const void *p) { return (kptr)(p); }

__device__ __forceinline__ cd cmul(cd a, cd b) {
  cd r;
  r.re = fma(a.re, b.re, -(a.im * b.im));
  r.im = fma(a.re, b.im, a.im * b.re);
  return r;
}

__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }

// Balanced product tree over register arrays re[LO..HI), im[LO..HI): depth log2(HI-LO)
// so consecutive complex multiplies are independent (ILP for the FP64 pipeline).
template <int LO, int HI>
struct TreeProd {
  static __device__ __forceinline__ cd run(const double *re, const double *im) {
    constexpr int MID = (LO + HI) / 2;
    return cmul(TreeProd<LO, MID>::run(re, im), TreeProd<MID, HI>::run(re, im));
  }
};
template <int LO>
struct TreeProd<LO, LO + 1> {
  static __device__ __forceinline__ cd run(const double *re, const double *im) {
    return {re[LO], im[LO]};
  }
};
template <int LO>
struct TreeProd<LO, LO> {
  static __device__ __forceinline__ cd run(const double *, const double *) { return {1.0, 0.0}; }
};

// acc += SIGN * a * b with the multiply fused into the accumulation: 4 FMAs instead of a
// complex multiply (4 instructions) followed by 2 adds.
template <int SIGN>
__device__ __forceinline__ void cmul_acc(double &ar, double &ai, cd a, cd b) {
  constexpr double s = SIGN;
  ar = fma(s * a.re, b.re, ar);
  ar = fma(-s * a.im, b.im, ar);
  ai = fma(s * a.re, b.im, ai);
  ai = fma(s * a.im, b.re, ai);
}

// acc += SIGN * prod(re[LO..HI) + i im[LO..HI)), the tree's last multiply fused.
template <int SIGN, int LO, int HI>
__device__ __forceinline__ void tree_prod_acc(double &ar, double &ai, const double *re,
                                              const double *im) {
  constexpr int MID = (LO + HI) / 2;
  cmul_acc<SIGN>(ar, ai, TreeProd<LO, MID>::run(re, im), TreeProd<MID, HI>::run(re, im));
}

// Butterfly sum over the 64 lanes of a wave; every lane receives the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Wave-wide sum without LDS traffic: permlane swaps for lane distances 32 and 16 (the
// swapped pair sums to v + v[lane ^ d] in every lane), then DPP rotations inside each
// row of 16 lanes. Every lane receives the total.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xf, 0xf, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double join64(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// permlane32_swap(a, b) -> {[a_lo | b_lo], [a_hi | b_hi]} (halves of 32 lanes), and
// permlane16_swap the same on rows of 16: the two results of (v, v) add up to
// v + v[lane ^ 32] (resp. ^ 16).
__device__ __forceinline__ double xor32_sum(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)u, (uint32_t)u, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32), false, false);
  return join64(lo[0], hi[0]) + join64(lo[1], hi[1]);
}
__device__ __forceinline__ double xor16_sum(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)u, (uint32_t)u, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32), false, false);
  return join64(lo[0], hi[0]) + join64(lo[1], hi[1]);
}

__device__ __forceinline__ double wave_allsum(double v) {
  v = xor32_sum(v);
  v = xor16_sum(v);
  v += dpp_mov<0x128>(v);  // row_ror:8
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x122>(v);  // row_ror:2
  v += dpp_mov<0x121>(v);  // row_ror:1
  return v;
}

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

}  // namespace tb

// Host-side error channel (defined in capi.hip).
namespace tbh {
int32_t fail(int32_t code, const char *fmt, ...);
int32_t check_hip(hipError_t e, const char *what);
// Tuning-table entry (0 = built-in plan) for TB_TUNE_* kernel `kernel` at `size` and
// `batch` (the table keeps one entry per batch bucket, see tune_bucket).
int32_t tuned(int kernel, int size, int64_t batch);
int tune_bucket(int64_t batch);
}  // namespace tbh
