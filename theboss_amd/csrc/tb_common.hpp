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

// Butterfly sum over the 64 lanes of a wave; every lane receives the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

}  // namespace tb

// Host-side error channel (defined in capi.hip).
namespace tbh {
int32_t fail(int32_t code, const char *fmt, ...);
int32_t check_hip(hipError_t e, const char *what);
}  // namespace tbh
