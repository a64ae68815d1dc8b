// FP32 column-sum Glynn (FP64 accumulation of the alternating sum): instantiates the
// launch templates for n = 2..TB_GLYNN_MAX_N.
#include "glynn_impl.hpp"

namespace tbgl {
int32_t glynn_dispatch_f32(int n, const tb::cpx<float> *A, int64_t mstride, const tb::cd *A64,
                           int64_t mstride64, int batch, tb::cd *out, tb::cd *partial,
                           int *handled, int wpm_cap, int c_pref, hipStream_t s) {
  return dispatch_glynn<float>(n, A, mstride, A64, mstride64, batch, out, partial, handled,
                               wpm_cap, c_pref, s);
}
}  // namespace tbgl
