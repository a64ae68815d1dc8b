// FP32 column-sum Glynn (FP64 accumulation of the alternating sum): instantiates the
// launch templates for n = 2..TB_GLYNN_MAX_N.
#include "glynn_impl.hpp"

namespace tbgl {
int32_t glynn_dispatch_f32(int n, const tb::cpx<float> *A, int64_t mstride, int batch,
                           tb::cd *out, tb::cd *partial, int c_pref, hipStream_t s) {
  return dispatch_glynn<float>(n, A, mstride, batch, out, partial, nullptr, c_pref, s);
}
}  // namespace tbgl
