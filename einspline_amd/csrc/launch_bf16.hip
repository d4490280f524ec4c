// launch_bf16.hip -- instantiations of the batched kernels for bf16s tables.
#define ESK_INSTANTIATE
#include "launch.hpp"

namespace esk {

hipError_t launch_bf16(const Launch &L, const Params &P, hipStream_t s) {
  if (use_mfma(L)) return launch_mfma16(L, P, false, s);
  switch (L.lane_bytes) {
    case 8: return go_cplx<bf16s, 8>(L, P, s);
    case 16: return go_cplx<bf16s, 16>(L, P, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace esk
