// launch_f32.hip -- instantiations of the batched kernels for float tables.
#define ESK_INSTANTIATE
#include "launch.hpp"

namespace esk {

hipError_t launch_f32(const Launch &L, const Params &P, hipStream_t s) {
  if (use_mfma(L)) return launch_mfma32(L, P, s);
  switch (L.lane_bytes) {
    case 8: return go_cplx<float, 8>(L, P, s);
    case 16: return go_cplx<float, 16>(L, P, s);
    case 32: return go_cplx<float, 32>(L, P, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace esk
