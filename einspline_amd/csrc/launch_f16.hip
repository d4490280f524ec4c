// launch_f16.hip -- instantiations of the batched kernels for f16s tables.
#define ESK_INSTANTIATE
#include "launch.hpp"

namespace esk {

hipError_t launch_f16(const Launch &L, const Params &P, hipStream_t s) {
  if (use_mfma(L)) return launch_mfma16(L, P, true, s);
  switch (L.lane_bytes) {
    case 8: return go_cplx<f16s, 8>(L, P, s);
    case 16: return go_cplx<f16s, 16>(L, P, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace esk
