// launch_f64.hip -- instantiations of the batched kernels for double tables.
#define ESK_INSTANTIATE
#include "launch.hpp"

namespace esk {

hipError_t launch_f64(const Launch &L, const Params &P, hipStream_t s) {
  switch (L.lane_bytes) {
    case 16: return go_cplx<double, 16>(L, P, s);
    case 32: return go_cplx<double, 32>(L, P, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace esk
