// launch_mfma.hip -- instantiations of the matrix-core vgh kernels
// (mfma_kernel.hpp): one wavefront per point, four per block.
#include "launch.hpp"
#include "mfma_kernel.hpp"

namespace esk {

hipError_t launch_mfma16(const Launch &L, const Params &P, bool f16, hipStream_t s) {
  (void)L;
  const dim3 block(256), grid(slab_blocks(P, 4));   // a wavefront per point, or per point and slab
  if (f16) {
    if (P.hpacked) hipLaunchKernelGGL((eval3d_mfma<f16s, true>), grid, block, 0, s, P);
    else hipLaunchKernelGGL((eval3d_mfma<f16s, false>), grid, block, 0, s, P);
  } else {
    if (P.hpacked) hipLaunchKernelGGL((eval3d_mfma<bf16s, true>), grid, block, 0, s, P);
    else hipLaunchKernelGGL((eval3d_mfma<bf16s, false>), grid, block, 0, s, P);
  }
  return hipGetLastError();
}

hipError_t launch_mfma32(const Launch &L, const Params &P, hipStream_t s) {
  (void)L;
  const dim3 block(256), grid(slab_blocks(P, 4));   // a wavefront per point, or per point and slab
  if (P.hpacked) hipLaunchKernelGGL((eval3d_mfma32<true>), grid, block, 0, s, P);
  else hipLaunchKernelGGL((eval3d_mfma32<false>), grid, block, 0, s, P);
  return hipGetLastError();
}

}  // namespace esk
