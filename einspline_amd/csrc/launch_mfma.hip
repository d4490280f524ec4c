// launch_mfma.hip -- instantiations of the matrix-core vgh kernels
// (mfma_kernel.hpp): one wavefront per point, four per block.
#include "launch.hpp"
#include "mfma_kernel.hpp"

namespace esk {

template <class S>
static void launch16(const Params &P, bool stream, dim3 grid, dim3 block, hipStream_t s) {
  if (P.hpacked) {
    if (stream) hipLaunchKernelGGL((eval3d_mfma<S, true, true>), grid, block, 0, s, P);
    else hipLaunchKernelGGL((eval3d_mfma<S, true, false>), grid, block, 0, s, P);
  } else {
    if (stream) hipLaunchKernelGGL((eval3d_mfma<S, false, true>), grid, block, 0, s, P);
    else hipLaunchKernelGGL((eval3d_mfma<S, false, false>), grid, block, 0, s, P);
  }
}

hipError_t launch_mfma16(const Launch &L, const Params &P, bool f16, hipStream_t s) {
  (void)L;
  const dim3 block(256), grid(slab_blocks(P, 4));   // a wavefront per point, or per point and slab
  // Whole 128-orbital slabs start 16-byte aligned in every output row when
  // the arrays do and each row stride is a multiple of four floats.
  auto a16 = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
  const bool stream = P.store == STORE_STAGED_NT && a16(P.out_v) && a16(P.out_g) && a16(P.out_h) &&
                      P.ld_v % 4 == 0 && P.ld_g % 4 == 0 && P.ld_h % 4 == 0;
  if (f16) launch16<f16s>(P, stream, grid, block, s);
  else launch16<bf16s>(P, stream, grid, block, s);
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
