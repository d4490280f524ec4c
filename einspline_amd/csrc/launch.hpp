// launch.hpp -- host-side dispatch from runtime parameters to one
// instantiation of esk::eval3d.  Each storage type is instantiated in its own
// translation unit (launch_<type>.hip) so the kernels build in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include "eval_kernel.hpp"

namespace esk {

enum Api { API_STRIDED = 0, API_PTRS_F32POS = 1, API_PTRS_F64POS = 2 };

struct Launch {
  int mode;          // MODE_*
  bool cplx;
  int api;           // Api
  int lane_bytes;    // 8, 16 or 32 (16-bit tables also 8..32)
  int block;         // work-items per block, multiple of 64, <= 256
  int kernel;        // 16-bit tables: 0 vector kernel, 1 matrix-core kernel (vgh, strided)
};

// vgh of a 16-bit table on the matrix cores (mfma_kernel.hpp); `f16` picks
// FP16 storage, else BF16.
hipError_t launch_mfma16(const Launch &L, const Params &P, bool f16, hipStream_t s);

// FP32 tables on v_mfma_f32_16x16x4_f32.
hipError_t launch_mfma32(const Launch &L, const Params &P, hipStream_t s);

// The matrix-core kernels serve strided vgh of real BF16, FP16 and FP32
// tables; everything else runs on the vector kernel.
inline bool use_mfma(const Launch &L) {
  return L.kernel == 1 && L.mode == MODE_VGH && L.api == API_STRIDED && !L.cplx;
}

hipError_t launch_f32(const Launch &L, const Params &P, hipStream_t s);
hipError_t launch_f64(const Launch &L, const Params &P, hipStream_t s);
hipError_t launch_bf16(const Launch &L, const Params &P, hipStream_t s);
hipError_t launch_f16(const Launch &L, const Params &P, hipStream_t s);
// 1-D and 2-D tables (FP32 or FP64 storage, 16-byte lanes), launch12.hip.
hipError_t launch12(const Launch &L, const Params &P, int dim, bool f64, hipStream_t s);

// Launch with `lds` bytes of dynamic LDS; more than 64 KB needs the
// kernel's limit raised (set on every such launch: it is per kernel).
template <class K>
hipError_t launch_lds(K kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, const Params &P) {
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, grid, block, lds, s, P);
  return hipGetLastError();
}

#ifdef ESK_INSTANTIATE
// Instantiation helpers, used only by launch_<type>.hip.
template <class S, int LB, int MODE, bool CPLX, bool UNIF>
hipError_t go_api(const Launch &L, const Params &P, hipStream_t s) {
  const long total = (long)P.npoints * P.chunks;
  dim3 block(L.block), grid((unsigned)((total + L.block - 1) / L.block));
  if (UNIF && P.slabs > 1) grid = dim3(slab_blocks(P, L.block / 64));
  const size_t lds = StageLds<S, LB, MODE, UNIF>::bytes(L.block, P.store);
  if (L.api == API_STRIDED) {
    if constexpr (MODE != MODE_QMC)
      return launch_lds(eval3d<S, LB, MODE, CPLX, UNIF, double, false>, grid, block, lds, s, P);
    return hipErrorInvalidValue;
  }
  if constexpr (LB == 16 && (MODE == MODE_V || MODE == MODE_VGH || MODE == MODE_QMC)) {
    if (L.api == API_PTRS_F32POS)
      return launch_lds(eval3d<S, LB, MODE, CPLX, UNIF, float, true>, grid, block, lds, s, P);
    if (L.api == API_PTRS_F64POS)
      return launch_lds(eval3d<S, LB, MODE, CPLX, UNIF, double, true>, grid, block, lds, s, P);
  }
  return hipErrorInvalidValue;
}

template <class S, int LB, int MODE, bool CPLX>
hipError_t go_unif(const Launch &L, const Params &P, hipStream_t s) {
  const bool unif = (P.chunks % 64) == 0 && (L.block % 64) == 0;
  return unif ? go_api<S, LB, MODE, CPLX, true>(L, P, s)
              : go_api<S, LB, MODE, CPLX, false>(L, P, s);
}

template <class S, int LB, bool CPLX>
hipError_t go_mode(const Launch &L, const Params &P, hipStream_t s) {
  switch (L.mode) {
    case MODE_V:   return go_unif<S, LB, MODE_V, CPLX>(L, P, s);
    case MODE_VG:  return go_unif<S, LB, MODE_VG, CPLX>(L, P, s);
    case MODE_VGL: return go_unif<S, LB, MODE_VGL, CPLX>(L, P, s);
    case MODE_VGH: return go_unif<S, LB, MODE_VGH, CPLX>(L, P, s);
    case MODE_QMC:
      if constexpr (LB == 16 && sizeof(S) >= 4) return go_unif<S, LB, MODE_QMC, CPLX>(L, P, s);
      break;
    case MODE_VGHGH:   // 20 accumulators per real: at most 16 bytes and 4 reals per lane
      if constexpr (LB <= 16 && LB / sizeof(S) <= 4) return go_unif<S, LB, MODE_VGHGH, CPLX>(L, P, s);
      break;
  }
  return hipErrorInvalidValue;
}

template <class S, int LB>
hipError_t go_cplx(const Launch &L, const Params &P, hipStream_t s) {
  if constexpr (sizeof(S) >= 4) {
    if (L.cplx) return go_mode<S, LB, true>(L, P, s);
  } else {
    if (L.cplx) return hipErrorInvalidValue;
  }
  return go_mode<S, LB, false>(L, P, s);
}
#endif

}  // namespace esk
