// launch12.hip -- dispatch of the 1-D and 2-D kernels (eval12_kernel.hpp)
// for FP32 and FP64 tables, 16-byte lanes.
#include "launch.hpp"
#include "eval12_kernel.hpp"

namespace esk {

namespace {

template <class S, int MODE, bool CPLX, bool UNIF, int DIM>
hipError_t go12_api(const Launch &L, const Params &P, hipStream_t s) {
  const long total = (long)P.npoints * P.chunks;
  dim3 block(L.block), grid((unsigned)((total + L.block - 1) / L.block));
  if (L.api == API_STRIDED) {
    hipLaunchKernelGGL((eval12<S, 16, MODE, CPLX, UNIF, DIM, double, false>), grid, block, 0, s, P);
    return hipGetLastError();
  }
  // pointer arrays: the 1-D CUDA-style value and vgl entry points
  if constexpr (DIM == 1 && (MODE == MODE_V || MODE == MODE_VGL)) {
    if (L.api == API_PTRS_F32POS) {
      hipLaunchKernelGGL((eval12<S, 16, MODE, CPLX, UNIF, DIM, float, true>), grid, block, 0, s, P);
      return hipGetLastError();
    }
    if (L.api == API_PTRS_F64POS) {
      hipLaunchKernelGGL((eval12<S, 16, MODE, CPLX, UNIF, DIM, double, true>), grid, block, 0, s, P);
      return hipGetLastError();
    }
  }
  return hipErrorInvalidValue;
}

template <class S, int MODE, bool CPLX, int DIM>
hipError_t go12_unif(const Launch &L, const Params &P, hipStream_t s) {
  const bool unif = (P.chunks % 64) == 0 && (L.block % 64) == 0;
  return unif ? go12_api<S, MODE, CPLX, true, DIM>(L, P, s) : go12_api<S, MODE, CPLX, false, DIM>(L, P, s);
}

template <class S, bool CPLX, int DIM>
hipError_t go12_mode(const Launch &L, const Params &P, hipStream_t s) {
  switch (L.mode) {
    case MODE_V:   return go12_unif<S, MODE_V, CPLX, DIM>(L, P, s);
    case MODE_VG:  return go12_unif<S, MODE_VG, CPLX, DIM>(L, P, s);
    case MODE_VGL: return go12_unif<S, MODE_VGL, CPLX, DIM>(L, P, s);
    case MODE_VGH: return go12_unif<S, MODE_VGH, CPLX, DIM>(L, P, s);
  }
  return hipErrorInvalidValue;
}

template <class S>
hipError_t go12(const Launch &L, const Params &P, int dim, hipStream_t s) {
  if (L.lane_bytes != 16) return hipErrorInvalidValue;
  if (dim == 1) return L.cplx ? go12_mode<S, true, 1>(L, P, s) : go12_mode<S, false, 1>(L, P, s);
  if (dim == 2) return L.cplx ? go12_mode<S, true, 2>(L, P, s) : go12_mode<S, false, 2>(L, P, s);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch12(const Launch &L, const Params &P, int dim, bool f64, hipStream_t s) {
  return f64 ? go12<double>(L, P, dim, s) : go12<float>(L, P, dim, s);
}

}  // namespace esk
