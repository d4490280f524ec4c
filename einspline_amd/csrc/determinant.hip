// determinant.hip -- Slater-determinant kernels that consume orbital rows,
// batched over matrices (one per walker):
//   det_ratios   ratio[m] = new_row_m . column `col` of Ainv_m
//                (src/cuda_walker.cu:150-185);
//   sm_prepare,  Sherman-Morrison update of Ainv after row k of A changes
//   sm_update    (src/determinant_update.h:8-104, src/determinant_update.cu:10-98);
//   fill_ratio   the spline row fill (src/cuda_walker.cu:69-95) with the ratio
//                taken from the values while they are still in registers.
//
// The update is HBM-bound: every element of Ainv has to be read and written
// once.  The reference reads Ainv twice (once for d = delta^T Ainv, once in
// the update) and writes it once.  Here a workgroup owns a slab of columns
// (16 lanes x 16 bytes wide) over all rows and keeps it in registers between
// the reduction and the update, so each element is read once and written
// once.  Every slab also needs column k of the old inverse; sm_prepare
// gathers it first into the caller's colk vector, so no slab reads a column
// that another slab is rewriting.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "determinant_update_hip.h"
#include "eval_kernel.hpp"
#include "runtime_internal.hpp"

namespace {

using esk::uniform;

constexpr int kThreads = 256;

template <class T>
__device__ __forceinline__ T fmaT(T a, T b, T c) { return esk::fmaA(a, b, c); }

// Sum over the 64 lanes of a wavefront (butterfly: every lane gets the total,
// in the same order on every lane).
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// ratios: one wavefront per matrix; lane l sums rows l, l+64, ...

template <class T>
__global__ void __launch_bounds__(kThreads) det_ratios(T *const *__restrict__ Ainv_list,
                                                       T *const *__restrict__ row_list, T *__restrict__ ratio,
                                                       int N, int rs, int col, int num) {
  const int m = uniform((int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
  if (m >= num) return;
  const int lane = threadIdx.x & 63;
  const T *Ainv = Ainv_list[m] + col;
  const T *u = row_list[m];
  T s = (T)0;
  for (int i0 = lane; i0 < N; i0 += 4 * 64) {
    T a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {            // four rows in flight per lane
      const int i = i0 + q * 64;
      a[q] = i < N ? Ainv[(long)i * rs] : (T)0;
      b[q] = i < N ? u[i] : (T)0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) s = fmaT(a[q], b[q], s);
  }
  s = wave_sum(s);
  if (lane == 0) ratio[m] = s;
}

// ---------------------------------------------------------------------------
// Sherman-Morrison

// One workgroup per matrix: colk = column k of Ainv; with A, delta = u - A[k]
// and A[k] = u.
template <class T>
__global__ void __launch_bounds__(kThreads) sm_prepare(T *const *__restrict__ A_list,
                                                       T *const *__restrict__ Ainv_list,
                                                       T *const *__restrict__ u_list,
                                                       T *const *__restrict__ delta_list,
                                                       T *const *__restrict__ colk_list, int N, int rs,
                                                       int k) {
  const int m = blockIdx.x;
  const T *Ainv = Ainv_list[m] + k;
  T *colk = colk_list[m];
  for (int i = threadIdx.x; i < N; i += kThreads) colk[i] = Ainv[(long)i * rs];
  if (A_list) {
    T *Ak = A_list[m] + (long)k * rs;
    const T *u = u_list[m];
    T *delta = delta_list[m];
    for (int i = threadIdx.x; i < N; i += kThreads) {
      const T ui = u[i];
      delta[i] = ui - Ak[i];
      Ak[i] = ui;
    }
  }
}

template <class T> struct Vec16;
template <> struct Vec16<float>  { typedef float4 type; static constexpr int n = 4; };
template <> struct Vec16<double> { typedef double2 type; static constexpr int n = 2; };

template <class T>
__device__ __forceinline__ void ld16(const T *p, T *r) {
  const typename Vec16<T>::type q = *(const typename Vec16<T>::type *)p;
  const T *e = (const T *)&q;
#pragma unroll
  for (int i = 0; i < Vec16<T>::n; ++i) r[i] = e[i];
}
template <class T>
__device__ __forceinline__ void st16(T *p, const T *r) {
  typename Vec16<T>::type q;
  T *e = (T *)&q;
#pragma unroll
  for (int i = 0; i < Vec16<T>::n; ++i) e[i] = r[i];
  *(typename Vec16<T>::type *)p = q;
}

// Slab update.  Block (x, m) owns columns [x*W, x*W + W) of matrix m; thread
// t handles VEC consecutive columns (lane group t % 16) of the rows
// t/16, t/16 + RG, ...  The first RPT of those rows stay in registers; rows
// past RPT*RG (N > 1024) are read again for the update.
//   d_j  = sum_i delta_i Ainv[i][j]                (per slab: rows, then row groups)
//   d_k  = sum_i delta_i colk_i                     (every block, same order)
//   Ainv[i][j] += d_j * (prefact * colk_i),  prefact = -1 / (1 + d_k)
// with the operand order of src/determinant_update.h:87-103.
// VECLD: the matrix base is 16-byte aligned and N, rs are multiples of VEC.
template <class T, int THREADS, int RPT, bool VECLD>
__device__ __forceinline__ void sm_slab(T *__restrict__ Ainv, const T *__restrict__ dl,
                                        const T *__restrict__ colk, int N, int rs, T (&part)[THREADS / 16][64],
                                        T *red, T *dsh) {
  constexpr int VEC = Vec16<T>::n;
  constexpr int CG = 16;
  constexpr int RG = THREADS / CG;
  constexpr int W = CG * VEC;
  const int t = threadIdx.x, cg = t % CG, rg = t / CG;
  const int c0 = blockIdx.x * W + cg * VEC;

  // d_k (uses only colk and delta)
  T s = (T)0;
  for (int i = t; i < N; i += THREADS) s = fmaT(dl[i], colk[i], s);
  s = wave_sum(s);
  if ((t & 63) == 0) red[t >> 6] = s;

  auto load = [&](int i, T *r) {
    const T *p = Ainv + (long)i * rs + c0;
    if constexpr (VECLD) {
      ld16(p, r);
    } else {
#pragma unroll
      for (int e = 0; e < VEC; ++e) r[e] = c0 + e < N ? p[e] : (T)0;
    }
  };
  auto store = [&](int i, const T *r) {
    T *p = Ainv + (long)i * rs + c0;
    if constexpr (VECLD) {
      st16(p, r);
    } else {
#pragma unroll
      for (int e = 0; e < VEC; ++e)
        if (c0 + e < N) p[e] = r[e];
    }
  };
  const bool mine = c0 < N;            // VECLD: then all VEC columns are < N

  T reg[RPT][VEC];
  T acc[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) acc[e] = (T)0;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int i = rg + r * RG;
    if (mine && i < N) {
      load(i, reg[r]);
      const T di = dl[i];
#pragma unroll
      for (int e = 0; e < VEC; ++e) acc[e] = fmaT(reg[r][e], di, acc[e]);
    }
  }
  for (int i = rg + RPT * RG; mine && i < N; i += RG) {
    T x[VEC];
    load(i, x);
    const T di = dl[i];
#pragma unroll
    for (int e = 0; e < VEC; ++e) acc[e] = fmaT(x[e], di, acc[e]);
  }
#pragma unroll
  for (int e = 0; e < VEC; ++e) part[rg][cg * VEC + e] = acc[e];
  __syncthreads();
  if (t < W) {
    T d = (T)0;
#pragma unroll 8
    for (int g = 0; g < RG; ++g) d += part[g][t];
    dsh[t] = d;
  }
  __syncthreads();
  T dk = (T)0;
#pragma unroll
  for (int w = 0; w < THREADS / 64; ++w) dk += red[w];
  const T prefact = (T)-1 / ((T)1 + dk);
  T dj[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) dj[e] = dsh[cg * VEC + e];

#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int i = rg + r * RG;
    if (mine && i < N) {
      const T ci = prefact * colk[i];
#pragma unroll
      for (int e = 0; e < VEC; ++e) reg[r][e] = fmaT(dj[e], ci, reg[r][e]);
      store(i, reg[r]);
    }
  }
  for (int i = rg + RPT * RG; mine && i < N; i += RG) {
    T x[VEC];
    load(i, x);
    const T ci = prefact * colk[i];
#pragma unroll
    for (int e = 0; e < VEC; ++e) x[e] = fmaT(dj[e], ci, x[e]);
    store(i, x);
  }
}

template <class T, int THREADS, int RPT>
__global__ void __launch_bounds__(THREADS) sm_update(T *const *__restrict__ Ainv_list,
                                                     T *const *__restrict__ dl_list,
                                                     T *const *__restrict__ colk_list, int N, int rs,
                                                     int vec_ok) {
  __shared__ T part[THREADS / 16][64];
  __shared__ T red[THREADS / 64];
  __shared__ T dsh[64];
  const int m = blockIdx.y;
  T *Ainv = Ainv_list[m];
  const T *dl = dl_list[m];
  const T *colk = colk_list[m];
  if (vec_ok && ((uintptr_t)Ainv & 15) == 0)
    sm_slab<T, THREADS, RPT, true>(Ainv, dl, colk, N, rs, part, red, dsh);
  else
    sm_slab<T, THREADS, RPT, false>(Ainv, dl, colk, N, rs, part, red, dsh);
}

// ---------------------------------------------------------------------------
// fused row fill + ratio: one wavefront per walker, lane l evaluates chunks
// l, l+64, ... of 16 bytes of the table row.  The values are computed with
// the operations, in the order, of the batched evaluation's value mode
// (eval_kernel.hpp), so the rows equal what the value entry points write.

template <class S, class PosT>
__global__ void __launch_bounds__(kThreads) fill_ratio(const esk::Params P,
                                                       typename esk::acc_of<S>::type *const *__restrict__ rows,
                                                       typename esk::acc_of<S>::type *const *__restrict__ Ainv_list,
                                                       typename esk::acc_of<S>::type *__restrict__ ratio,
                                                       int N, int rs, int col) {
  using A = typename esk::acc_of<S>::type;
  using L = esk::Loader<S, 16>;
  constexpr int V = L::V;
  const int p = uniform((int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
  if (p >= P.npoints) return;
  const int lane = threadIdx.x & 63;

  const PosT *pp = (const PosT *)P.pos + 3l * p;
  int ix, iy, iz;
  A tx, ty, tz;
  esk::locate<A>((double)pp[0], P.x0, P.xi, P.nx4, ix, tx);
  esk::locate<A>((double)pp[1], P.y0, P.yi, P.ny4, iy, ty);
  esk::locate<A>((double)pp[2], P.z0, P.zi, P.nz4, iz, tz);
  ix = uniform(ix); iy = uniform(iy); iz = uniform(iz);
  tx = uniform(tx); ty = uniform(ty); tz = uniform(tz);
  A a[4], b[4], c[4];
  esk::basis0(tx, a); esk::basis0(ty, b); esk::basis0(tz, c);
#pragma unroll
  for (int q = 0; q < 4; ++q) { a[q] = uniform(a[q]); b[q] = uniform(b[q]); c[q] = uniform(c[q]); }

  A *row = rows[p];
  const A *Acol = Ainv_list[p] + col;
  const bool vec = ((uintptr_t)row & 15) == 0;
  const S *cell = (const S *)P.coefs + (long)ix * P.xs + (long)iy * P.ys + (long)iz * P.zs;
  const long zs = P.zs;
  A s = (A)0;
  for (int ch = lane; ch < P.chunks; ch += 64) {
    const int e0 = ch * V;
    // this lane's entries of the Ainv column: independent of the table reads
    A av[V];
#pragma unroll
    for (int e = 0; e < V; ++e) av[e] = e0 + e < N ? Acol[(long)(e0 + e) * rs] : (A)0;

    const S *base = cell + e0;
    A acc[V];
#pragma unroll
    for (int e = 0; e < V; ++e) acc[e] = (A)0;
    A rb[2][4][V];
#pragma unroll
    for (int k = 0; k < 4; ++k) L::load(base + k * zs, rb[0][k]);
    constexpr bool YX = esk::contract_yx<S>::value;   // as the value entry point
    A tq[V];
#pragma unroll
    for (int st = 0; st < 16; ++st) {
      const int i = st >> 2, j = st & 3;
      if (st + 1 < 16) {
        const S *rp = base + ((st + 1) >> 2) * P.xs + ((st + 1) & 3) * P.ys;
#pragma unroll
        for (int k = 0; k < 4; ++k) L::load(rp + k * zs, rb[(st + 1) & 1][k]);
      }
      A (&r)[4][V] = rb[st & 1];
      A s0[V];
#pragma unroll
      for (int e = 0; e < V; ++e) {
        s0[e] = c[0] * r[0][e];
        s0[e] = esk::fmaA(c[1], r[1][e], s0[e]);
        s0[e] = esk::fmaA(c[2], r[2][e], s0[e]);
        s0[e] = esk::fmaA(c[3], r[3][e], s0[e]);
      }
      if constexpr (YX) {
#pragma unroll
        for (int e = 0; e < V; ++e) tq[e] = j == 0 ? b[j] * s0[e] : esk::fmaA(b[j], s0[e], tq[e]);
        if (j == 3) {
#pragma unroll
          for (int e = 0; e < V; ++e) acc[e] = esk::fmaA(a[i], tq[e], acc[e]);
        }
      } else {
        const A ab = a[i] * b[j];
#pragma unroll
        for (int e = 0; e < V; ++e) acc[e] = esk::fmaA(ab, s0[e], acc[e]);
      }
#pragma unroll
      for (int e = 0; e < V; ++e) esk::pin(acc[e]);
      __builtin_amdgcn_sched_barrier(0);
    }

    const int mrem = P.nreal - e0;
    if (mrem >= V) esk::store_block<A, V>(row, e0, acc, vec, false);
    else if (mrem > 0) esk::store_partial<A, V>(row + e0, acc, mrem);
#pragma unroll
    for (int e = 0; e < V; ++e)
      if (e0 + e < N) s = esk::fmaA(acc[e], av[e], s);
  }
  s = wave_sum(s);
  if (lane == 0) ratio[p] = s;
}

// ---------------------------------------------------------------------------
// host side

template <class T>
int update_inverse(T **A_list, T **Ainv_list, T **u_list, T **delta_list, T **colk_list, int N, int rs,
                   int k, int num, void *stream) {
  if (num < 0) return esk_fail(EINSPLINE_EINVAL, "update_inverse: negative matrix count");
  if (num == 0) return EINSPLINE_OK;
  if (!Ainv_list || !u_list || !colk_list || (A_list && !delta_list))
    return esk_fail(EINSPLINE_EINVAL, "update_inverse: NULL list");
  if (N < 1 || rs < N || k < 0 || k >= N)
    return esk_fail(EINSPLINE_EINVAL, "update_inverse: need 1 <= N <= rowstride and 0 <= k < N (N=%d, "
                    "rowstride=%d, k=%d)", N, rs, k);
  if (num > 65535) return esk_fail(EINSPLINE_EINVAL, "update_inverse: at most 65535 matrices per call");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(sm_prepare<T>, dim3(num), dim3(kThreads), 0, s, A_list, Ainv_list, u_list, delta_list,
                     colk_list, N, rs, k);
  constexpr int W = 16 * Vec16<T>::n;
  const int vec_ok = (N % Vec16<T>::n == 0) && (rs % Vec16<T>::n == 0);
  T **dl = A_list ? delta_list : u_list;
  const dim3 grid((unsigned)((N + W - 1) / W), (unsigned)num);
  if (N <= 256)
    hipLaunchKernelGGL((sm_update<T, 256, 16>), grid, dim3(256), 0, s, Ainv_list, dl, colk_list, N, rs, vec_ok);
  else
    hipLaunchKernelGGL((sm_update<T, 1024, 16>), grid, dim3(1024), 0, s, Ainv_list, dl, colk_list, N, rs,
                       vec_ok);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EINSPLINE_OK : esk_hip_fail(e, "update_inverse launch");
}

template <class T>
int calc_ratios(T **Ainv_list, T **row_list, T *ratio, int N, int rs, int col, int num, void *stream) {
  if (num < 0) return esk_fail(EINSPLINE_EINVAL, "calc_ratios: negative matrix count");
  if (num == 0) return EINSPLINE_OK;
  if (!Ainv_list || !row_list || !ratio) return esk_fail(EINSPLINE_EINVAL, "calc_ratios: NULL argument");
  if (N < 1 || rs < N || col < 0 || col >= N)
    return esk_fail(EINSPLINE_EINVAL, "calc_ratios: need 1 <= N <= rowstride and 0 <= elec < N (N=%d, "
                    "rowstride=%d, elec=%d)", N, rs, col);
  const unsigned blocks = (unsigned)((num + kThreads / 64 - 1) / (kThreads / 64));
  hipLaunchKernelGGL(det_ratios<T>, dim3(blocks), dim3(kThreads), 0, (hipStream_t)stream, Ainv_list, row_list,
                     ratio, N, rs, col, num);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EINSPLINE_OK : esk_hip_fail(e, "calc_ratios launch");
}

template <class A>
int fill_ratios(const multi_UBspline_3d_hip *t, type_code want, const void *pos, A **rows, A **Ainv_list,
                A *ratio, int N, int rs, int col, int num, void *stream) {
  if (num < 0) return esk_fail(EINSPLINE_EINVAL, "fill_ratios: negative walker count");
  if (!t || !t->coefs) return esk_fail(EINSPLINE_EINVAL, "fill_ratios: NULL device table");
  if (t->tcode != want) return esk_fail(EINSPLINE_EINVAL, "fill_ratios: device table has the wrong type");
  if (want == SINGLE_REAL && t->storage == EINSPLINE_STORE_F64)
    return esk_fail(EINSPLINE_EINVAL, "fill_ratios: F64 storage in a single-precision table");
  if (t->dim != 3) return esk_fail(EINSPLINE_EINVAL, "fill_ratios: this entry point takes a 3-D table");
  if (N < 1 || N > t->num_splines || rs < N || col < 0 || col >= N)
    return esk_fail(EINSPLINE_EINVAL, "fill_ratios: need 1 <= N <= num_splines, N <= rowstride and "
                    "0 <= elec < N (N=%d, num_splines=%d, rowstride=%d, elec=%d)", N, t->num_splines, rs, col);
  if (num == 0) return EINSPLINE_OK;
  if (!pos || !rows || !Ainv_list || !ratio) return esk_fail(EINSPLINE_EINVAL, "fill_ratios: NULL argument");
  esk::Launch L;
  esk::Params P;
  int rc = esk_prepare(t, num, L, P);
  if (rc) return rc;
  const int eb = t->storage == EINSPLINE_STORE_F64 ? 8 : t->storage == EINSPLINE_STORE_F32 ? 4 : 2;
  const int row_bytes = t->row_reals * eb;
  if (row_bytes % 16) return esk_fail(EINSPLINE_EINVAL, "fill_ratios: table rows must be whole 16-byte pieces");
  P.chunks = row_bytes / 16;
  P.pos = pos;
  const unsigned blocks = (unsigned)((num + kThreads / 64 - 1) / (kThreads / 64));
  const hipStream_t s = (hipStream_t)stream;
  if constexpr (sizeof(A) == 8) {
    hipLaunchKernelGGL((fill_ratio<double, double>), dim3(blocks), dim3(kThreads), 0, s, P, rows, Ainv_list,
                       ratio, N, rs, col);
  } else {
    switch (t->storage) {
      case EINSPLINE_STORE_F32:
        hipLaunchKernelGGL((fill_ratio<float, float>), dim3(blocks), dim3(kThreads), 0, s, P, rows, Ainv_list,
                           ratio, N, rs, col);
        break;
      case EINSPLINE_STORE_BF16:
        hipLaunchKernelGGL((fill_ratio<esk::bf16s, float>), dim3(blocks), dim3(kThreads), 0, s, P, rows, Ainv_list,
                           ratio, N, rs, col);
        break;
      case EINSPLINE_STORE_F16:
        hipLaunchKernelGGL((fill_ratio<esk::f16s, float>), dim3(blocks), dim3(kThreads), 0, s, P, rows, Ainv_list,
                           ratio, N, rs, col);
        break;
      default:
        return esk_fail(EINSPLINE_EINVAL, "fill_ratios: unsupported storage");
    }
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EINSPLINE_OK : esk_hip_fail(e, "fill_ratios launch");
}

}  // namespace

extern "C" int einspline_update_inverse_s(float **A_list, float **Ainv_list, float **u_list,
                                          float **Ainv_delta_list, float **Ainv_colk_list, int N, int rowstride,
                                          int k, int num_mats, void *stream) {
  return update_inverse<float>(A_list, Ainv_list, u_list, Ainv_delta_list, Ainv_colk_list, N, rowstride, k,
                               num_mats, stream);
}
extern "C" int einspline_update_inverse_d(double **A_list, double **Ainv_list, double **u_list,
                                          double **Ainv_delta_list, double **Ainv_colk_list, int N,
                                          int rowstride, int k, int num_mats, void *stream) {
  return update_inverse<double>(A_list, Ainv_list, u_list, Ainv_delta_list, Ainv_colk_list, N, rowstride, k,
                                num_mats, stream);
}
extern "C" int einspline_calc_ratios_s(float **Ainv_list, float **new_row_list, float *ratio_d, int N,
                                       int rowstride, int elec, int num_mats, void *stream) {
  return calc_ratios<float>(Ainv_list, new_row_list, ratio_d, N, rowstride, elec, num_mats, stream);
}
extern "C" int einspline_calc_ratios_d(double **Ainv_list, double **new_row_list, double *ratio_d, int N,
                                       int rowstride, int elec, int num_mats, void *stream) {
  return calc_ratios<double>(Ainv_list, new_row_list, ratio_d, N, rowstride, elec, num_mats, stream);
}
extern "C" int einspline_fill_ratios_s(const multi_UBspline_3d_s_hip *spline, const float *pos_d,
                                       float **new_row_list, float **Ainv_list, float *ratio_d, int N,
                                       int rowstride, int elec, int num, void *stream) {
  return fill_ratios<float>(spline, SINGLE_REAL, pos_d, new_row_list, Ainv_list, ratio_d, N, rowstride, elec,
                            num, stream);
}
extern "C" int einspline_fill_ratios_d(const multi_UBspline_3d_d_hip *spline, const double *pos_d,
                                       double **new_row_list, double **Ainv_list, double *ratio_d, int N,
                                       int rowstride, int elec, int num, void *stream) {
  return fill_ratios<double>(spline, DOUBLE_REAL, pos_d, new_row_list, Ainv_list, ratio_d, N, rowstride, elec,
                             num, stream);
}
