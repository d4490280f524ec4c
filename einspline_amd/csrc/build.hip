// build.hip -- solving a 3-D multi-spline table on the GPU.  Samples that
// are already in device memory become a device table with no host solve and
// no host-to-device copy of coefficients (the host path is
// set_multi_UBspline_3d_* then create_multi_UBspline_3d_*_hip,
// src/multi_bspline_create.c:323-1477).
//
// The solve is the host one (multi_create.c): the same factorised per-axis
// operators, exported by einspline_export_axes, applied along x, then y, then
// z, in double precision.  Every operation is the host's, in the host's
// order, with contraction off, so the table equals the host-built one bit for
// bit.  Steps, for a batch of orbitals at a time (workspace G, doubles,
// laid out [Nx][Ny][Nz][L] with L = orbitals x components):
//   scatter_samples  orbital-major samples -> G's interior elements, through
//                    an LDS tile so reads and writes are both coalesced;
//   axis_sweep       one thread per right-hand-side column, the recurrences
//                    run along the axis (consecutive threads take
//                    consecutive columns, so every step is a coalesced row);
//   store_table      G -> the table's rows in its storage type.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "convert.hpp"
#include "einspline_private.h"
#include "runtime_internal.hpp"

namespace {

constexpr double kAlpha = 1.0 / 6.0, kBeta = 2.0 / 3.0;
constexpr int kT = 64;                 // tile: 64 columns x 64 nodes

struct Geom {
  int Mx, My, Mz, Nx, Ny, Nz;
  long L;                              // reals per element of G in this batch
};

// samples data[o][node][c] (o < cnt, node < Mx*My*Mz) -> G[e(node)][o*NC + c]
template <class Src, int NC>
__global__ void __launch_bounds__(256) scatter_samples(const Src *__restrict__ data, double *__restrict__ G,
                                                       Geom g) {
  __shared__ double tile[kT][kT + 1];
  constexpr int TO = kT / NC;          // orbitals per tile
  const long Mvol = (long)g.Mx * g.My * g.Mz;
  const long node0 = (long)blockIdx.x * kT;
  const long q0 = (long)blockIdx.y * kT;
  const long o0 = q0 / NC;
  for (int idx = threadIdx.x; idx < TO * kT * NC; idx += 256) {
    const int ol = idx / (kT * NC), f = idx - ol * (kT * NC);
    const int nl = f / NC, c = f - nl * NC;
    const long node = node0 + nl, q = q0 + (long)ol * NC + c;
    double v = 0.0;
    if (node < Mvol && q < g.L) v = (double)data[((o0 + ol) * Mvol + node) * NC + c];
    tile[ol * NC + c][nl] = v;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < kT * kT; idx += 256) {
    const int nl = idx / kT, ql = idx - nl * kT;
    const long node = node0 + nl, q = q0 + ql;
    if (node >= Mvol || q >= g.L) continue;
    const long k = node % g.Mz, ij = node / g.Mz, j = ij % g.My, i = ij / g.My;
    const long e = ((i + 1) * g.Ny + (j + 1)) * g.Nz + (k + 1);
    G[e * g.L + q] = tile[ql][nl];
  }
}

struct AxisDev {
  int M, kind;
  double sgn, zinv, l0, r2, lv[2], rv[2];
  const double *arr;                   // inv, g, sub, w: M each
};

// The passes of one column (element accessor EL), as the host's axis solve.
template <class ElFn>
__device__ __forceinline__ void sweep_column(ElFn &EL, const AxisDev &op, int M, const double *inv,
                                             const double *gg, const double *sub, const double *w, long q,
                                             int ncomp) {
#pragma clang fp contract(off)
  if (op.kind != 2) {
    const int m = M - 1;
    double prev = EL(1) * inv[0];
    EL(1) = prev;
    for (int r = 1; r < m; ++r) {
      const double x = (EL(r + 1) - sub[r] * prev) * inv[r];
      EL(r + 1) = x;
      prev = x;
    }
    const double ye = prev;                           // element m after the forward pass
    double next = prev;
    for (int r = m - 2; r >= 0; --r) {
      const double x = EL(r + 1) - gg[r] * next;
      EL(r + 1) = x;
      next = x;
    }
    const double y0 = m >= 2 ? next : ye;             // element 1 after the backward pass
    const double sa = op.sgn * kAlpha;
    const double z = (EL(M) - kAlpha * ye - sa * y0) * op.zinv;
    EL(M) = z;
    double c1 = 0.0, c2 = 0.0;
    for (int r = 0; r < m; ++r) {
      const double x = EL(r + 1) - w[r] * z;
      EL(r + 1) = x;
      if (r == 0) c1 = x;
      if (r == 1) c2 = x;
    }
    if (m < 2) c2 = z;                                // M == 2: element 2 is the last unknown
    const double s = op.sgn;
    EL(0) = s * z;
    EL(M + 1) = s * c1;
    EL(M + 2) = s * c2;
    return;
  }
  // derivative conditions
  const int comp = (int)(q % ncomp);
  const double kl = op.l0 / kAlpha, kr = op.r2 / kAlpha;
  const double d0 = EL(1), dl = EL(M);
  EL(1) = op.lv[comp] - kl * d0;
  EL(M) = op.rv[comp] - kr * dl;
  double prev = EL(1) * inv[0];
  EL(1) = prev;
  for (int r = 1; r < M; ++r) {
    const double x = (EL(r + 1) - sub[r] * prev) * inv[r];
    EL(r + 1) = x;
    prev = x;
  }
  double next = prev;
  for (int r = M - 2; r >= 0; --r) {
    const double x = EL(r + 1) - gg[r] * next;
    EL(r + 1) = x;
    next = x;
  }
  const double c1 = EL(1), c2 = EL(2), cM = EL(M), cMm = EL(M - 1);
  EL(0) = (d0 - kBeta * c1 - kAlpha * c2) / kAlpha;
  EL(M + 1) = (dl - kAlpha * cMm - kBeta * cM) / kAlpha;
}

// One column per thread: elements e of the column are col[e*es], e < n (the
// axis length of G).  Columns: line * line_stride + q for line < nlines,
// q < len.  With LDS the thread copies its column into the block's dynamic
// LDS (element e of thread t at e * blockDim + t), runs the passes there and
// copies it back: one read and one write of G per axis instead of one per
// pass.  The arithmetic is the same in both forms.
template <bool LDS>
__global__ void __launch_bounds__(256) axis_sweep(double *__restrict__ G, AxisDev op, long nlines,
                                                  long line_stride, long es, long len, int ncomp, int n) {
#pragma clang fp contract(off)
  extern __shared__ double cl[];
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nlines * len) return;
  const long line = id / len, q = id - line * len;
  double *col = G + line * line_stride + q;
  const int M = op.M;
  const double *inv = op.arr, *gg = op.arr + M, *sub = op.arr + 2 * M, *w = op.arr + 3 * M;
  const int T = blockDim.x, t = threadIdx.x;
  auto EL = [&](long e) -> double & {
    if constexpr (LDS) return cl[e * T + t];
    else return col[e * es];
  };
  if constexpr (LDS) {
    // eight loads in flight per thread
    int e = 0;
    for (; e + 8 <= n; e += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = col[(long)(e + u) * es];
#pragma unroll
      for (int u = 0; u < 8; ++u) cl[(long)(e + u) * T + t] = v[u];
    }
    for (; e < n; ++e) cl[(long)e * T + t] = col[(long)e * es];
  }
  sweep_column(EL, op, M, inv, gg, sub, w, q, ncomp);
  if constexpr (LDS) {
    for (int e = 0; e < n; ++e) col[(long)e * es] = cl[(long)e * T + t];
  }
}

// G[e][q] (q < L) -> table row e, reals col0 + q, in the storage type
__global__ void __launch_bounds__(256) store_table(const double *__restrict__ G, void *table, long vol, long L,
                                                   long row, long col0, int storage, int *ovf_flag) {
  int ovf = 0;
  for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < vol * L; idx += (long)gridDim.x * 256) {
    const long e = idx / L, q = idx - e * L;
    const double v = G[idx];
    const long o = e * row + col0 + q;
    switch (storage) {
      case EINSPLINE_STORE_F64: ((double *)table)[o] = v; break;
      case EINSPLINE_STORE_F32: ((float *)table)[o] = (float)v; break;
      case EINSPLINE_STORE_BF16: ((uint16_t *)table)[o] = esk::f32_to_bf16(esk::f64_to_f32_odd(v), ovf); break;
      case EINSPLINE_STORE_F16: ((uint16_t *)table)[o] = esk::f32_to_f16(esk::f64_to_f32_odd(v), ovf); break;
    }
  }
  if (ovf) atomicOr(ovf_flag, 1);
}

unsigned grid1(long n) { return (unsigned)std::min<long>((n + 255) / 256, 1l << 30); }

constexpr size_t kSweepLds = 80 * 1024;   // LDS per sweep block: two blocks per CU

}  // namespace

extern "C" multi_UBspline_3d_hip *einspline_build_table_hip(void *spline, const void *data_d, int storage,
                                                           void *stream) {
  einspline_host3d *h = einspline_host_of(spline);
  if (!h || h->dim != 3) { esk_fail(EINSPLINE_EINVAL, "build_table: not a 3-D spline of this library"); return nullptr; }
  if (!data_d) { esk_fail(EINSPLINE_EINVAL, "build_table: NULL samples"); return nullptr; }
  const type_code kind = h->pub.s.tcode;
  const int ncomp = h->ncomp;
  const bool src_float = kind == SINGLE_REAL || kind == SINGLE_COMPLEX;
  einspline_storage st;
  type_code tcode = kind;
  if (storage < 0) {
    st = src_float ? EINSPLINE_STORE_F32 : EINSPLINE_STORE_F64;
  } else {
    st = (einspline_storage)storage;
    const bool ok = (st == EINSPLINE_STORE_F32 && kind != DOUBLE_REAL && kind != DOUBLE_COMPLEX) ||
                    (st == EINSPLINE_STORE_F64 && !src_float) ||
                    (kind == DOUBLE_REAL && (st == EINSPLINE_STORE_F32 || st == EINSPLINE_STORE_BF16 ||
                                             st == EINSPLINE_STORE_F16)) ||
                    (kind == DOUBLE_COMPLEX && st == EINSPLINE_STORE_F32);
    if (!ok) { esk_fail(EINSPLINE_EINVAL, "build_table: storage %d not available for this spline type", storage); return nullptr; }
    if (kind == DOUBLE_REAL && st != EINSPLINE_STORE_F64) tcode = SINGLE_REAL;
    if (kind == DOUBLE_COMPLEX && st == EINSPLINE_STORE_F32) tcode = SINGLE_COMPLEX;
  }
  einspline_axis_export ax[3];
  const int xrc = einspline_export_axes(spline, ax);
  if (xrc) {
    esk_fail(xrc == -2 ? EINSPLINE_ENOMEM : EINSPLINE_EINVAL, "build_table: unsupported boundary conditions");
    return nullptr;
  }
  struct AxFree { einspline_axis_export *a; ~AxFree() { for (int i = 0; i < 3; ++i) free(a[i].arrays); } } axfree{ax};

  const multi_UBspline_3d_s &pub = h->pub.s;     // grids sit at the same offsets in every 3-D struct
  const int N = h->nsplines, Nx = h->Nx, Ny = h->Ny, Nz = h->Nz;
  const size_t eb = st == EINSPLINE_STORE_F64 ? 8 : st == EINSPLINE_STORE_F32 ? 4 : 2;
  const size_t nreal = (size_t)N * ncomp;
  const size_t per32 = 32 / eb;
  const size_t row = (nreal + per32 - 1) / per32 * per32;
  const size_t vol = (size_t)Nx * Ny * Nz;
  const size_t bytes = vol * row * eb;
  const hipStream_t s = (hipStream_t)stream;
  // orbitals per pass: at most 256 MiB of double workspace (each pass's
  // kernels still stream hundreds of MB, enough to fill the GPU)
  const size_t per_orb = vol * ncomp * sizeof(double);
  const int per = (int)std::max<size_t>(1, std::min<size_t>((size_t)N, ((size_t)256 << 20) / per_orb));

  void *table = nullptr, *dax = nullptr;
  double *G = nullptr;
  int *dovf = nullptr;
  int rc = EINSPLINE_OK;
  hipError_t e;
  multi_UBspline_3d_hip meta;
  memset(&meta, 0, sizeof meta);
  if ((e = hipMalloc(&table, bytes)) != hipSuccess) {
    esk_fail(EINSPLINE_ENOMEM, "build_table: hipMalloc of %zu bytes: %s", bytes, hipGetErrorString(e));
    return nullptr;
  }
  const size_t axbytes = sizeof(double) * 4 * (size_t)(ax[0].M + ax[1].M + ax[2].M);
  if ((e = hipMalloc((void **)&G, per_orb * per)) != hipSuccess || (e = hipMalloc(&dax, axbytes)) != hipSuccess ||
      (e = hipMalloc((void **)&dovf, sizeof(int))) != hipSuccess)
    rc = esk_fail(EINSPLINE_ENOMEM, "build_table: workspace: %s", hipGetErrorString(e));
  AxisDev op[3];
  if (rc == EINSPLINE_OK) {
    size_t off = 0;
    for (int a = 0; a < 3; ++a) {
      const size_t n = 4 * (size_t)ax[a].M;
      if ((e = hipMemcpyAsync((double *)dax + off, ax[a].arrays, n * sizeof(double), hipMemcpyHostToDevice, s)) !=
          hipSuccess) {
        rc = esk_hip_fail(e, "build_table: operator copy");
        break;
      }
      op[a].M = ax[a].M; op[a].kind = ax[a].kind; op[a].sgn = ax[a].sgn; op[a].zinv = ax[a].zinv;
      op[a].l0 = ax[a].l0; op[a].r2 = ax[a].r2;
      op[a].lv[0] = ax[a].lv[0]; op[a].lv[1] = ax[a].lv[1]; op[a].rv[0] = ax[a].rv[0]; op[a].rv[1] = ax[a].rv[1];
      op[a].arr = (const double *)dax + off;
      off += n;
    }
  }
  if (rc == EINSPLINE_OK && ((e = hipMemsetAsync(table, 0, bytes, s)) != hipSuccess ||
                             (e = hipMemsetAsync(dovf, 0, sizeof(int), s)) != hipSuccess))
    rc = esk_hip_fail(e, "build_table: clear");
  // One sweep along an axis of n coefficients: columns in LDS when the
  // passes touch exactly the column's n elements and a block of at least 64
  // columns fits the budget (two blocks per CU), else from G directly.
  auto sweep = [&](const AxisDev &o, long nlines, long lstride, long es, long len, int n) {
    const long cols = nlines * len;
    const bool exact = (o.kind == 2 ? o.M + 2 : o.M + 3) == n;
    int T = 256;
    while (T > 64 && (size_t)n * T * sizeof(double) > kSweepLds) T /= 2;
    const size_t lds = (size_t)n * T * sizeof(double);
    if (exact && lds <= kSweepLds) {
      hipLaunchKernelGGL(axis_sweep<true>, dim3((unsigned)std::min<long>((cols + T - 1) / T, 1l << 30)), dim3(T),
                         lds, s, G, o, nlines, lstride, es, len, ncomp, n);
    } else {
      hipLaunchKernelGGL(axis_sweep<false>, dim3(grid1(cols)), dim3(256), 0, s, G, o, nlines, lstride, es, len,
                         ncomp, n);
    }
  };
  if (rc == EINSPLINE_OK &&
      (e = hipFuncSetAttribute((const void *)axis_sweep<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kSweepLds)) != hipSuccess)
    rc = esk_hip_fail(e, "build_table: LDS size");
  const Geom g0{ax[0].M, ax[1].M, ax[2].M, Nx, Ny, Nz, 0};
  const long Mvol = (long)g0.Mx * g0.My * g0.Mz;
  for (int o0 = 0; rc == EINSPLINE_OK && o0 < N; o0 += per) {
    const int cnt = std::min(per, N - o0);
    Geom g = g0;
    g.L = (long)cnt * ncomp;
    const long L = g.L;
    if ((e = hipMemsetAsync(G, 0, vol * L * sizeof(double), s)) != hipSuccess) { rc = esk_hip_fail(e, "build_table: clear"); break; }
    const dim3 tg((unsigned)((Mvol + kT - 1) / kT), (unsigned)((L + kT - 1) / kT));
    const size_t src_off = (size_t)o0 * Mvol * ncomp;
    if (src_float) {
      const float *src = (const float *)data_d + src_off;
      if (ncomp == 1) hipLaunchKernelGGL((scatter_samples<float, 1>), tg, dim3(256), 0, s, src, G, g);
      else hipLaunchKernelGGL((scatter_samples<float, 2>), tg, dim3(256), 0, s, src, G, g);
    } else {
      const double *src = (const double *)data_d + src_off;
      if (ncomp == 1) hipLaunchKernelGGL((scatter_samples<double, 1>), tg, dim3(256), 0, s, src, G, g);
      else hipLaunchKernelGGL((scatter_samples<double, 2>), tg, dim3(256), 0, s, src, G, g);
    }
    // x: one line of Ny*Nz*L columns; y: Nx lines of Nz*L; z: Nx*Ny lines of L
    const long plane = (long)Ny * Nz * L;
    sweep(op[0], 1l, 0l, plane, plane, Nx);
    sweep(op[1], (long)Nx, plane, (long)Nz * L, (long)Nz * L, Ny);
    sweep(op[2], (long)Nx * Ny, (long)Nz * L, L, L, Nz);
    hipLaunchKernelGGL(store_table, dim3(std::min<unsigned>(grid1((long)vol * L), 16384u)), dim3(256), 0, s, G,
                       table, (long)vol, L, (long)row, (long)o0 * ncomp, (int)st, dovf);
    if ((e = hipGetLastError()) != hipSuccess) rc = esk_hip_fail(e, "build_table: launch");
  }
  int ovf = 0;
  if (rc == EINSPLINE_OK && (e = hipMemcpyAsync(&ovf, dovf, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess)
    rc = esk_hip_fail(e, "build_table: flag");
  if ((e = hipStreamSynchronize(s)) != hipSuccess && rc == EINSPLINE_OK) rc = esk_hip_fail(e, "build_table");
  if (G) (void)hipFree(G);
  if (dax) (void)hipFree(dax);
  if (dovf) (void)hipFree(dovf);
  if (rc == EINSPLINE_OK && ovf) rc = esk_fail(EINSPLINE_ERANGE, "build_table: coefficient outside the range of the storage type");
  if (rc != EINSPLINE_OK) { (void)hipFree(table); return nullptr; }
  meta.z_stride = (intptr_t)row;
  meta.y_stride = (intptr_t)Nz * row;
  meta.x_stride = (intptr_t)Ny * Nz * row;
  meta.x_grid = pub.x_grid; meta.y_grid = pub.y_grid; meta.z_grid = pub.z_grid;
  meta.Nx = Nx; meta.Ny = Ny; meta.Nz = Nz;
  meta.num_splines = N;
  meta.row_reals = (int)row;
  meta.tcode = tcode;
  meta.storage = st;
  meta.bytes = bytes;
  meta.dim = 3;
  if ((e = hipGetDevice(&meta.device)) != hipSuccess) { esk_hip_fail(e, "hipGetDevice"); (void)hipFree(table); return nullptr; }
  multi_UBspline_3d_hip *t = esk_table_wrap(&meta, table);
  if (!t) (void)hipFree(table);
  return t;
}
