// eval12_kernel.hpp -- gfx950 kernels for batched 1-D and 2-D multi-spline
// evaluation.
//
// Same work split as the 3-D kernel: one lane owns LB contiguous bytes of an
// orbital row, and the lanes of a point are consecutive, so every stencil
// row is one coalesced load.  A 1-D point reads 4 rows, a 2-D point 16.  The
// contraction follows src/multi_bspline_eval_std_d_impl.h:37-180 (1-D) and
// :183-476 (2-D); derivatives are scaled by delta_inv as there.
//
// Output layouts (per point, in result elements):
//   1-D: vals[N], grads[N], lapl[N]; the 1-D "vgh" is the vgl (hess[N] is the
//        second derivative), as in the reference (std_d_impl.h:168-180);
//   2-D: vals[N], grads[2N] ([n][2]), lapl[N], hess[4N] (row-major 2x2).
// The pointer-array form used by the 1-D CUDA-style entry points writes the
// Laplacian either to its own array (out_h) or, when out_h is null, as the
// row `row_stride` reals after the gradient row.
#pragma once
#include "eval_kernel.hpp"

namespace esk {

// Accumulators: 1-D v, g, l; 2-D v, gx, gy, then lxx, lyy (vgl) or hxx,
// hxy, hyy (vgh).
template <int DIM, int MODE> struct Acc12 {
  static constexpr int N = DIM == 1 ? (MODE == MODE_V ? 1 : MODE == MODE_VG ? 2 : 3)
                                    : (MODE == MODE_V ? 1 : MODE == MODE_VG ? 3 : MODE == MODE_VGL ? 5 : 6);
};

template <class S, int LB, int MODE, bool CPLX, bool UNIF, int DIM, class PosT, bool PTRS>
__device__ __forceinline__ void eval12_chunk(const Params &P, int p, int ch) {
  using A = typename acc_of<S>::type;
  using L = Loader<S, LB>;
  constexpr int V = L::V;
  constexpr int NA = Acc12<DIM, MODE>::N;
  static_assert(DIM == 1 || DIM == 2, "1-D or 2-D");
  static_assert(!CPLX || (V % 2 == 0), "complex lanes hold whole orbitals");

  const PosT *pp = (const PosT *)P.pos + (long)DIM * p;
  int ix = 0, iy = 0;
  A tx = 0, ty = 0;
  locate<A>((double)pp[0], P.x0, P.xi, P.nx4, ix, tx);
  if constexpr (DIM == 2) locate<A>((double)pp[1], P.y0, P.yi, P.ny4, iy, ty);
  if (UNIF) {
    ix = uniform(ix); iy = uniform(iy);
    tx = uniform(tx); ty = uniform(ty);
  }
  A a[4], da[4], d2a[4], b[4], db[4], d2b[4];
  basis0(tx, a);
  if constexpr (MODE >= MODE_VG) basis1(tx, da);
  if constexpr (MODE >= MODE_VGL) basis2(tx, d2a);
  if constexpr (DIM == 2) {
    basis0(ty, b);
    if constexpr (MODE >= MODE_VG) basis1(ty, db);
    if constexpr (MODE >= MODE_VGL) basis2(ty, d2b);
  }
  if (UNIF) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = uniform(a[q]);
      if constexpr (MODE >= MODE_VG) da[q] = uniform(da[q]);
      if constexpr (MODE >= MODE_VGL) d2a[q] = uniform(d2a[q]);
      if constexpr (DIM == 2) {
        b[q] = uniform(b[q]);
        if constexpr (MODE >= MODE_VG) db[q] = uniform(db[q]);
        if constexpr (MODE >= MODE_VGL) d2b[q] = uniform(d2b[q]);
      }
    }
  }

  A acc[NA][V];
#pragma unroll
  for (int q = 0; q < NA; ++q)
#pragma unroll
    for (int e = 0; e < V; ++e) acc[q][e] = (A)0;

  const S *base = (const S *)P.coefs + (long)ix * P.xs + (long)iy * P.ys + (long)ch * V;
  if constexpr (DIM == 1) {
    A r[4][V];
#pragma unroll
    for (int i = 0; i < 4; ++i) L::load(base + i * P.xs, r[i]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e = 0; e < V; ++e) {
        acc[0][e] = fmaA(a[i], r[i][e], acc[0][e]);
        if constexpr (MODE >= MODE_VG) acc[1][e] = fmaA(da[i], r[i][e], acc[1][e]);
        if constexpr (MODE >= MODE_VGL) acc[2][e] = fmaA(d2a[i], r[i][e], acc[2][e]);
      }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      A r[4][V];
#pragma unroll
      for (int j = 0; j < 4; ++j) L::load(base + i * P.xs + j * P.ys, r[j]);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const A ab = a[i] * b[j];
#pragma unroll
        for (int e = 0; e < V; ++e) acc[0][e] = fmaA(ab, r[j][e], acc[0][e]);
        if constexpr (MODE >= MODE_VG) {
          const A dab = da[i] * b[j], adb = a[i] * db[j];
#pragma unroll
          for (int e = 0; e < V; ++e) {
            acc[1][e] = fmaA(dab, r[j][e], acc[1][e]);
            acc[2][e] = fmaA(adb, r[j][e], acc[2][e]);
          }
        }
        if constexpr (MODE == MODE_VGL) {
          const A d2ab = d2a[i] * b[j], ad2b = a[i] * d2b[j];
#pragma unroll
          for (int e = 0; e < V; ++e) {
            acc[3][e] = fmaA(d2ab, r[j][e], acc[3][e]);
            acc[4][e] = fmaA(ad2b, r[j][e], acc[4][e]);
          }
        }
        if constexpr (MODE == MODE_VGH) {
          const A d2ab = d2a[i] * b[j], dadb = da[i] * db[j], ad2b = a[i] * d2b[j];
#pragma unroll
          for (int e = 0; e < V; ++e) {
            acc[3][e] = fmaA(d2ab, r[j][e], acc[3][e]);
            acc[4][e] = fmaA(dadb, r[j][e], acc[4][e]);
            acc[5][e] = fmaA(ad2b, r[j][e], acc[5][e]);
          }
        }
      }
    }
  }

  // scale by delta_inv (rounded to A first, as the reference does)
  const A dxi = (A)P.xi, dyi = (A)P.yi;
#pragma unroll
  for (int e = 0; e < V; ++e) {
    if constexpr (DIM == 1) {
      if constexpr (MODE >= MODE_VG) acc[1][e] *= dxi;
      if constexpr (MODE >= MODE_VGL) acc[2][e] *= dxi * dxi;
    } else {
      if constexpr (MODE >= MODE_VG) { acc[1][e] *= dxi; acc[2][e] *= dyi; }
      if constexpr (MODE == MODE_VGL) {
        const A lx = acc[3][e] * (dxi * dxi), ly = acc[4][e] * (dyi * dyi);
        acc[3][e] = lx + ly;
      }
      if constexpr (MODE == MODE_VGH) {
        acc[3][e] *= dxi * dxi; acc[4][e] *= dxi * dyi; acc[5][e] *= dyi * dyi;
      }
    }
  }

  // ---- outputs ----
  A *ov, *og = nullptr, *oh = nullptr;
  if constexpr (PTRS) {
    ov = ((A *const *)P.out_v)[p];
    if (P.out_g) og = ((A *const *)P.out_g)[p];
    if (P.out_h) oh = ((A *const *)P.out_h)[p];
  } else {
    ov = (A *)P.out_v + (long)p * P.ld_v;
    if constexpr (MODE >= MODE_VG) og = (A *)P.out_g + (long)p * P.ld_g;
    if constexpr (MODE >= MODE_VGL) oh = (A *)P.out_h + (long)p * P.ld_h;
  }
  const int e0 = ch * V;
  const int m = P.nreal - e0;
  const bool full = m >= V;
  uintptr_t align = (uintptr_t)ov | (uintptr_t)og | (uintptr_t)oh;
  const bool vec = (align & 15) == 0;
  {
    A r[V];
#pragma unroll
    for (int e = 0; e < V; ++e) r[e] = acc[0][e];
    if (full) store_block<A, V>(ov, e0, r, vec, false); else store_partial<A, V>(ov + e0, r, m);
  }
  if constexpr (MODE >= MODE_VG) {
    A r[DIM * V];
#pragma unroll
    for (int e = 0; e < V; ++e)
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int o = CPLX ? (e >> 1) * 2 * DIM + d * 2 + (e & 1) : e * DIM + d;
        r[o] = acc[1 + d][e];
      }
    if (full) store_block<A, DIM * V>(og, DIM * e0, r, vec, false);
    else store_partial<A, DIM * V>(og + DIM * e0, r, DIM * m);
  }
  if constexpr (MODE == MODE_VGL || (MODE == MODE_VGH && DIM == 1)) {
    // Laplacian (1-D: the second derivative)
    constexpr int LI = DIM == 1 ? 2 : 3;
    A r[V];
#pragma unroll
    for (int e = 0; e < V; ++e) r[e] = acc[LI][e];
    A *ol = oh ? oh : og + P.row_stride;      // pointer form without a lapl array: next row
    const bool lvec = vec && (((uintptr_t)ol & 15) == 0);
    if (full) store_block<A, V>(ol, e0, r, lvec, false); else store_partial<A, V>(ol + e0, r, m);
  }
  if constexpr (MODE == MODE_VGH && DIM == 2) {
    constexpr int full4[4] = {3, 4, 4, 5};
    A r[4 * V];
#pragma unroll
    for (int e = 0; e < V; ++e)
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int o = CPLX ? (e >> 1) * 8 + h * 2 + (e & 1) : e * 4 + h;
        r[o] = acc[full4[h]][e];
      }
    if (full) store_block<A, 4 * V>(oh, 4 * e0, r, vec, false);
    else store_partial<A, 4 * V>(oh + 4 * e0, r, 4 * m);
  }
}

template <class S, int LB, int MODE, bool CPLX, bool UNIF, int DIM, class PosT, bool PTRS>
__global__ void __launch_bounds__(256) eval12(const Params P) {
  const unsigned gid = blockIdx.x * blockDim.x + threadIdx.x;
  int p = (int)(gid / (unsigned)P.chunks);
  if (UNIF) p = uniform(p);
  const int ch = (int)(gid - (unsigned)p * (unsigned)P.chunks);
  if (p >= P.npoints) return;
  eval12_chunk<S, LB, MODE, CPLX, UNIF, DIM, PosT, PTRS>(P, p, ch);
}

}  // namespace esk
