// mfma_kernel.hpp -- vgh of 16-bit (BF16 / FP16) tables on the matrix cores.
//
// For one point the outputs of every orbital are a small matrix product:
//   D[q][n] = sum_K W[q][K] P[K][n],   K < 64 (the stencil rows,
//   K = 16 i + 4 j + k), n < orbitals,
// where W[q][K] = X_q(i) Y_q(j) Z_q(k) is a product of basis values and
// P[K] is stencil row K.  The rows q are the point's outputs in the order
// they are stored per orbital: v, the 3 gradient entries, then the Hessian
// as stored (all 9 entries row-major, or the 6 of the packed form), so 13
// or 10 of the 16 rows of a tile are used; the duplicated Hessian rows cost
// nothing extra.  One wavefront takes one point and runs it as 16 x 16 x 32
// MFMA tiles: W is the A operand, the stencil rows are the B operand,
// accumulation is FP32.
//
// Operand layout of v_mfma_f32_16x16x32_{bf16,f16} (lane l, g = l / 16,
// r = l % 16): A holds W[r][8g .. 8g+7], B holds P[8g .. 8g+7][column r], D
// holds D[4g .. 4g+3][column r].  A lane's B column must therefore be eight
// stencil rows of one orbital, while memory holds rows of consecutive
// orbitals.  Lane (g, r) loads eight rows (8g .. 8g+7 of the 32-row chunk),
// 16 bytes each at orbitals 8r .. 8r+7 of a 128-orbital slab: four lane
// groups read four row blocks of 256 contiguous bytes.  Tile t (0..7) then
// takes column r to be orbital 8r + t, and the lane builds its B fragment
// from element t of each of its eight rows with byte permutes: an in-lane
// 8 x 8 transpose instead of an exchange between lanes.
//
// The weights are split into a 16-bit high part and a 16-bit remainder and
// both are applied (two MFMAs per tile), so they carry ~16 mantissa bits;
// the result's error is the table's own 16-bit rounding, as on the vector
// path.
//
// Results leave through LDS: lane (g, r) holds rows 4g .. 4g+3 of its
// orbitals, and writes them to the wavefront's staging area at their places
// in the output layout; once two slabs are staged the wavefront copies their
// v, gradient and Hessian blocks out as contiguous runs (16-byte stores
// where aligned).
#pragma once
#include <type_traits>
#include "eval_kernel.hpp"

namespace esk {

template <class S> struct Mfma16;
template <> struct Mfma16<bf16s> {
  typedef __bf16 h8 __attribute__((ext_vector_type(8)));
  __device__ static __forceinline__ __bf16 narrow(float f) { return (__bf16)f; }
  __device__ static __forceinline__ float back(__bf16 h) { return (float)h; }
  typedef float f4 __attribute__((ext_vector_type(4)));
  __device__ static __forceinline__ f4 mma(h8 a, h8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma16<f16s> {
  typedef _Float16 h8 __attribute__((ext_vector_type(8)));
  __device__ static __forceinline__ _Float16 narrow(float f) { return (_Float16)f; }
  __device__ static __forceinline__ float back(_Float16 h) { return (float)h; }
  typedef float f4 __attribute__((ext_vector_type(4)));
  __device__ static __forceinline__ f4 mma(h8 a, h8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

// Pack the 16-bit halves `t & 1` of x and y into one dword (x low, y high).
__device__ __forceinline__ uint32_t pick16(uint32_t x, uint32_t y, int hi) {
  return __builtin_amdgcn_perm(y, x, hi ? 0x07060302u : 0x05040100u);
}

// Derivative orders (x, y, z) of output row q, two bits each, and the row
// count: full Hessian rows v gx gy gz xx xy xz yx yy yz zx zy zz; packed rows
// v gx gy gz xx xy xz yy yz zz.
template <bool PACKED> struct MfmaRows;
template <> struct MfmaRows<false> {
  static constexpr int NQ = 13, HS = 9;
  static constexpr uint32_t OX = 0u | 1u << 2 | 2u << 8 | 1u << 10 | 1u << 12 | 1u << 14 | 1u << 20;
  static constexpr uint32_t OY = 1u << 4 | 1u << 10 | 1u << 14 | 2u << 16 | 1u << 18 | 1u << 22;
  static constexpr uint32_t OZ = 1u << 6 | 1u << 12 | 1u << 18 | 1u << 20 | 1u << 22 | 2u << 24;
};
template <> struct MfmaRows<true> {
  static constexpr int NQ = 10, HS = 6;
  static constexpr uint32_t OX = 1u << 2 | 2u << 8 | 1u << 10 | 1u << 12;
  static constexpr uint32_t OY = 1u << 4 | 1u << 10 | 2u << 14 | 1u << 16;
  static constexpr uint32_t OZ = 1u << 6 | 1u << 12 | 1u << 16 | 2u << 18;
};

// Floats of one wavefront's staging area: v, gradient and Hessian blocks of
// two 128-orbital slabs in output layout (the full Hessian, the larger case).
// Four wavefronts take 53 KB: three blocks fill a CU's 160 KB, as many as
// the registers allow.
constexpr int kMfmaStage = 256 * (1 + 3 + 9);

// The wavefront's point (blocks of four wavefronts, XCD runs and the
// traversal order as in eval3d): its cell and basis values (value, first and
// second derivative weights per axis).  False past the last point.
__device__ __forceinline__ bool mfma_point(const Params &P, int &p, int &slab, int &ix, int &iy, int &iz,
                                           float (&wx)[3][4], float (&wy)[3][4], float (&wz)[3][4]) {
  slab = -1;                          // all slabs of the point
  if (P.slabs > 1) {                  // orbital slabs outermost: one 256-orbital slab per wavefront
    if (!slab_wave(P, blockIdx.x, blockDim.x >> 6, threadIdx.x >> 6, p, slab)) return false;
    p = uniform(p);                   // wave-uniform by construction; said so to the compiler
    slab = uniform(slab);
  } else {
    const unsigned b = logical_block(blockIdx.x, gridDim.x, P.xcd_runs);
    p = uniform((int)(b * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    if (p >= P.npoints) return false;
  }
  if (P.perm) p = uniform(P.perm[p]);
  const double *pp = (const double *)P.pos + 3l * p;
  float tx, ty, tz;
  locate<float>(pp[0], P.x0, P.xi, P.nx4, ix, tx);
  locate<float>(pp[1], P.y0, P.yi, P.ny4, iy, ty);
  locate<float>(pp[2], P.z0, P.zi, P.nz4, iz, tz);
  ix = uniform(ix); iy = uniform(iy); iz = uniform(iz);
  tx = uniform(tx); ty = uniform(ty); tz = uniform(tz);
  basis0(tx, wx[0]); basis1(tx, wx[1]); basis2(tx, wx[2]);
  basis0(ty, wy[0]); basis1(ty, wy[1]); basis2(ty, wy[2]);
  basis0(tz, wz[0]); basis1(tz, wz[1]); basis2(tz, wz[2]);
  return true;
}

// Output scale of accumulator entries e = 0..3 of lane group g (output rows
// 4g + e): derivatives in physical units, times `oscale`.
template <class Rows>
__device__ __forceinline__ void mfma_scales(const Params &P, int g, float oscale, float (&sc)[4]) {
  const float dxi = (float)P.xi, dyi = (float)P.yi, dzi = (float)P.zi;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int qe = 4 * g + e;
    const int ex = (Rows::OX >> (2 * qe)) & 3, ey = (Rows::OY >> (2 * qe)) & 3, ez = (Rows::OZ >> (2 * qe)) & 3;
    float f = oscale;
    f = ex >= 1 ? f * dxi : f; f = ex == 2 ? f * dxi : f;
    f = ey >= 1 ? f * dyi : f; f = ey == 2 ? f * dyi : f;
    f = ez >= 1 ? f * dzi : f; f = ez == 2 ? f * dzi : f;
    sc[e] = f;
  }
}

// Choices are made between values, never between array elements: a
// conditional over two lvalues selects an address, and the array then lives
// in scratch memory.
__device__ __forceinline__ float by_order(int o, float w0, float w1, float w2) {
  return o == 0 ? w0 : o == 1 ? w1 : w2;
}

// Copy n floats of a staged block to dst: 16-byte pieces when dst is
// aligned.
__device__ __forceinline__ void mfma_copy_out(int lane, const float *src, float *dst, int n) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  int done = 0;
  if (((uintptr_t)dst & 15) == 0) {
    const int n4 = n >> 2;
    for (int i = lane; i < n4; i += 64) *(v4 *)(dst + 4 * i) = *(const v4 *)(src + 4 * i);
    done = 4 * n4;
  }
  for (int i = done + lane; i < n; i += 64) dst[i] = src[i];
}

// Streaming copy of N floats (a multiple of 128) from the staging area to
// 16-byte aligned dst: 16-byte pieces, consecutive lanes filling whole
// 128-byte lines, every instruction but a last half-wavefront one executed by
// all 64 lanes.
template <int N>
__device__ __forceinline__ void mfma_stream_out(int lane, const float *src, float *dst) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static_assert(N % 128 == 0, "whole lines for every group of eight lanes");
#pragma unroll
  for (int k = 0; k < N / 256; ++k) st_stream(*(const v4 *)(src + 256 * k + 4 * lane), dst, 1024u * k + 16u * lane);
  if constexpr (N % 256 != 0)
    if (lane < 32) st_stream(*(const v4 *)(src + N - 128 + 4 * lane), dst, 4u * (N - 128) + 16u * lane);
}

// Register budget: three wavefronts per SIMD (at most 168 VGPRs; the
// kernel needs about 150 without spills, while four would spill).
//
// STREAM: the launch guarantees that every whole slab's destinations are
// 16-byte aligned and asks for streaming stores, so whole slabs take the one
// fixed store sequence.  With the store variants chosen at run time inside
// the slab loop, the compiler's count of outstanding memory operations at
// the loop head is the most conservative over the variants' paths, and it
// then waits for the previous slabs' stores (vector-memory counts include
// stores) before the next rows.  With one path the waits are exact and the
// stores stay in flight.
template <class S, bool PACKED, bool STREAM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
eval3d_mfma(const Params P) {
  using M = Mfma16<S>;
  using Rows = MfmaRows<PACKED>;
  typedef typename M::h8 h8;
  typedef typename M::f4 f4;
  __shared__ __align__(16) float stage_all[4 * kMfmaStage];
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, r = lane & 15;
  float *const stage = stage_all + (threadIdx.x >> 6) * kMfmaStage;
  int p, slab, ix, iy, iz;
  float wx[3][4], wy[3][4], wz[3][4];
  if (!mfma_point(P, p, slab, ix, iy, iz, wx, wy, wz)) return;

  // FP16 holds magnitudes up to 65504 only: points outside the grid box
  // extrapolate with weights that grow like t^3 per axis, so products bounded
  // above 1024 are scaled by a power of two (exact) that brings the largest
  // below 1, and the outputs by its inverse.  Inside the box the bound is at
  // most 8 (no scaling); BF16 has the FP32 range.
  float wscale = 1.0f, oscale = 1.0f;
  if constexpr (std::is_same<S, f16s>::value) {
    float mx = 0.0f, my = 0.0f, mz = 0.0f;
#pragma unroll
    for (int o = 0; o < 3; ++o)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        mx = fmaxf(mx, fabsf(wx[o][u]));
        my = fmaxf(my, fabsf(wy[o][u]));
        mz = fmaxf(mz, fabsf(wz[o][u]));
      }
    const float bound = mx * my * mz;         // >= |W[q][K]| for every q, K
    if (bound > 1024.0f) {
      int e;
      (void)frexpf(bound, &e);                // bound < 2^e
      wscale = ldexpf(1.0f, -e);
      oscale = ldexpf(1.0f, e);
    }
  }

  // A operands (W split into high and low 16-bit parts) for the two K chunks.
  // Lane (g, r): row q = r, columns K = 32 c + 8 g + jj, i.e. x row
  // i = 2c + g/2, y row j = 2 (g & 1) + jj/4, z row k = jj & 3.  Every
  // selection below is a compare-and-select: no register array is indexed
  // with a lane-dependent value.
  const int q = r;
  const int ox = (Rows::OX >> (2 * q)) & 3, oy = (Rows::OY >> (2 * q)) & 3, oz = (Rows::OZ >> (2 * q)) & 3;
  const bool gx1 = (g >> 1) != 0, gy1 = (g & 1) != 0;
#define ESK_PICK(o, w, u) by_order(o, w[0][u], w[1][u], w[2][u])
  const float x0 = ESK_PICK(ox, wx, 0), x1 = ESK_PICK(ox, wx, 1), x2 = ESK_PICK(ox, wx, 2), x3 = ESK_PICK(ox, wx, 3);
  const float y0 = ESK_PICK(oy, wy, 0), y1 = ESK_PICK(oy, wy, 1), y2 = ESK_PICK(oy, wy, 2), y3 = ESK_PICK(oy, wy, 3);
  const float z0 = ESK_PICK(oz, wz, 0), z1 = ESK_PICK(oz, wz, 1), z2 = ESK_PICK(oz, wz, 2), z3 = ESK_PICK(oz, wz, 3);
#undef ESK_PICK
  const float xc0 = gx1 ? x1 : x0, xc1 = gx1 ? x3 : x2;     // x row 2c + g/2 for c = 0, 1
  const float yl = gy1 ? y2 : y0, yh = gy1 ? y3 : y1;       // y row 2 (g & 1) + jj/4
  h8 ahi[2], alo[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const float xi = c == 0 ? xc0 : xc1;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const float yj = jj < 4 ? yl : yh;
      const int kk = jj & 3;
      const float zk = kk == 0 ? z0 : kk == 1 ? z1 : kk == 2 ? z2 : z3;
      const float w = q < Rows::NQ ? xi * yj * zk * wscale : 0.0f;
      const auto h = M::narrow(w);
      ahi[c][jj] = h;
      alo[c][jj] = M::narrow(w - M::back(h));
    }
  }

  // Accumulator entry e of lane (g, r) is output row 4g + e of orbital
  // 8r + t (tile t) of the slab.
  float sc[4];
  mfma_scales<Rows>(P, g, oscale, sc);
  // The staging area holds two slabs in output layout: v block (256 floats),
  // gradient block (768), Hessian block (256 HS); slab `half` (0 or 1) of a
  // pair fills the second half of each block or the first.  Lanes g = 0
  // hold v and the gradient of their eight orbitals (8 and 24 consecutive
  // floats), lanes g > 0 Hessian entries 4 (g - 1) + e at stride HS per
  // orbital; each group's writes are its base address plus constant offsets.
  constexpr int HS = Rows::HS;
  float *const vg_base = stage + 8 * r;                            // v of orbital 8r
  float *const gr_base = stage + 256 + 24 * r;                     // gradient of orbital 8r
  float *const h_base = stage + 1024 + HS * 8 * r + 4 * (g - 1);

  const S *cell = (const S *)P.coefs + (long)ix * P.xs + (long)iy * P.ys + (long)iz * P.zs;
  const int row_reals = (int)P.zs;
  const int end = min(row_reals, P.nreal);            // slabs start below this
  float *ov = (float *)P.out_v + (long)p * P.ld_v;
  float *og = (float *)P.out_g + (long)p * P.ld_g;
  float *oh = (float *)P.out_h + (long)p * P.ld_h;

  // Eight stencil rows (K = 32 c + 8 g + jj) of the lane's eight orbitals
  // from o0: one lane offset (x row g/2, y rows 2 (g & 1) + .., orbital o0)
  // plus a wave-uniform offset per row.  Rows are padded to whole 32-byte
  // pieces, so a lane whose o0 is inside the row reads inside it; a lane
  // past the row reads the row's first orbitals instead.  Column n of a
  // product depends on column n of B only, so those lanes change nothing
  // but outputs of orbitals that are never stored, and every lane issues
  // every load: no branch around them, so the compiler's wait counts stay
  // exact and a half-slab in flight is not waited for early.  (Buffer loads
  // with the same offsets ran 30% slower in the locality order on MI355X.)
  // Each address is a wave-uniform base (the cell and the row, in scalar
  // registers) plus the lane's 32-bit byte offset, so the 16 loads of a slab
  // share one offset register; the host keeps the matrix kernel to tables
  // whose offsets fit (two x planes, two y rows and a row below 2^31 bytes).
  constexpr int EB = (int)sizeof(S);
  const char *const cellb = (const char *)cell;
  const uint32_t lane_off = (uint32_t)(((g >> 1) * P.xs + 2 * (g & 1) * P.ys) * EB);
  auto load_half = [&](uint4 (&R)[8], int s0, int c) __attribute__((always_inline)) {
    const int o0 = s0 + 8 * r;
    const uint32_t loff = lane_off + (uint32_t)((o0 < row_reals ? o0 : 0) * EB);
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const long soff = (2 * c * P.xs + (jj >> 2) * P.ys + (jj & 3) * P.zs) * EB;
      R[jj] = *(const uint4 *)(cellb + soff + loff);
    }
  };
  // Tile t takes element t of each row (an in-lane 8 x 8 transpose).
  auto mma_half = [&](const uint4 (&R)[8], f4 (&acc)[8], int c) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int d = t >> 1, hi = t & 1;
      uint32_t w4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t x = d == 0 ? R[2 * u].x : d == 1 ? R[2 * u].y : d == 2 ? R[2 * u].z : R[2 * u].w;
        const uint32_t y = d == 0 ? R[2 * u + 1].x : d == 1 ? R[2 * u + 1].y : d == 2 ? R[2 * u + 1].z : R[2 * u + 1].w;
        w4[u] = pick16(x, y, hi);
      }
      h8 bf;
      __builtin_memcpy(&bf, w4, sizeof bf);
      acc[t] = M::mma(c == 0 ? ahi[0] : ahi[1], bf, acc[t]);
      acc[t] = M::mma(c == 0 ? alo[0] : alo[1], bf, acc[t]);
    }
  };
  // Software pipeline over slabs: the first half's registers are refilled
  // with the next slab's rows once its products are issued, the second's
  // once the slab is staged in LDS, so the whole next slab is in flight
  // while this one is stored.  The last slab takes no refill, which keeps
  // every load unconditional.  Scheduling barriers keep the refills where
  // they are written: left to itself the scheduler moved them behind the
  // other half's products.
  uint4 R0[8], R1[8];
  // all slabs of the point, or only the wavefront's 256 orbitals (orbital
  // slabs outermost: one pair of 128-orbital blocks per wavefront)
  const int s_first = slab < 0 ? 0 : slab * 256;
  const int s_last = slab < 0 ? end : min(end, s_first + 256);
  if (s_first >= s_last) return;
  load_half(R0, s_first, 0);
  load_half(R1, s_first, 1);
  auto slab_step = [&](int s0, bool more, int half) __attribute__((always_inline)) {
    f4 acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    mma_half(R0, acc, 0);
    __builtin_amdgcn_sched_barrier(0);
    if (more) load_half(R0, s0 + 128, 0);
    __builtin_amdgcn_sched_barrier(0);
    mma_half(R1, acc, 1);
    // stage the slab (orbitals past the valid ones are staged but not
    // copied out)
    typedef float v4 __attribute__((ext_vector_type(4)));
    float *const vb = vg_base + 128 * half, *const gb = gr_base + 384 * half, *const hb = h_base + 128 * HS * half;
    if (g == 0) {
      *(v4 *)vb = v4{acc[0][0], acc[1][0], acc[2][0], acc[3][0]} * sc[0];
      *(v4 *)(vb + 4) = v4{acc[4][0], acc[5][0], acc[6][0], acc[7][0]} * sc[0];
#pragma unroll
      for (int t = 0; t < 8; t += 4) {
        const v4 a = {acc[t][1] * sc[1], acc[t][2] * sc[2], acc[t][3] * sc[3], acc[t + 1][1] * sc[1]};
        const v4 b = {acc[t + 1][2] * sc[2], acc[t + 1][3] * sc[3], acc[t + 2][1] * sc[1], acc[t + 2][2] * sc[2]};
        const v4 c = {acc[t + 2][3] * sc[3], acc[t + 3][1] * sc[1], acc[t + 3][2] * sc[2], acc[t + 3][3] * sc[3]};
        *(v4 *)(gb + 3 * t) = a;
        *(v4 *)(gb + 3 * t + 4) = b;
        *(v4 *)(gb + 3 * t + 8) = c;
      }
    } else {
#pragma unroll
      for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (4 * g + e < Rows::NQ) hb[HS * t + e] = acc[t][e] * sc[e];
    }
    wave_sync();
    // the accumulators are staged: their registers take the next rows
    __builtin_amdgcn_sched_barrier(0);
    if (more) load_half(R1, s0 + 128, 1);
    __builtin_amdgcn_sched_barrier(0);
  };
  // Copy out the staged orbitals [p0, p0 + m) (one or two slabs).  STREAM
  // launches write whole slabs with a fixed sequence of streaming 16-byte
  // stores (NS slabs: compile-time); everything else goes through
  // mfma_copy_out.
  auto store_whole = [&](int p0, auto ns) __attribute__((always_inline)) {
    constexpr int NS = decltype(ns)::value;
    mfma_stream_out<128 * NS>(lane, stage, ov + p0);
    mfma_stream_out<384 * NS>(lane, stage + 256, og + 3 * p0);
    mfma_stream_out<128 * HS * NS>(lane, stage + 1024, oh + HS * p0);
    wave_sync();                                      // reads done before the next slab's writes
  };
  auto store_some = [&](int p0, int m) __attribute__((always_inline)) {
    mfma_copy_out(lane, stage, ov + p0, m);
    mfma_copy_out(lane, stage + 256, og + 3 * p0, 3 * m);
    mfma_copy_out(lane, stage + 1024, oh + HS * p0, HS * m);
    wave_sync();
  };
  // Slabs go in pairs, stored together once both are staged: a point's
  // results leave in half as many bursts, each of whole 16-byte stores.
  // Every pair followed by more slabs is whole (its end is below s_last,
  // which is at most nreal) and takes one fixed store sequence.  The first
  // such pair is peeled off the loop, so that the loop is entered, as it is
  // repeated, with a pair's stores behind the next rows: with one path and
  // one entry state the compiler's wait counts at the loop head are exact
  // and the stores stay in flight while the next pair's first rows arrive.
  auto pair_step = [&](int p0) __attribute__((always_inline)) {
    slab_step(p0, true, 0);
    slab_step(p0 + 128, true, 1);
    if constexpr (STREAM) store_whole(p0, std::integral_constant<int, 2>{});
    else store_some(p0, 256);
  };
  const int nsl = (s_last - s_first + 127) >> 7;
  int k = 0;
  if (k + 2 < nsl) {
    pair_step(s_first);
    for (k = 2; k + 2 < nsl; k += 2) pair_step(s_first + 128 * k);
  }
  const int p0 = s_first + 128 * k;
  const int m = s_last - p0;                          // valid orbitals left: 1 .. 256
  if (nsl - k == 2) {
    slab_step(p0, true, 0);
    slab_step(p0 + 128, false, 1);
    if (STREAM && m == 256) store_whole(p0, std::integral_constant<int, 2>{});
    else store_some(p0, m);
  } else {
    slab_step(p0, false, 0);
    if (STREAM && m == 128) store_whole(p0, std::integral_constant<int, 1>{});
    else store_some(p0, m);
  }
}

// ---------------------------------------------------------------------------
// FP32 tables: the same product on v_mfma_f32_16x16x4_f32, FP32 throughout.
//
// Operand layout (lane l, g = l / 16, r = l % 16): A holds W[r][g], B holds
// P[g][column r], D holds D[4g .. 4g+3][column r].  The 64 stencil rows are
// taken in 16 chunks of four, chunk kc = 4 i + j holding rows (i, j, k = 0..3),
// so lane group g carries z row k = g.  Lane (g, r) loads 16 bytes (four
// orbitals, 4r .. 4r+3 of a 64-orbital slab) of its row of each chunk, and
// tile t (0..3) takes element t: column r is orbital 4r + t.  No data moves
// between or inside lanes before the products.
constexpr int kMfma32Stage = 64 * (1 + 3 + 9);

// Four wavefronts per SIMD: about 126 VGPRs, no spills (five would spill).
template <bool PACKED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
eval3d_mfma32(const Params P) {
  using Rows = MfmaRows<PACKED>;
  typedef float f4 __attribute__((ext_vector_type(4)));
  __shared__ __align__(16) float stage_all[4 * kMfma32Stage];
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, r = lane & 15;
  float *const stage = stage_all + (threadIdx.x >> 6) * kMfma32Stage;
  int p, slab, ix, iy, iz;
  float wx[3][4], wy[3][4], wz[3][4];
  if (!mfma_point(P, p, slab, ix, iy, iz, wx, wy, wz)) return;

  // A operand of chunk kc: W[q = r][(i, j, g)] with i = kc / 4, j = kc % 4
  const int q = r;
  const int ox = (Rows::OX >> (2 * q)) & 3, oy = (Rows::OY >> (2 * q)) & 3, oz = (Rows::OZ >> (2 * q)) & 3;
#define ESK_PICK(o, w, u) by_order(o, w[0][u], w[1][u], w[2][u])
  const float x0 = ESK_PICK(ox, wx, 0), x1 = ESK_PICK(ox, wx, 1), x2 = ESK_PICK(ox, wx, 2), x3 = ESK_PICK(ox, wx, 3);
  const float y0 = ESK_PICK(oy, wy, 0), y1 = ESK_PICK(oy, wy, 1), y2 = ESK_PICK(oy, wy, 2), y3 = ESK_PICK(oy, wy, 3);
  const float z0 = ESK_PICK(oz, wz, 0), z1 = ESK_PICK(oz, wz, 1), z2 = ESK_PICK(oz, wz, 2), z3 = ESK_PICK(oz, wz, 3);
#undef ESK_PICK
  const float zg = q < Rows::NQ ? (g == 0 ? z0 : g == 1 ? z1 : g == 2 ? z2 : z3) : 0.0f;
  float a[16];
#pragma unroll
  for (int kc = 0; kc < 16; ++kc) {
    const int i = kc >> 2, j = kc & 3;
    const float xi = i == 0 ? x0 : i == 1 ? x1 : i == 2 ? x2 : x3;
    const float yj = j == 0 ? y0 : j == 1 ? y1 : j == 2 ? y2 : y3;
    a[kc] = xi * yj * zg;
  }
  float sc[4];
  mfma_scales<Rows>(P, g, 1.0f, sc);
  // staging area in output layout for a 64-orbital slab: v (64 floats),
  // gradient (192), Hessian (64 HS)
  float *const vg_base = stage + 4 * r;                          // v of orbital 4r
  float *const gr_base = stage + 64 + 12 * r;                    // gradient of orbital 4r
  float *const h_base = stage + 256 + Rows::HS * 4 * r + 4 * (g - 1);

  const float *cell = (const float *)P.coefs + (long)ix * P.xs + (long)iy * P.ys + (long)iz * P.zs;
  const int row_reals = (int)P.zs;
  const int end = min(row_reals, P.nreal);
  float *ov = (float *)P.out_v + (long)p * P.ld_v;
  float *og = (float *)P.out_g + (long)p * P.ld_g;
  float *oh = (float *)P.out_h + (long)p * P.ld_h;
  const float *lane_row = cell + (long)g * P.zs;              // z row k = g
  // rows of chunks 8h .. 8h+7 at orbitals s0 + 4r (rows are padded to
  // whole 32-byte pieces, so a lane whose first orbital is inside the row
  // reads inside it; other lanes read zeros)
  auto load_half = [&](f4 (&R)[8], int s0, int h) {
    const int o0 = s0 + 4 * r;
    const bool have = o0 < row_reals;
    const float *base = lane_row + o0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int kc = 8 * h + u, i = kc >> 2, j = kc & 3;
      R[u] = have ? *(const f4 *)(base + i * P.xs + j * P.ys) : f4{0.f, 0.f, 0.f, 0.f};
    }
  };
  auto mma_half = [&](const f4 (&R)[8], f4 (&acc)[4], int h) {
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(h == 0 ? a[u] : a[8 + u], R[u][t], acc[t], 0, 0, 0);
  };

  f4 R0[8], R1[8];
  // all slabs of the point, or only the wavefront's slab (orbital slabs outermost)
  const int s_first = slab < 0 ? 0 : slab * 64;
  const int s_last = slab < 0 ? end : min(end, s_first + 64);
  load_half(R0, s_first, 0);
  for (int s0 = s_first; s0 < s_last; s0 += 64) {
    f4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    load_half(R1, s0, 1);
    mma_half(R0, acc, 0);
    if (s0 + 64 < s_last) load_half(R0, s0 + 64, 0);
    mma_half(R1, acc, 1);
    typedef float v4 __attribute__((ext_vector_type(4)));
    if (g == 0) {
      *(v4 *)vg_base = v4{acc[0][0], acc[1][0], acc[2][0], acc[3][0]} * sc[0];
      *(v4 *)gr_base = v4{acc[0][1] * sc[1], acc[0][2] * sc[2], acc[0][3] * sc[3], acc[1][1] * sc[1]};
      *(v4 *)(gr_base + 4) = v4{acc[1][2] * sc[2], acc[1][3] * sc[3], acc[2][1] * sc[1], acc[2][2] * sc[2]};
      *(v4 *)(gr_base + 8) = v4{acc[2][3] * sc[3], acc[3][1] * sc[1], acc[3][2] * sc[2], acc[3][3] * sc[3]};
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (4 * g + e < Rows::NQ) h_base[Rows::HS * t + e] = acc[t][e] * sc[e];
    }
    wave_sync();
    const int m = min(64, P.nreal - s0);              // valid orbitals of the slab
    mfma_copy_out(lane, stage, ov + s0, m);
    mfma_copy_out(lane, stage + 64, og + 3 * s0, 3 * m);
    mfma_copy_out(lane, stage + 256, oh + Rows::HS * s0, Rows::HS * m);
    wave_sync();                                      // reads done before the next slab's writes
  }
}

}  // namespace esk
