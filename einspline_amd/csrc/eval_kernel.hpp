// eval_kernel.hpp -- gfx950 kernels for batched 3-D multi-spline evaluation.
//
// One work-item ("lane") owns one (point, chunk) pair, where a chunk is
// LB contiguous bytes (V = LB / sizeof(storage) reals) of an orbital row.
// Lanes of the same point are consecutive, so for each of the 64 stencil
// rows a wavefront issues one fully coalesced load of 64*LB bytes.  The
// contraction follows src/multi_bspline_eval_std_d_impl.h:742-864 (the
// reference's std vgh loop) but is sum-factorised over the z axis: for every
// (i,j) the four contiguous k-rows are first reduced with the z basis
// (c, dc, d2c), then combined with the (i,j) products of the x/y basis.  That
// is 22 FMAs per coefficient quad for vgh instead of 40.
//
// Position -> cell follows the std evaluators: u = (x - start) * delta_inv
// in double, rounded to the accumulation type for float tables
// (src/multi_bspline_eval_std_s_impl.h:749-755), integer part by truncation,
// t = u - trunc(u).  The cell index is clamped to [0, N-4] with t adjusted,
// which is the identity for points inside [start, end] and keeps every
// access in bounds otherwise.
//
// Complex tables are handled as real tables of 2N interleaved reals (the
// basis is real); only the output permutation differs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace esk {

#ifndef ESK_WAVES_PER_EU
#define ESK_WAVES_PER_EU 4
#endif
// Contracting y before x: for each x row i the z-reduced partials of the
// four y rows are combined with the y basis into per-i partials, which the x
// basis then folds into the accumulators.  That takes fewer FMAs and no
// products of two uniform basis values, at the cost of the per-i partials'
// registers.  On MI355X it is 2-4% faster for FP32 tables where the launch
// is bound by FP32 issue (48^3 x 256, locality order) and no slower
// elsewhere, but 2-6% slower for the 16-bit tables' vector kernel, so it is
// used for 32- and 64-bit tables.  ESK_YX = 0 or 1 forces either
// contraction.
#ifndef ESK_YX
#define ESK_YX -1
#endif
#ifndef ESK_PREFETCH
#define ESK_PREFETCH 1
#endif


// MODE_QMC: value, lattice-transformed gradient and Laplacian in the layout
// of the reference's CUDA vgl (src/multi_bspline_cuda_s_impl.h:566-734).
// MODE_VGHGH: vgh plus the 27 third derivatives per orbital
// (src/multi_bspline_eval_std_d_impl.h:868-1049).
enum { MODE_V = 0, MODE_VG = 1, MODE_VGL = 2, MODE_VGH = 3, MODE_QMC = 4, MODE_VGHGH = 5 };

// Storage tags for 16-bit tables (accumulated in float).
struct bf16s { uint16_t b; };
struct f16s  { uint16_t b; };

template <class S> struct acc_of        { using type = S; };
template <>        struct acc_of<bf16s> { using type = float; };
template <>        struct acc_of<f16s>  { using type = float; };

// Whether tables of storage type S contract y before x (see ESK_YX).
template <class S> struct contract_yx {
  static constexpr bool value = ESK_YX < 0 ? sizeof(S) >= 4 : ESK_YX != 0;
};

// Kernel arguments (passed by value; ~200 bytes).
struct Params {
  const void *coefs;         // device table base
  long xs, ys, zs;           // strides in storage elements (reals)
  double x0, y0, z0;         // grid starts
  double xi, yi, zi;         // delta_inv per axis
  int nx4, ny4, nz4;         // N-4 per axis: largest valid cell index
  int chunks;                // lanes per point (row_reals / V)
  int nreal;                 // valid reals per point (N, or 2N complex)
  int npoints;
  const void *pos;           // positions, 3 per point (double or float)
  void *out_v, *out_g, *out_h;   // strided bases, or device arrays of per-point pointers
  long ld_v, ld_g, ld_h;     // per-point strides of strided outputs, in reals
  void *out_gh;              // MODE_VGHGH: third derivatives, 27 per orbital
  long ld_gh;
  int hpacked;               // vgh: 1 -> 6 unique Hessian entries per orbital, 0 -> full 9
  const void *linv;          // MODE_QMC: 3x3 row-major lattice matrix G (accumulation type)
  const void *sign;          // optional per-point factor applied to every output (or null)
  long row_stride;           // MODE_QMC: distance between gradient/Laplacian rows, in reals
  int store;                 // result store policy: STORE_STAGED_NT, STORE_STAGED, STORE_DIRECT
  const int *perm;           // optional traversal order: slot s evaluates point perm[s]
  int xcd_runs;              // 1: each XCD takes one contiguous run of blocks
  int slabs;                 // > 1: orbital slabs outermost (a wavefront per point and slab,
                             // slab-major); the number of slabs per point
};

// Result store policies.  STAGED: gradient/Hessian blocks of a wavefront go
// through LDS so every store instruction writes contiguous memory; NT:
// streaming stores where an instruction writes whole lines; DIRECT: each
// lane stores its own reals with ordinary stores.
enum { STORE_STAGED_NT = 0, STORE_STAGED = 1, STORE_DIRECT = 2 };

// ---------------------------------------------------------------------------
// helpers

template <class T> __device__ __forceinline__ T uniform(T v);
template <> __device__ __forceinline__ int uniform<int>(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}
template <> __device__ __forceinline__ float uniform<float>(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
template <> __device__ __forceinline__ double uniform<double>(double v) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ float fmaA(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaA(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float truncA(float a) { return __builtin_truncf(a); }
__device__ __forceinline__ double truncA(double a) { return __builtin_trunc(a); }

// Empty asm that claims to modify v: forces v to be computed at this point.
__device__ __forceinline__ void pin(float &v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void pin(double &v) { asm volatile("" : "+v"(v)); }

// Cell index and fraction along one axis.
template <class A>
__device__ __forceinline__ void locate(double x, double x0, double xinv, int n4,
                                       int &ix, A &t) {
  A u = (A)((x - x0) * xinv);
  A ip = truncA(u);
  A ipc = ip < (A)0 ? (A)0 : ip;          // NaN compares false -> kept, fixed below
  ipc = ipc > (A)n4 ? (A)n4 : ipc;
  if (!(ipc == ipc)) ipc = (A)0;
  t = u - ipc;
  ix = (int)ipc;
}

// Uniform cubic B-spline basis (power form, rows of the 4x4 basis matrix in
// src/bspline_data.c:143-193) and its first and second t-derivatives.
// Horner steps are explicit FMAs: left to contraction, the compiler may fuse
// a product in one kernel and not in another (a shared subexpression changes
// its choice), and every kernel must compute the same weights bit for bit.
template <class A>
__device__ __forceinline__ void basis0(A t, A w[4]) {
  const A s6 = (A)(1.0 / 6.0), h = (A)0.5, t3 = (A)(2.0 / 3.0);
  w[0] = fmaA(fmaA(fmaA(-s6, t, h), t, -h), t, s6);
  w[1] = fmaA(fmaA(h, t, (A)-1) * t, t, t3);
  w[2] = fmaA(fmaA(fmaA(-h, t, h), t, h), t, s6);
  w[3] = ((s6 * t) * t) * t;
}
template <class A>
__device__ __forceinline__ void basis1(A t, A w[4]) {
  const A h = (A)0.5, o5 = (A)1.5;
  w[0] = fmaA(fmaA(-h, t, (A)1), t, -h);
  w[1] = fmaA(o5, t, (A)-2) * t;
  w[2] = fmaA(fmaA(-o5, t, (A)1), t, h);
  w[3] = (h * t) * t;
}
template <class A>
__device__ __forceinline__ void basis2(A t, A w[4]) {
  w[0] = (A)1 - t;
  w[1] = fmaA((A)3, t, (A)-2);
  w[2] = fmaA((A)-3, t, (A)1);
  w[3] = t;
}

// Load V reals of storage type S (LB bytes) and widen to A.
template <class S, int LB> struct Loader;

template <int LB> struct Loader<float, LB> {
  static constexpr int V = LB / 4;
  __device__ __forceinline__ static void load(const float *p, float *r) {
    if constexpr (LB == 4) { r[0] = p[0]; }
    else if constexpr (LB == 8) { float2 q = *(const float2 *)p; r[0] = q.x; r[1] = q.y; }
    else {
#pragma unroll
      for (int c = 0; c < LB / 16; ++c) {
        float4 q = *((const float4 *)p + c);
        r[4 * c + 0] = q.x; r[4 * c + 1] = q.y; r[4 * c + 2] = q.z; r[4 * c + 3] = q.w;
      }
    }
  }
};
template <int LB> struct Loader<double, LB> {
  static constexpr int V = LB / 8;
  __device__ __forceinline__ static void load(const double *p, double *r) {
    if constexpr (LB == 8) { r[0] = p[0]; }
    else {
#pragma unroll
      for (int c = 0; c < LB / 16; ++c) {
        double2 q = *((const double2 *)p + c);
        r[2 * c + 0] = q.x; r[2 * c + 1] = q.y;
      }
    }
  }
};
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ float f16_lo(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
}
__device__ __forceinline__ float f16_hi(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
}
template <int LB> struct Loader<bf16s, LB> {
  static constexpr int V = LB / 2;
  __device__ __forceinline__ static void load(const bf16s *p, float *r) {
    if constexpr (LB == 4) { uint32_t w = *(const uint32_t *)p; r[0] = bf16_lo(w); r[1] = bf16_hi(w); }
    else if constexpr (LB == 8) {
      uint2 q = *(const uint2 *)p;
      r[0] = bf16_lo(q.x); r[1] = bf16_hi(q.x); r[2] = bf16_lo(q.y); r[3] = bf16_hi(q.y);
    } else {
#pragma unroll
      for (int c = 0; c < LB / 16; ++c) {
        uint4 q = *((const uint4 *)p + c);
        r[8 * c + 0] = bf16_lo(q.x); r[8 * c + 1] = bf16_hi(q.x);
        r[8 * c + 2] = bf16_lo(q.y); r[8 * c + 3] = bf16_hi(q.y);
        r[8 * c + 4] = bf16_lo(q.z); r[8 * c + 5] = bf16_hi(q.z);
        r[8 * c + 6] = bf16_lo(q.w); r[8 * c + 7] = bf16_hi(q.w);
      }
    }
  }
};
template <int LB> struct Loader<f16s, LB> {
  static constexpr int V = LB / 2;
  __device__ __forceinline__ static void load(const f16s *p, float *r) {
    if constexpr (LB == 4) { uint32_t w = *(const uint32_t *)p; r[0] = f16_lo(w); r[1] = f16_hi(w); }
    else if constexpr (LB == 8) {
      uint2 q = *(const uint2 *)p;
      r[0] = f16_lo(q.x); r[1] = f16_hi(q.x); r[2] = f16_lo(q.y); r[3] = f16_hi(q.y);
    } else {
#pragma unroll
      for (int c = 0; c < LB / 16; ++c) {
        uint4 q = *((const uint4 *)p + c);
        r[8 * c + 0] = f16_lo(q.x); r[8 * c + 1] = f16_hi(q.x);
        r[8 * c + 2] = f16_lo(q.y); r[8 * c + 3] = f16_hi(q.y);
        r[8 * c + 4] = f16_lo(q.z); r[8 * c + 5] = f16_hi(q.z);
        r[8 * c + 6] = f16_lo(q.w); r[8 * c + 7] = f16_hi(q.w);
      }
    }
  }
};

// Streaming stores: results are written once and never re-read by the
// kernel, so they bypass the caches that hold the coefficient table.
// Result stores.  `nt` selects non-temporal (streaming) stores, which keep
// the results from evicting the coefficient table from L2 and the Infinity
// Cache.  They are used only where every store instruction of a wavefront
// writes whole contiguous cache lines (a streaming store that covers part of
// a line is far slower than a cached one) and the destination is a
// wave-uniform base plus a per-lane offset: they are buffer stores with the
// non-temporal hint (st_stream).  A runtime choice between the compiler's
// nontemporal store builtin and a plain store does not work: the two stores
// are merged into one and the hint is dropped.
template <class T>
__device__ __forceinline__ void st1(T *p, T v) { *p = v; }

// Streaming store of 4, 8 or 16 bytes at byte offset `off` (< 2^31) from a
// wave-uniform base.
template <class T>
__device__ __forceinline__ void st_stream(T v, void *base, uint32_t off) {
  // raw buffer: stride 0, no range limit below 2^31 bytes, 32-bit data format
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(base, 0, 0x7fffffff, 0x00020000);
  constexpr int kNT = 2;                    // cache policy: nt
  if constexpr (sizeof(T) == 16) {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, off, 0, kNT);
  } else if constexpr (sizeof(T) == 8) {
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, off, 0, kNT);
  } else {
    static_assert(sizeof(T) == 4, "4, 8 or 16 bytes");
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, kNT);
  }
}

// Store K reals held in r[] to base + off (contiguous).  `vec` selects
// 16-byte stores (base + off must then be 16-byte aligned and K*sizeof(A) a
// multiple of 16); `nt` (base wave-uniform) streaming stores where a store
// instruction of the wavefront writes whole lines.
template <class A, int K>
__device__ __forceinline__ void store_block(A *base, int off, const A *r, bool vec, bool nt) {
  constexpr int PER16 = 16 / sizeof(A);
  A *const dst = base + off;
  if constexpr (K * sizeof(A) == 8) {
    if (vec) {
      typedef A v8 __attribute__((ext_vector_type(8 / sizeof(A))));
      v8 q;
#pragma unroll
      for (int e = 0; e < K; ++e) q[e] = r[e];
      if (nt) st_stream(q, base, (uint32_t)off * (uint32_t)sizeof(A));   // 64 lanes x 8 bytes: whole lines
      else *(v8 *)dst = q;
      return;
    }
  }
  if constexpr ((K * sizeof(A)) % 16 == 0) {
    if (vec) {
      typedef A v16 __attribute__((ext_vector_type(PER16)));
      // one piece per lane: consecutive lanes fill whole lines
      const bool full_lines = nt && (K * sizeof(A) == 16);
#pragma unroll
      for (int c = 0; c < K / PER16; ++c) {
        v16 q;
#pragma unroll
        for (int e = 0; e < PER16; ++e) q[e] = r[c * PER16 + e];
        if (full_lines) st_stream(q, base, (uint32_t)off * (uint32_t)sizeof(A));
        else ((v16 *)dst)[c] = q;
      }
      return;
    }
  }
#pragma unroll
  for (int e = 0; e < K; ++e) st1(dst + e, r[e]);
}

// Store only the first m (< K) reals.
template <class A, int K>
__device__ __forceinline__ void store_partial(A *dst, const A *r, int m) {
#pragma unroll
  for (int e = 0; e < K; ++e)
    if (e < m) st1(dst + e, r[e]);
}

// Logical block of hardware block b of nb.  Workgroups are dispatched to the
// 8 XCDs round-robin (block b runs on XCD b % 8), and each XCD has its own L2.
// With `runs`, XCD x takes the x-th contiguous run of logical blocks, so
// consecutive slots of a sorted traversal share that XCD's L2.  (Dealing
// chunks of 16-256 consecutive blocks round-robin instead, so that all XCDs
// walk one region of the sorted order together, measured the same or slower
// on MI355X.)
__device__ __forceinline__ unsigned logical_block(unsigned b, unsigned nb, int runs) {
  if (!runs) return b;
  const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u, l = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + l;
}

// Orbital slabs outermost: the point and slab of wavefront `wib` of block
// `blk` (P.slabs slabs per point, `wpb` wavefronts per block).  In the
// given order the waves run slab-major over all points.  With XCD runs
// (the locality order) XCD x = blk % 8 owns the x-th contiguous eighth of
// the sorted points and walks it slab-major, so each XCD's L2 holds one
// slab of its own region's rows at a time; the grid then has 8 x
// ceil(ceil(npoints / 8) x slabs / wpb) blocks (slab_blocks).  False for a
// wavefront with no work.
__device__ __forceinline__ bool slab_wave(const Params &P, unsigned blk, unsigned wpb, unsigned wib, int &p,
                                          int &slab) {
  unsigned pt, sl;
  if (P.xcd_runs) {
    const unsigned q = ((unsigned)P.npoints + 7u) >> 3;        // points per XCD
    const unsigned lw = (blk >> 3) * wpb + wib;                // wave within the XCD
    sl = lw / q;
    pt = (blk & 7u) * q + (lw - sl * q);
  } else {
    const unsigned w = blk * wpb + wib;
    sl = w / (unsigned)P.npoints;
    pt = w - sl * (unsigned)P.npoints;
  }
  p = uniform((int)pt);
  slab = uniform((int)sl);
  return sl < (unsigned)P.slabs && pt < (unsigned)P.npoints;
}

// Blocks of `wpb` wavefronts for `waves_per_point` x npoints wavefronts in
// the slab-major mapping of slab_wave.
inline unsigned slab_blocks(const Params &P, int wpb) {
  const long s = P.slabs > 1 ? P.slabs : 1;
  if (P.slabs > 1 && P.xcd_runs) {
    const long q = ((long)P.npoints + 7) / 8;
    return (unsigned)(8 * ((q * s + wpb - 1) / wpb));
  }
  return (unsigned)(((long)P.npoints * s + wpb - 1) / wpb);
}

// Make LDS writes of this wavefront visible to its other lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS staging area for the transposed result stores: dynamic LDS, sized by
// the launch (StageLds below) only when the store policy stages, so kernels
// that store directly keep the whole LDS for occupancy.
template <class A>
__device__ __forceinline__ A *lds_stage() {
  extern __shared__ __align__(16) unsigned char esk_stage_lds[];
  return (A *)esk_stage_lds;
}

// Store a wavefront's contiguous output region of 64*K reals, where lane l
// holds reals [l*K, (l+1)*K).  The values are transposed through the wave's
// LDS slice so that every store instruction writes 64 consecutive 16-byte
// pieces (1 KB) instead of 16 bytes at a 4*K-byte stride per lane.
template <class A, int K>
__device__ __forceinline__ void wave_store(A *base, int off, const A *r, A *lds, bool nt) {
  A *const dst = base + off;
  // largest piece (16 or 8 bytes) that divides a lane's K reals
  constexpr int PB = (K * sizeof(A)) % 16 == 0 ? 16 : 8;
  constexpr int PER = PB / (int)sizeof(A);
  static_assert((K * sizeof(A)) % PB == 0, "whole pieces per lane");
  typedef A vp __attribute__((ext_vector_type(PER)));
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < K / PER; ++c) {
    vp q;
#pragma unroll
    for (int e = 0; e < PER; ++e) q[e] = r[c * PER + e];
    *(vp *)(lds + lane * K + c * PER) = q;
  }
  wave_sync();
#pragma unroll
  for (int c = 0; c < K / PER; ++c) {
    const int piece = c * 64 + lane;
    const vp q = *(const vp *)(lds + piece * PER);
    if (nt) st_stream(q, base, (uint32_t)(off + piece * PER) * (uint32_t)sizeof(A));
    else ((vp *)dst)[piece] = q;
  }
  wave_sync();
}

// ---------------------------------------------------------------------------
// the kernel
//
// S       storage type (float, double, bf16s, f16s)
// LB      bytes of one row that a lane loads per stencil row (4..32)
// MODE    MODE_V / MODE_VG / MODE_VGL / MODE_VGH
// CPLX    complex orbitals (interleaved re/im reals)
// UNIF    every wavefront serves exactly one point (chunks % 64 == 0)
// PosT    position element type (double or float)
// PTRS    outputs are device arrays of per-point pointers (CUDA-style API)

// Occupancy target from an estimate of the live registers of one
// instantiation: accumulators, in-flight rows, z-reduced partials, basis
// weights (only when they cannot live in scalar registers) and addressing.
template <class S, int LB, int MODE, bool UNIF>
struct Occupancy {
  using A = typename acc_of<S>::type;
  static constexpr int V = LB / (int)sizeof(S);
  static constexpr int W = (int)sizeof(A) / 4;              // VGPRs per value
  static constexpr int NACC = MODE == MODE_V ? 1 : MODE == MODE_VG ? 4 : MODE == MODE_VGL ? 7
                           : MODE == MODE_VGHGH ? 20 : 10;
  static constexpr int NW = MODE == MODE_V ? 12 : MODE == MODE_VG ? 24 : 36;
  // double weights overflow the scalar register file, so half of them are
  // counted as vector registers even for one point per wavefront
  // per-x-row y partials of the y-then-x contraction
  static constexpr int NT = !contract_yx<S>::value ? 0 : MODE == MODE_V ? 1 : MODE == MODE_VG ? 3
                          : MODE == MODE_VGL ? 5 : MODE == MODE_VGHGH ? 10 : 6;
  static constexpr int regs = NACC * V * W + (ESK_PREFETCH + 1) * 4 * V * W + 3 * V * W + NT * V * W +
                              (UNIF ? (W == 2 ? NW : 0) : NW * W) + 24;
  static constexpr int waves = regs <= 116 ? ESK_WAVES_PER_EU : regs <= 160 ? 3 : regs <= 248 ? 2 : 1;
};

// Staged stores: used by one-point-per-wavefront instantiations with
// gradients; each wavefront needs kmax*64 reals of LDS (kmax: the widest
// staged block of one lane, 3V gradients or 9V Hessian entries).
template <class S, int LB, int MODE, bool UNIF>
struct StageLds {
  using A = typename acc_of<S>::type;
  static constexpr int V = LB / (int)sizeof(S);
  static constexpr bool on = UNIF && MODE >= MODE_VG && MODE != MODE_QMC && ((3 * V * sizeof(A)) % 8 == 0);
  static constexpr int kmax = (MODE == MODE_VGH || MODE == MODE_VGHGH ? 9 : 3) * V;
  static constexpr size_t wave_bytes = on ? (size_t)kmax * 64 * sizeof(A) : 0;
  // dynamic LDS of a launch of `block` work-items with store policy `store`
  static size_t bytes(int block, int store) {
    return store != STORE_DIRECT ? (size_t)(block / 64) * wave_bytes : 0;
  }
};

// Evaluate chunk `ch` of point `p` (0 <= p < P.npoints, 0 <= ch < P.chunks).
// With UNIF, p must be the same for the whole wavefront.  PF: stencil steps
// whose rows are requested ahead of the one being reduced (15 requests all
// 64 rows at once: the shortest latency for one point, where throughput
// does not matter); the results do not depend on it.
template <class S, int LB, int MODE, bool CPLX, bool UNIF, class PosT, bool PTRS, int PF = ESK_PREFETCH>
__device__ __forceinline__ void eval_chunk(const Params &P, int p, int ch) {
  using A = typename acc_of<S>::type;
  using L = Loader<S, LB>;
  constexpr int V = L::V;
  static_assert(!CPLX || (V % 2 == 0), "complex lanes hold whole orbitals");

  const PosT *pp = (const PosT *)P.pos + 3l * p;
  const double x = (double)pp[0], y = (double)pp[1], z = (double)pp[2];
  int ix, iy, iz;
  A tx, ty, tz;
  locate<A>(x, P.x0, P.xi, P.nx4, ix, tx);
  locate<A>(y, P.y0, P.yi, P.ny4, iy, ty);
  locate<A>(z, P.z0, P.zi, P.nz4, iz, tz);
  if (UNIF) {
    ix = uniform(ix); iy = uniform(iy); iz = uniform(iz);
    tx = uniform(tx); ty = uniform(ty); tz = uniform(tz);
  }

  A a[4], b[4], c[4], da[4], db[4], dc[4], d2a[4], d2b[4], d2c[4];
  basis0(tx, a); basis0(ty, b); basis0(tz, c);
  if constexpr (MODE >= MODE_VG) { basis1(tx, da); basis1(ty, db); basis1(tz, dc); }
  if constexpr (MODE >= MODE_VGL) { basis2(tx, d2a); basis2(ty, d2b); basis2(tz, d2c); }
  if (UNIF) {
    // one point per wavefront: keep the basis in scalar registers
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = uniform(a[q]); b[q] = uniform(b[q]); c[q] = uniform(c[q]);
      if constexpr (MODE >= MODE_VG) { da[q] = uniform(da[q]); db[q] = uniform(db[q]); dc[q] = uniform(dc[q]); }
      if constexpr (MODE >= MODE_VGL) { d2a[q] = uniform(d2a[q]); d2b[q] = uniform(d2b[q]); d2c[q] = uniform(d2c[q]); }
    }
  }

  // accumulators
  constexpr bool HESS = MODE == MODE_VGH || MODE == MODE_QMC || MODE == MODE_VGHGH;
  constexpr bool THIRD = MODE == MODE_VGHGH;
  constexpr int NACC = MODE == MODE_V ? 1 : MODE == MODE_VG ? 4 : MODE == MODE_VGL ? 7 : THIRD ? 20 : 10;
  A acc[NACC][V];
#pragma unroll
  for (int q = 0; q < NACC; ++q)
#pragma unroll
    for (int e = 0; e < V; ++e) acc[q][e] = (A)0;

  const S *base = (const S *)P.coefs + (long)ix * P.xs + (long)iy * P.ys + (long)iz * P.zs +
                  (long)ch * V;
  const long zs = P.zs;
  auto load_row = [&](int nx, int k, A *dst) {
    L::load(base + (nx >> 2) * P.xs + (nx & 3) * P.ys + k * zs, dst);
  };

  // 16 (i,j) steps, fully unrolled so the basis indices are constants.  The
  // rows of step st+PF are requested while step st is reduced, and a
  // scheduling barrier after each step stops the compiler from hoisting all
  // 64 row loads (which would need ~256 VGPRs and kill occupancy).
  static_assert(PF >= 1 && PF <= 15, "prefetch depth");
  A rb[PF + 1][4][V];
#pragma unroll
  for (int st = 0; st < PF; ++st) {
#pragma unroll
    for (int k = 0; k < 4; ++k) load_row(st, k, rb[st][k]);
  }
  constexpr bool YX = contract_yx<S>::value;
  // y partials per x row: b s0, db s0, d2b s0, b s1, db s1, b s2 (+ d3b s0,
  // d2b s1, db s2, b s3 for third derivatives); unused without YX
  constexpr int NT = MODE == MODE_V ? 1 : MODE == MODE_VG ? 3 : MODE == MODE_VGL ? 5 : THIRD ? 10 : 6;
  A tq[NT][V];
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    const int i = st >> 2, j = st & 3;
    if (st + PF < 16) {
      const int nx = st + PF;
#pragma unroll
      for (int k = 0; k < 4; ++k) load_row(nx, k, rb[nx % (PF + 1)][k]);
    }
    A (&r)[4][V] = rb[st % (PF + 1)];

    A s0[V], s1[V], s2[V], s3[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
      if constexpr (THIRD) {
        // third-derivative z weights are the constants -1, 3, -3, 1
        s3[e] = fmaA((A)3, r[1][e] - r[2][e], r[3][e] - r[0][e]);
      }
      s0[e] = c[0] * r[0][e];
      s0[e] = fmaA(c[1], r[1][e], s0[e]);
      s0[e] = fmaA(c[2], r[2][e], s0[e]);
      s0[e] = fmaA(c[3], r[3][e], s0[e]);
      if constexpr (MODE >= MODE_VG) {
        s1[e] = dc[0] * r[0][e];
        s1[e] = fmaA(dc[1], r[1][e], s1[e]);
        s1[e] = fmaA(dc[2], r[2][e], s1[e]);
        s1[e] = fmaA(dc[3], r[3][e], s1[e]);
      }
      if constexpr (MODE >= MODE_VGL) {
        s2[e] = d2c[0] * r[0][e];
        s2[e] = fmaA(d2c[1], r[1][e], s2[e]);
        s2[e] = fmaA(d2c[2], r[2][e], s2[e]);
        s2[e] = fmaA(d2c[3], r[3][e], s2[e]);
      }
    }
    if constexpr (YX) {
      {
      // y step: fold row j into the per-i partials (basis values straight from
      // scalar registers); the first row of each i initialises them
      constexpr A D3[4] = {(A)-1, (A)3, (A)-3, (A)1};
      A w[NT];
      int q = 0;
      w[q++] = b[j];
      if constexpr (MODE >= MODE_VG) { w[q++] = db[j]; }
      if constexpr (MODE == MODE_VGL) { w[q++] = d2b[j]; }
      if constexpr (HESS) { w[q++] = d2b[j]; }
      if constexpr (THIRD) { w[q++] = D3[j]; }
      // source partial of each y partial: 0 -> s0, 1 -> s1, 2 -> s2, 3 -> s3
      constexpr int NS0 = MODE == MODE_V ? 1 : MODE == MODE_VG ? 2 : MODE == MODE_VGL ? 3 : THIRD ? 4 : 3;
#pragma unroll
      for (int e = 0; e < V; ++e) {
#pragma unroll
        for (int u = 0; u < NS0; ++u) {           // b s0, db s0, d2b s0, d3b s0
          if (j == 0) tq[u][e] = w[u] * s0[e];
          else tq[u][e] = fmaA(w[u], s0[e], tq[u][e]);
        }
        if constexpr (MODE >= MODE_VG) {
          // b s1, (db s1), (d2b s1)
          constexpr int NS1 = MODE == MODE_VG ? 1 : MODE == MODE_VGL ? 1 : THIRD ? 3 : 2;
#pragma unroll
          for (int u = 0; u < NS1; ++u) {
            const A wu = u == 0 ? b[j] : u == 1 ? db[j] : d2b[j];
            if (j == 0) tq[NS0 + u][e] = wu * s1[e];
            else tq[NS0 + u][e] = fmaA(wu, s1[e], tq[NS0 + u][e]);
          }
          if constexpr (MODE >= MODE_VGL) {
            // b s2, (db s2)
            constexpr int NS2 = THIRD ? 2 : 1;
#pragma unroll
            for (int u = 0; u < NS2; ++u) {
              const A wu = u == 0 ? b[j] : db[j];
              if (j == 0) tq[NS0 + NS1 + u][e] = wu * s2[e];
              else tq[NS0 + NS1 + u][e] = fmaA(wu, s2[e], tq[NS0 + NS1 + u][e]);
            }
            if constexpr (THIRD) {
              if (j == 0) tq[NS0 + NS1 + NS2][e] = b[j] * s3[e];
              else tq[NS0 + NS1 + NS2][e] = fmaA(b[j], s3[e], tq[NS0 + NS1 + NS2][e]);
            }
          }
        }
      }
    }
    if (j == 3) {
      // x step: fold the per-i partials into the accumulators.  Partial
      // index: V/VG/VGL/VGH: 0 b s0, 1 db s0, 2 d2b s0 (VGL, VGH), then s1
      // terms, then s2 terms; VGHGH: 0 b s0, 1 db s0, 2 d2b s0, 3 d3b s0,
      // 4 b s1, 5 db s1, 6 d2b s1, 7 b s2, 8 db s2, 9 b s3.
      constexpr A D3[4] = {(A)-1, (A)3, (A)-3, (A)1};
      constexpr int BS0 = 0, DBS0 = 1, D2BS0 = 2;
      constexpr int BS1 = MODE == MODE_VG ? 2 : THIRD ? 4 : 3;
      constexpr int DBS1 = THIRD ? 5 : 4;
      constexpr int BS2 = MODE == MODE_VGL ? 4 : THIRD ? 7 : 5;
#pragma unroll
      for (int e = 0; e < V; ++e) {
        acc[0][e] = fmaA(a[i], tq[BS0][e], acc[0][e]);                    // v
        if constexpr (MODE >= MODE_VG) {
          acc[1][e] = fmaA(da[i], tq[BS0][e], acc[1][e]);                 // d/dx
          acc[2][e] = fmaA(a[i], tq[DBS0][e], acc[2][e]);                 // d/dy
          acc[3][e] = fmaA(a[i], tq[BS1][e], acc[3][e]);                  // d/dz
        }
        if constexpr (MODE == MODE_VGL) {
          acc[4][e] = fmaA(d2a[i], tq[BS0][e], acc[4][e]);                // d2/dx2
          acc[5][e] = fmaA(a[i], tq[D2BS0][e], acc[5][e]);                // d2/dy2
          acc[6][e] = fmaA(a[i], tq[BS2][e], acc[6][e]);                  // d2/dz2
        }
        if constexpr (HESS) {
          acc[4][e] = fmaA(d2a[i], tq[BS0][e], acc[4][e]);                // xx
          acc[5][e] = fmaA(da[i], tq[DBS0][e], acc[5][e]);                // xy
          acc[6][e] = fmaA(da[i], tq[BS1][e], acc[6][e]);                 // xz
          acc[7][e] = fmaA(a[i], tq[D2BS0][e], acc[7][e]);                // yy
          acc[8][e] = fmaA(a[i], tq[DBS1][e], acc[8][e]);                 // yz
          acc[9][e] = fmaA(a[i], tq[BS2][e], acc[9][e]);                  // zz
        }
        if constexpr (THIRD) {
          acc[10][e] = fmaA(D3[i], tq[BS0][e], acc[10][e]);               // xxx
          acc[11][e] = fmaA(d2a[i], tq[DBS0][e], acc[11][e]);             // xxy
          acc[12][e] = fmaA(d2a[i], tq[BS1][e], acc[12][e]);              // xxz
          acc[13][e] = fmaA(da[i], tq[D2BS0][e], acc[13][e]);             // xyy
          acc[14][e] = fmaA(da[i], tq[DBS1][e], acc[14][e]);              // xyz
          acc[15][e] = fmaA(da[i], tq[BS2][e], acc[15][e]);               // xzz
          acc[16][e] = fmaA(a[i], tq[3][e], acc[16][e]);                  // yyy
          acc[17][e] = fmaA(a[i], tq[6][e], acc[17][e]);                  // yyz
          acc[18][e] = fmaA(a[i], tq[8][e], acc[18][e]);                  // yzz
          acc[19][e] = fmaA(a[i], tq[9][e], acc[19][e]);                  // zzz
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NT; ++q)
#pragma unroll
      for (int e = 0; e < V; ++e) pin(tq[q][e]);
    } else {
    const A ab = a[i] * b[j];
    if constexpr (MODE == MODE_V) {
#pragma unroll
      for (int e = 0; e < V; ++e) acc[0][e] = fmaA(ab, s0[e], acc[0][e]);
    } else {
      const A dab = da[i] * b[j], adb = a[i] * db[j];
#pragma unroll
      for (int e = 0; e < V; ++e) {
        acc[0][e] = fmaA(ab, s0[e], acc[0][e]);    // v
        acc[1][e] = fmaA(dab, s0[e], acc[1][e]);   // d/dx
        acc[2][e] = fmaA(adb, s0[e], acc[2][e]);   // d/dy
        acc[3][e] = fmaA(ab, s1[e], acc[3][e]);    // d/dz
      }
      if constexpr (MODE == MODE_VGL) {
        const A d2ab = d2a[i] * b[j], ad2b = a[i] * d2b[j];
#pragma unroll
        for (int e = 0; e < V; ++e) {
          acc[4][e] = fmaA(d2ab, s0[e], acc[4][e]);  // d2/dx2
          acc[5][e] = fmaA(ad2b, s0[e], acc[5][e]);  // d2/dy2
          acc[6][e] = fmaA(ab, s2[e], acc[6][e]);    // d2/dz2
        }
      }
      if constexpr (HESS) {
        const A d2ab = d2a[i] * b[j], dadb = da[i] * db[j], ad2b = a[i] * d2b[j];
#pragma unroll
        for (int e = 0; e < V; ++e) {
          acc[4][e] = fmaA(d2ab, s0[e], acc[4][e]);  // xx
          acc[5][e] = fmaA(dadb, s0[e], acc[5][e]);  // xy
          acc[6][e] = fmaA(dab, s1[e], acc[6][e]);   // xz
          acc[7][e] = fmaA(ad2b, s0[e], acc[7][e]);  // yy
          acc[8][e] = fmaA(adb, s1[e], acc[8][e]);   // yz
          acc[9][e] = fmaA(ab, s2[e], acc[9][e]);    // zz
        }
      }
      if constexpr (THIRD) {
        // d3a[i], d3b[j] are the constants -1, 3, -3, 1
        constexpr A D3[4] = {(A)-1, (A)3, (A)-3, (A)1};
        const A d3ab = D3[i] * b[j], d2adb = d2a[i] * db[j], d2ab = d2a[i] * b[j];
        const A dad2b = da[i] * d2b[j], dadb = da[i] * db[j], dab = da[i] * b[j];
        const A ad3b = a[i] * D3[j], ad2b = a[i] * d2b[j], adb = a[i] * db[j];
#pragma unroll
        for (int e = 0; e < V; ++e) {
          acc[10][e] = fmaA(d3ab, s0[e], acc[10][e]);   // xxx
          acc[11][e] = fmaA(d2adb, s0[e], acc[11][e]);  // xxy
          acc[12][e] = fmaA(d2ab, s1[e], acc[12][e]);   // xxz
          acc[13][e] = fmaA(dad2b, s0[e], acc[13][e]);  // xyy
          acc[14][e] = fmaA(dadb, s1[e], acc[14][e]);   // xyz
          acc[15][e] = fmaA(dab, s2[e], acc[15][e]);    // xzz
          acc[16][e] = fmaA(ad3b, s0[e], acc[16][e]);   // yyy
          acc[17][e] = fmaA(ad2b, s1[e], acc[17][e]);   // yyz
          acc[18][e] = fmaA(adb, s2[e], acc[18][e]);    // yzz
          acc[19][e] = fmaA(ab, s3[e], acc[19][e]);     // zzz
        }
      }
    }
    }
    // Materialise the accumulators here: without this the optimiser sinks
    // the reductions past later loads and keeps every loaded row alive.
#pragma unroll
    for (int q = 0; q < NACC; ++q)
#pragma unroll
      for (int e = 0; e < V; ++e) pin(acc[q][e]);
    __builtin_amdgcn_sched_barrier(0);
  }

  // scale derivatives by delta_inv (rounded to A first, as the reference does)
  const A dxi = (A)P.xi, dyi = (A)P.yi, dzi = (A)P.zi;
  if constexpr (MODE >= MODE_VG) {
#pragma unroll
    for (int e = 0; e < V; ++e) {
      acc[1][e] *= dxi; acc[2][e] *= dyi; acc[3][e] *= dzi;
    }
  }
  if constexpr (MODE == MODE_VGL) {
    const A sxx = dxi * dxi, syy = dyi * dyi, szz = dzi * dzi;
#pragma unroll
    for (int e = 0; e < V; ++e) {
      A lx = acc[4][e] * sxx, ly = acc[5][e] * syy, lz = acc[6][e] * szz;
      acc[4][e] = lx + ly + lz;
    }
  }
  if constexpr (HESS) {
    const A sxx = dxi * dxi, sxy = dxi * dyi, sxz = dxi * dzi;
    const A syy = dyi * dyi, syz = dyi * dzi, szz = dzi * dzi;
#pragma unroll
    for (int e = 0; e < V; ++e) {
      acc[4][e] *= sxx; acc[5][e] *= sxy; acc[6][e] *= sxz;
      acc[7][e] *= syy; acc[8][e] *= syz; acc[9][e] *= szz;
    }
  }
  if constexpr (THIRD) {
    const A f[10] = {dxi * dxi * dxi, dxi * dxi * dyi, dxi * dxi * dzi, dxi * dyi * dyi,
                     dxi * dyi * dzi, dxi * dzi * dzi, dyi * dyi * dyi, dyi * dyi * dzi,
                     dyi * dzi * dzi, dzi * dzi * dzi};
#pragma unroll
    for (int q = 0; q < 10; ++q)
#pragma unroll
      for (int e = 0; e < V; ++e) acc[10 + q][e] *= f[q];
  }

  // ---- outputs ----
  if constexpr (MODE == MODE_QMC) {
    // value, G*grad and tr((G^T G) H) as rows of the caller's grad_lapl block
    const A *L = (const A *)P.linv;
    A G[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) G[q] = L[q];
    A GG[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c2 = 0; c2 < 3; ++c2)
        GG[3 * r + c2] = G[0 * 3 + r] * G[0 * 3 + c2] + G[1 * 3 + r] * G[1 * 3 + c2] +
                         G[2 * 3 + r] * G[2 * 3 + c2];
    const A sg = P.sign ? ((const A *)P.sign)[p] : (A)1;
    A *ov = ((A *const *)P.out_v)[p];
    A *ogl = ((A *const *)P.out_g)[p];
    const int e0 = ch * V;
    const long rs = P.row_stride;
#pragma unroll
    for (int e = 0; e < V; ++e) {
      if (e0 + e >= P.nreal) break;
      const A g0 = acc[1][e], g1 = acc[2][e], g2 = acc[3][e];
      const A h00 = acc[4][e], h01 = acc[5][e], h02 = acc[6][e];
      const A h11 = acc[7][e], h12 = acc[8][e], h22 = acc[9][e];
      const A lap = GG[0] * h00 + GG[3] * h01 + GG[6] * h02 + GG[1] * h01 + GG[4] * h11 +
                    GG[7] * h12 + GG[2] * h02 + GG[5] * h12 + GG[8] * h22;
      st1(ov + e0 + e, sg * acc[0][e]);
      st1(ogl + e0 + e + 0 * rs, sg * (G[0] * g0 + G[1] * g1 + G[2] * g2));
      st1(ogl + e0 + e + 1 * rs, sg * (G[3] * g0 + G[4] * g1 + G[5] * g2));
      st1(ogl + e0 + e + 2 * rs, sg * (G[6] * g0 + G[7] * g1 + G[8] * g2));
      st1(ogl + e0 + e + 3 * rs, sg * lap);
    }
    return;
  }
  A *ov, *og = nullptr, *oh = nullptr;
  if constexpr (PTRS) {
    ov = ((A *const *)P.out_v)[p];
    if constexpr (MODE >= MODE_VG) og = ((A *const *)P.out_g)[p];
    if constexpr (MODE >= MODE_VGL) oh = ((A *const *)P.out_h)[p];
  } else {
    ov = (A *)P.out_v + (long)p * P.ld_v;
    if constexpr (MODE >= MODE_VG) og = (A *)P.out_g + (long)p * P.ld_g;
    if constexpr (MODE >= MODE_VGL) oh = (A *)P.out_h + (long)p * P.ld_h;
  }
  uintptr_t align = (uintptr_t)ov;
  if constexpr (MODE >= MODE_VG) align |= (uintptr_t)og;
  if constexpr (MODE >= MODE_VGL) align |= (uintptr_t)oh;
  const bool vec = (align & 15) == 0;

  const int e0 = ch * V;                 // first real of this lane
  const int m = P.nreal - e0;            // valid reals in this lane (may exceed V)
  const bool full = m >= V;
  // streaming stores need the wave-uniform output base of one point per wave
  const bool nt = UNIF && P.store == STORE_STAGED_NT;

  // One point per wavefront, 64 full chunks and aligned outputs: the
  // wavefront's gradient and Hessian blocks are each one contiguous region,
  // written through LDS with fully coalesced stores.
  constexpr bool STAGE = StageLds<S, LB, MODE, UNIF>::on;
  const int lane = threadIdx.x & 63;
  const int wave_e0 = (ch - lane) * V;   // first real of the wavefront (UNIF)
  bool staged = false;
  A *lds = nullptr;
  if constexpr (STAGE) {
    staged = vec && P.store != STORE_DIRECT && (wave_e0 + 64 * V <= P.nreal);
    constexpr int KMAX = StageLds<S, LB, MODE, UNIF>::kmax;
    lds = lds_stage<A>() + (threadIdx.x >> 6) * (KMAX * 64);
  }

  // values (and the Laplacian) keep the real index layout
  {
    A r[V];
    const A sg = (MODE == MODE_V && P.sign) ? ((const A *)P.sign)[p] : (A)1;
#pragma unroll
    for (int e = 0; e < V; ++e) r[e] = sg * acc[0][e];
    if (full) store_block<A, V>(ov, e0, r, vec, nt); else store_partial<A, V>(ov + e0, r, m);
  }
  if constexpr (MODE == MODE_VGL) {
    A r[V];
#pragma unroll
    for (int e = 0; e < V; ++e) r[e] = acc[4][e];
    if (full) store_block<A, V>(oh, e0, r, vec, nt); else store_partial<A, V>(oh + e0, r, m);
  }
  // Complex orbitals: the CPU layout interleaves re/im per entry.  Through
  // pointer arrays the reference's CUDA kernels treat a complex table as 2N
  // reals, so entries are [q][re,im][3] and [q][re,im][6]
  // (src/multi_bspline_cuda_z_impl.h:205-228, src/multi_bspline_cuda_c_impl.h:389-412).
  constexpr bool CPU_CPLX = CPLX && !PTRS;
  if constexpr (MODE >= MODE_VG) {
    // gradient: real orbital e -> [e][3]; complex orbital q -> [q][3][re,im]
    // (the CPU layout), except through pointer arrays (CPU_CPLX false)
    A r[3 * V];
#pragma unroll
    for (int e = 0; e < V; ++e)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int o = CPU_CPLX ? (e >> 1) * 6 + d * 2 + (e & 1) : e * 3 + d;
        r[o] = acc[1 + d][e];
      }
    if constexpr (STAGE) {
      if (staged) wave_store<A, 3 * V>(og, 3 * wave_e0, r, lds, nt);
    }
    if (!staged) {
      if (full) store_block<A, 3 * V>(og, 3 * e0, r, vec, nt);
      else store_partial<A, 3 * V>(og + 3 * e0, r, 3 * m);
    }
  }
  if constexpr (MODE == MODE_VGH || THIRD) {
    // acc index of each Hessian entry, row-major 3x3 and packed upper triangle
    constexpr int full9[9] = {4, 5, 6, 5, 7, 8, 6, 8, 9};
    constexpr int pack6[6] = {4, 5, 6, 7, 8, 9};
    if (P.hpacked) {
      A r[6 * V];
#pragma unroll
      for (int e = 0; e < V; ++e)
#pragma unroll
        for (int h = 0; h < 6; ++h) {
          const int o = CPU_CPLX ? (e >> 1) * 12 + h * 2 + (e & 1) : e * 6 + h;
          r[o] = acc[pack6[h]][e];
        }
      if constexpr (STAGE) {
        if (staged) wave_store<A, 6 * V>(oh, 6 * wave_e0, r, lds, nt);
      }
      if (!staged) {
        if (full) store_block<A, 6 * V>(oh, 6 * e0, r, vec, nt);
        else store_partial<A, 6 * V>(oh + 6 * e0, r, 6 * m);
      }
    } else {
      A r[9 * V];
#pragma unroll
      for (int e = 0; e < V; ++e)
#pragma unroll
        for (int h = 0; h < 9; ++h) {
          const int o = CPU_CPLX ? (e >> 1) * 18 + h * 2 + (e & 1) : e * 9 + h;
          r[o] = acc[full9[h]][e];
        }
      if constexpr (STAGE) {
        if (staged) wave_store<A, 9 * V>(oh, 9 * wave_e0, r, lds, nt);
      }
      if (!staged) {
        if (full) store_block<A, 9 * V>(oh, 9 * e0, r, vec, nt);
        else store_partial<A, 9 * V>(oh + 9 * e0, r, 9 * m);
      }
    }
  }
  if constexpr (THIRD) {
    // full symmetric 3x3x3 tensor per orbital: slot (a,b,c) holds the entry of
    // the sorted triple (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz -> acc 10..19)
    constexpr int sym27[27] = {10, 11, 12, 11, 13, 14, 12, 14, 15,
                               11, 13, 14, 13, 16, 17, 14, 17, 18,
                               12, 14, 15, 14, 17, 18, 15, 18, 19};
    A *ogh = (A *)P.out_gh + (long)p * P.ld_gh;
    A r[27 * V];
#pragma unroll
    for (int e = 0; e < V; ++e)
#pragma unroll
      for (int h = 0; h < 27; ++h) {
        const int o = CPLX ? (e >> 1) * 54 + h * 2 + (e & 1) : e * 27 + h;
        r[o] = acc[sym27[h]][e];
      }
    const bool gvec = vec && ((uintptr_t)ogh & 15) == 0;
    if (full) store_block<A, 27 * V>(ogh, 27 * e0, r, gvec, false);
    else store_partial<A, 27 * V>(ogh + 27 * e0, r, 27 * m);
  }
}

template <class S, int LB, int MODE, bool CPLX, bool UNIF, class PosT, bool PTRS>
__global__ void __launch_bounds__(256, (Occupancy<S, LB, MODE, UNIF>::waves)) eval3d(const Params P) {
  // npoints * chunks < 2^31 is checked on the host, so 32-bit math suffices
  const unsigned b = logical_block(blockIdx.x, gridDim.x, P.xcd_runs);
  const unsigned gid = b * blockDim.x + threadIdx.x;
  int p, ch;
  if (UNIF && P.slabs > 1) {
    // orbital slabs outermost (the same 64-lane piece of the point's chunks
    // per slab): slab_wave() maps the wavefront to its point and slab
    int slab;
    if (!slab_wave(P, blockIdx.x, blockDim.x >> 6, threadIdx.x >> 6, p, slab)) return;
    ch = slab * 64 + (int)(threadIdx.x & 63u);
  } else {
    p = (int)(gid / (unsigned)P.chunks);
    if (UNIF) p = uniform(p);
    ch = (int)(gid - (unsigned)p * (unsigned)P.chunks);
    if (p >= P.npoints) return;
  }
  if (P.perm) {
    p = P.perm[p];
    if (UNIF) p = uniform(p);
  }
  eval_chunk<S, LB, MODE, CPLX, UNIF, PosT, PTRS>(P, p, ch);
}

}  // namespace esk
