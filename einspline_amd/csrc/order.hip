// order.hip -- locality traversal of a batch of points.
//
// With random points, each coefficient row is read by every point whose 4x4x4
// stencil covers it, at random times during the launch.  When the table is
// larger than the caches, most of those reads go to HBM.  Grouping the points
// by the Morton code of their cell makes points that share rows run close
// together in time.  The eval kernel then reads `perm[slot]` to find its point
// and maps blocks so that each XCD (with its own L2) walks one contiguous run
// of the grouped order.
//
// The grouping is a counting sort on a coarsened Morton key (about two points
// per bin, at most 2^11 bins), in two launches:
//   bin_points    key of every point and its rank in its bin (atomic count);
//   scan_scatter  every block computes the exclusive prefix sum of the counts
//                 in LDS (8 KB of L2 hits) and places its own points:
//                 perm[offset[bin] + rank] = point.
// No block waits for another and there is no device-wide fence (on a
// multi-XCD GPU a device-scope release writes back the L2).  The counts are
// double-buffered: while one call uses one buffer, its block 0 clears the
// other for the next call.
// The order inside a bin depends on atomic timing.  Results do not: each
// point's outputs are written to its own index by the same instruction
// sequence, so the traversal changes only which cache serves the reads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "runtime_internal.hpp"

namespace {

constexpr int kMaxBinBits = 11;            // 2048 bins: 8 KB of counts in LDS
constexpr int kBins = 1 << kMaxBinBits;
constexpr int kThreads = 256;

// Spread the low 10 bits of v so that bit k lands at bit 3k.
__device__ __forceinline__ unsigned spread3(unsigned v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ unsigned cell_of(double x, double x0, double xi, int n4, int shift) {
  const double u = (x - x0) * xi;
  int c = 0;
  if (u == u) c = u <= 0.0 ? 0 : (u >= (double)n4 ? n4 : (int)u);   // NaN -> cell 0
  return (unsigned)c >> shift;
}

// Inclusive scan of one value per lane across a wavefront.
__device__ __forceinline__ int wave_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

__global__ void __launch_bounds__(kThreads) zero_ints(int *__restrict__ p, int n) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <class PosT>
__global__ void __launch_bounds__(kThreads) bin_points(const PosT *__restrict__ pos, int n, esk::Params P,
                                                       int cell_shift, int bin_shift,
                                                       int *__restrict__ counts, int *__restrict__ bin,
                                                       int *__restrict__ rank) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const PosT *p = pos + 3l * i;
  const unsigned cx = cell_of((double)p[0], P.x0, P.xi, P.nx4, cell_shift);
  const unsigned cy = cell_of((double)p[1], P.y0, P.yi, P.ny4, cell_shift);
  const unsigned cz = cell_of((double)p[2], P.z0, P.zi, P.nz4, cell_shift);
  const int b = (int)(((spread3(cx) << 2) | (spread3(cy) << 1) | spread3(cz)) >> bin_shift);
  bin[i] = b;
  rank[i] = atomicAdd(&counts[b], 1);
}

// Exclusive prefix sum of counts[0, nbins) into offs[] (LDS), by every block
// of T threads: rounds of T consecutive bins (one per thread, no LDS bank
// conflicts), each scanned across the wavefront with shuffles and across the
// block's wavefronts through LDS, with a running carry between rounds.
template <int T>
__device__ __forceinline__ void block_offsets(const int *__restrict__ counts, int nbins, int *offs,
                                              int *wave_sum) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < (kBins + T - 1) / T; ++k) {
    const int b = k * T + t;
    if (b < nbins) offs[b] = counts[b];
  }
  __syncthreads();
  int carry = 0;
  for (int k = 0; k < nbins; k += T) {
    const int b = k + t;
    const int v = b < nbins ? offs[b] : 0;
    const int incl = wave_scan(v);
    if ((t & 63) == 63) wave_sum[t >> 6] = incl;
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < T / 64; ++w) {
      const int ws = wave_sum[w];
      if (w < (t >> 6)) before += ws;
      total += ws;
    }
    if (b < nbins) offs[b] = carry + before + incl - v;
    carry += total;
    __syncthreads();                                      // wave_sum is rewritten next round
  }
}

// Blocks of 1024 threads: the prefix over at most 2048 bins takes two rounds,
// and a quarter as many blocks repeat it.
constexpr int kScanThreads = 1024;

__global__ void __launch_bounds__(kScanThreads) scan_scatter(const int *__restrict__ counts,
                                                             int *__restrict__ next_counts, int nbins,
                                                             const int *__restrict__ bin,
                                                             const int *__restrict__ rank, int n,
                                                             int *__restrict__ perm) {
  __shared__ int offs[kBins];
  __shared__ int wave_sum[kScanThreads / 64];
  if (blockIdx.x == 0) {
    // the next call counts into the other buffer; clear it now
    for (int k = threadIdx.x; k < kBins; k += kScanThreads) next_counts[k] = 0;
  }
  block_offsets<kScanThreads>(counts, nbins, offs, wave_sum);
  const int i = blockIdx.x * kScanThreads + threadIdx.x;
  if (i < n) perm[offs[bin[i]] + rank[i]] = i;
}

int bits_for(int n4) {
  int b = 0;
  while ((1 << b) <= n4) ++b;
  return std::max(b, 1);
}

}  // namespace

struct EskOrderWs {
  std::mutex mu;
  int device = -1;
  int *buf = nullptr;           // counts[2][kBins], bin[cap], rank[cap], perm[cap]
  size_t cap = 0;               // points the buffers hold
  int parity = 0;               // which counts buffer the next call uses (the other is zero)
  hipEvent_t done = nullptr;    // recorded after the last launch that read the buffers
};

void esk_order_free(EskOrderWs *w) {
  if (!w) return;
  if (w->done) (void)hipEventSynchronize(w->done);
  if (w->buf) (void)hipFree(w->buf);
  if (w->done) (void)hipEventDestroy(w->done);
  delete w;
}

namespace {

// The two sorting launches on caller-provided buffers: counts (nbins ints,
// zero on entry), next_counts (cleared for a later call), bin, rank and
// perm (num ints each).
hipError_t sort_launch(const esk::Params &P, const void *pos, bool pos_f32, int num, hipStream_t s,
                       int *counts, int *next_counts, int *bin, int *rank, int *pm) {
  // key: Morton code of the cell (30 bits at most; coarser cells on huge grids),
  // then the top bin_bits of it
  const int bmax = std::max(bits_for(P.nx4), std::max(bits_for(P.ny4), bits_for(P.nz4)));
  const int cell_shift = bmax > 10 ? bmax - 10 : 0;
  const int key_bits = 3 * (bmax - cell_shift);
  int want = 0;
  while ((1 << (want + 1)) <= num) ++want;                 // about two points per bin
  const int bin_bits = std::max(0, std::min(std::min(want, key_bits), kMaxBinBits));
  const int bin_shift = key_bits - bin_bits;
  const int nbins = 1 << bin_bits;
  const int blocks = (num + kThreads - 1) / kThreads;
  if (pos_f32)
    hipLaunchKernelGGL(bin_points<float>, dim3(blocks), dim3(kThreads), 0, s, (const float *)pos, num, P,
                       cell_shift, bin_shift, counts, bin, rank);
  else
    hipLaunchKernelGGL(bin_points<double>, dim3(blocks), dim3(kThreads), 0, s, (const double *)pos, num, P,
                       cell_shift, bin_shift, counts, bin, rank);
  hipLaunchKernelGGL(scan_scatter, dim3((unsigned)((num + kScanThreads - 1) / kScanThreads)), dim3(kScanThreads), 0,
                     s, counts, next_counts, nbins, bin, rank, num, pm);
  return hipGetLastError();
}

}  // namespace

int esk_order_begin(EskOrderWs **slot, const esk::Params &P, const void *pos, bool pos_f32,
                    int num, hipStream_t s, const int **perm) {
  hipError_t e;
  EskOrderWs *w;
  {
    static std::mutex create_mu;
    std::lock_guard<std::mutex> g(create_mu);
    if (!*slot) *slot = new EskOrderWs;
    w = *slot;
  }
  w->mu.lock();                 // released by esk_order_end
  auto bail = [&](int rc) { w->mu.unlock(); return rc; };
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (w->done && w->device != dev)
    return bail(esk_fail(EINSPLINE_EINVAL, "locality order: table used from another device"));
  if (!w->done) {
    if ((e = hipEventCreateWithFlags(&w->done, hipEventDisableTiming)) != hipSuccess)
      return bail(esk_hip_fail(e, "hipEventCreate"));
    w->device = dev;
    // recorded once so that every later wait and synchronise has a completed event to see
    if ((e = hipEventRecord(w->done, s)) != hipSuccess) return bail(esk_hip_fail(e, "hipEventRecord"));
  }
  // order this stream after the previous user of the buffers
  if ((e = hipStreamWaitEvent(s, w->done, 0)) != hipSuccess) return bail(esk_hip_fail(e, "hipStreamWaitEvent"));
  constexpr size_t kHead = 2 * (size_t)kBins;            // two count buffers
  const size_t need = ((size_t)num + 63) & ~(size_t)63;
  if (need > w->cap) {
    // earlier launches on any stream may still read the old buffers
    (void)hipEventSynchronize(w->done);
    if (w->buf) (void)hipFree(w->buf);
    w->buf = nullptr;
    w->cap = 0;
    const size_t ints = kHead + 3 * need;
    if ((e = hipMalloc(&w->buf, ints * sizeof(int))) != hipSuccess)
      return bail(esk_fail(EINSPLINE_ENOMEM, "locality order: cannot allocate %zu bytes", ints * sizeof(int)));
    w->cap = need;
    hipLaunchKernelGGL(zero_ints, dim3((unsigned)((kHead + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                       w->buf, (int)kHead);
    if ((e = hipGetLastError()) != hipSuccess) return bail(esk_hip_fail(e, "zero launch"));
  }
  int *counts = w->buf + (size_t)w->parity * kBins, *next_counts = w->buf + (size_t)(1 - w->parity) * kBins;
  int *bin = w->buf + kHead, *rank = bin + w->cap, *pm = rank + w->cap;
  if ((e = sort_launch(P, pos, pos_f32, num, s, counts, next_counts, bin, rank, pm)) != hipSuccess) {
    w->cap = 0;                 // the counts may be dirty: reallocate and clear on the next call
    return bail(esk_hip_fail(e, "locality order launch"));
  }
  w->parity = 1 - w->parity;
  *perm = pm;
  return EINSPLINE_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free sorting: the same two launches on a file-scope device
// array instead of a table's workspace, for callers that cannot allocate.
// Layout: counts[kBins], spare counts[kBins], then bin, rank and perm for
// at most EINSPLINE_NOALLOC_MAX_POINTS points each (12.6 MB per GPU, held
// until the process ends).  Every call on a GPU shares that GPU's copy, so
// calls that overlap on one GPU race; each call clears the counts itself.

constexpr size_t kScratchInts = 2 * (size_t)kBins + 3 * (size_t)EINSPLINE_NOALLOC_MAX_POINTS;
static __device__ __align__(256) int scratch[kScratchInts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int *g_scratch_ptr[kMaxDevices];   // device address of `scratch` per device, once looked up
}  // namespace

// The current device's address of `scratch`, looked up once per device with
// that device current.  In host code the array's own name is a host-side
// stand-in, not a device address: only this pointer may reach kernels.
static int scratch_of_current_device(int **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return esk_hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return esk_fail(EINSPLINE_EINVAL, "device %d out of range", dev);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return esk_hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return esk_fail(EINSPLINE_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = (int *)p;
  }
  *out = g_scratch_ptr[dev];
  return EINSPLINE_OK;
}

int esk_order_begin_scratch(const esk::Params &P, const void *pos, bool pos_f32, int num, hipStream_t s,
                            const int **perm) {
  if (num < 0 || num > EINSPLINE_NOALLOC_MAX_POINTS)
    return esk_fail(EINSPLINE_EINVAL, "locality order without allocation: %d points, at most %d", num,
                    EINSPLINE_NOALLOC_MAX_POINTS);
  int *base = nullptr;
  int rc = scratch_of_current_device(&base);
  if (rc) return rc;
  int *counts = base, *spare = base + kBins;
  int *bin = base + 2 * (size_t)kBins, *rank = bin + EINSPLINE_NOALLOC_MAX_POINTS,
      *pm = rank + EINSPLINE_NOALLOC_MAX_POINTS;
  // the previous call left its counts behind: clear them with a kernel (no
  // memset node, so the call can also be captured into a graph)
  hipLaunchKernelGGL(zero_ints, dim3((unsigned)((kBins + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                     counts, kBins);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = sort_launch(P, pos, pos_f32, num, s, counts, spare, bin, rank, pm);
  if (e != hipSuccess) return esk_hip_fail(e, "locality order launch");
  *perm = pm;
  return EINSPLINE_OK;
}

void esk_order_end(EskOrderWs *w, hipStream_t s) {
  (void)hipEventRecord(w->done, s);
  w->mu.unlock();
}
