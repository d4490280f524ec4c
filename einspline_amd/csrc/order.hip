// order.hip -- locality traversal of a batch of points.
//
// With random points, each coefficient row is read by every point whose 4x4x4
// stencil covers it, at random times during the launch.  When the table is
// larger than the caches, most of those reads go to HBM.  Sorting the points
// by the Morton code of their cell makes points that share rows run close
// together in time.  The eval kernel then reads `perm[slot]` to find its point
// and maps blocks so that each XCD (with its own L2) walks one contiguous run
// of the sorted order.
//
// Results do not depend on the order: each point's outputs are written to its
// own index by the same instruction sequence, so the traversal changes only
// which cache serves the reads.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>

#include "runtime_internal.hpp"

namespace {

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

template <class PosT>
__global__ void __launch_bounds__(256) cell_keys(const PosT *__restrict__ pos, int n, esk::Params P,
                                                 int shift, unsigned *__restrict__ keys,
                                                 int *__restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PosT *p = pos + 3l * i;
  const unsigned cx = cell_of((double)p[0], P.x0, P.xi, P.nx4, shift);
  const unsigned cy = cell_of((double)p[1], P.y0, P.yi, P.ny4, shift);
  const unsigned cz = cell_of((double)p[2], P.z0, P.zi, P.nz4, shift);
  keys[i] = (spread3(cx) << 2) | (spread3(cy) << 1) | spread3(cz);
  idx[i] = i;
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
  char *buf = nullptr;          // keys[2][cap], idx[2][cap], sort scratch
  size_t cap = 0;               // points the buffers hold
  size_t temp_bytes = 0;        // sort scratch bytes
  hipEvent_t done = nullptr;    // recorded after the last launch that read the buffers
};

void esk_order_free(EskOrderWs *w) {
  if (!w) return;
  if (w->done) (void)hipEventSynchronize(w->done);
  if (w->buf) (void)hipFree(w->buf);
  if (w->done) (void)hipEventDestroy(w->done);
  delete w;
}

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
  const int bx = bits_for(P.nx4), by = bits_for(P.ny4), bz = bits_for(P.nz4);
  const int bmax = std::max(bx, std::max(by, bz));
  const int shift = bmax > 10 ? bmax - 10 : 0;   // 30-bit keys: coarser cells on huge grids
  const int end_bit = 3 * (bmax - shift);
  size_t temp = 0;
  e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned *)nullptr, (unsigned *)nullptr,
                                         (const int *)nullptr, (int *)nullptr, num, 0, end_bit, s);
  if (e != hipSuccess) { w->mu.unlock(); return esk_hip_fail(e, "sort size query"); }
  temp = (temp + 255) & ~(size_t)255;
  const size_t need_pts = ((size_t)num + 63) & ~(size_t)63;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (w->done && w->device != dev) {
    w->mu.unlock();
    return esk_fail(EINSPLINE_EINVAL, "locality order: table used from another device");
  }
  if (!w->done) {
    if ((e = hipEventCreateWithFlags(&w->done, hipEventDisableTiming)) != hipSuccess) {
      w->mu.unlock();
      return esk_hip_fail(e, "hipEventCreate");
    }
    w->device = dev;
    // recorded once so that every later wait and synchronise has a completed event to see
    if ((e = hipEventRecord(w->done, s)) != hipSuccess) {
      w->mu.unlock();
      return esk_hip_fail(e, "hipEventRecord");
    }
  }
  if (need_pts > w->cap || temp > w->temp_bytes) {
    // earlier launches on any stream may still read the old buffers
    (void)hipEventSynchronize(w->done);
    if (w->buf) (void)hipFree(w->buf);
    w->buf = nullptr;
    const size_t cap = std::max(need_pts, w->cap), tb = std::max(temp, w->temp_bytes);
    if ((e = hipMalloc(&w->buf, cap * 16 + tb)) != hipSuccess) {
      w->cap = w->temp_bytes = 0;
      w->mu.unlock();
      return esk_fail(EINSPLINE_ENOMEM, "locality order: cannot allocate %zu bytes", cap * 16 + tb);
    }
    w->cap = cap;
    w->temp_bytes = tb;
  }
  // order this stream after the previous user of the buffers
  if ((e = hipStreamWaitEvent(s, w->done, 0)) != hipSuccess) { w->mu.unlock(); return esk_hip_fail(e, "hipStreamWaitEvent"); }
  unsigned *k0 = (unsigned *)w->buf, *k1 = k0 + w->cap;
  int *i0 = (int *)(k1 + w->cap), *i1 = i0 + w->cap;
  void *tmp = (void *)(i1 + w->cap);
  const int blocks = (num + 255) / 256;
  if (pos_f32)
    hipLaunchKernelGGL(cell_keys<float>, dim3(blocks), dim3(256), 0, s, (const float *)pos, num, P, shift, k0, i0);
  else
    hipLaunchKernelGGL(cell_keys<double>, dim3(blocks), dim3(256), 0, s, (const double *)pos, num, P, shift, k0, i0);
  if ((e = hipGetLastError()) != hipSuccess) { w->mu.unlock(); return esk_hip_fail(e, "cell_keys launch"); }
  size_t tb = w->temp_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, i0, i1, num, 0, end_bit, s);
  if (e != hipSuccess) { w->mu.unlock(); return esk_hip_fail(e, "radix sort"); }
  *perm = i1;
  return EINSPLINE_OK;
}

void esk_order_end(EskOrderWs *w, hipStream_t s) {
  (void)hipEventRecord(w->done, s);
  w->mu.unlock();
}
