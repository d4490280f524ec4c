// point_service.hip -- low-latency single-point evaluation.
//
// einspline's original API evaluates one point per call
// (eval_multi_UBspline_3d_{s,d,c,z}[_vg|_vgl|_vgh], src/multi_bspline_eval_*.h),
// and callers invoke it in tight loops.  A kernel launch plus a stream
// synchronisation alone costs about 11 us on MI355X, so calls on 3-D tables
// are served by a resident kernel instead: one block whose first wavefront
// polls a mailbox in pinned host memory; the block evaluates the posted
// point with the batch kernel's per-chunk code (eval_chunk, same
// arithmetic, so the same bits) and writes the results straight into
// pinned host memory.  The host posts
// a request by writing the mailbox and bumping its sequence number, then
// spins on the "done" number; no HIP call is on that path.
//
// Lifetime: the kernel exits by itself after `idle` without a request (and
// when the host sets `stop`, at process exit).  The host does not track
// whether it is running: after a request has waited for a while it asks
// the stream, and launches the kernel again if it has exited, so a request
// posted while the kernel was leaving is still served.  Every wait is
// bounded on both sides.
//
// While the kernel is resident it holds one block slot (four wavefronts) and a
// device-wide synchronisation (hipDeviceSynchronize, and calls that imply
// one) waits for it to go idle: EINSPLINE_POINT_SERVICE_IDLE_US sets the
// idle time (default 2000 us) and EINSPLINE_POINT_SERVICE=0 turns the
// service off (every call is then one launch and one synchronisation).
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <mutex>

#include "eval_kernel.hpp"
#include "runtime_internal.hpp"

namespace {

// The mailbox: written by the host (request fields, then `req`), read by
// the kernel; `done` written by the kernel after the results.
struct alignas(64) Mailbox {
  unsigned req;          // sequence number of the posted request
  unsigned done;         // sequence number of the last request served
  unsigned stop;         // nonzero: the kernel exits
  int kind;              // F64 << 3 | complex << 2 | mode (MODE_V .. MODE_VGH)
  double pos[3];
  esk::Params P;         // the batch kernel's arguments for one point (pos unset)
};

constexpr int kParamWords = (int)(sizeof(esk::Params) / 4);
static_assert(sizeof(esk::Params) % 4 == 0, "Params copied as 32-bit words");

// The service evaluates with one real (one complex number) per lane and
// all 64 stencil rows requested at once (half of them for complex FP64,
// to stay within 128 row registers): the shortest latency for one point.
template <class S, bool CPLX> struct ServeTile {
  static constexpr int LB = (CPLX ? 2 : 1) * (int)sizeof(S);
  static constexpr int PF = CPLX && sizeof(S) == 8 ? 7 : 15;
};

template <class S, int MODE, bool CPLX>
__device__ __forceinline__ void serve_point(const esk::Params &P, int tid) {
  using T = ServeTile<S, CPLX>;
  for (int ch = tid; ch < P.chunks; ch += 256)
    esk::eval_chunk<S, T::LB, MODE, CPLX, false, double, false, T::PF>(P, 0, ch);
}

template <class S, bool CPLX>
__device__ __forceinline__ void serve_mode(const esk::Params &P, int mode, int tid) {
  switch (mode) {
    case esk::MODE_V: serve_point<S, esk::MODE_V, CPLX>(P, tid); break;
    case esk::MODE_VG: serve_point<S, esk::MODE_VG, CPLX>(P, tid); break;
    case esk::MODE_VGL: serve_point<S, esk::MODE_VGL, CPLX>(P, tid); break;
    default: serve_point<S, esk::MODE_VGH, CPLX>(P, tid); break;
  }
}

// Four wavefronts: the first polls the mailbox and copies a request into
// LDS, all four evaluate it.  `idle` is in ticks of the 100 MHz real-time
// counter.
__global__ void __launch_bounds__(256) point_service_kernel(Mailbox *mb, unsigned long long idle) {
  __shared__ unsigned words[kParamWords];
  __shared__ double spos[3];
  __shared__ int cmd[2];                     // [0]: 1 serve, 0 leave; [1]: request kind
  const int tid = threadIdx.x, lane = tid & 63;
  unsigned last = __hip_atomic_load(&mb->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  unsigned long long t_last = __builtin_amdgcn_s_memrealtime();
  unsigned s = last;
  for (;;) {
    if (tid < 64) {
      bool go = false;
      for (unsigned poll = 1;; ++poll) {
        // one read of host memory per poll (each is a PCIe round trip)
        s = (unsigned)esk::uniform((int)__hip_atomic_load(&mb->req, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM));
        if (s != last) { go = true; break; }
        if (__builtin_amdgcn_s_memrealtime() - t_last > idle) break;
        if ((poll & 63) == 0 && __hip_atomic_load(&mb->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) break;
        __builtin_amdgcn_s_sleep(2);
      }
      if (go) {
        const unsigned *src = (const unsigned *)&mb->P;
        for (int i = lane; i < kParamWords; i += 64) words[i] = src[i];
        if (lane < 3) spos[lane] = mb->pos[lane];
        if (lane == 0) cmd[1] = mb->kind;
      }
      if (lane == 0) cmd[0] = go ? 1 : 0;
    }
    __syncthreads();
    if (!cmd[0]) break;                      // the same for the whole block
    esk::Params P;
    __builtin_memcpy(&P, words, sizeof P);
    P.pos = spos;
    const int kind = cmd[1], mode = kind & 3;
    if (kind & 8) {
      if (kind & 4) serve_mode<double, true>(P, mode, tid);
      else serve_mode<double, false>(P, mode, tid);
    } else {
      if (kind & 4) serve_mode<float, true>(P, mode, tid);
      else serve_mode<float, false>(P, mode, tid);
    }
    // every wavefront's results reach host memory before the sequence
    // number does (and the request in LDS is no longer read)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&mb->done, s, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    last = s;
    t_last = __builtin_amdgcn_s_memrealtime();
  }
}

// Host side, one service per device.
struct Service {
  std::mutex mu;
  Mailbox *mb = nullptr;          // host address (pinned, coherent, mapped)
  Mailbox *mb_dev = nullptr;      // the device's address of it
  unsigned char *out = nullptr;   // result staging, host address
  unsigned char *out_dev = nullptr;
  size_t out_bytes = 0;
  hipStream_t stream = nullptr;
  unsigned seq = 0;
  bool launched = false;
  bool broken = false;
};
constexpr int kMaxDevices = 64;
Service g_svc[kMaxDevices];
std::once_flag g_exit_once;

// Settings: from EINSPLINE_POINT_SERVICE / EINSPLINE_POINT_SERVICE_IDLE_US
// at first use, or einspline_set_point_service.
std::mutex g_set_mu;
int g_enabled = -1;                 // -1: not read yet
long g_idle_us = 2000;

void read_env_locked() {
  if (g_enabled >= 0) return;
  const char *e = getenv("EINSPLINE_POINT_SERVICE");
  g_enabled = e && e[0] == '0' ? 0 : 1;
  const char *i = getenv("EINSPLINE_POINT_SERVICE_IDLE_US");
  if (i) g_idle_us = atol(i);
}

long clamp_idle(long us) { return us < 100 ? 100 : us > 1000000 ? 1000000 : us; }

bool enabled(unsigned long long *idle_ticks) {
  std::lock_guard<std::mutex> g(g_set_mu);
  read_env_locked();
  *idle_ticks = (unsigned long long)clamp_idle(g_idle_us) * 100ull;    // 100 MHz counter
  return g_enabled != 0;
}

// Ask a resident kernel to leave and wait for it (caller holds v.mu).
void stop_locked(Service &v) {
  if (!v.launched || !v.mb) return;
  __atomic_store_n(&v.mb->stop, 1u, __ATOMIC_RELEASE);
  (void)hipStreamSynchronize(v.stream);
  __atomic_store_n(&v.mb->stop, 0u, __ATOMIC_RELEASE);
  v.launched = false;
}

// At exit: every resident kernel leaves before the runtime goes away.
void stop_all() {
  for (int d = 0; d < kMaxDevices; ++d) {
    std::lock_guard<std::mutex> g(g_svc[d].mu);
    stop_locked(g_svc[d]);
  }
}

int setup(Service &v, size_t out_bytes) {
  hipError_t e;
  if (!v.mb) {
    e = hipHostMalloc((void **)&v.mb, sizeof(Mailbox), hipHostMallocMapped | hipHostMallocCoherent);
    if (e != hipSuccess) return esk_hip_fail(e, "hipHostMalloc point-service mailbox");
    memset(v.mb, 0, sizeof(Mailbox));
    e = hipHostGetDevicePointer((void **)&v.mb_dev, v.mb, 0);
    if (e != hipSuccess) return esk_hip_fail(e, "hipHostGetDevicePointer mailbox");
    e = hipStreamCreateWithFlags(&v.stream, hipStreamNonBlocking);
    if (e != hipSuccess) return esk_hip_fail(e, "hipStreamCreate point service");
  }
  if (v.out_bytes < out_bytes) {
    // no request is in flight (the caller holds the lock and waits for each)
    if (v.out) (void)hipHostFree(v.out);
    v.out = v.out_dev = nullptr;
    v.out_bytes = 0;
    const size_t want = out_bytes > ((size_t)64 << 10) ? out_bytes : ((size_t)64 << 10);
    e = hipHostMalloc((void **)&v.out, want, hipHostMallocMapped | hipHostMallocCoherent);
    if (e != hipSuccess) return esk_hip_fail(e, "hipHostMalloc point-service results");
    e = hipHostGetDevicePointer((void **)&v.out_dev, v.out, 0);
    if (e != hipSuccess) return esk_hip_fail(e, "hipHostGetDevicePointer results");
    v.out_bytes = want;
  }
  return EINSPLINE_OK;
}

int launch(Service &v, unsigned long long idle_ticks) {
  std::call_once(g_exit_once, [] { atexit(stop_all); });
  hipLaunchKernelGGL(point_service_kernel, dim3(1), dim3(256), 0, v.stream, v.mb_dev, idle_ticks);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return esk_hip_fail(e, "point-service launch");
  v.launched = true;
  return EINSPLINE_OK;
}

}  // namespace

// Serve one point of a 3-D table through the resident kernel: 1 when done
// (results copied to vals/grads/lh), 0 when the call is not one the service
// takes (the caller then launches), negative on error.  Layout of the
// results as for a batch of one point: values (N), gradients (3N), then the
// Laplacian (N) or the full Hessian (9N).
int esk_point_service_eval(const multi_UBspline_3d_hip *t, int mode, double x, double y, double z, void *vals,
                           void *grads, void *lh) {
  unsigned long long idle = 0;
  if (!enabled(&idle) || !t || t->dim != 3 || mode < esk::MODE_V || mode > esk::MODE_VGH) return 0;
  const bool f64 = t->storage == EINSPLINE_STORE_F64;
  if (!f64 && t->storage != EINSPLINE_STORE_F32) return 0;
  const size_t eb = f64 ? 8 : 4;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return esk_hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return 0;
  Service &v = g_svc[dev];
  std::lock_guard<std::mutex> guard(v.mu);
  if (v.broken) return 0;

  esk::Launch L;
  esk::Params P;
  int rc = esk_prepare(t, 1, L, P);
  if (rc) return rc;
  const bool cplx = t->tcode == SINGLE_COMPLEX || t->tcode == DOUBLE_COMPLEX;
  const size_t nr = (size_t)P.nreal;                         // reals per output row
  const size_t nv = nr * eb, ng = mode >= esk::MODE_VG ? 3 * nv : 0;
  const size_t nh = mode == esk::MODE_VGH ? 9 * nv : mode == esk::MODE_VGL ? nv : 0;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_g = up(nv), o_h = o_g + up(ng), total = o_h + up(nh);
  if ((rc = setup(v, total))) return rc;

  P.chunks = cplx ? t->row_reals / 2 : t->row_reals;          // one real (complex number) per lane
  P.npoints = 1;
  P.pos = nullptr;                                           // the kernel's copy of the position
  P.out_v = v.out_dev;
  P.out_g = ng ? v.out_dev + o_g : nullptr;
  P.out_h = nh ? v.out_dev + o_h : nullptr;
  P.ld_v = P.nreal;
  P.ld_g = 3l * P.nreal;
  P.ld_h = mode == esk::MODE_VGL ? (long)P.nreal : 9l * P.nreal;
  P.hpacked = 0;
  P.store = esk::STORE_DIRECT;
  P.perm = nullptr;
  P.xcd_runs = 0;
  P.slabs = 0;

  Mailbox *mb = v.mb;
  mb->pos[0] = x; mb->pos[1] = y; mb->pos[2] = z;
  mb->kind = (f64 ? 8 : 0) | (cplx ? 4 : 0) | mode;
  memcpy(&mb->P, &P, sizeof P);
  const unsigned seq = ++v.seq;
  __atomic_store_n(&mb->req, seq, __ATOMIC_RELEASE);
  if (!v.launched && (rc = launch(v, idle))) return rc;

  // Wait for the result.  A resident kernel answers within microseconds;
  // after 30 us ask the stream whether the kernel is still there (it leaves
  // after its idle time) and launch it again if not.
  using clk = std::chrono::steady_clock;
  clk::time_point t_post = clk::now(), t_check = t_post;
  for (unsigned it = 1;; ++it) {
    if (__atomic_load_n(&mb->done, __ATOMIC_ACQUIRE) == seq) break;
    if ((it & 63) == 0) {
      const clk::time_point now = clk::now();
      if (now - t_check > std::chrono::microseconds(30)) {
        t_check = now;
        e = hipStreamQuery(v.stream);
        if (e == hipSuccess) {
          if (__atomic_load_n(&mb->done, __ATOMIC_ACQUIRE) == seq) break;
          v.launched = false;
          if ((rc = launch(v, idle))) return rc;
        } else if (e != hipErrorNotReady) {
          v.broken = true;
          return esk_hip_fail(e, "point service");
        }
        if (now - t_post > std::chrono::seconds(10)) {
          v.broken = true;
          return esk_fail(EINSPLINE_EHIP, "point service: no answer in 10 s");
        }
      }
    }
    __builtin_ia32_pause();
  }
  memcpy(vals, v.out, nv);
  if (ng) memcpy(grads, v.out + o_g, ng);
  if (nh) memcpy(lh, v.out + o_h, nh);
  return 1;
}

extern "C" int einspline_set_point_service(int enable, int idle_us) {
  if (enable < 0 || enable > 1) return esk_fail(EINSPLINE_EINVAL, "enable must be 0 or 1");
  {
    std::lock_guard<std::mutex> g(g_set_mu);
    read_env_locked();
    g_enabled = enable;
    if (idle_us > 0) g_idle_us = clamp_idle(idle_us);
  }
  // resident kernels leave; the next call starts one with the new settings
  stop_all();
  return EINSPLINE_OK;
}

extern "C" int einspline_point_service_resident(void) {
  int n = 0;
  for (int d = 0; d < kMaxDevices; ++d) {
    Service &v = g_svc[d];
    std::lock_guard<std::mutex> g(v.mu);
    if (!v.launched) continue;
    const hipError_t e = hipStreamQuery(v.stream);
    if (e == hipErrorNotReady) ++n;
    else v.launched = false;
  }
  return n;
}
