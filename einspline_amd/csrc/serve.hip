// serve.hip -- serving runtime (include/einspline_serve.h).
//
// One worker thread owns the GPU side.  It keeps `slots` batches in flight,
// each on its own stream with its own pinned staging buffer, device buffer
// and instantiated hipGraph.  A batch is a list of segments (one per
// request) grouped by kernel variant; the graph holds one persistent kernel
// per non-empty group.  The persistent kernels read the batch descriptor
// from device memory, so one captured graph serves every batch with the same
// set of groups, whatever its size.  Each wavefront walks a strided range
// of "wave work units" (64 lanes of one segment) until the group is done,
// so every wavefront reaches the exit.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "einspline_serve.h"
#include "einspline_private.h"
#include "eval_kernel.hpp"
#include "runtime_internal.hpp"
#include "serve_policy.hpp"

namespace {

using clk = std::chrono::steady_clock;

// kernel groups: storage (F32, BF16, F16) x (one point per wavefront or not)
constexpr int kGroups = 6;
constexpr int kMaxSeg = 256;

struct Seg {
  esk::Params P;   // table, positions and outputs of one request
  int wave0;       // first wave work unit of the segment within its group
  int waves;
};

// Re-admission copy: device memory from pinned host memory through its
// device address, 16 bytes per lane and then any tail bytes.  A copy kernel
// runs at the PCIe rate whatever DMA engine a stream would be given (on
// MI355X, DMA copies of re-admitted tables ran at 57 or 30 GB/s from one
// process to the next); it takes kPullBlocks workgroups while it runs.
constexpr int kPullBlocks = 128;
__global__ void __launch_bounds__(256) pull_table(const unsigned char *__restrict__ src,
                                                  unsigned char *__restrict__ dst, size_t bytes) {
  const size_t n16 = bytes / 16;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t first = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint4 *s16 = (const uint4 *)src;
  uint4 *d16 = (uint4 *)dst;
  for (size_t i = first; i < n16; i += stride) d16[i] = s16[i];
  for (size_t i = 16 * n16 + first; i < bytes; i += stride) dst[i] = src[i];
}

struct BatchDesc {
  int nseg[kGroups];
  int seg0[kGroups];
  int total_waves[kGroups];
  int pad[2];
  Seg seg[kMaxSeg];
};

template <class S, bool UNIF>
__global__ void __launch_bounds__(256, (esk::Occupancy<S, 16, esk::MODE_VGH, UNIF>::waves))
serve_vgh(const BatchDesc *__restrict__ B, int g) {
  const int lane = threadIdx.x & 63;
  const int wv = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nw = (int)((gridDim.x * blockDim.x) >> 6);
  const int total = B->total_waves[g];
  const int s0 = B->seg0[g], ns = B->nseg[g];
  int s = 0;
  for (int w = esk::uniform(wv); w < total; w += nw) {
    while (s + 1 < ns && B->seg[s0 + s + 1].wave0 <= w) ++s;
    s = esk::uniform(s);
    const Seg &sg = B->seg[s0 + s];
    const int chunks = sg.P.chunks;
    const int lid = (w - sg.wave0) * 64 + lane;
    int p = lid / chunks;
    if (UNIF) p = esk::uniform(p);
    const int ch = lid - p * chunks;
    if (p < sg.P.npoints)
      esk::eval_chunk<S, 16, esk::MODE_VGH, false, UNIF, double, false>(sg.P, p, ch);
  }
}

typedef void (*serve_fn)(const BatchDesc *, int);

serve_fn group_kernel(int g) {
  switch (g) {
    case 0: return serve_vgh<float, true>;
    case 1: return serve_vgh<float, false>;
    case 2: return serve_vgh<esk::bf16s, true>;
    case 3: return serve_vgh<esk::bf16s, false>;
    case 4: return serve_vgh<esk::f16s, true>;
    default: return serve_vgh<esk::f16s, false>;
  }
}

// Dynamic LDS of a group kernel: room for staged stores in every wavefront
// (requests keep their table's store policy).
size_t group_lds(int g) {
  switch (g) {
    case 0: return esk::StageLds<float, 16, esk::MODE_VGH, true>::bytes(256, esk::STORE_STAGED);
    case 2: return esk::StageLds<esk::bf16s, 16, esk::MODE_VGH, true>::bytes(256, esk::STORE_STAGED);
    case 4: return esk::StageLds<esk::f16s, 16, esk::MODE_VGH, true>::bytes(256, esk::STORE_STAGED);
    default: return 0;   // several points per wavefront: no staging
  }
}

int storage_group(einspline_storage st) {
  return st == EINSPLINE_STORE_F32 ? 0 : st == EINSPLINE_STORE_BF16 ? 2 : 4;
}

struct Table {
  void *host = nullptr;
  bool from_double = false;
  einspline_storage storage = EINSPLINE_STORE_F32;
  size_t bytes = 0;                 // device bytes once admitted
  multi_UBspline_3d_hip *dev = nullptr;
  bool loading = false;             // upload in progress on the uploader thread
  size_t reserved = 0;              // bytes reserved for the upload in progress
  void *pinned = nullptr;           // storage-format copy in pinned host memory
  multi_UBspline_3d_hip meta{};     // geometry of the device table (valid with pinned)
  uint64_t last_use = 0;
  int inflight = 0;                 // requests of this table on the GPU
  int queued = 0;                   // requests of this table in the pending queue
  bool draining = false;            // an absent table waits for this one to go idle
  clk::time_point drain_since{};    // when it was first asked to drain: later requests wait
  int nsplines = 0;
};

// A resident table asked to make room for an absent one ("draining") serves
// the requests that were queued for it when the drain began, all of them and
// no later ones; then it gets no new work, goes idle and is evicted.  Every
// request is thus served in the first residency of its table that begins
// after it arrives, however long that residency's queue is, and a residency
// ends as soon as its queue at that moment is served.  Serving each table's
// queue whole instead of request by request keeps the number of uploads down
// when two tables that do not fit together are both in demand.

struct Req {
  uint64_t id = 0;
  int table = 0;
  int npts = 0;
  std::vector<double> pos;
  void *v = nullptr, *g = nullptr, *h = nullptr;
  clk::time_point t0;
  double latency = -1.0;
  bool done = false;
  bool collect = true;              // false: nobody waits, drop at completion
  bool taken = false;               // in the batch being formed
  bool failed = false;              // its batch failed: the outputs are not valid
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  unsigned char *hbuf = nullptr;    // pinned: descriptor + positions
  unsigned char *dbuf = nullptr;
  size_t bytes = 0;
  std::map<unsigned, hipGraphExec_t> graphs;   // by non-empty group mask
  bool busy = false;
  std::vector<Req *> reqs;
};

}  // namespace

struct einspline_server {
  einspline_server_config cfg{};
  int device = 0;
  int blocks[kGroups] = {};         // persistent grid per group
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> pending;
  std::unordered_map<uint64_t, Req *> live;
  std::vector<Table> tables;
  std::vector<Slot> slots;
  std::thread worker, uploader;
  // Re-admissions run pull_table on a non-blocking stream of their own.
  hipStream_t upload_stream = nullptr;
  std::deque<int> uploads;          // tables waiting for the uploader thread
  std::vector<multi_UBspline_3d_hip *> frees;   // evicted device tables, freed by the uploader
  std::condition_variable cv_upload;
  bool stop = false;
  bool stop_uploader = false;       // set after the worker has drained
  int worker_error = 0;             // first failure of a batch or an upload: sticky
  std::string worker_msg;           // its message (errors are recorded per thread)
  uint64_t next_id = 1, use_clock = 0;
  size_t budget = 0, used = 0;
  einspline_server_stats st{};
  std::vector<double> lat_log;      // latencies of completed requests (seconds)

  int init(const einspline_server_config *c);
  void run();
  void upload_loop();
  void poll_completions();
  bool launch_batch(Slot &s, std::unique_lock<std::mutex> &lk);
  int admit(int tid, bool may_evict, const std::vector<char> &held, std::vector<char> &drain_now);
  hipGraphExec_t graph_for(Slot &s, unsigned mask);
  // Record the first failure (the calling thread's last error message goes
  // with it), under the lock.
  void set_error(int code) {
    if (worker_error) return;
    worker_error = code;
    worker_msg = einspline_hip_last_error();
  }
  // Hand the sticky failure to a caller's thread: its code and message.
  int error_for_caller() const { return esk_fail(worker_error, "%s", worker_msg.c_str()); }
};

// ---------------------------------------------------------------------------

static size_t predicted_bytes(const Table &t) {
  // same padding rule as the device upload: rows padded to 32 bytes
  const int eb = t.storage == EINSPLINE_STORE_F32 ? 4 : 2;
  int Nx, Ny, Nz, N;
  if (t.from_double) {
    const multi_UBspline_3d_d *s = (const multi_UBspline_3d_d *)t.host;
    Nz = (int)(s->y_stride / s->z_stride); Ny = (int)(s->x_stride / s->y_stride);
    N = s->num_splines;
  } else {
    const multi_UBspline_3d_s *s = (const multi_UBspline_3d_s *)t.host;
    Nz = (int)(s->y_stride / s->z_stride); Ny = (int)(s->x_stride / s->y_stride);
    N = s->num_splines;
  }
  const einspline_host3d *h = einspline_host_of(t.host);
  Nx = h ? h->Nx : 0;
  const size_t per32 = 32 / eb;
  const size_t row = ((size_t)N + per32 - 1) / per32 * per32;
  return (size_t)Nx * Ny * Nz * row * eb;
}

int einspline_server::init(const einspline_server_config *c) {
  cfg = *c;
  if (cfg.slots <= 0) cfg.slots = 4;
  if (cfg.max_batch_points <= 0) cfg.max_batch_points = 16384;
  if (cfg.max_batch_requests <= 0) cfg.max_batch_requests = 256;
  cfg.max_batch_requests = std::min(cfg.max_batch_requests, kMaxSeg);
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return esk_hip_fail(e, "hipGetDevice");
  if (cfg.hbm_budget == 0) {
    size_t fr = 0, tot = 0;
    e = hipMemGetInfo(&fr, &tot);
    if (e != hipSuccess) return esk_hip_fail(e, "hipMemGetInfo");
    cfg.hbm_budget = fr / 10 * 9;
  }
  budget = cfg.hbm_budget;
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return esk_hip_fail(e, "hipGetDeviceProperties");
  for (int g = 0; g < kGroups; ++g) {
    const size_t lds = group_lds(g);
    if (lds > 65536) {
      e = hipFuncSetAttribute((const void *)group_kernel(g), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
      if (e != hipSuccess) return esk_hip_fail(e, "hipFuncSetAttribute");
    }
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)group_kernel(g), 256, lds);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    blocks[g] = prop.multiProcessorCount * per_cu;
  }
  const size_t bytes = sizeof(BatchDesc) + 16 + (size_t)cfg.max_batch_points * 3 * sizeof(double);
  slots.resize(cfg.slots);
  for (Slot &s : slots) {
    s.bytes = bytes;
    if ((e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipHostMalloc((void **)&s.hbuf, bytes, hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc((void **)&s.dbuf, bytes)) != hipSuccess)
      return esk_hip_fail(e, "server slot setup");
  }
  if ((e = hipStreamCreateWithFlags(&upload_stream, hipStreamNonBlocking)) != hipSuccess)
    return esk_hip_fail(e, "server upload stream");
  worker = std::thread([this] { run(); });
  uploader = std::thread([this] { upload_loop(); });
  return EINSPLINE_OK;
}

hipGraphExec_t einspline_server::graph_for(Slot &s, unsigned mask) {
  auto it = s.graphs.find(mask);
  if (it != s.graphs.end()) return it->second;
  // The graph is built node by node rather than captured: while a stream
  // captures, the HIP runtime refuses legacy-stream work from other threads
  // of the process (a caller's hipMemcpy), which a server cannot forbid.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipGraphCreate(&graph, 0) != hipSuccess) return nullptr;
  const BatchDesc *d = (const BatchDesc *)s.dbuf;
  hipGraphNode_t prev = nullptr;
  hipError_t e = hipSuccess;
  for (int g = 0; g < kGroups && e == hipSuccess; ++g) {
    if (!(mask & (1u << g))) continue;
    int group = g;
    void *args[] = {(void *)&d, (void *)&group};
    hipKernelNodeParams p = {};
    p.func = (void *)group_kernel(g);
    p.gridDim = dim3(blocks[g]);
    p.blockDim = dim3(256);
    p.sharedMemBytes = (unsigned)group_lds(g);
    p.kernelParams = args;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    // one group after another, as launches on one stream would run
    e = hipGraphAddKernelNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &p);
    prev = node;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return nullptr;
  s.graphs[mask] = exec;
  return exec;
}

// Make table `tid` resident.  Returns 1 when resident, 0 when deferred
// (upload queued or in progress, or waiting for room).  The choice is
// esk::plan_admission's: evict the cheapest set of idle tables that makes
// room, or, when that set holds a table with work in flight (or in the
// batch being formed: `held`), mark it in `drain_now` so that it is drained
// (it serves what was queued for it before the drain began) and evicted once
// idle.  Only the oldest waiting request of
// a pass may evict or drain (`may_evict`); later ones upload only into free
// room.  Evicted tables are handed to the uploader thread, which frees them
// outside the lock; uploads run there too, so the worker keeps launching
// batches for resident tables meanwhile.
int einspline_server::admit(int tid, bool may_evict, const std::vector<char> &held,
                            std::vector<char> &drain_now) {
  Table &t = tables[tid];
  if (t.dev) return 1;
  if (t.loading) return 0;
  std::vector<esk::AdmitTable> view(tables.size());
  for (size_t i = 0; i < tables.size(); ++i) {
    const Table &c = tables[i];
    view[i].bytes = c.dev ? c.bytes : c.loading ? c.reserved : predicted_bytes(c);
    view[i].state = c.dev ? 1 : c.loading ? 2 : 0;
    view[i].busy = c.inflight > 0 || held[i];
    view[i].pending = c.queued;
    view[i].last_use = c.last_use;
  }
  std::vector<int> victims, drain;
  const int d = esk::plan_admission(view, tid, used, budget, victims, drain);
  if (d != esk::ADMIT_UPLOAD && d != esk::ADMIT_DEFER) return 0;
  if (!may_evict && (!victims.empty() || !drain.empty())) return 0;
  if (d == esk::ADMIT_DEFER) {
    for (int i : drain) drain_now[i] = 1;
    return 0;
  }
  for (int i : victims) {
    Table &v = tables[i];
    frees.push_back(v.dev);
    v.dev = nullptr;
    v.draining = false;
    used -= v.bytes;
    st.evictions++;
  }
  const size_t need = predicted_bytes(t);
  used += need;
  t.reserved = need;
  t.loading = true;
  uploads.push_back(tid);
  cv_upload.notify_one();
  return 0;
}

void einspline_server::upload_loop() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_upload.wait(lk, [&] { return stop_uploader || !uploads.empty() || !frees.empty(); });
    if (!frees.empty()) {
      // free evicted tables before allocating for the next upload
      std::vector<multi_UBspline_3d_hip *> f;
      f.swap(frees);
      lk.unlock();
      for (multi_UBspline_3d_hip *d : f) destroy_multi_UBspline_3d_hip(d);
      lk.lock();
      continue;
    }
    if (uploads.empty()) {
      if (stop_uploader) return;
      continue;
    }
    const int tid = uploads.front();
    uploads.pop_front();
    Table &t = tables[tid];
    void *host = t.host;
    const bool from_double = t.from_double;
    const einspline_storage storage = t.storage;
    void *pinned = t.pinned;
    const multi_UBspline_3d_hip meta = t.meta;
    lk.unlock();
    multi_UBspline_3d_hip *d = nullptr;
    void *new_pinned = nullptr;
    int err = EINSPLINE_OK;
    if (pinned) {
      // re-admission: the converted table comes straight from pinned memory
      void *dc = nullptr, *src = nullptr;
      hipError_t e = hipHostGetDevicePointer(&src, pinned, 0);
      if (e == hipSuccess) e = hipMalloc(&dc, meta.bytes);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(pull_table, dim3(kPullBlocks), dim3(256), 0, upload_stream,
                           (const unsigned char *)src, (unsigned char *)dc, meta.bytes);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(upload_stream);
      }
      if (e != hipSuccess) err = esk_hip_fail(e, "server: re-admitting a table");
      else if (!(d = esk_table_wrap(&meta, dc))) err = EINSPLINE_ENOMEM;
      if (!d && dc) (void)hipFree(dc);
    } else {
      d = from_double ? create_multi_UBspline_3d_s_hip_store((const multi_UBspline_3d_d *)host, storage)
                      : create_multi_UBspline_3d_s_hip((multi_UBspline_3d_s *)host);
      if (!d) err = EINSPLINE_ENOMEM;   // the create call's message says what failed
      // keep a storage-format host copy so later re-admissions skip conversion
      if (d && hipHostMalloc(&new_pinned, d->bytes, hipHostMallocDefault) == hipSuccess) {
        if (hipMemcpy(new_pinned, d->coefs, d->bytes, hipMemcpyDeviceToHost) != hipSuccess) {
          (void)hipHostFree(new_pinned);
          new_pinned = nullptr;
        }
      } else {
        new_pinned = nullptr;
      }
    }
    lk.lock();
    Table &u = tables[tid];
    used -= u.reserved;
    u.reserved = 0;
    u.loading = false;
    if (d) {
      u.dev = d;
      u.bytes = d->bytes;
      used += d->bytes;
      st.admissions++;
      if (new_pinned) {
        u.pinned = new_pinned;
        u.meta = *d;
        u.meta.coefs = nullptr;
      }
    } else {
      // Requests for this table cannot be served: fail the server rather
      // than retry the upload for ever (waiters and submitters get the error).
      set_error(err);
      cv_done.notify_all();
    }
    cv_work.notify_all();
  }
}

void einspline_server::poll_completions() {
  for (Slot &s : slots) {
    if (!s.busy) continue;
    hipError_t e = hipEventQuery(s.done);
    if (e == hipErrorNotReady) continue;
    const bool bad = e != hipSuccess;
    if (bad) set_error(esk_hip_fail(e, "server batch"));
    const auto now = clk::now();
    for (Req *r : s.reqs) {
      r->latency = std::chrono::duration<double>(now - r->t0).count();
      r->done = true;
      r->failed |= bad;
      tables[r->table].inflight--;
      st.completed++;
      st.evals += (double)r->npts * tables[r->table].nsplines;
      lat_log.push_back(r->latency);
      if (!r->collect) { live.erase(r->id); delete r; }
    }
    s.reqs.clear();
    s.busy = false;
  }
}

// Form and launch one batch on slot s from the pending queue.
bool einspline_server::launch_batch(Slot &s, std::unique_lock<std::mutex> &lk) {
  std::vector<char> held(tables.size(), 0), drain_now(tables.size(), 0);
  std::vector<Req *> take;
  int points = 0;
  const auto now = clk::now();
  // Pass 1: requests of resident tables in arrival order, over the whole
  // queue, so a table serves everything queued for it; a table being
  // drained only what arrived before its drain began.  Only this thread
  // removes from `pending`; submitters append.
  for (Req *r : pending) {
    if ((int)take.size() >= cfg.max_batch_requests || points >= cfg.max_batch_points) break;
    const Table &t = tables[r->table];
    if (!t.dev || points + r->npts > cfg.max_batch_points) continue;
    if (t.draining && r->t0 > t.drain_since) {
      st.deferrals++;
      continue;
    }
    held[r->table] = 1;
    r->taken = true;
    take.push_back(r);
    points += r->npts;
  }
  if (!take.empty())
    pending.erase(std::remove_if(pending.begin(), pending.end(), [](Req *r) { return r->taken; }),
                  pending.end());
  // Pass 2: admission.  The oldest request whose table is absent may evict
  // or drain; later ones only upload into free room.
  bool planned = false;
  std::vector<char> seen(tables.size(), 0);
  for (Req *r : pending) {
    if (tables[r->table].dev || seen[r->table]) continue;
    seen[r->table] = 1;
    (void)admit(r->table, !planned, held, drain_now);
    planned = true;
    st.deferrals++;
  }
  for (size_t i = 0; i < tables.size(); ++i) {
    Table &t = tables[i];
    if (drain_now[i] && !t.draining) t.drain_since = now;
    t.draining = drain_now[i] != 0;
  }
  if (take.empty()) return false;
  // Kernel arguments of each request (its table is resident and, with the
  // request counted in flight below, stays resident until the batch ends).
  std::vector<std::pair<int, Req *>> order;
  std::vector<esk::Params> params(take.size());
  for (size_t i = 0; i < take.size(); ++i) {
    Req *r = take[i];
    Table &t = tables[r->table];
    esk::Launch L;
    esk_prepare(t.dev, r->npts, L, params[i]);
    // serving kernels use 16-byte lanes
    params[i].chunks = t.dev->row_reals * (t.storage == EINSPLINE_STORE_F32 ? 4 : 2) / 16;
    const int g = storage_group(t.storage) + ((params[i].chunks % 64) ? 1 : 0);
    order.push_back({g, r});
    params[i].npoints = r->npts;
    t.queued--;
    t.inflight++;
    t.last_use = ++use_clock;
  }
  s.reqs = take;
  s.busy = true;
  st.batches++;
  // The descriptor, the position copy and the launches run without the
  // lock, so submitters are not held up by them.  Only this thread forms
  // batches and touches a slot, and a busy slot's requests are not read by
  // others until poll_completions (this thread) marks them done.
  lk.unlock();
  std::vector<size_t> idx(take.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return order[a].first < order[b].first; });
  // Layout of the copy: header, the batch's segments, then its positions
  // (16-byte aligned); only these bytes go to the device.
  BatchDesc *B = (BatchDesc *)s.hbuf;
  const size_t pos_off = (offsetof(BatchDesc, seg) + take.size() * sizeof(Seg) + 15) & ~(size_t)15;
  double *hpos = (double *)(s.hbuf + pos_off);
  const double *dpos = (const double *)(s.dbuf + pos_off);
  memset(B, 0, offsetof(BatchDesc, seg));
  int off = 0;
  for (size_t k = 0; k < idx.size(); ++k) {
    const size_t i = idx[k];
    Req *r = order[i].second;
    const int g = order[i].first;
    if (B->nseg[g] == 0) B->seg0[g] = (int)k;
    Seg &sg = B->seg[k];
    sg.P = params[i];
    memcpy(hpos + 3 * (size_t)off, r->pos.data(), sizeof(double) * 3 * r->npts);
    sg.P.pos = dpos + 3 * (size_t)off;
    sg.P.out_v = r->v; sg.P.out_g = r->g; sg.P.out_h = r->h;
    sg.P.ld_v = sg.P.nreal;
    sg.P.ld_g = 3l * sg.P.nreal;
    sg.P.ld_h = (cfg.hess_packed ? 6l : 9l) * sg.P.nreal;
    sg.P.hpacked = cfg.hess_packed ? 1 : 0;
    sg.wave0 = B->total_waves[g];
    sg.waves = (int)(((long)r->npts * sg.P.chunks + 63) / 64);
    B->total_waves[g] += sg.waves;
    B->nseg[g]++;
    off += r->npts;
  }
  unsigned mask = 0;
  for (int g = 0; g < kGroups; ++g)
    if (B->nseg[g]) mask |= 1u << g;
  hipGraphExec_t exec = graph_for(s, mask);
  const size_t bytes = pos_off + sizeof(double) * 3 * (size_t)off;
  hipError_t e = exec ? hipMemcpyAsync(s.dbuf, s.hbuf, bytes, hipMemcpyHostToDevice, s.stream)
                      : hipErrorInvalidValue;
  if (e == hipSuccess) e = hipGraphLaunch(exec, s.stream);
  if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
  lk.lock();
  if (e != hipSuccess) {
    set_error(esk_hip_fail(e, "server launch"));
    for (Req *r : s.reqs) r->failed = true;
  }
  return true;
}

void einspline_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    poll_completions();
    cv_done.notify_all();
    bool any_busy = false;
    for (Slot &s : slots) any_busy |= s.busy;
    // After a failure no batch is formed: the batches in flight finish and
    // the requests still queued are released by einspline_server_destroy.
    if (stop && !any_busy && (pending.empty() || worker_error)) break;
    Slot *free_slot = nullptr;
    if (!worker_error)
      for (Slot &s : slots)
        if (!s.busy) { free_slot = &s; break; }
    bool launched = false;
    if (free_slot && !pending.empty()) launched = launch_batch(*free_slot, lk);
    if (!launched) {
      if (any_busy) {
        // completions are polled: yield briefly without sleeping long
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
      } else {
        // idle, or waiting for an upload: submit and the uploader notify
        cv_work.wait_for(lk, std::chrono::milliseconds(1));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// C API

extern "C" einspline_server *einspline_server_create(const einspline_server_config *cfg) {
  einspline_server_config def{};
  einspline_server *s = new einspline_server();
  if (s->init(cfg ? cfg : &def) != EINSPLINE_OK) {
    einspline_server_destroy(s);    // releases whatever init set up (no thread runs yet)
    return nullptr;
  }
  return s;
}

extern "C" void einspline_server_destroy(einspline_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop_uploader = true;
  }
  s->cv_upload.notify_all();
  if (s->uploader.joinable()) s->uploader.join();
  for (Slot &sl : s->slots) {
    for (auto &kv : sl.graphs) (void)hipGraphExecDestroy(kv.second);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.hbuf) (void)hipHostFree(sl.hbuf);
    if (sl.dbuf) (void)hipFree(sl.dbuf);
  }
  if (s->upload_stream) (void)hipStreamDestroy(s->upload_stream);
  for (multi_UBspline_3d_hip *d : s->frees) destroy_multi_UBspline_3d_hip(d);
  for (Table &t : s->tables) {
    if (t.dev) destroy_multi_UBspline_3d_hip(t.dev);
    if (t.pinned) (void)hipHostFree(t.pinned);
  }
  for (auto &kv : s->live) delete kv.second;
  delete s;
}

extern "C" int einspline_server_register_table(einspline_server *s, void *host,
                                               einspline_storage storage, int *table_id) {
  if (!s || !host || !table_id) return esk_fail(EINSPLINE_EINVAL, "NULL argument");
  einspline_host3d *h = einspline_host_of(host);
  if (!h) return esk_fail(EINSPLINE_EINVAL, "spline was not created by this library");
  if (h->dim != 3) return esk_fail(EINSPLINE_EINVAL, "serving takes 3-D tables");
  Table t;
  t.host = host;
  t.storage = storage;
  const type_code tc = h->pub.s.tcode;
  if (tc == SINGLE_REAL && storage == EINSPLINE_STORE_F32) t.from_double = false;
  else if (tc == DOUBLE_REAL && storage != EINSPLINE_STORE_F64) t.from_double = true;
  else return esk_fail(EINSPLINE_EINVAL, "serving supports real float tables (F32) or double tables stored as F32/BF16/F16");
  t.nsplines = h->nsplines;
  std::lock_guard<std::mutex> g(s->mu);
  if (predicted_bytes(t) > s->budget)
    return esk_fail(EINSPLINE_EBUDGET, "table of %zu bytes exceeds the HBM budget of %zu bytes",
                    predicted_bytes(t), s->budget);
  s->tables.push_back(t);
  *table_id = (int)s->tables.size() - 1;
  return EINSPLINE_OK;
}

static int submit_impl(einspline_server *s, int tid, const double *pos, int npts, void *v, void *g,
                       void *h, uint64_t *ticket) {
  if (!s || !pos || !v || !g || !h) return esk_fail(EINSPLINE_EINVAL, "NULL argument");
  if (npts < 1 || npts > s->cfg.max_batch_points)
    return esk_fail(EINSPLINE_EINVAL, "npts must be in 1..%d", s->cfg.max_batch_points);
  Req *r = new Req();
  r->npts = npts;
  r->pos.assign(pos, pos + 3 * (size_t)npts);
  r->v = v; r->g = g; r->h = h;
  r->collect = ticket != nullptr;
  {
    std::lock_guard<std::mutex> lg(s->mu);
    if (tid < 0 || tid >= (int)s->tables.size()) {
      delete r;
      return esk_fail(EINSPLINE_EINVAL, "unknown table id %d", tid);
    }
    if (s->worker_error) {
      delete r;
      return s->error_for_caller();
    }
    r->table = tid;
    r->id = s->next_id++;
    r->t0 = clk::now();
    s->live[r->id] = r;
    s->pending.push_back(r);
    s->tables[tid].queued++;
    s->st.submitted++;
    if (ticket) *ticket = r->id;
  }
  s->cv_work.notify_one();
  return EINSPLINE_OK;
}

extern "C" int einspline_server_submit(einspline_server *s, int tid, const double *pos, int npts,
                                       void *v, void *g, void *h, uint64_t *ticket) {
  return submit_impl(s, tid, pos, npts, v, g, h, ticket);
}

static int finish(einspline_server *s, uint64_t id, double *lat, bool block) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(id);
  if (it == s->live.end()) return esk_fail(EINSPLINE_EINVAL, "unknown or already collected ticket");
  Req *r = it->second;
  if (!r->done) {
    if (!block) return 0;
    s->cv_done.wait(lk, [&] { return r->done || s->worker_error; });
    if (!r->done) return s->error_for_caller();
  }
  if (lat) *lat = r->latency;
  const bool failed = r->failed;
  s->live.erase(it);
  delete r;
  return failed ? s->error_for_caller() : 1;
}

extern "C" int einspline_server_test(einspline_server *s, uint64_t t, double *lat) {
  if (!s) return esk_fail(EINSPLINE_EINVAL, "NULL server");
  return finish(s, t, lat, false);
}

extern "C" int einspline_server_wait(einspline_server *s, uint64_t t, double *lat) {
  if (!s) return esk_fail(EINSPLINE_EINVAL, "NULL server");
  int rc = finish(s, t, lat, true);
  return rc == 1 ? EINSPLINE_OK : rc;
}

extern "C" int einspline_server_drain(einspline_server *s) {
  if (!s) return esk_fail(EINSPLINE_EINVAL, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] {
    if (s->worker_error) return true;
    if (!s->pending.empty()) return false;
    for (Slot &sl : s->slots)
      if (sl.busy) return false;
    return true;
  });
  return s->worker_error ? s->error_for_caller() : EINSPLINE_OK;
}

extern "C" int einspline_server_get_stats(einspline_server *s, einspline_server_stats *out) {
  if (!s || !out) return esk_fail(EINSPLINE_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->st;
  out->resident_tables = 0;
  for (Table &t : s->tables) out->resident_tables += t.dev != nullptr;
  out->resident_bytes = s->used;
  return EINSPLINE_OK;
}

extern "C" int einspline_server_table_resident(einspline_server *s, int tid) {
  if (!s) return esk_fail(EINSPLINE_EINVAL, "NULL server");
  std::lock_guard<std::mutex> g(s->mu);
  if (tid < 0 || tid >= (int)s->tables.size()) return esk_fail(EINSPLINE_EINVAL, "unknown table");
  return s->tables[tid].dev != nullptr;
}

extern "C" int einspline_server_plan_admission(const einspline_admission_table *tables, int ntables,
                                               int table_id, size_t used, size_t budget, int *victims,
                                               int *nvictims, int *drain, int *ndrain) {
  if (!tables || ntables < 1 || table_id < 0 || table_id >= ntables || !victims || !nvictims || !drain ||
      !ndrain)
    return esk_fail(EINSPLINE_EINVAL, "plan_admission: bad arguments");
  std::vector<esk::AdmitTable> view(ntables);
  for (int i = 0; i < ntables; ++i) {
    if (tables[i].state < 0 || tables[i].state > 2)
      return esk_fail(EINSPLINE_EINVAL, "plan_admission: table %d has state %d", i, tables[i].state);
    view[i] = {tables[i].bytes, tables[i].state, tables[i].busy != 0, tables[i].pending, tables[i].last_use};
  }
  std::vector<int> v, d;
  const int rc = esk::plan_admission(view, table_id, used, budget, v, d);
  *nvictims = (int)v.size();
  *ndrain = (int)d.size();
  std::copy(v.begin(), v.end(), victims);
  std::copy(d.begin(), d.end(), drain);
  if (rc == esk::ADMIT_NEVER)
    return esk_fail(EINSPLINE_EBUDGET, "table of %zu bytes exceeds the budget of %zu bytes", tables[table_id].bytes,
                    budget);
  return rc == esk::ADMIT_RESIDENT ? EINSPLINE_ADMIT_RESIDENT
         : rc == esk::ADMIT_UPLOAD ? EINSPLINE_ADMIT_UPLOAD : EINSPLINE_ADMIT_DEFER;
}

// ---------------------------------------------------------------------------
// native load generator (benchmark helper)

extern "C" int einspline_server_run_load(einspline_server *s, const einspline_load *ld,
                                         double *latencies, size_t cap, size_t *nreq,
                                         double *wall_s) {
  if (!s || !ld || ld->ntables < 1 || !ld->table_ids) return esk_fail(EINSPLINE_EINVAL, "bad load");
  std::mt19937_64 rng(ld->seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::exponential_distribution<double> gap(ld->rate_hz > 0 ? ld->rate_hz : 1.0);
  const int maxp = ld->max_points > 0 ? ld->max_points : 2048;
  // Positions come from a pool of uniform random points drawn up front (a
  // request takes a run of them at a random offset), so that generating the
  // load costs little next to serving it.
  const size_t pool_pts = std::max<size_t>(1 << 16, 8 * (size_t)maxp);
  std::vector<double> pool(3 * pool_pts);
  for (size_t i = 0; i < pool.size(); ++i) pool[i] = U(rng) * ld->box[i % 3];
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->lat_log.clear();
  }
  const auto t_start = clk::now();
  double t_next = 0.0;
  for (;;) {
    if (ld->rate_hz > 0) t_next += gap(rng);
    if (t_next >= ld->duration_s) break;
    if (ld->rate_hz > 0) {
      while (std::chrono::duration<double>(clk::now() - t_start).count() < t_next) {
      }
    } else if (std::chrono::duration<double>(clk::now() - t_start).count() >= ld->duration_s) {
      break;
    }
    // closed loop at saturation: bound the queue
    if (ld->rate_hz <= 0) {
      for (;;) {
        size_t q;
        {
          std::lock_guard<std::mutex> g(s->mu);
          q = s->pending.size();
        }
        if (q < (size_t)s->cfg.max_batch_requests * s->slots.size()) break;
        std::this_thread::yield();
      }
    }
    const int k = (int)(U(rng) * ld->ntables) % ld->ntables;
    // log-uniform point count on [1, maxp]
    const int np = std::min(maxp, std::max(1, (int)std::exp(U(rng) * std::log((double)maxp + 1.0))));
    const size_t at = (size_t)(U(rng) * (double)(pool_pts - (size_t)np + 1)) % (pool_pts - (size_t)np + 1);
    int rc = submit_impl(s, ld->table_ids[k], pool.data() + 3 * at, np, ld->vals_d[k], ld->grads_d[k],
                         ld->hess_d[k], nullptr);
    if (rc) return rc;
  }
  int rc = einspline_server_drain(s);
  const double wall = std::chrono::duration<double>(clk::now() - t_start).count();
  std::lock_guard<std::mutex> g(s->mu);
  const size_t m = std::min(cap, s->lat_log.size());
  if (latencies) std::copy(s->lat_log.begin(), s->lat_log.begin() + m, latencies);
  if (nreq) *nreq = s->lat_log.size();
  if (wall_s) *wall_s = wall;
  return rc;
}
