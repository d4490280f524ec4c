// runtime.hip -- device tables, batched entry points, the single-point
// path and a minimal device runtime, all behind the C ABI declared in
// include/multi_bspline_hip.h and include/multi_bspline_eval_*.h.
#include <hip/hip_runtime.h>

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <thread>
#include <mutex>
#include <tuple>
#include <vector>

#include "multi_bspline.h"
#include "einspline_private.h"
#include "launch.hpp"
#include "runtime_internal.hpp"
#include "convert.hpp"
#include "mirror_cache.hpp"

using esk::Launch;
using esk::Params;

// ---------------------------------------------------------------------------
// errors

static thread_local char g_err[512] = "no error";

int esk_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define fail esk_fail

int esk_hip_fail(hipError_t e, const char *what) {
  return fail(EINSPLINE_EHIP, "%s: %s", what, hipGetErrorString(e));
}
#define hip_fail esk_hip_fail

extern "C" const char *einspline_hip_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------
// storage conversion (device side, round to nearest even)

static size_t storage_bytes(einspline_storage s) {
  switch (s) {
    case EINSPLINE_STORE_F32: return 4;
    case EINSPLINE_STORE_F64: return 8;
    case EINSPLINE_STORE_BF16:
    case EINSPLINE_STORE_F16: return 2;
  }
  return 0;
}

namespace {

// Copy `nrows` rows of `nreal` source reals (float or double) into table
// rows of `row` storage elements starting at row r0, zero-padding each row.
template <class Src>
__global__ void __launch_bounds__(256) convert_rows(const Src *__restrict__ src, void *table,
                                                   size_t r0, size_t nrows, int nreal, int row,
                                                   int storage, int *ovf_flag) {
  const size_t total = nrows * (size_t)row;
  int ovf = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / row, q = i - r * row;
    const double v = q < (size_t)nreal ? (double)src[r * nreal + q] : 0.0;
    const size_t o = (r0 + r) * row + q;
    switch (storage) {
      case EINSPLINE_STORE_F32: ((float *)table)[o] = (float)v; break;
      case EINSPLINE_STORE_F64: ((double *)table)[o] = v; break;
      case EINSPLINE_STORE_BF16:
        ((uint16_t *)table)[o] = esk::f32_to_bf16(sizeof(Src) == 8 ? esk::f64_to_f32_odd(v) : (float)v, ovf);
        break;
      case EINSPLINE_STORE_F16:
        ((uint16_t *)table)[o] = esk::f32_to_f16(sizeof(Src) == 8 ? esk::f64_to_f32_odd(v) : (float)v, ovf);
        break;
    }
  }
  if (ovf) atomicOr(ovf_flag, 1);
}

struct TableExt {
  multi_UBspline_3d_hip pub;   // must stay first
  int lane_bytes;              // tile: bytes per lane per row (0 = default)
  int block;                   // tile: work-items per block (0 = default)
  int store_policy;            // esk::STORE_* for result stores
  int point_order;             // EINSPLINE_ORDER_*
  int kernel;                  // EINSPLINE_KERNEL_* (16-bit tables, vgh)
  EskOrderWs *order_ws;        // sort workspace of the locality order (lazily made)
};

// Batches at least this large use the locality order when the table asks
// for it; smaller ones gain nothing from sorting.
const int kOrderMinPoints = 1024;

// host memcpy split over a few threads (large chunks are memory-bound)
void par_memcpy(void *dst, const void *src, size_t n) {
  const size_t per = (size_t)8 << 20;
  const int nt = (int)std::min<size_t>(8, (n + per - 1) / per);
  if (nt <= 1) { memcpy(dst, src, n); return; }
  std::vector<std::thread> th;
  const size_t piece = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const size_t a = t * piece, b = std::min(n, a + piece);
    if (a < b) th.emplace_back([=] { memcpy((char *)dst + a, (const char *)src + a, b - a); });
  }
  for (auto &x : th) x.join();
}

}  // namespace

// Upload `src` (host table: Nx*Ny*Nz contiguous rows of `nreal` reals,
// float if src_float else double) into a new device table of storage `st`.
// Rows go through two pinned staging buffers: the host copy of one chunk
// overlaps the transfer and on-device conversion of the previous one.
static multi_UBspline_3d_hip *upload(const void *src, int src_float, size_t src_row,
                                     int Nx, int Ny, int Nz, int nsplines, int ncomp,
                                     const Ugrid &xg, const Ugrid &yg, const Ugrid &zg,
                                     type_code tcode, einspline_storage st) {
  const size_t eb = storage_bytes(st);
  if (!eb) { fail(EINSPLINE_EINVAL, "unknown storage type %d", (int)st); return nullptr; }
  const size_t nreal = (size_t)nsplines * ncomp;
  if (src_row != nreal) { fail(EINSPLINE_EINVAL, "host rows must be contiguous"); return nullptr; }
  // pad rows to 32 bytes so every supported lane width divides them
  const size_t per32 = 32 / eb;
  const size_t row = (nreal + per32 - 1) / per32 * per32;
  const size_t nrows = (size_t)Nx * Ny * Nz;
  const size_t bytes = nrows * row * eb;
  const size_t sb = src_float ? 4 : 8;

  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) { hip_fail(e, "hipGetDevice"); return nullptr; }
  const size_t chunk_bytes = (size_t)64 << 20;
  const size_t rows_per = std::max<size_t>(1, std::min(nrows, chunk_bytes / (nreal * sb)));
  const size_t stage = rows_per * nreal * sb;
  void *d = nullptr, *dstage[2] = {nullptr, nullptr}, *hstage[2] = {nullptr, nullptr};
  int *dflag = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = EINSPLINE_OK;
  e = hipMalloc(&d, bytes);
  if (e != hipSuccess) {
    fail(EINSPLINE_ENOMEM, "hipMalloc of %zu bytes for a spline table: %s", bytes,
         hipGetErrorString(e));
    return nullptr;
  }
  if ((e = hipMalloc(&dstage[0], stage)) != hipSuccess || (e = hipMalloc(&dstage[1], stage)) != hipSuccess ||
      (e = hipHostMalloc(&hstage[0], stage, hipHostMallocDefault)) != hipSuccess ||
      (e = hipHostMalloc(&hstage[1], stage, hipHostMallocDefault)) != hipSuccess ||
      (e = hipMalloc((void **)&dflag, sizeof(int))) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming)) != hipSuccess ||
      (e = hipMemsetAsync(dflag, 0, sizeof(int), s)) != hipSuccess)
    rc = fail(EINSPLINE_ENOMEM, "upload staging: %s", hipGetErrorString(e));
  int k = 0;
  for (size_t r0 = 0; rc == EINSPLINE_OK && r0 < nrows; r0 += rows_per, ++k) {
    const int b = k & 1;
    const size_t nr = std::min(rows_per, nrows - r0);
    if (k >= 2 && (e = hipEventSynchronize(ev[b])) != hipSuccess) { rc = hip_fail(e, "upload"); break; }
    par_memcpy(hstage[b], (const char *)src + r0 * nreal * sb, nr * nreal * sb);
    e = hipMemcpyAsync(dstage[b], hstage[b], nr * nreal * sb, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) { rc = hip_fail(e, "upload copy"); break; }
    const size_t total = nr * row;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 4096);
    if (src_float)
      hipLaunchKernelGGL(convert_rows<float>, dim3(blocks), dim3(256), 0, s, (const float *)dstage[b], d,
                         r0, nr, (int)nreal, (int)row, (int)st, dflag);
    else
      hipLaunchKernelGGL(convert_rows<double>, dim3(blocks), dim3(256), 0, s, (const double *)dstage[b], d,
                         r0, nr, (int)nreal, (int)row, (int)st, dflag);
    if ((e = hipGetLastError()) != hipSuccess || (e = hipEventRecord(ev[b], s)) != hipSuccess) {
      rc = hip_fail(e, "upload convert");
      break;
    }
  }
  int ovf = 0;
  if (s && (e = hipStreamSynchronize(s)) != hipSuccess && rc == EINSPLINE_OK) rc = hip_fail(e, "upload");
  if (rc == EINSPLINE_OK && (e = hipMemcpy(&ovf, dflag, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
    rc = hip_fail(e, "upload flag");
  for (int b = 0; b < 2; ++b) {
    if (dstage[b]) (void)hipFree(dstage[b]);
    if (hstage[b]) (void)hipHostFree(hstage[b]);
    if (ev[b]) (void)hipEventDestroy(ev[b]);
  }
  if (dflag) (void)hipFree(dflag);
  if (s) (void)hipStreamDestroy(s);
  if (rc == EINSPLINE_OK && ovf)
    rc = fail(EINSPLINE_ERANGE, "coefficient outside the range of the storage type");
  if (rc != EINSPLINE_OK) {
    (void)hipFree(d);
    return nullptr;
  }
  TableExt *t = (TableExt *)calloc(1, sizeof(TableExt));
  if (!t) { (void)hipFree(d); fail(EINSPLINE_ENOMEM, "out of host memory"); return nullptr; }
  multi_UBspline_3d_hip *p = &t->pub;
  p->coefs = d;
  p->z_stride = (intptr_t)row;
  p->y_stride = (intptr_t)Nz * row;
  p->x_stride = (intptr_t)Ny * Nz * row;
  p->x_grid = xg; p->y_grid = yg; p->z_grid = zg;
  p->Nx = Nx; p->Ny = Ny; p->Nz = Nz;
  p->num_splines = nsplines;
  p->row_reals = (int)row;
  p->tcode = tcode;
  p->storage = st;
  p->bytes = bytes;
  p->device = dev;
  p->dim = 3;
  t->lane_bytes = 0;
  t->block = 0;
  t->store_policy = esk::STORE_STAGED_NT;
  return p;
}

// A device table around existing device memory, with the geometry of `meta`
// (used by the serving runtime to re-admit a table from its host copy).
multi_UBspline_3d_hip *esk_table_wrap(const multi_UBspline_3d_hip *meta, void *dcoefs) {
  TableExt *t = (TableExt *)calloc(1, sizeof(TableExt));
  if (!t) { fail(EINSPLINE_ENOMEM, "out of host memory"); return nullptr; }
  t->pub = *meta;
  t->pub.coefs = dcoefs;
  return &t->pub;
}

template <class H>
static void dims_of(const H *s, int &Nx, int &Ny, int &Nz) {
  // z_stride == num_splines for tables from create_multi_UBspline_3d_*
  Nz = (int)(s->y_stride / s->z_stride);
  Ny = (int)(s->x_stride / s->y_stride);
  const einspline_host3d *h = einspline_host_of((void *)s);
  if (h) Nx = h->Nx;
  else Nx = s->x_grid.num + ((s->xBC.lCode == PERIODIC || s->xBC.lCode == ANTIPERIODIC) ? 3 : 2);
}

#define UPLOAD_FROM(S, IS_FLOAT, NCOMP, TC, ST)                                       \
  do {                                                                                \
    if (!S || !S->coefs) { fail(EINSPLINE_EINVAL, "NULL spline"); return nullptr; }   \
    const einspline_host3d *hh_ = einspline_host_of((void *)S);                        \
    if (hh_ && hh_->dim != 3) { fail(EINSPLINE_EINVAL, "spline is not 3-D"); return nullptr; } \
    int Nx, Ny, Nz;                                                                   \
    dims_of(S, Nx, Ny, Nz);                                                           \
    return upload(S->coefs, IS_FLOAT, (size_t)S->z_stride * NCOMP, Nx, Ny, Nz,        \
                  S->num_splines, NCOMP, S->x_grid, S->y_grid, S->z_grid, TC, ST);    \
  } while (0)

extern "C" multi_UBspline_3d_s_hip *create_multi_UBspline_3d_s_hip(multi_UBspline_3d_s *s) {
  UPLOAD_FROM(s, 1, 1, SINGLE_REAL, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_3d_s_hip *create_multi_UBspline_3d_s_hip_conv(multi_UBspline_3d_d *s) {
  UPLOAD_FROM(s, 0, 1, SINGLE_REAL, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_3d_s_hip *create_multi_UBspline_3d_s_hip_store(const multi_UBspline_3d_d *s,
                                                                        einspline_storage st) {
  if (st == EINSPLINE_STORE_F64) { fail(EINSPLINE_EINVAL, "use create_multi_UBspline_3d_d_hip for F64"); return nullptr; }
  UPLOAD_FROM(s, 0, 1, SINGLE_REAL, st);
}
extern "C" multi_UBspline_3d_d_hip *create_multi_UBspline_3d_d_hip(multi_UBspline_3d_d *s) {
  UPLOAD_FROM(s, 0, 1, DOUBLE_REAL, EINSPLINE_STORE_F64);
}
extern "C" multi_UBspline_3d_c_hip *create_multi_UBspline_3d_c_hip(multi_UBspline_3d_c *s) {
  UPLOAD_FROM(s, 1, 2, SINGLE_COMPLEX, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_3d_c_hip *create_multi_UBspline_3d_c_hip_conv(multi_UBspline_3d_z *s) {
  UPLOAD_FROM(s, 0, 2, SINGLE_COMPLEX, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_3d_z_hip *create_multi_UBspline_3d_z_hip(multi_UBspline_3d_z *s) {
  UPLOAD_FROM(s, 0, 2, DOUBLE_COMPLEX, EINSPLINE_STORE_F64);
}

// 1-D and 2-D host tables: the same device layout with Ny (and Nz) = 1.
static const Ugrid kNoAxis = {0.0, 1.0, 1, 1.0, 0.0};

template <class H>
static multi_UBspline_3d_hip *upload12(const H *s, int dim, int is_float, int ncomp, type_code tc,
                                       einspline_storage st, const Ugrid &yg, intptr_t row_elems,
                                       int Ny) {
  if (!s || !s->coefs) { fail(EINSPLINE_EINVAL, "NULL spline"); return nullptr; }
  const einspline_host3d *h = einspline_host_of((void *)s);
  if (h && h->dim != dim) { fail(EINSPLINE_EINVAL, "spline has %d dimensions, not %d", h->dim, dim); return nullptr; }
  const int Nx = h ? h->Nx
                   : s->x_grid.num + ((s->xBC.lCode == PERIODIC || s->xBC.lCode == ANTIPERIODIC) ? 3 : 2);
  multi_UBspline_3d_hip *t = upload(s->coefs, is_float, (size_t)row_elems * ncomp, Nx, Ny, 1,
                                    s->num_splines, ncomp, s->x_grid, yg, kNoAxis, tc, st);
  if (t) t->dim = dim;
  return t;
}
#define UPLOAD_1D(S, IS_FLOAT, NCOMP, TC, ST) \
  return upload12(S, 1, IS_FLOAT, NCOMP, TC, ST, kNoAxis, S ? S->x_stride : 0, 1)
#define UPLOAD_2D(S, IS_FLOAT, NCOMP, TC, ST)                                             \
  return upload12(S, 2, IS_FLOAT, NCOMP, TC, ST, S ? S->y_grid : kNoAxis, S ? S->y_stride : 0, \
                  S ? (int)(S->x_stride / S->y_stride) : 0)

extern "C" multi_UBspline_1d_s_hip *create_multi_UBspline_1d_s_hip(multi_UBspline_1d_s *s) {
  UPLOAD_1D(s, 1, 1, SINGLE_REAL, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_1d_s_hip *create_multi_UBspline_1d_s_hip_conv(multi_UBspline_1d_d *s) {
  UPLOAD_1D(s, 0, 1, SINGLE_REAL, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_1d_d_hip *create_multi_UBspline_1d_d_hip(multi_UBspline_1d_d *s) {
  UPLOAD_1D(s, 0, 1, DOUBLE_REAL, EINSPLINE_STORE_F64);
}
extern "C" multi_UBspline_1d_c_hip *create_multi_UBspline_1d_c_hip(multi_UBspline_1d_c *s) {
  UPLOAD_1D(s, 1, 2, SINGLE_COMPLEX, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_1d_c_hip *create_multi_UBspline_1d_c_hip_conv(multi_UBspline_1d_z *s) {
  UPLOAD_1D(s, 0, 2, SINGLE_COMPLEX, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_1d_z_hip *create_multi_UBspline_1d_z_hip(multi_UBspline_1d_z *s) {
  UPLOAD_1D(s, 0, 2, DOUBLE_COMPLEX, EINSPLINE_STORE_F64);
}
extern "C" multi_UBspline_2d_s_hip *create_multi_UBspline_2d_s_hip(multi_UBspline_2d_s *s) {
  UPLOAD_2D(s, 1, 1, SINGLE_REAL, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_2d_s_hip *create_multi_UBspline_2d_s_hip_conv(multi_UBspline_2d_d *s) {
  UPLOAD_2D(s, 0, 1, SINGLE_REAL, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_2d_d_hip *create_multi_UBspline_2d_d_hip(multi_UBspline_2d_d *s) {
  UPLOAD_2D(s, 0, 1, DOUBLE_REAL, EINSPLINE_STORE_F64);
}
extern "C" multi_UBspline_2d_c_hip *create_multi_UBspline_2d_c_hip(multi_UBspline_2d_c *s) {
  UPLOAD_2D(s, 1, 2, SINGLE_COMPLEX, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_2d_c_hip *create_multi_UBspline_2d_c_hip_conv(multi_UBspline_2d_z *s) {
  UPLOAD_2D(s, 0, 2, SINGLE_COMPLEX, EINSPLINE_STORE_F32);
}
extern "C" multi_UBspline_2d_z_hip *create_multi_UBspline_2d_z_hip(multi_UBspline_2d_z *s) {
  UPLOAD_2D(s, 0, 2, DOUBLE_COMPLEX, EINSPLINE_STORE_F64);
}

extern "C" void destroy_multi_UBspline_3d_hip(multi_UBspline_3d_hip *t) {
  if (!t) return;
  esk_order_free(((TableExt *)t)->order_ws);
  if (t->coefs) (void)hipFree(t->coefs);
  free((TableExt *)t);
}

extern "C" void einspline_mirror_release(void *mirror) {
  // the host spline is being destroyed: no call may still use its mirror
  esk::mirror_free(mirror, [](void *m) { destroy_multi_UBspline_3d_hip((multi_UBspline_3d_hip *)m); });
}

extern "C" int einspline_set_tile(multi_UBspline_3d_hip *t, int lane_bytes, int block) {
  if (!t) return fail(EINSPLINE_EINVAL, "NULL table");
  const bool wide = t->storage == EINSPLINE_STORE_F64;
  if (lane_bytes && lane_bytes != 8 && lane_bytes != 16 && lane_bytes != 32)
    return fail(EINSPLINE_EINVAL, "lane_bytes must be 8, 16 or 32");
  if (wide && lane_bytes == 8) return fail(EINSPLINE_EINVAL, "F64 tables need lane_bytes >= 16");
  const bool narrow = t->storage == EINSPLINE_STORE_BF16 || t->storage == EINSPLINE_STORE_F16;
  if (narrow && lane_bytes == 32)
    return fail(EINSPLINE_EINVAL, "16-bit tables take lane_bytes 8 or 16");
  if (block && (block < 64 || block > 256 || block % 64))
    return fail(EINSPLINE_EINVAL, "block_threads must be 64, 128, 192 or 256");
  TableExt *x = (TableExt *)t;
  x->lane_bytes = lane_bytes;
  x->block = block;
  return EINSPLINE_OK;
}

extern "C" int einspline_set_store_policy(multi_UBspline_3d_hip *t, int policy) {
  if (!t) return fail(EINSPLINE_EINVAL, "NULL table");
  if (policy < 0 || policy > 2) return fail(EINSPLINE_EINVAL, "store policy must be 0, 1 or 2");
  ((TableExt *)t)->store_policy = policy;
  return EINSPLINE_OK;
}

extern "C" int einspline_set_point_order(multi_UBspline_3d_hip *t, int order) {
  if (!t) return fail(EINSPLINE_EINVAL, "NULL table");
  if (order < 0 || order > (EINSPLINE_ORDER_LOCALITY | EINSPLINE_ORDER_SLABS))
    return fail(EINSPLINE_EINVAL, "point order must be EINSPLINE_ORDER_GIVEN or _LOCALITY, optionally "
                                  "with EINSPLINE_ORDER_SLABS");
  ((TableExt *)t)->point_order = order;
  return EINSPLINE_OK;
}

extern "C" int einspline_set_kernel(multi_UBspline_3d_hip *t, int kernel) {
  if (!t) return fail(EINSPLINE_EINVAL, "NULL table");
  if (kernel != EINSPLINE_KERNEL_VECTOR && kernel != EINSPLINE_KERNEL_MATRIX)
    return fail(EINSPLINE_EINVAL, "kernel must be EINSPLINE_KERNEL_VECTOR or EINSPLINE_KERNEL_MATRIX");
  ((TableExt *)t)->kernel = kernel;
  return EINSPLINE_OK;
}

extern "C" int einspline_get_kernel(const multi_UBspline_3d_hip *t) {
  if (!t) return fail(EINSPLINE_EINVAL, "NULL table");
  return ((const TableExt *)t)->kernel;
}

extern "C" int einspline_get_tuning(const multi_UBspline_3d_hip *t, int *lane_bytes, int *block,
                                    int *store_policy, int *point_order) {
  if (!t) return fail(EINSPLINE_EINVAL, "NULL table");
  const TableExt *x = (const TableExt *)t;
  if (lane_bytes) *lane_bytes = x->lane_bytes ? x->lane_bytes : 16;
  if (block) *block = x->block ? x->block : 256;
  if (store_policy) *store_policy = x->store_policy;
  if (point_order) *point_order = x->point_order;
  return EINSPLINE_OK;
}

// ---------------------------------------------------------------------------
// autotuner

namespace {
std::mutex g_tune_mu;
// The best tile depends on the table's extent as well as its row: a table
// that fits the Infinity Cache favours other tiles than one that streams
// from HBM.
struct TuneKey {
  int storage, tcode, mode, row_bytes, sample, nx, ny, nz;
  bool operator<(const TuneKey &o) const {
    return std::tie(storage, tcode, mode, row_bytes, sample, nx, ny, nz) <
           std::tie(o.storage, o.tcode, o.mode, o.row_bytes, o.sample, o.nx, o.ny, o.nz);
  }
};
struct TuneVal {
  int lane_bytes, block, store, order, kernel;
};
std::map<TuneKey, TuneVal> g_tune;
bool g_tune_loaded = false;

void tune_load() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char *f = getenv("EINSPLINE_TUNE_CACHE");
  if (!f) return;
  FILE *fp = fopen(f, "r");
  if (!fp) return;
  // one entry per line, 13 integers (8 of the key, then the choice); lines
  // of any other shape (older files with fewer fields) are skipped
  char line[256];
  while (fgets(line, sizeof line, fp)) {
    TuneKey k;
    TuneVal v;
    char extra;
    if (sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d %d %d %c", &k.storage, &k.tcode, &k.mode,
               &k.row_bytes, &k.sample, &k.nx, &k.ny, &k.nz, &v.lane_bytes, &v.block, &v.store, &v.order,
               &v.kernel, &extra) == 13)
      g_tune[k] = v;
  }
  fclose(fp);
}

void tune_save() {
  const char *f = getenv("EINSPLINE_TUNE_CACHE");
  if (!f) return;
  FILE *fp = fopen(f, "w");
  if (!fp) return;
  for (auto &kv : g_tune)
    fprintf(fp, "%d %d %d %d %d %d %d %d %d %d %d %d %d\n", kv.first.storage, kv.first.tcode, kv.first.mode,
            kv.first.row_bytes, kv.first.sample, kv.first.nx, kv.first.ny, kv.first.nz, kv.second.lane_bytes,
            kv.second.block, kv.second.store, kv.second.order, kv.second.kernel);
  fclose(fp);
}
}  // namespace

static int run_strided(const multi_UBspline_3d_hip *t, int mode, const double *pos, int num,
                       void *v, void *g, void *h, int packed, void *stream, void *gh = nullptr,
                       bool noalloc = false);
static int result_reals(const multi_UBspline_3d_hip *t);

// Time every candidate (tile x store policy x point order, and the
// matrix-core kernel where it applies) of `mode` on `sample` host positions,
// apply the fastest to the table and return it.  Synchronises the device.
static int tune_on(multi_UBspline_3d_hip *t, int mode, const double *hpos, int sample, TuneVal &won,
                   float &best_out) {
  const int eb = (int)storage_bytes(t->storage);
  const size_t nreal = (size_t)result_reals(t);
  const size_t es = t->storage == EINSPLINE_STORE_F64 ? 8 : 4;
  void *dp = nullptr, *dv = nullptr, *dg = nullptr, *dh = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t st = nullptr;
  int rc = EINSPLINE_OK;
  hipError_t e;
  if ((e = hipMalloc(&dp, 3 * (size_t)sample * 8)) != hipSuccess ||
      (e = hipMalloc(&dv, sample * nreal * es)) != hipSuccess ||
      (e = hipMalloc(&dg, 3 * sample * nreal * es)) != hipSuccess ||
      (e = hipMalloc(&dh, 9 * sample * nreal * es)) != hipSuccess ||
      (e = hipMemcpy(dp, hpos, 3 * (size_t)sample * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
    rc = hip_fail(e, "autotune setup");
  TableExt *ext = (TableExt *)t;
  const int old_lb = ext->lane_bytes, old_blk = ext->block, old_store = ext->store_policy;
  const int old_order = ext->point_order, old_kernel = ext->kernel;
  int best_lb = 16, best_blk = 256, best_nt = esk::STORE_STAGED_NT, best_order = EINSPLINE_ORDER_GIVEN;
  int best_kernel = EINSPLINE_KERNEL_VECTOR;
  ext->kernel = EINSPLINE_KERNEL_VECTOR;
  float best = -1.0f, default_ms = -1.0f;
  const int lbs_f64[] = {16, 32}, lbs_f32[] = {8, 16, 32}, lbs_16[] = {8, 16};
  const bool narrow = t->storage == EINSPLINE_STORE_BF16 || t->storage == EINSPLINE_STORE_F16;
  const int *lbs = t->storage == EINSPLINE_STORE_F64 ? lbs_f64 : narrow ? lbs_16 : lbs_f32;
  const int nlb = narrow || t->storage == EINSPLINE_STORE_F64 ? 2 : 3;
  const int blks[] = {64, 128, 256};
  // point orders: given, given in orbital slabs, then (batches large enough
  // to sort) locality, locality in orbital slabs
  const int orders[4] = {EINSPLINE_ORDER_GIVEN, EINSPLINE_ORDER_SLABS, EINSPLINE_ORDER_LOCALITY,
                         EINSPLINE_ORDER_LOCALITY | EINSPLINE_ORDER_SLABS};
  const int norders = sample >= kOrderMinPoints ? 4 : 2;
  for (int oi = 0; oi < norders; ++oi)
  for (int nt = 0; nt <= 2; ++nt)
  for (int a = 0; rc == EINSPLINE_OK && a < nlb; ++a) {
    if ((t->row_reals * eb) % lbs[a]) continue;
    for (int b = 0; rc == EINSPLINE_OK && b < 3; ++b) {
      ext->lane_bytes = lbs[a];
      ext->block = blks[b];
      ext->store_policy = nt;
      const int order = orders[oi];
      ext->point_order = order;
      auto launch = [&]() {
        return run_strided(t, mode, (const double *)dp, sample, dv, dg, dh, 0, st);
      };
      if ((rc = launch()) != EINSPLINE_OK) break;       // warm-up
      // best of 3 trials of 5 launches: launch-to-launch noise is a few percent
      float ms = -1.0f;
      for (int trial = 0; trial < 3 && rc == EINSPLINE_OK; ++trial) {
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < 5 && rc == EINSPLINE_OK; ++r) rc = launch();
        (void)hipEventRecord(e1, st);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) { rc = hip_fail(e, "autotune"); break; }
        float t_ms = 0.0f;
        (void)hipEventElapsedTime(&t_ms, e0, e1);
        t_ms /= 5.0f;
        if (ms < 0.0f || t_ms < ms) ms = t_ms;
      }
      if (rc != EINSPLINE_OK) break;
      if (lbs[a] == 16 && blks[b] == 256 && nt == esk::STORE_STAGED_NT && order == 0) default_ms = ms;
      if (best < 0.0f || ms < best) {
        best = ms; best_lb = lbs[a]; best_blk = blks[b]; best_nt = nt; best_order = order;
      }
    }
  }
  // real BF16, FP16 and FP32 tables, vgh: the matrix-core kernel in either
  // order, streaming or ordinary stores (it has no tile)
  const bool cplx = t->tcode == SINGLE_COMPLEX || t->tcode == DOUBLE_COMPLEX;
  if (t->storage != EINSPLINE_STORE_F64 && !cplx && mode == esk::MODE_VGH) {
    for (int cand = 0; rc == EINSPLINE_OK && cand < 2 * norders; ++cand) {
      const int order = orders[cand >> 1], nt = cand & 1 ? esk::STORE_STAGED : esk::STORE_STAGED_NT;
      ext->kernel = EINSPLINE_KERNEL_MATRIX;
      ext->point_order = order;
      ext->store_policy = nt;
      auto launch = [&]() {
        return run_strided(t, mode, (const double *)dp, sample, dv, dg, dh, 0, st);
      };
      if ((rc = launch()) != EINSPLINE_OK) break;
      float ms = -1.0f;
      for (int trial = 0; trial < 3 && rc == EINSPLINE_OK; ++trial) {
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < 5 && rc == EINSPLINE_OK; ++r) rc = launch();
        (void)hipEventRecord(e1, st);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) { rc = hip_fail(e, "autotune"); break; }
        float t_ms = 0.0f;
        (void)hipEventElapsedTime(&t_ms, e0, e1);
        t_ms /= 5.0f;
        if (ms < 0.0f || t_ms < ms) ms = t_ms;
      }
      if (rc == EINSPLINE_OK && (best < 0.0f || ms < best)) {
        best = ms; best_kernel = EINSPLINE_KERNEL_MATRIX; best_order = order;
        best_lb = 16; best_blk = 256; best_nt = nt;
      }
    }
  }
  ext->lane_bytes = old_lb;
  ext->block = old_blk;
  ext->store_policy = old_store;
  ext->point_order = old_order;
  ext->kernel = old_kernel;
  // keep the default (16 B x 256, streaming stores, given order) unless the
  // winner is at least 3% faster
  if (default_ms > 0.0f && best > 0.97f * default_ms) {
    best = default_ms;
    best_lb = 16;
    best_blk = 256;
    best_nt = esk::STORE_STAGED_NT;
    best_order = EINSPLINE_ORDER_GIVEN;
    best_kernel = EINSPLINE_KERNEL_VECTOR;
  }
  if (dp) (void)hipFree(dp);
  if (dv) (void)hipFree(dv);
  if (dg) (void)hipFree(dg);
  if (dh) (void)hipFree(dh);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (rc != EINSPLINE_OK) return rc;
  einspline_set_tile(t, best_lb, best_blk);
  einspline_set_store_policy(t, best_nt);
  einspline_set_point_order(t, best_order);
  einspline_set_kernel(t, best_kernel);
  won = {best_lb, best_blk, best_nt, best_order, best_kernel};
  best_out = best;
  return EINSPLINE_OK;
}

extern "C" int einspline_autotune(multi_UBspline_3d_hip *t, int mode, int sample, int *lb_out,
                                  int *blk_out, int *nt_out, float *ms_out) {
  // nt_out receives the chosen store policy
  if (!t || mode < 0 || mode > 3) return fail(EINSPLINE_EINVAL, "bad autotune arguments");
  if (t->dim != 3) return fail(EINSPLINE_EINVAL, "the autotuner takes 3-D tables");
  if (sample <= 0) sample = 8192;
  const int eb = (int)storage_bytes(t->storage);
  const TuneKey key{(int)t->storage, (int)t->tcode, mode, t->row_reals * eb, sample, t->Nx, t->Ny, t->Nz};
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    tune_load();
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      einspline_set_tile(t, it->second.lane_bytes, it->second.block);
      einspline_set_store_policy(t, it->second.store);
      einspline_set_point_order(t, it->second.order);
      einspline_set_kernel(t, it->second.kernel);
      if (lb_out) *lb_out = it->second.lane_bytes;
      if (blk_out) *blk_out = it->second.block;
      if (nt_out) *nt_out = it->second.store;
      if (ms_out) *ms_out = -1.0f;
      return EINSPLINE_OK;
    }
  }
  // sample points: deterministic, uniform in the box
  std::vector<double> pos(3 * (size_t)sample);
  uint64_t x = 0x9e3779b97f4a7c15ull;
  const Ugrid *gr[3] = {&t->x_grid, &t->y_grid, &t->z_grid};
  for (size_t i = 0; i < pos.size(); ++i) {
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    const double u = (double)(x >> 11) * (1.0 / 9007199254740992.0);
    const Ugrid *g = gr[i % 3];
    pos[i] = g->start + u * (g->end - g->start) * (1.0 - 1e-12);
  }
  TuneVal won;
  float best = -1.0f;
  const int rc = tune_on(t, mode, pos.data(), sample, won, best);
  if (rc != EINSPLINE_OK) return rc;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[key] = won;
    tune_save();
  }
  if (lb_out) *lb_out = won.lane_bytes;
  if (blk_out) *blk_out = won.block;
  if (nt_out) *nt_out = won.store;
  if (ms_out) *ms_out = best;
  return EINSPLINE_OK;
}

extern "C" int einspline_autotune_points(multi_UBspline_3d_hip *t, int mode, const double *pos, int npoints,
                                         int *lb_out, int *blk_out, int *nt_out, float *ms_out) {
  if (!t || mode < 0 || mode > 3 || !pos || npoints <= 0)
    return fail(EINSPLINE_EINVAL, "bad autotune_points arguments");
  if (t->dim != 3) return fail(EINSPLINE_EINVAL, "the autotuner takes 3-D tables");
  // the first kTunePointsMax positions stand for the rest (bounds the
  // tuner's output buffers)
  constexpr int kTunePointsMax = 65536;
  TuneVal won;
  float best = -1.0f;
  const int rc = tune_on(t, mode, pos, npoints < kTunePointsMax ? npoints : kTunePointsMax, won, best);
  if (rc != EINSPLINE_OK) return rc;
  if (lb_out) *lb_out = won.lane_bytes;
  if (blk_out) *blk_out = won.block;
  if (nt_out) *nt_out = won.store;
  if (ms_out) *ms_out = best;
  return EINSPLINE_OK;
}

// ---------------------------------------------------------------------------
// launching

static int result_reals(const multi_UBspline_3d_hip *t) {
  return t->num_splines * ((t->tcode == SINGLE_COMPLEX || t->tcode == DOUBLE_COMPLEX) ? 2 : 1);
}

// Fill the table-dependent kernel arguments and the tile.
int esk_prepare(const multi_UBspline_3d_hip *t, int num, Launch &L, Params &P) {
  if (!t || !t->coefs) return fail(EINSPLINE_EINVAL, "NULL device table");
  if (num < 0) return fail(EINSPLINE_EINVAL, "negative point count");
  const TableExt *x = (const TableExt *)t;
  const size_t eb = storage_bytes(t->storage);
  int lb = x->lane_bytes ? x->lane_bytes : 16;
  const int row_bytes = t->row_reals * (int)eb;
  // narrow rows: use the widest lane that still gives every lane work
  while (lb > 8 && row_bytes % lb) lb /= 2;
  if (t->storage == EINSPLINE_STORE_F64 && lb < 16) lb = 16;
  L.lane_bytes = lb;
  L.block = x->block ? x->block : 256;
  L.kernel = x->kernel;
  // the 16-bit matrix kernel addresses a stencil row as the cell plus a
  // 32-bit lane offset (one x plane, two y rows and an orbital offset)
  if (L.kernel == EINSPLINE_KERNEL_MATRIX && eb == 2 &&
      ((long long)t->x_stride + 2ll * t->y_stride + t->z_stride) * 2 >= (1ll << 31))
    L.kernel = EINSPLINE_KERNEL_VECTOR;
  L.cplx = t->tcode == SINGLE_COMPLEX || t->tcode == DOUBLE_COMPLEX;
  memset(&P, 0, sizeof P);
  P.coefs = t->coefs;
  P.xs = t->x_stride; P.ys = t->y_stride; P.zs = t->z_stride;
  P.x0 = t->x_grid.start; P.y0 = t->y_grid.start; P.z0 = t->z_grid.start;
  P.xi = t->x_grid.delta_inv; P.yi = t->y_grid.delta_inv; P.zi = t->z_grid.delta_inv;
  P.nx4 = t->Nx - 4; P.ny4 = t->Ny - 4; P.nz4 = t->Nz - 4;
  P.chunks = row_bytes / lb;
  P.nreal = result_reals(t);
  P.npoints = num;
  P.store = x->store_policy;
  if ((long long)num * P.chunks >= (1ll << 31))
    return fail(EINSPLINE_EINVAL, "batch too large: points * lanes per point must be < 2^31");
  return EINSPLINE_OK;
}

static hipError_t dispatch(const multi_UBspline_3d_hip *t, const Launch &L, const Params &P,
                           hipStream_t s) {
  if (P.npoints == 0) return hipSuccess;
  switch (t->storage) {
    case EINSPLINE_STORE_F32: return esk::launch_f32(L, P, s);
    case EINSPLINE_STORE_F64: return esk::launch_f64(L, P, s);
    case EINSPLINE_STORE_BF16: return esk::launch_bf16(L, P, s);
    case EINSPLINE_STORE_F16: return esk::launch_f16(L, P, s);
  }
  return hipErrorInvalidValue;
}

// Sort the batch for the locality order when the table asks for it.  Not
// under stream capture: the sort workspace is shared by the table's calls.
static int order_begin(const multi_UBspline_3d_hip *t, Params &P, const void *pos, bool pos_f32,
                       int num, hipStream_t s, EskOrderWs **ws) {
  *ws = nullptr;
  TableExt *x = (TableExt *)t;
  if (!(x->point_order & EINSPLINE_ORDER_LOCALITY) || num < kOrderMinPoints) return EINSPLINE_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return EINSPLINE_OK;
  int rc = esk_order_begin(&x->order_ws, P, pos, pos_f32, num, s, &P.perm);
  if (rc) return rc;
  P.xcd_runs = 1;
  *ws = x->order_ws;
  return EINSPLINE_OK;
}

// 1-D and 2-D tables (FP32 / FP64 storage): strided or pointer-array
// outputs, 16-byte lanes, points in the given order.  pos holds dim
// coordinates per point.  row_stride (reals) places the Laplacian row after
// the gradient row when the pointer form has no Laplacian array.
static int run12(const multi_UBspline_3d_hip *t, int mode, int api, const void *pos, int num,
                 void *v, void *g, void *h, long row_stride, hipStream_t stream) {
  if (t->storage != EINSPLINE_STORE_F32 && t->storage != EINSPLINE_STORE_F64)
    return fail(EINSPLINE_EINVAL, "1-D and 2-D tables are stored as F32 or F64");
  if (mode < esk::MODE_V || mode > esk::MODE_VGH)
    return fail(EINSPLINE_EINVAL, "mode not available for %d-D tables", t->dim);
  Launch L;
  Params P;
  int rc = esk_prepare(t, num, L, P);
  if (rc) return rc;
  if (num == 0) return EINSPLINE_OK;
  if (!pos || !v || (mode >= esk::MODE_VG && !g) ||
      (mode >= esk::MODE_VGL && !h && !(api != esk::API_STRIDED && row_stride > 0)))
    return fail(EINSPLINE_EINVAL, "NULL position or output pointer");
  const int eb = (int)storage_bytes(t->storage);
  L.lane_bytes = 16;
  P.chunks = t->row_reals * eb / 16;
  if ((long long)num * P.chunks >= (1ll << 31))
    return fail(EINSPLINE_EINVAL, "batch too large: points * lanes per point must be < 2^31");
  L.mode = mode;
  L.api = api;
  P.pos = pos;
  P.out_v = v; P.out_g = g; P.out_h = h;
  P.ld_v = P.nreal;
  P.ld_g = (long)t->dim * P.nreal;
  P.ld_h = (mode == esk::MODE_VGL || t->dim == 1) ? (long)P.nreal : 4l * P.nreal;
  P.row_stride = row_stride;
  hipError_t e = esk::launch12(L, P, t->dim, t->storage == EINSPLINE_STORE_F64, stream);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return EINSPLINE_OK;
}

static int run_strided(const multi_UBspline_3d_hip *t, int mode, const double *pos, int num,
                       void *v, void *g, void *h, int packed, void *stream, void *gh, bool noalloc) {
  if (t && t->dim != 3) {
    if (packed) return fail(EINSPLINE_EINVAL, "packed Hessians are for 3-D tables");
    return run12(t, mode, esk::API_STRIDED, pos, num, v, g, h, 0, (hipStream_t)stream);
  }
  Launch L;
  Params P;
  int rc = esk_prepare(t, num, L, P);
  if (rc) return rc;
  if (num == 0) return EINSPLINE_OK;
  if (!pos || !v || (mode >= esk::MODE_VG && !g) || (mode >= esk::MODE_VGL && !h) ||
      (mode == esk::MODE_VGHGH && !gh))
    return fail(EINSPLINE_EINVAL, "NULL position or output pointer");
  if (mode == esk::MODE_VGHGH) {
    // the vghgh kernels exist for lanes of at most 16 bytes and 4 reals
    const int eb = (int)storage_bytes(t->storage);
    while (L.lane_bytes > 16 || L.lane_bytes / eb > 4) {
      P.chunks *= 2;
      L.lane_bytes /= 2;
    }
    packed = 0;
    P.out_gh = gh;
    P.ld_gh = 27l * P.nreal;
  }
  L.mode = mode;
  L.api = esk::API_STRIDED;
  P.pos = pos;
  P.out_v = v; P.out_g = g; P.out_h = h;
  P.ld_v = P.nreal;
  P.ld_g = 3l * P.nreal;
  P.ld_h = mode == esk::MODE_VGL ? (long)P.nreal : (packed ? 6l : 9l) * P.nreal;
  if ((long long)num * P.chunks >= (1ll << 31))
    return fail(EINSPLINE_EINVAL, "batch too large: points * lanes per point must be < 2^31");
  P.hpacked = packed ? 1 : 0;
  if (((const TableExt *)t)->point_order & EINSPLINE_ORDER_SLABS) {
    // orbital slabs outermost: a wavefront per (point, slab), slab-major
    if (esk::use_mfma(L)) {
      // the matrix kernels' slab width: 64 orbitals for FP32; for 16-bit
      // tables one pair of 128-orbital blocks, which a wavefront stages and
      // stores in one burst
      const int w = t->storage == EINSPLINE_STORE_F32 ? 64 : 256;
      const int end = std::min(t->row_reals, P.nreal);
      P.slabs = (end + w - 1) / w;
    } else if (P.chunks % 64 == 0 && L.block % 64 == 0) {
      P.slabs = P.chunks / 64;
    }
    if (P.slabs < 2) P.slabs = 0;
  }
  EskOrderWs *ws = nullptr;
  if (noalloc) {
    // the locality order sorts into the file-scope scratch array instead of
    // the table's workspace (also under stream capture: only kernels are enqueued)
    if ((((const TableExt *)t)->point_order & EINSPLINE_ORDER_LOCALITY) && num >= kOrderMinPoints) {
      if ((rc = esk_order_begin_scratch(P, pos, false, num, (hipStream_t)stream, &P.perm))) return rc;
      P.xcd_runs = 1;
    }
  } else if ((rc = order_begin(t, P, pos, false, num, (hipStream_t)stream, &ws))) {
    return rc;
  }
  hipError_t e = dispatch(t, L, P, (hipStream_t)stream);
  if (ws) esk_order_end(ws, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return EINSPLINE_OK;
}

extern "C" int einspline_eval_v_batch(const multi_UBspline_3d_hip *t, const double *pos, int num,
                                      void *v, void *stream) {
  return run_strided(t, esk::MODE_V, pos, num, v, nullptr, nullptr, 0, stream);
}
extern "C" int einspline_eval_vg_batch(const multi_UBspline_3d_hip *t, const double *pos, int num,
                                       void *v, void *g, void *stream) {
  return run_strided(t, esk::MODE_VG, pos, num, v, g, nullptr, 0, stream);
}
extern "C" int einspline_eval_vgl_batch(const multi_UBspline_3d_hip *t, const double *pos, int num,
                                        void *v, void *g, void *l, void *stream) {
  return run_strided(t, esk::MODE_VGL, pos, num, v, g, l, 0, stream);
}
extern "C" int einspline_eval_vgh_batch(const multi_UBspline_3d_hip *t, const double *pos, int num,
                                        void *v, void *g, void *h, int packed, void *stream) {
  return run_strided(t, esk::MODE_VGH, pos, num, v, g, h, packed, stream);
}
extern "C" int einspline_eval_vgh_batch_noalloc(const multi_UBspline_3d_hip *t, const double *pos, int num,
                                                void *v, void *g, void *h, int packed, void *stream) {
  if (num > EINSPLINE_NOALLOC_MAX_POINTS)
    return fail(EINSPLINE_EINVAL, "einspline_eval_vgh_batch_noalloc: %d points, at most %d", num,
                EINSPLINE_NOALLOC_MAX_POINTS);
  if (t && t->dim != 3) return fail(EINSPLINE_EINVAL, "einspline_eval_vgh_batch_noalloc takes 3-D tables");
  return run_strided(t, esk::MODE_VGH, pos, num, v, g, h, packed, stream, nullptr, true);
}
extern "C" int einspline_eval_vghgh_batch(const multi_UBspline_3d_hip *t, const double *pos, int num,
                                          void *v, void *g, void *h, void *gh, void *stream) {
  return run_strided(t, esk::MODE_VGHGH, pos, num, v, g, h, 0, stream, gh);
}

// CUDA-style: per-point pointer arrays, packed Hessian, synchronous.
static int run_ptrs(const multi_UBspline_3d_hip *t, type_code want, int mode, const void *pos,
                    bool pos_f32, void **v, void **g, void **h, int num,
                    const void *linv = nullptr, const void *sign = nullptr, int row_stride = 0) {
  if (t && t->tcode != want)
    return fail(EINSPLINE_EINVAL, "device table has the wrong type for this entry point");
  if (t && want == SINGLE_REAL && t->storage == EINSPLINE_STORE_F64)
    return fail(EINSPLINE_EINVAL, "F64 storage in a single-precision table");
  if (t && t->dim != 3) return fail(EINSPLINE_EINVAL, "this entry point takes a 3-D table");
  Launch L;
  Params P;
  int rc = esk_prepare(t, num, L, P);
  if (rc) return rc;
  if (num == 0) return EINSPLINE_OK;
  if (!pos || !v || (mode == esk::MODE_VGH && (!g || !h)) || (mode == esk::MODE_QMC && (!g || !linv)))
    return fail(EINSPLINE_EINVAL, "NULL position, Linv or output pointer array");
  if (mode == esk::MODE_QMC && row_stride < t->num_splines)
    return fail(EINSPLINE_EINVAL, "row_stride %d is smaller than num_splines %d", row_stride,
                t->num_splines);
  L.lane_bytes = 16;            // the pointer-array kernels exist for 16-byte lanes
  if (t->row_reals * (int)storage_bytes(t->storage) % 16)
    return fail(EINSPLINE_EINVAL, "row not a multiple of 16 bytes");
  P.chunks = t->row_reals * (int)storage_bytes(t->storage) / 16;
  L.mode = mode;
  L.api = pos_f32 ? esk::API_PTRS_F32POS : esk::API_PTRS_F64POS;
  P.pos = pos;
  P.out_v = v; P.out_g = g; P.out_h = h;
  P.hpacked = 1;
  P.linv = linv;
  P.sign = sign;
  P.row_stride = (long)row_stride * (L.cplx ? 2 : 1);
  EskOrderWs *ws;
  if ((rc = order_begin(t, P, pos, pos_f32, num, nullptr, &ws))) return rc;
  hipError_t e = dispatch(t, L, P, nullptr);
  if (ws) esk_order_end(ws, nullptr);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  return EINSPLINE_OK;
}

extern "C" int eval_multi_multi_UBspline_3d_s_hip(multi_UBspline_3d_s_hip *t, float *pos,
                                                  float *vals[], int num) {
  return run_ptrs(t, SINGLE_REAL, esk::MODE_V, pos, true, (void **)vals, nullptr, nullptr, num);
}
extern "C" int eval_multi_multi_UBspline_3d_s_vgh_hip(multi_UBspline_3d_s_hip *t, float *pos,
                                                      float *vals[], float *grads[],
                                                      float *hess[], int num) {
  return run_ptrs(t, SINGLE_REAL, esk::MODE_VGH, pos, true, (void **)vals, (void **)grads,
                  (void **)hess, num);
}
extern "C" int eval_multi_multi_UBspline_3d_d_hip(multi_UBspline_3d_d_hip *t, double *pos,
                                                  double *vals[], int num) {
  return run_ptrs(t, DOUBLE_REAL, esk::MODE_V, pos, false, (void **)vals, nullptr, nullptr, num);
}
extern "C" int eval_multi_multi_UBspline_3d_d_vgh_hip(multi_UBspline_3d_d_hip *t, double *pos,
                                                      double *vals[], double *grads[],
                                                      double *hess[], int num) {
  return run_ptrs(t, DOUBLE_REAL, esk::MODE_VGH, pos, false, (void **)vals, (void **)grads,
                  (void **)hess, num);
}
extern "C" int eval_multi_multi_UBspline_3d_c_hip(multi_UBspline_3d_c_hip *t, float *pos,
                                                  complex_float *vals[], int num) {
  return run_ptrs(t, SINGLE_COMPLEX, esk::MODE_V, pos, true, (void **)vals, nullptr, nullptr, num);
}
extern "C" int eval_multi_multi_UBspline_3d_c_vgh_hip(multi_UBspline_3d_c_hip *t, float *pos,
                                                      complex_float *vals[],
                                                      complex_float *grads[],
                                                      complex_float *hess[], int num) {
  return run_ptrs(t, SINGLE_COMPLEX, esk::MODE_VGH, pos, true, (void **)vals, (void **)grads,
                  (void **)hess, num);
}
extern "C" int eval_multi_multi_UBspline_3d_z_hip(multi_UBspline_3d_z_hip *t, double *pos,
                                                  complex_double *vals[], int num) {
  return run_ptrs(t, DOUBLE_COMPLEX, esk::MODE_V, pos, false, (void **)vals, nullptr, nullptr, num);
}
extern "C" int eval_multi_multi_UBspline_3d_z_vgh_hip(multi_UBspline_3d_z_hip *t, double *pos,
                                                      complex_double *vals[],
                                                      complex_double *grads[],
                                                      complex_double *hess[], int num) {
  return run_ptrs(t, DOUBLE_COMPLEX, esk::MODE_VGH, pos, false, (void **)vals, (void **)grads,
                  (void **)hess, num);
}

// QMC-shaped evaluation (src/multi_bspline_eval_cuda.h:67-79, 96-99, 112-115,
// 131-134): value (optionally times a per-point sign), the gradient
// transformed by the 3x3 lattice matrix Linv, and the Laplacian
// tr((Linv^T Linv) H), written as rows grad_lapl[i] + d*row_stride (d = 0..2
// gradient, 3 Laplacian; row_stride counts result elements).
extern "C" int eval_multi_multi_UBspline_3d_s_sign_hip(multi_UBspline_3d_s_hip *t, float *pos,
                                                       float *sign, float *vals[], int num) {
  if (!sign) return fail(EINSPLINE_EINVAL, "NULL sign array");
  return run_ptrs(t, SINGLE_REAL, esk::MODE_V, pos, true, (void **)vals, nullptr, nullptr, num,
                  nullptr, sign);
}
extern "C" int eval_multi_multi_UBspline_3d_s_vgl_hip(multi_UBspline_3d_s_hip *t, float *pos,
                                                      float *Linv, float *vals[],
                                                      float *grad_lapl[], int num, int row_stride) {
  return run_ptrs(t, SINGLE_REAL, esk::MODE_QMC, pos, true, (void **)vals, (void **)grad_lapl,
                  nullptr, num, Linv, nullptr, row_stride);
}
extern "C" int eval_multi_multi_UBspline_3d_s_vgl_sign_hip(multi_UBspline_3d_s_hip *t, float *pos,
                                                           float *sign, float *Linv, float *vals[],
                                                           float *grad_lapl[], int num,
                                                           int row_stride) {
  if (!sign) return fail(EINSPLINE_EINVAL, "NULL sign array");
  return run_ptrs(t, SINGLE_REAL, esk::MODE_QMC, pos, true, (void **)vals, (void **)grad_lapl,
                  nullptr, num, Linv, sign, row_stride);
}
extern "C" int eval_multi_multi_UBspline_3d_d_vgl_hip(multi_UBspline_3d_d_hip *t, double *pos,
                                                      double *Linv, double *vals[],
                                                      double *grad_lapl[], int num, int row_stride) {
  return run_ptrs(t, DOUBLE_REAL, esk::MODE_QMC, pos, false, (void **)vals, (void **)grad_lapl,
                  nullptr, num, Linv, nullptr, row_stride);
}
extern "C" int eval_multi_multi_UBspline_3d_c_vgl_hip(multi_UBspline_3d_c_hip *t, float *pos,
                                                      float *Linv, complex_float *vals[],
                                                      complex_float *grad_lapl[], int num,
                                                      int row_stride) {
  return run_ptrs(t, SINGLE_COMPLEX, esk::MODE_QMC, pos, true, (void **)vals, (void **)grad_lapl,
                  nullptr, num, Linv, nullptr, row_stride);
}
extern "C" int eval_multi_multi_UBspline_3d_z_vgl_hip(multi_UBspline_3d_z_hip *t, double *pos,
                                                      double *Linv, complex_double *vals[],
                                                      complex_double *grad_lapl[], int num,
                                                      int row_stride) {
  return run_ptrs(t, DOUBLE_COMPLEX, esk::MODE_QMC, pos, false, (void **)vals, (void **)grad_lapl,
                  nullptr, num, Linv, nullptr, row_stride);
}

// 1-D CUDA-style entry points (src/multi_bspline_eval_cuda.h:10-51): one
// coordinate per point, per-point output pointers, synchronous.
static int run_ptrs12(const multi_UBspline_3d_hip *t, type_code want, int mode, const void *pos,
                      bool pos_f32, void **v, void **g, void **h, int num, int row_stride) {
  if (t && t->tcode != want)
    return fail(EINSPLINE_EINVAL, "device table has the wrong type for this entry point");
  if (t && t->dim != 1) return fail(EINSPLINE_EINVAL, "this entry point takes a 1-D table");
  if (!t) return fail(EINSPLINE_EINVAL, "NULL device table");
  if (mode == esk::MODE_VGL && !h && row_stride < t->num_splines)
    return fail(EINSPLINE_EINVAL, "row_stride %d is smaller than num_splines %d", row_stride,
                t->num_splines);
  const int ncomp = (want == SINGLE_COMPLEX || want == DOUBLE_COMPLEX) ? 2 : 1;
  int rc = run12(t, mode, pos_f32 ? esk::API_PTRS_F32POS : esk::API_PTRS_F64POS, pos, num, v, g, h,
                 (long)row_stride * ncomp, nullptr);
  if (rc || num == 0) return rc;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  return EINSPLINE_OK;
}

extern "C" int eval_multi_multi_UBspline_1d_s_hip(multi_UBspline_1d_s_hip *t, float *pos, float *vals[],
                                                  int num) {
  return run_ptrs12(t, SINGLE_REAL, esk::MODE_V, pos, true, (void **)vals, nullptr, nullptr, num, 0);
}
extern "C" int eval_multi_multi_UBspline_1d_s_vgl_hip(multi_UBspline_1d_s_hip *t, float *pos,
                                                      float *vals[], float *grads[], float *lapl[],
                                                      int num) {
  if (!lapl) return fail(EINSPLINE_EINVAL, "NULL Laplacian pointer array");
  return run_ptrs12(t, SINGLE_REAL, esk::MODE_VGL, pos, true, (void **)vals, (void **)grads,
                    (void **)lapl, num, 0);
}
extern "C" int eval_multi_multi_UBspline_1d_d_hip(multi_UBspline_1d_d_hip *t, double *pos,
                                                  double *vals[], int num) {
  return run_ptrs12(t, DOUBLE_REAL, esk::MODE_V, pos, false, (void **)vals, nullptr, nullptr, num, 0);
}
extern "C" int eval_multi_multi_UBspline_1d_d_vgl_hip(multi_UBspline_1d_d_hip *t, double *pos,
                                                      double *vals[], double *grad_lapl[], int num,
                                                      int row_stride) {
  return run_ptrs12(t, DOUBLE_REAL, esk::MODE_VGL, pos, false, (void **)vals, (void **)grad_lapl,
                    nullptr, num, row_stride);
}
extern "C" int eval_multi_multi_UBspline_1d_c_hip(multi_UBspline_1d_c_hip *t, float *pos,
                                                  complex_float *vals[], int num) {
  return run_ptrs12(t, SINGLE_COMPLEX, esk::MODE_V, pos, true, (void **)vals, nullptr, nullptr, num, 0);
}
extern "C" int eval_multi_multi_UBspline_1d_c_vgl_hip(multi_UBspline_1d_c_hip *t, float *pos,
                                                      complex_float *vals[], complex_float *grad_lapl[],
                                                      int num, int row_stride) {
  return run_ptrs12(t, SINGLE_COMPLEX, esk::MODE_VGL, pos, true, (void **)vals, (void **)grad_lapl,
                    nullptr, num, row_stride);
}
extern "C" int eval_multi_multi_UBspline_1d_z_hip(multi_UBspline_1d_z_hip *t, double *pos,
                                                  complex_double *vals[], int num) {
  return run_ptrs12(t, DOUBLE_COMPLEX, esk::MODE_V, pos, false, (void **)vals, nullptr, nullptr, num, 0);
}
extern "C" int eval_multi_multi_UBspline_1d_z_vgl_hip(multi_UBspline_1d_z_hip *t, double *pos,
                                                      complex_double *vals[], complex_double *grad_lapl[],
                                                      int num, int row_stride) {
  return run_ptrs12(t, DOUBLE_COMPLEX, esk::MODE_VGL, pos, false, (void **)vals, (void **)grad_lapl,
                    nullptr, num, row_stride);
}

// ---------------------------------------------------------------------------
// single-point path: einspline's original API, one point per call

namespace {

// Per-thread staging: pinned host memory mapped into the device's address
// space, holding the position and the results, and the thread's stream.
// The kernel reads the position from it and writes the results straight
// into it, so a call is one launch and no copies.  Deliberately not freed
// at thread exit: the HIP runtime may already be shut down when
// thread-local destructors of the main thread run.
struct Staging {
  unsigned char *host = nullptr;   // host address
  unsigned char *dev = nullptr;    // the device's address of the same memory
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  int device = -1;
};
thread_local Staging g_stage;

[[noreturn]] void die(const char *fn) {
  fprintf(stderr, "%s: %s\n", fn, g_err);
  abort();
}

int stage_reserve(size_t bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (g_stage.device != dev) {
    // the previous device's stream is idle: every call waits for its launch
    if (g_stage.host) (void)hipHostFree(g_stage.host);
    if (g_stage.stream) (void)hipStreamDestroy(g_stage.stream);
    g_stage = Staging();
    g_stage.device = dev;
  }
  if (!g_stage.stream) {
    e = hipStreamCreateWithFlags(&g_stage.stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  }
  if (g_stage.bytes >= bytes) return EINSPLINE_OK;
  if (g_stage.host) (void)hipHostFree(g_stage.host);
  g_stage.host = g_stage.dev = nullptr;
  g_stage.bytes = 0;
  const size_t want = std::max(bytes, (size_t)64 << 10);
  e = hipHostMalloc((void **)&g_stage.host, want, hipHostMallocMapped | hipHostMallocCoherent);
  if (e != hipSuccess) return hip_fail(e, "hipHostMalloc staging");
  e = hipHostGetDevicePointer((void **)&g_stage.dev, g_stage.host, 0);
  if (e != hipSuccess) return hip_fail(e, "hipHostGetDevicePointer staging");
  g_stage.bytes = want;
  return EINSPLINE_OK;
}

void destroy_mirror_table(void *m) { destroy_multi_UBspline_3d_hip((multi_UBspline_3d_hip *)m); }

// Take a reference to the current device mirror of a host table, building
// it first when the host coefficients changed (mirror_cache.hpp).
esk::MirrorRef *mirror_acquire(void *spline, const char *fn) {
  einspline_host3d *h = einspline_host_of(spline);
  if (!h) { fail(EINSPLINE_EINVAL, "spline was not created by this library"); die(fn); }
  auto build = [h]() -> void * {
    const int key = h->dim * 4 + (int)h->pub.s.tcode;
    switch (key) {
      case 3 * 4 + SINGLE_REAL: return create_multi_UBspline_3d_s_hip(&h->pub.s);
      case 3 * 4 + DOUBLE_REAL: return create_multi_UBspline_3d_d_hip(&h->pub.d);
      case 3 * 4 + SINGLE_COMPLEX: return create_multi_UBspline_3d_c_hip(&h->pub.c);
      case 3 * 4 + DOUBLE_COMPLEX: return create_multi_UBspline_3d_z_hip(&h->pub.z);
      case 2 * 4 + SINGLE_REAL: return create_multi_UBspline_2d_s_hip(&h->pub.s2);
      case 2 * 4 + DOUBLE_REAL: return create_multi_UBspline_2d_d_hip(&h->pub.d2);
      case 2 * 4 + SINGLE_COMPLEX: return create_multi_UBspline_2d_c_hip(&h->pub.c2);
      case 2 * 4 + DOUBLE_COMPLEX: return create_multi_UBspline_2d_z_hip(&h->pub.z2);
      case 1 * 4 + SINGLE_REAL: return create_multi_UBspline_1d_s_hip(&h->pub.s1);
      case 1 * 4 + DOUBLE_REAL: return create_multi_UBspline_1d_d_hip(&h->pub.d1);
      case 1 * 4 + SINGLE_COMPLEX: return create_multi_UBspline_1d_c_hip(&h->pub.c1);
      case 1 * 4 + DOUBLE_COMPLEX: return create_multi_UBspline_1d_z_hip(&h->pub.z1);
    }
    fail(EINSPLINE_EINVAL, "spline of unknown dimension or type");
    return nullptr;
  };
  esk::MirrorRef *r = esk::mirror_acquire(&h->lock, &h->dirty, &h->mirror, build, destroy_mirror_table);
  if (!r) die(fn);
  return r;
}

void mirror_release(void *spline, esk::MirrorRef *r) {
  esk::mirror_release(&einspline_host_of(spline)->lock, r, destroy_mirror_table);
}

// Evaluate one point and copy the result blocks to the caller: values (N),
// gradients (3N), Laplacian (N) or Hessian (9N).  3-D tables go to the
// resident polling kernel (point_service.hip) when it takes the call.
// Otherwise the point is evaluated into the thread's mapped staging memory:
// one launch and one stream synchronisation, no copies on the GPU side.
// (Measured on MI355X: an empty launch plus synchronisation takes 11 us, a
// launch that writes 13 KB into mapped host memory 12.6 us, the same with a
// device buffer and a copy back 14.4 us and with a position copy up as well
// 16.9 us; polling the stream with hipStreamQuery is slower, 19-23 us.)
void eval_point(void *spline, const char *fn, int mode, double x, double y, double z,
                void *vals, void *grads, void *lh, void *gh = nullptr) {
  esk::MirrorRef *mr = mirror_acquire(spline, fn);
  const multi_UBspline_3d_hip *t = (const multi_UBspline_3d_hip *)mr->table;
  if (gh == nullptr) {
    const int served = esk_point_service_eval(t, mode, x, y, z, vals, grads, lh);
    if (served < 0) die(fn);
    if (served > 0) {
      mirror_release(spline, mr);
      return;
    }
  }
  const size_t elt = (t->storage == EINSPLINE_STORE_F64 ? 8 : 4) *
                     ((t->tcode == SINGLE_COMPLEX || t->tcode == DOUBLE_COMPLEX) ? 2 : 1);
  const size_t N = (size_t)t->num_splines, dim = (size_t)t->dim;
  const size_t nv = N * elt, ng = mode >= esk::MODE_VG ? dim * N * elt : 0;
  const size_t hh = dim == 3 ? 9 : dim == 2 ? 4 : 1;    // 1-D vgh is the vgl
  const size_t nh = mode == esk::MODE_VGH || mode == esk::MODE_VGHGH ? hh * N * elt
                  : mode == esk::MODE_VGL ? N * elt : 0;
  const size_t n3 = mode == esk::MODE_VGHGH ? 27 * N * elt : 0;
  // layout in the staging buffer: [pos | vals | grads | lapl/hess | third], 256-byte aligned parts
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_v = 256, o_g = o_v + up(nv), o_h = o_g + up(ng), o_3 = o_h + up(nh),
               total = o_3 + up(n3);
  if (stage_reserve(total)) die(fn);
  unsigned char *hb = g_stage.host, *db = g_stage.dev;
  double *hp = (double *)hb;
  hp[0] = x; hp[1] = y; hp[2] = z;
  int rc = run_strided(t, mode, (const double *)db, 1, db + o_v, ng ? db + o_g : nullptr,
                       nh ? db + o_h : nullptr, 0, g_stage.stream, n3 ? db + o_3 : nullptr);
  if (rc) die(fn);
  const hipError_t e = hipStreamSynchronize(g_stage.stream);
  if (e != hipSuccess) { hip_fail(e, "single-point evaluation"); die(fn); }
  memcpy(vals, hb + o_v, nv);
  if (ng) memcpy(grads, hb + o_g, ng);
  if (nh) memcpy(lh, hb + o_h, nh);
  if (n3) memcpy(gh, hb + o_3, n3);
  mirror_release(spline, mr);
}

}  // namespace

#define POINT_API(T, TY)                                                                       \
  extern "C" void eval_multi_UBspline_3d_##T(multi_UBspline_3d_##T *s, double x, double y,    \
                                             double z, TY *v) {                                \
    eval_point(s, "eval_multi_UBspline_3d_" #T, esk::MODE_V, x, y, z, v, nullptr, nullptr);   \
  }                                                                                            \
  extern "C" void eval_multi_UBspline_3d_##T##_vg(multi_UBspline_3d_##T *s, double x,         \
                                                  double y, double z, TY *v, TY *g) {          \
    eval_point(s, "eval_multi_UBspline_3d_" #T "_vg", esk::MODE_VG, x, y, z, v, g, nullptr);  \
  }                                                                                            \
  extern "C" void eval_multi_UBspline_3d_##T##_vgl(multi_UBspline_3d_##T *s, double x,        \
                                                   double y, double z, TY *v, TY *g, TY *l) {  \
    eval_point(s, "eval_multi_UBspline_3d_" #T "_vgl", esk::MODE_VGL, x, y, z, v, g, l);      \
  }                                                                                            \
  extern "C" void eval_multi_UBspline_3d_##T##_vgh(multi_UBspline_3d_##T *s, double x,        \
                                                   double y, double z, TY *v, TY *g, TY *h) {  \
    eval_point(s, "eval_multi_UBspline_3d_" #T "_vgh", esk::MODE_VGH, x, y, z, v, g, h);      \
  }                                                                                            \
  extern "C" void eval_multi_UBspline_3d_##T##_vghgh(multi_UBspline_3d_##T *s, double x,      \
                                                     double y, double z, TY *v, TY *g, TY *h, \
                                                     TY *gh) {                                 \
    eval_point(s, "eval_multi_UBspline_3d_" #T "_vghgh", esk::MODE_VGHGH, x, y, z, v, g, h, gh); \
  }

POINT_API(s, float)
POINT_API(d, double)
POINT_API(c, complex_float)
POINT_API(z, complex_double)

// 1-D and 2-D (src/multi_bspline_eval_std_d_impl.h:37-476 and the s/c/z files)
#define POINT_API12(T, TY)                                                                      \
  extern "C" void eval_multi_UBspline_1d_##T(multi_UBspline_1d_##T *s, double x, TY *v) {       \
    eval_point(s, "eval_multi_UBspline_1d_" #T, esk::MODE_V, x, 0, 0, v, nullptr, nullptr);     \
  }                                                                                             \
  extern "C" void eval_multi_UBspline_1d_##T##_vg(multi_UBspline_1d_##T *s, double x, TY *v,   \
                                                  TY *g) {                                      \
    eval_point(s, "eval_multi_UBspline_1d_" #T "_vg", esk::MODE_VG, x, 0, 0, v, g, nullptr);    \
  }                                                                                             \
  extern "C" void eval_multi_UBspline_1d_##T##_vgl(multi_UBspline_1d_##T *s, double x, TY *v,  \
                                                   TY *g, TY *l) {                              \
    eval_point(s, "eval_multi_UBspline_1d_" #T "_vgl", esk::MODE_VGL, x, 0, 0, v, g, l);        \
  }                                                                                             \
  extern "C" void eval_multi_UBspline_1d_##T##_vgh(multi_UBspline_1d_##T *s, double x, TY *v,  \
                                                   TY *g, TY *h) {                              \
    eval_point(s, "eval_multi_UBspline_1d_" #T "_vgh", esk::MODE_VGH, x, 0, 0, v, g, h);        \
  }                                                                                             \
  extern "C" void eval_multi_UBspline_2d_##T(multi_UBspline_2d_##T *s, double x, double y,     \
                                             TY *v) {                                           \
    eval_point(s, "eval_multi_UBspline_2d_" #T, esk::MODE_V, x, y, 0, v, nullptr, nullptr);     \
  }                                                                                             \
  extern "C" void eval_multi_UBspline_2d_##T##_vg(multi_UBspline_2d_##T *s, double x, double y, \
                                                  TY *v, TY *g) {                               \
    eval_point(s, "eval_multi_UBspline_2d_" #T "_vg", esk::MODE_VG, x, y, 0, v, g, nullptr);    \
  }                                                                                             \
  extern "C" void eval_multi_UBspline_2d_##T##_vgl(multi_UBspline_2d_##T *s, double x,         \
                                                   double y, TY *v, TY *g, TY *l) {             \
    eval_point(s, "eval_multi_UBspline_2d_" #T "_vgl", esk::MODE_VGL, x, y, 0, v, g, l);        \
  }                                                                                             \
  extern "C" void eval_multi_UBspline_2d_##T##_vgh(multi_UBspline_2d_##T *s, double x,         \
                                                   double y, TY *v, TY *g, TY *h) {             \
    eval_point(s, "eval_multi_UBspline_2d_" #T "_vgh", esk::MODE_VGH, x, y, 0, v, g, h);        \
  }

POINT_API12(s, float)
POINT_API12(d, double)
POINT_API12(c, complex_float)
POINT_API12(z, complex_double)

// ---------------------------------------------------------------------------
// plain copy kernel: the bandwidth ceiling and the counter calibration for
// measured HBM traffic (a copy of B bytes reads B and writes B)

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) stream_copy_kernel(const u32x4 *__restrict__ src,
                                                          u32x4 *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

extern "C" int einspline_hip_stream_copy(void *dst, const void *src, size_t bytes, void *stream) {
  if (!dst || !src || bytes % 16) return fail(EINSPLINE_EINVAL, "copy needs 16-byte multiples");
  const size_t n = bytes / 16;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
  hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                     (const u32x4 *)src, (u32x4 *)dst, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "stream copy launch");
}

// (two instantiations: a runtime choice between the nontemporal and the
// plain store is merged by the compiler into one plain store)
template <bool NT>
__global__ void __launch_bounds__(256) stream_fill_kernel(u32x4 *__restrict__ dst, unsigned v, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const u32x4 q = {v, v, v, v};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if constexpr (NT) __builtin_nontemporal_store(q, dst + i);
    else dst[i] = q;
  }
}

extern "C" int einspline_hip_stream_fill(void *dst, unsigned value, size_t bytes, int nontemporal,
                                         void *stream) {
  if (!dst || bytes % 16) return fail(EINSPLINE_EINVAL, "fill needs 16-byte multiples");
  const size_t n = bytes / 16;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
  if (nontemporal)
    hipLaunchKernelGGL(stream_fill_kernel<true>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (u32x4 *)dst, value, n);
  else
    hipLaunchKernelGGL(stream_fill_kernel<false>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (u32x4 *)dst, value, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "stream fill launch");
}

// ---------------------------------------------------------------------------
// minimal device runtime

extern "C" int einspline_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
extern "C" void *einspline_hip_malloc(size_t bytes) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) { hip_fail(e, "hipMalloc"); return nullptr; }
  return p;
}
extern "C" int einspline_hip_free(void *p) {
  hipError_t e = hipFree(p);
  return e == hipSuccess ? 0 : hip_fail(e, "hipFree");
}
extern "C" int einspline_hip_memcpy_htod(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? 0 : hip_fail(e, "hipMemcpy H2D");
}
extern "C" int einspline_hip_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? 0 : hip_fail(e, "hipMemcpy D2H");
}
extern "C" int einspline_hip_memset(void *p, int v, size_t bytes) {
  hipError_t e = hipMemset(p, v, bytes);
  return e == hipSuccess ? 0 : hip_fail(e, "hipMemset");
}
extern "C" int einspline_hip_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? 0 : hip_fail(e, "hipDeviceSynchronize");
}
extern "C" void *einspline_hip_stream_create(void) {
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) { hip_fail(e, "hipStreamCreate"); return nullptr; }
  return s;
}
extern "C" int einspline_hip_stream_destroy(void *s) {
  hipError_t e = hipStreamDestroy((hipStream_t)s);
  return e == hipSuccess ? 0 : hip_fail(e, "hipStreamDestroy");
}
extern "C" int einspline_hip_stream_synchronize(void *s) {
  hipError_t e = hipStreamSynchronize((hipStream_t)s);
  return e == hipSuccess ? 0 : hip_fail(e, "hipStreamSynchronize");
}
extern "C" void *einspline_hip_event_create(void) {
  hipEvent_t ev = nullptr;
  hipError_t e = hipEventCreate(&ev);
  if (e != hipSuccess) { hip_fail(e, "hipEventCreate"); return nullptr; }
  return ev;
}
extern "C" int einspline_hip_event_destroy(void *ev) {
  hipError_t e = hipEventDestroy((hipEvent_t)ev);
  return e == hipSuccess ? 0 : hip_fail(e, "hipEventDestroy");
}
extern "C" int einspline_hip_event_record(void *ev, void *s) {
  hipError_t e = hipEventRecord((hipEvent_t)ev, (hipStream_t)s);
  return e == hipSuccess ? 0 : hip_fail(e, "hipEventRecord");
}
extern "C" int einspline_hip_event_synchronize(void *ev) {
  hipError_t e = hipEventSynchronize((hipEvent_t)ev);
  return e == hipSuccess ? 0 : hip_fail(e, "hipEventSynchronize");
}
extern "C" float einspline_hip_event_elapsed_ms(void *a, void *b) {
  float ms = -1.0f;
  hipError_t e = hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b);
  if (e != hipSuccess) { hip_fail(e, "hipEventElapsedTime"); return -1.0f; }
  return ms;
}
