// table_io.hip -- device tables on disk, so that a serving process can
// restart without solving or converting its tables again (the reference has
// no table file).  The file holds the table as the kernels read it: a fixed
// header, then the coefficient rows in the storage type, padded to the
// device row length, byte for byte as they sit in HBM.
//
// Layout (little-endian, 512-byte header):
//     0  char[8]  "ESPLTBL1"
//     8  u32      header bytes (512)
//    12  u32      format version (1)
//    16  i32 x 8  dim, tcode, storage, num_splines, row_reals, Nx, Ny, Nz
//    48  3 x 40   per axis: f64 start, f64 end, i32 num, i32 0, f64 delta, f64 delta_inv
//   168  i32 x 4  tuning: lane bytes, block, store policy, point order
//   184  u64      coefficient bytes = Nx * Ny * Nz * row_reals * element bytes
//   192  u64      hash of the coefficient bytes
//   200  u64      hash of header bytes [0, 200)
//   208  zeros to 512
//   512  coefficients
// Loading checks the magic, version, both hashes and every size before the
// table is used.
#include <hip/hip_runtime.h>

#include <errno.h>
#include <fcntl.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "runtime_internal.hpp"

namespace {

constexpr char kMagic[8] = {'E', 'S', 'P', 'L', 'T', 'B', 'L', '1'};
constexpr uint32_t kHeaderBytes = 512, kVersion = 1;
constexpr size_t kChunk = (size_t)32 << 20;      // staging chunk

struct AxisRec {
  double start, end;
  int32_t num, zero;
  double delta, delta_inv;
};
struct Header {
  char magic[8];
  uint32_t header_bytes, version;
  int32_t dim, tcode, storage, num_splines, row_reals, Nx, Ny, Nz;
  AxisRec axis[3];
  int32_t lane_bytes, block, store_policy, point_order;
  uint64_t coef_bytes, coef_hash, head_hash;
};
static_assert(sizeof(AxisRec) == 40, "axis record layout");
static_assert(offsetof(Header, axis) == 48 && offsetof(Header, lane_bytes) == 168 &&
                  offsetof(Header, coef_bytes) == 184 && offsetof(Header, head_hash) == 200,
              "header layout");

// Four-lane 64-bit multiply-rotate hash over 8-byte words (lanes take words
// round-robin, so the four chains run in parallel); the tail is zero-padded.
struct Hash {
  uint64_t h[4] = {0x243f6a8885a308d3ull, 0x13198a2e03707344ull, 0xa4093822299f31d0ull,
                   0x082efa98ec4e6c89ull};
  uint64_t n = 0;          // bytes hashed
  int lane = 0;
  static uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
  void word(uint64_t w) {
    uint64_t &x = h[lane];
    x = rotl(x ^ (w * 0x9e3779b97f4a7c15ull), 29) * 0xbf58476d1ce4e5b9ull;
    lane = (lane + 1) & 3;
  }
  // Feed bytes; all but the last call must pass a multiple of 8 bytes.
  void add(const void *p, size_t bytes) {
    const unsigned char *b = (const unsigned char *)p;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
      uint64_t w;
      memcpy(&w, b + i, 8);
      word(w);
    }
    if (i < bytes) {
      uint64_t w = 0;
      memcpy(&w, b + i, bytes - i);
      word(w);
    }
    n += bytes;
  }
  uint64_t done() const {
    uint64_t x = n * 0xd6e8feb86659fd93ull;
    for (int i = 0; i < 4; ++i) x = rotl(x ^ h[i], 23) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
  }
};

// Every axis a table of dimension `dim` uses needs at least the 4 planes of
// one stencil, and a grid consistent with its extent: N = num + 3 with
// delta_inv = num / (end - start) (periodic and antiperiodic conditions), or
// N = num + 2 with delta_inv = (num - 1) / (end - start) (the others).
// Axes past `dim` are single planes.  Returns NULL or what is wrong.
const char *check_axes(const Header &hd) {
  const int32_t n[3] = {hd.Nx, hd.Ny, hd.Nz};
  for (int a = 0; a < 3; ++a) {
    if (a >= hd.dim) {
      if (n[a] != 1) return "an axis past the table's dimension has more than one plane";
      continue;
    }
    if (n[a] < 4) return "an axis has fewer than 4 planes";
    const AxisRec &g = hd.axis[a];
    const double ext = g.end - g.start;
    // (create accepts end < start: a reversed axis has a negative delta_inv)
    if (!std::isfinite(g.start) || !std::isfinite(g.end) || !std::isfinite(ext) || ext == 0.0)
      return "an axis has no extent";
    if (!std::isfinite(g.delta_inv) || g.delta_inv == 0.0) return "an axis has a bad delta_inv";
    double intervals;
    if (n[a] == g.num + 3) intervals = g.num;
    else if (n[a] == g.num + 2) intervals = g.num - 1;
    else return "an axis's plane count does not match its grid";
    if (!(intervals >= 1.0) || fabs(g.delta_inv * ext - intervals) > 1e-9 * intervals)
      return "an axis's delta_inv does not match its grid";
  }
  return nullptr;
}

uint64_t head_hash(const Header &hd) {
  Hash hs;
  hs.add(&hd, offsetof(Header, head_hash));
  return hs.done();
}

size_t elem_bytes(int storage) {
  switch (storage) {
    case EINSPLINE_STORE_F32: return 4;
    case EINSPLINE_STORE_F64: return 8;
    case EINSPLINE_STORE_BF16: case EINSPLINE_STORE_F16: return 2;
  }
  return 0;
}

bool write_all(int fd, const void *p, size_t n) {
  const char *c = (const char *)p;
  while (n) {
    const ssize_t w = write(fd, c, n);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) return false;
    c += w;
    n -= (size_t)w;
  }
  return true;
}
bool read_all(int fd, void *p, size_t n) {
  char *c = (char *)p;
  while (n) {
    const ssize_t r = read(fd, c, n);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    c += r;
    n -= (size_t)r;
  }
  return true;
}

// Pinned double buffer and a private stream for chunked transfers.
struct Staging {
  void *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipStream_t s = nullptr;
  ~Staging() {
    if (s) (void)hipStreamSynchronize(s);
    for (int i = 0; i < 2; ++i) {
      if (buf[i]) (void)hipHostFree(buf[i]);
      if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
    if (s) (void)hipStreamDestroy(s);
  }
  hipError_t init(size_t bytes) {
    hipError_t e;
    for (int i = 0; i < 2; ++i) {
      if ((e = hipHostMalloc(&buf[i], bytes, hipHostMallocDefault)) != hipSuccess) return e;
      if ((e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming)) != hipSuccess) return e;
    }
    return hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  }
};

}  // namespace

extern "C" int einspline_table_save(const multi_UBspline_3d_hip *t, const char *path) {
  if (!t || !t->coefs || !path) return esk_fail(EINSPLINE_EINVAL, "table_save: NULL table or path");
  const size_t eb = elem_bytes(t->storage);
  const size_t bytes = (size_t)t->Nx * t->Ny * t->Nz * t->row_reals * eb;
  if (!eb || bytes != t->bytes) return esk_fail(EINSPLINE_EINVAL, "table_save: inconsistent table sizes");
  Header hd;
  memset(&hd, 0, sizeof hd);
  memcpy(hd.magic, kMagic, 8);
  hd.header_bytes = kHeaderBytes;
  hd.version = kVersion;
  hd.dim = t->dim; hd.tcode = (int32_t)t->tcode; hd.storage = (int32_t)t->storage;
  hd.num_splines = t->num_splines; hd.row_reals = t->row_reals;
  hd.Nx = t->Nx; hd.Ny = t->Ny; hd.Nz = t->Nz;
  const Ugrid *g[3] = {&t->x_grid, &t->y_grid, &t->z_grid};
  for (int a = 0; a < 3; ++a)
    hd.axis[a] = AxisRec{g[a]->start, g[a]->end, g[a]->num, 0, g[a]->delta, g[a]->delta_inv};
  int lb = 0, blk = 0, pol = 0, ord = 0;
  if (einspline_get_tuning(t, &lb, &blk, &pol, &ord) == EINSPLINE_OK) {
    hd.lane_bytes = lb; hd.block = blk; hd.store_policy = pol; hd.point_order = ord;
  }
  hd.coef_bytes = bytes;

  const std::string tmp = std::string(path) + ".tmp";
  const int fd = open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) return esk_fail(EINSPLINE_EINVAL, "table_save: cannot create %s: %s", tmp.c_str(), strerror(errno));
  int rc = EINSPLINE_OK;
  {
    unsigned char pad[kHeaderBytes] = {0};
    if (!write_all(fd, pad, kHeaderBytes)) rc = esk_fail(EINSPLINE_EINVAL, "table_save: write failed");
    Staging st;
    hipError_t e = hipSuccess;
    if (rc == EINSPLINE_OK && (e = st.init(std::min(kChunk, std::max<size_t>(bytes, 8)))) != hipSuccess)
      rc = esk_hip_fail(e, "table_save staging");
    // device -> pinned (chunk k+1 in flight while chunk k is hashed and written)
    Hash hs;
    const size_t nchunks = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t k) {
      const size_t off = k * kChunk, n = std::min(kChunk, bytes - off);
      hipError_t x = hipMemcpyAsync(st.buf[k & 1], (const char *)t->coefs + off, n, hipMemcpyDeviceToHost, st.s);
      if (x == hipSuccess) x = hipEventRecord(st.ev[k & 1], st.s);
      return x;
    };
    if (rc == EINSPLINE_OK && nchunks && (e = issue(0)) != hipSuccess) rc = esk_hip_fail(e, "table_save copy");
    for (size_t k = 0; rc == EINSPLINE_OK && k < nchunks; ++k) {
      if ((e = hipEventSynchronize(st.ev[k & 1])) != hipSuccess) { rc = esk_hip_fail(e, "table_save copy"); break; }
      if (k + 1 < nchunks && (e = issue(k + 1)) != hipSuccess) { rc = esk_hip_fail(e, "table_save copy"); break; }
      const size_t n = std::min(kChunk, bytes - k * kChunk);
      hs.add(st.buf[k & 1], n);
      if (!write_all(fd, st.buf[k & 1], n)) rc = esk_fail(EINSPLINE_EINVAL, "table_save: write failed");
    }
    if (rc == EINSPLINE_OK) {
      hd.coef_hash = hs.done();
      hd.head_hash = head_hash(hd);
      unsigned char head[kHeaderBytes] = {0};
      memcpy(head, &hd, sizeof hd);
      if (lseek(fd, 0, SEEK_SET) != 0 || !write_all(fd, head, kHeaderBytes) || fsync(fd) != 0)
        rc = esk_fail(EINSPLINE_EINVAL, "table_save: write failed");
    }
  }
  if (close(fd) != 0 && rc == EINSPLINE_OK) rc = esk_fail(EINSPLINE_EINVAL, "table_save: close failed");
  if (rc == EINSPLINE_OK && rename(tmp.c_str(), path) != 0)
    rc = esk_fail(EINSPLINE_EINVAL, "table_save: cannot rename to %s: %s", path, strerror(errno));
  if (rc != EINSPLINE_OK) unlink(tmp.c_str());
  return rc;
}

extern "C" multi_UBspline_3d_hip *einspline_table_load(const char *path) {
  if (!path) { esk_fail(EINSPLINE_EINVAL, "table_load: NULL path"); return nullptr; }
  const int fd = open(path, O_RDONLY);
  if (fd < 0) { esk_fail(EINSPLINE_EINVAL, "table_load: cannot open %s: %s", path, strerror(errno)); return nullptr; }
  struct Closer { int fd; ~Closer() { close(fd); } } closer{fd};
  unsigned char head[kHeaderBytes];
  Header hd;
  if (!read_all(fd, head, kHeaderBytes)) { esk_fail(EINSPLINE_EINVAL, "table_load: %s: short header", path); return nullptr; }
  memcpy(&hd, head, sizeof hd);
  if (memcmp(hd.magic, kMagic, 8) != 0) { esk_fail(EINSPLINE_EINVAL, "table_load: %s: not a table file", path); return nullptr; }
  if (hd.version != kVersion || hd.header_bytes != kHeaderBytes) {
    esk_fail(EINSPLINE_EINVAL, "table_load: %s: format version %u not supported", path, hd.version);
    return nullptr;
  }
  if (head_hash(hd) != hd.head_hash) { esk_fail(EINSPLINE_EINVAL, "table_load: %s: header hash mismatch", path); return nullptr; }
  const size_t eb = elem_bytes(hd.storage);
  const bool cplx = hd.tcode == SINGLE_COMPLEX || hd.tcode == DOUBLE_COMPLEX;
  if (!eb || hd.dim < 1 || hd.dim > 3 || hd.tcode < SINGLE_REAL || hd.tcode > DOUBLE_COMPLEX ||
      hd.num_splines < 1 || hd.row_reals < 1 ||
      (size_t)hd.row_reals < (size_t)hd.num_splines * (cplx ? 2 : 1) ||
      ((size_t)hd.row_reals * eb) % 32 != 0) {   // the kernels read rows in whole 16- and 32-byte pieces
    esk_fail(EINSPLINE_EINVAL, "table_load: %s: inconsistent header", path);
    return nullptr;
  }
  const char *bad_axis = check_axes(hd);
  if (bad_axis) {
    esk_fail(EINSPLINE_EINVAL, "table_load: %s: inconsistent header: %s", path, bad_axis);
    return nullptr;
  }
  uint64_t want = 0;
  if (__builtin_mul_overflow((uint64_t)hd.Nx, (uint64_t)hd.Ny, &want) ||
      __builtin_mul_overflow(want, (uint64_t)hd.Nz, &want) ||
      __builtin_mul_overflow(want, (uint64_t)hd.row_reals, &want) ||
      __builtin_mul_overflow(want, (uint64_t)eb, &want) || hd.coef_bytes != want) {
    esk_fail(EINSPLINE_EINVAL, "table_load: %s: inconsistent header: coefficient bytes", path);
    return nullptr;
  }
  const off_t fsize = lseek(fd, 0, SEEK_END);
  if (fsize != (off_t)(kHeaderBytes + hd.coef_bytes) || lseek(fd, kHeaderBytes, SEEK_SET) != (off_t)kHeaderBytes) {
    esk_fail(EINSPLINE_EINVAL, "table_load: %s: file is %lld bytes, header says %llu", path, (long long)fsize,
             (unsigned long long)(kHeaderBytes + hd.coef_bytes));
    return nullptr;
  }
  const size_t bytes = hd.coef_bytes;
  multi_UBspline_3d_hip meta;
  memset(&meta, 0, sizeof meta);
  meta.z_stride = hd.row_reals;
  meta.y_stride = (intptr_t)hd.Nz * hd.row_reals;
  meta.x_stride = (intptr_t)hd.Ny * hd.Nz * hd.row_reals;
  Ugrid *g[3] = {&meta.x_grid, &meta.y_grid, &meta.z_grid};
  for (int a = 0; a < 3; ++a) {
    g[a]->start = hd.axis[a].start; g[a]->end = hd.axis[a].end; g[a]->num = hd.axis[a].num;
    g[a]->delta = hd.axis[a].delta; g[a]->delta_inv = hd.axis[a].delta_inv;
  }
  meta.Nx = hd.Nx; meta.Ny = hd.Ny; meta.Nz = hd.Nz;
  meta.num_splines = hd.num_splines;
  meta.row_reals = hd.row_reals;
  meta.tcode = (type_code)hd.tcode;
  meta.storage = (einspline_storage)hd.storage;
  meta.bytes = bytes;
  meta.dim = hd.dim;
  hipError_t e = hipGetDevice(&meta.device);
  if (e != hipSuccess) { esk_hip_fail(e, "hipGetDevice"); return nullptr; }
  void *d = nullptr;
  if ((e = hipMalloc(&d, bytes)) != hipSuccess) {
    esk_fail(EINSPLINE_ENOMEM, "table_load: hipMalloc of %zu bytes: %s", bytes, hipGetErrorString(e));
    return nullptr;
  }
  int rc = EINSPLINE_OK;
  {
    Staging st;
    if ((e = st.init(std::min(kChunk, bytes))) != hipSuccess) rc = esk_hip_fail(e, "table_load staging");
    Hash hs;
    const size_t nchunks = (bytes + kChunk - 1) / kChunk;
    for (size_t k = 0; rc == EINSPLINE_OK && k < nchunks; ++k) {
      const int b = (int)(k & 1);
      const size_t off = k * kChunk, n = std::min(kChunk, bytes - off);
      // the copy issued from this buffer two chunks ago must be done before it is refilled
      if (k >= 2 && (e = hipEventSynchronize(st.ev[b])) != hipSuccess) { rc = esk_hip_fail(e, "table_load copy"); break; }
      if (!read_all(fd, st.buf[b], n)) { rc = esk_fail(EINSPLINE_EINVAL, "table_load: %s: read failed", path); break; }
      hs.add(st.buf[b], n);
      if ((e = hipMemcpyAsync((char *)d + off, st.buf[b], n, hipMemcpyHostToDevice, st.s)) != hipSuccess ||
          (e = hipEventRecord(st.ev[b], st.s)) != hipSuccess) {
        rc = esk_hip_fail(e, "table_load copy");
        break;
      }
    }
    if (rc == EINSPLINE_OK && (e = hipStreamSynchronize(st.s)) != hipSuccess) rc = esk_hip_fail(e, "table_load copy");
    if (rc == EINSPLINE_OK && hs.done() != hd.coef_hash)
      rc = esk_fail(EINSPLINE_EINVAL, "table_load: %s: coefficient hash mismatch (file damaged)", path);
  }
  if (rc != EINSPLINE_OK) { (void)hipFree(d); return nullptr; }
  multi_UBspline_3d_hip *t = esk_table_wrap(&meta, d);
  if (!t) { (void)hipFree(d); return nullptr; }
  // restore the saved tuning (all zero means defaults)
  if (hd.lane_bytes || hd.block) (void)einspline_set_tile(t, hd.lane_bytes, hd.block);
  (void)einspline_set_store_policy(t, hd.store_policy);
  (void)einspline_set_point_order(t, hd.point_order);
  return t;
}
