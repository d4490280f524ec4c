// nub.hip -- 1-D complex-double multi-splines on non-uniform grids,
// evaluated on the GPU (src/multi_nubspline_eval_std_z_impl.h:32-113).
//
// Per point: the interval from the grid (bisection over the points of a
// general grid; the closed-form inverse of a center or log grid, corrected
// against the points), the four basis functions and their derivatives by
// de Boor's triangular scheme on the basis knots (nub_basis.h, the host's
// code), then every orbital is the weighted sum of coefficient rows
// i..i+3.  One work-item per (point, orbital): a row of orbitals is
// contiguous, so the four row reads of a wavefront are coalesced.  The
// interval and basis are recomputed by every work-item of a point (a few
// dozen flops, all from L1-resident data).  The device table carries one
// zero row past the last coefficient row: a point at or past the last grid
// point is evaluated on interval num_points-1, whose fourth row lies past
// the table (its weight is zero at the last point itself).
#include <hip/hip_runtime.h>

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "einspline_private.h"
#include "nub_basis.h"
#include "runtime_internal.hpp"

struct multi_NUBspline_1d_hip {
  double *coefs;        // device: (nrows + 1) rows of N complex doubles, last row zero
  int N, nrows;         // orbitals, coefficient rows (num_points + 2)
  int code, npts;       // grid_type, grid points
  double center, aInv, bInv, even_half;
  int half_points;
  double ainv, startinv;
  double *points;       // device grid points
  double *xvals;        // device basis knots (npts + 5)
  // single-point staging: position, then 3 N complex results
  double *stage_d;
  double *stage_h;      // pinned
  hipStream_t stream;
};

namespace {

struct NubArgs {
  const double *coefs;
  int N, code, npts, half_points;
  double center, aInv, bInv, even_half, ainv, startinv;
  const double *points, *xvals;
};

// floor(f) as an index guess in [0, n], 0 for NaN
__device__ int guess_of(double f, int n) {
  if (!(f == f) || f < 0.0) return 0;
  return f > (double)n ? n : (int)f;
}

__device__ int nub_interval(const NubArgs &g, double x) {
  int guess = -1;
  if (g.code == CENTER) {
    const double u = x - g.center;
    guess = guess_of(floor(copysign(log1p(fabs(u) * g.aInv) * g.bInv, u) + g.half_points - g.even_half), g.npts);
  } else if (g.code == LOG) {
    guess = guess_of(floor(x > 0.0 ? log(x * g.startinv) * g.ainv : -1.0), g.npts);
  }
  return esk_nub_interval(g.points, g.npts, x, guess);
}

template <int MODE>
__global__ void __launch_bounds__(256) nub1d_eval(NubArgs g, const double *__restrict__ pos, int num,
                                                  double2 *__restrict__ vals, double2 *__restrict__ grads,
                                                  double2 *__restrict__ lapl) {
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  const long p = gid / g.N;
  const int n = (int)(gid - p * g.N);
  if (p >= num) return;
  const double x = pos[p];
  const int i = nub_interval(g, x);
  double w[3][4];
  esk_nub_basis(g.xvals, i, x, MODE, w);
  const double *a = w[0], *da = w[1], *d2a = w[2];
  const double2 *c = (const double2 *)g.coefs + (long)i * g.N + n;
  double2 r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = c[(long)k * g.N];
  const long o = p * g.N + n;
  if (MODE == 0) {
    // vals = a0 c0 + a1 c1 + a2 c2 + a3 c3, left to right
    double2 v;
    v.x = a[0] * r[0].x + a[1] * r[1].x + a[2] * r[2].x + a[3] * r[3].x;
    v.y = a[0] * r[0].y + a[1] * r[1].y + a[2] * r[2].y + a[3] * r[3].y;
    vals[o] = v;
    return;
  }
  // vg / vgl: accumulated from zero, one row at a time
  double2 v = {0.0, 0.0}, gr = {0.0, 0.0}, l = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v.x += a[k] * r[k].x;  v.y += a[k] * r[k].y;
    gr.x += da[k] * r[k].x; gr.y += da[k] * r[k].y;
    if (MODE >= 2) { l.x += d2a[k] * r[k].x; l.y += d2a[k] * r[k].y; }
  }
  vals[o] = v;
  grads[o] = gr;
  if (MODE >= 2) lapl[o] = l;
}

NubArgs args_of(const multi_NUBspline_1d_hip *t) {
  NubArgs g;
  g.coefs = t->coefs;
  g.N = t->N; g.code = t->code; g.npts = t->npts; g.half_points = t->half_points;
  g.center = t->center; g.aInv = t->aInv; g.bInv = t->bInv; g.even_half = t->even_half;
  g.ainv = t->ainv; g.startinv = t->startinv;
  g.points = t->points; g.xvals = t->xvals;
  return g;
}

int launch(const multi_NUBspline_1d_hip *t, const double *pos, int num, void *v, void *g, void *l, int mode,
           hipStream_t s) {
  const long total = (long)num * t->N;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  const NubArgs a = args_of(t);
  switch (mode) {
    case 0: hipLaunchKernelGGL(nub1d_eval<0>, dim3(blocks), dim3(256), 0, s, a, pos, num, (double2 *)v, nullptr, nullptr); break;
    case 1: hipLaunchKernelGGL(nub1d_eval<1>, dim3(blocks), dim3(256), 0, s, a, pos, num, (double2 *)v, (double2 *)g, nullptr); break;
    default: hipLaunchKernelGGL(nub1d_eval<2>, dim3(blocks), dim3(256), 0, s, a, pos, num, (double2 *)v, (double2 *)g, (double2 *)l); break;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EINSPLINE_OK : esk_hip_fail(e, "nub eval launch");
}

void free_table(multi_NUBspline_1d_hip *t) {
  if (!t) return;
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  (void)hipFree(t->coefs);
  (void)hipFree(t->points);
  (void)hipFree(t->xvals);
  (void)hipFree(t->stage_d);
  if (t->stage_h) (void)hipHostFree(t->stage_h);
  if (t->stream) (void)hipStreamDestroy(t->stream);
  free(t);
}

}  // namespace

extern "C" multi_NUBspline_1d_hip *create_multi_NUBspline_1d_z_hip(multi_NUBspline_1d_z *spline) {
  einspline_nub1d *h = einspline_nub_of(spline);
  if (!h || h->ncomp != 2) { esk_fail(EINSPLINE_EINVAL, "create_multi_NUBspline_1d_z_hip: not a 1-D z spline of this library"); return nullptr; }
  const NUgrid *grid = spline->x_grid;
  const NUBasis *basis = spline->x_basis;
  multi_NUBspline_1d_hip *t = (multi_NUBspline_1d_hip *)calloc(1, sizeof(multi_NUBspline_1d_hip));
  if (!t) { esk_fail(EINSPLINE_ENOMEM, "out of host memory"); return nullptr; }
  t->N = spline->num_splines;
  t->nrows = h->nrows;
  t->npts = grid->num_points;
  t->code = grid->code == CENTER || grid->code == LOG ? (int)grid->code : (int)GENERAL;
  if (grid->code == CENTER) {
    const center_grid *c = (const center_grid *)grid;
    t->center = c->center; t->aInv = c->aInv; t->bInv = c->bInv; t->even_half = c->even_half;
    t->half_points = c->half_points;
  } else if (grid->code == LOG) {
    const log_grid *l = (const log_grid *)grid;
    t->ainv = l->ainv; t->startinv = l->startinv;
  }
  const size_t rowb = (size_t)t->N * 2 * sizeof(double);
  const size_t cbytes = (size_t)(t->nrows + 1) * rowb;
  hipError_t e;
  if ((e = hipMalloc((void **)&t->coefs, cbytes)) != hipSuccess ||
      (e = hipMalloc((void **)&t->points, sizeof(double) * t->npts)) != hipSuccess ||
      (e = hipMalloc((void **)&t->xvals, sizeof(double) * (t->npts + 5))) != hipSuccess ||
      (e = hipMalloc((void **)&t->stage_d, 256 + 3 * rowb)) != hipSuccess ||
      (e = hipHostMalloc((void **)&t->stage_h, 256 + 3 * rowb, hipHostMallocDefault)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMemcpy(t->coefs, spline->coefs, (size_t)t->nrows * rowb, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemset((char *)t->coefs + (size_t)t->nrows * rowb, 0, rowb)) != hipSuccess ||
      (e = hipMemcpy(t->points, grid->points, sizeof(double) * t->npts, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(t->xvals, basis->xVals, sizeof(double) * (t->npts + 5), hipMemcpyHostToDevice)) != hipSuccess) {
    esk_hip_fail(e, "create_multi_NUBspline_1d_z_hip");
    free_table(t);
    return nullptr;
  }
  return t;
}

extern "C" void destroy_multi_NUBspline_1d_hip(multi_NUBspline_1d_hip *t) { free_table(t); }

extern "C" void einspline_nub_mirror_release(void *mirror) { free_table((multi_NUBspline_1d_hip *)mirror); }

extern "C" int einspline_eval_nub_1d_z_batch(const multi_NUBspline_1d_hip *t, const double *pos_d, int num,
                                             void *vals_d, void *grads_d, void *lapl_d, int mode, void *stream) {
  if (!t) return esk_fail(EINSPLINE_EINVAL, "NULL non-uniform table");
  if (num < 0 || mode < 0 || mode > 2) return esk_fail(EINSPLINE_EINVAL, "bad point count or mode");
  if (num == 0) return EINSPLINE_OK;
  if (!pos_d || !vals_d || (mode >= 1 && !grads_d) || (mode >= 2 && !lapl_d))
    return esk_fail(EINSPLINE_EINVAL, "NULL position or output pointer");
  if ((long long)num * t->N >= (1ll << 40)) return esk_fail(EINSPLINE_EINVAL, "batch too large");
  return launch(t, pos_d, num, vals_d, grads_d, lapl_d, mode, (hipStream_t)stream);
}

// ---- single-point API (device mirror refreshed after set_*) -------------

namespace {

[[noreturn]] void die(const char *fn) {
  fprintf(stderr, "%s: %s\n", fn, einspline_hip_last_error());
  abort();
}

// Called with h->lock held.
const multi_NUBspline_1d_hip *nub_mirror_locked(einspline_nub1d *h, multi_NUBspline_1d_z *spline) {
  if (h->dirty || !h->mirror) {
    if (h->mirror) free_table((multi_NUBspline_1d_hip *)h->mirror);
    h->mirror = create_multi_NUBspline_1d_z_hip(spline);
    if (!h->mirror) return nullptr;
    h->dirty = 0;
  }
  return (const multi_NUBspline_1d_hip *)h->mirror;
}

void nub_point(multi_NUBspline_1d_z *spline, const char *fn, int mode, double x, void *v, void *g, void *l) {
  einspline_nub1d *h = einspline_nub_of(spline);
  if (!h || h->ncomp != 2) {
    fprintf(stderr, "%s: spline was not created by this library\n", fn);
    abort();
  }
  // the mirror and its staging area are used under the spline's lock
  pthread_mutex_lock(&h->lock);
  const multi_NUBspline_1d_hip *t = nub_mirror_locked(h, spline);
  int rc = t ? EINSPLINE_OK : EINSPLINE_EHIP;
  if (rc == EINSPLINE_OK) {
    const size_t rowb = (size_t)t->N * 2 * sizeof(double);
    char *hb = (char *)t->stage_h, *db = (char *)t->stage_d;
    *(double *)hb = x;
    hipError_t e = hipMemcpyAsync(db, hb, sizeof(double), hipMemcpyHostToDevice, t->stream);
    if (e != hipSuccess) rc = esk_hip_fail(e, "position copy");
    if (rc == EINSPLINE_OK)
      rc = launch(t, (const double *)db, 1, db + 256, db + 256 + rowb, db + 256 + 2 * rowb, mode, t->stream);
    const size_t nout = (size_t)(mode + 1) * rowb;
    if (rc == EINSPLINE_OK &&
        ((e = hipMemcpyAsync(hb + 256, db + 256, nout, hipMemcpyDeviceToHost, t->stream)) != hipSuccess ||
         (e = hipStreamSynchronize(t->stream)) != hipSuccess))
      rc = esk_hip_fail(e, "result copy");
    if (rc == EINSPLINE_OK) {
      memcpy(v, hb + 256, rowb);
      if (mode >= 1) memcpy(g, hb + 256 + rowb, rowb);
      if (mode >= 2) memcpy(l, hb + 256 + 2 * rowb, rowb);
    }
  }
  pthread_mutex_unlock(&h->lock);
  if (rc) die(fn);
}

}  // namespace

extern "C" void eval_multi_NUBspline_1d_z(multi_NUBspline_1d_z *s, double x, complex_double *vals) {
  nub_point(s, "eval_multi_NUBspline_1d_z", 0, x, vals, nullptr, nullptr);
}
extern "C" void eval_multi_NUBspline_1d_z_vg(multi_NUBspline_1d_z *s, double x, complex_double *vals,
                                             complex_double *grads) {
  nub_point(s, "eval_multi_NUBspline_1d_z_vg", 1, x, vals, grads, nullptr);
}
extern "C" void eval_multi_NUBspline_1d_z_vgl(multi_NUBspline_1d_z *s, double x, complex_double *vals,
                                              complex_double *grads, complex_double *lapl) {
  nub_point(s, "eval_multi_NUBspline_1d_z_vgl", 2, x, vals, grads, lapl);
}
extern "C" void eval_multi_NUBspline_1d_z_vgh(multi_NUBspline_1d_z *s, double x, complex_double *vals,
                                              complex_double *grads, complex_double *hess) {
  nub_point(s, "eval_multi_NUBspline_1d_z_vgh", 2, x, vals, grads, hess);
}
