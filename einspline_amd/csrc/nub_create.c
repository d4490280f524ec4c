/* nub_create.c -- non-uniform grids, their cubic B-spline basis, and the
 * 1-D non-uniform multi-splines (create / set / destroy).
 *
 * Follows the reference's definitions:
 *   grids:  src/nugrid.c:25-172 (general grid: bisection reverse map; center
 *           grid: points c +- a expm1(b k), reverse map by log1p; log grid:
 *           points start exp(a i), reverse map by log);
 *   basis:  src/nubasis.c:24-685 (two extra knots before the grid and three
 *           after; Cox-de Boor recursion over dxInv[3i+j] = 1/(x_{i+j+1}-x_i));
 *   solves: src/nubspline_create.c:389-574 (tridiagonal elimination with the
 *           two derivative rows, or the cyclic system with a last column);
 *   multi:  src/multi_nubspline_create.c:643-700, :898-965 (one orbital per
 *           set call, complex orbitals as two real solves).
 * The steps of the recursion and of the eliminations are the reference's,
 * in the same order, so results match it to the last bit in double.  Unlike
 * the reference, every reverse map clamps its index to [0, num_points-1] (the
 * center and log maps can otherwise leave the table for points outside the
 * grid).
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "einspline_private.h"

/* ------------------------------------------------------------------------ */
/* grids                                                                     */

static int clamp_index(int i, int n) { return i < 0 ? 0 : (i > n - 1 ? n - 1 : i); }

static int general_reverse(void *gp, double x) {
  const NUgrid *g = (const NUgrid *)gp;
  const int n = g->num_points;
  const double *p = g->points;
  if (x <= p[0]) return 0;
  if (x >= p[n - 1]) return n - 1;
  int lo = 0, hi = n - 1;                 /* p[lo] <= x < p[hi] */
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (p[mid] > x) hi = mid; else lo = mid;
  }
  return lo;
}

static int center_reverse(void *gp, double x) {
  const center_grid *g = (const center_grid *)gp;
  const double u = x - g->center;
  const double k = copysign(log1p(fabs(u) * g->aInv) * g->bInv, u);
  const double f = floor(g->half_points + k - g->even_half);
  if (!(f == f)) return 0;
  return clamp_index(f < -1.0 ? -1 : (f > (double)g->num_points ? g->num_points : (int)f), g->num_points);
}

static int log_reverse(void *gp, double x) {
  const log_grid *g = (const log_grid *)gp;
  const double f = floor(g->ainv * log(x * g->startinv));
  if (!(f == f)) return 0;
  return clamp_index(f < -1.0 ? -1 : (f > (double)g->num_points ? g->num_points : (int)f), g->num_points);
}

NUgrid *create_general_grid(double *points, int num_points) {
  if (!points || num_points < 2) return NULL;
  NUgrid *g = calloc(1, sizeof(NUgrid));
  if (!g) return NULL;
  g->points = malloc(sizeof(double) * (size_t)num_points);
  if (!g->points) { free(g); return NULL; }
  memcpy(g->points, points, sizeof(double) * (size_t)num_points);
  g->code = GENERAL;
  g->start = points[0];
  g->end = points[num_points - 1];
  g->num_points = num_points;
  g->reverse_map = general_reverse;
  return g;
}

NUgrid *create_center_grid(double start, double end, double ratio, int num_points) {
  if (!(ratio > 1.0) || num_points < 4) return NULL;
  center_grid *g = calloc(1, sizeof(center_grid));
  if (!g) return NULL;
  g->points = malloc(sizeof(double) * (size_t)num_points);
  if (!g->points) { free(g); return NULL; }
  g->code = CENTER;
  g->start = start;
  g->end = end;
  g->center = 0.5 * (start + end);
  g->num_points = num_points;
  g->half_points = num_points / 2;
  g->odd = (num_points % 2) == 1;
  g->b = log(ratio) / (double)(g->half_points - 1);
  g->bInv = 1.0 / g->b;
  const int h = g->half_points;
  if (g->odd) {
    /* points at c + sign(k) a expm1(b |k|), k = -h..h */
    g->even_half = 0.0;
    g->odd_one = 1;
    g->a = 0.5 * (end - start) / expm1(g->b * h);
    g->aInv = 1.0 / g->a;
    for (int k = -h; k <= h; ++k) {
      const double sign = k < 0 ? -1.0 : 1.0;
      g->points[k + h] = sign * g->a * expm1(g->b * abs(k)) + g->center;
    }
  } else {
    /* points at c + sign(k) a expm1(b |k + 1/2|), k = -h..h-1 */
    g->even_half = 0.5;
    g->odd_one = 0;
    g->a = 0.5 * (end - start) / expm1(g->b * (-0.5 + h));
    g->aInv = 1.0 / g->a;
    for (int k = -h; k < h; ++k) {
      const double sign = k < 0 ? -1.0 : 1.0;
      g->points[k + h] = sign * g->a * expm1(g->b * fabs(0.5 + k)) + g->center;
    }
  }
  g->reverse_map = center_reverse;
  return (NUgrid *)g;
}

NUgrid *create_log_grid(double start, double end, int num_points) {
  if (!(start > 0.0) || !(end > start) || num_points < 2) return NULL;
  log_grid *g = calloc(1, sizeof(log_grid));
  if (!g) return NULL;
  g->points = malloc(sizeof(double) * (size_t)num_points);
  if (!g->points) { free(g); return NULL; }
  g->code = LOG;
  g->start = start;
  g->end = end;
  g->num_points = num_points;
  g->a = 1.0 / (double)(num_points - 1) * log(end / start);
  g->ainv = 1.0 / g->a;
  g->startinv = 1.0 / start;
  for (int i = 0; i < num_points; ++i) g->points[i] = start * exp(g->a * (double)i);
  g->reverse_map = log_reverse;
  return (NUgrid *)g;
}

void destroy_grid(NUgrid *grid) {
  if (!grid) return;
  free(grid->points);
  free(grid);
}

/* ------------------------------------------------------------------------ */
/* basis                                                                     */

NUBasis *create_NUBasis(NUgrid *grid, bool periodic) {
  if (!grid || grid->num_points < 4) return NULL;
  NUBasis *b = calloc(1, sizeof(NUBasis));
  if (!b) return NULL;
  const int N = grid->num_points;
  b->grid = grid;
  b->periodic = periodic;
  b->xVals = malloc(sizeof(double) * (size_t)(N + 5));
  b->dxInv = malloc(sizeof(double) * 3 * (size_t)(N + 2));
  if (!b->xVals || !b->dxInv) { free(b->xVals); free(b->dxInv); free(b); return NULL; }
  const double *g = grid->points;
  double *xv = b->xVals;
  for (int i = 0; i < N; ++i) xv[i + 2] = g[i];
  if (!periodic) {
    /* extend with the end spacings */
    xv[0] = g[0] - 2.0 * (g[1] - g[0]);
    xv[1] = g[0] - 1.0 * (g[1] - g[0]);
    xv[N + 2] = g[N - 1] + 1.0 * (g[N - 1] - g[N - 2]);
    xv[N + 3] = g[N - 1] + 2.0 * (g[N - 1] - g[N - 2]);
    xv[N + 4] = g[N - 1] + 3.0 * (g[N - 1] - g[N - 2]);
  } else {
    /* wrap the spacings from the other end */
    xv[1] = g[0] - (g[N - 1] - g[N - 2]);
    xv[0] = g[0] - (g[N - 1] - g[N - 3]);
    xv[N + 2] = g[N - 1] + (g[1] - g[0]);
    xv[N + 3] = g[N - 1] + (g[2] - g[0]);
    xv[N + 4] = g[N - 1] + (g[3] - g[0]);
  }
  for (int i = 0; i < N + 2; ++i)
    for (int j = 0; j < 3; ++j) b->dxInv[3 * i + j] = 1.0 / (xv[i + j + 1] - xv[i]);
  return b;
}

void destroy_NUBasis(NUBasis *basis) {
  if (!basis) return;
  free(basis->xVals);
  free(basis->dxInv);
  free(basis);
}

/* The four non-zero cubic basis functions on interval i at x (and their
 * first and second derivatives when d1 / d2 are non-NULL): degree-1 values
 * b1, degree-2 values b2, then degree 3, each level from the one below. */
static void nub_funcs(const NUBasis *basis, int i, double x, double b[4], double *d1, double *d2) {
  const double *dx = basis->dxInv, *xv = basis->xVals;
  const int i2 = i + 2;
  double b1[2], b2[3];
  b1[0] = (xv[i2 + 1] - x) * dx[3 * (i + 2) + 0];
  b1[1] = (x - xv[i2]) * dx[3 * (i + 2) + 0];
  b2[0] = (xv[i2 + 1] - x) * dx[3 * (i + 1) + 1] * b1[0];
  b2[1] = ((x - xv[i2 - 1]) * dx[3 * (i + 1) + 1] * b1[0] + (xv[i2 + 2] - x) * dx[3 * (i + 2) + 1] * b1[1]);
  b2[2] = (x - xv[i2]) * dx[3 * (i + 2) + 1] * b1[1];
  b[0] = (xv[i2 + 1] - x) * dx[3 * (i) + 2] * b2[0];
  b[1] = ((x - xv[i2 - 2]) * dx[3 * (i) + 2] * b2[0] + (xv[i2 + 2] - x) * dx[3 * (i + 1) + 2] * b2[1]);
  b[2] = ((x - xv[i2 - 1]) * dx[3 * (i + 1) + 2] * b2[1] + (xv[i2 + 3] - x) * dx[3 * (i + 2) + 2] * b2[2]);
  b[3] = (x - xv[i2]) * dx[3 * (i + 2) + 2] * b2[2];
  if (d1) {
    d1[0] = -3.0 * (dx[3 * (i) + 2] * b2[0]);
    d1[1] = 3.0 * (dx[3 * (i) + 2] * b2[0] - dx[3 * (i + 1) + 2] * b2[1]);
    d1[2] = 3.0 * (dx[3 * (i + 1) + 2] * b2[1] - dx[3 * (i + 2) + 2] * b2[2]);
    d1[3] = 3.0 * (dx[3 * (i + 2) + 2] * b2[2]);
  }
  if (d2) {
    d2[0] = 6.0 * (+dx[3 * (i + 0) + 2] * dx[3 * (i + 1) + 1] * b1[0]);
    d2[1] = 6.0 * (-dx[3 * (i + 1) + 1] * (dx[3 * (i + 0) + 2] + dx[3 * (i + 1) + 2]) * b1[0] +
                   dx[3 * (i + 1) + 2] * dx[3 * (i + 2) + 1] * b1[1]);
    d2[2] = 6.0 * (+dx[3 * (i + 1) + 2] * dx[3 * (i + 1) + 1] * b1[0] -
                   dx[3 * (i + 2) + 1] * (dx[3 * (i + 1) + 2] + dx[3 * (i + 2) + 2]) * b1[1]);
    d2[3] = 6.0 * (+dx[3 * (i + 2) + 2] * dx[3 * (i + 2) + 1] * b1[1]);
  }
}

static int nub_index(const NUBasis *basis, double x) {
  return basis->grid->reverse_map(basis->grid, x);
}

int get_NUBasis_funcs_d(NUBasis *basis, double x, double bf[4]) {
  const int i = nub_index(basis, x);
  nub_funcs(basis, i, x, bf, NULL, NULL);
  return i;
}
void get_NUBasis_funcs_di(NUBasis *basis, int i, double bf[4]) {
  nub_funcs(basis, i, basis->grid->points[i], bf, NULL, NULL);
}
int get_NUBasis_dfuncs_d(NUBasis *basis, double x, double bf[4], double dbf[4]) {
  const int i = nub_index(basis, x);
  nub_funcs(basis, i, x, bf, dbf, NULL);
  return i;
}
void get_NUBasis_dfuncs_di(NUBasis *basis, int i, double bf[4], double dbf[4]) {
  nub_funcs(basis, i, basis->grid->points[i], bf, dbf, NULL);
}
int get_NUBasis_d2funcs_d(NUBasis *basis, double x, double bf[4], double dbf[4], double d2bf[4]) {
  const int i = nub_index(basis, x);
  nub_funcs(basis, i, x, bf, dbf, d2bf);
  return i;
}
void get_NUBasis_d2funcs_di(NUBasis *basis, int i, double bf[4], double dbf[4], double d2bf[4]) {
  nub_funcs(basis, i, basis->grid->points[i], bf, dbf, d2bf);
}

static void to_float(const double *a, float *b, int n) {
  for (int k = 0; k < n; ++k) b[k] = (float)a[k];
}
int get_NUBasis_funcs_s(NUBasis *basis, double x, float bf[4]) {
  double b[4];
  const int i = get_NUBasis_funcs_d(basis, x, b);
  to_float(b, bf, 4);
  return i;
}
void get_NUBasis_funcs_si(NUBasis *basis, int i, float bf[4]) {
  double b[4];
  get_NUBasis_funcs_di(basis, i, b);
  to_float(b, bf, 4);
}
int get_NUBasis_dfuncs_s(NUBasis *basis, double x, float bf[4], float dbf[4]) {
  double b[4], d[4];
  const int i = get_NUBasis_dfuncs_d(basis, x, b, d);
  to_float(b, bf, 4); to_float(d, dbf, 4);
  return i;
}
void get_NUBasis_dfuncs_si(NUBasis *basis, int i, float bf[4], float dbf[4]) {
  double b[4], d[4];
  get_NUBasis_dfuncs_di(basis, i, b, d);
  to_float(b, bf, 4); to_float(d, dbf, 4);
}
int get_NUBasis_d2funcs_s(NUBasis *basis, double x, float bf[4], float dbf[4], float d2bf[4]) {
  double b[4], d[4], e[4];
  const int i = get_NUBasis_d2funcs_d(basis, x, b, d, e);
  to_float(b, bf, 4); to_float(d, dbf, 4); to_float(e, d2bf, 4);
  return i;
}
void get_NUBasis_d2funcs_si(NUBasis *basis, int i, float bf[4], float dbf[4], float d2bf[4]) {
  double b[4], d[4], e[4];
  get_NUBasis_d2funcs_di(basis, i, b, d, e);
  to_float(b, bf, 4); to_float(d, dbf, 4); to_float(e, d2bf, 4);
}

/* ------------------------------------------------------------------------ */
/* 1-D solves (one real orbital: data[i*ds], coefficients p[i*ps])          */

/* Rows: the left condition, the M data rows, the right condition; each row
 * holds three band entries and the right-hand side (bands[4r .. 4r+3]). */
static int nub_solve_deriv(NUBasis *basis, const double *data, int ds, double *p, int ps,
                           const double left[4], const double right[4]) {
  const int M = basis->grid->num_points, N = M + 2;
  double *bd = malloc(sizeof(double) * 4 * (size_t)N);
  if (!bd) return -1;
  for (int k = 0; k < 4; ++k) {
    bd[k] = left[k];
    bd[4 * (N - 1) + k] = right[k];
  }
  for (int i = 0; i < M; ++i) {
    get_NUBasis_funcs_di(basis, i, bd + 4 * (i + 1));
    bd[4 * (i + 1) + 3] = data[(long)ds * i];
  }
  /* rows 0 and 1: the left condition spans three coefficients */
  bd[1] /= bd[0];
  bd[2] /= bd[0];
  bd[3] /= bd[0];
  bd[0] = 1.0;
  bd[4 + 1] -= bd[4 + 0] * bd[1];
  bd[4 + 2] -= bd[4 + 0] * bd[2];
  bd[4 + 3] -= bd[4 + 0] * bd[3];
  bd[0] = 0.0;
  bd[4 + 2] /= bd[4 + 1];
  bd[4 + 3] /= bd[4 + 1];
  bd[4 + 1] = 1.0;
  for (int r = 2; r < N - 1; ++r) {
    double *row = bd + 4 * r;
    const double *up = row - 4;
    row[1] -= row[0] * up[2];
    row[3] -= row[0] * up[3];
    row[2] /= row[1];
    row[3] /= row[1];
    row[0] = 0.0;
    row[1] = 1.0;
  }
  /* last row: the right condition spans the last three coefficients */
  double *last = bd + 4 * (M + 1);
  last[1] -= last[0] * bd[4 * (M - 1) + 2];
  last[3] -= last[0] * bd[4 * (M - 1) + 3];
  last[2] -= last[1] * bd[4 * M + 2];
  last[3] -= last[1] * bd[4 * M + 3];
  last[3] /= last[2];
  last[2] = 1.0;
  p[(long)ps * (M + 1)] = last[3];
  for (int r = M; r > 0; --r) p[(long)ps * r] = bd[4 * r + 3] - bd[4 * r + 2] * p[(long)ps * (r + 1)];
  p[0] = bd[3] - bd[1] * p[ps] - bd[2] * p[(long)ps * 2];
  free(bd);
  return 0;
}

/* Cyclic system of the M = num_points - 1 distinct points: eliminated down
 * the band while the wrap entries collect in a last column. */
static int nub_solve_periodic(NUBasis *basis, const double *data, int ds, double *p, int ps) {
  const int M = basis->grid->num_points - 1;
  double *bd = malloc(sizeof(double) * 4 * (size_t)M);
  double *lc = malloc(sizeof(double) * (size_t)M);
  if (!bd || !lc) { free(bd); free(lc); return -1; }
  for (int i = 0; i < M; ++i) {
    get_NUBasis_funcs_di(basis, i, bd + 4 * i);
    bd[4 * i + 3] = data[(long)i * ds];
  }
  double *lr = bd + 4 * (M - 1);         /* the last row */
  bd[2] /= bd[1];
  bd[0] /= bd[1];
  bd[3] /= bd[1];
  bd[1] = 1.0;
  lr[1] -= lr[2] * bd[0];
  lr[3] -= lr[2] * bd[3];
  lr[2] = -lr[2] * bd[2];
  lc[0] = bd[0];
  for (int r = 1; r < M - 1; ++r) {
    double *row = bd + 4 * r;
    const double *up = row - 4;
    row[1] -= row[0] * up[2];
    row[3] -= row[0] * up[3];
    lc[r] = -row[0] * lc[r - 1];
    row[0] = 0.0;
    row[2] /= row[1];
    row[3] /= row[1];
    lc[r] /= row[1];
    row[1] = 1.0;
    if (r < M - 2) {
      lr[3] -= lr[2] * row[3];
      lr[1] -= lr[2] * lc[r];
      lr[2] = -lr[2] * row[2];
    }
  }
  /* the last row's [2] and [0] entries now sit in the same column */
  lr[0] += lr[2];
  lr[1] -= lr[0] * (bd[4 * (M - 2) + 2] + lc[M - 2]);
  lr[3] -= lr[0] * bd[4 * (M - 2) + 3];
  lr[3] /= lr[1];
  p[(long)ps * M] = lr[3];
  for (int r = M - 2; r >= 0; --r)
    p[(long)ps * (r + 1)] = bd[4 * r + 3] - bd[4 * r + 2] * p[(long)ps * (r + 2)] - lc[r] * p[(long)ps * M];
  p[0] = p[(long)ps * M];
  p[(long)ps * (M + 1)] = p[ps];
  p[(long)ps * (M + 2)] = p[(long)ps * 2];
  free(bd);
  free(lc);
  return 0;
}

static int nub_find_coefs(NUBasis *basis, BCtype_d bc, const double *data, int ds, double *coefs, int cs) {
  if (bc.lCode == PERIODIC) return nub_solve_periodic(basis, data, ds, coefs, cs);
  const int M = basis->grid->num_points;
  double bf[4], dbf[4], left[4], right[4];
  if (bc.lCode == FLAT || bc.lCode == NATURAL) bc.lVal = 0.0;
  if (bc.rCode == FLAT || bc.rCode == NATURAL) bc.rVal = 0.0;
  if (bc.lCode == FLAT || bc.lCode == DERIV1) get_NUBasis_dfuncs_di(basis, 0, bf, left);
  else if (bc.lCode == NATURAL || bc.lCode == DERIV2) get_NUBasis_d2funcs_di(basis, 0, bf, dbf, left);
  else return -2;
  left[3] = bc.lVal;
  if (bc.rCode == FLAT || bc.rCode == DERIV1) get_NUBasis_dfuncs_di(basis, M - 1, bf, right);
  else if (bc.rCode == NATURAL || bc.rCode == DERIV2) get_NUBasis_d2funcs_di(basis, M - 1, bf, dbf, right);
  else return -2;
  right[3] = bc.rVal;
  return nub_solve_deriv(basis, data, ds, coefs, cs, left, right);
}

/* ------------------------------------------------------------------------ */
/* 1-D multi-splines                                                         */

einspline_nub1d *einspline_nub_of(void *spline) {
  if (!spline) return NULL;
  einspline_nub1d *h = (einspline_nub1d *)spline;
  return h->magic == EINSPLINE_NUB_MAGIC ? h : NULL;
}

static einspline_nub1d *nub_create(NUgrid *grid, int lcode, int ncomp, int num_splines, const char *fn) {
  if (!grid || num_splines < 1 || grid->num_points < 4) {
    fprintf(stderr, "%s: need a grid of at least 4 points and num_splines >= 1\n", fn);
    return NULL;
  }
  einspline_nub1d *h = calloc(1, sizeof(einspline_nub1d));
  if (!h) return NULL;
  NUBasis *basis = create_NUBasis(grid, lcode == PERIODIC);
  const int nrows = grid->num_points + 2;
  void *coefs = calloc((size_t)nrows * num_splines * ncomp, sizeof(double));
  if (!basis || !coefs) {
    destroy_NUBasis(basis);
    free(coefs);
    free(h);
    fprintf(stderr, "Out of memory in %s.\n", fn);
    return NULL;
  }
  h->magic = EINSPLINE_NUB_MAGIC;
  h->ncomp = ncomp;
  h->nrows = nrows;
  h->dirty = 1;
  pthread_mutex_init(&h->lock, NULL);
  /* the d and z structs differ after x_stride (BCtype_z is larger) */
  if (ncomp == 2) {
    multi_NUBspline_1d_z *z = &h->pub.z;
    z->spcode = MULTI_NU1D;
    z->tcode = DOUBLE_COMPLEX;
    z->coefs = (complex_double *)coefs;
    z->x_stride = num_splines;
    z->num_splines = num_splines;
    z->x_grid = grid;
    z->x_basis = basis;
  } else {
    multi_NUBspline_1d_d *d = &h->pub.d;
    d->spcode = MULTI_NU1D;
    d->tcode = DOUBLE_REAL;
    d->coefs = (double *)coefs;
    d->x_stride = num_splines;
    d->num_splines = num_splines;
    d->x_grid = grid;
    d->x_basis = basis;
  }
  return h;
}

multi_NUBspline_1d_d *create_multi_NUBspline_1d_d(NUgrid *x_grid, BCtype_d xBC, int num_splines) {
  einspline_nub1d *h = nub_create(x_grid, xBC.lCode, 1, num_splines, "create_multi_NUBspline_1d_d");
  if (!h) return NULL;
  h->pub.d.xBC = xBC;
  return &h->pub.d;
}

multi_NUBspline_1d_z *create_multi_NUBspline_1d_z(NUgrid *x_grid, BCtype_z xBC, int num_splines) {
  einspline_nub1d *h = nub_create(x_grid, xBC.lCode, 2, num_splines, "create_multi_NUBspline_1d_z");
  if (!h) return NULL;
  h->pub.z.xBC = xBC;
  return &h->pub.z;
}

static void mark_dirty(einspline_nub1d *h) {
  pthread_mutex_lock(&h->lock);
  h->dirty = 1;
  pthread_mutex_unlock(&h->lock);
}

static void set_real(void *spline, int n, const double *data, int ds, BCtype_d bc, int comp, const char *fn) {
  einspline_nub1d *h = einspline_nub_of(spline);
  if (!h) { fprintf(stderr, "%s: spline was not created by this library\n", fn); return; }
  const int ncomp = h->ncomp;
  const int nspl = ncomp == 2 ? h->pub.z.num_splines : h->pub.d.num_splines;
  NUBasis *basis = ncomp == 2 ? h->pub.z.x_basis : h->pub.d.x_basis;
  double *base = ncomp == 2 ? (double *)h->pub.z.coefs : h->pub.d.coefs;
  const intptr_t xs = ncomp == 2 ? h->pub.z.x_stride : h->pub.d.x_stride;
  if (n < 0 || n >= nspl) { fprintf(stderr, "%s: spline index out of range\n", fn); return; }
  double *coefs = base + (long)n * ncomp + comp;
  const int rc = nub_find_coefs(basis, bc, data, ds, coefs, (int)(xs * ncomp));
  if (rc == -1) { fprintf(stderr, "Out of memory in %s.\n", fn); abort(); }
  if (rc == -2) { fprintf(stderr, "%s: unsupported boundary condition\n", fn); return; }
  mark_dirty(h);
}

void set_multi_NUBspline_1d_d(multi_NUBspline_1d_d *spline, int num, double *data) {
  if (!spline) return;
  set_real(spline, num, data, 1, spline->xBC, 0, "set_multi_NUBspline_1d_d");
}

void set_multi_NUBspline_1d_d_BC(multi_NUBspline_1d_d *spline, int num, double *data, BCtype_d xBC) {
  set_real(spline, num, data, 1, xBC, 0, "set_multi_NUBspline_1d_d_BC");
}

static void set_cplx(multi_NUBspline_1d_z *spline, int num, complex_double *data, BCtype_z bc, const char *fn) {
  if (!spline) return;
  BCtype_d re = {bc.lCode, bc.rCode, bc.lVal_r, bc.rVal_r};
  BCtype_d im = {bc.lCode, bc.rCode, bc.lVal_i, bc.rVal_i};
  set_real(spline, num, (const double *)data, 2, re, 0, fn);
  set_real(spline, num, (const double *)data + 1, 2, im, 1, fn);
}

void set_multi_NUBspline_1d_z(multi_NUBspline_1d_z *spline, int num, complex_double *data) {
  if (!spline) return;
  set_cplx(spline, num, data, spline->xBC, "set_multi_NUBspline_1d_z");
}

void set_multi_NUBspline_1d_z_BC(multi_NUBspline_1d_z *spline, int num, complex_double *data, BCtype_z xBC) {
  set_cplx(spline, num, data, xBC, "set_multi_NUBspline_1d_z_BC");
}

void destroy_multi_NUBspline(void *spline) {
  einspline_nub1d *h = einspline_nub_of(spline);
  if (!h) return;
  if (h->mirror) einspline_nub_mirror_release(h->mirror);
  if (h->ncomp == 2) {
    destroy_NUBasis(h->pub.z.x_basis);
    free(h->pub.z.coefs);
  } else {
    destroy_NUBasis(h->pub.d.x_basis);
    free(h->pub.d.coefs);
  }
  pthread_mutex_destroy(&h->lock);
  h->magic = 0;
  free(h);
}
