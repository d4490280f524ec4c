/* nub_create.c -- non-uniform grids, their cubic B-spline basis, and the
 * 1-D non-uniform multi-splines (create / set / destroy).
 *
 * Interfaces and definitions follow the reference (src/nugrid.h, nubasis.h,
 * multi_nubspline_create.h): the same grid point formulas, the same knot
 * sequence (grid points with two knots before and three after, the end
 * spacings repeated, or wrapped for periodic bases), the same coefficient
 * layout and boundary conditions.  The computations are this library's:
 *   - reverse maps: a closed-form first guess for center and log grids,
 *     corrected against the stored points (so the interval brackets x even
 *     where log / log1p round), bisection on general grids;
 *   - basis: de Boor's triangular scheme (nub_basis.h), knot differences
 *     formed on the fly;
 *   - interpolation: the collocation system is factored once per set call
 *     (Thomas factors of its tridiagonal part after folding the derivative
 *     rows' corner entries; periodic systems by bordering the last unknown)
 *     and applied to every real component of the data.
 * Results agree with the reference's to rounding, not bit for bit.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "einspline_private.h"
#include "nub_basis.h"

/* ------------------------------------------------------------------------ */
/* grids                                                                     */

static int general_interval(void *gp, double x) {
  const NUgrid *g = (const NUgrid *)gp;
  return esk_nub_interval(g->points, g->num_points, x, -1);
}

/* floor(f) as an index guess in [0, n], 0 for NaN */
static int guess_of(double f, int n) {
  if (!(f == f) || f < 0.0) return 0;
  return f > (double)n ? n : (int)f;
}

static int center_interval(void *gp, double x) {
  const center_grid *g = (const center_grid *)gp;
  /* invert x = center +- a expm1(b |k - half_points + even_half|) */
  const double u = x - g->center;
  const double k = copysign(log1p(fabs(u) * g->aInv) * g->bInv, u) + g->half_points - g->even_half;
  return esk_nub_interval(g->points, g->num_points, x, guess_of(floor(k), g->num_points));
}

static int log_interval(void *gp, double x) {
  const log_grid *g = (const log_grid *)gp;
  /* invert x = start exp(a i) */
  const double k = x > 0.0 ? log(x * g->startinv) * g->ainv : -1.0;
  return esk_nub_interval(g->points, g->num_points, x, guess_of(floor(k), g->num_points));
}

static NUgrid *grid_alloc(size_t bytes, int n) {
  NUgrid *g = calloc(1, bytes);
  if (!g) return NULL;
  g->points = malloc(sizeof(double) * (size_t)n);
  if (!g->points) { free(g); return NULL; }
  g->num_points = n;
  return g;
}

NUgrid *create_general_grid(double *points, int num_points) {
  if (!points || num_points < 2) return NULL;
  for (int i = 1; i < num_points; ++i)
    if (!(points[i] > points[i - 1])) return NULL;      /* strictly increasing */
  NUgrid *g = grid_alloc(sizeof(NUgrid), num_points);
  if (!g) return NULL;
  memcpy(g->points, points, sizeof(double) * (size_t)num_points);
  g->code = GENERAL;
  g->start = points[0];
  g->end = points[num_points - 1];
  g->reverse_map = general_interval;
  return g;
}

/* Points at center + sign(k) a expm1(b |k|): k = -h..h for an odd count,
 * k = -h+1/2 .. h-1/2 for an even one, h = num_points / 2; b = log(ratio) /
 * (h - 1) and a puts the outermost points at start and end. */
NUgrid *create_center_grid(double start, double end, double ratio, int num_points) {
  if (!(ratio > 1.0) || num_points < 4 || !(end > start)) return NULL;
  center_grid *g = (center_grid *)grid_alloc(sizeof(center_grid), num_points);
  if (!g) return NULL;
  const int h = num_points / 2;
  g->code = CENTER;
  g->start = start;
  g->end = end;
  g->center = 0.5 * (start + end);
  g->half_points = h;
  g->odd = (num_points & 1) != 0;
  g->odd_one = g->odd ? 1 : 0;
  g->even_half = g->odd ? 0.0 : 0.5;
  g->b = log(ratio) / (double)(h - 1);
  g->bInv = 1.0 / g->b;
  const double kmax = (double)h - g->even_half;           /* |k| of the outermost points */
  g->a = 0.5 * (end - start) / expm1(g->b * kmax);
  g->aInv = 1.0 / g->a;
  for (int i = 0; i < num_points; ++i) {
    const double k = (double)(i - h) + g->even_half;       /* signed position from the centre */
    const double r = g->a * expm1(g->b * fabs(k));
    g->points[i] = (k < 0.0 ? -r : r) + g->center;
  }
  g->reverse_map = center_interval;
  return (NUgrid *)g;
}

/* points[i] = start exp(a i), a = log(end / start) / (num_points - 1). */
NUgrid *create_log_grid(double start, double end, int num_points) {
  if (!(start > 0.0) || !(end > start) || num_points < 2) return NULL;
  log_grid *g = (log_grid *)grid_alloc(sizeof(log_grid), num_points);
  if (!g) return NULL;
  g->code = LOG;
  g->start = start;
  g->end = end;
  g->a = log(end / start) / (double)(num_points - 1);
  g->ainv = 1.0 / g->a;
  g->startinv = 1.0 / start;
  for (int i = 0; i < num_points; ++i) g->points[i] = start * exp(g->a * (double)i);
  g->reverse_map = log_interval;
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
  double *kn = b->xVals;
  memcpy(kn + 2, g, sizeof(double) * (size_t)N);
  /* two knots before the grid and three after: the end spacing repeated,
   * or, for a periodic basis, the spacings of the other end */
  for (int k = 1; k <= 2; ++k)
    kn[2 - k] = g[0] - (periodic ? g[N - 1] - g[N - 1 - k] : (double)k * (g[1] - g[0]));
  for (int k = 1; k <= 3; ++k)
    kn[N + 1 + k] = g[N - 1] + (periodic ? g[k] - g[0] : (double)k * (g[N - 1] - g[N - 2]));
  /* the documented table of inverse knot spans (not used by the evaluation) */
  for (int i = 0; i < N + 2; ++i)
    for (int j = 0; j < 3; ++j) b->dxInv[3 * i + j] = 1.0 / (kn[i + j + 1] - kn[i]);
  return b;
}

void destroy_NUBasis(NUBasis *basis) {
  if (!basis) return;
  free(basis->xVals);
  free(basis->dxInv);
  free(basis);
}

/* Interval of x and the functions (and derivatives up to nder) there. */
static int basis_at(const NUBasis *basis, double x, int nder, double out[3][4]) {
  const int i = basis->grid->reverse_map(basis->grid, x);
  esk_nub_basis(basis->xVals, i, x, nder, out);
  return i;
}

static void basis_at_point(const NUBasis *basis, int i, int nder, double out[3][4]) {
  esk_nub_basis(basis->xVals, i, basis->grid->points[i], nder, out);
}

static void copy4d(double *dst, const double *src) { memcpy(dst, src, 4 * sizeof(double)); }
static void copy4f(float *dst, const double *src) {
  for (int k = 0; k < 4; ++k) dst[k] = (float)src[k];
}

int get_NUBasis_funcs_d(NUBasis *basis, double x, double bf[4]) {
  double o[3][4];
  const int i = basis_at(basis, x, 0, o);
  copy4d(bf, o[0]);
  return i;
}
void get_NUBasis_funcs_di(NUBasis *basis, int i, double bf[4]) {
  double o[3][4];
  basis_at_point(basis, i, 0, o);
  copy4d(bf, o[0]);
}
int get_NUBasis_dfuncs_d(NUBasis *basis, double x, double bf[4], double dbf[4]) {
  double o[3][4];
  const int i = basis_at(basis, x, 1, o);
  copy4d(bf, o[0]); copy4d(dbf, o[1]);
  return i;
}
void get_NUBasis_dfuncs_di(NUBasis *basis, int i, double bf[4], double dbf[4]) {
  double o[3][4];
  basis_at_point(basis, i, 1, o);
  copy4d(bf, o[0]); copy4d(dbf, o[1]);
}
int get_NUBasis_d2funcs_d(NUBasis *basis, double x, double bf[4], double dbf[4], double d2bf[4]) {
  double o[3][4];
  const int i = basis_at(basis, x, 2, o);
  copy4d(bf, o[0]); copy4d(dbf, o[1]); copy4d(d2bf, o[2]);
  return i;
}
void get_NUBasis_d2funcs_di(NUBasis *basis, int i, double bf[4], double dbf[4], double d2bf[4]) {
  double o[3][4];
  basis_at_point(basis, i, 2, o);
  copy4d(bf, o[0]); copy4d(dbf, o[1]); copy4d(d2bf, o[2]);
}

int get_NUBasis_funcs_s(NUBasis *basis, double x, float bf[4]) {
  double o[3][4];
  const int i = basis_at(basis, x, 0, o);
  copy4f(bf, o[0]);
  return i;
}
void get_NUBasis_funcs_si(NUBasis *basis, int i, float bf[4]) {
  double o[3][4];
  basis_at_point(basis, i, 0, o);
  copy4f(bf, o[0]);
}
int get_NUBasis_dfuncs_s(NUBasis *basis, double x, float bf[4], float dbf[4]) {
  double o[3][4];
  const int i = basis_at(basis, x, 1, o);
  copy4f(bf, o[0]); copy4f(dbf, o[1]);
  return i;
}
void get_NUBasis_dfuncs_si(NUBasis *basis, int i, float bf[4], float dbf[4]) {
  double o[3][4];
  basis_at_point(basis, i, 1, o);
  copy4f(bf, o[0]); copy4f(dbf, o[1]);
}
int get_NUBasis_d2funcs_s(NUBasis *basis, double x, float bf[4], float dbf[4], float d2bf[4]) {
  double o[3][4];
  const int i = basis_at(basis, x, 2, o);
  copy4f(bf, o[0]); copy4f(dbf, o[1]); copy4f(d2bf, o[2]);
  return i;
}
void get_NUBasis_d2funcs_si(NUBasis *basis, int i, float bf[4], float dbf[4], float d2bf[4]) {
  double o[3][4];
  basis_at_point(basis, i, 2, o);
  copy4f(bf, o[0]); copy4f(dbf, o[1]); copy4f(d2bf, o[2]);
}

/* ------------------------------------------------------------------------ */
/* interpolation: the collocation operator, factored once per set call      */

/* Interpolation at the grid points of a basis with boundary rows (end
 * derivatives) or a periodic wrap.  Unknowns: the M + 2 coefficients
 * c_0 .. c_{M+1} (derivative conditions, M = num_points) or the M =
 * num_points - 1 distinct ones c_1 .. c_M (periodic).
 *   Derivative conditions: the boundary rows L (columns 0..2) and R
 *   (columns M-1..M+1) are pivot rows for the outer columns: data row 0
 *   minus a multiple of L loses column 0, data row M-1 minus a multiple of
 *   R loses column M+1, leaving a tridiagonal system in c_1 .. c_M; c_0 and
 *   c_{M+1} follow from L and R.  (L and R never have a zero outer entry:
 *   the outermost function of an end interval has non-zero first and second
 *   derivatives at the grid's end point.)
 *   Periodic: a cyclic tridiagonal system, solved by bordering its last
 *   unknown.
 * The tridiagonal part is kept as Thomas factors: multiplier lo[r], inverse
 * pivot iv[r], super entry up[r]. */
typedef struct {
  int periodic;
  int M, n;            /* data points; unknowns of the tridiagonal part */
  double *lo, *iv, *up;
  /* derivative conditions: the boundary rows and the multiples of them
   * taken from data rows 0 and M-1 */
  double L[3], R[3], ml, mr;
  /* periodic bordering: w = T^-1 p, the last row's entries q0 (col 0) and
   * qe (col n-1), 1 / Schur complement */
  double *w, q0, qe, zinv;
} nub_op;

static void op_free(nub_op *op) {
  free(op->lo);
  free(op->iv);
  free(op->up);
  free(op->w);
}

/* Thomas factors of the tridiagonal rows (sub, dia, sup) of length n, in place:
 * sub becomes the multipliers, dia the inverse pivots. */
static void thomas_factor(int n, double *sub, double *dia, const double *sup) {
  dia[0] = 1.0 / dia[0];
  for (int r = 1; r < n; ++r) {
    sub[r] *= dia[r - 1];
    dia[r] = 1.0 / (dia[r] - sub[r] * sup[r - 1]);
  }
}

/* Solve in place for x[r * xs], r < n. */
static void thomas_apply(int n, const double *lo, const double *iv, const double *up, double *x, long xs) {
  for (int r = 1; r < n; ++r) x[r * xs] -= lo[r] * x[(r - 1) * xs];
  x[(n - 1) * xs] *= iv[n - 1];
  for (int r = n - 2; r >= 0; --r) x[r * xs] = (x[r * xs] - up[r] * x[(r + 1) * xs]) * iv[r];
}

/* The boundary row of a derivative condition at grid point `pt` (interval
 * `pt`): the order-1 (DERIV1, FLAT) or order-2 (DERIV2, NATURAL) basis
 * derivatives; -1 for a code that is not a derivative condition. */
static int deriv_row(const NUBasis *basis, int code, int pt, double row[3]) {
  int order;
  if (code == DERIV1 || code == FLAT) order = 1;
  else if (code == DERIV2 || code == NATURAL) order = 2;
  else return -1;
  double o[3][4];
  basis_at_point(basis, pt, order, o);
  for (int k = 0; k < 3; ++k) row[k] = o[order][k];
  return 0;
}

static int op_build(nub_op *op, const NUBasis *basis, int lcode, int rcode) {
  memset(op, 0, sizeof *op);
  const int P = basis->grid->num_points;
  op->periodic = lcode == PERIODIC;
  op->M = op->periodic ? P - 1 : P;
  const int M = op->M;
  op->n = op->periodic ? M - 1 : M;
  const int n = op->n;
  op->lo = calloc((size_t)n + 1, sizeof(double));
  op->iv = calloc((size_t)n + 1, sizeof(double));
  op->up = calloc((size_t)n + 1, sizeof(double));
  if (op->periodic) op->w = calloc((size_t)n + 1, sizeof(double));
  if (!op->lo || !op->iv || !op->up || (op->periodic && !op->w)) { op_free(op); return -1; }
  double o[3][4];
  if (!op->periodic) {
    if (deriv_row(basis, lcode, 0, op->L) || deriv_row(basis, rcode, M - 1, op->R)) { op_free(op); return -2; }
    /* core row r (unknown c_{r+1}) is data row r: b0 c_r + b1 c_{r+1} + b2 c_{r+2} */
    for (int r = 0; r < M; ++r) {
      basis_at_point(basis, r, 0, o);
      op->lo[r] = o[0][0];
      op->iv[r] = o[0][1];
      op->up[r] = o[0][2];
    }
    /* data row 0 minus ml L has no c_0; data row M-1 minus mr R no c_{M+1} */
    op->ml = op->lo[0] / op->L[0];
    op->iv[0] -= op->ml * op->L[1];
    op->up[0] -= op->ml * op->L[2];
    op->lo[0] = 0.0;
    op->mr = op->up[M - 1] / op->R[2];
    op->lo[M - 1] -= op->mr * op->R[0];
    op->iv[M - 1] -= op->mr * op->R[1];
    op->up[M - 1] = 0.0;
    thomas_factor(n, op->lo, op->iv, op->up);
    return 0;
  }
  /* periodic: row j is b0 u_{j-1} + b1 u_j + b2 u_{j+1} (indices mod M) */
  double sub_last = 0.0, dia_last = 1.0, sup_last = 0.0;
  for (int j = 0; j < M; ++j) {
    basis_at_point(basis, j, 0, o);
    if (j < n) {
      op->lo[j] = o[0][0];
      op->iv[j] = o[0][1];
      op->up[j] = o[0][2];
    } else {
      sub_last = o[0][0];
      dia_last = o[0][1];
      sup_last = o[0][2];
    }
  }
  /* the last unknown's column in rows 0 .. n-1: row 0's wrap and row n-1's super entry */
  op->w[0] = op->lo[0];
  op->w[n - 1] += op->up[n - 1];
  op->q0 = sup_last;            /* last row, column 0 (wrap) */
  op->qe = sub_last;            /* last row, column n-1 */
  op->lo[0] = 0.0;
  op->up[n - 1] = 0.0;
  thomas_factor(n, op->lo, op->iv, op->up);
  thomas_apply(n, op->lo, op->iv, op->up, op->w, 1);
  op->zinv = 1.0 / (dia_last - op->q0 * op->w[0] - op->qe * op->w[n - 1]);
  return 0;
}

/* Coefficients p[r * ps] (num_points + 2 of them) of one real component:
 * data[j * ds] at the grid points, end values lval / rval. */
static void op_apply(const nub_op *op, const double *data, long ds, double lval, double rval, double *p,
                     long ps) {
  const int M = op->M, n = op->n;
  if (!op->periodic) {
    double *u = p + ps;             /* c_1 .. c_M */
    for (int j = 0; j < M; ++j) u[j * ps] = data[j * ds];
    u[0] -= op->ml * lval;
    u[(long)(M - 1) * ps] -= op->mr * rval;
    thomas_apply(n, op->lo, op->iv, op->up, u, ps);
    p[0] = (lval - op->L[1] * p[ps] - op->L[2] * p[2 * ps]) / op->L[0];
    p[(long)(M + 1) * ps] =
        (rval - op->R[0] * p[(long)(M - 1) * ps] - op->R[1] * p[(long)M * ps]) / op->R[2];
    return;
  }
  /* u_r = c_{r+1}: solve the bordered system on p[ps ..] */
  double *u = p + ps;
  for (int j = 0; j < n; ++j) u[j * ps] = data[j * ds];
  thomas_apply(n, op->lo, op->iv, op->up, u, ps);
  const double last = (data[(long)(M - 1) * ds] - op->q0 * u[0] - op->qe * u[(n - 1) * ps]) * op->zinv;
  for (int j = 0; j < n; ++j) u[j * ps] -= op->w[j] * last;
  u[(long)(M - 1) * ps] = last;
  /* the wrapped copies: c_0 = c_M, c_{M+1} = c_1, c_{M+2} = c_2 */
  p[0] = p[(long)M * ps];
  p[(long)(M + 1) * ps] = p[ps];
  p[(long)(M + 2) * ps] = p[2 * ps];
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

/* Orbital n from `ncomp`-component data (interleaved): one factorisation,
 * applied to each real component with its own end values. */
static void set_orbital(void *spline, int n, const double *data, int lcode, int rcode, const double *lval,
                        const double *rval, const char *fn) {
  einspline_nub1d *h = einspline_nub_of(spline);
  if (!h) { fprintf(stderr, "%s: spline was not created by this library\n", fn); return; }
  const int ncomp = h->ncomp;
  const int nspl = ncomp == 2 ? h->pub.z.num_splines : h->pub.d.num_splines;
  NUBasis *basis = ncomp == 2 ? h->pub.z.x_basis : h->pub.d.x_basis;
  double *base = ncomp == 2 ? (double *)h->pub.z.coefs : h->pub.d.coefs;
  const intptr_t xs = ncomp == 2 ? h->pub.z.x_stride : h->pub.d.x_stride;
  if (n < 0 || n >= nspl) { fprintf(stderr, "%s: spline index out of range\n", fn); return; }
  if ((lcode == PERIODIC) != (rcode == PERIODIC) || (lcode == PERIODIC) != basis->periodic) {
    fprintf(stderr, "%s: boundary conditions do not match the spline's periodicity\n", fn);
    return;
  }
  nub_op op;
  const int rc = op_build(&op, basis, lcode, rcode);
  if (rc == -1) { fprintf(stderr, "Out of memory in %s.\n", fn); abort(); }
  if (rc == -2) { fprintf(stderr, "%s: unsupported boundary condition\n", fn); return; }
  for (int c = 0; c < ncomp; ++c) {
    /* FLAT and NATURAL fix a zero end derivative */
    const double lv = (lcode == FLAT || lcode == NATURAL) ? 0.0 : lval[c];
    const double rv = (rcode == FLAT || rcode == NATURAL) ? 0.0 : rval[c];
    op_apply(&op, data + c, ncomp, lv, rv, base + (long)n * ncomp + c, (long)xs * ncomp);
  }
  op_free(&op);
  mark_dirty(h);
}

static void set_real(multi_NUBspline_1d_d *spline, int num, const double *data, BCtype_d bc, const char *fn) {
  if (!spline || !data) return;
  set_orbital(spline, num, data, bc.lCode, bc.rCode, &bc.lVal, &bc.rVal, fn);
}

void set_multi_NUBspline_1d_d(multi_NUBspline_1d_d *spline, int num, double *data) {
  if (!spline) return;
  set_real(spline, num, data, spline->xBC, "set_multi_NUBspline_1d_d");
}

void set_multi_NUBspline_1d_d_BC(multi_NUBspline_1d_d *spline, int num, double *data, BCtype_d xBC) {
  set_real(spline, num, data, xBC, "set_multi_NUBspline_1d_d_BC");
}

static void set_cplx(multi_NUBspline_1d_z *spline, int num, complex_double *data, BCtype_z bc, const char *fn) {
  if (!spline || !data) return;
  const double lv[2] = {bc.lVal_r, bc.lVal_i}, rv[2] = {bc.rVal_r, bc.rVal_i};
  set_orbital(spline, num, (const double *)data, bc.lCode, bc.rCode, lv, rv, fn);
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
