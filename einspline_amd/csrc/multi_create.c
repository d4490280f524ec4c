/* multi_create.c -- host construction of uniform multi-splines.
 *
 * Implements create_multi_UBspline_{1,2,3}d_{s,d,c,z},
 * set_multi_UBspline_{1,2,3}d_* (+ 3-D block variants, 1-D _BC variants)
 * and destroy_Bspline with the behaviour of src/multi_bspline_create.c:
 * 88-373 (s), :389-732 (c), :767-1073 (d), :1091-1477 (z), :1480-1485
 * (destroy).
 *
 * Interpolation conditions (one axis, M data points d_0..d_{M-1}):
 *   every data point:  c_{i}/6 + 2 c_{i+1}/3 + c_{i+2}/6 = d_i
 *   PERIODIC:          c_0 = c_M, c_{M+1} = c_1, c_{M+2} = c_2      (M+3 coefs)
 *   ANTIPERIODIC:      same with a sign flip on every wrapped coefficient
 *   DERIV1 / FLAT:     (c_2 - c_0) * delta_inv / 2          = value  (M+2 coefs)
 *   DERIV2 / NATURAL:  (c_0 - 2 c_1 + c_2) * delta_inv^2    = value
 * (src/bspline_create.c:1115-1198 sets up these rows; FLAT and NATURAL
 * force value = 0.)  A 2-D or 3-D table is the tensor product: the 1-D
 * problem is solved along x, then y, then z.  A 1-D or 2-D table is held as
 * a 3-D one whose missing axes have extent 1 and are not solved.
 *
 * The solve is organised differently from the reference (which eliminates
 * a banded matrix per line, src/bspline_create.c:97-145, :958-1055): each
 * axis operator is factorised once per set call (Thomas factors of the
 * tridiagonal core, plus a bordering vector for the periodic wrap), and the
 * factors are then applied to whole planes of right-hand sides at a time,
 * so the inner loops run over contiguous memory.  Everything is in double
 * precision; float tables are rounded once at the end.
 */
#include <stddef.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>

#include "multi_bspline_create.h"
#include "einspline_private.h"

#define ALPHA (1.0 / 6.0)
#define BETA  (2.0 / 3.0)

/* ------------------------------------------------------------------------ */
/* one axis: factorised interpolation operator                               */

typedef struct {
  int M;          /* data points */
  int n;          /* coefficients along the axis */
  int kind;       /* 0 periodic, 1 antiperiodic, 2 derivative conditions */
  /* Thomas factors of the tridiagonal core (length M-1 periodic, M deriv) */
  double *inv;    /* 1 / pivot of each row */
  double *g;      /* super-diagonal / pivot */
  double *sub;    /* sub-diagonal of each row */
  /* periodic bordering */
  double *w;      /* core^{-1} * (coupling column of the last unknown) */
  double sgn;     /* +1 periodic, -1 antiperiodic */
  double zinv;    /* 1 / Schur complement of the last unknown */
  /* derivative conditions: coefficients of the left/right rows */
  double l0, l1, l2, r0, r1, r2;
  int lcode, rcode;
} axis_op;

static void axis_free(axis_op *op) {
  free(op->inv); free(op->g); free(op->sub); free(op->w);
  memset(op, 0, sizeof(*op));
}

/* Forward/back substitution with precomputed factors on one column of
 * length m (rhs overwritten by the solution). */
static void thomas1(const axis_op *op, int m, double *x) {
  x[0] *= op->inv[0];
  for (int r = 1; r < m; ++r) x[r] = (x[r] - op->sub[r] * x[r - 1]) * op->inv[r];
  for (int r = m - 2; r >= 0; --r) x[r] -= op->g[r] * x[r + 1];
}

static int bc_rows(int code, double dinv, double *c0, double *c1, double *c2) {
  switch (code) {
    case FLAT: case DERIV1:
      *c0 = -0.5 * dinv; *c1 = 0.0; *c2 = 0.5 * dinv; return 0;
    case NATURAL: case DERIV2:
      *c0 = dinv * dinv; *c1 = -2.0 * dinv * dinv; *c2 = dinv * dinv; return 0;
    default:
      return -1;
  }
}

static int axis_init(axis_op *op, const Ugrid *grid, int lcode, int rcode) {
  memset(op, 0, sizeof(*op));
  const int M = grid->num;
  op->M = M;
  op->lcode = lcode; op->rcode = rcode;
  if (lcode == PERIODIC || lcode == ANTIPERIODIC) {
    op->kind = lcode == PERIODIC ? 0 : 1;
    op->sgn = lcode == PERIODIC ? 1.0 : -1.0;
    op->n = M + 3;
    const int m = M - 1;               /* core: unknowns x_0..x_{M-2} */
    op->inv = calloc((size_t)m, sizeof(double));
    op->g = calloc((size_t)m, sizeof(double));
    op->sub = calloc((size_t)m, sizeof(double));
    op->w = calloc((size_t)m, sizeof(double));
    if (!op->inv || !op->g || !op->sub || !op->w) { axis_free(op); return -2; }
    for (int r = 0; r < m; ++r) {
      op->sub[r] = r ? ALPHA : 0.0;
      const double piv = BETA - (r ? ALPHA * op->g[r - 1] : 0.0);
      op->inv[r] = 1.0 / piv;
      op->g[r] = ALPHA / piv;
    }
    /* coupling of x_{M-1} into the core rows: row 0 through the wrap,
     * row M-2 as its ordinary neighbour */
    op->w[0] += op->sgn * ALPHA;
    op->w[m - 1] += ALPHA;
    thomas1(op, m, op->w);
    const double schur = BETA - ALPHA * op->w[m - 1] - op->sgn * ALPHA * op->w[0];
    op->zinv = 1.0 / schur;
    return 0;
  }
  /* derivative-type conditions on both ends */
  op->kind = 2;
  op->n = M + 2;
  const double dinv = grid->delta_inv;
  if (bc_rows(lcode, dinv, &op->l0, &op->l1, &op->l2) ||
      bc_rows(rcode, dinv, &op->r0, &op->r1, &op->r2))
    return -1;
  /* Core unknowns c_1..c_M.  Row 1: left condition with c_0 eliminated by
   * the first data row; row M: right condition with c_{M+1} eliminated by
   * the last data row; rows 2..M-1: data rows. */
  const int m = M;
  op->inv = calloc((size_t)m, sizeof(double));
  op->g = calloc((size_t)m, sizeof(double));
  op->sub = calloc((size_t)m, sizeof(double));
  if (!op->inv || !op->g || !op->sub) { axis_free(op); return -2; }
  for (int r = 0; r < m; ++r) {
    double sub = ALPHA, dia = BETA, sup = ALPHA;
    if (r == 0) { sub = 0.0; dia = op->l1 - op->l0 * BETA / ALPHA; sup = op->l2 - op->l0; }
    if (r == m - 1) { sub = op->r0 - op->r2; dia = op->r1 - op->r2 * BETA / ALPHA; sup = 0.0; }
    const double piv = dia - (r ? sub * op->g[r - 1] : 0.0);
    op->sub[r] = sub;
    op->inv[r] = 1.0 / piv;
    op->g[r] = sup / piv;
  }
  return 0;
}

/* Apply the operator along one axis.  Element e (0..n-1) of the axis is the
 * vector G[e*es .. e*es+len) of independent right-hand sides; on entry
 * elements 1..M hold the data, on exit elements 0..n-1 hold coefficients.
 * lval/rval give the derivative-condition values for component (q % ncomp)
 * of a vector position q. */
static void axis_apply(const axis_op *op, double *G, ptrdiff_t es, ptrdiff_t len,
                       const double *lval, const double *rval, int ncomp) {
  const int M = op->M;
#define EL(e) (G + (ptrdiff_t)(e) * es)
  if (op->kind != 2) {
    const int m = M - 1;
    /* core solve on elements 1..M-1 */
    {
      double *x0 = EL(1);
      const double i0 = op->inv[0];
      for (ptrdiff_t q = 0; q < len; ++q) x0[q] *= i0;
    }
    for (int r = 1; r < m; ++r) {
      double *xr = EL(r + 1);
      const double *xp = EL(r);
      const double s = op->sub[r], iv = op->inv[r];
      for (ptrdiff_t q = 0; q < len; ++q) xr[q] = (xr[q] - s * xp[q]) * iv;
    }
    for (int r = m - 2; r >= 0; --r) {
      double *xr = EL(r + 1);
      const double *xn = EL(r + 2);
      const double gg = op->g[r];
      for (ptrdiff_t q = 0; q < len; ++q) xr[q] -= gg * xn[q];
    }
    /* last unknown x_{M-1} (element M), then correct the core */
    {
      double *z = EL(M);
      const double *y0 = EL(1), *ye = EL(m);
      const double sa = op->sgn * ALPHA, zi = op->zinv;
      for (ptrdiff_t q = 0; q < len; ++q) z[q] = (z[q] - ALPHA * ye[q] - sa * y0[q]) * zi;
      for (int r = 0; r < m; ++r) {
        double *xr = EL(r + 1);
        const double wr = op->w[r];
        for (ptrdiff_t q = 0; q < len; ++q) xr[q] -= wr * z[q];
      }
    }
    /* wrapped copies */
    {
      const double s = op->sgn;
      double *c0 = EL(0), *cm1 = EL(M + 1), *cm2 = EL(M + 2);
      const double *cm = EL(M), *c1 = EL(1), *c2 = EL(2);
      for (ptrdiff_t q = 0; q < len; ++q) {
        c0[q] = s * cm[q];
        cm1[q] = s * c1[q];
        cm2[q] = s * c2[q];
      }
    }
    return;
  }
  /* derivative conditions */
  {
    double *e0 = EL(0), *e1 = EL(1), *eM = EL(M), *eM1 = EL(M + 1);
    const double kl = op->l0 / ALPHA, kr = op->r2 / ALPHA;
    for (ptrdiff_t q = 0; q < len; ++q) {
      const double d0 = e1[q], dl = eM[q];
      e0[q] = d0;        /* kept to recover c_0 */
      eM1[q] = dl;       /* kept to recover c_{M+1} */
      e1[q] = lval[q % ncomp] - kl * d0;
      eM[q] = rval[q % ncomp] - kr * dl;
    }
  }
  {
    double *x0 = EL(1);
    const double i0 = op->inv[0];
    for (ptrdiff_t q = 0; q < len; ++q) x0[q] *= i0;
  }
  for (int r = 1; r < M; ++r) {
    double *xr = EL(r + 1);
    const double *xp = EL(r);
    const double s = op->sub[r], iv = op->inv[r];
    for (ptrdiff_t q = 0; q < len; ++q) xr[q] = (xr[q] - s * xp[q]) * iv;
  }
  for (int r = M - 2; r >= 0; --r) {
    double *xr = EL(r + 1);
    const double *xn = EL(r + 2);
    const double gg = op->g[r];
    for (ptrdiff_t q = 0; q < len; ++q) xr[q] -= gg * xn[q];
  }
  {
    double *e0 = EL(0), *eM1 = EL(M + 1);
    const double *c1 = EL(1), *c2 = EL(2), *cM = EL(M), *cMm = EL(M - 1);
    for (ptrdiff_t q = 0; q < len; ++q) {
      e0[q] = (e0[q] - BETA * c1[q] - ALPHA * c2[q]) / ALPHA;
      eM1[q] = (eM1[q] - ALPHA * cMm[q] - BETA * cM[q]) / ALPHA;
    }
  }
#undef EL
}

/* ------------------------------------------------------------------------ */
/* objects                                                                   */

einspline_host3d *einspline_host_of(void *spline) {
  if (!spline) return NULL;
  einspline_host3d *h = (einspline_host3d *)spline;
  if (h->magic != EINSPLINE_HOST_MAGIC) return NULL;
  return h;
}

static int axis_extent(const Ugrid *g, bc_code l, int *N) {
  if (g->num < 2) return -1;
  if (!(g->end > g->start) && !(g->end < g->start)) return -1;
  *N = (l == PERIODIC || l == ANTIPERIODIC) ? g->num + 3 : g->num + 2;
  return 0;
}

static void fill_grid(Ugrid *g, int N) {
  g->delta = (g->end - g->start) / (double)(N - 3);
  g->delta_inv = 1.0 / g->delta;
}

/* Shared part of the create functions.  `pub` points at the public struct
 * inside the returned object; the caller fills BCs and the typed coefs
 * pointer from `table`.  Axes past `dim` pass NULL grids (extent 1). */
static einspline_host3d *create_common(const char *fn, int dim, Ugrid *xg, Ugrid *yg, Ugrid *zg,
                                       bc_code xl, bc_code yl, bc_code zl, int num_splines,
                                       size_t elt, int ncomp, void **table) {
  int Nx = 1, Ny = 1, Nz = 1;
  if (num_splines < 1 || axis_extent(xg, xl, &Nx) || (dim >= 2 && axis_extent(yg, yl, &Ny)) ||
      (dim >= 3 && axis_extent(zg, zl, &Nz))) {
    fprintf(stderr, "%s: invalid grid (num >= 2, end != start) or num_splines < 1\n", fn);
    return NULL;
  }
  fill_grid(xg, Nx);
  if (dim >= 2) fill_grid(yg, Ny);
  if (dim >= 3) fill_grid(zg, Nz);
  einspline_host3d *h = calloc(1, sizeof(*h));
  if (!h) {
    fprintf(stderr, "Out of memory allocating spline in %s.\n", fn);
    abort();
  }
  const size_t bytes = elt * (size_t)Nx * Ny * Nz * (size_t)num_splines;
  void *t = NULL;
  if (posix_memalign(&t, 256, bytes ? bytes : 256) != 0 || !t) {
    fprintf(stderr, "Out of memory allocating spline coefficients in %s.\n", fn);
    abort();
  }
  memset(t, 0, bytes);
  *table = t;
  h->magic = EINSPLINE_HOST_MAGIC;
  h->dim = dim;
  h->Nx = Nx; h->Ny = Ny; h->Nz = Nz;
  h->ncomp = ncomp;
  h->nsplines = num_splines;
  h->dirty = 1;
  h->mirror = NULL;
  pthread_mutex_init(&h->lock, NULL);
  return h;
}

#define SET_COMMON_FIELDS(S, TC)                                              \
  S->spcode = MULTI_U3D;                                                      \
  S->tcode = TC;                                                              \
  S->x_grid = x_grid; S->y_grid = y_grid; S->z_grid = z_grid;                 \
  S->xBC = xBC; S->yBC = yBC; S->zBC = zBC;                                   \
  S->num_splines = num_splines;                                               \
  S->z_stride = num_splines;                                                  \
  S->y_stride = (intptr_t)h->Nz * num_splines;                                \
  S->x_stride = (intptr_t)h->Ny * h->Nz * num_splines;

multi_UBspline_3d_s *create_multi_UBspline_3d_s(Ugrid x_grid, Ugrid y_grid, Ugrid z_grid,
                                                BCtype_s xBC, BCtype_s yBC, BCtype_s zBC,
                                                int num_splines) {
  void *t;
  einspline_host3d *h = create_common("create_multi_UBspline_3d_s", 3, &x_grid, &y_grid, &z_grid,
                                      xBC.lCode, yBC.lCode, zBC.lCode, num_splines,
                                      sizeof(float), 1, &t);
  if (!h) return NULL;
  multi_UBspline_3d_s *s = &h->pub.s;
  SET_COMMON_FIELDS(s, SINGLE_REAL)
  s->coefs = (float *)t;
  return s;
}

multi_UBspline_3d_d *create_multi_UBspline_3d_d(Ugrid x_grid, Ugrid y_grid, Ugrid z_grid,
                                                BCtype_d xBC, BCtype_d yBC, BCtype_d zBC,
                                                int num_splines) {
  void *t;
  einspline_host3d *h = create_common("create_multi_UBspline_3d_d", 3, &x_grid, &y_grid, &z_grid,
                                      xBC.lCode, yBC.lCode, zBC.lCode, num_splines,
                                      sizeof(double), 1, &t);
  if (!h) return NULL;
  multi_UBspline_3d_d *s = &h->pub.d;
  SET_COMMON_FIELDS(s, DOUBLE_REAL)
  s->coefs = (double *)t;
  return s;
}

multi_UBspline_3d_c *create_multi_UBspline_3d_c(Ugrid x_grid, Ugrid y_grid, Ugrid z_grid,
                                                BCtype_c xBC, BCtype_c yBC, BCtype_c zBC,
                                                int num_splines) {
  void *t;
  einspline_host3d *h = create_common("create_multi_UBspline_3d_c", 3, &x_grid, &y_grid, &z_grid,
                                      xBC.lCode, yBC.lCode, zBC.lCode, num_splines,
                                      2 * sizeof(float), 2, &t);
  if (!h) return NULL;
  multi_UBspline_3d_c *s = &h->pub.c;
  SET_COMMON_FIELDS(s, SINGLE_COMPLEX)
  s->coefs = (complex_float *)t;
  s->lapl3 = calloc(3 * (size_t)num_splines, sizeof(complex_float));
  if (!s->lapl3) {
    fprintf(stderr, "Out of memory allocating spline coefficients in create_multi_UBspline_3d_c.\n");
    abort();
  }
  return s;
}

multi_UBspline_3d_z *create_multi_UBspline_3d_z(Ugrid x_grid, Ugrid y_grid, Ugrid z_grid,
                                                BCtype_z xBC, BCtype_z yBC, BCtype_z zBC,
                                                int num_splines) {
  void *t;
  einspline_host3d *h = create_common("create_multi_UBspline_3d_z", 3, &x_grid, &y_grid, &z_grid,
                                      xBC.lCode, yBC.lCode, zBC.lCode, num_splines,
                                      2 * sizeof(double), 2, &t);
  if (!h) return NULL;
  multi_UBspline_3d_z *s = &h->pub.z;
  SET_COMMON_FIELDS(s, DOUBLE_COMPLEX)
  s->coefs = (complex_double *)t;
  s->lapl3 = calloc(3 * (size_t)num_splines, sizeof(complex_double));
  if (!s->lapl3) {
    fprintf(stderr, "Out of memory allocating spline coefficients in create_multi_UBspline_3d_z.\n");
    abort();
  }
  return s;
}

/* 1-D and 2-D: the same object with Ny (and Nz) = 1. */
#define CREATE_1D(T, TY, BCT, TC, ELT, NC)                                                     \
  multi_UBspline_1d_##T *create_multi_UBspline_1d_##T(Ugrid x_grid, BCT xBC, int num_splines) { \
    void *t;                                                                                  \
    einspline_host3d *h = create_common("create_multi_UBspline_1d_" #T, 1, &x_grid, NULL, NULL, \
                                        xBC.lCode, PERIODIC, PERIODIC, num_splines, ELT, NC, &t); \
    if (!h) return NULL;                                                                      \
    multi_UBspline_1d_##T *s = &h->pub.T##1;                                                  \
    s->spcode = MULTI_U1D;                                                                    \
    s->tcode = TC;                                                                            \
    s->x_grid = x_grid;                                                                       \
    s->xBC = xBC;                                                                             \
    s->num_splines = num_splines;                                                             \
    s->x_stride = num_splines;                                                                \
    s->coefs = (TY *)t;                                                                       \
    return s;                                                                                 \
  }
CREATE_1D(s, float, BCtype_s, SINGLE_REAL, sizeof(float), 1)
CREATE_1D(d, double, BCtype_d, DOUBLE_REAL, sizeof(double), 1)
CREATE_1D(c, complex_float, BCtype_c, SINGLE_COMPLEX, 2 * sizeof(float), 2)
CREATE_1D(z, complex_double, BCtype_z, DOUBLE_COMPLEX, 2 * sizeof(double), 2)

#define CREATE_2D(T, TY, BCT, TC, ELT, NC, LAPL2)                                              \
  multi_UBspline_2d_##T *create_multi_UBspline_2d_##T(Ugrid x_grid, Ugrid y_grid, BCT xBC,     \
                                                      BCT yBC, int num_splines) {             \
    void *t;                                                                                  \
    einspline_host3d *h = create_common("create_multi_UBspline_2d_" #T, 2, &x_grid, &y_grid,   \
                                        NULL, xBC.lCode, yBC.lCode, PERIODIC, num_splines,    \
                                        ELT, NC, &t);                                         \
    if (!h) return NULL;                                                                      \
    multi_UBspline_2d_##T *s = &h->pub.T##2;                                                  \
    s->spcode = MULTI_U2D;                                                                    \
    s->tcode = TC;                                                                            \
    s->x_grid = x_grid; s->y_grid = y_grid;                                                   \
    s->xBC = xBC; s->yBC = yBC;                                                               \
    s->num_splines = num_splines;                                                             \
    s->y_stride = num_splines;                                                                \
    s->x_stride = (intptr_t)h->Ny * num_splines;                                              \
    s->coefs = (TY *)t;                                                                       \
    LAPL2                                                                                     \
    return s;                                                                                 \
  }
#define NO_LAPL2
#define WITH_LAPL2(TY)                                                                        \
  s->lapl2 = calloc(2 * (size_t)num_splines, sizeof(TY));                                    \
  if (!s->lapl2) {                                                                            \
    fprintf(stderr, "Out of memory allocating spline coefficients in create_multi_UBspline_2d.\n"); \
    abort();                                                                                  \
  }
CREATE_2D(s, float, BCtype_s, SINGLE_REAL, sizeof(float), 1, NO_LAPL2)
CREATE_2D(d, double, BCtype_d, DOUBLE_REAL, sizeof(double), 1, NO_LAPL2)
CREATE_2D(c, complex_float, BCtype_c, SINGLE_COMPLEX, 2 * sizeof(float), 2, WITH_LAPL2(complex_float))
CREATE_2D(z, complex_double, BCtype_z, DOUBLE_COMPLEX, 2 * sizeof(double), 2, WITH_LAPL2(complex_double))

/* ------------------------------------------------------------------------ */
/* set                                                                       */

typedef struct {
  bc_code l, r;
  double lv[2], rv[2];     /* per component (re, im) */
} axis_bc;

/* Solve `count` orbitals starting at `first`.  src holds count arrays of
 * Mx*My*Mz elements (My = Mz = 1 past the table's dim), each element
 * `ncomp` reals of type given by src_is_float.  The result is written into
 * the table (float or double reals). */
static void solve_block(einspline_host3d *h, const Ugrid *grids, const axis_bc *bcs,
                        int first, int count, const void *src, int src_is_float,
                        void *table, int table_is_float, const char *fn) {
  const int ncomp = h->ncomp, dim = h->dim;
  const int Nx = h->Nx, Ny = h->Ny, Nz = h->Nz;
  const int Mx = grids[0].num, My = dim >= 2 ? grids[1].num : 1, Mz = dim >= 3 ? grids[2].num : 1;
  const int oy = dim >= 2, oz = dim >= 3;     /* halo offset of the first data node */
  const int N = h->nsplines;
  if (first < 0 || count < 1 || first + count > N) {
    fprintf(stderr, "%s: spline index out of range\n", fn);
    return;
  }
  axis_op ops[3];
  memset(ops, 0, sizeof ops);
  for (int a = 0; a < dim; ++a) {
    int rc = axis_init(&ops[a], &grids[a], bcs[a].l, bcs[a].r);
    if (rc == -2) {
      fprintf(stderr, "Out of memory in %s.\n", fn);
      abort();
    }
    if (rc) {
      fprintf(stderr, "%s: unsupported boundary-condition combination\n", fn);
      for (int b = 0; b < a; ++b) axis_free(&ops[b]);
      return;
    }
  }
  /* orbitals per pass: keep the double scratch grid under ~256 MB */
  const size_t vol = (size_t)Nx * Ny * Nz;
  size_t per = (32u << 20) / (vol * (size_t)ncomp);
  if (per < 1) per = 1;
  if (per > (size_t)count) per = (size_t)count;
  double *G = malloc(sizeof(double) * vol * per * ncomp);
  if (!G) {
    fprintf(stderr, "Out of memory in %s.\n", fn);
    abort();
  }
  const size_t Mvol = (size_t)Mx * My * Mz;
  for (int o0 = 0; o0 < count; o0 += (int)per) {
    const int cnt = (count - o0) < (int)per ? (count - o0) : (int)per;
    const ptrdiff_t L = (ptrdiff_t)cnt * ncomp;
    memset(G, 0, sizeof(double) * vol * L);
    /* data node (i,j,k) -> element (i+1, j+1, k+1) on solved axes */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Mx; ++i)
      for (int j = 0; j < My; ++j)
        for (int k = 0; k < Mz; ++k) {
          double *dst = G + ((((size_t)(i + 1) * Ny + (j + oy)) * Nz + (k + oz)) * L);
          const size_t node = ((size_t)i * My + j) * Mz + k;
          for (int o = 0; o < cnt; ++o)
            for (int c = 0; c < ncomp; ++c) {
              const size_t si = ((size_t)(o0 + o) * Mvol + node) * ncomp + c;
              dst[o * ncomp + c] = src_is_float ? (double)((const float *)src)[si]
                                                : ((const double *)src)[si];
            }
        }
    /* x sweep: each x plane is one element */
    {
      const ptrdiff_t plane = (ptrdiff_t)Ny * Nz * L;
      const int nth = plane > 4096 ? 64 : 1;
      /* pieces start on whole elements so that q % ncomp is the component */
      const ptrdiff_t piece = ((plane / ncomp + nth - 1) / nth) * ncomp;
#pragma omp parallel for schedule(static) if (nth > 1)
      for (int t = 0; t < nth; ++t) {
        const ptrdiff_t q0 = (ptrdiff_t)t * piece;
        const ptrdiff_t q1 = q0 + piece < plane ? q0 + piece : plane;
        if (q1 > q0)
          axis_apply(&ops[0], G + q0, plane, q1 - q0, bcs[0].lv, bcs[0].rv, ncomp);
      }
    }
    /* y sweep, per x plane */
    if (dim >= 2) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < Nx; ++i)
        axis_apply(&ops[1], G + (size_t)i * Ny * Nz * L, (ptrdiff_t)Nz * L, (ptrdiff_t)Nz * L,
                   bcs[1].lv, bcs[1].rv, ncomp);
    }
    /* z sweep, per (x, y) line */
    if (dim >= 3) {
#pragma omp parallel for schedule(static)
      for (int ij = 0; ij < Nx * Ny; ++ij)
        axis_apply(&ops[2], G + (size_t)ij * Nz * L, L, L, bcs[2].lv, bcs[2].rv, ncomp);
    }
    /* scatter into the table column(s) */
    const size_t row = (size_t)N * ncomp;   /* reals per table row */
#pragma omp parallel for schedule(static)
    for (long e = 0; e < (long)vol; ++e) {
      const double *g = G + (size_t)e * L;
      const size_t off = (size_t)e * row + (size_t)(first + o0) * ncomp;
      if (table_is_float) {
        float *t = (float *)table + off;
        for (ptrdiff_t q = 0; q < L; ++q) t[q] = (float)g[q];
      } else {
        double *t = (double *)table + off;
        for (ptrdiff_t q = 0; q < L; ++q) t[q] = g[q];
      }
    }
  }
  free(G);
  for (int a = 0; a < dim; ++a) axis_free(&ops[a]);
  pthread_mutex_lock(&h->lock);
  h->dirty = 1;
  pthread_mutex_unlock(&h->lock);
}

static void bcs_real(axis_bc *b, bc_code l, bc_code r, double lv, double rv) {
  b->l = l; b->r = r;
  b->lv[0] = b->lv[1] = (l == FLAT || l == NATURAL) ? 0.0 : lv;
  b->rv[0] = b->rv[1] = (r == FLAT || r == NATURAL) ? 0.0 : rv;
}
static void bcs_cplx(axis_bc *b, bc_code l, bc_code r, double lr, double li, double rr,
                     double ri) {
  b->l = l; b->r = r;
  const int lz = (l == FLAT || l == NATURAL), rz = (r == FLAT || r == NATURAL);
  b->lv[0] = lz ? 0.0 : lr; b->lv[1] = lz ? 0.0 : li;
  b->rv[0] = rz ? 0.0 : rr; b->rv[1] = rz ? 0.0 : ri;
}

#define CHECK_OBJ(H, S, FN)                                                   \
  einspline_host3d *H = einspline_host_of(S);                                 \
  if (!H) {                                                                   \
    fprintf(stderr, "%s: spline was not created by this library\n", FN);     \
    return;                                                                   \
  }

#define CHECK_DIM(H, D, FN)                                                   \
  if ((H)->dim != (D)) {                                                      \
    fprintf(stderr, "%s: spline has %d dimensions\n", FN, (H)->dim);          \
    return;                                                                   \
  }

void set_multi_UBspline_3d_s_block(multi_UBspline_3d_s *s, int first, int count,
                                   const float *data) {
  CHECK_OBJ(h, s, "set_multi_UBspline_3d_s")
  CHECK_DIM(h, 3, "set_multi_UBspline_3d_s")
  Ugrid g[3] = {s->x_grid, s->y_grid, s->z_grid};
  axis_bc b[3];
  bcs_real(&b[0], s->xBC.lCode, s->xBC.rCode, s->xBC.lVal, s->xBC.rVal);
  bcs_real(&b[1], s->yBC.lCode, s->yBC.rCode, s->yBC.lVal, s->yBC.rVal);
  bcs_real(&b[2], s->zBC.lCode, s->zBC.rCode, s->zBC.lVal, s->zBC.rVal);
  solve_block(h, g, b, first, count, data, 1, s->coefs, 1, "set_multi_UBspline_3d_s");
}

void set_multi_UBspline_3d_d_block(multi_UBspline_3d_d *s, int first, int count,
                                   const double *data) {
  CHECK_OBJ(h, s, "set_multi_UBspline_3d_d")
  CHECK_DIM(h, 3, "set_multi_UBspline_3d_d")
  Ugrid g[3] = {s->x_grid, s->y_grid, s->z_grid};
  axis_bc b[3];
  bcs_real(&b[0], s->xBC.lCode, s->xBC.rCode, s->xBC.lVal, s->xBC.rVal);
  bcs_real(&b[1], s->yBC.lCode, s->yBC.rCode, s->yBC.lVal, s->yBC.rVal);
  bcs_real(&b[2], s->zBC.lCode, s->zBC.rCode, s->zBC.lVal, s->zBC.rVal);
  solve_block(h, g, b, first, count, data, 0, s->coefs, 0, "set_multi_UBspline_3d_d");
}

void set_multi_UBspline_3d_c_block(multi_UBspline_3d_c *s, int first, int count,
                                   const complex_float *data) {
  CHECK_OBJ(h, s, "set_multi_UBspline_3d_c")
  CHECK_DIM(h, 3, "set_multi_UBspline_3d_c")
  Ugrid g[3] = {s->x_grid, s->y_grid, s->z_grid};
  axis_bc b[3];
  bcs_cplx(&b[0], s->xBC.lCode, s->xBC.rCode, s->xBC.lVal_r, s->xBC.lVal_i, s->xBC.rVal_r, s->xBC.rVal_i);
  bcs_cplx(&b[1], s->yBC.lCode, s->yBC.rCode, s->yBC.lVal_r, s->yBC.lVal_i, s->yBC.rVal_r, s->yBC.rVal_i);
  bcs_cplx(&b[2], s->zBC.lCode, s->zBC.rCode, s->zBC.lVal_r, s->zBC.lVal_i, s->zBC.rVal_r, s->zBC.rVal_i);
  solve_block(h, g, b, first, count, data, 1, s->coefs, 1, "set_multi_UBspline_3d_c");
}

void set_multi_UBspline_3d_z_block(multi_UBspline_3d_z *s, int first, int count,
                                   const complex_double *data) {
  CHECK_OBJ(h, s, "set_multi_UBspline_3d_z")
  CHECK_DIM(h, 3, "set_multi_UBspline_3d_z")
  Ugrid g[3] = {s->x_grid, s->y_grid, s->z_grid};
  axis_bc b[3];
  bcs_cplx(&b[0], s->xBC.lCode, s->xBC.rCode, s->xBC.lVal_r, s->xBC.lVal_i, s->xBC.rVal_r, s->xBC.rVal_i);
  bcs_cplx(&b[1], s->yBC.lCode, s->yBC.rCode, s->yBC.lVal_r, s->yBC.lVal_i, s->yBC.rVal_r, s->yBC.rVal_i);
  bcs_cplx(&b[2], s->zBC.lCode, s->zBC.rCode, s->zBC.lVal_r, s->zBC.lVal_i, s->zBC.rVal_r, s->zBC.rVal_i);
  solve_block(h, g, b, first, count, data, 0, s->coefs, 0, "set_multi_UBspline_3d_z");
}

void set_multi_UBspline_3d_s(multi_UBspline_3d_s *s, int n, float *data) {
  set_multi_UBspline_3d_s_block(s, n, 1, data);
}
void set_multi_UBspline_3d_d(multi_UBspline_3d_d *s, int n, double *data) {
  set_multi_UBspline_3d_d_block(s, n, 1, data);
}
void set_multi_UBspline_3d_c(multi_UBspline_3d_c *s, int n, complex_float *data) {
  set_multi_UBspline_3d_c_block(s, n, 1, data);
}
void set_multi_UBspline_3d_z(multi_UBspline_3d_z *s, int n, complex_double *data) {
  set_multi_UBspline_3d_z_block(s, n, 1, data);
}

/* 1-D and 2-D sets. */
static int same_extent_class(bc_code a, bc_code b) {
  const int pa = (a == PERIODIC || a == ANTIPERIODIC), pb = (b == PERIODIC || b == ANTIPERIODIC);
  return pa == pb;
}

static void set1d_real(multi_UBspline_1d_d *s_d, multi_UBspline_1d_s *s_s, int n, const void *data,
                       bc_code l, bc_code r, double lv, double rv, const char *fn) {
  void *sp = s_d ? (void *)s_d : (void *)s_s;
  CHECK_OBJ(h, sp, fn)
  CHECK_DIM(h, 1, fn)
  const Ugrid g0 = s_d ? s_d->x_grid : s_s->x_grid;
  const bc_code tl = s_d ? s_d->xBC.lCode : s_s->xBC.lCode;
  if (!same_extent_class(l, tl)) {
    fprintf(stderr, "%s: boundary condition changes the table extent\n", fn);
    return;
  }
  Ugrid g[3] = {g0, g0, g0};
  axis_bc b[3];
  bcs_real(&b[0], l, r, lv, rv);
  b[1] = b[2] = b[0];
  solve_block(h, g, b, n, 1, data, s_s != NULL, s_d ? (void *)s_d->coefs : (void *)s_s->coefs,
              s_s != NULL, fn);
}

static void set1d_cplx(multi_UBspline_1d_z *s_z, multi_UBspline_1d_c *s_c, int n, const void *data,
                       bc_code l, bc_code r, double lr, double li, double rr, double ri,
                       const char *fn) {
  void *sp = s_z ? (void *)s_z : (void *)s_c;
  CHECK_OBJ(h, sp, fn)
  CHECK_DIM(h, 1, fn)
  const Ugrid g0 = s_z ? s_z->x_grid : s_c->x_grid;
  const bc_code tl = s_z ? s_z->xBC.lCode : s_c->xBC.lCode;
  if (!same_extent_class(l, tl)) {
    fprintf(stderr, "%s: boundary condition changes the table extent\n", fn);
    return;
  }
  Ugrid g[3] = {g0, g0, g0};
  axis_bc b[3];
  bcs_cplx(&b[0], l, r, lr, li, rr, ri);
  b[1] = b[2] = b[0];
  solve_block(h, g, b, n, 1, data, s_c != NULL, s_z ? (void *)s_z->coefs : (void *)s_c->coefs,
              s_c != NULL, fn);
}

void set_multi_UBspline_1d_s(multi_UBspline_1d_s *s, int n, float *data) {
  set1d_real(NULL, s, n, data, s->xBC.lCode, s->xBC.rCode, s->xBC.lVal, s->xBC.rVal,
             "set_multi_UBspline_1d_s");
}
void set_multi_UBspline_1d_d(multi_UBspline_1d_d *s, int n, double *data) {
  set1d_real(s, NULL, n, data, s->xBC.lCode, s->xBC.rCode, s->xBC.lVal, s->xBC.rVal,
             "set_multi_UBspline_1d_d");
}
void set_multi_UBspline_1d_d_BC(multi_UBspline_1d_d *s, int n, double *data, BCtype_d xBC) {
  set1d_real(s, NULL, n, data, xBC.lCode, xBC.rCode, xBC.lVal, xBC.rVal, "set_multi_UBspline_1d_d_BC");
}
void set_multi_UBspline_1d_c(multi_UBspline_1d_c *s, int n, complex_float *data) {
  set1d_cplx(NULL, s, n, data, s->xBC.lCode, s->xBC.rCode, s->xBC.lVal_r, s->xBC.lVal_i,
             s->xBC.rVal_r, s->xBC.rVal_i, "set_multi_UBspline_1d_c");
}
void set_multi_UBspline_1d_z(multi_UBspline_1d_z *s, int n, complex_double *data) {
  set1d_cplx(s, NULL, n, data, s->xBC.lCode, s->xBC.rCode, s->xBC.lVal_r, s->xBC.lVal_i,
             s->xBC.rVal_r, s->xBC.rVal_i, "set_multi_UBspline_1d_z");
}
void set_multi_UBspline_1d_z_BC(multi_UBspline_1d_z *s, int n, complex_double *data, BCtype_z xBC) {
  set1d_cplx(s, NULL, n, data, xBC.lCode, xBC.rCode, xBC.lVal_r, xBC.lVal_i, xBC.rVal_r, xBC.rVal_i,
             "set_multi_UBspline_1d_z_BC");
}

#define SET_2D_REAL(T, TY, ISF)                                                                \
  void set_multi_UBspline_2d_##T(multi_UBspline_2d_##T *s, int n, TY *data) {                  \
    CHECK_OBJ(h, s, "set_multi_UBspline_2d_" #T)                                                \
    CHECK_DIM(h, 2, "set_multi_UBspline_2d_" #T)                                                \
    Ugrid g[3] = {s->x_grid, s->y_grid, s->y_grid};                                            \
    axis_bc b[3];                                                                              \
    bcs_real(&b[0], s->xBC.lCode, s->xBC.rCode, s->xBC.lVal, s->xBC.rVal);                     \
    bcs_real(&b[1], s->yBC.lCode, s->yBC.rCode, s->yBC.lVal, s->yBC.rVal);                     \
    b[2] = b[1];                                                                               \
    solve_block(h, g, b, n, 1, data, ISF, s->coefs, ISF, "set_multi_UBspline_2d_" #T);         \
  }
#define SET_2D_CPLX(T, TY, ISF)                                                                \
  void set_multi_UBspline_2d_##T(multi_UBspline_2d_##T *s, int n, TY *data) {                  \
    CHECK_OBJ(h, s, "set_multi_UBspline_2d_" #T)                                                \
    CHECK_DIM(h, 2, "set_multi_UBspline_2d_" #T)                                                \
    Ugrid g[3] = {s->x_grid, s->y_grid, s->y_grid};                                            \
    axis_bc b[3];                                                                              \
    bcs_cplx(&b[0], s->xBC.lCode, s->xBC.rCode, s->xBC.lVal_r, s->xBC.lVal_i, s->xBC.rVal_r,   \
             s->xBC.rVal_i);                                                                   \
    bcs_cplx(&b[1], s->yBC.lCode, s->yBC.rCode, s->yBC.lVal_r, s->yBC.lVal_i, s->yBC.rVal_r,   \
             s->yBC.rVal_i);                                                                   \
    b[2] = b[1];                                                                               \
    solve_block(h, g, b, n, 1, data, ISF, s->coefs, ISF, "set_multi_UBspline_2d_" #T);         \
  }
SET_2D_REAL(s, float, 1)
SET_2D_REAL(d, double, 0)
SET_2D_CPLX(c, complex_float, 1)
SET_2D_CPLX(z, complex_double, 0)

/* ------------------------------------------------------------------------ */
/* operator export for the GPU solve                                         */

int einspline_export_axes(void *spline, einspline_axis_export ax[3]) {
  einspline_host3d *h = einspline_host_of(spline);
  if (!h || h->dim != 3) return -1;
  Ugrid g[3];
  axis_bc b[3];
  switch (h->pub.s.tcode) {
    case SINGLE_REAL: {
      const multi_UBspline_3d_s *S = &h->pub.s;
      g[0] = S->x_grid; g[1] = S->y_grid; g[2] = S->z_grid;
      bcs_real(&b[0], S->xBC.lCode, S->xBC.rCode, S->xBC.lVal, S->xBC.rVal);
      bcs_real(&b[1], S->yBC.lCode, S->yBC.rCode, S->yBC.lVal, S->yBC.rVal);
      bcs_real(&b[2], S->zBC.lCode, S->zBC.rCode, S->zBC.lVal, S->zBC.rVal);
      break;
    }
    case DOUBLE_REAL: {
      const multi_UBspline_3d_d *S = &h->pub.d;
      g[0] = S->x_grid; g[1] = S->y_grid; g[2] = S->z_grid;
      bcs_real(&b[0], S->xBC.lCode, S->xBC.rCode, S->xBC.lVal, S->xBC.rVal);
      bcs_real(&b[1], S->yBC.lCode, S->yBC.rCode, S->yBC.lVal, S->yBC.rVal);
      bcs_real(&b[2], S->zBC.lCode, S->zBC.rCode, S->zBC.lVal, S->zBC.rVal);
      break;
    }
    case SINGLE_COMPLEX: {
      const multi_UBspline_3d_c *S = &h->pub.c;
      g[0] = S->x_grid; g[1] = S->y_grid; g[2] = S->z_grid;
      bcs_cplx(&b[0], S->xBC.lCode, S->xBC.rCode, S->xBC.lVal_r, S->xBC.lVal_i, S->xBC.rVal_r, S->xBC.rVal_i);
      bcs_cplx(&b[1], S->yBC.lCode, S->yBC.rCode, S->yBC.lVal_r, S->yBC.lVal_i, S->yBC.rVal_r, S->yBC.rVal_i);
      bcs_cplx(&b[2], S->zBC.lCode, S->zBC.rCode, S->zBC.lVal_r, S->zBC.lVal_i, S->zBC.rVal_r, S->zBC.rVal_i);
      break;
    }
    case DOUBLE_COMPLEX: {
      const multi_UBspline_3d_z *S = &h->pub.z;
      g[0] = S->x_grid; g[1] = S->y_grid; g[2] = S->z_grid;
      bcs_cplx(&b[0], S->xBC.lCode, S->xBC.rCode, S->xBC.lVal_r, S->xBC.lVal_i, S->xBC.rVal_r, S->xBC.rVal_i);
      bcs_cplx(&b[1], S->yBC.lCode, S->yBC.rCode, S->yBC.lVal_r, S->yBC.lVal_i, S->yBC.rVal_r, S->yBC.rVal_i);
      bcs_cplx(&b[2], S->zBC.lCode, S->zBC.rCode, S->zBC.lVal_r, S->zBC.lVal_i, S->zBC.rVal_r, S->zBC.rVal_i);
      break;
    }
    default:
      return -1;
  }
  memset(ax, 0, 3 * sizeof(*ax));
  for (int a = 0; a < 3; ++a) {
    axis_op op;
    const int rc = axis_init(&op, &g[a], b[a].l, b[a].r);
    if (rc) {
      for (int k = 0; k < a; ++k) free(ax[k].arrays);
      return rc;
    }
    const int M = op.M;
    double *arr = calloc((size_t)4 * M, sizeof(double));
    if (!arr) {
      axis_free(&op);
      for (int k = 0; k < a; ++k) free(ax[k].arrays);
      return -2;
    }
    const int m = op.kind == 2 ? M : M - 1;       /* rows of the factorised core */
    memcpy(arr, op.inv, sizeof(double) * m);
    memcpy(arr + M, op.g, sizeof(double) * m);
    memcpy(arr + 2 * M, op.sub, sizeof(double) * m);
    if (op.w) memcpy(arr + 3 * M, op.w, sizeof(double) * m);
    ax[a].M = M; ax[a].n = op.n; ax[a].kind = op.kind;
    ax[a].sgn = op.sgn; ax[a].zinv = op.zinv;
    ax[a].l0 = op.l0; ax[a].r2 = op.r2;
    memcpy(ax[a].lv, b[a].lv, sizeof b[a].lv);
    memcpy(ax[a].rv, b[a].rv, sizeof b[a].rv);
    ax[a].arrays = arr;
    axis_free(&op);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */

void einspline_hip_invalidate(void *spline) {
  einspline_host3d *h = einspline_host_of(spline);
  if (!h) return;
  pthread_mutex_lock(&h->lock);
  h->dirty = 1;
  pthread_mutex_unlock(&h->lock);
}

/* src/bspline_create.c:1864-1877 dispatches on sp_code; this library
 * creates only multi-splines, which carry a private tail.  Anything else is
 * treated as a plain {codes, coefs} object. */
void destroy_Bspline(void *spline) {
  if (!spline) return;
  einspline_host3d *h = einspline_host_of(spline);
  if (!h) {
    Bspline *b = (Bspline *)spline;
    free(b->coefs);
    free(b);
    return;
  }
  if (h->mirror) einspline_mirror_release(h->mirror);
  h->mirror = NULL;
  free(h->pub.s.coefs);
  if (h->dim == 3 && h->pub.s.tcode == SINGLE_COMPLEX) free(h->pub.c.lapl3);
  if (h->dim == 3 && h->pub.s.tcode == DOUBLE_COMPLEX) free(h->pub.z.lapl3);
  if (h->dim == 2 && h->pub.s.tcode == SINGLE_COMPLEX) free(h->pub.c2.lapl2);
  if (h->dim == 2 && h->pub.s.tcode == DOUBLE_COMPLEX) free(h->pub.z2.lapl2);
  h->magic = 0;
  pthread_mutex_destroy(&h->lock);
  free(h);
}
