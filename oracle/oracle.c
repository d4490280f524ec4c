/* oracle.c -- CPU restatement of einspline's 3-D multi-spline path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the
 * HIP library in einspline_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it, and only as a checker or as the
 * timed CPU baseline -- never as the product path.
 *
 * It restates, with the same arithmetic order, the reference's default
 * ("std") build:
 *   - 1-D interpolating solves: src/bspline_create.c:46-89 (deriv, float),
 *     :97-145 (periodic, float), :154-205 (antiperiodic, float),
 *     :237-318 (find_coefs_1d_s, float solve of the default build),
 *     :958-1001, :1009-1055, :1063-1111, :1115-1198 (the double versions);
 *   - 3-D set: src/multi_bspline_create.c:323-373 (s), :650-732 (c),
 *     :1020-1073 (d), :1392-1477 (z): sweeps along x, then y, then z;
 *   - evaluation: src/multi_bspline_eval_std_d_impl.h:482-533 (v),
 *     :537-622 (vg), :627-737 (vgl), :742-864 (vgh); the s/c/z files are the
 *     same loops in float / complex float / complex double
 *     (std_s_impl.h:480-862, std_c_impl.h:477-860, std_z_impl.h:481-864);
 *   - basis matrices: src/bspline_data.c:143-200.
 * Complex splines are evaluated as two real evaluations with the same
 * weights, which is what real*complex arithmetic in the reference computes.
 *
 * Build with -ffp-contract=off so every product and sum rounds as written
 * (the reference's default x86-64 build has no FMA).
 *
 * The reference itself cannot be built here (its headers include an
 * autoconf-generated config.h that the snapshot does not contain), so this
 * restatement is pinned by the closed-form and brute-force checks in tests/.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

/* ---- basis matrices (power form rows, applied to [t^3, t^2, t, 1]) ---- */
static const double A_d[16] = {-1.0 / 6.0, 3.0 / 6.0, -3.0 / 6.0, 1.0 / 6.0,
                               3.0 / 6.0, -6.0 / 6.0, 0.0 / 6.0, 4.0 / 6.0,
                               -3.0 / 6.0, 3.0 / 6.0, 3.0 / 6.0, 1.0 / 6.0,
                               1.0 / 6.0, 0.0 / 6.0, 0.0 / 6.0, 0.0 / 6.0};
static const double dA_d[16] = {0.0, -0.5, 1.0, -0.5, 0.0, 1.5, -2.0, 0.0,
                                0.0, -1.5, 1.0, 0.5, 0.0, 0.5, 0.0, 0.0};
static const double d2A_d[16] = {0.0, 0.0, -1.0, 1.0, 0.0, 0.0, 3.0, -2.0,
                                 0.0, 0.0, -3.0, 1.0, 0.0, 0.0, 1.0, 0.0};
static const float A_f[16] = {-1.0 / 6.0, 3.0 / 6.0, -3.0 / 6.0, 1.0 / 6.0,
                              3.0 / 6.0, -6.0 / 6.0, 0.0 / 6.0, 4.0 / 6.0,
                              -3.0 / 6.0, 3.0 / 6.0, 3.0 / 6.0, 1.0 / 6.0,
                              1.0 / 6.0, 0.0 / 6.0, 0.0 / 6.0, 0.0 / 6.0};
static const float dA_f[16] = {0.0, -0.5, 1.0, -0.5, 0.0, 1.5, -2.0, 0.0,
                               0.0, -1.5, 1.0, 0.5, 0.0, 0.5, 0.0, 0.0};
static const float d2A_f[16] = {0.0, 0.0, -1.0, 1.0, 0.0, 0.0, 3.0, -2.0,
                                0.0, 0.0, -3.0, 1.0, 0.0, 0.0, 1.0, 0.0};

/* ---- table geometry (src/multi_bspline_create.c:268-300 for 3-D; the
 * 1-D and 2-D creates, :767-820 and :840-901, use the same per-axis rules) */
int oracle_table_init(otable *t, const int M[3], const double start[3], const double end[3],
                      const int lcode[3], const int rcode[3], const double lval[6],
                      const double rval[6], int N, int ncomp, int dim) {
  memset(t, 0, sizeof *t);
  if (N < 1 || (ncomp != 1 && ncomp != 2) || dim < 1 || dim > 3) return -1;
  int Nn[3] = {1, 1, 1};
  for (int a = 0; a < 3; ++a) {
    t->M[a] = 1;
    if (a >= dim) continue;          /* inactive axis: one row, never offset */
    if (M[a] < 2) return -1;
    t->M[a] = M[a];
    t->start[a] = start[a];
    t->end[a] = end[a];
    t->lcode[a] = lcode[a];
    t->rcode[a] = rcode[a];
    t->lval[a][0] = lval[2 * a]; t->lval[a][1] = lval[2 * a + 1];
    t->rval[a][0] = rval[2 * a]; t->rval[a][1] = rval[2 * a + 1];
    Nn[a] = (lcode[a] == OR_PERIODIC || lcode[a] == OR_ANTIPERIODIC) ? M[a] + 3 : M[a] + 2;
    t->delta[a] = (end[a] - start[a]) / (double)(Nn[a] - 3);
    t->delta_inv[a] = 1.0 / t->delta[a];
  }
  for (int a = 0; a < 3; ++a) t->Nn[a] = Nn[a];
  t->N = N;
  t->ncomp = ncomp;
  t->dim = dim;
  t->zs = dim >= 3 ? N : 0;
  t->ys = dim >= 2 ? (long)Nn[2] * N : 0;
  t->xs = (long)Nn[1] * Nn[2] * N;
  return 0;
}

size_t oracle_table_elems(const otable *t) {
  return (size_t)t->Nn[0] * t->Nn[1] * t->Nn[2] * t->N;
}

/* ---- 1-D solvers, generated for float and double ---------------------- */
/* bands: 4 values per row, [sub, diag, super, rhs] */
#define DEFINE_SOLVERS(T, SFX)                                                     \
  static void periodic_##SFX(T *bands, T *coefs, int M, long cs, T sgn) {          \
    T *last = malloc(sizeof(T) * (size_t)M);                                       \
    if (sgn < 0) { bands[0] *= -1; bands[4 * (M - 1) + 2] *= -1; }                 \
    T *b0 = bands, *bl = bands + 4 * (M - 1);                                      \
    b0[2] /= b0[1]; b0[0] /= b0[1]; b0[3] /= b0[1]; b0[1] = 1;                     \
    bl[1] -= bl[2] * b0[0];                                                        \
    bl[3] -= bl[2] * b0[3];                                                        \
    bl[2] = -bl[2] * b0[2];                                                        \
    last[0] = b0[0];                                                               \
    for (int r = 1; r < M - 1; ++r) {                                              \
      T *br = bands + 4 * r, *bp = bands + 4 * (r - 1);                            \
      br[1] -= br[0] * bp[2];                                                      \
      br[3] -= br[0] * bp[3];                                                      \
      last[r] = -br[0] * last[r - 1];                                              \
      br[0] = 0;                                                                   \
      br[2] /= br[1];                                                              \
      br[3] /= br[1];                                                              \
      last[r] /= br[1];                                                            \
      br[1] = 1;                                                                   \
      if (r < M - 2) {                                                             \
        bl[3] -= bl[2] * br[3];                                                    \
        bl[1] -= bl[2] * last[r];                                                  \
        bl[2] = -bl[2] * br[2];                                                    \
      }                                                                            \
    }                                                                              \
    T *bm = bands + 4 * (M - 2);                                                   \
    bl[0] += bl[2];                                                                \
    bl[1] -= bl[0] * (bm[2] + last[M - 2]);                                        \
    bl[3] -= bl[0] * bm[3];                                                        \
    bl[3] /= bl[1];                                                                \
    coefs[M * cs] = bl[3];                                                         \
    for (int r = M - 2; r >= 0; --r)                                               \
      coefs[(r + 1) * cs] = bands[4 * r + 3] - bands[4 * r + 2] * coefs[(r + 2) * cs] - \
                            last[r] * coefs[M * cs];                               \
    coefs[0] = sgn * coefs[M * cs];                                                \
    coefs[(M + 1) * cs] = sgn * coefs[1 * cs];                                     \
    coefs[(M + 2) * cs] = sgn * coefs[2 * cs];                                     \
    free(last);                                                                    \
  }                                                                                \
  static void deriv_##SFX(T *bands, T *coefs, int M, long cs) {                    \
    T *b0 = bands, *b1 = bands + 4;                                                \
    b0[1] /= b0[0]; b0[2] /= b0[0]; b0[3] /= b0[0]; b0[0] = 1;                     \
    b1[1] -= b1[0] * b0[1];                                                        \
    b1[2] -= b1[0] * b0[2];                                                        \
    b1[3] -= b1[0] * b0[3];                                                        \
    b0[0] = 0;                                                                     \
    b1[2] /= b1[1];                                                                \
    b1[3] /= b1[1];                                                                \
    b1[1] = 1;                                                                     \
    for (int r = 2; r < M + 1; ++r) {                                              \
      T *br = bands + 4 * r, *bp = bands + 4 * (r - 1);                            \
      br[1] -= br[0] * bp[2];                                                      \
      br[3] -= br[0] * bp[3];                                                      \
      br[2] /= br[1];                                                              \
      br[3] /= br[1];                                                              \
      br[0] = 0;                                                                   \
      br[1] = 1;                                                                   \
    }                                                                              \
    T *bl = bands + 4 * (M + 1), *bm1 = bands + 4 * (M - 1), *bm = bands + 4 * M;  \
    bl[1] -= bl[0] * bm1[2];                                                       \
    bl[3] -= bl[0] * bm1[3];                                                       \
    bl[2] -= bl[1] * bm[2];                                                        \
    bl[3] -= bl[1] * bm[3];                                                        \
    bl[3] /= bl[2];                                                                \
    bl[2] = 1;                                                                     \
    coefs[(M + 1) * cs] = bl[3];                                                   \
    for (int r = M; r > 0; --r)                                                    \
      coefs[r * cs] = bands[4 * r + 3] - bands[4 * r + 2] * coefs[cs * (r + 1)];   \
    coefs[0] = b0[3] - b0[1] * coefs[1 * cs] - b0[2] * coefs[2 * cs];              \
  }                                                                                \
  /* data[i*ds] -> coefs[j*cs], one axis */                                       \
  static void line_##SFX(const otable *t, int axis, int comp, const T *data, long ds, \
                         T *coefs, long cs) {                                      \
    const int M = t->M[axis];                                                      \
    const int lc = t->lcode[axis], rc = t->rcode[axis];                            \
    const T basis[3] = {(T)(1.0 / 6.0), (T)(2.0 / 3.0), (T)(1.0 / 6.0)};           \
    if (lc == OR_PERIODIC || lc == OR_ANTIPERIODIC) {                              \
      T *bands = malloc(sizeof(T) * 4 * (size_t)M);                                \
      for (int i = 0; i < M; ++i) {                                                \
        bands[4 * i + 0] = basis[0]; bands[4 * i + 1] = basis[1];                  \
        bands[4 * i + 2] = basis[2]; bands[4 * i + 3] = data[i * ds];              \
      }                                                                            \
      periodic_##SFX(bands, coefs, M, cs, lc == OR_PERIODIC ? (T)1 : (T)-1);       \
      free(bands);                                                                 \
      return;                                                                      \
    }                                                                              \
    const double di = t->delta_inv[axis];                                          \
    T L[4], R[4];                                                                  \
    double lv = t->lval[axis][comp], rv = t->rval[axis][comp];                     \
    if (lc == OR_FLAT || lc == OR_NATURAL) lv = 0.0;                               \
    if (rc == OR_FLAT || rc == OR_NATURAL) rv = 0.0;                               \
    if (lc == OR_FLAT || lc == OR_DERIV1) {                                        \
      L[0] = (T)(-0.5 * di); L[1] = (T)(0.0 * di); L[2] = (T)(0.5 * di);           \
    } else {                                                                       \
      L[0] = (T)(1.0 * di * di); L[1] = (T)(-2.0 * di * di); L[2] = (T)(1.0 * di * di); \
    }                                                                              \
    L[3] = (T)lv;                                                                  \
    if (rc == OR_FLAT || rc == OR_DERIV1) {                                        \
      R[0] = (T)(-0.5 * di); R[1] = (T)(0.0 * di); R[2] = (T)(0.5 * di);           \
    } else {                                                                       \
      R[0] = (T)(1.0 * di * di); R[1] = (T)(-2.0 * di * di); R[2] = (T)(1.0 * di * di); \
    }                                                                              \
    R[3] = (T)rv;                                                                  \
    T *bands = malloc(sizeof(T) * 4 * (size_t)(M + 2));                            \
    for (int i = 0; i < 4; ++i) { bands[i] = L[i]; bands[4 * (M + 1) + i] = R[i]; } \
    for (int i = 0; i < M; ++i) {                                                  \
      bands[4 * (i + 1) + 0] = basis[0]; bands[4 * (i + 1) + 1] = basis[1];        \
      bands[4 * (i + 1) + 2] = basis[2]; bands[4 * (i + 1) + 3] = data[i * ds];    \
    }                                                                              \
    deriv_##SFX(bands, coefs, M, cs);                                              \
    free(bands);                                                                   \
  }                                                                                \
  /* Set of orbital n: data[M0][M1][M2] (ncomp reals per element; axes past  \
   * dim have M = 1).  Sweeps x, then y, then z, as the reference's 1-D, 2-D  \
   * and 3-D sets do (multi_bspline_create.c:822-836, :903-935, :1020-1073). */ \
  static void setnd_##SFX(const otable *t, T *coefs, int n, const T *data) {        \
    const int nc = t->ncomp, dim = t->dim;                                         \
    const int My = t->M[1], Mz = t->M[2];                                          \
    const int Nx = t->Nn[0], Ny = t->Nn[1], Nz = t->Nn[2];                         \
    const long es = (long)t->N;  /* elements per grid node */                     \
    T *c = coefs + (long)n * nc;                                                    \
    for (int comp = 0; comp < nc; ++comp) {                                        \
      for (int iy = 0; iy < My; ++iy)                                              \
        for (int iz = 0; iz < Mz; ++iz)                                            \
          line_##SFX(t, 0, comp, data + (long)nc * (iy * Mz + iz) + comp,          \
                     (long)nc * My * Mz, c + (long)nc * (iy * Nz + iz) * es + comp, \
                     (long)nc * Ny * Nz * es);                                     \
      if (dim >= 2)                                                                \
        for (int ix = 0; ix < Nx; ++ix)                                            \
          for (int iz = 0; iz < Mz; ++iz) {                                        \
            T *p = c + (long)nc * ((long)ix * Ny * Nz + iz) * es + comp;           \
            line_##SFX(t, 1, comp, p, (long)nc * Nz * es, p, (long)nc * Nz * es);  \
          }                                                                        \
      if (dim >= 3)                                                                \
        for (int ix = 0; ix < Nx; ++ix)                                            \
          for (int iy = 0; iy < Ny; ++iy) {                                        \
            T *p = c + (long)nc * (((long)ix * Ny + iy) * Nz) * es + comp;         \
            line_##SFX(t, 2, comp, p, (long)nc * es, p, (long)nc * es);            \
          }                                                                        \
    }                                                                              \
  }

DEFINE_SOLVERS(float, f)
DEFINE_SOLVERS(double, d)

/* The y sweep above visits z rows [0, Mz) only; the reference also sweeps
 * the z-halo rows [Mz, Nz), whose inputs are not yet set and whose outputs
 * the z sweep then overwrites, so the tables are identical. */

void oracle_set_s(const otable *t, float *coefs, int n, const float *data) {
  setnd_f(t, coefs, n, data);
}
void oracle_set_d(const otable *t, double *coefs, int n, const double *data) {
  setnd_d(t, coefs, n, data);
}
void oracle_set_c(const otable *t, float *coefs, int n, const float *data) {
  setnd_f(t, coefs, n, data);
}
void oracle_set_z(const otable *t, double *coefs, int n, const double *data) {
  setnd_d(t, coefs, n, data);
}

/* ---- evaluation -------------------------------------------------------- */
/* One loop for every dimension: an axis past dim has the single weight 1 and
 * zero derivatives, so its products reduce exactly (x*1 == x) to the 1-D and
 * 2-D reference loops (std_d_impl.h:37-180 and :183-476), and to the 3-D ones
 * (:482-864) for dim 3.  vghgh follows :868-1049: third-derivative weights
 * are the constant column of d3A (bspline_data.c:195-200).
 * Returns 0, or -1 (outputs NaN) when the stencil would leave the table,
 * where the reference reads outside its allocation. */
#define DEFINE_EVAL(T, SFX, MODF, A, DA, D2A)                                      \
  static int eval_##SFX(const otable *t, const T *coefs, int mode, double x, double y, \
                        double z, T *vals, T *grads, T *lh, T *gh) {               \
    const int nc = t->ncomp, N = t->N, NR = N * nc, dim = t->dim;                  \
    const double xx[3] = {x, y, z};                                                \
    T tt[3] = {0, 0, 0};                                                           \
    int ii[3] = {0, 0, 0}, ok = 1;                                                 \
    for (int d = 0; d < dim; ++d) {                                                \
      T u = (xx[d] - t->start[d]) * t->delta_inv[d], ip;                           \
      tt[d] = MODF(u, &ip);                                                        \
      ii[d] = (int)ip;                                                             \
      ok = ok && ii[d] >= 0 && ii[d] + 3 < t->Nn[d];                               \
    }                                                                              \
    const int ng = dim, nh = mode == OR_VGHGH ? 9 : mode == OR_VGH ? (dim == 1 ? 1 : dim * dim) \
                                                 : mode == OR_VGL ? 1 : 0;         \
    const int ngh = mode == OR_VGHGH ? 27 : 0;                                     \
    if (!ok) {                                                                     \
      for (int q = 0; q < NR; ++q) vals[q] = (T)NAN;                               \
      if (mode >= OR_VG) for (int q = 0; q < ng * NR; ++q) grads[q] = (T)NAN;      \
      for (int q = 0; q < nh * NR; ++q) lh[q] = (T)NAN;                            \
      for (int q = 0; q < ngh * NR; ++q) gh[q] = (T)NAN;                           \
      return -1;                                                                   \
    }                                                                              \
    /* weights per axis; an inactive axis has a = {1}, derivatives 0 */          \
    T a[3][4], da[3][4], d2a[3][4], d3a[3][4];                                     \
    int nk[3];                                                                     \
    for (int d = 0; d < 3; ++d) {                                                  \
      nk[d] = d < dim ? 4 : 1;                                                     \
      if (d >= dim) {                                                              \
        a[d][0] = 1; da[d][0] = 0; d2a[d][0] = 0; d3a[d][0] = 0;                   \
        continue;                                                                  \
      }                                                                            \
      const T tp[4] = {tt[d] * tt[d] * tt[d], tt[d] * tt[d], tt[d], 1};            \
      for (int r = 0; r < 4; ++r) {                                                \
        a[d][r] = (A[4 * r + 0] * tp[0] + A[4 * r + 1] * tp[1] +                   \
                   A[4 * r + 2] * tp[2] + A[4 * r + 3] * tp[3]);                   \
        da[d][r] = (DA[4 * r + 0] * tp[0] + DA[4 * r + 1] * tp[1] +                \
                    DA[4 * r + 2] * tp[2] + DA[4 * r + 3] * tp[3]);                \
        d2a[d][r] = (D2A[4 * r + 0] * tp[0] + D2A[4 * r + 1] * tp[1] +             \
                     D2A[4 * r + 2] * tp[2] + D2A[4 * r + 3] * tp[3]);             \
        d3a[d][r] = (T)D3_COL[r] * tp[3];                                          \
      }                                                                            \
    }                                                                              \
    /* hessian slot of each unique pair (xx xy xz yy yz zz) in the dim x dim  \
     * row-major block; 1-D vgh is vgl (std_d_impl.h:168-180) */                 \
    const int hs3[6] = {0, 1, 2, 4, 5, 8}, hs2[6] = {0, 1, -1, 3, -1, -1};          \
    const int *hs = dim == 3 ? hs3 : hs2;                                          \
    /* gradhess slots of the unique triples (xxx xxy xxz xyy xyz xzz yyy yyz  \
     * yzz zzz), std_d_impl.h:1006-1015 */                                        \
    const int gs[10] = {0, 1, 2, 4, 5, 8, 13, 14, 17, 26};                         \
    T *l3 = NULL;                                                                  \
    const int lap = mode == OR_VGL || (mode == OR_VGH && dim == 1);                \
    if (lap) l3 = calloc((size_t)3 * NR, sizeof(T));                               \
    for (int q = 0; q < NR; ++q) vals[q] = 0;                                      \
    if (mode >= OR_VG) for (int q = 0; q < ng * NR; ++q) grads[q] = 0;             \
    if (!lap) for (int q = 0; q < nh * NR; ++q) lh[q] = 0;                         \
    for (int q = 0; q < ngh * NR; ++q) gh[q] = 0;                                  \
    const int ix = ii[0], iy = ii[1], iz = ii[2];                                  \
    for (int i = 0; i < nk[0]; ++i)                                                \
      for (int j = 0; j < nk[1]; ++j)                                              \
        for (int k = 0; k < nk[2]; ++k) {                                          \
          const T abc = a[0][i] * a[1][j] * a[2][k];                               \
          const T *cf = coefs + ((ix + i) * t->xs + (iy + j) * t->ys + (iz + k) * t->zs) * nc; \
          if (mode == OR_V) {                                                      \
            for (int q = 0; q < NR; ++q) vals[q] += abc * cf[q];                   \
            continue;                                                              \
          }                                                                        \
          T g[3], h[6], g3[10];                                                    \
          g[0] = da[0][i] * a[1][j] * a[2][k];                                     \
          g[1] = a[0][i] * da[1][j] * a[2][k];                                     \
          g[2] = a[0][i] * a[1][j] * da[2][k];                                     \
          h[0] = d2a[0][i] * a[1][j] * a[2][k];                                    \
          h[1] = da[0][i] * da[1][j] * a[2][k];                                    \
          h[2] = da[0][i] * a[1][j] * da[2][k];                                    \
          h[3] = a[0][i] * d2a[1][j] * a[2][k];                                    \
          h[4] = a[0][i] * da[1][j] * da[2][k];                                    \
          h[5] = a[0][i] * a[1][j] * d2a[2][k];                                    \
          if (mode == OR_VGHGH) {                                                  \
            g3[0] = d3a[0][i] * a[1][j] * a[2][k];                                 \
            g3[1] = d2a[0][i] * da[1][j] * a[2][k];                                \
            g3[2] = d2a[0][i] * a[1][j] * da[2][k];                                \
            g3[3] = da[0][i] * d2a[1][j] * a[2][k];                                \
            g3[4] = da[0][i] * da[1][j] * da[2][k];                                \
            g3[5] = da[0][i] * a[1][j] * d2a[2][k];                                \
            g3[6] = a[0][i] * d3a[1][j] * a[2][k];                                 \
            g3[7] = a[0][i] * d2a[1][j] * da[2][k];                                \
            g3[8] = a[0][i] * da[1][j] * d2a[2][k];                                \
            g3[9] = a[0][i] * a[1][j] * d3a[2][k];                                 \
          }                                                                        \
          for (int q = 0; q < NR; ++q) {                                           \
            const int n = q / nc, r = q % nc;                                      \
            const T cv = cf[q];                                                    \
            vals[q] += abc * cv;                                                   \
            for (int d = 0; d < ng; ++d) grads[(ng * n + d) * nc + r] += g[d] * cv; \
            if (lap) {                                                             \
              for (int d = 0; d < dim; ++d)                                        \
                l3[(3 * n + d) * nc + r] += h[d == 0 ? 0 : d == 1 ? 3 : 5] * cv;    \
            } else if (mode >= OR_VGH) {                                           \
              for (int u = 0; u < 6; ++u)                                          \
                if (hs[u] >= 0) lh[(nh * n + hs[u]) * nc + r] += h[u] * cv;        \
            }                                                                      \
            if (mode == OR_VGHGH)                                                  \
              for (int u = 0; u < 10; ++u) gh[(27 * n + gs[u]) * nc + r] += g3[u] * cv; \
          }                                                                        \
        }                                                                          \
    if (mode >= OR_VG) {                                                           \
      const T di[3] = {(T)t->delta_inv[0], (T)t->delta_inv[1], (T)t->delta_inv[2]}; \
      for (int q = 0; q < NR; ++q) {                                               \
        const int n = q / nc, r = q % nc;                                          \
        for (int d = 0; d < ng; ++d) grads[(ng * n + d) * nc + r] *= di[d];        \
        if (lap) {                                                                 \
          T sum = 0;                                                               \
          for (int d = 0; d < dim; ++d) {                                          \
            l3[(3 * n + d) * nc + r] *= di[d] * di[d];                             \
            sum = d == 0 ? l3[(3 * n) * nc + r] : sum + l3[(3 * n + d) * nc + r];  \
          }                                                                        \
          lh[q] = sum;                                                             \
        } else if (mode >= OR_VGH && dim == 3) {                                   \
          T *hh = lh + 9 * n * nc + r;                                             \
          hh[0 * nc] *= di[0] * di[0];                                             \
          hh[4 * nc] *= di[1] * di[1];                                             \
          hh[8 * nc] *= di[2] * di[2];                                             \
          hh[1 * nc] *= di[0] * di[1];                                             \
          hh[2 * nc] *= di[0] * di[2];                                             \
          hh[5 * nc] *= di[1] * di[2];                                             \
          hh[3 * nc] = hh[1 * nc];                                                 \
          hh[6 * nc] = hh[2 * nc];                                                 \
          hh[7 * nc] = hh[5 * nc];                                                 \
        } else if (mode == OR_VGH && dim == 2) {                                   \
          T *hh = lh + 4 * n * nc + r;                                             \
          hh[0 * nc] *= di[0] * di[0];                                             \
          hh[1 * nc] *= di[0] * di[1];                                             \
          hh[3 * nc] *= di[1] * di[1];                                             \
          hh[2 * nc] = hh[1 * nc];                                                 \
        }                                                                          \
        if (mode == OR_VGHGH) {                                                    \
          T *G = gh + 27 * n * nc + r;                                             \
          G[0 * nc] *= di[0] * di[0] * di[0];                                      \
          G[1 * nc] *= di[0] * di[0] * di[1];                                      \
          G[2 * nc] *= di[0] * di[0] * di[2];                                      \
          G[4 * nc] *= di[0] * di[1] * di[1];                                      \
          G[5 * nc] *= di[0] * di[1] * di[2];                                      \
          G[8 * nc] *= di[0] * di[2] * di[2];                                      \
          G[13 * nc] *= di[1] * di[1] * di[1];                                     \
          G[14 * nc] *= di[1] * di[1] * di[2];                                     \
          G[17 * nc] *= di[1] * di[2] * di[2];                                     \
          G[26 * nc] *= di[2] * di[2] * di[2];                                     \
          /* every slot (a,b,c) holds the entry of the sorted triple */           \
          for (int u = 0; u < 27; ++u) {                                           \
            int e[3] = {u / 9, (u / 3) % 3, u % 3}, tmp;                           \
            if (e[0] > e[1]) { tmp = e[0]; e[0] = e[1]; e[1] = tmp; }              \
            if (e[1] > e[2]) { tmp = e[1]; e[1] = e[2]; e[2] = tmp; }              \
            if (e[0] > e[1]) { tmp = e[0]; e[0] = e[1]; e[1] = tmp; }              \
            G[u * nc] = G[(9 * e[0] + 3 * e[1] + e[2]) * nc];                      \
          }                                                                        \
        }                                                                          \
      }                                                                            \
    }                                                                              \
    free(l3);                                                                      \
    return 0;                                                                      \
  }

/* constant column of the third-derivative basis (src/bspline_data.c:195-200) */
static const double D3_COL[4] = {-1.0, 3.0, -3.0, 1.0};

DEFINE_EVAL(float, f, modff, A_f, dA_f, d2A_f)
DEFINE_EVAL(double, d, modf, A_d, dA_d, d2A_d)

int oracle_eval_s(const otable *t, const float *coefs, int mode, double x, double y, double z,
                  float *v, float *g, float *lh, float *gh) {
  return eval_f(t, coefs, mode, x, y, z, v, g, lh, gh);
}
int oracle_eval_d(const otable *t, const double *coefs, int mode, double x, double y, double z,
                  double *v, double *g, double *lh, double *gh) {
  return eval_d(t, coefs, mode, x, y, z, v, g, lh, gh);
}

/* Many points, OpenMP over points.  Point p reads pos[dim*p ...] and writes
 * vals + p*NR, grads + p*dim*NR, lh + p*nh*NR, gh + p*27*NR (NR = N*ncomp,
 * nh as in eval_). */
int oracle_eval_batch(const otable *t, int is_double, const void *coefs, int mode,
                      const double *pos, int npts, void *vals, void *grads, void *lh,
                      void *gh, int nthreads) {
  const long NR = (long)t->N * t->ncomp;
  const int dim = t->dim;
  const long nh = mode == OR_VGHGH ? 9 : mode == OR_VGH ? (dim == 1 ? 1 : dim * dim)
                : mode == OR_VGL ? 1 : 0;
  int bad = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : bad)
#endif
  for (int p = 0; p < npts; ++p) {
    double x[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) x[d] = pos[(long)dim * p + d];
    if (is_double) {
      double *v = (double *)vals + p * NR;
      double *g = grads ? (double *)grads + dim * p * NR : NULL;
      double *h = lh ? (double *)lh + nh * p * NR : NULL;
      double *G = gh ? (double *)gh + 27 * p * NR : NULL;
      bad += eval_d(t, (const double *)coefs, mode, x[0], x[1], x[2], v, g, h, G) != 0;
    } else {
      float *v = (float *)vals + p * NR;
      float *g = grads ? (float *)grads + dim * p * NR : NULL;
      float *h = lh ? (float *)lh + nh * p * NR : NULL;
      float *G = gh ? (float *)gh + 27 * p * NR : NULL;
      bad += eval_f(t, (const float *)coefs, mode, x[0], x[1], x[2], v, g, h, G) != 0;
    }
  }
  (void)nthreads;
  return bad;
}

/* Basis weights of one axis at fraction t (for the closed-form checks). */
void oracle_basis_d(double tt, double a[4], double da[4], double d2a[4]) {
  const double tp[4] = {tt * tt * tt, tt * tt, tt, 1.0};
  for (int r = 0; r < 4; ++r) {
    a[r] = A_d[4 * r] * tp[0] + A_d[4 * r + 1] * tp[1] + A_d[4 * r + 2] * tp[2] + A_d[4 * r + 3] * tp[3];
    da[r] = dA_d[4 * r] * tp[0] + dA_d[4 * r + 1] * tp[1] + dA_d[4 * r + 2] * tp[2] + dA_d[4 * r + 3] * tp[3];
    d2a[r] = d2A_d[4 * r] * tp[0] + d2A_d[4 * r + 1] * tp[1] + d2A_d[4 * r + 2] * tp[2] + d2A_d[4 * r + 3] * tp[3];
  }
}

void oracle_basis3_d(double d3a[4]) {
  for (int r = 0; r < 4; ++r) d3a[r] = D3_COL[r];
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------
 * Slater-determinant kernels that consume the orbital rows.
 *
 * Row k of A (row-major, stride rs) changes; Ainv is its inverse.
 *   oracle_update_inverse_*: Sherman-Morrison update, as the pair
 *     update_inverse_cuda1/2 of src/determinant_update.h:8-104 computes it:
 *     delta = u - A[k] (A given; A[k] = u is written) or delta = u (A NULL,
 *     src/determinant_update.cu:10-98); d_j = sum_i Ainv[i][j] delta_i summed
 *     over i = 0..N-1 (:30-55); colk = column k of the old Ainv;
 *     prefact = -1 / (1 + d_k); Ainv[i][j] += d_j * (prefact * colk_i)
 *     (:80-103).  The host form src/determinant_update.cu:164-176 is the
 *     same update on the transpose.
 *   oracle_det_ratio_*: sum_i Ainv[i][col] * u_i in row order, the ratio
 *     det(A') / det(A) that calc_ratios computes (src/cuda_walker.cu:150-185).
 */
#define ORACLE_SM(T, NAME)                                                        \
  void NAME(T *A, T *Ainv, const T *u, int N, long rs, int k) {                  \
    T *delta = (T *)malloc(sizeof(T) * (size_t)N);                               \
    T *d = (T *)malloc(sizeof(T) * (size_t)N);                                   \
    T *colk = (T *)malloc(sizeof(T) * (size_t)N);                                \
    for (int i = 0; i < N; ++i) {                                                \
      delta[i] = A ? u[i] - A[k * rs + i] : u[i];                                \
      colk[i] = Ainv[i * rs + k];                                                \
    }                                                                            \
    if (A)                                                                       \
      for (int i = 0; i < N; ++i) A[k * rs + i] = u[i];                          \
    for (int j = 0; j < N; ++j) {                                                \
      T s = 0;                                                                   \
      for (int i = 0; i < N; ++i) s += Ainv[i * rs + j] * delta[i];              \
      d[j] = s;                                                                  \
    }                                                                            \
    const T prefact = (T)-1 / ((T)1 + d[k]);                                     \
    for (int i = 0; i < N; ++i) {                                                \
      const T c = prefact * colk[i];                                             \
      for (int j = 0; j < N; ++j) Ainv[i * rs + j] += d[j] * c;                  \
    }                                                                            \
    free(delta); free(d); free(colk);                                            \
  }
ORACLE_SM(float, oracle_update_inverse_f)
ORACLE_SM(double, oracle_update_inverse_d)

#define ORACLE_RATIO(T, NAME)                                                     \
  T NAME(const T *Ainv, const T *u, int N, long rs, int col) {                   \
    T s = 0;                                                                     \
    for (int i = 0; i < N; ++i) s += Ainv[i * rs + col] * u[i];                  \
    return s;                                                                    \
  }
ORACLE_RATIO(float, oracle_det_ratio_f)
ORACLE_RATIO(double, oracle_det_ratio_d)
