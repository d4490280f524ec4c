/* oracle.h -- CPU restatement of einspline's 1-D, 2-D and 3-D multi-spline create/eval.
 * TEST INFRASTRUCTURE ONLY (see oracle.c). */
#ifndef EINSPLINE_ORACLE_H
#define EINSPLINE_ORACLE_H
#include <stddef.h>

/* boundary codes, same values as bc_code in src/bspline_base.h:49 */
enum { OR_PERIODIC = 0, OR_DERIV1, OR_DERIV2, OR_FLAT, OR_NATURAL, OR_ANTIPERIODIC };
/* evaluation modes */
enum { OR_V = 0, OR_VG = 1, OR_VGL = 2, OR_VGH = 3, OR_VGHGH = 4 };

typedef struct {
  int M[3];              /* data points per axis (1 on axes >= dim) */
  int Nn[3];             /* coefficients per axis */
  double start[3], end[3], delta[3], delta_inv[3];
  int lcode[3], rcode[3];
  double lval[3][2], rval[3][2];   /* derivative-condition values (re, im) */
  int N;                 /* orbitals */
  int ncomp;             /* 1 real, 2 complex */
  long xs, ys, zs;       /* strides in elements (complex element = ncomp reals) */
  int dim;               /* 1, 2 or 3 */
} otable;

int oracle_table_init(otable *t, const int M[3], const double start[3], const double end[3],
                      const int lcode[3], const int rcode[3], const double lval[6],
                      const double rval[6], int N, int ncomp, int dim);
size_t oracle_table_elems(const otable *t);

void oracle_set_s(const otable *t, float *coefs, int n, const float *data);
void oracle_set_d(const otable *t, double *coefs, int n, const double *data);
void oracle_set_c(const otable *t, float *coefs, int n, const float *data);
void oracle_set_z(const otable *t, double *coefs, int n, const double *data);

/* Output blocks per orbital: vals 1, grads dim, lh 1 (vgl) or dim*dim (vgh; 1 for
 * dim 1, whose vgh is its vgl), gh 27 (vghgh, dim 3 only); complex: x2 interleaved. */
int oracle_eval_s(const otable *t, const float *coefs, int mode, double x, double y, double z,
                  float *v, float *g, float *lh, float *gh);
int oracle_eval_d(const otable *t, const double *coefs, int mode, double x, double y, double z,
                  double *v, double *g, double *lh, double *gh);
/* pos holds dim coordinates per point. */
int oracle_eval_batch(const otable *t, int is_double, const void *coefs, int mode,
                      const double *pos, int npts, void *vals, void *grads, void *lh,
                      void *gh, int nthreads);
void oracle_basis_d(double t, double a[4], double da[4], double d2a[4]);
void oracle_basis3_d(double d3a[4]);
int oracle_max_threads(void);

/* Slater-determinant kernels (row-major matrices with row stride rs):
 * Sherman-Morrison update of Ainv after row k of A becomes u (A non-NULL; A
 * is updated) or changes by u (A NULL), and the ratio sum_i Ainv[i][col] u_i. */
void oracle_update_inverse_f(float *A, float *Ainv, const float *u, int N, long rs, int k);
void oracle_update_inverse_d(double *A, double *Ainv, const double *u, int N, long rs, int k);
float oracle_det_ratio_f(const float *Ainv, const float *u, int N, long rs, int col);
double oracle_det_ratio_d(const double *Ainv, const double *u, int N, long rs, int col);
#endif
