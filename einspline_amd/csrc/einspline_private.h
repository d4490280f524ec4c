/* einspline_private.h -- library-internal layout shared by the C host code
 * (create/set/destroy) and the HIP runtime. */
#ifndef EINSPLINE_PRIVATE_H
#define EINSPLINE_PRIVATE_H

#include <pthread.h>
#include <stdint.h>
#include "multi_bspline_structs.h"
#include "multi_nubspline.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EINSPLINE_HOST_MAGIC 0x45534833u /* "ESH3" */

/* Every multi-spline returned by create_multi_UBspline_{1,2,3}d_* is the
 * public struct followed by this private tail, so a pointer to the public
 * struct is a pointer to the whole object.  All public structs share the
 * prefix {spcode, tcode, coefs}. */
typedef struct einspline_host3d {
  union {
    multi_UBspline_3d_s s;
    multi_UBspline_3d_d d;
    multi_UBspline_3d_c c;
    multi_UBspline_3d_z z;
    multi_UBspline_2d_s s2;
    multi_UBspline_2d_d d2;
    multi_UBspline_2d_c c2;
    multi_UBspline_2d_z z2;
    multi_UBspline_1d_s s1;
    multi_UBspline_1d_d d1;
    multi_UBspline_1d_c c1;
    multi_UBspline_1d_z z1;
  } pub;
  uint32_t magic;
  int dim;                 /* 1, 2 or 3 */
  int Nx, Ny, Nz;          /* coefficient-table extent per axis (1 past dim) */
  int ncomp;               /* reals per element: 1 real, 2 complex */
  int nsplines;            /* num_splines (its offset differs between the union members) */
  int dirty;               /* host coefs changed since the device mirror was built */
  void *mirror;            /* multi_UBspline_3d_hip*, built on first single-point eval */
  pthread_mutex_t lock;    /* guards dirty/mirror */
} einspline_host3d;

/* Returns the private object behind a public pointer, or NULL if the
 * pointer was not created by this library. */
einspline_host3d *einspline_host_of(void *spline);

/* One axis of a 3-D host spline's interpolation operator, for the GPU
 * solve (build.hip).  `arrays` holds 4*M doubles: the Thomas factors inv,
 * g, sub and the periodic bordering vector w (multi_create.c axis_op). */
typedef struct {
  int M, n, kind;          /* data points, coefficients, 0 periodic / 1 anti / 2 derivative */
  double sgn, zinv;        /* periodic wrap sign, 1 / Schur complement */
  double l0, r2;           /* derivative rows: coefficient of c_0 (left), c_{M+1} (right) */
  double lv[2], rv[2];     /* boundary values per component */
  double *arrays;          /* malloc'd; release with free() */
} einspline_axis_export;

/* Fill ax[0..2] for a 3-D host spline.  Returns 0, -1 (not 3-D or an
 * unsupported boundary pair) or -2 (out of memory). */
int einspline_export_axes(void *spline, einspline_axis_export ax[3]);

/* Non-uniform 1-D multi-splines (nub_create.c): the public struct, then
 * this tail.  `mirror` is the device copy (multi_NUBspline_1d_hip*) used by
 * the single-point evaluation; set_* marks it stale. */
#define EINSPLINE_NUB_MAGIC 0x4e554231u /* "NUB1" */
typedef struct einspline_nub1d {
  union {
    multi_NUBspline_1d_d d;
    multi_NUBspline_1d_z z;
  } pub;
  uint32_t magic;
  int ncomp;               /* 1 (d) or 2 (z) */
  int nrows;               /* coefficient rows: num_points + 2 */
  int dirty;
  void *mirror;
  pthread_mutex_t lock;
} einspline_nub1d;

einspline_nub1d *einspline_nub_of(void *spline);
/* Implemented by the HIP runtime: releases a non-uniform device mirror. */
void einspline_nub_mirror_release(void *mirror);

/* Implemented by the HIP runtime: releases a device mirror. */
void einspline_mirror_release(void *mirror);

#ifdef __cplusplus
}
#endif

#endif
