/*
 * multi_bspline_structs.h -- 1-D, 2-D and 3-D uniform multi-spline
 * descriptors, binary-compatible with the reference library.
 *
 * Replaces: src/multi_bspline_structs.h:29-38 (1d_s), :40-49 (2d_s),
 * :51-60 (3d_s), :66-75 (1d_d), :77-86 (2d_d), :88-97 (3d_d),
 * :104-113 (1d_c), :115-126 (2d_c), :128-139 (3d_c), :145-154 (1d_z),
 * :156-167 (2d_z), :169-180 (3d_z).
 *
 * 1-D and 2-D tables use the same scheme with fewer axes: coefs[ix][n]
 * (x_stride = num_splines) and coefs[ix][iy][n] (y_stride = num_splines,
 * x_stride = Ny*y_stride).
 *
 * Coefficients are stored as coefs[ix][iy][iz][n] with the orbital index n
 * fastest: element (ix,iy,iz,n) lives at
 *   coefs[ix*x_stride + iy*y_stride + iz*z_stride + n]
 * with z_stride = num_splines, y_stride = Nz*z_stride, x_stride = Ny*y_stride
 * and N{x,y,z} = M+3 for PERIODIC/ANTIPERIODIC axes, M+2 otherwise
 * (M = grid.num).  Strides count elements of the spline's own type (one
 * complex number is one element for c/z).
 */
#ifndef EINSPLINE_MULTI_BSPLINE_STRUCTS_H
#define EINSPLINE_MULTI_BSPLINE_STRUCTS_H

#include <stdint.h>
#include "bspline_base.h"

typedef struct {
  spline_code spcode;
  type_code   tcode;
  float      *coefs;
  intptr_t    x_stride, y_stride, z_stride;
  Ugrid       x_grid, y_grid, z_grid;
  BCtype_s    xBC, yBC, zBC;
  int         num_splines;
} multi_UBspline_3d_s;

typedef struct {
  spline_code spcode;
  type_code   tcode;
  double     *coefs;
  intptr_t    x_stride, y_stride, z_stride;
  Ugrid       x_grid, y_grid, z_grid;
  BCtype_d    xBC, yBC, zBC;
  int         num_splines;
} multi_UBspline_3d_d;

typedef struct {
  spline_code    spcode;
  type_code      tcode;
  complex_float *coefs;
  intptr_t       x_stride, y_stride, z_stride;
  Ugrid          x_grid, y_grid, z_grid;
  BCtype_c       xBC, yBC, zBC;
  int            num_splines;
  /* Scratch of 3*num_splines elements.  The reference's vgl writes its
   * Laplacian partials here; this library allocates it for layout
   * compatibility only and never writes it, so evaluation is re-entrant. */
  complex_float *lapl3;
} multi_UBspline_3d_c;

typedef struct {
  spline_code     spcode;
  type_code       tcode;
  complex_double *coefs;
  intptr_t        x_stride, y_stride, z_stride;
  Ugrid           x_grid, y_grid, z_grid;
  BCtype_z        xBC, yBC, zBC;
  int             num_splines;
  complex_double *lapl3;   /* see multi_UBspline_3d_c::lapl3 */
} multi_UBspline_3d_z;

/* ---- 1-D and 2-D ----------------------------------------------------- */

typedef struct {
  spline_code spcode;
  type_code   tcode;
  float      *coefs;
  intptr_t    x_stride;
  Ugrid       x_grid;
  BCtype_s    xBC;
  int         num_splines;
} multi_UBspline_1d_s;

typedef struct {
  spline_code spcode;
  type_code   tcode;
  float      *coefs;
  intptr_t    x_stride, y_stride;
  Ugrid       x_grid, y_grid;
  BCtype_s    xBC, yBC;
  int         num_splines;
} multi_UBspline_2d_s;

typedef struct {
  spline_code spcode;
  type_code   tcode;
  double     *coefs;
  intptr_t    x_stride;
  Ugrid       x_grid;
  BCtype_d    xBC;
  int         num_splines;
} multi_UBspline_1d_d;

typedef struct {
  spline_code spcode;
  type_code   tcode;
  double     *coefs;
  intptr_t    x_stride, y_stride;
  Ugrid       x_grid, y_grid;
  BCtype_d    xBC, yBC;
  int         num_splines;
} multi_UBspline_2d_d;

typedef struct {
  spline_code    spcode;
  type_code      tcode;
  complex_float *coefs;
  intptr_t       x_stride;
  Ugrid          x_grid;
  BCtype_c       xBC;
  int            num_splines;
} multi_UBspline_1d_c;

typedef struct {
  spline_code    spcode;
  type_code      tcode;
  complex_float *coefs;
  intptr_t       x_stride, y_stride;
  Ugrid          x_grid, y_grid;
  BCtype_c       xBC, yBC;
  int            num_splines;
  complex_float *lapl2;    /* allocated for layout compatibility, never written */
} multi_UBspline_2d_c;

typedef struct {
  spline_code     spcode;
  type_code       tcode;
  complex_double *coefs;
  intptr_t        x_stride;
  Ugrid           x_grid;
  BCtype_z        xBC;
  int             num_splines;
} multi_UBspline_1d_z;

typedef struct {
  spline_code     spcode;
  type_code       tcode;
  complex_double *coefs;
  intptr_t        x_stride, y_stride;
  Ugrid           x_grid, y_grid;
  BCtype_z        xBC, yBC;
  int             num_splines;
  complex_double *lapl2;   /* allocated for layout compatibility, never written */
} multi_UBspline_2d_z;

#endif /* EINSPLINE_MULTI_BSPLINE_STRUCTS_H */
