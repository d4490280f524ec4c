/*
 * multi_bspline_structs.h -- 3-D uniform multi-spline descriptors,
 * binary-compatible with the reference library.
 *
 * Replaces: src/multi_bspline_structs.h:51-60 (3d_s), :88-97 (3d_d),
 * :128-139 (3d_c), :169-180 (3d_z).
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

#endif /* EINSPLINE_MULTI_BSPLINE_STRUCTS_H */
