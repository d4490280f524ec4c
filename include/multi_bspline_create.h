/*
 * multi_bspline_create.h -- construction of 1-D, 2-D and 3-D uniform
 * multi-splines.
 *
 * Replaces: src/multi_bspline_create.h:36-174 (1d/2d/3d for s, d, c, z);
 * destroy via src/bspline_base.h:98-102.
 *
 * create_* allocates the descriptor and an (uninitialised) coefficient
 * table; set_* solves the interpolation problem for one orbital and writes
 * its column of the table.  The solve is always carried out in double
 * precision (single-precision tables are rounded once at the end).
 *
 * Error behaviour follows the reference: out-of-memory prints to stderr and
 * aborts.  Invalid arguments (num < 2 on an axis, num_splines < 1, spline
 * index out of range) print to stderr; create_* then returns NULL and set_*
 * leaves the table unchanged.
 */
#ifndef EINSPLINE_MULTI_BSPLINE_CREATE_H
#define EINSPLINE_MULTI_BSPLINE_CREATE_H

#include "bspline_base.h"
#include "multi_bspline_structs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/multi_bspline_create.c:251-321 */
multi_UBspline_3d_s *
create_multi_UBspline_3d_s(Ugrid x_grid, Ugrid y_grid, Ugrid z_grid,
                           BCtype_s xBC, BCtype_s yBC, BCtype_s zBC,
                           int num_splines);
/* src/multi_bspline_create.c:323-373; data is float[Mx][My][Mz] (C order) */
void set_multi_UBspline_3d_s(multi_UBspline_3d_s *spline, int spline_num,
                             float *data);

/* src/multi_bspline_create.c:941-1018 */
multi_UBspline_3d_d *
create_multi_UBspline_3d_d(Ugrid x_grid, Ugrid y_grid, Ugrid z_grid,
                           BCtype_d xBC, BCtype_d yBC, BCtype_d zBC,
                           int num_splines);
/* src/multi_bspline_create.c:1020-1073 */
void set_multi_UBspline_3d_d(multi_UBspline_3d_d *spline, int spline_num,
                             double *data);

/* src/multi_bspline_create.c:576-648 */
multi_UBspline_3d_c *
create_multi_UBspline_3d_c(Ugrid x_grid, Ugrid y_grid, Ugrid z_grid,
                           BCtype_c xBC, BCtype_c yBC, BCtype_c zBC,
                           int num_splines);
/* src/multi_bspline_create.c:650-732; data is complex_float[Mx][My][Mz] */
void set_multi_UBspline_3d_c(multi_UBspline_3d_c *spline, int spline_num,
                             complex_float *data);

/* src/multi_bspline_create.c:1317-1390 */
multi_UBspline_3d_z *
create_multi_UBspline_3d_z(Ugrid x_grid, Ugrid y_grid, Ugrid z_grid,
                           BCtype_z xBC, BCtype_z yBC, BCtype_z zBC,
                           int num_splines);
/* src/multi_bspline_create.c:1392-1477 */
void set_multi_UBspline_3d_z(multi_UBspline_3d_z *spline, int spline_num,
                             complex_double *data);

/* Block setters with no reference counterpart: solve `count` consecutive
 * orbitals starting at `first` in one pass.  data holds `count` arrays of
 * Mx*My*Mz values back to back (the layout `count` calls of set_* would
 * read).  Results are identical to `count` set_* calls. */
void set_multi_UBspline_3d_s_block(multi_UBspline_3d_s *spline, int first,
                                   int count, const float *data);
void set_multi_UBspline_3d_d_block(multi_UBspline_3d_d *spline, int first,
                                   int count, const double *data);
void set_multi_UBspline_3d_c_block(multi_UBspline_3d_c *spline, int first,
                                   int count, const complex_float *data);
void set_multi_UBspline_3d_z_block(multi_UBspline_3d_z *spline, int first,
                                   int count, const complex_double *data);

/* ---- 1-D and 2-D ------------------------------------------------------
 * Same rules per axis as 3-D.  data is T[Mx] (1-D) or T[Mx][My] (2-D). */

/* src/multi_bspline_create.c:88-139, :141-147 */
multi_UBspline_1d_s *create_multi_UBspline_1d_s(Ugrid x_grid, BCtype_s xBC, int num_splines);
void set_multi_UBspline_1d_s(multi_UBspline_1d_s *spline, int spline_num, float *data);
/* src/multi_bspline_create.c:152-214, :216-250 */
multi_UBspline_2d_s *create_multi_UBspline_2d_s(Ugrid x_grid, Ugrid y_grid, BCtype_s xBC,
                                                BCtype_s yBC, int num_splines);
void set_multi_UBspline_2d_s(multi_UBspline_2d_s *spline, int spline_num, float *data);

/* src/multi_bspline_create.c:767-820, :822-828 */
multi_UBspline_1d_d *create_multi_UBspline_1d_d(Ugrid x_grid, BCtype_d xBC, int num_splines);
void set_multi_UBspline_1d_d(multi_UBspline_1d_d *spline, int spline_num, double *data);
/* src/multi_bspline_create.c:830-837: set with this orbital's own boundary conditions */
void set_multi_UBspline_1d_d_BC(multi_UBspline_1d_d *spline, int spline_num, double *data,
                                BCtype_d xBC);
/* src/multi_bspline_create.c:840-901, :903-935 */
multi_UBspline_2d_d *create_multi_UBspline_2d_d(Ugrid x_grid, Ugrid y_grid, BCtype_d xBC,
                                                BCtype_d yBC, int num_splines);
void set_multi_UBspline_2d_d(multi_UBspline_2d_d *spline, int spline_num, double *data);

/* src/multi_bspline_create.c:389-432, :434-454 */
multi_UBspline_1d_c *create_multi_UBspline_1d_c(Ugrid x_grid, BCtype_c xBC, int num_splines);
void set_multi_UBspline_1d_c(multi_UBspline_1d_c *spline, int spline_num, complex_float *data);
/* src/multi_bspline_create.c:456-520, :522-574 */
multi_UBspline_2d_c *create_multi_UBspline_2d_c(Ugrid x_grid, Ugrid y_grid, BCtype_c xBC,
                                                BCtype_c yBC, int num_splines);
void set_multi_UBspline_2d_c(multi_UBspline_2d_c *spline, int spline_num, complex_float *data);

/* src/multi_bspline_create.c:1091-1136, :1138-1166 */
multi_UBspline_1d_z *create_multi_UBspline_1d_z(Ugrid x_grid, BCtype_z xBC, int num_splines);
void set_multi_UBspline_1d_z(multi_UBspline_1d_z *spline, int spline_num, complex_double *data);
/* src/multi_bspline_create.c:1168-1197 */
void set_multi_UBspline_1d_z_BC(multi_UBspline_1d_z *spline, int spline_num,
                                complex_double *data, BCtype_z xBC);
/* src/multi_bspline_create.c:1199-1255, :1257-1315 */
multi_UBspline_2d_z *create_multi_UBspline_2d_z(Ugrid x_grid, Ugrid y_grid, BCtype_z xBC,
                                                BCtype_z yBC, int num_splines);
void set_multi_UBspline_2d_z(multi_UBspline_2d_z *spline, int spline_num, complex_double *data);

#ifdef __cplusplus
}
#endif

#endif /* EINSPLINE_MULTI_BSPLINE_CREATE_H */
