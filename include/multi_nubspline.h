/*
 * multi_nubspline.h -- multi-orbital splines on non-uniform grids.
 *
 * Replaces src/multi_nubspline.h, src/multi_nubspline_structs.h:145-162 and
 * src/multi_nubspline_create.h / src/multi_nubspline_eval_z.h for the 1-D
 * splines.  The reference can evaluate only 1-D complex-double non-uniform
 * multi-splines (src/multi_nubspline_eval_std_z_impl.h:32-113; its 2-D and
 * 3-D bodies are commented out), so this library provides the 1-D d and z
 * create/set and the 1-D z evaluation, run on the GPU.
 *
 * Layout, as in the reference's non-SSE build: coefs[ix * x_stride + n],
 * x_stride = num_splines, ix < x_grid->num_points + 2.
 */
#ifndef EINSPLINE_MULTI_NUBSPLINE_H
#define EINSPLINE_MULTI_NUBSPLINE_H

#include <stdint.h>
#include "bspline_base.h"
#include "nubasis.h"

typedef struct {
  spline_code spcode;
  type_code tcode;
  double *coefs;
  intptr_t x_stride;
  BCtype_d xBC;
  int num_splines;
  NUgrid *x_grid;
  NUBasis *x_basis;
} multi_NUBspline_1d_d;

typedef struct {
  spline_code spcode;
  type_code tcode;
  complex_double *coefs;
  intptr_t x_stride;
  BCtype_z xBC;
  int num_splines;
  NUgrid *x_grid;
  NUBasis *x_basis;
} multi_NUBspline_1d_z;

#ifdef __cplusplus
extern "C" {
#endif

/* src/multi_nubspline_create.c:643-700, :898-965.  The grid is borrowed,
 * not copied: it must outlive the spline. */
multi_NUBspline_1d_d *create_multi_NUBspline_1d_d(NUgrid *x_grid, BCtype_d xBC, int num_splines);
void set_multi_NUBspline_1d_d(multi_NUBspline_1d_d *spline, int spline_num, double *data);
void set_multi_NUBspline_1d_d_BC(multi_NUBspline_1d_d *spline, int spline_num, double *data, BCtype_d xBC);
multi_NUBspline_1d_z *create_multi_NUBspline_1d_z(NUgrid *x_grid, BCtype_z xBC, int num_splines);
void set_multi_NUBspline_1d_z(multi_NUBspline_1d_z *spline, int spline_num, complex_double *data);
void set_multi_NUBspline_1d_z_BC(multi_NUBspline_1d_z *spline, int spline_num, complex_double *data,
                                 BCtype_z xBC);
/* Frees the spline, its coefficients, its basis and its device mirror (not
 * the grid). */
void destroy_multi_NUBspline(void *spline);

/* src/multi_nubspline_eval_std_z_impl.h:32-113: values, d/dx, d2/dx2 of
 * every orbital at x; the 1-D vgh is the vgl.  Evaluated on the GPU through
 * a device mirror that set_* refreshes. */
void eval_multi_NUBspline_1d_z(multi_NUBspline_1d_z *spline, double x, complex_double *vals);
void eval_multi_NUBspline_1d_z_vg(multi_NUBspline_1d_z *spline, double x, complex_double *vals,
                                  complex_double *grads);
void eval_multi_NUBspline_1d_z_vgl(multi_NUBspline_1d_z *spline, double x, complex_double *vals,
                                   complex_double *grads, complex_double *lapl);
void eval_multi_NUBspline_1d_z_vgh(multi_NUBspline_1d_z *spline, double x, complex_double *vals,
                                   complex_double *grads, complex_double *hess);

/* ---- device side (no counterpart in the reference) ---------------------- */
typedef struct multi_NUBspline_1d_hip multi_NUBspline_1d_hip;
/* Device copy of the coefficients, grid and basis of a 1-D z spline. */
multi_NUBspline_1d_hip *create_multi_NUBspline_1d_z_hip(multi_NUBspline_1d_z *spline);
void destroy_multi_NUBspline_1d_hip(multi_NUBspline_1d_hip *table);
/* Batched evaluation at pos_d[0..num) (device doubles): mode 0 v, 1 vg,
 * 2 vgl (= vgh); outputs [num][num_splines] complex doubles (grads_d,
 * lapl_d unused below their mode).  Enqueued on `stream`, no wait. */
int einspline_eval_nub_1d_z_batch(const multi_NUBspline_1d_hip *table, const double *pos_d, int num,
                                  void *vals_d, void *grads_d, void *lapl_d, int mode, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* EINSPLINE_MULTI_NUBSPLINE_H */
