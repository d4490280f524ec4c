/*
 * multi_bspline_eval_z.h -- single-point evaluation of 3-D uniform
 * double-precision complex multi-splines.
 *
 * Replaces: src/multi_bspline_eval_z.h:86-109 (3-D prototypes).
 *
 * Each call evaluates all num_splines orbitals at one point (x,y,z) and
 * writes vals[N], grads[3N] (gradient of orbital n at 3n..3n+2), lapl[N]
 * (Laplacian) or hess[9N] (row-major symmetric 3x3 Hessian of orbital n at
 * 9n..9n+8).  Output pointers are host memory.  The work runs on the GPU:
 * the first call after a set_* (or einspline_hip_invalidate) uploads the
 * table, then each call is one kernel launch plus a copy of the results.
 *
 * Points must lie in [start, end] on every axis, as in the reference;
 * points outside are evaluated by extending the edge cell's polynomial
 * (memory-safe, otherwise unspecified).  Errors print to stderr and abort,
 * as the reference's GPU path does.
 *
 * The reference declares these without extern "C" and builds them twice
 * (C and C++ symbols); here they are C symbols for both languages.
 */
#ifndef EINSPLINE_MULTI_BSPLINE_EVAL_Z_H
#define EINSPLINE_MULTI_BSPLINE_EVAL_Z_H

#include "bspline_base.h"
#include "multi_bspline_structs.h"

#ifdef __cplusplus
extern "C" {
#endif

void eval_multi_UBspline_3d_z(multi_UBspline_3d_z *spline,
                               double x, double y, double z,
                               complex_double *vals);

void eval_multi_UBspline_3d_z_vg(multi_UBspline_3d_z *spline,
                                  double x, double y, double z,
                                  complex_double *vals, complex_double *grads);

void eval_multi_UBspline_3d_z_vgl(multi_UBspline_3d_z *spline,
                                   double x, double y, double z,
                                   complex_double *vals, complex_double *grads, complex_double *lapl);

void eval_multi_UBspline_3d_z_vgh(multi_UBspline_3d_z *spline,
                                   double x, double y, double z,
                                   complex_double *vals, complex_double *grads, complex_double *hess);

/* Value, gradient, Hessian and the 27 third derivatives (gradhess[27n + 9a +
 * 3b + c] = d3 f_n / dx_a dx_b dx_c, full symmetric tensor).
 * Replaces: src/multi_bspline_eval_std_z_impl.h:868-1049 (sse: _sse_z_impl.h:1347). */
void eval_multi_UBspline_3d_z_vghgh(multi_UBspline_3d_z *spline,
                                     double x, double y, double z,
                                     complex_double *vals, complex_double *grads, complex_double *hess, complex_double *gradhess);

/* ---- 1-D and 2-D (replaces src/multi_bspline_eval_z.h:28-84; src/multi_bspline_eval_std_z_impl.h:38-479) ----
 * 1-D: vals[N], grads[N], lapl[N]; the 1-D "vgh" is the vgl (hess[N] holds
 * the second derivative), as in the reference.  2-D: vals[N], grads[2N]
 * ([n][2]), lapl[N], hess[4N] (row-major symmetric 2x2 per orbital). */
void eval_multi_UBspline_1d_z(multi_UBspline_1d_z *spline, double x, complex_double *vals);
void eval_multi_UBspline_1d_z_vg(multi_UBspline_1d_z *spline, double x, complex_double *vals, complex_double *grads);
void eval_multi_UBspline_1d_z_vgl(multi_UBspline_1d_z *spline, double x, complex_double *vals, complex_double *grads,
                                  complex_double *lapl);
void eval_multi_UBspline_1d_z_vgh(multi_UBspline_1d_z *spline, double x, complex_double *vals, complex_double *grads,
                                  complex_double *hess);
void eval_multi_UBspline_2d_z(multi_UBspline_2d_z *spline, double x, double y, complex_double *vals);
void eval_multi_UBspline_2d_z_vg(multi_UBspline_2d_z *spline, double x, double y, complex_double *vals,
                                 complex_double *grads);
void eval_multi_UBspline_2d_z_vgl(multi_UBspline_2d_z *spline, double x, double y, complex_double *vals,
                                  complex_double *grads, complex_double *lapl);
void eval_multi_UBspline_2d_z_vgh(multi_UBspline_2d_z *spline, double x, double y, complex_double *vals,
                                  complex_double *grads, complex_double *hess);

#ifdef __cplusplus
}
#endif

#endif
