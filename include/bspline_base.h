/*
 * bspline_base.h -- basic einspline types, binary-compatible with the
 * reference library.
 *
 * Replaces: src/bspline_base.h:49-102 (enums bc_code / spline_code /
 * type_code, BCtype_{s,d,c,z}, Ugrid, Bspline, destroy_Bspline).
 *
 * Field order, field types and enumerator values are kept identical so that
 * an object built against the reference header has the same layout here.
 * The reference includes an autoconf-generated config.h from this header;
 * this one is self-contained.
 */
#ifndef EINSPLINE_BSPLINE_BASE_H
#define EINSPLINE_BSPLINE_BASE_H

#ifdef __cplusplus
#include <complex>
typedef std::complex<float>  complex_float;
typedef std::complex<double> complex_double;
#else
#include <complex.h>
typedef float complex  complex_float;
typedef double complex complex_double;
#endif

/* Postfix convention: s = float, d = double, c = complex float,
 * z = complex double. */

typedef enum { PERIODIC, DERIV1, DERIV2, FLAT, NATURAL, ANTIPERIODIC } bc_code;

typedef enum {
  U1D,        U2D,        U3D,
  NU1D,       NU2D,       NU3D,
  MULTI_U1D,  MULTI_U2D,  MULTI_U3D,
  MULTI_NU1D, MULTI_NU2D, MULTI_NU3D
} spline_code;

typedef enum { SINGLE_REAL, DOUBLE_REAL, SINGLE_COMPLEX, DOUBLE_COMPLEX } type_code;

/* Boundary condition of one axis: left/right code plus the derivative
 * value used by DERIV1/DERIV2 (ignored for PERIODIC/ANTIPERIODIC, forced to
 * zero for FLAT/NATURAL). */
typedef struct { bc_code lCode, rCode; float  lVal, rVal; } BCtype_s;
typedef struct { bc_code lCode, rCode; double lVal, rVal; } BCtype_d;
typedef struct { bc_code lCode, rCode; float  lVal_r, lVal_i, rVal_r, rVal_i; } BCtype_c;
typedef struct { bc_code lCode, rCode; double lVal_r, lVal_i, rVal_r, rVal_i; } BCtype_z;

/* Uniform grid of `num` data points on [start, end].  delta and delta_inv
 * are filled in by the create_* routines. */
typedef struct {
  double start, end;
  int num;
  double delta, delta_inv;
} Ugrid;

/* Common prefix of every spline object. */
typedef struct {
  spline_code sp_code;
  type_code   t_code;
  void       *coefs;
} Bspline;

#ifdef __cplusplus
extern "C"
#endif
void destroy_Bspline(void *spline);

#endif /* EINSPLINE_BSPLINE_BASE_H */
