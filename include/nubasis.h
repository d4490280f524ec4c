/*
 * nubasis.h -- cubic B-spline basis on a non-uniform grid.
 *
 * Replaces src/nubasis.h:25-145 (the std functions).  xVals holds the grid
 * points with two extra knots before and three after; dxInv[3*i+j] =
 * 1 / (xVals[i+j+1] - xVals[i]).  The *_d / *_s functions find the interval
 * of x with the grid's reverse map, return its index i and fill the four
 * basis functions (and first / second derivatives) that are non-zero there:
 * the value at x is sum_k b[k] * coefs[i+k].  The *_di / *_si forms do the
 * same at grid point i.
 */
#ifndef EINSPLINE_NUBASIS_H
#define EINSPLINE_NUBASIS_H

#include <stdbool.h>
#include "nugrid.h"

typedef struct {
  NUgrid *grid;
  double *xVals;
  double *dxInv;
  bool periodic;
} NUBasis;

#ifdef __cplusplus
extern "C" {
#endif

NUBasis *create_NUBasis(NUgrid *grid, bool periodic);
void destroy_NUBasis(NUBasis *basis);

int get_NUBasis_funcs_s(NUBasis *basis, double x, float bfuncs[4]);
void get_NUBasis_funcs_si(NUBasis *basis, int i, float bfuncs[4]);
int get_NUBasis_dfuncs_s(NUBasis *basis, double x, float bfuncs[4], float dbfuncs[4]);
void get_NUBasis_dfuncs_si(NUBasis *basis, int i, float bfuncs[4], float dbfuncs[4]);
int get_NUBasis_d2funcs_s(NUBasis *basis, double x, float bfuncs[4], float dbfuncs[4], float d2bfuncs[4]);
void get_NUBasis_d2funcs_si(NUBasis *basis, int i, float bfuncs[4], float dbfuncs[4], float d2bfuncs[4]);

int get_NUBasis_funcs_d(NUBasis *basis, double x, double bfuncs[4]);
void get_NUBasis_funcs_di(NUBasis *basis, int i, double bfuncs[4]);
int get_NUBasis_dfuncs_d(NUBasis *basis, double x, double bfuncs[4], double dbfuncs[4]);
void get_NUBasis_dfuncs_di(NUBasis *basis, int i, double bfuncs[4], double dbfuncs[4]);
int get_NUBasis_d2funcs_d(NUBasis *basis, double x, double bfuncs[4], double dbfuncs[4], double d2bfuncs[4]);
void get_NUBasis_d2funcs_di(NUBasis *basis, int i, double bfuncs[4], double dbfuncs[4], double d2bfuncs[4]);

#ifdef __cplusplus
}
#endif

#endif /* EINSPLINE_NUBASIS_H */
