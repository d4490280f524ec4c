/*
 * nugrid.h -- non-uniform grids (general, center-refined, logarithmic).
 *
 * Replaces src/nugrid.h:25-95.  The structs keep the reference's layout (the
 * public prefix {code, start, end, points, num_points, reverse_map} is shared
 * by every grid), so code that reads grid->points or calls
 * grid->reverse_map(grid, x) works unchanged.  reverse_map returns the index
 * i of the interval [points[i], points[i+1]) holding x.
 */
#ifndef EINSPLINE_NUGRID_H
#define EINSPLINE_NUGRID_H

#include <stdbool.h>

typedef enum { LINEAR, GENERAL, CENTER, LOG } grid_type;

typedef struct {
  grid_type code;
  double start, end;
  double *points;
  int num_points;
  int (*reverse_map)(void *grid, double x);
} NUgrid;

/* Points cluster around the centre: the spacing grows geometrically by
 * `ratio` overall from the centre to either end. */
typedef struct {
  grid_type code;
  double start, end;
  double *points;
  int num_points;
  int (*reverse_map)(void *grid, double x);
  double a, aInv, b, bInv, center, even_half;
  int half_points, odd_one;
  bool odd;
} center_grid;

/* points[i] = start * exp(a i), a = log(end / start) / (num_points - 1). */
typedef struct {
  grid_type code;
  double start, end;
  double *points;
  int num_points;
  int (*reverse_map)(void *grid, double x);
  double a, ainv, startinv;
} log_grid;

#ifdef __cplusplus
extern "C" {
#endif

NUgrid *create_center_grid(double start, double end, double ratio, int num_points);
NUgrid *create_log_grid(double start, double end, int num_points);
NUgrid *create_general_grid(double *points, int num_points);
void destroy_grid(NUgrid *grid);

#ifdef __cplusplus
}
#endif

#endif /* EINSPLINE_NUGRID_H */
