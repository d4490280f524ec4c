/* nub_basis.h -- cubic B-splines on a non-uniform knot sequence, for the
 * host (nub_create.c) and the device (nub.hip).
 *
 * The knot sequence is the basis's xVals: the grid points with two knots
 * before and three after.  B-spline j has support [xVals[j], xVals[j+4]);
 * coefficient c of a spline multiplies B-spline c - 1, so on grid interval
 * i (knot span s = i + 2) the four functions that do not vanish multiply
 * coefficients i .. i + 3.
 *
 * The functions and their derivatives come from the triangular scheme of
 * de Boor's recurrence in the form of Piegl and Tiller, "The NURBS Book"
 * (2nd ed., algorithms A2.2 and A2.3): one pass builds the basis values of
 * degree 0 to 3 in the upper half of a 4 x 4 table and the knot differences
 * they divide by in the lower half; derivatives are then combinations of
 * the degree-2 and degree-1 values with those differences.  Knot differences
 * are formed from xVals on the fly.  Evaluated outside its span the result
 * is the span's polynomial piece.
 */
#ifndef EINSPLINE_NUB_BASIS_H
#define EINSPLINE_NUB_BASIS_H

#ifdef __HIPCC__
#define ESK_NUB_FN __host__ __device__ static inline
#else
#define ESK_NUB_FN static inline
#endif

/* out[d][k], d = 0 .. nder (nder <= 2): the d-th derivative of the function
 * multiplying coefficient i + k on grid interval i, at x. */
ESK_NUB_FN void esk_nub_basis(const double *knots, int i, double x, int nder, double out[3][4]) {
  const int s = i + 2;
  double left[4], right[4];
  double t[4][4]; /* t[r][j], r <= j: degree-j value of function r; t[j][r], r < j: knot differences */
  t[0][0] = 1.0;
  for (int j = 1; j <= 3; ++j) {
    left[j] = x - knots[s + 1 - j];
    right[j] = knots[s + j] - x;
    double carry = 0.0;
    for (int r = 0; r < j; ++r) {
      t[j][r] = right[r + 1] + left[j - r];
      const double q = t[r][j - 1] / t[j][r];
      t[r][j] = carry + right[r + 1] * q;
      carry = left[j - r] * q;
    }
    t[j][j] = carry;
  }
  for (int k = 0; k < 4; ++k) out[0][k] = t[k][3];
  /* derivatives: for each function r, the coefficients a of the
   * differences of lower-degree values, two rows used alternately */
  for (int r = 0; r < 4; ++r) {
    double a[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    int cur = 0, nxt = 1;
    a[0][0] = 1.0;
    for (int k = 1; k <= nder; ++k) {
      const int rk = r - k, pk = 3 - k;
      double d = 0.0;
      if (r >= k) {
        a[nxt][0] = a[cur][0] / t[pk + 1][rk];
        d = a[nxt][0] * t[rk][pk];
      }
      const int j1 = rk >= -1 ? 1 : -rk;
      const int j2 = r - 1 <= pk ? k - 1 : 3 - r;
      for (int j = j1; j <= j2; ++j) {
        a[nxt][j] = (a[cur][j] - a[cur][j - 1]) / t[pk + 1][rk + j];
        d += a[nxt][j] * t[rk + j][pk];
      }
      if (r <= pk) {
        a[nxt][k] = -a[cur][k - 1] / t[pk + 1][r];
        d += a[nxt][k] * t[r][pk];
      }
      out[k][r] = d;
      cur = 1 - cur;
      nxt = 1 - nxt;
    }
  }
  /* the recurrence's factors p, p (p - 1) for the first and second derivative */
  for (int k = 0; k < 4; ++k) {
    if (nder >= 1) out[1][k] *= 3.0;
    if (nder >= 2) out[2][k] *= 6.0;
  }
}

/* Grid interval of x on points[0 .. n): the i with points[i] <= x <
 * points[i + 1], 0 below the grid and n - 1 at or past its last point.
 * `guess` (any value; clamped) is where the search starts: the closed-form
 * inverse of a center or log grid lands on i or next to it, so the walk is
 * a step or two; otherwise (guess < 0) the interval is bisected. */
ESK_NUB_FN int esk_nub_interval(const double *points, int n, double x, int guess) {
  if (!(x > points[0])) return 0;      /* also NaN */
  if (x >= points[n - 1]) return n - 1;
  if (guess >= 0) {
    int i = guess > n - 2 ? n - 2 : guess;
    while (i > 0 && points[i] > x) --i;
    while (i < n - 2 && points[i + 1] <= x) ++i;
    return i;
  }
  /* points[lo] <= x < points[lo + len] */
  int lo = 0, len = n - 1;
  while (len > 1) {
    const int half = len >> 1;
    if (points[lo + half] <= x) {
      lo += half;
      len -= half;
    } else {
      len = half;
    }
  }
  return lo;
}

#endif /* EINSPLINE_NUB_BASIS_H */
