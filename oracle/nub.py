"""Oracle for non-uniform 1-D multi-splines (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module, as the checker.  A pure-Python/numpy
restatement of the reference's non-uniform path, step for step, in double:
  * grids: src/nugrid.c:25-172 (general grid and its bisection; center grid
    c +- a expm1(b k); log grid start exp(a i)) -- the reverse maps clamp to
    [0, num_points - 1] as the library does;
  * basis: src/nubasis.c:24-60 (knots: two before the grid, three after;
    dxInv[3i+j] = 1/(xVals[i+j+1] - xVals[i])) and :520-560 (the Cox-de Boor
    recursion for the functions and their first and second derivatives);
  * solves: src/nubspline_create.c:389-574 (derivative-condition and cyclic
    eliminations) and src/multi_nubspline_create.c:938-965 (complex orbitals as
    two real solves);
  * evaluation: src/multi_nubspline_eval_std_z_impl.h:32-113.
"""
from __future__ import annotations

import math

import numpy as np

PERIODIC, DERIV1, DERIV2, FLAT, NATURAL = 0, 1, 2, 3, 4


class Grid:
    def __init__(self, points, reverse):
        self.points = np.asarray(points, float)
        self._reverse = reverse

    def reverse(self, x):
        n = len(self.points)
        i = self._reverse(x)
        return 0 if i < 0 else (n - 1 if i > n - 1 else i)


def general_grid(points):
    p = np.asarray(points, float)

    def rev(x):
        n = len(p)
        if x <= p[0]:
            return 0
        if x >= p[n - 1]:
            return n - 1
        lo, hi = 0, n - 1
        while hi - lo > 1:
            mid = (lo + hi) >> 1
            if p[mid] > x:
                hi = mid
            else:
                lo = mid
        return lo
    return Grid(p, rev)


def center_grid(start, end, ratio, n):
    c = 0.5 * (start + end)
    h = n // 2
    b = math.log(ratio) / float(h - 1)
    pts = np.empty(n)
    if n % 2:
        even_half = 0.0
        a = 0.5 * (end - start) / math.expm1(b * h)
        for k in range(-h, h + 1):
            pts[k + h] = (-1.0 if k < 0 else 1.0) * a * math.expm1(b * abs(k)) + c
    else:
        even_half = 0.5
        a = 0.5 * (end - start) / math.expm1(b * (-0.5 + h))
        for k in range(-h, h):
            pts[k + h] = (-1.0 if k < 0 else 1.0) * a * math.expm1(b * abs(0.5 + k)) + c
    ainv, binv = 1.0 / a, 1.0 / b

    def rev(x):
        u = x - c
        k = math.copysign(math.log1p(abs(u) * ainv) * binv, u)
        return int(math.floor(h + k - even_half))
    return Grid(pts, rev)


def log_grid(start, end, n):
    a = 1.0 / float(n - 1) * math.log(end / start)
    ainv, sinv = 1.0 / a, 1.0 / start
    pts = np.array([start * math.exp(a * float(i)) for i in range(n)])

    def rev(x):
        return int(math.floor(ainv * math.log(x * sinv)))
    return Grid(pts, rev)


class Basis:
    def __init__(self, grid: Grid, periodic: bool):
        g = grid.points
        N = len(g)
        xv = np.empty(N + 5)
        xv[2:N + 2] = g
        if not periodic:
            xv[0] = g[0] - 2.0 * (g[1] - g[0])
            xv[1] = g[0] - 1.0 * (g[1] - g[0])
            xv[N + 2] = g[N - 1] + 1.0 * (g[N - 1] - g[N - 2])
            xv[N + 3] = g[N - 1] + 2.0 * (g[N - 1] - g[N - 2])
            xv[N + 4] = g[N - 1] + 3.0 * (g[N - 1] - g[N - 2])
        else:
            xv[1] = g[0] - (g[N - 1] - g[N - 2])
            xv[0] = g[0] - (g[N - 1] - g[N - 3])
            xv[N + 2] = g[N - 1] + (g[1] - g[0])
            xv[N + 3] = g[N - 1] + (g[2] - g[0])
            xv[N + 4] = g[N - 1] + (g[3] - g[0])
        dx = np.empty(3 * (N + 2))
        for i in range(N + 2):
            for j in range(3):
                dx[3 * i + j] = 1.0 / (xv[i + j + 1] - xv[i])
        self.grid, self.xv, self.dx, self.periodic = grid, xv, dx, periodic

    def funcs(self, i, x):
        """(a, da, d2a): the four non-zero functions on interval i at x."""
        xv, dx = self.xv.tolist(), self.dx.tolist()
        i2 = i + 2
        b10 = (xv[i2 + 1] - x) * dx[3 * (i + 2) + 0]
        b11 = (x - xv[i2]) * dx[3 * (i + 2) + 0]
        b20 = (xv[i2 + 1] - x) * dx[3 * (i + 1) + 1] * b10
        b21 = ((x - xv[i2 - 1]) * dx[3 * (i + 1) + 1] * b10 + (xv[i2 + 2] - x) * dx[3 * (i + 2) + 1] * b11)
        b22 = (x - xv[i2]) * dx[3 * (i + 2) + 1] * b11
        a = [(xv[i2 + 1] - x) * dx[3 * i + 2] * b20,
             ((x - xv[i2 - 2]) * dx[3 * i + 2] * b20 + (xv[i2 + 2] - x) * dx[3 * (i + 1) + 2] * b21),
             ((x - xv[i2 - 1]) * dx[3 * (i + 1) + 2] * b21 + (xv[i2 + 3] - x) * dx[3 * (i + 2) + 2] * b22),
             (x - xv[i2]) * dx[3 * (i + 2) + 2] * b22]
        da = [-3.0 * (dx[3 * i + 2] * b20),
              3.0 * (dx[3 * i + 2] * b20 - dx[3 * (i + 1) + 2] * b21),
              3.0 * (dx[3 * (i + 1) + 2] * b21 - dx[3 * (i + 2) + 2] * b22),
              3.0 * (dx[3 * (i + 2) + 2] * b22)]
        d2a = [6.0 * (+dx[3 * (i + 0) + 2] * dx[3 * (i + 1) + 1] * b10),
               6.0 * (-dx[3 * (i + 1) + 1] * (dx[3 * (i + 0) + 2] + dx[3 * (i + 1) + 2]) * b10
                      + dx[3 * (i + 1) + 2] * dx[3 * (i + 2) + 1] * b11),
               6.0 * (+dx[3 * (i + 1) + 2] * dx[3 * (i + 1) + 1] * b10
                      - dx[3 * (i + 2) + 1] * (dx[3 * (i + 1) + 2] + dx[3 * (i + 2) + 2]) * b11),
               6.0 * (+dx[3 * (i + 2) + 2] * dx[3 * (i + 2) + 1] * b11)]
        return a, da, d2a

    def at(self, x):
        i = self.grid.reverse(x)
        return (i,) + self.funcs(i, x)


def _solve_deriv(basis: Basis, data, left, right):
    M = len(basis.grid.points)
    N = M + 2
    bd = [[0.0] * 4 for _ in range(N)]
    bd[0] = list(left)
    bd[N - 1] = list(right)
    for i in range(M):
        a, _, _ = basis.funcs(i, float(basis.grid.points[i]))
        bd[i + 1] = [a[0], a[1], a[2], float(data[i])]
    r0, r1 = bd[0], bd[1]
    r0[1] /= r0[0]; r0[2] /= r0[0]; r0[3] /= r0[0]; r0[0] = 1.0
    r1[1] -= r1[0] * r0[1]; r1[2] -= r1[0] * r0[2]; r1[3] -= r1[0] * r0[3]
    r0[0] = 0.0
    r1[2] /= r1[1]; r1[3] /= r1[1]; r1[1] = 1.0
    for r in range(2, N - 1):
        row, up = bd[r], bd[r - 1]
        row[1] -= row[0] * up[2]
        row[3] -= row[0] * up[3]
        row[2] /= row[1]
        row[3] /= row[1]
        row[0] = 0.0
        row[1] = 1.0
    last = bd[M + 1]
    last[1] -= last[0] * bd[M - 1][2]
    last[3] -= last[0] * bd[M - 1][3]
    last[2] -= last[1] * bd[M][2]
    last[3] -= last[1] * bd[M][3]
    last[3] /= last[2]
    last[2] = 1.0
    p = [0.0] * N
    p[M + 1] = last[3]
    for r in range(M, 0, -1):
        p[r] = bd[r][3] - bd[r][2] * p[r + 1]
    p[0] = bd[0][3] - bd[0][1] * p[1] - bd[0][2] * p[2]
    return np.array(p)


def _solve_periodic(basis: Basis, data):
    M = len(basis.grid.points) - 1
    bd = []
    for i in range(M):
        a, _, _ = basis.funcs(i, float(basis.grid.points[i]))
        bd.append([a[0], a[1], a[2], float(data[i])])
    lc = [0.0] * M
    lr, b0 = bd[M - 1], bd[0]
    b0[2] /= b0[1]; b0[0] /= b0[1]; b0[3] /= b0[1]; b0[1] = 1.0
    lr[1] -= lr[2] * b0[0]
    lr[3] -= lr[2] * b0[3]
    lr[2] = -lr[2] * b0[2]
    lc[0] = b0[0]
    for r in range(1, M - 1):
        row, up = bd[r], bd[r - 1]
        row[1] -= row[0] * up[2]
        row[3] -= row[0] * up[3]
        lc[r] = -row[0] * lc[r - 1]
        row[0] = 0.0
        row[2] /= row[1]
        row[3] /= row[1]
        lc[r] /= row[1]
        row[1] = 1.0
        if r < M - 2:
            lr[3] -= lr[2] * row[3]
            lr[1] -= lr[2] * lc[r]
            lr[2] = -lr[2] * row[2]
    lr[0] += lr[2]
    lr[1] -= lr[0] * (bd[M - 2][2] + lc[M - 2])
    lr[3] -= lr[0] * bd[M - 2][3]
    lr[3] /= lr[1]
    p = [0.0] * (M + 3)
    p[M] = lr[3]
    for r in range(M - 2, -1, -1):
        p[r + 1] = bd[r][3] - bd[r][2] * p[r + 2] - lc[r] * p[M]
    p[0] = p[M]
    p[M + 1] = p[1]
    p[M + 2] = p[2]
    return np.array(p)


def find_coefs(basis: Basis, lcode, rcode, lval, rval, data):
    """Coefficients (num_points + 2) of one real orbital."""
    if lcode == PERIODIC:
        return _solve_periodic(basis, data)
    M = len(basis.grid.points)
    if lcode in (FLAT, NATURAL):
        lval = 0.0
    if rcode in (FLAT, NATURAL):
        rval = 0.0
    g0, gl = float(basis.grid.points[0]), float(basis.grid.points[M - 1])
    _, d1, d2 = basis.funcs(0, g0)
    left = list(d1 if lcode in (FLAT, DERIV1) else d2)
    left[3] = lval
    _, d1, d2 = basis.funcs(M - 1, gl)
    right = list(d1 if rcode in (FLAT, DERIV1) else d2)
    right[3] = rval
    return _solve_deriv(basis, data, left, right)


class Spline1D:
    """Multi-orbital 1-D non-uniform spline, coefs[num_points + 2, N]
    (complex for kind z)."""

    def __init__(self, kind, grid: Grid, bc, N):
        self.kind, self.grid, self.N = kind, grid, N
        self.bc = tuple(bc) + (0.0,) * (4 - len(bc))
        self.basis = Basis(grid, bc[0] == PERIODIC)
        dt = np.complex128 if kind == "z" else np.float64
        self.coefs = np.zeros((len(grid.points) + 2, N), dt)

    def set(self, n, data):
        lc, rc, lv, rv = self.bc
        data = np.asarray(data)
        if self.kind == "z":
            lv, rv = complex(lv), complex(rv)
            re = find_coefs(self.basis, lc, rc, lv.real, rv.real, data.real)
            im = find_coefs(self.basis, lc, rc, lv.imag, rv.imag, data.imag)
            self.coefs[:, n] = re + 1j * im
        else:
            self.coefs[:, n] = find_coefs(self.basis, lc, rc, float(lv), float(rv), data)

    def eval(self, mode, x):
        """src/multi_nubspline_eval_std_z_impl.h: 'v' sums a0 c0 + .. + a3 c3
        left to right; 'vg' / 'vgl' accumulate row by row from zero."""
        i, a, da, d2a = self.basis.at(float(x))
        rows = [self.coefs[i + k] if i + k < len(self.coefs) else np.zeros(self.N, self.coefs.dtype)
                for k in range(4)]
        if mode == "v":
            return (a[0] * rows[0] + a[1] * rows[1] + a[2] * rows[2] + a[3] * rows[3],)
        v = np.zeros(self.N, self.coefs.dtype)
        g = np.zeros(self.N, self.coefs.dtype)
        l = np.zeros(self.N, self.coefs.dtype)
        for k in range(4):
            v = v + a[k] * rows[k]
            g = g + da[k] * rows[k]
            l = l + d2a[k] * rows[k]
        return (v, g) if mode == "vg" else (v, g, l)
