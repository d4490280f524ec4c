"""Closed forms that pin the oracle (TEST INFRASTRUCTURE ONLY).

Nothing here follows the reference's code: each function restates the
mathematics the reference implements, in a form chosen so that a plausible
slip in the restatement (a transposed basis row, a wrong stencil offset, a
dropped halo copy, a wrong delta_inv power) gives a different answer.

* B3 kernel sums: the uniform cubic B-spline is
    B3(s) = (2-|s|)^3/6 for 1 <= |s| < 2,  (4 - 6 s^2 + 3|s|^3)/6 for |s| < 1,
  and a spline with stored coefficients c_j is f(u) = sum_j c_j B3(u - j + 1),
  u = (x - start) * delta_inv (the stored index j is the data index + 1).
  The value is the full tensor-product sum over the whole table (1, 2 or 3
  axes), with no stencil selection at all; derivatives up to the third come
  from the piecewise-polynomial derivatives of B3.
* Interpolation maps: dense per-axis linear maps from data to coefficients,
  built from the interpolation conditions and solved with numpy.linalg.
* FFT periodic solve: eigenvalues (4 + 2 cos(2 pi k / M)) / 6 of the
  circulant interpolation matrix.
* Plane waves: analytic value, gradient and Hessian.
"""
from __future__ import annotations

import numpy as np

PERIODIC, DERIV1, DERIV2, FLAT, NATURAL, ANTIPERIODIC = range(6)


def b3(s):
    s = np.abs(np.asarray(s, float))
    return np.where(s < 1, (4 - 6 * s**2 + 3 * s**3) / 6, np.where(s < 2, (2 - s) ** 3 / 6, 0.0))


def db3(s):
    s = np.asarray(s, float)
    a = np.abs(s)
    sg = np.sign(s)
    return sg * np.where(a < 1, (-12 * a + 9 * a**2) / 6, np.where(a < 2, -(2 - a) ** 2 / 2, 0.0))


def d2b3(s):
    a = np.abs(np.asarray(s, float))
    return np.where(a < 1, (-12 + 18 * a) / 6, np.where(a < 2, (2 - a), 0.0))


def d3b3(s):
    """Third derivative: piecewise constant, 3 sign(s) inside |s| < 1 and
    -sign(s) on 1 <= |s| < 2 (undefined at the knots)."""
    s = np.asarray(s, float)
    a = np.abs(s)
    return np.sign(s) * np.where(a < 1, 3.0, np.where(a < 2, -1.0, 0.0))


def axis_weights(x, start, dinv, n):
    """Weights of all n stored coefficients along one axis at coordinate x:
    value and first, second, third derivative (derivatives in physical units)."""
    u = (x - start) * dinv
    j = np.arange(n)
    s = u - j + 1
    return b3(s), db3(s) * dinv, d2b3(s) * dinv**2, d3b3(s) * dinv**3


def eval_dense(coefs, start, dinv, pts):
    """Full tensor-product evaluation in dim = len(start) dimensions.
    coefs: (N_0, .., N_{dim-1}, R) reals or complex.  Returns v (P,R),
    g (P,R,dim), H (P,R,dim,dim) and T (P,R,dim,dim,dim), T the third
    derivatives."""
    dim = len(start)
    shape = coefs.shape[:dim]
    pts = np.asarray(pts, float).reshape(-1, dim)
    P, R = len(pts), coefs.shape[dim]
    v = np.zeros((P, R), coefs.dtype)
    g = np.zeros((P, R, dim), coefs.dtype)
    H = np.zeros((P, R, dim, dim), coefs.dtype)
    T = np.zeros((P, R, dim, dim, dim), coefs.dtype)
    for p in range(P):
        w = [axis_weights(pts[p, a], start[a], dinv[a], shape[a]) for a in range(dim)]

        def con(*axes):
            """Contract with d/dx_{axes[0]} d/dx_{axes[1]} ... (order per axis)."""
            order = [list(axes).count(a) for a in range(dim)]
            out = coefs
            for a in range(dim):
                out = np.tensordot(w[a][order[a]], out, axes=(0, 0))
            return out

        v[p] = con()
        for a in range(dim):
            g[p, :, a] = con(a)
            for b2 in range(dim):
                H[p, :, a, b2] = con(a, b2)
                for c in range(dim):
                    T[p, :, a, b2, c] = con(a, b2, c)
    return v, g, H, T


def interp_map(M, lcode, rcode, dinv):
    """Dense map (n x (M+2)) from [data_0..data_{M-1}, lval, rval] to the n
    stored coefficients of one axis, built from the interpolation conditions."""
    al, be = 1.0 / 6.0, 2.0 / 3.0
    if lcode in (PERIODIC, ANTIPERIODIC):
        n = M + 3
        s = 1.0 if lcode == PERIODIC else -1.0
        A = np.zeros((n, n))
        B = np.zeros((n, M + 2))
        for i in range(M):
            A[i, i], A[i, i + 1], A[i, i + 2] = al, be, al
            B[i, i] = 1.0
        # wrapped copies: c_0 = s c_M, c_{M+1} = s c_1, c_{M+2} = s c_2
        for r, (a, b) in enumerate([(0, M), (M + 1, 1), (M + 2, 2)]):
            A[M + r, a], A[M + r, b] = 1.0, -s
        return np.linalg.solve(A, B)
    n = M + 2

    def row(code):
        if code in (FLAT, DERIV1):
            return np.array([-0.5, 0.0, 0.5]) * dinv
        return np.array([1.0, -2.0, 1.0]) * dinv**2

    A = np.zeros((n, n))
    B = np.zeros((n, M + 2))
    A[0, 0:3] = row(lcode)
    B[0, M] = 0.0 if lcode in (FLAT, NATURAL) else 1.0
    for i in range(M):
        A[i + 1, i:i + 3] = (al, be, al)
        B[i + 1, i] = 1.0
    A[n - 1, n - 3:n] = row(rcode)
    B[n - 1, M + 1] = 0.0 if rcode in (FLAT, NATURAL) else 1.0
    return np.linalg.solve(A, B)


def solve_dense(data, bcs, dinv, lr_vals=None):
    """Coefficients for data (M_0, .., M_{dim-1}[, R]), dim = len(bcs), by
    applying the dense per-axis maps along x, then y, then z.
    lr_vals[a] = (lval, rval)."""
    dim = len(bcs)
    data = np.asarray(data)
    squeeze = data.ndim == dim
    if squeeze:
        data = data[..., None]
    out = data
    for a in range(dim):
        M = data.shape[a]
        P = interp_map(M, bcs[a][0], bcs[a][1], dinv[a])
        lv, rv = (lr_vals[a] if lr_vals is not None else (0.0, 0.0))
        ext_shape = list(out.shape)
        ext_shape[a] = M + 2
        ext = np.zeros(ext_shape, dtype=np.result_type(out.dtype, np.asarray(lv).dtype, float))
        sl = [slice(None)] * (dim + 1)
        sl[a] = slice(0, M)
        ext[tuple(sl)] = out
        sl[a] = M
        ext[tuple(sl)] = lv
        sl[a] = M + 1
        ext[tuple(sl)] = rv
        out = np.moveaxis(np.tensordot(P, np.moveaxis(ext, a, 0), axes=(1, 0)), 0, a)
    return out[..., 0] if squeeze else out


def solve_fft_periodic(data):
    """Periodic coefficients by FFT; data (M_0, .., M_{dim-1}).  Returns
    (M_0+3, .., M_{dim-1}+3)."""
    d = np.asarray(data, float)
    den = np.ones(d.shape)
    for a, M in enumerate(d.shape):
        lam = (4 + 2 * np.cos(2 * np.pi * np.arange(M) / M)) / 6
        sh = [1] * d.ndim
        sh[a] = M
        den = den * lam.reshape(sh)
    x = np.real(np.fft.ifftn(np.fft.fftn(d) / den))
    # stored index j holds x[(j - 1) mod M]
    idx = [(np.arange(M + 3) - 1) % M for M in d.shape]
    return x[np.ix_(*idx)]


def plane_waves(rng, nwaves, kmax, ncomplex=False):
    """Random plane-wave orbital: integer k with |k_i| <= kmax, N(0,1)
    amplitudes.  Returns (K (m,3), a, b) for f = sum a cos(2pi k.r) + b sin(...)
    (real) or (K, c) for f = sum c exp(2 pi i k.r) (complex)."""
    K = rng.integers(-kmax, kmax + 1, size=(nwaves, 3))
    if ncomplex:
        c = rng.standard_normal(nwaves) + 1j * rng.standard_normal(nwaves)
        return K, c
    return K, rng.standard_normal(nwaves), rng.standard_normal(nwaves)


def pw_eval(K, coef, pts, L=1.0):
    """Analytic v, grad, Hessian of sum_m coef_m exp(2 pi i k_m.r / L) at pts
    (complex); for the real form pass coef = a - 1j b and take the real part."""
    kk = 2 * np.pi * np.asarray(K, float) / L            # (m,3)
    ph = np.exp(1j * (pts @ kk.T))                       # (P,m)
    v = ph @ coef
    g = np.einsum("pm,m,md->pd", ph, coef, 1j * kk)
    H = np.einsum("pm,m,md,me->pde", ph, coef, 1j * kk, 1j * kk)
    return v, g, H


def pw_grid(K, coef, M, L=1.0):
    """Plane-wave samples on the periodic grid r = i*L/M, shape (M, M, M)."""
    x = np.arange(M) * L / M
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    v, _, _ = pw_eval(K, coef, r, L)
    return v.reshape(M, M, M)
