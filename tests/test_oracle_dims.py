"""Pins for the oracle's 1-D and 2-D splines and its 3-D vghgh, each against
a closed form that does not share the oracle's code path:

* third-derivative basis vs the piecewise-constant third derivative of B3;
* vghgh (value .. third derivatives) vs the full tensor-product B3 sum, and
  its value/gradient/Hessian blocks bitwise equal to vgh;
* 1-D / 2-D set vs dense interpolation maps and the FFT periodic solve, the
  interpolation property at the grid nodes, and eval vs the B3 sum, for
  real and complex tables.
"""
import numpy as np
import pytest

import oracle
from oracle import brute


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_third_derivative_basis_is_the_b3_third_derivative(rng):
    d3a = oracle.basis3()
    for t in rng.random(20):
        s = t - np.arange(4) + 1            # offsets of the 4 stencil rows
        np.testing.assert_allclose(d3a, brute.d3b3(s), atol=1e-15)


@pytest.mark.parametrize("kind", ["d", "z"])
def test_vghgh_matches_brute_force_third_derivatives(kind, rng):
    M, start, end = (6, 5, 7), (3.1, 8.7, 4.5), (9.1, 12.7, 9.3)
    bcs = [(0, 0), (4, 4), (3, 3)]
    t = oracle.Table(kind, M, start, end, 4, bc=bcs)
    c = rng.uniform(-1, 1, t.coefs.shape)
    t.coefs[...] = c
    pts = start + (np.array(end) - start) * rng.random((30, 3)) * 0.999
    cplx = t.coefs.view(np.complex128) if kind == "z" else t.coefs
    v, g, H, T = brute.eval_dense(cplx, start, t.delta_inv, pts)
    ov, og, oh, ogh = t.eval_batch("vghgh", pts)
    assert rel(ov, v) < 1e-13 and rel(og, g) < 1e-13 and rel(oh, H) < 1e-13
    assert rel(ogh, T) < 1e-12
    # the full 27-entry tensor is symmetric
    for perm in ((0, 1, 2, 4, 3), (0, 1, 3, 2, 4), (0, 1, 4, 3, 2)):
        np.testing.assert_array_equal(ogh, ogh.transpose(perm))
    # value, gradient and Hessian are the vgh results exactly
    for a, b in zip(t.eval_batch("vgh", pts), (ov, og, oh)):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
@pytest.mark.parametrize("dim", [1, 2])
def test_low_dim_set_matches_dense_maps(kind, dim, rng):
    M = (9, 7)[:dim]
    start, end = (1.5, -2.0)[:dim], (4.0, 3.5)[:dim]
    bcs = [(oracle.OR_DERIV1, oracle.OR_DERIV2, 0.3 - 0.2j, -0.7 + 0.1j), (oracle.OR_FLAT, oracle.OR_NATURAL)][:dim]
    if kind in "sd":
        bcs = [(b[0], b[1], np.real(b[2]), np.real(b[3])) if len(b) > 2 else b for b in bcs]
    N = 3
    t = oracle.Table(kind, M, start, end, N, bc=bcs)
    cplx = kind in "cz"
    data = rng.uniform(-1, 1, (N,) + M) + (1j * rng.uniform(-1, 1, (N,) + M) if cplx else 0)
    lr = [(b[2], b[3]) if len(b) > 2 else (0.0, 0.0) for b in bcs]
    for n in range(N):
        t.set(n, data[n])
    view = t.coefs.view(np.complex64 if kind == "c" else np.complex128) if cplx else t.coefs
    tol = 1e-5 if kind in "sc" else 1e-12
    for n in range(N):
        ref = brute.solve_dense(data[n].astype(np.complex128 if cplx else float), [b[:2] for b in bcs],
                                t.delta_inv, lr)
        assert rel(view[..., n], ref) < tol


@pytest.mark.parametrize("dim", [1, 2])
def test_low_dim_periodic_set_matches_fft(dim, rng):
    M = (12, 10)[:dim]
    t = oracle.Table("d", M, (0.0,) * dim, (1.0,) * dim, 2)
    data = rng.uniform(-1, 1, (2,) + M)
    for n in range(2):
        t.set(n, data[n])
        assert rel(t.coefs[..., n], brute.solve_fft_periodic(data[n])) < 1e-13


@pytest.mark.parametrize("kind", ["d", "z"])
@pytest.mark.parametrize("dim", [1, 2])
def test_low_dim_eval_matches_brute_force(kind, dim, rng):
    M = (8, 6)[:dim]
    start, end = (0.5, -1.0)[:dim], (3.0, 2.0)[:dim]
    bcs = [(0, 0), (oracle.OR_NATURAL, oracle.OR_FLAT)][:dim]
    t = oracle.Table(kind, M, start, end, 5, bc=bcs)
    t.coefs[...] = rng.uniform(-1, 1, t.coefs.shape)
    pts = np.array(start) + (np.array(end) - np.array(start)) * rng.random((25, dim)) * 0.999
    cview = t.coefs.view(np.complex128) if kind == "z" else t.coefs
    v, g, H, _ = brute.eval_dense(cview, start, t.delta_inv, pts)
    ov, og, oh = t.eval_batch("vgh", pts)
    assert rel(ov, v) < 1e-13 and rel(og, g) < 1e-13
    if dim == 1:
        # 1-D vgh is vgl: the "Hessian" is the second derivative
        assert rel(oh, H[..., 0, 0]) < 1e-13
    else:
        assert rel(oh, H) < 1e-13
    ov2, og2, ol = t.eval_batch("vgl", pts)
    assert rel(ol, np.trace(H, axis1=2, axis2=3)) < 1e-13
    np.testing.assert_array_equal(ov2, ov)
    np.testing.assert_array_equal(og2, og)
    (ov3,) = t.eval_batch("v", pts)
    np.testing.assert_array_equal(ov3, ov)


@pytest.mark.parametrize("kind", ["s", "c"])
@pytest.mark.parametrize("dim", [1, 2])
def test_low_dim_interpolates_the_data(kind, dim, rng):
    M = (11, 9)[:dim]
    t = oracle.Table(kind, M, (0.0,) * dim, (1.0,) * dim, 4)
    cplx = kind == "c"
    data = rng.uniform(-1, 1, (4,) + M) + (1j * rng.uniform(-1, 1, (4,) + M) if cplx else 0)
    for n in range(4):
        t.set(n, data[n])
    nodes = np.stack(np.meshgrid(*[np.arange(m) / m for m in M], indexing="ij"), -1).reshape(-1, dim)
    (v,) = t.eval_batch("v", nodes)
    assert rel(v, data.reshape(4, -1).T) < 3e-6
