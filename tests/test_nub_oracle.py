"""Non-uniform 1-D splines on the CPU: the oracle (oracle/nub.py) pinned by
what the mathematics fixes, and the library's host code (grids, basis,
set_multi_NUBspline_1d_*), which computes them its own way (de Boor's
triangular scheme, a factored collocation operator), checked against the
oracle to rounding (1e-14 relative for the basis, 1e-13 for coefficients).

Pins of the oracle:
  * on a uniform grid the non-uniform basis is the uniform cubic B-spline
    basis (oracle.basis, itself pinned by closed forms), derivatives scaled
    by 1/h and 1/h^2;
  * partition of unity, derivatives of the basis against central
    differences, on random non-uniform grids;
  * a cubic polynomial with its exact end derivatives (DERIV1) is reproduced
    everywhere; every BC kind interpolates its data at the grid points and
    meets its end condition; periodic splines are C2 across the wrap;
  * grid formulas: center grids are symmetric with the stated end points,
    log grids have a constant ratio, reverse maps bracket x.
"""
import math

import numpy as np
import pytest

import einspline_amd as ea
import oracle
from oracle import nub

rng0 = np.random.default_rng(2026)


def random_grid(n, rng, lo=-1.0, hi=2.0):
    steps = rng.uniform(0.3, 1.7, n - 1)
    pts = np.concatenate([[0.0], np.cumsum(steps)])
    return lo + (hi - lo) * pts / pts[-1]


def test_uniform_grid_basis_is_the_uniform_basis():
    h = 0.25
    g = nub.general_grid(np.arange(12) * h)
    b = nub.Basis(g, False)
    for x in np.linspace(0.0, 11 * h - 1e-9, 57):
        i, a, da, d2a = b.at(x)
        t = x / h - i
        ua, uda, ud2a = oracle.basis(t)
        np.testing.assert_allclose(a, ua, atol=1e-14)
        np.testing.assert_allclose(da, np.asarray(uda) / h, atol=1e-12)
        np.testing.assert_allclose(d2a, np.asarray(ud2a) / h**2, atol=1e-10)


@pytest.mark.parametrize("periodic", [False, True])
def test_partition_of_unity_and_derivatives(periodic):
    rng = np.random.default_rng(5)
    g = nub.general_grid(random_grid(15, rng))
    b = nub.Basis(g, periodic)
    for x in rng.uniform(g.points[0], g.points[-1], 40):
        i, a, da, d2a = b.at(x)
        assert abs(sum(a) - 1.0) < 1e-14
        assert abs(sum(da)) < 1e-11 and abs(sum(d2a)) < 1e-9
        eps = 1e-6
        ap = b.funcs(i, x + eps)[0]
        am = b.funcs(i, x - eps)[0]
        fd = (np.asarray(ap) - np.asarray(am)) / (2 * eps)
        np.testing.assert_allclose(da, fd, atol=1e-6)
        dp = b.funcs(i, x + eps)[1]
        dm = b.funcs(i, x - eps)[1]
        np.testing.assert_allclose(d2a, (np.asarray(dp) - np.asarray(dm)) / (2 * eps), atol=1e-5)


def test_cubic_is_reproduced_exactly():
    rng = np.random.default_rng(9)
    pts = random_grid(13, rng, 0.5, 3.0)
    c = [0.3, -1.2, 0.7, 0.25]
    f = lambda x: ((c[3] * x + c[2]) * x + c[1]) * x + c[0]
    df = lambda x: (3 * c[3] * x + 2 * c[2]) * x + c[1]
    d2f = lambda x: 6 * c[3] * x + 2 * c[2]
    s = nub.Spline1D("d", nub.general_grid(pts), (nub.DERIV1, nub.DERIV1, df(pts[0]), df(pts[-1])), 1)
    s.set(0, f(pts))
    for x in np.linspace(pts[0], pts[-1], 61):
        v, g, l = s.eval("vgl", x)
        assert abs(v[0] - f(x)) < 1e-12 and abs(g[0] - df(x)) < 1e-10 and abs(l[0] - d2f(x)) < 1e-8


@pytest.mark.parametrize("bc", [(nub.NATURAL, nub.NATURAL), (nub.FLAT, nub.DERIV2, 0.0, 1.5),
                                (nub.DERIV1, nub.FLAT, -0.4, 0.0), (nub.PERIODIC, nub.PERIODIC)])
def test_interpolation_and_end_conditions(bc):
    rng = np.random.default_rng(13)
    pts = random_grid(17, rng)
    s = nub.Spline1D("d", nub.general_grid(pts), bc, 2)
    M = len(pts) - 1 if bc[0] == nub.PERIODIC else len(pts)
    data = rng.uniform(-1, 1, (2, M))
    for n in range(2):
        s.set(n, data[n])
    for k in range(M):
        v, = s.eval("v", pts[k])
        np.testing.assert_allclose(v, data[:, k], atol=1e-13)
    lc, rc = bc[0], bc[1]
    lv = bc[2] if len(bc) > 2 else 0.0
    rv = bc[3] if len(bc) > 3 else 0.0
    if lc == nub.PERIODIC:
        # C2 across the wrap: value, slope and curvature agree at both ends
        v0, g0, l0 = s.eval("vgl", pts[0])
        v1, g1, l1 = s.eval("vgl", pts[-1] - 1e-12)
        np.testing.assert_allclose(v0, v1, atol=1e-9)
        np.testing.assert_allclose(g0, g1, atol=1e-7)
        np.testing.assert_allclose(l0, l1, atol=1e-4)
        return
    _, g0, l0 = s.eval("vgl", pts[0])
    _, g1, l1 = s.eval("vgl", pts[-1] - 1e-13)
    left = g0 if lc in (nub.DERIV1, nub.FLAT) else l0
    right = g1 if rc in (nub.DERIV1, nub.FLAT) else l1
    np.testing.assert_allclose(left, [lv] * 2, atol=1e-9)
    np.testing.assert_allclose(right, [rv] * 2, atol=1e-6)


def test_grid_formulas():
    g = nub.center_grid(-2.0, 3.0, 6.0, 21)
    p = g.points
    assert abs(p[0] + 2.0) < 1e-13 and abs(p[-1] - 3.0) < 1e-13
    np.testing.assert_allclose(p - 0.5, -(p[::-1] - 0.5), atol=1e-13)
    sp = np.diff(p)
    assert abs(sp[-1] / sp[len(sp) // 2] - 6.0) < 0.8          # spacing grows about `ratio` to the ends
    ge = nub.center_grid(-2.0, 3.0, 6.0, 20)
    assert abs(ge.points[0] + 2.0) < 1e-13 and abs(ge.points[-1] - 3.0) < 1e-13
    lg = nub.log_grid(0.01, 10.0, 30)
    r = lg.points[1:] / lg.points[:-1]
    np.testing.assert_allclose(r, r[0], rtol=1e-12)
    for grid in (g, ge, lg, nub.general_grid(random_grid(11, rng0))):
        q = grid.points
        for x in np.random.default_rng(1).uniform(q[0], q[-1], 200):
            i = grid.reverse(x)
            assert q[i] <= x * (1 + 1e-14) + 1e-14 and (i == len(q) - 1 or x < q[i + 1] * (1 + 1e-14) + 1e-14)


# ---- the library's host code against the oracle -----------------------------

def lib_grid(kind):
    if kind == "general":
        pts = random_grid(14, np.random.default_rng(3))
        return ea.NUGrid.general(pts), nub.general_grid(pts)
    if kind == "center":
        return ea.NUGrid.center(-1.0, 1.5, 5.0, 19), nub.center_grid(-1.0, 1.5, 5.0, 19)
    return ea.NUGrid.log(0.05, 4.0, 16), nub.log_grid(0.05, 4.0, 16)


@pytest.mark.parametrize("kind", ["general", "center", "log"])
def test_library_grids_and_basis_match_the_oracle(kind):
    lg, og = lib_grid(kind)
    np.testing.assert_allclose(lg.points, og.points, rtol=1e-15, atol=0)
    og = nub.general_grid(lg.points)  # same points from here on
    q = lg.points
    for periodic in (False, True):
        b = ea.lib.create_NUBasis(lg.ptr, periodic)
        ob = nub.Basis(og, periodic)
        # the documented knots and inverse-span table
        bs = b.contents
        np.testing.assert_array_equal(np.ctypeslib.as_array(bs.xVals, (len(q) + 5,)), ob.xv)
        np.testing.assert_allclose(np.ctypeslib.as_array(bs.dxInv, (3 * (len(q) + 2),)), ob.dx, rtol=1e-15)
        xs = np.concatenate([np.random.default_rng(4).uniform(q[0], q[-1], 50), q, [q[0] - 0.1, q[-1] + 0.1]])
        for x in xs:
            a, da, d2a = (np.zeros(4) for _ in range(3))
            i = ea.lib.get_NUBasis_d2funcs_d(b, float(x), a.ctypes.data, da.ctypes.data, d2a.ctypes.data)
            # the interval brackets x (first and last clamp), as the oracle's does
            assert i == og.reverse(x)
            oa, oda, od2a = ob.funcs(i, float(x))
            for got, ref in ((a, oa), (da, oda), (d2a, od2a)):
                np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14 * max(1.0, np.abs(ref).max()))
        ea.lib.destroy_NUBasis(b)
    lg.destroy()


@pytest.mark.parametrize("bc", [(ea.PERIODIC, ea.PERIODIC), (ea.NATURAL, ea.FLAT),
                                (ea.DERIV1, ea.DERIV2, 0.3 - 0.2j, -1.1 + 0.5j)])
@pytest.mark.parametrize("kind", ["d", "z"])
def test_library_set_matches_the_oracle(kind, bc):
    pts = random_grid(16, np.random.default_rng(8))
    lg = ea.NUGrid.general(pts)
    bcv = bc if kind == "z" or len(bc) == 2 else (bc[0], bc[1], bc[2].real, bc[3].real)
    s = ea.NUBSpline1D(kind, lg, bcv, 5)
    o = nub.Spline1D(kind, nub.general_grid(pts), bcv, 5)
    M = len(pts) - 1 if bc[0] == ea.PERIODIC else len(pts)
    r = np.random.default_rng(1)
    for n in range(5):
        d = r.uniform(-1, 1, M) + (1j * r.uniform(-1, 1, M) if kind == "z" else 0)
        s.set(n, d)
        o.set(n, d)
    c = s.coefs()
    assert np.abs(c - o.coefs).max() <= 1e-13 * np.abs(o.coefs).max()
    s.destroy()
    lg.destroy()


@pytest.mark.parametrize("bc", [(ea.NATURAL, ea.NATURAL), (ea.DERIV2, ea.DERIV2, 0.7, -0.3),
                                (ea.FLAT, ea.FLAT), (ea.DERIV1, ea.DERIV2, 2.0, 1.0)])
def test_uniform_grid_set_is_the_uniform_solve(bc):
    """On a uniform grid the non-uniform spline is the uniform one: the
    library's 1-D uniform set (create_multi_UBspline_1d_d, its own solver)
    gives the same coefficients for every end condition, including the
    second-derivative ones whose first collocation row would otherwise meet
    a zero pivot."""
    n = 11
    pts = np.linspace(0.0, 2.0, n)
    lg = ea.NUGrid.general(pts)
    s = ea.NUBSpline1D("d", lg, bc, 1)
    u = ea.HostSpline("d", (n,), (0.0,), (2.0,), 1, bc=[bc])
    d = np.random.default_rng(3).uniform(-1, 1, n)
    s.set(0, d)
    u.set(0, d)
    np.testing.assert_allclose(s.coefs()[:, 0], u.coefs()[:, 0], rtol=0, atol=1e-13)
    s.destroy()
    u.destroy()
    lg.destroy()


def test_library_rejects_bad_grids():
    with pytest.raises(ea.EinsplineError):
        ea.NUGrid.center(0.0, 1.0, 0.5, 10)           # ratio must exceed 1
    with pytest.raises(ea.EinsplineError):
        ea.NUGrid.log(-1.0, 1.0, 10)
    g = ea.NUGrid.general([0.0, 1.0, 2.0])
    with pytest.raises(ea.EinsplineError):
        ea.NUBSpline1D("z", g, (ea.NATURAL, ea.NATURAL), 2)   # fewer than 4 points
    g.destroy()
    assert math.isfinite(1.0)
