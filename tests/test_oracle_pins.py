"""Pins for the oracle (oracle/oracle.c) against closed forms.

The reference ships no known-answer vectors and cannot be compiled here
(its headers need an autoconf-generated config.h), so every function of the
oracle is checked against mathematics it does not share code with:
textbook B3 kernel sums, dense interpolation maps solved by numpy, an FFT
solve, analytic plane waves, and the golden fixtures in tests/golden/
(written by tests/golden/make_golden.py from those closed forms).
"""
import os

import numpy as np
import pytest

import oracle
from oracle import brute

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


# ---- basis (src/bspline_data.c:143-200) ----

def test_basis_closed_form():
    for t in np.linspace(0.0, 1.0, 37):
        a, da, d2a = oracle.basis(t)
        closed = [(1 - t) ** 3 / 6, (3 * t**3 - 6 * t**2 + 4) / 6,
                  (-3 * t**3 + 3 * t**2 + 3 * t + 1) / 6, t**3 / 6]
        np.testing.assert_allclose(a, closed, atol=1e-15)
        # weights are B3 at the four stencil offsets
        np.testing.assert_allclose(a, brute.b3(t - np.arange(4) + 1), atol=1e-15)
        np.testing.assert_allclose(da, brute.db3(t - np.arange(4) + 1), atol=1e-14)
        np.testing.assert_allclose(d2a, brute.d2b3(t - np.arange(4) + 1), atol=1e-14)
        assert abs(a.sum() - 1) < 1e-15 and abs(da.sum()) < 1e-15 and abs(d2a.sum()) < 1e-14


def test_basis_derivatives_match_finite_differences():
    h = 1e-6
    for t in (0.1, 0.37, 0.5, 0.93):
        a_p, da_p, _ = oracle.basis(t + h)
        a_m, da_m, _ = oracle.basis(t - h)
        _, da, d2a = oracle.basis(t)
        np.testing.assert_allclose((a_p - a_m) / (2 * h), da, atol=1e-8)
        np.testing.assert_allclose((da_p - da_m) / (2 * h), d2a, atol=1e-8)


# ---- periodic solve (src/bspline_create.c:1009-1055) ----

def test_set_d_periodic_matches_fft(rng):
    M = (7, 5, 6)
    t = oracle.Table("d", M, (0, 0, 0), (1, 1, 1), 3)
    data = rng.uniform(-1, 1, (3,) + M)
    for n in range(3):
        t.set(n, data[n])
    for n in range(3):
        ref = brute.solve_fft_periodic(data[n])
        assert rel(t.coefs[..., n], ref) < 1e-13


BCS = [
    ("periodic", [(0, 0)] * 3, None),
    ("antiperiodic", [(5, 5)] * 3, None),
    ("natural", [(4, 4)] * 3, None),
    ("flat", [(3, 3)] * 3, None),
    ("deriv1", [(1, 1, 0.7, -0.3), (1, 1, 0.2, 0.5), (1, 1, -1.0, 0.25)], None),
    ("deriv2", [(2, 2, 0.3, -0.6), (2, 2, 1.5, 0.1), (2, 2, 0.0, 2.0)], None),
    ("mixed", [(0, 0), (3, 4), (1, 2, 0.4, -0.2)], None),
]


@pytest.mark.parametrize("name,bcs,_", BCS, ids=[b[0] for b in BCS])
def test_set_d_matches_dense_maps(name, bcs, _, rng):
    M = (6, 7, 5)
    start, end = (0.5, -1.0, 2.0), (2.0, 1.5, 3.0)
    t = oracle.Table("d", M, start, end, 2, bc=bcs)
    data = rng.uniform(-1, 1, (2,) + M)
    for n in range(2):
        t.set(n, data[n])
    lr = [(b[2], b[3]) if len(b) > 2 else (0.0, 0.0) for b in bcs]
    for n in range(2):
        ref = brute.solve_dense(data[n], [b[:2] for b in bcs], t.delta_inv, lr)
        assert rel(t.coefs[..., n], ref) < 1e-12, name


def test_set_s_is_float_solve_of_d(rng):
    """The default build solves float tables in float (src/bspline_create.c:237-318);
    the result agrees with the double solve to float precision."""
    M = (9, 8, 7)
    ts = oracle.Table("s", M, (0, 0, 0), (1, 1, 1), 2)
    td = oracle.Table("d", M, (0, 0, 0), (1, 1, 1), 2)
    data = rng.uniform(-1, 1, (2,) + M).astype(np.float32)
    for n in range(2):
        ts.set(n, data[n])
        td.set(n, data[n].astype(np.float64))
    e = rel(ts.coefs, td.coefs)
    assert 0 < e < 2e-6


def test_set_complex_is_two_real_solves(rng):
    M = (5, 6, 7)
    bcs = [(0, 0), (1, 1, 0.3 + 0.2j, -0.1 + 0.5j), (4, 4)]
    tz = oracle.Table("z", M, (0, 0, 0), (1, 2, 3), 2, bc=bcs)
    data = rng.uniform(-1, 1, (2,) + M) + 1j * rng.uniform(-1, 1, (2,) + M)
    for n in range(2):
        tz.set(n, data[n])
    cz = tz.coefs.view(np.complex128)
    for part, f in (("re", np.real), ("im", np.imag)):
        bd = [(0, 0), (1, 1, f(0.3 + 0.2j), f(-0.1 + 0.5j)), (4, 4)]
        td = oracle.Table("d", M, (0, 0, 0), (1, 2, 3), 2, bc=bd)
        for n in range(2):
            td.set(n, f(data[n]))
        assert rel(f(cz), td.coefs) < 1e-14, part


# ---- evaluation (src/multi_bspline_eval_std_d_impl.h:482-864) ----

def random_points(rng, start, end, P):
    s, e = np.array(start), np.array(end)
    return s + (e - s) * rng.random((P, 3)) * 0.999999


@pytest.mark.parametrize("bcs", [[(0, 0)] * 3, [(4, 4), (3, 3), (1, 1, 0.5, -0.5)]])
def test_eval_d_matches_brute_force(bcs, rng):
    M = (6, 5, 7)
    start, end = (3.1, 8.7, 4.5), (9.1, 12.7, 9.3)       # test_multi.c:709-715 style boxes
    N = 5
    t = oracle.Table("d", M, start, end, N, bc=bcs)
    t.coefs[...] = rng.uniform(-1, 1, t.coefs.shape)    # arbitrary coefficients
    pts = random_points(rng, start, end, 40)
    v, g, H, _ = brute.eval_dense(t.coefs, start, t.delta_inv, pts)
    ov, og, oh = t.eval_batch("vgh", pts)
    assert rel(ov, v) < 1e-13 and rel(og, g) < 1e-13 and rel(oh, H) < 1e-13
    ov2, og2, ol = t.eval_batch("vgl", pts)
    assert rel(ol, np.trace(H, axis1=2, axis2=3)) < 1e-13
    np.testing.assert_array_equal(ov2, ov)
    ov3, og3 = t.eval_batch("vg", pts)
    np.testing.assert_array_equal(og3, og)
    (ov4,) = t.eval_batch("v", pts)
    np.testing.assert_array_equal(ov4, ov)


def test_eval_complex_is_two_real_evals(rng):
    M = (5, 6, 4)
    tz = oracle.Table("z", M, (0, 0, 0), (1, 1, 1), 3)
    td_re = oracle.Table("d", M, (0, 0, 0), (1, 1, 1), 3)
    td_im = oracle.Table("d", M, (0, 0, 0), (1, 1, 1), 3)
    tz.coefs[...] = rng.uniform(-1, 1, tz.coefs.shape)
    c = tz.coefs.view(np.complex128)
    td_re.coefs[...] = c.real
    td_im.coefs[...] = c.imag
    pts = random_points(rng, (0, 0, 0), (1, 1, 1), 25)
    for mode in ("v", "vg", "vgl", "vgh"):
        rz = tz.eval_batch(mode, pts)
        rr = td_re.eval_batch(mode, pts)
        ri = td_im.eval_batch(mode, pts)
        for a, b, cc in zip(rz, rr, ri):
            np.testing.assert_array_equal(a.real, b)
            np.testing.assert_array_equal(a.imag, cc)


def test_eval_s_close_to_d(rng):
    M = (8, 8, 8)
    ts = oracle.Table("s", M, (0, 0, 0), (1, 1, 1), 7)
    td = oracle.Table("d", M, (0, 0, 0), (1, 1, 1), 7)
    ts.coefs[...] = rng.uniform(-1, 1, ts.coefs.shape).astype(np.float32)
    td.coefs[...] = ts.coefs
    pts = random_points(rng, (0, 0, 0), (1, 1, 1), 50)
    for a, b in zip(ts.eval_batch("vgh", pts), td.eval_batch("vgh", pts)):
        assert rel(a, b) < 2e-6


def test_interpolation_property(rng):
    """At every grid node the spline reproduces the data."""
    M = (6, 7, 5)
    start, end = (0.0, -1.0, 2.0), (3.0, 1.0, 4.0)
    t = oracle.Table("d", M, start, end, 2)
    data = rng.uniform(-1, 1, (2,) + M)
    for n in range(2):
        t.set(n, data[n])
    dl = (np.array(end) - np.array(start)) / np.array(M)
    idx = np.array([(i, j, k) for i in range(M[0]) for j in range(M[1]) for k in range(M[2])])
    pts = np.array(start) + idx * dl + 1e-13          # just inside the cell
    (v,) = t.eval_batch("v", pts)
    ref = data[:, idx[:, 0], idx[:, 1], idx[:, 2]].T
    assert rel(v, ref) < 1e-11


def test_constant_data_gives_zero_derivatives():
    M = (5, 5, 5)
    t = oracle.Table("d", M, (0, 0, 0), (1, 1, 1), 2)
    for n in range(2):
        t.set(n, np.full(M, 2.5 + n))
    pts = np.random.default_rng(3).random((20, 3)) * 0.99
    v, g, H = t.eval_batch("vgh", pts)
    np.testing.assert_allclose(v, np.array([2.5, 3.5])[None, :] + 0 * v, atol=1e-13)
    assert np.max(np.abs(g)) < 1e-11 and np.max(np.abs(H)) < 1e-9


def test_plane_wave_converges_fourth_order(rng):
    K, a, b = brute.plane_waves(rng, 4, 2)
    coef = a - 1j * b
    pts = rng.random((200, 3)) * 0.999
    errs = []
    for M in (16, 32):
        t = oracle.Table("d", (M, M, M), (0, 0, 0), (1, 1, 1), 1)
        t.set(0, brute.pw_grid(K, coef, M).real)
        (v,) = t.eval_batch("v", pts)
        ref = brute.pw_eval(K, coef, pts)[0].real
        errs.append(np.max(np.abs(v[:, 0] - ref)))
    assert errs[1] < errs[0] / 10          # O(h^4): ideally 16x


def test_outside_table_is_flagged():
    t = oracle.Table("d", (5, 5, 5), (0, 0, 0), (1, 1, 1), 1)
    (v,) = t.eval_batch("v", np.array([[1.5, 0.5, 0.5]]))
    assert t.bad == 1 and np.isnan(v).all()


def test_golden_fixtures():
    """Known-answer vectors written from the closed forms (make_golden.py)."""
    files = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    assert files
    for f in files:
        z = np.load(os.path.join(GOLDEN, f))
        kind = str(z["kind"])
        bcs = [tuple(b) for b in z["bcs"].tolist()]
        t = oracle.Table(kind, tuple(z["M"]), tuple(z["start"]), tuple(z["end"]),
                         int(z["N"]), bc=bcs)
        if "data" in z:
            for n in range(int(z["N"])):
                t.set(n, z["data"][n])
            tol = 1e-12 if kind in "dz" else 3e-6
            assert rel(t.coefs, z["coefs"]) < tol, f
        else:
            t.coefs[...] = z["coefs"]
        tol = 1e-12 if kind in "dz" else 5e-6
        v, g, H = t.eval_batch("vgh", z["pts"])
        for got, want, nm in ((v, z["v"], "v"), (g, z["g"], "g"), (H, z["H"], "H")):
            assert rel(got, want) < tol, (f, nm)
