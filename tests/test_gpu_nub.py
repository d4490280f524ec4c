"""1-D complex-double non-uniform multi-splines on the GPU
(eval_multi_NUBspline_1d_z*, einspline_eval_nub_1d_z_batch) against the
oracle (oracle/nub.py, pinned in test_nub_oracle.py) on the library's own
coefficients.  The device evaluates the basis with the host library's
triangular scheme (not the oracle's recursion), so results agree to
rounding: 1e-12 normwise per point.  At a grid point the two may pick
neighbouring intervals only where the spline is C2, which the same bar
covers."""
import numpy as np
import pytest

import einspline_amd as ea
from oracle import nub

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ea.device_count() >= 1, "no HIP device visible: GPU tests must run on a GPU"


def grids(kind, rng):
    if kind == "general":
        steps = rng.uniform(0.3, 1.7, 23)
        pts = np.concatenate([[0.0], np.cumsum(steps)]) / steps.sum() * 3.0 - 1.0
        return ea.NUGrid.general(pts)
    if kind == "center":
        return ea.NUGrid.center(-1.0, 2.0, 8.0, 25)
    return ea.NUGrid.log(0.02, 3.0, 22)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


BCS = [(ea.PERIODIC, ea.PERIODIC), (ea.NATURAL, ea.NATURAL), (ea.DERIV1, ea.FLAT, 0.5 - 0.25j, 0.0),
       (ea.DERIV2, ea.DERIV1, 1.0 + 1.0j, -2.0)]


@pytest.mark.parametrize("gkind", ["general", "center", "log"])
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("N", [1, 37, 128])
def test_batch_matches_oracle(gkind, bc, N, rng):
    g = grids(gkind, rng)
    s = ea.NUBSpline1D("z", g, bc, N)
    M = g.ptr.contents.num_points - (1 if bc[0] == ea.PERIODIC else 0)
    for n in range(N):
        s.set(n, rng.uniform(-1, 1, M) + 1j * rng.uniform(-1, 1, M))
    o = nub.Spline1D("z", nub.general_grid(g.points), bc, N)
    o.coefs[...] = s.coefs()
    q = g.points
    xs = np.concatenate([rng.uniform(q[0], q[-1], 300), q, [q[-1], q[0]]])
    tab = ea.lib.create_multi_NUBspline_1d_z_hip(s.ptr)
    assert tab
    P = len(xs)
    pos = ea.DeviceBuffer.from_array(xs)
    bufs = [ea.DeviceBuffer(P * N * 16) for _ in range(3)]
    for mode, name in ((0, "v"), (1, "vg"), (2, "vgl")):
        ea.check(ea.lib.einspline_eval_nub_1d_z_batch(tab, pos.ptr, P, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                                      mode, None), "nub batch")
        ea.synchronize()
        got = [b.download(np.complex128, (P, N)) for b in bufs[:mode + 1]]
        refs = [o.eval(name, x) for x in xs]
        for q in range(mode + 1):
            ref = np.array([r[q] for r in refs])
            # per point, against that point's largest output (or, where the
            # quantity vanishes at a point, e.g. a natural end's second
            # derivative, the largest over all points)
            scale = np.maximum(np.abs(ref).max(1), 1e-3 * np.abs(ref).max())
            err = np.abs(got[q] - ref).max(1) / scale
            assert err.max() < TOL, (gkind, name, float(err.max()), float(xs[err.argmax()]))
    ea.lib.destroy_multi_NUBspline_1d_hip(tab)
    s.destroy()
    g.destroy()


def test_single_point_api_and_refresh(rng):
    g = grids("general", rng)
    N = 20
    s = ea.NUBSpline1D("z", g, (ea.NATURAL, ea.FLAT), N)
    M = g.ptr.contents.num_points
    data = rng.uniform(-1, 1, (N, M)) + 1j * rng.uniform(-1, 1, (N, M))
    for n in range(N):
        s.set(n, data[n])
    o = nub.Spline1D("z", nub.general_grid(g.points), (ea.NATURAL, ea.FLAT), N)
    o.coefs[...] = s.coefs()
    for x in rng.uniform(g.points[0], g.points[-1], 10):
        for mode in ("v", "vg", "vgl", "vgh"):
            got = s.eval(mode, x)
            ref = o.eval("vgl" if mode == "vgh" else mode, x)
            for a, r in zip(got, ref):
                assert rel(a, r) < TOL
    # interpolation at the nodes through the GPU, and the mirror follows set_*
    for k in (0, M // 2, M - 1):
        np.testing.assert_allclose(s.eval("v", g.points[k])[0], data[:, k], atol=1e-12)
    s.set(3, np.zeros(M))
    assert np.abs(s.eval("v", g.points[5])[0][3]) < 1e-14
    s.destroy()
    g.destroy()


def test_batch_argument_errors(rng):
    g = grids("log", rng)
    s = ea.NUBSpline1D("z", g, (ea.NATURAL, ea.NATURAL), 4)
    tab = ea.lib.create_multi_NUBspline_1d_z_hip(s.ptr)
    assert ea.lib.einspline_eval_nub_1d_z_batch(None, None, 1, None, None, None, 0, None) == ea.E_INVAL
    assert ea.lib.einspline_eval_nub_1d_z_batch(tab, None, 1, None, None, None, 0, None) == ea.E_INVAL
    assert ea.lib.einspline_eval_nub_1d_z_batch(tab, None, 0, None, None, None, 0, None) == ea.E_OK
    assert ea.lib.einspline_eval_nub_1d_z_batch(tab, None, 1, None, None, None, 3, None) == ea.E_INVAL
    ea.lib.destroy_multi_NUBspline_1d_hip(tab)
    d = ea.NUBSpline1D("d", g, (ea.NATURAL, ea.NATURAL), 4)
    assert not ea.lib.create_multi_NUBspline_1d_z_hip(C_cast(d))
    s.destroy()
    d.destroy()
    g.destroy()


def C_cast(spline):
    import ctypes as C
    return C.cast(spline.ptr, C.POINTER(ea.multi_NUBspline_1d_z))
