"""The task's workload configurations at full size, on the GPU.

Every configuration is evaluated in one batched launch at its real shape;
parity with the oracle is checked on a sample of the points (all of them
for the small FP64 case), plus size-independent properties of the whole
batch: finiteness, Hessian symmetry, and the Laplacian equal to the trace of
the Hessian from a separate vgh launch.
"""
import numpy as np
import pytest

import einspline_amd as ea
import oracle

pytestmark = pytest.mark.gpu


def normwise(got, ref):
    P = ref.shape[0]
    d = np.abs(got - ref).reshape(P, -1).max(1)
    s = np.abs(ref).reshape(P, -1).max(1)
    return float(np.max(d / np.maximum(s, 1e-300)))


def plane_wave_data(rng, M, count, complex_=False, nwaves=8, kmax=4):
    x = np.arange(M) / M
    out = np.empty((count, M, M, M), np.complex128 if complex_ else np.float64)
    for o in range(count):
        K = rng.integers(-kmax, kmax + 1, size=(nwaves, 3))
        c = rng.standard_normal(nwaves) + 1j * rng.standard_normal(nwaves)
        acc = np.zeros((M, M, M), np.complex128)
        for m in range(nwaves):
            e = [np.exp(2j * np.pi * K[m, a] * x) for a in range(3)]
            acc += c[m] * (e[0][:, None, None] * e[1][None, :, None] * e[2][None, None, :])
        out[o] = acc if complex_ else acc.real
    return out


def run(tab, mode, pts):
    P = len(pts)
    pos = ea.DeviceBuffer.from_array(pts)
    outs = tab.alloc_outputs(mode, P)
    tab.eval_batch(mode, pos, P, outs)
    ea.synchronize()
    return tab.download(mode, P, outs)


def test_c1_double_32_n64_vgh_all_points():
    """multi_UBspline_3d_d, periodic 32^3, 64 orbitals of U[-1,1] data, vgh at
    4096 uniform points, FP64: every point against the oracle at 1e-12."""
    rng = np.random.default_rng(1)
    M, N, P = 32, 64, 4096
    h = ea.HostSpline("d", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    h.set_block(0, rng.uniform(-1, 1, (N, M, M, M)))
    tab = ea.DeviceTable(h)
    pts = rng.random((P, 3))
    v, g, H = run(tab, "vgh", pts)
    o = oracle.Table("d", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    o.coefs[...] = h.coefs()
    rv, rg, rH = o.eval_batch("vgh", pts)
    assert normwise(v.reshape(rv.shape), rv) < 1e-12
    assert normwise(g.reshape(rg.shape), rg) < 1e-12
    assert normwise(H.reshape(rH.shape), rH) < 1e-12


def test_c2_float_48_n256_vgh_32768_points():
    rng = np.random.default_rng(2)
    M, N, P = 48, 256, 32768
    h = ea.HostSpline("s", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    for first in range(0, N, 64):
        h.set_block(first, plane_wave_data(rng, M, 64))
    tab = ea.DeviceTable(h)
    pts = rng.random((P, 3))
    v, g, H = run(tab, "vgh", pts)
    assert np.isfinite(v).all() and np.isfinite(g).all() and np.isfinite(H).all()
    H = H.reshape(P, N, 3, 3)
    np.testing.assert_array_equal(H, np.swapaxes(H, -1, -2))
    # vgl Laplacian of the same points equals the Hessian trace
    _, _, lap = run(tab, "vgl", pts)
    tr = np.trace(H, axis1=2, axis2=3)
    assert np.max(np.abs(lap.reshape(P, N) - tr)) / np.max(np.abs(tr)) < 1e-5
    o = oracle.Table("s", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    o.coefs[...] = h.coefs()
    idx = rng.choice(P, 512, replace=False)
    rv, rg, rH = o.eval_batch("vgh", pts[idx])
    assert normwise(v.reshape(P, N)[idx], rv) < 1e-5
    assert normwise(g.reshape(P, N, 3)[idx], rg) < 1e-5
    assert normwise(H[idx], rH) < 1e-5


@pytest.mark.parametrize("kind,tol", [("z", 1e-12), ("c", 1e-5)])
def test_c3_complex_64_n128_vgl_65536_points(kind, tol):
    rng = np.random.default_rng(3)
    M, N, P = 64, 128, 65536
    h = ea.HostSpline(kind, (M, M, M), (0, 0, 0), (1, 1, 1), N)
    for first in range(0, N, 32):
        h.set_block(first, plane_wave_data(rng, M, 32, complex_=True))
    tab = ea.DeviceTable(h)
    pts = rng.random((P, 3))
    v, g, lap = run(tab, "vgl", pts)
    assert np.isfinite(v.view(np.float32 if kind == "c" else np.float64)).all()
    o = oracle.Table(kind, (M, M, M), (0, 0, 0), (1, 1, 1), N)
    o.coefs[...] = h.coefs().reshape(o.coefs.shape)
    idx = rng.choice(P, 256, replace=False)
    rv, rg, rl = o.eval_batch("vgl", pts[idx])
    assert normwise(v.reshape(P, N)[idx], rv) < tol
    assert normwise(g.reshape(P, N, 3)[idx], rg) < tol
    assert normwise(lap.reshape(P, N)[idx], rl) < tol


@pytest.mark.parametrize("storage,vtol,htol", [(ea.STORE_F32, 1e-5, 1e-4),
                                                (ea.STORE_BF16, 1e-2, 5e-2),
                                                (ea.STORE_F16, 2e-3, 1e-2)])
def test_c4_precision_form_error_bounds(storage, vtol, htol):
    """Precision form at reduced size (32^3, 64 smooth plane-wave orbitals):
    built in FP64, stored narrow, errors against the FP64 evaluation."""
    rng = np.random.default_rng(4)
    M, N, P = 32, 64, 2048
    h = ea.HostSpline("d", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    h.set_block(0, plane_wave_data(rng, M, N, kmax=2))
    tab = ea.DeviceTable(h, storage=storage)
    pts = rng.random((P, 3))
    v, g, H = run(tab, "vgh", pts)
    o = oracle.Table("d", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    o.coefs[...] = h.coefs()
    rv, rg, rH = o.eval_batch("vgh", pts)

    def err(a, b):
        return np.max(np.abs(a.reshape(b.shape) - b)) / np.max(np.abs(b))

    assert err(v, rv) < vtol
    assert err(g, rg) < htol
    assert err(H, rH) < htol
