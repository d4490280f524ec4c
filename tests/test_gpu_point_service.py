"""Single-point calls served by the resident polling kernel.

The kernel evaluates with the batch kernel's per-chunk code, so its results
must be the batch path's bits for every kind and mode; the launch-per-call
path (service off) must give the same bits too.  The kernel leaves after its
idle time and the next call starts it again; a device synchronisation after
that returns.
"""
import time

import numpy as np
import pytest

import einspline_amd as ea
import oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[point service] {msg}", flush=True)


@pytest.fixture(autouse=True)
def service_defaults():
    ea.set_point_service(True, 2000)
    yield
    ea.set_point_service(True, 2000)


def make(kind, M, N, rng):
    h = ea.HostSpline(kind, M, (0, 0, 0), (1, 1, 1), N)
    data = rng.uniform(-1, 1, (N,) + tuple(M))
    if kind in "cz":
        data = data + 1j * rng.uniform(-1, 1, (N,) + tuple(M))
    h.set_block(0, data)
    return h


def batch(h, mode, pts):
    """The batch path on the single-point API's own device mirror layout:
    a device table built from the same host table."""
    tab = ea.DeviceTable(h)
    P = len(pts)
    pos = ea.DeviceBuffer.from_array(np.ascontiguousarray(pts, np.float64))
    outs = tab.alloc_outputs(mode, P)
    tab.eval_batch(mode, pos, P, outs)
    ea.synchronize()
    res = tab.download(mode, P, outs)
    tab.destroy()
    return res


def as_reals(a):
    a = np.asarray(a)
    return a.view(a.real.dtype) if np.iscomplexobj(a) else a


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_service_bits_equal_batch(kind, rng):
    M, N, P = (12, 11, 10), 70, 9
    h = make(kind, M, N, rng)
    pts = rng.random((P, 3)) * (1 - 1e-9)
    pts[0] = (0.0, 0.0, 0.0)
    for mode in ("v", "vg", "vgl", "vgh"):
        say(f"{kind} {mode}")
        ref = batch(h, mode, pts)
        for p in range(P):
            got = h.eval(mode, *pts[p])
            for q, (g, r) in enumerate(zip(got, ref)):
                r = np.asarray(r).reshape(P, -1)[p]
                np.testing.assert_array_equal(as_reals(g).ravel(), as_reals(r).ravel(), err_msg=f"{mode} {p} {q}")


def test_service_off_gives_the_same_bits(rng):
    h = make("s", (10, 10, 10), 256, rng)
    pts = rng.random((5, 3))
    on = [h.eval("vgh", *x) for x in pts]
    ea.set_point_service(False)
    assert ea.point_service_resident() == 0
    off = [h.eval("vgh", *x) for x in pts]
    assert ea.point_service_resident() == 0
    for a, b in zip(on, off):
        for u, w in zip(a, b):
            np.testing.assert_array_equal(u, w)


def test_idle_exit_and_restart(rng):
    ea.set_point_service(True, 5000)
    h = make("d", (9, 9, 9), 40, rng)
    o = oracle.Table("d", (9, 9, 9), (0, 0, 0), (1, 1, 1), 40)
    o.coefs[...] = h.coefs().reshape(o.coefs.shape)
    x = (0.31, 0.52, 0.77)
    ref = o.eval_batch("vgh", np.array([x]))
    for rounds in range(3):
        v, g, H = h.eval("vgh", *x)
        assert ea.point_service_resident() == 1
        np.testing.assert_allclose(v, ref[0][0], rtol=0, atol=1e-12 * np.abs(ref[0]).max())
        np.testing.assert_allclose(H.reshape(-1), ref[2][0].reshape(-1), rtol=0, atol=1e-11 * np.abs(ref[2]).max())
        time.sleep(0.05)                       # well past the 5 ms idle time
        assert ea.point_service_resident() == 0, "the kernel stays past its idle time"
    ea.synchronize()                           # nothing left on the device


def test_many_calls_and_device_sync(rng):
    """A long run of calls with batch work on another stream in between."""
    h = make("s", (16, 16, 16), 256, rng)
    tab = ea.DeviceTable(h)
    pts = rng.random((64, 3))
    pos = ea.DeviceBuffer.from_array(pts)
    outs = tab.alloc_outputs("vgh", 64)
    s = ea.Stream()
    first = [h.eval("vgh", *x) for x in pts]
    for rep in range(20):
        tab.eval_batch("vgh", pos, 64, outs, False, s)
        for i in (0, 17, 63):
            got = h.eval("vgh", *pts[i])
            for u, w in zip(got, first[i]):
                np.testing.assert_array_equal(u, w)
    s.synchronize()
    ref = tab.download("vgh", 64, outs)
    for i in (0, 17, 63):
        np.testing.assert_array_equal(first[i][0].ravel(), ref[0].reshape(64, -1)[i])
    t0 = time.perf_counter()
    ea.synchronize()                           # waits for the idle exit (2 ms)
    assert time.perf_counter() - t0 < 5.0
    tab.destroy()
