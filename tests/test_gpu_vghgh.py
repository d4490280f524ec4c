"""vghgh on the GPU (value, gradient, Hessian and the 27 third derivatives per
orbital, src/multi_bspline_eval_std_d_impl.h:868-1049) against the oracle,
batched (einspline_eval_vghgh_batch) and through the single-point API
(eval_multi_UBspline_3d_X_vghgh).  Tolerances as for the other modes:
1e-5 normwise for float tables, 1e-12 for double tables."""
import numpy as np
import pytest

import einspline_amd as ea
from test_gpu_parity import TOL, make_pair, normwise, points

pytestmark = pytest.mark.gpu


def gpu_vghgh(tab, pts):
    P, N = len(pts), tab.N
    pos = ea.DeviceBuffer.from_array(np.ascontiguousarray(pts, np.float64))
    outs = tab.alloc_outputs("vghgh", P)
    tab.eval_batch("vghgh", pos, P, outs)
    ea.synchronize()
    v, g, h, t = tab.download("vghgh", P, outs)
    return v.reshape(P, N), g.reshape(P, N, 3), h.reshape(P, N, 3, 3), t.reshape(P, N, 3, 3, 3)


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
@pytest.mark.parametrize("N", [21, 64, 256])
def test_vghgh_batch_matches_oracle(kind, N, rng):
    M = (9, 8, 10)
    start, end = (3.1, 8.7, 4.5), (9.1, 12.7, 9.3)
    bcs = [(0, 0), (4, 4), (3, 3)]
    h, o = make_pair(kind, M, N, rng, start, end, bcs=bcs)
    tab = ea.DeviceTable(h)
    pts = points(rng, 150, start, end)
    got = gpu_vghgh(tab, pts)
    ref = o.eval_batch("vghgh", pts)
    for a, r in zip(got, ref):
        assert normwise(a, r) < TOL[kind]
    # symmetric tensor
    t = got[3]
    for perm in ((0, 1, 3, 2, 4), (0, 1, 2, 4, 3), (0, 1, 4, 3, 2)):
        np.testing.assert_array_equal(t, t.transpose(perm))


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_vghgh_lower_blocks_equal_vgh(kind, rng):
    """Value, gradient and Hessian of vghgh are the vgh results bit for bit
    (the same accumulations in the same order)."""
    from test_gpu_parity import gpu_eval
    h, _ = make_pair(kind, (8, 8, 8), 64, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 100)
    v, g, hh, _ = gpu_vghgh(tab, pts)
    rv, rg, rh = gpu_eval(tab, "vgh", pts)
    np.testing.assert_array_equal(v, rv)
    np.testing.assert_array_equal(g, rg)
    np.testing.assert_array_equal(hh, rh)


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_vghgh_single_point_api(kind, rng):
    h, o = make_pair(kind, (7, 6, 8), 24, rng)
    for x in points(rng, 3):
        got = h.eval("vghgh", *x)
        ref = o.eval_batch("vghgh", x[None])
        for a, r in zip(got, ref):
            assert normwise(a[None], r) < TOL[kind]


def test_vghgh_tiles_and_order_do_not_change_results(rng):
    h, _ = make_pair("s", (12, 12, 12), 256, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 2000)
    base = gpu_vghgh(tab, pts)
    for lb, blk in ((8, 64), (32, 256), (16, 128)):      # 32-byte requests run as 16
        tab.set_tile(lb, blk)
        for order in (ea.ORDER_GIVEN, ea.ORDER_LOCALITY):
            tab.set_point_order(order)
            for a, b in zip(gpu_vghgh(tab, pts), base):
                np.testing.assert_array_equal(a, b)


def test_vghgh_bf16_table_close_to_f64(rng):
    """A BF16-stored table evaluates vghgh with FP32 accumulation."""
    h, o = make_pair("d", (10, 10, 10), 64, rng)
    tab = ea.DeviceTable(h, storage=ea.STORE_BF16)
    pts = points(rng, 50)
    got = gpu_vghgh(tab, pts)
    ref = o.eval_batch("vghgh", pts)
    # 8-bit mantissa coefficients: relative error well under 1e-1 on every block
    for a, r in zip(got, ref):
        assert normwise(a, r) < 5e-2
