"""Locality traversal (einspline_set_point_order): batches sorted by cell give
bitwise the same results as the given order, for every entry point that
takes a batch, on several streams, across workspace growth, on grids large
enough to need coarsened sort keys, and with points outside the box."""
import numpy as np
import pytest

import einspline_amd as ea
import oracle
from test_gpu_parity import TOL, gpu_eval, make_pair, normwise, points

pytestmark = pytest.mark.gpu


def both_orders(tab, mode, pts, packed=False):
    tab.set_point_order(ea.ORDER_GIVEN)
    given = gpu_eval(tab, mode, pts, packed)
    tab.set_point_order(ea.ORDER_LOCALITY)
    local = gpu_eval(tab, mode, pts, packed)
    tab.set_point_order(ea.ORDER_GIVEN)
    return given, local


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
@pytest.mark.parametrize("mode", ["v", "vg", "vgl", "vgh"])
def test_locality_order_is_bitwise_identical(kind, mode, rng):
    h, o = make_pair(kind, (12, 10, 9), 64, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 3000)
    given, local = both_orders(tab, mode, pts)
    for a, b in zip(local, given):
        np.testing.assert_array_equal(a, b)
    ref = o.eval_batch(mode, pts[:300])
    for a, r in zip(local, ref):
        assert normwise(a[:300], r) < TOL[kind]


def test_locality_order_with_tiles_policies_and_packed(rng):
    h, _ = make_pair("s", (16, 16, 16), 256, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 4100)
    base = gpu_eval(tab, "vgh", pts)
    base_p = gpu_eval(tab, "vgh", pts, packed=True)
    tab.set_point_order(ea.ORDER_LOCALITY)
    for pol in (0, 1, 2):
        tab.set_store_policy(pol)
        for lb, blk in ((8, 256), (16, 64), (32, 128)):
            tab.set_tile(lb, blk)
            for a, b in zip(gpu_eval(tab, "vgh", pts), base):
                np.testing.assert_array_equal(a, b)
    for a, b in zip(gpu_eval(tab, "vgh", pts, packed=True), base_p):
        np.testing.assert_array_equal(a, b)


def test_points_outside_the_box_and_nan(rng):
    """Keys clamp like the kernel's cell index; a NaN position sorts into
    cell 0 and still gets exactly the given-order result."""
    h, _ = make_pair("d", (8, 8, 8), 16, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 2000)
    pts[::7] = pts[::7] * 5 - 2            # far outside on both sides
    pts[5] = [np.nan, 0.5, 0.5]
    given, local = both_orders(tab, "vgh", pts)
    for a, b in zip(local, given):
        np.testing.assert_array_equal(a, b)


def test_huge_axis_uses_coarsened_keys(rng):
    """An axis of more than 1024 cells needs the key shift; results unchanged."""
    h, o = make_pair("s", (1500, 5, 4), 8, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 5000)
    given, local = both_orders(tab, "vgl", pts)
    for a, b in zip(local, given):
        np.testing.assert_array_equal(a, b)
    ref = o.eval_batch("vgl", pts[:100])
    for a, r in zip(local, ref):
        assert normwise(a[:100], r) < TOL["s"]


def test_workspace_growth_and_two_streams(rng):
    h, _ = make_pair("s", (14, 14, 14), 128, rng)
    tab = ea.DeviceTable(h)
    ref_tab = ea.DeviceTable(h)
    streams = [ea.Stream(), ea.Stream()]
    tab.set_point_order(ea.ORDER_LOCALITY)
    jobs = []
    for i, P in enumerate((1500, 6000, 1100, 70000, 3000, 900)):
        pts = points(rng, P)
        pos = ea.DeviceBuffer.from_array(pts)
        outs = tab.alloc_outputs("vgh", P)
        tab.eval_batch("vgh", pos, P, outs, stream=streams[i % 2])
        jobs.append((pts, pos, outs))
    for s in streams:
        s.synchronize()
    for pts, pos, outs in jobs:
        got = tab.download("vgh", len(pts), outs)
        want = gpu_eval(ref_tab, "vgh", pts)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a.reshape(b.shape), b)


@pytest.mark.parametrize("kind", ["s", "z"])
def test_pointer_array_api_in_locality_order(kind, rng):
    M, N, P = (9, 9, 9), 32, 2500
    h, _ = make_pair(kind, M, N, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, P)
    fpos = kind in "sc"
    pos = ea.DeviceBuffer.from_array(pts.astype(np.float32) if fpos else pts)
    nr = tab.nreal
    es = np.dtype(tab.result_dtype).itemsize
    f = getattr(ea.lib, f"eval_multi_multi_UBspline_3d_{kind}_vgh_hip")
    res = []
    for order in (ea.ORDER_GIVEN, ea.ORDER_LOCALITY):
        tab.set_point_order(order)
        vb, gb, hb = (ea.DeviceBuffer(P * k * nr * es) for k in (1, 3, 6))
        ptrs = lambda b, k: ea.DeviceBuffer.from_array(
            np.array([b.ptr + p * k * nr * es for p in range(P)], np.uint64))
        vp, gp, hp = ptrs(vb, 1), ptrs(gb, 3), ptrs(hb, 6)
        ea.check(f(tab.ptr, pos.ptr, vp.ptr, gp.ptr, hp.ptr, P), "vgh_hip")
        res.append([b.download(tab.result_dtype, (P, k * nr)) for b, k in ((vb, 1), (gb, 3), (hb, 6))])
    for a, b in zip(*res):
        np.testing.assert_array_equal(a, b)


def test_autotune_reports_an_order_and_keeps_results(rng):
    h, _ = make_pair("s", (20, 20, 20), 128, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 3000)
    before = gpu_eval(tab, "vgh", pts)
    tab.autotune("vgh", 4096)
    lb, blk, pol, order = tab.tuning()
    assert order in (0, 1, 2, 3)          # given / locality, optionally with orbital slabs
    # the matrix-core kernel (a candidate for vgh) sums in another order
    for a, b in zip(gpu_eval(tab, "vgh", pts), before):
        if tab.kernel == ea.KERNEL_VECTOR:
            np.testing.assert_array_equal(a, b)
        else:
            assert normwise(a, b) < TOL["s"]
    with pytest.raises(ea.EinsplineError):
        tab.set_point_order(7)


def test_autotune_on_clustered_caller_points(rng):
    """einspline_autotune_points times the candidates on the caller's own
    positions: points piled into a few cells (where the locality order makes
    every CU of an XCD read the same rows at once) still tune to a valid
    configuration, results keep their bar, and bad arguments are refused."""
    h, _ = make_pair("s", (24, 24, 24), 128, rng)
    tab = ea.DeviceTable(h)
    pts = rng.random((8192, 3)) * 0.1          # about 3 cells per axis
    before = gpu_eval(tab, "vgh", pts)
    lb, blk, pol, ms = tab.autotune_points("vgh", pts)
    assert ms > 0.0 and (lb, blk, pol) == tab.tuning()[:3]
    assert tab.tuning()[3] in (0, 1, 2, 3)
    for a, b in zip(gpu_eval(tab, "vgh", pts), before):
        if tab.kernel == ea.KERNEL_VECTOR:
            np.testing.assert_array_equal(a, b)
        else:
            assert normwise(a, b) < TOL["s"]
    with pytest.raises(ea.EinsplineError):
        tab.autotune_points("vgh", np.zeros((0, 3)))


@pytest.mark.parametrize("storage,kernel,tile,N", [(ea.STORE_F32, ea.KERNEL_VECTOR, (8, 256, 2), 384),
                                                   (ea.STORE_F32, ea.KERNEL_VECTOR, (16, 128, 0), 512),
                                                   (ea.STORE_BF16, ea.KERNEL_VECTOR, (8, 64, 1), 512),
                                                   (ea.STORE_BF16, ea.KERNEL_MATRIX, (16, 256, 0), 300),
                                                   (ea.STORE_F16, ea.KERNEL_MATRIX, (16, 256, 1), 300),
                                                   (ea.STORE_F16, ea.KERNEL_MATRIX, (16, 256, 0), 512),
                                                   (ea.STORE_BF16, ea.KERNEL_MATRIX, (16, 256, 0), 1000),
                                                   (ea.STORE_F32, ea.KERNEL_MATRIX, (16, 256, 1), 300)])
def test_orbital_slab_orders_are_bitwise_identical(storage, kernel, tile, N, rng):
    """EINSPLINE_ORDER_SLABS (orbital slabs outermost) with either point order
    gives the bits of the plain given order.  Rows span 2-5 wavefront widths
    (vector kernel: 128-192 lanes per point; 16-bit matrix kernel: 256-orbital
    slabs, 300 and 1000 orbitals ending in a partial slab, 512 in two whole
    ones; FP32 matrix kernel: 64-orbital slabs); 5000 points, so the locality
    sort runs and the per-XCD split of the points is uneven."""
    P = 5000
    h = ea.HostSpline("d", (11, 9, 12), (0, 0, 0), (1, 1, 1), N)
    h.set_block(0, rng.uniform(-1, 1, (N, 11, 9, 12)))
    tab = ea.DeviceTable(h, storage=storage)
    tab.set_kernel(kernel)
    tab.set_tile(tile[0], tile[1])
    tab.set_store_policy(tile[2])
    pts = points(rng, P)
    tab.set_point_order(ea.ORDER_GIVEN)
    base = gpu_eval(tab, "vgh", pts)
    for order in (ea.ORDER_SLABS, ea.ORDER_LOCALITY | ea.ORDER_SLABS, ea.ORDER_LOCALITY):
        tab.set_point_order(order)
        assert tab.tuning()[3] == order
        for a, b in zip(gpu_eval(tab, "vgh", pts), base):
            np.testing.assert_array_equal(a, b)
        # packed Hessians too
        for a, b in zip(gpu_eval(tab, "vgh", pts, packed=True), gpu_eval_given_packed(tab, pts)):
            np.testing.assert_array_equal(a, b)


def gpu_eval_given_packed(tab, pts):
    order = tab.tuning()[3]
    tab.set_point_order(ea.ORDER_GIVEN)
    out = gpu_eval(tab, "vgh", pts, packed=True)
    tab.set_point_order(order)
    return out
