"""einspline_eval_vgh_batch_noalloc against einspline_eval_vgh_batch.

The allocation-free entry point sorts the locality order in a file-scope
device array instead of the table's workspace; its results must equal the
allocating entry point's bit for bit, on two calls in a row with the second
one smaller (the array keeps the first call's contents).
"""
import numpy as np
import pytest

import einspline_amd as ea

pytestmark = pytest.mark.gpu


def _table(kind_storage, rng):
    M, N = (20, 18, 22), 96
    if kind_storage == "f32":
        h = ea.HostSpline("s", M, (0, 0, 0), (1, 1, 1), N)
        h.set_block(0, rng.uniform(-1, 1, (N,) + M))
        tab = ea.DeviceTable(h)
    else:
        h = ea.HostSpline("d", M, (0, 0, 0), (1, 1, 1), N)
        h.set_block(0, rng.uniform(-1, 1, (N,) + M))
        tab = ea.DeviceTable(h, storage=ea.STORE_BF16)
        tab.set_kernel(ea.KERNEL_MATRIX)
    tab.set_point_order(ea.ORDER_LOCALITY)
    return h, tab


def _run(tab, fn, pos, P, packed):
    outs = tab.alloc_outputs("vgh", P, packed)[:3]
    for b in outs:
        b.zero()
    ea.check(fn(tab.ptr, pos.ptr, P, outs[0].ptr, outs[1].ptr, outs[2].ptr, 1 if packed else 0, None), fn.__name__)
    ea.synchronize()
    return [x.view(np.uint32) for x in tab.download("vgh", P, outs, packed)]


@pytest.mark.parametrize("kind_storage", ["f32", "bf16"])
@pytest.mark.parametrize("packed", [False, True])
def test_noalloc_equals_allocating_entry_point(kind_storage, packed):
    rng = np.random.default_rng(2718)
    print(f"[noalloc] {kind_storage} packed={packed}", flush=True)
    h, tab = _table(kind_storage, rng)
    for P in (6000, 1500):                 # the second call is smaller
        pts = rng.random((P, 3))
        pos = ea.DeviceBuffer.from_array(pts)
        got = _run(tab, ea.lib.einspline_eval_vgh_batch_noalloc, pos, P, packed)
        ref = _run(tab, ea.lib.einspline_eval_vgh_batch, pos, P, packed)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        # and against the given order (no sort at all)
        tab.set_point_order(ea.ORDER_GIVEN)
        given = _run(tab, ea.lib.einspline_eval_vgh_batch, pos, P, packed)
        tab.set_point_order(ea.ORDER_LOCALITY)
        for a, b in zip(got, given):
            np.testing.assert_array_equal(a, b)


def test_noalloc_refuses_batches_above_the_limit():
    rng = np.random.default_rng(1)
    h, tab = _table("f32", rng)
    rc = ea.lib.einspline_eval_vgh_batch_noalloc(tab.ptr, None, ea.NOALLOC_MAX_POINTS + 1, None, None, None,
                                                 0, None)
    assert rc == ea.E_INVAL
    assert b"at most" in ea.lib.einspline_hip_last_error()
