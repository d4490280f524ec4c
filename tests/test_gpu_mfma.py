"""vgh of BF16 / FP16 / FP32 tables on the matrix cores (einspline_set_kernel
with EINSPLINE_KERNEL_MATRIX) against the oracle evaluated in FP64 on the
table's own coefficients, and against the vector kernel on the same table.

Error budget: the 16-bit kernel splits every basis-weight product into a
16-bit high part and a 16-bit remainder (two products per tile), so each
weight carries a relative error of at most 2^-18 (BF16) or 2^-24 (FP16) plus
FP32 accumulation; against the FP64 evaluation of the same 16-bit table the
normwise error is therefore bounded well under 1e-4 (BF16) and 1e-5 (FP16).
The FP32 kernel forms the weights and products in FP32 like the vector
kernel (only the summation order differs): the FP32 bar of the other parity
tests, 1e-5.  The vector kernel computes the weights in FP32, so the two
kernels agree to the same bars."""
import numpy as np
import pytest

import einspline_amd as ea
import oracle

pytestmark = pytest.mark.gpu

TOL = {ea.STORE_BF16: 1e-4, ea.STORE_F16: 1e-5, ea.STORE_F32: 1e-5}
STORAGES = [ea.STORE_BF16, ea.STORE_F16, ea.STORE_F32]


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ea.device_count() >= 1, "no HIP device visible: GPU tests must run on a GPU"


def normwise(got, ref):
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    P = ref.shape[0]
    d = np.abs(got - ref).reshape(P, -1).max(1)
    s = np.abs(ref).reshape(P, -1).max(1)
    return float(np.max(d / np.maximum(s, 1e-300)))


def table_values(tab):
    """The device table's coefficients as FP64, shaped (Nx, Ny, Nz, N)."""
    t = tab.ptr.contents
    rows = t.Nx * t.Ny * t.Nz
    raw = np.empty((rows, t.row_reals), np.float32 if t.storage == ea.STORE_F32 else np.uint16)
    ea.check(ea.lib.einspline_hip_memcpy_dtoh(raw.ctypes.data, t.coefs, raw.nbytes), "dtoh")
    raw = raw[:, :t.num_splines]
    if t.storage == ea.STORE_F32:
        vals = raw.astype(np.float64)
    elif t.storage == ea.STORE_BF16:
        vals = (raw.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    else:
        vals = raw.view(np.float16).astype(np.float64)
    return vals.reshape(t.Nx, t.Ny, t.Nz, t.num_splines)


def build(storage, M, N, rng, bcs=None, start=(0, 0, 0), end=(1, 1, 1)):
    h = ea.HostSpline("d", M, start, end, N, bc=bcs)
    h.set_block(0, rng.uniform(-1, 1, (N,) + tuple(M)))
    tab = ea.DeviceTable(h, storage=storage)
    o = oracle.Table("d", M, start, end, N, bc=bcs)
    o.coefs[...] = table_values(tab).reshape(o.coefs.shape)
    return h, tab, o


def eval_vgh(tab, pts, kernel, packed=False):
    tab.set_kernel(kernel)
    P = len(pts)
    pos = ea.DeviceBuffer.from_array(np.ascontiguousarray(pts, np.float64))
    outs = tab.alloc_outputs("vgh", P, packed)
    tab.eval_batch("vgh", pos, P, outs, packed)
    ea.synchronize()
    v, g, h = tab.download("vgh", P, outs, packed)[:3]
    N = tab.N
    return v.reshape(P, N), g.reshape(P, N, 3), h.reshape(P, N, 6) if packed else h.reshape(P, N, 3, 3)


def full_to_packed(H):
    return H[..., [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]


@pytest.mark.parametrize("storage", STORAGES)
# 128-orbital slabs go in pairs: N covers one partial slab (5, 40), one whole
# (128), a whole pair with or without a partial one (130, 200, 300), two pairs
# (512), the pair loop once (768) and twice (898, 1000, partial last pair);
# N % 4 != 0 (5, 130, 898) takes the plain-store variant.
@pytest.mark.parametrize("N", [5, 40, 128, 130, 200, 300, 512, 768, 898, 1000])
@pytest.mark.parametrize("packed", [False, True])
def test_matrix_kernel_matches_oracle_and_vector_kernel(storage, N, packed, rng):
    M = (9, 10, 11)
    h, tab, o = build(storage, M, N, rng)
    pts = rng.random((300, 3)) * (1 - 1e-9)
    rv, rg, rH = o.eval_batch("vgh", pts)
    if packed:
        rH = full_to_packed(rH)
    mv, mg, mH = eval_vgh(tab, pts, ea.KERNEL_MATRIX, packed)
    vv, vg, vH = eval_vgh(tab, pts, ea.KERNEL_VECTOR, packed)
    tol = TOL[storage]
    for got, ref, name in ((mv, rv, "v"), (mg, rg, "g"), (mH, rH, "h")):
        assert normwise(got, ref) < tol, (name, normwise(got, ref))
    for got, ref in ((mv, vv), (mg, vg), (mH, vH)):
        assert normwise(got, ref) < tol
    if not packed:
        # the full Hessian is written symmetric
        np.testing.assert_array_equal(mH, np.swapaxes(mH, 2, 3))


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("N", [512, 1000])
def test_matrix_kernel_store_policies_agree(storage, N, rng):
    """Streaming stores (aligned outputs) and the cached and direct policies
    write the same bits, whole pairs and a partial last pair alike."""
    h, tab, o = build(storage, (9, 10, 11), N, rng)
    pts = rng.random((257, 3))
    res = []
    for policy in (0, 1, 2):
        tab.set_store_policy(policy)
        res.append(eval_vgh(tab, pts, ea.KERNEL_MATRIX))
    for other in res[1:]:
        for x, y in zip(res[0], other):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("storage", STORAGES)
def test_matrix_kernel_locality_order_is_bitwise_identical(storage, rng):
    h, tab, o = build(storage, (16, 16, 16), 96, rng)
    pts = rng.random((5000, 3))
    tab.set_point_order(ea.ORDER_GIVEN)
    a = eval_vgh(tab, pts, ea.KERNEL_MATRIX)
    tab.set_point_order(ea.ORDER_LOCALITY)
    b = eval_vgh(tab, pts, ea.KERNEL_MATRIX)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    idx = rng.choice(len(pts), 200, replace=False)
    rv, rg, rH = o.eval_batch("vgh", pts[idx])
    assert normwise(b[0][idx], rv) < TOL[storage]
    assert normwise(b[2][idx], rH) < TOL[storage]


@pytest.mark.parametrize("storage", STORAGES)
def test_matrix_kernel_nonperiodic_and_edge_points(storage, rng):
    bcs = [(ea.NATURAL, ea.FLAT), (ea.DERIV1, ea.DERIV2, 0.5, -1.0), (ea.PERIODIC, ea.PERIODIC)]
    start, end = (-1.0, 0.5, 2.0), (1.0, 3.0, 2.5)
    M = (7, 8, 6)
    h, tab, o = build(storage, M, 33, rng, bcs=bcs, start=start, end=end)
    s, e = np.array(start), np.array(end)
    inner = s + (e - s) * rng.random((200, 3)) * (1 - 1e-9)
    corners = np.array([[s[0], s[1], s[2]], [e[0] - 1e-12, e[1] - 1e-12, e[2] - 1e-12],
                        [s[0], e[1] - 1e-12, s[2]]])
    pts = np.concatenate([inner, corners])
    rv, rg, rH = o.eval_batch("vgh", pts)
    mv, mg, mH = eval_vgh(tab, pts, ea.KERNEL_MATRIX)
    tol = TOL[storage]
    assert normwise(mv, rv) < tol and normwise(mg, rg) < tol and normwise(mH, rH) < tol
    # points outside the box take the edge cell, as in the vector kernel (clamped, finite)
    out = np.array([[s[0] - 0.3, e[1] + 0.4, s[2] - 1.0], [e[0] + 5.0, s[1] - 5.0, e[2] + 0.1]])
    a = eval_vgh(tab, out, ea.KERNEL_MATRIX)
    b = eval_vgh(tab, out, ea.KERNEL_VECTOR)
    for x, y in zip(a, b):
        assert np.isfinite(x).all()
        assert normwise(x, y) < 1e-2


def test_kernel_setting_api(rng):
    h, tab, o = build(ea.STORE_BF16, (6, 6, 6), 16, rng)
    assert tab.kernel == ea.KERNEL_VECTOR
    tab.set_kernel(ea.KERNEL_MATRIX)
    assert tab.kernel == ea.KERNEL_MATRIX
    with pytest.raises(ea.EinsplineError):
        tab.set_kernel(7)
    assert tab.kernel == ea.KERNEL_MATRIX
    # other modes ignore the setting: results equal the vector kernel's bit for bit
    pts = rng.random((64, 3))
    pos = ea.DeviceBuffer.from_array(pts)
    res = []
    for k in (ea.KERNEL_MATRIX, ea.KERNEL_VECTOR):
        tab.set_kernel(k)
        outs = tab.alloc_outputs("vgl", 64)
        tab.eval_batch("vgl", pos, 64, outs)
        ea.synchronize()
        res.append(tab.download("vgl", 64, outs))
    for x, y in zip(*res):
        np.testing.assert_array_equal(x, y)
    # FP64 and complex tables accept and ignore the setting
    for kind in ("d", "c"):
        hk = ea.HostSpline(kind, (6, 6, 6), (0, 0, 0), (1, 1, 1), 8)
        data = rng.uniform(-1, 1, (8, 6, 6, 6))
        hk.set_block(0, data + (1j * data[::-1] if kind == "c" else 0))
        tk = ea.DeviceTable(hk)
        P = len(pts)
        posk = ea.DeviceBuffer.from_array(pts)
        res = []
        for k in (ea.KERNEL_VECTOR, ea.KERNEL_MATRIX):
            tk.set_kernel(k)
            outs = tk.alloc_outputs("vgh", P)
            tk.eval_batch("vgh", posk, P, outs)
            ea.synchronize()
            res.append(tk.download("vgh", P, outs))
        for x, y in zip(*res):
            np.testing.assert_array_equal(x, y)
    assert ea.lib.einspline_set_kernel(None, 0) == ea.E_INVAL
    assert ea.lib.einspline_get_kernel(None) == ea.E_INVAL


@pytest.mark.parametrize("storage", [ea.STORE_BF16, ea.STORE_F32])
def test_autotune_may_pick_the_matrix_kernel_and_keeps_results(storage, rng):
    h, tab, o = build(storage, (24, 24, 24), 256, rng)
    tab.autotune("vgh", 8192)
    k = tab.kernel
    assert k in (ea.KERNEL_VECTOR, ea.KERNEL_MATRIX)
    pts = rng.random((256, 3))
    rv, rg, rH = o.eval_batch("vgh", pts)
    tab_k = tab.kernel
    v, g, H = eval_vgh(tab, pts, tab_k)
    assert normwise(v, rv) < TOL[storage] and normwise(H, rH) < TOL[storage]
