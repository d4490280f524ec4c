"""GPU parity: the HIP path (through the C ABI) against the oracle on the
same coefficient tables and points.

Tolerances are normwise per point and per quantity,
    max_n |gpu - oracle| / max_n |oracle|,
with 1e-5 for float tables (s, c) and 1e-12 for double tables (d, z), the
bars the task sets for matching the reference.  Tables are shared exactly:
the oracle evaluates the very coefficients the library built, so eval parity
is separated from create parity (tested in test_create_host.py).
"""
import ctypes as C

import numpy as np
import pytest

import einspline_amd as ea
import oracle

pytestmark = pytest.mark.gpu

TOL = {"s": 1e-5, "c": 1e-5, "d": 1e-12, "z": 1e-12}


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    n = ea.device_count()
    assert n >= 1, "no HIP device visible: GPU tests must run on a GPU"


def normwise(got, ref):
    """Largest per-point normwise relative error."""
    got, ref = np.asarray(got), np.asarray(ref)
    P = ref.shape[0]
    d = np.abs(got - ref).reshape(P, -1).max(1)
    s = np.abs(ref).reshape(P, -1).max(1)
    return float(np.max(d / np.maximum(s, 1e-300)))


def make_pair(kind, M, N, rng, start=(0, 0, 0), end=(1, 1, 1), bcs=None):
    h = ea.HostSpline(kind, M, start, end, N, bc=bcs)
    data = rng.uniform(-1, 1, (N,) + tuple(M))
    if kind in "cz":
        data = data + 1j * rng.uniform(-1, 1, (N,) + tuple(M))
    h.set_block(0, data)
    o = oracle.Table(kind, M, start, end, N, bc=bcs)
    o.coefs[...] = h.coefs().reshape(o.coefs.shape)
    return h, o


def points(rng, P, start=(0, 0, 0), end=(1, 1, 1)):
    s, e = np.array(start, float), np.array(end, float)
    return s + (e - s) * rng.random((P, 3)) * (1 - 1e-9)


def gpu_eval(tab, mode, pts, packed=False):
    P = len(pts)
    pos = ea.DeviceBuffer.from_array(np.ascontiguousarray(pts, np.float64))
    outs = tab.alloc_outputs(mode, P, packed)
    tab.eval_batch(mode, pos, P, outs, packed)
    ea.synchronize()
    res = tab.download(mode, P, outs, packed)
    N = tab.N
    shaped = [res[0].reshape(P, N)]
    if mode != "v":
        shaped.append(res[1].reshape(P, N, 3))
    if mode == "vgl":
        shaped.append(res[2].reshape(P, N))
    if mode == "vgh":
        shaped.append(res[2].reshape(P, N, 6) if packed else res[2].reshape(P, N, 3, 3))
    return shaped


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
@pytest.mark.parametrize("mode", ["v", "vg", "vgl", "vgh"])
@pytest.mark.parametrize("N", [21, 64, 256])
def test_batch_matches_oracle(kind, mode, N, rng):
    M = (9, 8, 10)
    start, end = (3.1, 8.7, 4.5), (9.1, 12.7, 9.3)
    h, o = make_pair(kind, M, N, rng, start, end)
    tab = ea.DeviceTable(h)
    pts = points(rng, 300, start, end)
    got = gpu_eval(tab, mode, pts)
    ref = o.eval_batch(mode, pts)
    for g, r in zip(got, ref):
        assert g.shape == r.shape
        assert normwise(g, r) < TOL[kind], (kind, mode, N)


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_reference_test_geometry(kind, rng):
    """The reference's own multi-vs-single test geometry
    (src/test_multi.c:707-716): periodic 73 x 91 x 29 grid on the box
    [3.1, 9.1] x [8.7, 12.7] x [4.5, 9.3], 21 orbitals of random data; every
    mode through the batched path in both point orders (2048 points, so the
    locality sort runs), the CUDA-style vgh entry point, and the single-point
    API at a few points."""
    M, N = (73, 91, 29), 21
    start, end = (3.1, 8.7, 4.5), (9.1, 12.7, 9.3)
    h, o = make_pair(kind, M, N, rng, start, end)
    tab = ea.DeviceTable(h)
    pts = points(rng, 2048, start, end)
    for mode in ("v", "vg", "vgl", "vgh"):
        ref = o.eval_batch(mode, pts)
        got = gpu_eval(tab, mode, pts)
        for g, r in zip(got, ref):
            assert normwise(g, r) < TOL[kind], (kind, mode)
        tab.set_point_order(ea.ORDER_LOCALITY)
        for a, b in zip(gpu_eval(tab, mode, pts), got):
            np.testing.assert_array_equal(a, b)
        tab.set_point_order(ea.ORDER_GIVEN)
    # CUDA-style pointer arrays (packed Hessian; float positions for s/c)
    P = 256
    fpos = kind in "sc"
    sub = pts[:P]
    used = sub.astype(np.float32).astype(np.float64) if fpos else sub
    pos = ea.DeviceBuffer.from_array(sub.astype(np.float32) if fpos else sub)
    nr, es = tab.nreal, np.dtype(tab.result_dtype).itemsize
    vb, gb, hb = (ea.DeviceBuffer(P * k * nr * es) for k in (1, 3, 6))
    ptrs = lambda b, k: ea.DeviceBuffer.from_array(np.array([b.ptr + p * k * nr * es for p in range(P)], np.uint64))
    vp, gp, hp = ptrs(vb, 1), ptrs(gb, 3), ptrs(hb, 6)
    ea.check(getattr(ea.lib, f"eval_multi_multi_UBspline_3d_{kind}_vgh_hip")(tab.ptr, pos.ptr, vp.ptr, gp.ptr,
                                                                            hp.ptr, P), "vgh_hip")
    rv = o.eval_batch("v", used)[0]
    v = vb.download(tab.result_dtype, (P, nr))
    if kind in "cz":
        v = v.view(np.complex64 if kind == "c" else np.complex128)
    assert normwise(v.reshape(P, N), rv) < TOL[kind]
    # single-point API
    for p in range(3):
        for mode in ("v", "vgh"):
            ref = o.eval_batch(mode, pts[p:p + 1])
            for g, r in zip(h.eval(mode, *pts[p]), ref):
                assert normwise(g[None], r) < TOL[kind], (kind, mode, p)


@pytest.mark.parametrize("kind", ["s", "d"])
def test_nonperiodic_tables(kind, rng):
    bcs = [(ea.NATURAL, ea.NATURAL), (ea.FLAT, ea.DERIV1, 0.0, 0.5), (ea.ANTIPERIODIC, ea.ANTIPERIODIC)]
    M = (7, 9, 6)
    h, o = make_pair(kind, M, 33, rng, (0, -1, 2), (1, 1, 3), bcs)
    tab = ea.DeviceTable(h)
    pts = points(rng, 200, (0, -1, 2), (1, 1, 3))
    for mode in ("v", "vgl", "vgh"):
        for g, r in zip(gpu_eval(tab, mode, pts), o.eval_batch(mode, pts)):
            assert normwise(g, r) < TOL[kind]


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_single_point_api_matches_oracle(kind, rng):
    M, N = (6, 7, 8), 13
    h, o = make_pair(kind, M, N, rng, (0, 0, 0), (2, 1, 3))
    pts = points(rng, 6, (0, 0, 0), (2, 1, 3))
    for mode in ("v", "vg", "vgl", "vgh"):
        ref = o.eval_batch(mode, pts)
        for p in range(len(pts)):
            got = h.eval(mode, *pts[p])
            for g, r in zip(got, ref):
                assert normwise(g[None], r[p:p + 1]) < TOL[kind], (kind, mode)


def test_single_point_sees_new_data_after_set(rng):
    M, N = (6, 6, 6), 4
    h, o = make_pair("d", M, N, rng)
    x = (0.3, 0.4, 0.5)
    v0 = h.eval("v", *x)[0]
    h.set(1, rng.uniform(-1, 1, M))             # marks the device mirror stale
    o.coefs[...] = h.coefs()
    v1 = h.eval("v", *x)[0]
    ref = o.eval_batch("v", np.array([x]))[0][0]
    assert v1[1] != v0[1]
    assert normwise(v1[None], ref[None]) < 1e-12
    # direct writes to coefs need an explicit invalidate
    h.coefs()[..., 2] *= 2.0
    h.invalidate()
    o.coefs[...] = h.coefs()
    v2 = h.eval("v", *x)[0]
    assert normwise(v2[None], o.eval_batch("v", np.array([x]))[0]) < 1e-12


def test_packed_hessian_equals_full(rng):
    h, _ = make_pair("s", (8, 8, 8), 64, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 128)
    v, g, H = gpu_eval(tab, "vgh", pts)
    vp, gp, Hp = gpu_eval(tab, "vgh", pts, packed=True)
    np.testing.assert_array_equal(v, vp)
    np.testing.assert_array_equal(g, gp)
    iu = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for q, (a, b) in enumerate(iu):
        np.testing.assert_array_equal(Hp[..., q], H[..., a, b])
    np.testing.assert_array_equal(H, np.swapaxes(H, -1, -2))


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_cuda_style_pointer_arrays(kind, rng):
    """eval_multi_multi_UBspline_3d_X{,_vgh}_hip: device arrays of per-point
    pointers, packed 6-entry Hessian, float positions for s/c
    (src/multi_bspline_eval_cuda.h:62-137).  Complex tables are written as
    the reference's CUDA kernels write them, as 2N reals: value re, im;
    gradient [re: x y z][im: x y z]; Hessian [re: 6][im: 6] per orbital
    (src/multi_bspline_cuda_z_impl.h:205-228)."""
    M, N, P = (7, 7, 7), 24, 50
    h, o = make_pair(kind, M, N, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, P)
    fpos = kind in "sc"
    pts_used = pts.astype(np.float32).astype(np.float64) if fpos else pts
    pos = ea.DeviceBuffer.from_array(pts.astype(np.float32) if fpos else pts)
    nr = tab.nreal
    es = np.dtype(tab.result_dtype).itemsize
    vb, gb, hb = (ea.DeviceBuffer(P * k * nr * es) for k in (1, 3, 6))
    ptrs = lambda b, k: ea.DeviceBuffer.from_array(
        np.array([b.ptr + p * k * nr * es for p in range(P)], np.uint64))
    vp, gp, hp = ptrs(vb, 1), ptrs(gb, 3), ptrs(hb, 6)
    f = getattr(ea.lib, f"eval_multi_multi_UBspline_3d_{kind}_vgh_hip")
    ea.check(f(tab.ptr, pos.ptr, vp.ptr, gp.ptr, hp.ptr, P), "vgh_hip")
    v = vb.download(tab.result_dtype, (P, nr))
    g = gb.download(tab.result_dtype, (P, 3 * nr))
    hh = hb.download(tab.result_dtype, (P, 6 * nr))
    rv, rg, rh = o.eval_batch("vgh", pts_used)
    if kind in "cz":
        cdt = np.complex64 if kind == "c" else np.complex128
        v = v.view(cdt)
        g2 = g.reshape(P, N, 2, 3)
        g = g2[:, :, 0, :] + 1j * g2[:, :, 1, :]
        h2 = hh.reshape(P, N, 2, 6)
        hh = h2[:, :, 0, :] + 1j * h2[:, :, 1, :]
    assert normwise(v.reshape(P, N), rv) < TOL[kind]
    assert normwise(g.reshape(P, N, 3), rg) < TOL[kind]
    iu = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    ref6 = np.stack([rh[..., a, b] for a, b in iu], -1)
    assert normwise(hh.reshape(P, N, 6), ref6) < TOL[kind]
    # value-only variant
    f = getattr(ea.lib, f"eval_multi_multi_UBspline_3d_{kind}_hip")
    vb.zero()
    ea.check(f(tab.ptr, pos.ptr, vp.ptr, P), "v_hip")
    v = vb.download(tab.result_dtype, (P, nr))
    if kind in "cz":
        v = v.view(cdt)
    assert normwise(v.reshape(P, N), rv) < TOL[kind]


@pytest.mark.parametrize("kind,lbs", [("s", (8, 16, 32)), ("d", (16, 32)), ("c", (8, 16, 32))])
def test_tiles_are_bitwise_identical(kind, lbs, rng):
    h, _ = make_pair(kind, (8, 9, 7), 96, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 77)
    base = gpu_eval(tab, "vgh", pts)
    for pol in (0, 1, 2):
        tab.set_store_policy(pol)
        for lb in lbs:
            for blk in (64, 128, 256):
                tab.set_tile(lb, blk)
                for a, b in zip(gpu_eval(tab, "vgh", pts), base):
                    np.testing.assert_array_equal(a, b)
    tab.set_tile(0, 0)
    tab.set_store_policy(0)


def test_results_independent_of_batch(rng):
    """A point's result is bitwise the same alone and inside any batch."""
    h, _ = make_pair("s", (10, 10, 10), 256, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 500)
    full = gpu_eval(tab, "vgh", pts)
    for idx in ([0], [499], list(range(100, 137)), [3, 250, 17]):
        part = gpu_eval(tab, "vgh", pts[idx])
        for a, b in zip(part, full):
            np.testing.assert_array_equal(a, b[idx])


def test_edge_points_and_empty_batch(rng):
    M = (6, 6, 6)
    h, o = make_pair("d", M, 8, rng)
    tab = ea.DeviceTable(h)
    # grid nodes, the start corner and points just below the end
    pts = np.array([[0, 0, 0], [0.5, 0.5, 0.5], [1 - 1e-12, 1 - 1e-12, 1 - 1e-12],
                    [1 / 6, 2 / 6, 5 / 6]], float)
    for g, r in zip(gpu_eval(tab, "vgh", pts), o.eval_batch("vgh", pts)):
        assert normwise(g, r) < 1e-12
    # x == end: the reference reads one plane past the table; here the edge
    # cell is extended, which equals the limit from inside
    at_end = gpu_eval(tab, "vgh", np.array([[1.0, 1.0, 1.0]]))
    near = o.eval_batch("vgh", np.array([[1 - 1e-13] * 3]))
    for g, r in zip(at_end, near):
        assert normwise(g, r) < 1e-9
    # points outside the box stay finite (memory-safe extrapolation)
    far = gpu_eval(tab, "v", np.array([[-3.0, 0.5, 7.0]]))
    assert np.isfinite(far[0]).all()
    # empty batch is a no-op
    pos = ea.DeviceBuffer(24)
    outs = tab.alloc_outputs("v", 1)
    tab.eval_batch("v", pos, 0, outs)
    ea.synchronize()


def test_errors_are_reported(rng):
    h, _ = make_pair("s", (5, 5, 5), 4, rng)
    tab = ea.DeviceTable(h)
    with pytest.raises(ea.EinsplineError):
        tab.set_tile(12, 0)
    rc = ea.lib.einspline_eval_v_batch(tab.ptr, None, 5, None, None)
    assert rc == ea.E_INVAL
    # wrong-type device table for a typed entry point
    hd, _ = make_pair("d", (5, 5, 5), 4, rng)
    tabd = ea.DeviceTable(hd)
    rc = ea.lib.eval_multi_multi_UBspline_3d_s_hip(tabd.ptr, None, None, 1)
    assert rc == ea.E_INVAL


def test_autotune_picks_a_valid_tile_and_keeps_results(rng):
    h, o = make_pair("s", (10, 9, 8), 256, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, 200)
    before = gpu_eval(tab, "vgh", pts)
    lb, blk, nt, ms = tab.autotune("vgh", 4096)
    assert lb in (8, 16, 32) and blk in (64, 128, 256) and nt in (0, 1, 2) and ms > 0
    after = gpu_eval(tab, "vgh", pts)
    # tiles and orders never change results; the matrix-core kernel (which
    # the tuner may pick for vgh) sums in another order, within the FP32 bar
    for a, b in zip(after, before):
        if tab.kernel == ea.KERNEL_VECTOR:
            np.testing.assert_array_equal(a, b)
        else:
            assert normwise(a, b) < TOL["s"]
    # second call on the same shape hits the cache
    tab2 = ea.DeviceTable(h)
    lb2, blk2, nt2, ms2 = tab2.autotune("vgh", 4096)
    assert (lb2, blk2, nt2) == (lb, blk, nt) and ms2 < 0


def _bf16_rne_from_f64(x):
    """Round float64 values (normal bf16 range) to bf16 bit patterns, nearest even."""
    b = np.asarray(x, np.float64).view(np.uint64)
    sign = (b >> np.uint64(63)).astype(np.uint32)
    lsb = (b >> np.uint64(45)) & np.uint64(1)
    r = (b & np.uint64(0x7FFFFFFFFFFFFFFF)) + np.uint64((1 << 44) - 1) + lsb
    e = (r >> np.uint64(52)).astype(np.int64) - 1023 + 127
    m = ((r >> np.uint64(45)) & np.uint64(0x7F)).astype(np.uint32)
    return ((sign << 15) | (e.astype(np.uint32) << 7) | m).astype(np.uint16)


@pytest.mark.parametrize("storage", [ea.STORE_BF16, ea.STORE_F16, ea.STORE_F32])
def test_storage_conversion_rounds_to_nearest_even(storage, rng):
    """The device table holds each FP64 coefficient rounded once to the
    storage type (no double rounding through float)."""
    M, N = (4, 4, 4), 5
    h = ea.HostSpline("d", M, (0, 0, 0), (1, 1, 1), N)
    c = h.coefs()
    vals = rng.uniform(0.5, 2.0, c.shape) * np.where(rng.random(c.shape) < 0.5, -1, 1)
    # include values sitting just off a bf16 / f16 rounding midpoint
    vals.flat[0] = 1.0 + 2.0**-8 + 2.0**-30      # just above a bf16 midpoint: rounds up
    vals.flat[1] = 1.0 + 2.0**-8 - 2.0**-30      # just below: rounds down
    vals.flat[2] = 1.0 + 2.0**-11 + 2.0**-40     # f16 midpoint + tiny: rounds up
    c[...] = vals
    tab = ea.DeviceTable(h, storage=storage)
    t = tab.ptr.contents
    rows = t.Nx * t.Ny * t.Nz
    if storage == ea.STORE_F32:
        raw = np.empty((rows, t.row_reals), np.float32)
        ea.check(ea.lib.einspline_hip_memcpy_dtoh(raw.ctypes.data, t.coefs, raw.nbytes), "dtoh")
        np.testing.assert_array_equal(raw[:, :N], vals.reshape(rows, N).astype(np.float32))
        assert not raw[:, N:].any()
        return
    raw = np.empty((rows, t.row_reals), np.uint16)
    ea.check(ea.lib.einspline_hip_memcpy_dtoh(raw.ctypes.data, t.coefs, raw.nbytes), "dtoh")
    flat = vals.reshape(rows, N)
    if storage == ea.STORE_BF16:
        want = _bf16_rne_from_f64(flat)
    else:
        want = flat.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(raw[:, :N], want)
    assert not raw[:, N:].any()


def test_f16_overflow_is_reported(rng):
    h = ea.HostSpline("d", (4, 4, 4), (0, 0, 0), (1, 1, 1), 2)
    h.coefs()[...] = 1.0
    h.coefs()[1, 1, 1, 0] = 70000.0
    with pytest.raises(ea.EinsplineError) as e:
        ea.DeviceTable(h, storage=ea.STORE_F16)
    assert "range" in str(e.value)


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_qmc_vgl_with_lattice_and_row_stride(kind, rng):
    """eval_multi_multi_UBspline_3d_X_vgl_hip: value, Linv-transformed
    gradient and tr((Linv^T Linv) H) as SoA rows row_stride apart
    (src/multi_bspline_cuda_s_impl.h:566-734)."""
    M, N, P = (7, 8, 6), 40, 33
    h, o = make_pair(kind, M, N, rng)
    tab = ea.DeviceTable(h)
    pts = points(rng, P)
    fpos = kind in "sc"
    real = np.float32 if fpos else np.float64
    pts_used = pts.astype(np.float32).astype(np.float64) if fpos else pts
    pos = ea.DeviceBuffer.from_array(pts.astype(real))
    G = np.array([[1.2, 0.1, -0.3], [0.0, 0.9, 0.2], [0.05, -0.1, 1.1]])
    linv = ea.DeviceBuffer.from_array(G.astype(real).ravel())
    nc = 2 if kind in "cz" else 1
    es = np.dtype(real).itemsize * nc
    rs = N + 5                                    # row stride > N
    vb = ea.DeviceBuffer(P * N * es)
    gb = ea.DeviceBuffer(P * 4 * rs * es)
    gb.zero()
    vp = ea.DeviceBuffer.from_array(np.array([vb.ptr + p * N * es for p in range(P)], np.uint64))
    gp = ea.DeviceBuffer.from_array(np.array([gb.ptr + p * 4 * rs * es for p in range(P)], np.uint64))
    f = getattr(ea.lib, f"eval_multi_multi_UBspline_3d_{kind}_vgl_hip")
    ea.check(f(tab.ptr, pos.ptr, linv.ptr, vp.ptr, gp.ptr, P, rs), "vgl_hip")
    cdt = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}[kind]
    v = vb.download(real, (P, N * nc)).view(cdt)
    gl = gb.download(real, (P, 4 * rs * nc)).view(cdt).reshape(P, 4, rs)
    rv, rg, rH = o.eval_batch("vgh", pts_used)
    Gd = G.astype(real).astype(np.float64)
    g_ref = np.einsum("dj,pnj->pnd", Gd, rg)
    lap_ref = np.einsum("ab,pnab->pn", Gd.T @ Gd, rH)
    tol = TOL[kind] * 10
    assert normwise(v, rv) < TOL[kind]
    assert normwise(np.moveaxis(gl[:, :3, :N], 1, 2), g_ref) < tol
    assert normwise(gl[:, 3, :N], lap_ref) < tol
    assert not gl[:, :, N:].any()                 # padding between rows untouched
    if kind == "s":
        sign = ea.DeviceBuffer.from_array(np.where(rng.random(P) < 0.5, -1.0, 1.0).astype(np.float32))
        sg = sign.download(np.float32, (P,))
        f = ea.lib.eval_multi_multi_UBspline_3d_s_vgl_sign_hip
        ea.check(f(tab.ptr, pos.ptr, sign.ptr, linv.ptr, vp.ptr, gp.ptr, P, rs), "vgl_sign_hip")
        v2 = vb.download(real, (P, N))
        gl2 = gb.download(real, (P, 4, rs))
        np.testing.assert_array_equal(v2, sg[:, None] * v)
        np.testing.assert_array_equal(gl2[:, :, :N], sg[:, None, None] * gl[:, :, :N])
        ea.check(ea.lib.eval_multi_multi_UBspline_3d_s_sign_hip(tab.ptr, pos.ptr, sign.ptr, vp.ptr, P),
                 "sign_hip")
        np.testing.assert_array_equal(vb.download(real, (P, N)), sg[:, None] * v)


def test_gpu_against_golden_fixtures():
    """The GPU path against the closed-form known-answer vectors directly
    (tests/golden, written by make_golden.py without any oracle code)."""
    import os
    gdir = os.path.join(os.path.dirname(__file__), "golden")
    for f in sorted(x for x in os.listdir(gdir) if x.endswith(".npz")):
        z = np.load(os.path.join(gdir, f))
        kind = str(z["kind"])
        bcs = [tuple(int(c) for c in b) for b in z["bcs"].tolist()]
        h = ea.HostSpline(kind, tuple(z["M"]), tuple(z["start"]), tuple(z["end"]), int(z["N"]), bc=bcs)
        h.coefs()[...] = z["coefs"].reshape(h.coefs().shape)
        h.invalidate()
        tab = ea.DeviceTable(h)
        v, g, H = gpu_eval(tab, "vgh", z["pts"])
        tol = 1e-12 if kind in "dz" else 5e-6
        for got, want in ((v, z["v"]), (g, z["g"]), (H, z["H"])):
            err = np.max(np.abs(got - want)) / np.max(np.abs(want))
            assert err < tol, (f, err)
