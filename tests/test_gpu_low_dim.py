"""1-D and 2-D multi-splines on the GPU against the oracle: the strided batch
entry points, the single-point API (eval_multi_UBspline_{1,2}d_X*,
src/multi_bspline_eval_std_d_impl.h:37-476), and the 1-D CUDA-style pointer
forms (src/multi_bspline_eval_cuda.h:10-51).  Tolerances as for 3-D:
1e-5 normwise for float tables, 1e-12 for double tables."""
import numpy as np
import pytest

import einspline_amd as ea
import oracle
from test_gpu_parity import TOL, normwise

pytestmark = pytest.mark.gpu

BCS = [(ea.PERIODIC, ea.PERIODIC), (ea.NATURAL, ea.FLAT)]


def make_pair(kind, M, N, rng, start, end, bcs):
    h = ea.HostSpline(kind, M, start, end, N, bc=bcs)
    data = rng.uniform(-1, 1, (N,) + tuple(M))
    if kind in "cz":
        data = data + 1j * rng.uniform(-1, 1, (N,) + tuple(M))
    h.set_block(0, data)
    o = oracle.Table(kind, M, start, end, N, bc=bcs)
    o.coefs[...] = h.coefs().reshape(o.coefs.shape)
    return h, o


def points(rng, P, start, end):
    s, e = np.array(start, float), np.array(end, float)
    return s + (e - s) * rng.random((P, len(start))) * (1 - 1e-9)


def gpu_eval(tab, mode, pts):
    P, D, N = len(pts), tab.dim, tab.N
    pos = ea.DeviceBuffer.from_array(np.ascontiguousarray(pts, np.float64))
    outs = tab.alloc_outputs(mode, P)
    tab.eval_batch(mode, pos, P, outs)
    ea.synchronize()
    res = tab.download(mode, P, outs)
    shaped = [res[0].reshape(P, N)]
    if mode != "v":
        shaped.append(res[1].reshape(P, N, D))
    if mode == "vgl" or (mode == "vgh" and D == 1):
        shaped.append(res[2].reshape(P, N))
    elif mode == "vgh":
        shaped.append(res[2].reshape(P, N, D, D))
    return shaped


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
@pytest.mark.parametrize("mode", ["v", "vg", "vgl", "vgh"])
@pytest.mark.parametrize("N", [5, 64, 200])
def test_low_dim_batch_matches_oracle(dim, kind, mode, N, rng):
    M, start, end = (11, 9)[:dim], (1.5, -2.0)[:dim], (4.0, 3.5)[:dim]
    h, o = make_pair(kind, M, N, rng, start, end, BCS[:dim])
    tab = ea.DeviceTable(h)
    assert tab.dim == dim
    pts = points(rng, 300, start, end)
    got = gpu_eval(tab, mode, pts)
    ref = o.eval_batch(mode, pts)
    assert len(got) == len(ref)
    for a, r in zip(got, ref):
        assert normwise(a, r) < TOL[kind]


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_low_dim_single_point_api(dim, kind, rng):
    M, start, end = (8, 7)[:dim], (0.0, 0.5)[:dim], (2.0, 1.5)[:dim]
    h, o = make_pair(kind, M, 24, rng, start, end, BCS[:dim])
    for x in points(rng, 3, start, end):
        for mode in ("v", "vg", "vgl", "vgh"):
            got = h.eval(mode, *x)
            ref = o.eval_batch(mode, x[None])
            for a, r in zip(got, ref):
                assert normwise(a[None], r) < TOL[kind]
    # the mirror follows set_*
    h.set(0, np.zeros(M))
    x = points(rng, 1, start, end)[0]
    assert not h.eval("v", *x)[0][0]


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_1d_cuda_style_pointer_forms(kind, rng):
    M, N, P = 13, 40, 70
    start, end = (0.0,), (1.0,)
    h, o = make_pair(kind, (M,), N, rng, start, end, [BCS[0]])
    tab = ea.DeviceTable(h)
    x = points(rng, P, start, end)[:, 0]
    fpos = kind in "sc"
    x_used = x.astype(np.float32).astype(np.float64) if fpos else x
    pos = ea.DeviceBuffer.from_array(x.astype(np.float32) if fpos else x)
    nr, es = tab.nreal, np.dtype(tab.result_dtype).itemsize
    cdt = {"s": None, "d": None, "c": np.complex64, "z": np.complex128}[kind]
    view = (lambda a: a.view(cdt)) if cdt else (lambda a: a)
    rv, rg, rl = o.eval_batch("vgl", x_used[:, None])
    ptrs = lambda b, k: ea.DeviceBuffer.from_array(
        np.array([b.ptr + p * k * nr * es for p in range(P)], np.uint64))
    # values
    vb = ea.DeviceBuffer(P * nr * es)
    vp = ptrs(vb, 1)
    ea.check(getattr(ea.lib, f"eval_multi_multi_UBspline_1d_{kind}_hip")(tab.ptr, pos.ptr, vp.ptr, P), "v")
    assert normwise(view(vb.download(tab.result_dtype, (P, nr))), rv) < TOL[kind]
    if kind == "s":
        gb, lb = ea.DeviceBuffer(P * nr * es), ea.DeviceBuffer(P * nr * es)
        gp, lp = ptrs(gb, 1), ptrs(lb, 1)
        ea.check(ea.lib.eval_multi_multi_UBspline_1d_s_vgl_hip(tab.ptr, pos.ptr, vp.ptr, gp.ptr, lp.ptr, P),
                 "vgl")
        g = gb.download(np.float32, (P, N))
        l = lb.download(np.float32, (P, N))
    else:
        nc = nr // N
        rs = N + 3                                 # gradient row, then the Laplacian row
        per = 2 * rs * nc                          # reals per point
        glb = ea.DeviceBuffer(P * per * es)
        glp = ea.DeviceBuffer.from_array(np.array([glb.ptr + p * per * es for p in range(P)], np.uint64))
        f = getattr(ea.lib, f"eval_multi_multi_UBspline_1d_{kind}_vgl_hip")
        ea.check(f(tab.ptr, pos.ptr, vp.ptr, glp.ptr, P, rs), "vgl")
        gl = glb.download(tab.result_dtype, (P, per))
        g = gl[:, :nr]
        l = gl[:, rs * nc: rs * nc + nr]
    assert normwise(view(vb.download(tab.result_dtype, (P, nr))), rv) < TOL[kind]
    assert normwise(view(np.ascontiguousarray(g)), rg[..., 0]) < TOL[kind]
    assert normwise(view(np.ascontiguousarray(l)), rl) < TOL[kind]


@pytest.mark.parametrize("dim", [1, 2])
def test_low_dim_conv_tables(dim, rng):
    """create_multi_UBspline_{1,2}d_{s,c}_hip_conv: double tables stored as float."""
    M, start, end = (12, 10)[:dim], (0.0, 0.0)[:dim], (1.0, 1.0)[:dim]
    for kind in ("d", "z"):
        h, o = make_pair(kind, M, 32, rng, start, end, BCS[:dim])
        tab = ea.DeviceTable(h, storage="conv")
        assert tab.result_dtype == np.float32
        pts = points(rng, 100, start, end)
        for a, r in zip(gpu_eval(tab, "vgh", pts), o.eval_batch("vgh", pts)):
            assert normwise(a, r) < 1e-5


def test_low_dim_tiles_and_rejected_modes(rng):
    h, _ = make_pair("s", (30, 20), 256, rng, (0.0, 0.0), (1.0, 1.0), BCS)
    tab = ea.DeviceTable(h)
    pts = points(rng, 500, (0.0, 0.0), (1.0, 1.0))
    base = gpu_eval(tab, "vgh", pts)
    for blk in (64, 128):
        tab.set_tile(16, blk)
        tab.set_point_order(ea.ORDER_LOCALITY)        # 1-D/2-D use the given order
        for a, b in zip(gpu_eval(tab, "vgh", pts), base):
            np.testing.assert_array_equal(a, b)
    pos = ea.DeviceBuffer.from_array(pts)
    outs = tab.alloc_outputs("vgh", 500)
    rc = ea.lib.einspline_eval_vgh_batch(tab.ptr, pos.ptr, 500, outs[0].ptr, outs[1].ptr,
                                         outs[2].ptr, 1, None)
    assert rc == ea.E_INVAL                               # packed Hessians are 3-D only
    with pytest.raises(ea.EinsplineError):
        tab.autotune("vgh", 1024)
    big = ea.DeviceBuffer(500 * 27 * 256 * 4)
    rc = ea.lib.einspline_eval_vghgh_batch(tab.ptr, pos.ptr, 500, outs[0].ptr, outs[1].ptr,
                                           outs[2].ptr, big.ptr, None)
    assert rc == ea.E_INVAL
