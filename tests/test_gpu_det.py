"""Slater-determinant kernels on the GPU (include/determinant_update_hip.h)
against the oracle and against what the mathematics fixes:

  * einspline_update_inverse_{s,d}: Sherman-Morrison update, both forms
    (src/determinant_update.h:8-104 with A; src/determinant_update.cu:10-98
    with the change of row), aligned and unaligned matrices, row strides
    past N, sizes across the register-cached and re-read ranges;
  * einspline_calc_ratios_{s,d} (src/cuda_walker.cu:150-185);
  * einspline_fill_ratios_{s,d}: the row fill (src/cuda_walker.cu:69-95)
    fused with the ratio; its rows must equal, bit for bit, what the value
    entry point writes for the same table and positions.

Tolerances: normwise relative 1e-5 for float, 1e-12 for double (the
summation order of the GPU reductions differs from the oracle's).
"""
import numpy as np
import pytest

import einspline_amd as ea
import oracle
from test_gpu_parity import make_pair, points

pytestmark = pytest.mark.gpu

TOL = {"s": 1e-5, "d": 1e-12}
DT = {"s": np.float32, "d": np.float64}


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ea.device_count() >= 1, "no HIP device visible: GPU tests must run on a GPU"


def rel(got, ref):
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


class Batch:
    """num matrices of N x rs (row-major) in one device buffer, with a device
    list of per-matrix pointers; `shift` elements of offset make every base
    unaligned."""

    def __init__(self, host, shift=0):
        self.host = np.ascontiguousarray(host)
        num = host.shape[0]
        self.es = host.dtype.itemsize
        self.per = int(np.prod(host.shape[1:]))
        flat = np.zeros(num * self.per + shift, host.dtype)
        flat[shift:] = self.host.reshape(-1)
        self.buf = ea.DeviceBuffer.from_array(flat)
        self.shift = shift
        self.list = ea.DeviceBuffer.from_array(
            np.array([self.buf.ptr + (shift + m * self.per) * self.es for m in range(num)], np.uint64))

    def get(self):
        flat = self.buf.download(self.host.dtype, (self.host.size + self.shift,))
        return flat[self.shift:].reshape(self.host.shape)


def make_mats(kind, num, N, rs, rng):
    """Well-conditioned A, their inverses (padded columns hold markers) and
    new rows close to the old ones."""
    A = np.full((num, N, rs), 7.0)
    Ainv = np.full((num, N, rs), -3.0)
    for m in range(num):
        a = np.eye(N) + 0.5 * rng.uniform(-1, 1, (N, N)) / np.sqrt(N)
        A[m, :, :N] = a
        Ainv[m, :, :N] = np.linalg.inv(a)
    return A.astype(DT[kind]), Ainv.astype(DT[kind])


CASES = [(1, 0), (3, 1), (17, 0), (64, 0), (100, 4), (128, 0), (256, 0), (257, 3),
         (300, 0), (1024, 0), (1100, 2)]


@pytest.mark.parametrize("kind", ["s", "d"])
@pytest.mark.parametrize("form", ["row", "delta"])
@pytest.mark.parametrize("N,pad", CASES)
def test_update_inverse_matches_oracle(kind, form, N, pad, rng):
    rs = N + pad
    num = 5 if N <= 300 else 2
    k = (N * 2) // 3
    A, Ainv = make_mats(kind, num, N, rs, rng)
    change = (0.3 * rng.uniform(-1, 1, (num, N)) / np.sqrt(N))
    u = np.zeros((num, N), A.dtype)
    u[:] = (A[:, k, :N] + change) if form == "row" else change
    dA, dI, du = Batch(A), Batch(Ainv), Batch(u)
    delta, colk = Batch(np.zeros((num, N), A.dtype)), Batch(np.zeros((num, N), A.dtype))
    f = getattr(ea.lib, f"einspline_update_inverse_{kind}")
    alist = dA.list.ptr if form == "row" else None
    ea.check(f(alist, dI.list.ptr, du.list.ptr, delta.list.ptr, colk.list.ptr, N, rs, k, num, None),
             "update_inverse")
    ea.synchronize()
    gI, gA, gc = dI.get(), dA.get(), colk.get()
    for m in range(num):
        oA, oI = A[m].copy(), Ainv[m].copy()
        oracle.update_inverse(oA if form == "row" else None, oI, np.ascontiguousarray(
            np.pad(u[m], (0, rs - N))), k, N=N)
        assert rel(gI[m, :, :N], oI[:, :N]) < TOL[kind], m
        assert (gI[m, :, N:] == -3.0).all()                 # padding untouched
        np.testing.assert_array_equal(gc[m], Ainv[m, :, k])  # column k of the old inverse
        np.testing.assert_array_equal(gA[m], oA)             # row k written (row form) or A untouched
        if form == "row":
            np.testing.assert_array_equal(delta.get()[m], u[m] - A[m, k, :N])
    if kind == "d":
        A2 = A[0, :, :N].copy()
        A2[k] += change[0]
        assert np.abs(gI[0, :, :N] @ A2 - np.eye(N)).max() < 1e-10


@pytest.mark.parametrize("kind", ["s", "d"])
def test_update_inverse_unaligned_bases(kind, rng):
    """Matrix bases off 16-byte alignment take the per-element path."""
    N, num, k = 96, 4, 5
    A, Ainv = make_mats(kind, num, N, N, rng)
    u = (A[:, k] + 0.2 * rng.uniform(-1, 1, (num, N)) / np.sqrt(N)).astype(A.dtype)
    dA, dI, du = Batch(A, shift=1), Batch(Ainv, shift=1), Batch(u, shift=1)
    delta, colk = Batch(np.zeros((num, N), A.dtype)), Batch(np.zeros((num, N), A.dtype))
    f = getattr(ea.lib, f"einspline_update_inverse_{kind}")
    ea.check(f(dA.list.ptr, dI.list.ptr, du.list.ptr, delta.list.ptr, colk.list.ptr, N, N, k, num, None), "upd")
    ea.synchronize()
    gI = dI.get()
    for m in range(num):
        oA, oI = A[m].copy(), Ainv[m].copy()
        oracle.update_inverse(oA, oI, u[m].copy(), k)
        assert rel(gI[m], oI) < TOL[kind]


def test_sequence_of_updates_stays_an_inverse(rng):
    """Replacing every row in turn (one particle-by-particle sweep) keeps
    Ainv A = I: errors do not build up."""
    N, num = 128, 3
    A, Ainv = make_mats("d", num, N, N, rng)
    dA, dI = Batch(A), Batch(Ainv)
    delta, colk = Batch(np.zeros((num, N))), Batch(np.zeros((num, N)))
    Ahost = A.copy()
    for k in range(N):
        u = Ahost[:, k] + 0.3 * rng.uniform(-1, 1, (num, N)) / np.sqrt(N)
        Ahost[:, k] = u
        du = Batch(u)
        ea.check(ea.lib.einspline_update_inverse_d(dA.list.ptr, dI.list.ptr, du.list.ptr, delta.list.ptr,
                                                   colk.list.ptr, N, N, k, num, None), "upd")
    ea.synchronize()
    gI, gA = dI.get(), dA.get()
    np.testing.assert_array_equal(gA, Ahost)
    for m in range(num):
        assert np.abs(gI[m] @ Ahost[m] - np.eye(N)).max() < 1e-9


@pytest.mark.parametrize("kind", ["s", "d"])
@pytest.mark.parametrize("N,pad", [(1, 0), (5, 3), (64, 0), (128, 0), (200, 1), (1000, 0)])
def test_calc_ratios_matches_oracle(kind, N, pad, rng):
    rs, num = N + pad, 37
    A, Ainv = make_mats(kind, num, N, rs, rng)
    col = N // 2
    rows = (A[:, col, :N] + 0.4 * rng.uniform(-1, 1, (num, N)) / np.sqrt(N)).astype(A.dtype)
    dI, dr = Batch(Ainv), Batch(rows)
    out = ea.DeviceBuffer(num * A.dtype.itemsize)
    f = getattr(ea.lib, f"einspline_calc_ratios_{kind}")
    ea.check(f(dI.list.ptr, dr.list.ptr, out.ptr, N, rs, col, num, None), "calc_ratios")
    got = out.download(A.dtype, (num,))
    for m in range(num):
        want = oracle.det_ratio(Ainv[m], np.ascontiguousarray(np.pad(rows[m], (0, rs - N))), col, N=N)
        scale = np.abs(Ainv[m, :N, col] * rows[m]).sum()
        assert abs(got[m] - want) <= TOL[kind] * scale
        if kind == "d":
            A2 = A[m, :, :N].astype(np.float64)
            A2[col] = rows[m]
            det = np.linalg.det(A2) / np.linalg.det(A[m, :, :N])
            assert abs(got[m] - det) < 1e-10 * max(1, abs(det))


def fill_setup(kind, storage, M, NS, N, rs, num, rng):
    """16-bit storage is converted from a double host table into a table that
    evaluates in float (kind s)."""
    if storage is None:
        h, _ = make_pair(kind, M, NS, rng)
        tab = ea.DeviceTable(h)
    else:
        h, _ = make_pair("d", M, NS, rng)
        tab = ea.DeviceTable(h, storage=storage)
    pts = points(rng, num)
    fpos = kind == "s"
    pos = ea.DeviceBuffer.from_array(pts.astype(np.float32) if fpos else pts)
    _, Ainv = make_mats(kind, num, N, rs, rng)
    return h, tab, pos, Ainv


@pytest.mark.parametrize("kind,storage", [("s", None), ("s", ea.STORE_BF16), ("s", ea.STORE_F16), ("d", None)])
@pytest.mark.parametrize("NS,N,pad", [(64, 64, 0), (64, 50, 2), (256, 256, 0), (40, 33, 0), (600, 520, 8)])
def test_fill_ratios_rows_bitwise_and_ratios(kind, storage, NS, N, pad, rng):
    rs, num, elec = N + pad, 70, N - 1
    h, tab, pos, Ainv = fill_setup(kind, storage, (9, 8, 7), NS, N, rs, num, rng)
    dt = DT[kind]
    es = np.dtype(dt).itemsize
    rows = ea.DeviceBuffer(num * NS * es)
    rlist = ea.DeviceBuffer.from_array(np.array([rows.ptr + m * NS * es for m in range(num)], np.uint64))
    dI = Batch(Ainv)
    ratio = ea.DeviceBuffer(num * es)
    f = getattr(ea.lib, f"einspline_fill_ratios_{kind}")
    ea.check(f(tab.ptr, pos.ptr, rlist.ptr, dI.list.ptr, ratio.ptr, N, rs, elec, num, None), "fill_ratios")
    # the value entry point on the same table and positions
    ref_rows = ea.DeviceBuffer(num * NS * es)
    ref_list = ea.DeviceBuffer.from_array(np.array([ref_rows.ptr + m * NS * es for m in range(num)], np.uint64))
    ea.check(getattr(ea.lib, f"eval_multi_multi_UBspline_3d_{kind}_hip")(tab.ptr, pos.ptr, ref_list.ptr, num),
             "eval")
    ea.synchronize()
    got_rows = rows.download(dt, (num, NS))
    np.testing.assert_array_equal(got_rows, ref_rows.download(dt, (num, NS)))
    got = ratio.download(dt, (num,))
    for m in range(num):
        u = np.ascontiguousarray(np.pad(got_rows[m, :N], (0, rs - N)))
        want = oracle.det_ratio(Ainv[m], u, elec, N=N)
        scale = np.abs(Ainv[m, :N, elec] * got_rows[m, :N]).sum()
        assert abs(got[m] - want) <= TOL[kind] * max(scale, 1e-30)


def test_fill_ratios_rejects_mismatches(rng):
    h, tab, pos, Ainv = fill_setup("s", None, (6, 6, 6), 16, 16, 16, 4, rng)
    f = ea.lib.einspline_fill_ratios_s
    dummy = Batch(Ainv)
    assert f(tab.ptr, pos.ptr, dummy.list.ptr, dummy.list.ptr, dummy.buf.ptr, 17, 17, 0, 4, None) == ea.E_INVAL
    assert f(tab.ptr, pos.ptr, dummy.list.ptr, dummy.list.ptr, dummy.buf.ptr, 16, 16, 16, 4, None) == ea.E_INVAL
    assert ea.lib.einspline_fill_ratios_d(tab.ptr, pos.ptr, dummy.list.ptr, dummy.list.ptr, dummy.buf.ptr,
                                          16, 16, 0, 4, None) == ea.E_INVAL     # float table
    assert f(tab.ptr, pos.ptr, dummy.list.ptr, dummy.list.ptr, dummy.buf.ptr, 16, 16, 0, 0, None) == ea.E_OK


def test_particle_move_round_trip(rng):
    """One QMC particle move: fill + ratio, then the update with the filled
    row; the ratio equals det(A')/det(A) and the new inverse inverts A'."""
    NS = N = 64
    num, elec = 20, 9
    h, _ = make_pair("d", (10, 9, 8), NS, rng)
    tab = ea.DeviceTable(h)
    pts0 = points(rng, N)                      # one configuration of N electrons, shared
    # A[i][j] = orbital j at electron i, evaluated on the GPU
    P0 = ea.DeviceBuffer.from_array(pts0)
    out = tab.alloc_outputs("v", N)
    tab.eval_batch("v", P0, N, out)
    ea.synchronize()
    A1 = tab.download("v", N, out)[0].reshape(N, NS)
    A1 = A1 + 3 * np.eye(N)                   # keep it well conditioned
    A = np.repeat(A1[None], num, 0)
    Ainv = np.repeat(np.linalg.inv(A1)[None], num, 0)
    dA, dI = Batch(A), Batch(Ainv)
    newpos = ea.DeviceBuffer.from_array(points(rng, num))
    rows = Batch(np.zeros((num, NS)))
    ratio = ea.DeviceBuffer(num * 8)
    ea.check(ea.lib.einspline_fill_ratios_d(tab.ptr, newpos.ptr, rows.list.ptr, dI.list.ptr, ratio.ptr,
                                            N, N, elec, num, None), "fill")
    delta, colk = Batch(np.zeros((num, N))), Batch(np.zeros((num, N)))
    ea.check(ea.lib.einspline_update_inverse_d(dA.list.ptr, dI.list.ptr, rows.list.ptr, delta.list.ptr,
                                               colk.list.ptr, N, N, elec, num, None), "update")
    ea.synchronize()
    u = rows.get()
    r = ratio.download(np.float64, (num,))
    gI = dI.get()
    for m in range(num):
        A2 = A1.copy()
        A2[elec] = u[m]
        want = np.linalg.det(A2) / np.linalg.det(A1)
        assert abs(r[m] - want) < 1e-10 * max(1.0, abs(want))
        if abs(want) > 1e-3:
            assert np.abs(gI[m] @ A2 - np.eye(N)).max() < 1e-8
