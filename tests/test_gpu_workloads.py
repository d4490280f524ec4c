"""The two large workload configurations at their full size, and the
headline configuration at the tile and point order the benchmark times.

* Precision form: periodic 96^3 table of 512 smooth plane-wave orbitals,
  built in FP64, stored FP32 / BF16 / FP16 with FP32 accumulation, vgh at
  131072 uniform points, each storage at its autotuned kernel, tile and
  point order.  Checked on a 512-point sample against the oracle's FP64
  evaluation of the table's own stored values (parity of the evaluation:
  1e-5 for FP32 and FP16, 1e-4 for BF16 on the matrix cores, whose weights
  carry 2^-18), against the FP64 table (the precision form's error report,
  with bars that hold at this size), and on the whole batch through
  properties: finite, symmetric Hessians, the same bits in the other point
  order.
* Serving form: the three real tables (48^3 x 256 FP32, 64^3 x 512 BF16,
  96^3 x 1024 FP32; 4.42 GB on the device) all resident and under a 4.3 GB
  budget; 1200 requests of log-uniform 1..2048 points, each bitwise equal to
  the same request evaluated alone with einspline_eval_vgh_batch, plus an
  oracle sample.
* Headline: 48^3 x 256 FP32 vgh at 32768 points, at the autotuned tile and
  at 16-byte lanes x 256 with staged stores in the locality order, against
  the oracle on 512 points and bitwise against the default tile in the
  given order on every point.
"""

import numpy as np
import pytest

import einspline_amd as ea
import oracle
from einspline_amd.serve import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ea.device_count() >= 1, "no HIP device visible: GPU tests must run on a GPU"


def say(msg):
    print(f"[workloads] {msg}", flush=True)


def plane_waves(rng, M, count, nwaves=8, kmax=4):
    """count orbitals f(r) = Re sum_m c_m exp(2 pi i k_m.r), integer |k_i| <=
    kmax, c_m ~ N(0,1) + i N(0,1), sampled at r = i/M; (count, M, M, M)."""
    x = np.arange(M) / M
    K = rng.integers(-kmax, kmax + 1, size=(count, nwaves, 3))
    c = rng.standard_normal((count, nwaves)) + 1j * rng.standard_normal((count, nwaves))
    E = np.exp(2j * np.pi * K[..., None] * x)                       # (count, nwaves, 3, M)
    XY = c[..., None, None] * E[:, :, 0, :, None] * E[:, :, 1, None, :]
    XY = XY.reshape(count, nwaves, M * M).transpose(0, 2, 1)        # (count, M*M, nwaves)
    return np.matmul(XY, E[:, :, 2, :]).real.reshape(count, M, M, M)


def normwise(got, ref):
    """Largest per-point normwise relative error max_n|d| / max_n|ref|."""
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    P = ref.shape[0]
    d = np.abs(got - ref).reshape(P, -1).max(1)
    s = np.abs(ref).reshape(P, -1).max(1)
    return float(np.max(d / np.maximum(s, 1e-300)))


def rel_max(got, ref):
    """max|d| / max|ref| over the whole sample (the precision form's report)."""
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    return float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


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
    return np.ascontiguousarray(vals.reshape(t.Nx, t.Ny, t.Nz, t.num_splines))


def eval_vgh(tab, pos, P):
    outs = tab.alloc_outputs("vgh", P)
    tab.eval_batch("vgh", pos, P, outs)
    ea.synchronize()
    v, g, h = tab.download("vgh", P, outs)[:3]
    for b in outs[:3]:
        b.free()
    N = tab.N
    return v.reshape(P, N), g.reshape(P, N, 3), h.reshape(P, N, 3, 3)


# ---------------------------------------------------------------------------
# precision form: 96^3, 512 orbitals, 131072 points

C4_M, C4_N, C4_P, C4_SAMPLE = 96, 512, 131072, 512


@pytest.fixture(scope="module")
def c4():
    rng = np.random.default_rng(4)
    M, N = C4_M, C4_N
    h = ea.HostSpline("d", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    for first in range(0, N, 32):
        h.set_block(first, plane_waves(rng, M, 32))
    say("precision form: FP64 table 96^3 x 512 built on the host")
    pts = rng.random((C4_P, 3))
    idx = np.sort(rng.choice(C4_P, C4_SAMPLE, replace=False))
    o = oracle.Table("d", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    o.coefs = h.coefs()                       # the FP64 table itself, shared
    ref = o.eval_batch("vgh", pts[idx])
    yield h, pts, idx, ref
    h.destroy()


# (parity bar of the vector kernel vs the stored table's own values, error
# bars vs the FP64 table for v, g, h): the latter hold with a factor ~2.5 to
# spare at this size
C4_BARS = {
    ea.STORE_F32: (1e-5, (1e-5, 1e-5, 1e-5)),
    ea.STORE_BF16: (1e-5, (3e-3, 2.5e-2, 2.5e-1)),
    ea.STORE_F16: (1e-5, (4e-4, 3e-3, 3e-2)),
}

# The matrix-core kernel applies every weight as a 16-bit high part plus a
# 16-bit remainder: the pair carries the weight to a relative 2^-18 (BF16,
# 8-bit significands) or 2^-22 (FP16); FP32 tables take FP32 weights (2^-24
# per rounding).  Each output sums its 64 stencil products (128 with the two
# weight parts) in four matrix instructions of 32 products; taken as an adder
# that aligns one instruction's products to the largest and keeps a 24-bit
# significand, each instruction adds at most 32 x 2^-23 of its largest
# product, 2^-16 of the sum of magnitudes over the four.  The FP32 products
# of three basis values add a few 2^-24.
MATRIX_ACC_EPS = 2.0**-16 + 8 * 2.0**-24
SPLIT_EPS = {ea.STORE_BF16: 2.0**-18, ea.STORE_F16: 2.0**-22, ea.STORE_F32: 2.0**-24}


def _basis_abs(t, order):
    """|d^order/dt^order a_i(t)| of the four uniform cubic B-spline pieces."""
    t = np.asarray(t, np.float64)
    if order == 0:
        b = [(1 - t)**3 / 6, (3 * t**3 - 6 * t**2 + 4) / 6, (-3 * t**3 + 3 * t**2 + 3 * t + 1) / 6, t**3 / 6]
    elif order == 1:
        b = [-(1 - t)**2 / 2, (9 * t**2 - 12 * t) / 6, (-9 * t**2 + 6 * t + 3) / 6, t**2 / 2]
    else:
        b = [1 - t, 3 * t - 2, 1 - 3 * t, t]
    return np.abs(np.stack(b, -1))


def weight_bound(coefs, pts, M, start, end):
    """S[q, n] = sum over the 64 stencil rows of |weight_q| |coefficient_n|
    for every output q of vgh at each point (periodic table of M cells per
    axis): the scale that bounds the rounding error of any order of
    summation.  Returns (P, N), (P, N, 3), (P, N, 3, 3)."""
    pts = np.asarray(pts, np.float64)
    dinv = np.array(M, float) / (np.array(end, float) - np.array(start, float))
    u = (pts - np.array(start, float)) * dinv
    ix = np.clip(np.floor(u).astype(int), 0, np.array(coefs.shape[:3]) - 4)
    t = u - ix
    P, N = len(pts), coefs.shape[3]
    W = [[_basis_abs(t[:, a], o) * dinv[a]**o for o in range(3)] for a in range(3)]   # [axis][order] (P, 4)
    Sv = np.empty((P, N))
    Sg = np.empty((P, N, 3))
    SH = np.empty((P, N, 3, 3))
    off = np.arange(4)
    for p in range(P):
        C = np.abs(coefs[np.ix_(ix[p, 0] + off, ix[p, 1] + off, ix[p, 2] + off)])       # (4, 4, 4, N)

        def s(ox, oy, oz):
            return np.einsum("i,j,k,ijkn->n", W[0][ox][p], W[1][oy][p], W[2][oz][p], C)
        Sv[p] = s(0, 0, 0)
        for a in range(3):
            o = [0, 0, 0]
            o[a] = 1
            Sg[p, :, a] = s(*o)
            for b in range(3):
                o = [0, 0, 0]
                o[a] += 1
                o[b] += 1
                SH[p, :, a, b] = s(*o)
    return Sv, Sg, SH


@pytest.mark.parametrize("storage", [ea.STORE_F32, ea.STORE_BF16, ea.STORE_F16])
def test_precision_form_full_workload(c4, storage):
    h, pts, idx, (rv, rg, rH) = c4
    P, N = len(pts), h.N
    say(f"precision form storage {storage}: upload and autotune")
    tab = ea.DeviceTable(h, storage=storage)
    tab.autotune("vgh", P)
    tuning, kernel = tab.tuning(), tab.kernel
    say(f"  tuned tile/policy/order {tuning}, kernel {kernel}")
    pos = ea.DeviceBuffer.from_array(pts)
    v, g, H = eval_vgh(tab, pos, P)
    # whole batch: finite, symmetric Hessians
    assert np.isfinite(v).all() and np.isfinite(g).all() and np.isfinite(H).all()
    np.testing.assert_array_equal(H, np.swapaxes(H, -1, -2))
    # whole batch: the given order without orbital slabs gives the same bits
    # (the tuned order is the locality order, slabs outermost, or both)
    tab.set_point_order(0 if tuning[3] else 1)
    v2, g2, H2 = eval_vgh(tab, pos, P)
    tab.set_point_order(tuning[3])
    np.testing.assert_array_equal(v2, v)
    np.testing.assert_array_equal(g2, g)
    np.testing.assert_array_equal(H2, H)
    del v2, g2, H2
    # parity: the oracle's FP64 evaluation of the table's own stored values
    o = oracle.Table("d", (C4_M,) * 3, (0, 0, 0), (1, 1, 1), N)
    o.coefs = table_values(tab)
    sv, sg, sH = o.eval_batch("vgh", pts[idx])
    errs = [normwise(v[idx], sv), normwise(g[idx], sg), normwise(H[idx], sH)]
    say(f"  parity vs the stored table (normwise per point): {errs}")
    if kernel == ea.KERNEL_MATRIX:
        # every output within the a-priori bound of the split-weight products
        Sv, Sg, SH = weight_bound(o.coefs, pts[idx], (C4_M,) * 3, (0, 0, 0), (1, 1, 1))
        eps = SPLIT_EPS[storage] + MATRIX_ACC_EPS
        ratio = max(float(np.max(np.abs(np.asarray(a, np.float64) - r) / (eps * S + 1e-300)))
                    for a, r, S in ((v[idx], sv, Sv), (g[idx], sg, Sg), (H[idx], sH, SH)))
        say(f"  matrix kernel: largest error / a-priori bound = {ratio:.3f}")
        assert ratio <= 1.0
    else:
        assert max(errs) < 1e-5
    del o
    _, (ev, eg, eh) = C4_BARS[storage]
    # the precision form's error against the FP64 table
    errs64 = [rel_max(v[idx], rv), rel_max(g[idx], rg), rel_max(H[idx], rH)]
    say(f"  error vs the FP64 table (max rel): {errs64}")
    assert errs64[0] < ev and errs64[1] < eg and errs64[2] < eh
    tab.destroy()


# ---------------------------------------------------------------------------
# serving form: the three real tables


class _At:
    """A device address inside an arena (what Server.submit reads .ptr of)."""

    def __init__(self, ptr):
        self.ptr = ptr


@pytest.fixture(scope="module")
def c5():
    rng = np.random.default_rng(55)
    hosts = []
    for kind, M, N in (("s", 48, 256), ("d", 64, 512), ("s", 96, 1024)):
        hh = ea.HostSpline(kind, (M, M, M), (0, 0, 0), (1, 1, 1), N)
        cf = hh.coefs()
        dt = cf.dtype
        for i in range(cf.shape[0]):          # plane by plane: random coefficients, no solve
            cf[i] = rng.uniform(-1, 1, cf.shape[1:]).astype(dt)
        hosts.append(hh)
    storages = [ea.STORE_F32, ea.STORE_BF16, ea.STORE_F32]
    say("serving: three host tables filled")
    # the requests: table uniform, log-uniform 1..2048 points, uniform positions
    reqs = []
    for _ in range(1200):
        k = int(rng.integers(0, 3))
        n = int(min(2048, max(1, int(np.exp(rng.uniform(0, np.log(2049)))))))
        reqs.append((k, rng.random((n, 3))))
    yield hosts, storages, reqs
    for hh in hosts:
        hh.destroy()


def _arena_layout(hosts, reqs):
    """Per request (v, g, h) offsets in one float arena; total floats."""
    offs, at = [], 0
    for k, pts in reqs:
        n, N = len(pts), hosts[k].N
        o = []
        for c in (1, 3, 9):
            o.append(at)
            at += n * N * c
            at = (at + 63) // 64 * 64
        offs.append(o)
    return offs, at


def _alone(hosts, storages, reqs, offs, total):
    """Every request evaluated by itself with einspline_eval_vgh_batch (the
    default vector kernel, given order) into an arena laid out like the
    server's; returns the arena as uint32 words."""
    arena = ea.DeviceBuffer(total * 4)
    arena.zero()
    tabs = [ea.DeviceTable(h, storage=None if h.kind == "s" else st) for h, st in zip(hosts, storages)]
    for (k, pts), o in zip(reqs, offs):
        n = len(pts)
        pos = ea.DeviceBuffer.from_array(pts)
        ea.check(ea.lib.einspline_eval_vgh_batch(tabs[k].ptr, pos.ptr, n, arena.ptr + 4 * o[0],
                                                 arena.ptr + 4 * o[1], arena.ptr + 4 * o[2], 0, None),
                 "einspline_eval_vgh_batch")
        ea.synchronize()
        pos.free()
    out = arena.download(np.uint32, (total,))
    arena.free()
    return out, tabs


def _serve(srv, ids, hosts, reqs, offs, total):
    arena = ea.DeviceBuffer(total * 4)
    arena.zero()
    tickets = []
    for (k, pts), o in zip(reqs, offs):
        tickets.append(srv.submit(ids[k], pts, _At(arena.ptr + 4 * o[0]), _At(arena.ptr + 4 * o[1]),
                                  _At(arena.ptr + 4 * o[2])))
    lats = [srv.wait(t) for t in tickets]
    assert all(lt > 0 for lt in lats)
    out = arena.download(np.uint32, (total,))
    arena.free()
    return out, np.array(lats)


def test_serving_three_real_tables_mix_equals_alone(c5):
    hosts, storages, reqs = c5
    offs, total = _arena_layout(hosts, reqs)
    say(f"serving: {len(reqs)} requests, {sum(len(p) for _, p in reqs)} points, "
        f"{total * 4 / 1e9:.2f} GB of results")
    alone, tabs = _alone(hosts, storages, reqs, offs, total)
    tbytes = [t.bytes for t in tabs]
    say(f"serving: alone done; device table bytes {tbytes} (total {sum(tbytes) / 1e9:.3f} GB)")
    assert sum(tbytes) > 4.3e9 > max(tbytes)
    # oracle sample: the first small requests of each table
    for tix in range(3):
        h = hosts[tix]
        picks = [i for i, (k, p) in enumerate(reqs) if k == tix and len(p) <= 64][:4]
        if storages[tix] == ea.STORE_F32:
            o = oracle.Table("s", h.M, (0, 0, 0), (1, 1, 1), h.N)
            o.coefs = h.coefs()
        else:
            o = oracle.Table("d", h.M, (0, 0, 0), (1, 1, 1), h.N)
            o.coefs = table_values(tabs[tix])
        for i in picks:
            k, pts = reqs[i]
            n, N = len(pts), h.N
            rv, rg, rH = o.eval_batch("vgh", pts)
            got_v = alone[offs[i][0]:offs[i][0] + n * N].view(np.float32).reshape(n, N)
            got_h = alone[offs[i][2]:offs[i][2] + 9 * n * N].view(np.float32).reshape(n, N, 3, 3)
            assert normwise(got_v, rv) < 1e-5 and normwise(got_h, rH) < 1e-5, (tix, i)
        del o
    for t in tabs:
        t.destroy()
    for name, budget in (("all resident", 0), ("4.3 GB budget", int(4.3e9))):
        srv = Server(hbm_budget=budget, slots=4, max_batch_points=16384, max_batch_requests=128)
        ids = [srv.register(h, st) for h, st in zip(hosts, storages)]
        mixed, lats = _serve(srv, ids, hosts, reqs, offs, total)
        st = srv.stats()
        say(f"serving {name}: p50 {np.percentile(lats, 50) * 1e3:.2f} ms, p99 "
            f"{np.percentile(lats, 99) * 1e3:.2f} ms, batches {st['batches']}, admissions "
            f"{st['admissions']}, evictions {st['evictions']}, resident {st['resident_bytes'] / 1e9:.3f} GB")
        srv.close()
        assert st["completed"] == len(reqs)
        if budget:
            assert st["resident_bytes"] <= budget
        else:
            assert st["evictions"] == 0 and st["admissions"] == 3
        bad = np.flatnonzero(mixed != alone)
        assert bad.size == 0, f"{name}: {bad.size} result words differ from the request evaluated alone"


# ---------------------------------------------------------------------------
# headline configuration at the benchmarked tile and order


def test_headline_at_benchmarked_tile_and_locality_order():
    rng = np.random.default_rng(2)
    M, N, P = 48, 256, 32768
    h = ea.HostSpline("s", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    for first in range(0, N, 64):
        h.set_block(first, plane_waves(rng, M, 64))
    pts = rng.random((P, 3))
    pos = ea.DeviceBuffer.from_array(pts)
    tab = ea.DeviceTable(h)
    base = eval_vgh(tab, pos, P)          # default tile, given order
    o = oracle.Table("s", (M, M, M), (0, 0, 0), (1, 1, 1), N)
    o.coefs[...] = h.coefs()
    idx = rng.choice(P, 512, replace=False)
    ref = o.eval_batch("vgh", pts[idx])
    configs = [("autotuned", None),
               ("headline tile: 8 B x 256, direct stores, given order", (8, 256, 2, ea.ORDER_GIVEN)),
               ("locality profile tile: 16 B x 64, staged streaming stores, locality",
                (16, 64, 0, ea.ORDER_LOCALITY)),
               ("16 B x 256, staged stores, locality", (16, 256, 1, ea.ORDER_LOCALITY))]
    for name, cfg in configs:
        if cfg is None:
            tab.autotune("vgh", P)
        else:
            tab.set_kernel(ea.KERNEL_VECTOR)
            tab.set_tile(cfg[0], cfg[1])
            tab.set_store_policy(cfg[2])
            tab.set_point_order(cfg[3])
        say(f"headline {name}: tuning {tab.tuning()}, kernel {tab.kernel}")
        got = eval_vgh(tab, pos, P)
        for a, r in zip(got, ref):
            assert normwise(a[idx], r) < 1e-5, name
        if tab.kernel == ea.KERNEL_VECTOR:
            for a, b in zip(got, base):
                np.testing.assert_array_equal(a, b)
