"""Serving runtime on the GPU: continuous batching of uneven requests for
several resident tables.  Each request's result must be identical to
evaluating it alone with einspline_eval_vgh_batch (bitwise: the same
per-point instruction sequence runs in both), and close to the oracle."""
import numpy as np
import pytest

import einspline_amd as ea
import oracle
from einspline_amd.serve import Server

pytestmark = pytest.mark.gpu


def tables(rng):
    """T0: s 12^3 x 64 F32 (one point per wavefront); T1: d 10^3 x 96 stored
    BF16; T2: s 9^3 x 200 F32 (several lanes groups per wavefront)."""
    out = []
    for kind, M, N in (("s", 12, 64), ("d", 10, 96), ("s", 9, 200)):
        h = ea.HostSpline(kind, (M, M, M), (0, 0, 0), (1, 1, 1), N)
        h.set_block(0, rng.uniform(-1, 1, (N, M, M, M)))
        out.append(h)
    return out


def alone(host, storage, pts):
    tab = ea.DeviceTable(host, storage=None if host.kind == "s" else storage)
    P = len(pts)
    pos = ea.DeviceBuffer.from_array(pts)
    outs = tab.alloc_outputs("vgh", P)
    tab.eval_batch("vgh", pos, P, outs)
    ea.synchronize()
    return tab.download("vgh", P, outs)


def run_mix(srv, ids, hosts, storages, rng, nreq=120, table=None):
    """Submit nreq requests (for random tables, or all for `table`), wait for
    all of them and check each against the same points evaluated alone."""
    reqs = []
    for r in range(nreq):
        k = int(rng.integers(0, 3)) if table is None else table
        n = int(rng.integers(1, 300))
        pts = rng.random((n, 3))
        N = hosts[k].N
        bufs = [ea.DeviceBuffer(n * N * c * 4) for c in (1, 3, 9)]
        t = srv.submit(ids[k], pts, *bufs)
        reqs.append((k, pts, bufs, t))
    lats = [srv.wait(t) for (_, _, _, t) in reqs]
    assert all(l > 0 for l in lats)
    for k, pts, bufs, _ in reqs:
        n, N = len(pts), hosts[k].N
        got = [b.download(np.float32, (n, c * N)) for b, c in zip(bufs, (1, 3, 9))]
        ref = alone(hosts[k], storages[k], pts)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
    return reqs


def test_mix_equals_alone_and_oracle(rng):
    hosts = tables(rng)
    storages = [ea.STORE_F32, ea.STORE_BF16, ea.STORE_F32]
    srv = Server(slots=3, max_batch_points=2048, max_batch_requests=32)
    ids = [srv.register(h, s) for h, s in zip(hosts, storages)]
    reqs = run_mix(srv, ids, hosts, storages, rng)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] >= 2
    assert st["resident_tables"] == 3 and st["evictions"] == 0
    # FP32 tables also match the oracle
    for k, pts, bufs, _ in reqs[:20]:
        if storages[k] != ea.STORE_F32:
            continue
        h = hosts[k]
        o = oracle.Table("s", h.M, (0, 0, 0), (1, 1, 1), h.N)
        o.coefs[...] = h.coefs()
        rv, rg, rh = o.eval_batch("vgh", pts)
        v = bufs[0].download(np.float32, rv.shape)
        e = np.max(np.abs(v - rv)) / np.max(np.abs(rv))
        assert e < 1e-5
    srv.close()


def test_eviction_under_budget(rng):
    hosts = tables(rng)
    storages = [ea.STORE_F32, ea.STORE_BF16, ea.STORE_F32]
    sizes = []
    for h, s in zip(hosts, storages):
        t = ea.DeviceTable(h, storage=None if h.kind == "s" else s)
        sizes.append(t.bytes)
        t.destroy()
    budget = sorted(sizes)[-1] + sorted(sizes)[-2]     # room for two of the three
    srv = Server(hbm_budget=budget, slots=2, max_batch_points=1024, max_batch_requests=8)
    ids = [srv.register(h, s) for h, s in zip(hosts, storages)]
    # Each table in turn, twice, every request waited for before the next
    # table's: the third table evicts one of the first two, and the second
    # pass finds at least one table absent again, whatever the victims.
    for _ in range(2):
        for k in range(3):
            run_mix(srv, ids, hosts, storages, rng, nreq=10, table=k)
            assert srv.stats()["resident_bytes"] <= budget
    st = srv.stats()
    assert st["evictions"] >= 2 and st["admissions"] >= 4
    assert st["completed"] == 60
    # then a mix of all three, queued together
    run_mix(srv, ids, hosts, storages, rng, nreq=60)
    assert srv.stats()["resident_bytes"] <= budget
    srv.close()


def test_budget_and_argument_errors(rng):
    hosts = tables(rng)
    srv = Server(hbm_budget=1 << 20, slots=1)
    with pytest.raises(ea.EinsplineError) as e:
        srv.register(hosts[2], ea.STORE_F32)          # larger than the budget
    assert e.value.code == ea.E_BUDGET
    with pytest.raises(ea.EinsplineError):
        srv.register(hosts[0], ea.STORE_BF16)         # BF16 needs a double table
    srv.close()
    srv = Server(slots=1, max_batch_points=64)
    tid = srv.register(hosts[0], ea.STORE_F32)
    b = ea.DeviceBuffer(1 << 20)
    with pytest.raises(ea.EinsplineError):
        srv.submit(tid, np.zeros((65, 3)), b, b, b)  # more points than a batch holds
    with pytest.raises(ea.EinsplineError):
        srv.submit(7, np.zeros((1, 3)), b, b, b)     # unknown table
    srv.close()


def test_load_generator_smoke(rng):
    hosts = tables(rng)
    storages = [ea.STORE_F32, ea.STORE_BF16, ea.STORE_F32]
    srv = Server(slots=4, max_batch_points=4096, max_batch_requests=64)
    ids = [srv.register(h, s) for h, s in zip(hosts, storages)]
    sinks = [[ea.DeviceBuffer(512 * h.N * c * 4) for c in (1, 3, 9)] for h in hosts]
    lat, wall = srv.run_load(ids, sinks, rate_hz=2000.0, duration_s=0.5, max_points=512)
    assert len(lat) > 100 and np.all(lat > 0)
    st = srv.stats()
    assert st["completed"] == st["submitted"]
    srv.close()


def test_concurrent_submitters(rng):
    """Several host threads submit and wait at once (ctypes drops the GIL)."""
    import threading

    hosts = tables(rng)
    storages = [ea.STORE_F32, ea.STORE_BF16, ea.STORE_F32]
    srv = Server(slots=4, max_batch_points=4096, max_batch_requests=64)
    ids = [srv.register(h, s) for h, s in zip(hosts, storages)]
    errors = []

    def worker(seed):
        r = np.random.default_rng(seed)
        try:
            for _ in range(25):
                k = int(r.integers(0, 3))
                n = int(r.integers(1, 200))
                pts = r.random((n, 3))
                N = hosts[k].N
                bufs = [ea.DeviceBuffer(n * N * c * 4) for c in (1, 3, 9)]
                srv.wait(srv.submit(ids[k], pts, *bufs))
                v = bufs[0].download(np.float32, (n, N))
                ref = alone(hosts[k], storages[k], pts)[0]
                np.testing.assert_array_equal(v, ref)
        except Exception as exc:      # surfaced below
            errors.append(exc)

    th = [threading.Thread(target=worker, args=(s,)) for s in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors[0]
    assert srv.stats()["completed"] == 100
    srv.close()
