"""Table files on the GPU: save -> load gives the same coefficient bytes,
the same tuning and bitwise the same results, for every kind, storage type
and dimension; a damaged file is refused; the file's hashes follow the
documented format (restated in test_table_io.py)."""
import struct

import numpy as np
import pytest

import einspline_amd as ea
from test_gpu_parity import make_pair, points
from test_table_io import HEADER, table_hash

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ea.device_count() >= 1, "no HIP device visible: GPU tests must run on a GPU"


def coef_bytes(tab):
    t = tab.ptr.contents
    out = np.empty(t.bytes, np.uint8)
    ea.check(ea.lib.einspline_hip_memcpy_dtoh(out.ctypes.data, t.coefs, t.bytes), "dtoh")
    return out


def eval_all(tab, pts, mode):
    pos = ea.DeviceBuffer.from_array(np.ascontiguousarray(pts))
    outs = tab.alloc_outputs(mode, len(pts))
    tab.eval_batch(mode, pos, len(pts), outs)
    ea.synchronize()
    return tab.download(mode, len(pts), outs)


CASES = [("s", None, 3), ("d", None, 3), ("c", None, 3), ("z", None, 3),
         ("d", ea.STORE_BF16, 3), ("d", ea.STORE_F16, 3), ("d", ea.STORE_F32, 3),
         ("s", None, 1), ("z", None, 2)]


@pytest.mark.parametrize("kind,storage,dim", CASES)
def test_round_trip_is_exact(kind, storage, dim, rng, tmp_path):
    M = (9, 8, 7)[:dim]
    h, _ = make_pair(kind, M, 40, rng, (0.0, 0.5, -1.0)[:dim], (1.0, 2.0, 0.0)[:dim],
                     [(ea.PERIODIC, ea.PERIODIC), (ea.NATURAL, ea.FLAT), (ea.PERIODIC, ea.PERIODIC)][:dim])
    tab = ea.DeviceTable(h, storage=storage) if storage is not None else ea.DeviceTable(h)
    if dim == 3:
        tab.set_tile(8 if kind in "sc" or storage not in (None, ea.STORE_F64) else 16, 128)
        tab.set_store_policy(2)
        tab.set_point_order(ea.ORDER_LOCALITY)
    path = tmp_path / "t.tbl"
    tab.save(path)
    raw = path.read_bytes()
    assert len(raw) == HEADER + tab.bytes
    # the documented hashes
    cb, ch, hh = struct.unpack_from("<QQQ", raw, 184)
    assert cb == tab.bytes and ch == table_hash(raw[HEADER:]) and hh == table_hash(raw[:200])
    back = ea.DeviceTable.load(path)
    assert back.tuning() == tab.tuning()
    assert (back.dim, back.N, back.storage, back.tcode) == (tab.dim, tab.N, tab.storage, tab.tcode)
    np.testing.assert_array_equal(coef_bytes(back), coef_bytes(tab))
    lo, hi = (0.0, 0.5, -1.0)[:dim], (1.0, 2.0, 0.0)[:dim]
    pts = np.asarray(lo) + (np.asarray(hi) - np.asarray(lo)) * rng.random((1500, dim)) * (1 - 1e-9)
    if dim == 3:
        pts = points(rng, 1500, lo, hi)
    for a, b in zip(eval_all(back, pts, "vgh"), eval_all(tab, pts, "vgh")):
        np.testing.assert_array_equal(a, b)


def test_damaged_files_are_refused(rng, tmp_path):
    h, _ = make_pair("s", (8, 8, 8), 16, rng)
    tab = ea.DeviceTable(h)
    path = tmp_path / "t.tbl"
    tab.save(path)
    raw = bytearray(path.read_bytes())
    raw[HEADER + 1000] ^= 0x10                          # one bit in the coefficients
    bad = tmp_path / "bad.tbl"
    bad.write_bytes(bytes(raw))
    with pytest.raises(ea.EinsplineError) as e:
        ea.DeviceTable.load(bad)
    assert "hash" in str(e.value)
    bad.write_bytes(bytes(raw[:-1]))
    with pytest.raises(ea.EinsplineError):
        ea.DeviceTable.load(bad)
    # the intact file still loads, and saving over it replaces it whole
    ea.DeviceTable.load(path)
    tab.save(path)
    assert not (tmp_path / "t.tbl.tmp").exists()
