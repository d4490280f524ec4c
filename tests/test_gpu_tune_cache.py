"""The autotuner's cache file ($EINSPLINE_TUNE_CACHE): one line of thirteen
integers per (storage, type, mode, row bytes, sample, Nx, Ny, Nz) key --
lane bytes, block, store policy, point order, kernel -- written after a
tuning run and
read back by a later process, which then skips the timing (reported ms < 0)
and applies the stored choice; lines of any other shape (an older format)
are ignored.  Each step runs in its own process, since the cache is read
once per process."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path.insert(0, {root!r})
import numpy as np
import einspline_amd as ea
assert ea.device_count() >= 1
h = ea.HostSpline("d", (10, 10, 10), (0, 0, 0), (1, 1, 1), 64)
h.set_block(0, np.random.default_rng(3).uniform(-1, 1, (64, 10, 10, 10)))
t = ea.DeviceTable(h, storage=ea.STORE_BF16)
lb, blk, nt, ms = t.autotune("vgh", 2048)
print("RESULT", lb, blk, nt, ms, *t.tuning(), t.kernel, flush=True)
"""


def run(env_file):
    env = dict(os.environ, EINSPLINE_TUNE_CACHE=env_file)
    out = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT)], env=env, capture_output=True,
                         text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    line = [l for l in out.stdout.splitlines() if l.startswith("RESULT")][-1].split()[1:]
    return [float(x) for x in line]


def test_cache_written_and_reused(tmp_path):
    f = str(tmp_path / "tune.txt")
    with open(f, "w") as fh:
        fh.write("1 2 3 4 5 6 7 8 9\n")          # lines of the older nine- and ten-field shapes
        fh.write("2 0 3 128 2048 16 256 1 1 1\n")
        fh.write("garbage\n")
    first = run(f)
    assert first[3] > 0                          # timed
    lines = [l.split() for l in open(f).read().splitlines() if l.strip()]
    assert lines and all(len(l) == 13 for l in lines)
    lb, blk, nt, _, tlb, tblk, tnt, torder, kern = first
    key_line = lines[-1]
    assert [int(x) for x in key_line[5:8]] == [13, 13, 13]      # the table's extent (10 + 3 halo planes)
    assert [int(x) for x in key_line[8:]] == [int(tlb), int(tblk), int(tnt), int(torder), int(kern)]
    second = run(f)
    assert second[3] < 0                         # served from the cache
    assert second[4:] == first[4:]               # same tile, policy, order and kernel applied
