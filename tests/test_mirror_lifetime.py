"""The single-point API's device-mirror lifetime on the CPU: the host-only
bookkeeping of einspline_amd/csrc/mirror_cache.hpp, driven by a C++ stress
program (tests/cpp/mirror_lifetime.cpp) built with AddressSanitizer: six
threads evaluate while a seventh keeps marking the spline dirty, as
concurrent set_* calls do.  A mirror stays alive while a call holds it and
every replaced mirror is freed exactly once."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_mirror_survives_concurrent_set_and_is_freed_once(tmp_path):
    exe = tmp_path / "mirror_lifetime"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-pthread",
                    "-I", os.path.join(ROOT, "einspline_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "mirror_lifetime.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-3000:]
    assert "mirror lifetime ok" in out.stdout
