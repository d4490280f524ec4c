"""Measure the device-to-device copy bandwidth with the library's plain copy
kernel (einspline_hip_stream_copy).  Used for the achievable-bandwidth
ceiling and to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE readings: a copy
of B bytes moves B bytes each way.

    python -m einspline_amd.stream_copy [--mib 1024] [--reps 20]
"""
import argparse
import json

import einspline_amd as ea


def measure(mib=1024, reps=20):
    nbytes = mib << 20
    a, b = ea.DeviceBuffer(nbytes), ea.DeviceBuffer(nbytes)
    a.zero()
    s = ea.Stream()
    for _ in range(3):
        ea.check(ea.lib.einspline_hip_stream_copy(b.ptr, a.ptr, nbytes, s.ptr), "copy")
    s.synchronize()
    e0, e1 = ea.Event(), ea.Event()
    e0.record(s)
    for _ in range(reps):
        ea.check(ea.lib.einspline_hip_stream_copy(b.ptr, a.ptr, nbytes, s.ptr), "copy")
    e1.record(s)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    return {"bytes_per_copy": nbytes, "ms": ms, "GBs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    print(json.dumps(measure(a.mib, a.reps)))
