"""Measure the device-to-device copy bandwidth with the library's plain copy
kernel (einspline_hip_stream_copy), and the write-only bandwidth with its
fill kernel (einspline_hip_stream_fill).  Used for the achievable-bandwidth
ceilings and to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE readings: a copy
of B bytes moves B bytes each way.

    python -m einspline_amd.stream_copy [--mib 1024] [--reps 20] [--fill]
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


def measure_fill(mib=1024, reps=20, nontemporal=False):
    nbytes = mib << 20
    b = ea.DeviceBuffer(nbytes)
    s = ea.Stream()
    nt = 1 if nontemporal else 0
    for _ in range(3):
        ea.check(ea.lib.einspline_hip_stream_fill(b.ptr, 0, nbytes, nt, s.ptr), "fill")
    s.synchronize()
    e0, e1 = ea.Event(), ea.Event()
    e0.record(s)
    for _ in range(reps):
        ea.check(ea.lib.einspline_hip_stream_fill(b.ptr, 0, nbytes, nt, s.ptr), "fill")
    e1.record(s)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    return {"bytes_per_fill": nbytes, "nontemporal": bool(nontemporal), "ms": ms,
            "GBs_write": nbytes / (ms * 1e-3) / 1e9}


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--fill", action="store_true", help="also measure write-only bandwidth")
    a = ap.parse_args()
    out = measure(a.mib, a.reps)
    if a.fill:
        out["fill"] = measure_fill(a.mib, a.reps, False)
        out["fill_nontemporal"] = measure_fill(a.mib, a.reps, True)
    print(json.dumps(out))
