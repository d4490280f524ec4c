"""Writes the known-answer fixtures in this folder from closed forms.

The reference ships no known-answer vectors, and its own code cannot be
built here, so these vectors come from the mathematics alone
(oracle/brute.py): dense interpolation maps solved with numpy, and the full
tensor-product B3 sum for values, gradients and Hessians.  They pin the
oracle (tests/test_oracle_pins.py) and, through it, the GPU path.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import brute  # noqa: E402


def dinv_of(M, start, end, bcs):
    n = [m + (3 if b[0] in (0, 5) else 2) for m, b in zip(M, bcs)]
    return np.array([1.0 / ((e - s) / (k - 3)) for s, e, k in zip(start, end, n)])


def case(name, kind, M, start, end, N, bcs, seed, P=24):
    rng = np.random.default_rng(seed)
    cplx = kind in "cz"
    data = rng.uniform(-1, 1, (N,) + M)
    if cplx:
        data = data + 1j * rng.uniform(-1, 1, (N,) + M)
    if kind in "sc":
        data = data.astype(np.complex64 if cplx else np.float32)
    dinv = dinv_of(M, start, end, bcs)
    lr = [(b[2], b[3]) if len(b) > 2 else (0.0, 0.0) for b in bcs]
    co = np.stack([brute.solve_dense(data[n].astype(np.complex128 if cplx else np.float64),
                                     [b[:2] for b in bcs], dinv, lr) for n in range(N)], -1)
    if kind in "sc":
        co = co.astype(np.complex64 if cplx else np.float32)
    s, e = np.array(start), np.array(end)
    pts = s + (e - s) * rng.random((P, 3)) * 0.999999
    v, g, H, _ = brute.eval_dense(co.astype(np.complex128 if cplx else np.float64), start, dinv, pts)
    if cplx:
        co_store = np.ascontiguousarray(co).view(np.float32 if kind == "c" else np.float64)
    else:
        co_store = co
    np.savez(os.path.join(HERE, name), kind=kind, M=np.array(M), start=np.array(start),
             end=np.array(end), N=N, bcs=np.array([b[:2] for b in bcs]), data=data,
             coefs=co_store, pts=pts, v=v, g=g, H=H)


if __name__ == "__main__":
    case("periodic_d.npz", "d", (5, 6, 4), (3.1, 8.7, 4.5), (9.1, 12.7, 9.3), 3, [(0, 0)] * 3, 1)
    case("natural_flat_d.npz", "d", (6, 5, 5), (0.0, -1.0, 2.0), (1.0, 1.0, 3.5), 2,
         [(4, 4), (3, 3), (3, 4)], 2)
    case("antiperiodic_d.npz", "d", (4, 5, 6), (0.0, 0.0, 0.0), (1.0, 2.0, 1.5), 2,
         [(5, 5), (0, 0), (5, 5)], 3)
    case("periodic_z.npz", "z", (5, 4, 6), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3, [(0, 0)] * 3, 4)
    case("periodic_s.npz", "s", (7, 6, 5), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 5, [(0, 0)] * 3, 5)
    case("periodic_c.npz", "c", (5, 5, 6), (-1.0, 0.0, 1.0), (1.0, 2.0, 2.0), 2, [(0, 0)] * 3, 6)
    print("written")
