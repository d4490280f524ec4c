"""oracle -- the correctness oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package, and only as the checker (or the timed CPU baseline).
The product path (einspline_amd) never touches it.

* liboracle.so (oracle.c): C restatement of the reference's default-build
  create/set and std evaluation (src/multi_bspline_create.c,
  src/bspline_create.c, src/multi_bspline_eval_std_*_impl.h).
* liboracle.so also restates the Slater-determinant kernels that consume
  the orbital rows (src/determinant_update.h, src/cuda_walker.cu):
  update_inverse() and det_ratio() below.
* brute.py: numpy closed forms used to pin the restatement (textbook B3
  kernel sums, FFT periodic solve, analytic plane waves).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
lib = C.CDLL(LIB_PATH)

OR_PERIODIC, OR_DERIV1, OR_DERIV2, OR_FLAT, OR_NATURAL, OR_ANTIPERIODIC = range(6)
MODE = {"v": 0, "vg": 1, "vgl": 2, "vgh": 3, "vghgh": 4}


class otable(C.Structure):
    _fields_ = [("M", C.c_int * 3), ("Nn", C.c_int * 3),
                ("start", C.c_double * 3), ("end", C.c_double * 3),
                ("delta", C.c_double * 3), ("delta_inv", C.c_double * 3),
                ("lcode", C.c_int * 3), ("rcode", C.c_int * 3),
                ("lval", C.c_double * 6), ("rval", C.c_double * 6),
                ("N", C.c_int), ("ncomp", C.c_int),
                ("xs", C.c_long), ("ys", C.c_long), ("zs", C.c_long), ("dim", C.c_int)]


_vp, _i, _d = C.c_void_p, C.c_int, C.c_double


def _bind(l):
    """Declare the prototypes of an oracle library (this build's or a
    native build of the same oracle.c)."""
    l.oracle_table_init.restype = _i
    l.oracle_table_init.argtypes = [C.POINTER(otable), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i]
    for k in "sdcz":
        f = getattr(l, f"oracle_set_{k}")
        f.restype = None
        f.argtypes = [C.POINTER(otable), _vp, _i, _vp]
    for k in "sd":
        f = getattr(l, f"oracle_eval_{k}")
        f.restype = _i
        f.argtypes = [C.POINTER(otable), _vp, _i, _d, _d, _d, _vp, _vp, _vp, _vp]
    l.oracle_eval_batch.restype = _i
    l.oracle_eval_batch.argtypes = [C.POINTER(otable), _i, _vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _i]
    l.oracle_basis_d.restype = None
    l.oracle_basis_d.argtypes = [_d, _vp, _vp, _vp]
    l.oracle_basis3_d.restype = None
    l.oracle_basis3_d.argtypes = [_vp]
    l.oracle_max_threads.restype = _i
    for t, ct in (("f", C.c_float), ("d", C.c_double)):
        f = getattr(l, f"oracle_update_inverse_{t}")
        f.restype = None
        f.argtypes = [_vp, _vp, _vp, _i, C.c_long, _i]
        f = getattr(l, f"oracle_det_ratio_{t}")
        f.restype = ct
        f.argtypes = [_vp, _vp, _i, C.c_long, _i]
    return l


lib = _bind(lib)


def load_library(path):
    """Another build of oracle.c (for example the CPU baseline's native
    -O3 -march=native build), prototyped like the default one."""
    return _bind(C.CDLL(path))


def basis(t: float):
    a, da, d2a = (np.zeros(4) for _ in range(3))
    lib.oracle_basis_d(float(t), a.ctypes.data, da.ctypes.data, d2a.ctypes.data)
    return a, da, d2a


def basis3():
    d3a = np.zeros(4)
    lib.oracle_basis3_d(d3a.ctypes.data)
    return d3a


def max_threads() -> int:
    return lib.oracle_max_threads()


class Table:
    """Oracle spline table of kind s/d/c/z (numpy-owned coefficients), in
    len(M) = 1, 2 or 3 dimensions."""

    def __init__(self, kind, M, start, end, N, bc=None, library=None):
        self.lib = library or lib
        self.kind = kind
        self.dim = dim = len(M)
        M, start, end = (tuple(v) + (1,) * (3 - dim) for v in (M, start, end))
        self.ncomp = 2 if kind in "cz" else 1
        self.real = np.float32 if kind in "sc" else np.float64
        bc = list(bc or [(OR_PERIODIC, OR_PERIODIC)] * dim) + [(OR_PERIODIC, OR_PERIODIC)] * (3 - dim)
        lc = np.array([b[0] for b in bc], np.int32)
        rc = np.array([b[1] for b in bc], np.int32)
        lv = np.zeros(6)
        rv = np.zeros(6)
        for a, b in enumerate(bc):
            if len(b) > 2:
                lv[2 * a], lv[2 * a + 1] = np.real(b[2]), np.imag(b[2])
                rv[2 * a], rv[2 * a + 1] = np.real(b[3]), np.imag(b[3])
        self.t = otable()
        # keep every argument array alive across the call
        m_a = np.array(M, np.int32)
        s_a = np.array(start, float)
        e_a = np.array(end, float)
        rc_ = self.lib.oracle_table_init(C.byref(self.t), m_a.ctypes.data, s_a.ctypes.data,
                                    e_a.ctypes.data, lc.ctypes.data, rc.ctypes.data,
                                    lv.ctypes.data, rv.ctypes.data, int(N), self.ncomp, dim)
        assert rc_ == 0, "invalid oracle table"
        self.N = int(N)
        self.M = tuple(M[:dim])
        self.shape = tuple(self.t.Nn[:dim])
        self.coefs = np.zeros(self.shape + (self.N * self.ncomp,), self.real)

    @property
    def delta_inv(self):
        return np.array(self.t.delta_inv[:self.dim])

    def set(self, n, data):
        data = np.asarray(data)
        if self.ncomp == 2:
            cd = np.complex64 if self.kind == "c" else np.complex128
            data = np.ascontiguousarray(data, dtype=cd)
        else:
            data = np.ascontiguousarray(data, dtype=self.real)
        assert data.size == int(np.prod(self.M))
        getattr(self.lib, f"oracle_set_{self.kind}")(C.byref(self.t), self.coefs.ctypes.data, int(n),
                                                 data.ctypes.data)

    def eval_batch(self, mode, pos, nthreads=0):
        """pos: (P, dim) float64.  Returns (vals, grads, lapl/hess[, gradhess])
        shaped (P, N), (P, N, dim), (P, N) or (P, N, dim, dim) (1-D vgh: (P, N)),
        (P, N, 3, 3, 3); complex for c/z."""
        dim = self.dim
        pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, dim))
        P = pos.shape[0]
        NR = self.N * self.ncomp
        m = MODE[mode]
        assert mode != "vghgh" or dim == 3
        v = np.zeros((P, NR), self.real)
        g = np.zeros((P, dim * NR), self.real) if m >= 1 else None
        nh = {"vgl": 1, "vgh": 1 if dim == 1 else dim * dim, "vghgh": 9}.get(mode, 0)
        h = np.zeros((P, nh * NR), self.real) if nh else None
        gh = np.zeros((P, 27 * NR), self.real) if mode == "vghgh" else None
        ptr = lambda a: a.ctypes.data if a is not None else None
        bad = self.lib.oracle_eval_batch(C.byref(self.t), 1 if self.real == np.float64 else 0,
                                    self.coefs.ctypes.data, m, pos.ctypes.data, P, v.ctypes.data,
                                    ptr(g), ptr(h), ptr(gh), int(nthreads))
        self.bad = bad
        return shape_results(mode, self.N, self.ncomp, v, g, h, gh, dim)


def shape_results(mode, N, ncomp, v, g, h, gh=None, dim=3):
    """Reshape flat per-point result blocks (CPU layout) to (P, N, ...)."""
    P = v.shape[0]

    def cx(a):
        if ncomp == 2:
            return a.view(np.complex64 if a.dtype == np.float32 else np.complex128)
        return a

    out = [cx(v).reshape(P, N)]
    if g is not None:
        out.append(cx(g).reshape(P, N, dim))
    if h is not None:
        if mode == "vgl" or dim == 1:
            out.append(cx(h).reshape(P, N))
        else:
            out.append(cx(h).reshape(P, N, dim, dim))
    if gh is not None:
        out.append(cx(gh).reshape(P, N, 3, 3, 3))
    return tuple(out)


def _det_args(Ainv, *vecs):
    Ainv = np.asarray(Ainv)
    assert Ainv.ndim == 2 and Ainv.flags.c_contiguous and Ainv.dtype in (np.float32, np.float64)
    t = "f" if Ainv.dtype == np.float32 else "d"
    for v in vecs:
        assert v is None or (v.dtype == Ainv.dtype and v.flags.c_contiguous)
    return t, Ainv.shape[1]


def update_inverse(A, Ainv, u, k, N=None):
    """Sherman-Morrison update in place (src/determinant_update.h:8-104): row
    k of A (modified) becomes u, or, with A None, changes by u.  Matrices are
    row-major with row stride = their second extent; N defaults to it."""
    t, rs = _det_args(Ainv, A, u)
    N = rs if N is None else N
    getattr(lib, f"oracle_update_inverse_{t}")(A.ctypes.data if A is not None else None,
                                                Ainv.ctypes.data, u.ctypes.data, N, rs, k)


def det_ratio(Ainv, u, col, N=None):
    """sum_i Ainv[i][col] * u[i] (src/cuda_walker.cu:150-185)."""
    t, rs = _det_args(Ainv, u)
    N = rs if N is None else N
    return getattr(lib, f"oracle_det_ratio_{t}")(Ainv.ctypes.data, u.ctypes.data, N, rs, col)
