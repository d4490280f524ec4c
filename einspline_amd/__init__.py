"""einspline_amd -- ctypes front end of libeinspline.so.

The library itself is C (create/set, src/multi_bspline_create.c) and HIP for
gfx950 (evaluation); this module only declares its C ABI for Python callers
(tests, benchmarks, serving drivers) and wraps device buffers and streams.
It does not compute anything itself: if the shared library is missing,
importing this module raises.

The C declarations mirror include/*.h; see those headers for semantics.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libeinspline.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C einspline_amd/csrc` "
        "(or __graft_entry__.build())"
    )


class _PrototypedLib(C.CDLL):
    """CDLL whose functions refuse to run until their prototype is declared.
    Without argtypes ctypes passes a Python int as a C int, masked to 32
    bits, which silently truncates 64-bit device addresses."""

    def __init__(self, path):
        super().__init__(path)
        base = self._FuncPtr

        class _Fn(base):
            _flags_ = base._flags_
            _restype_ = base._restype_

            def __call__(self, *args):
                if self.argtypes is None:
                    name = getattr(self, "__name__", "library function")
                    raise TypeError(f"{name}: called before its ctypes prototype was declared")
                return base.__call__(self, *args)

        self._FuncPtr = _Fn


lib = _PrototypedLib(LIB_PATH)

# ---------------------------------------------------------------------------
# types (include/bspline_base.h, include/multi_bspline_structs.h)

PERIODIC, DERIV1, DERIV2, FLAT, NATURAL, ANTIPERIODIC = range(6)
SINGLE_REAL, DOUBLE_REAL, SINGLE_COMPLEX, DOUBLE_COMPLEX = range(4)
MULTI_U1D, MULTI_U2D, MULTI_U3D = 6, 7, 8
STORE_F32, STORE_F64, STORE_BF16, STORE_F16 = range(4)
MODES = ("v", "vg", "vgl", "vgh")          # autotuner modes 0..3; "vghgh" is batch-only
NOALLOC_MAX_POINTS = 1 << 20               # EINSPLINE_NOALLOC_MAX_POINTS
ORDER_GIVEN, ORDER_LOCALITY, ORDER_SLABS = 0, 1, 2   # ORDER_SLABS is a flag added to either order
KERNEL_VECTOR, KERNEL_MATRIX = 0, 1

E_OK, E_INVAL, E_NOMEM, E_HIP, E_RANGE, E_BUDGET, E_BUSY = 0, -1, -2, -3, -4, -5, -6


class Ugrid(C.Structure):
    _fields_ = [("start", C.c_double), ("end", C.c_double), ("num", C.c_int),
                ("delta", C.c_double), ("delta_inv", C.c_double)]


class BCtype_s(C.Structure):
    _fields_ = [("lCode", C.c_int), ("rCode", C.c_int), ("lVal", C.c_float), ("rVal", C.c_float)]


class BCtype_d(C.Structure):
    _fields_ = [("lCode", C.c_int), ("rCode", C.c_int), ("lVal", C.c_double), ("rVal", C.c_double)]


class BCtype_c(C.Structure):
    _fields_ = [("lCode", C.c_int), ("rCode", C.c_int), ("lVal_r", C.c_float), ("lVal_i", C.c_float),
                ("rVal_r", C.c_float), ("rVal_i", C.c_float)]


class BCtype_z(C.Structure):
    _fields_ = [("lCode", C.c_int), ("rCode", C.c_int), ("lVal_r", C.c_double), ("lVal_i", C.c_double),
                ("rVal_r", C.c_double), ("rVal_i", C.c_double)]


def _multi(bc, lapl3):
    f = [("spcode", C.c_int), ("tcode", C.c_int), ("coefs", C.c_void_p),
         ("x_stride", C.c_ssize_t), ("y_stride", C.c_ssize_t), ("z_stride", C.c_ssize_t),
         ("x_grid", Ugrid), ("y_grid", Ugrid), ("z_grid", Ugrid),
         ("xBC", bc), ("yBC", bc), ("zBC", bc), ("num_splines", C.c_int)]
    if lapl3:
        f.append(("lapl3", C.c_void_p))
    return f


class multi_UBspline_3d_s(C.Structure):
    _fields_ = _multi(BCtype_s, False)


class multi_UBspline_3d_d(C.Structure):
    _fields_ = _multi(BCtype_d, False)


class multi_UBspline_3d_c(C.Structure):
    _fields_ = _multi(BCtype_c, True)


class multi_UBspline_3d_z(C.Structure):
    _fields_ = _multi(BCtype_z, True)


def _multi12(bc, dim, lapl2):
    axes = "xy"[:dim]
    f = [("spcode", C.c_int), ("tcode", C.c_int), ("coefs", C.c_void_p)]
    f += [(a + "_stride", C.c_ssize_t) for a in axes]
    f += [(a + "_grid", Ugrid) for a in axes]
    f += [(a + "BC", bc) for a in axes]
    f += [("num_splines", C.c_int)]
    if lapl2:
        f.append(("lapl2", C.c_void_p))
    return f


# 1-D and 2-D descriptors, keyed (dim, kind)
STRUCTS12 = {}
for _d in (1, 2):
    for _k, _bc in (("s", BCtype_s), ("d", BCtype_d), ("c", BCtype_c), ("z", BCtype_z)):
        STRUCTS12[(_d, _k)] = type(f"multi_UBspline_{_d}d_{_k}", (C.Structure,),
                                   {"_fields_": _multi12(_bc, _d, _d == 2 and _k in "cz")})


class multi_UBspline_3d_hip(C.Structure):
    _fields_ = [("coefs", C.c_void_p),
                ("x_stride", C.c_ssize_t), ("y_stride", C.c_ssize_t), ("z_stride", C.c_ssize_t),
                ("x_grid", Ugrid), ("y_grid", Ugrid), ("z_grid", Ugrid),
                ("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
                ("num_splines", C.c_int), ("row_reals", C.c_int),
                ("tcode", C.c_int), ("storage", C.c_int),
                ("bytes", C.c_size_t), ("device", C.c_int), ("dim", C.c_int)]


KINDS = {
    # kind: (struct, BC type, real ctype, numpy real dtype, ncomp)
    "s": (multi_UBspline_3d_s, BCtype_s, C.c_float, np.float32, 1),
    "d": (multi_UBspline_3d_d, BCtype_d, C.c_double, np.float64, 1),
    "c": (multi_UBspline_3d_c, BCtype_c, C.c_float, np.float32, 2),
    "z": (multi_UBspline_3d_z, BCtype_z, C.c_double, np.float64, 2),
}

# ---------------------------------------------------------------------------
# prototypes

_vp, _i, _d, _sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t


def _proto(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


for k, (S, BC, RT, _, _) in KINDS.items():
    _proto(f"create_multi_UBspline_3d_{k}", C.POINTER(S), Ugrid, Ugrid, Ugrid, BC, BC, BC, _i)
    _proto(f"set_multi_UBspline_3d_{k}", None, C.POINTER(S), _i, _vp)
    _proto(f"set_multi_UBspline_3d_{k}_block", None, C.POINTER(S), _i, _i, _vp)
    _proto(f"eval_multi_UBspline_3d_{k}", None, C.POINTER(S), _d, _d, _d, _vp)
    _proto(f"eval_multi_UBspline_3d_{k}_vg", None, C.POINTER(S), _d, _d, _d, _vp, _vp)
    _proto(f"eval_multi_UBspline_3d_{k}_vgl", None, C.POINTER(S), _d, _d, _d, _vp, _vp, _vp)
    _proto(f"eval_multi_UBspline_3d_{k}_vgh", None, C.POINTER(S), _d, _d, _d, _vp, _vp, _vp)
    _proto(f"eval_multi_UBspline_3d_{k}_vghgh", None, C.POINTER(S), _d, _d, _d, _vp, _vp, _vp, _vp)
    _proto(f"create_multi_UBspline_3d_{k}_hip", C.POINTER(multi_UBspline_3d_hip), C.POINTER(S))
    for d in (1, 2):
        S12 = STRUCTS12[(d, k)]
        _proto(f"create_multi_UBspline_{d}d_{k}", C.POINTER(S12), *([Ugrid] * d), *([BC] * d), _i)
        _proto(f"set_multi_UBspline_{d}d_{k}", None, C.POINTER(S12), _i, _vp)
        _proto(f"create_multi_UBspline_{d}d_{k}_hip", C.POINTER(multi_UBspline_3d_hip), C.POINTER(S12))
        for m, nout in (("", 1), ("_vg", 2), ("_vgl", 3), ("_vgh", 3)):
            _proto(f"eval_multi_UBspline_{d}d_{k}{m}", None, C.POINTER(S12), *([_d] * d), *([_vp] * nout))
for d in (1, 2):
    _proto(f"create_multi_UBspline_{d}d_s_hip_conv", C.POINTER(multi_UBspline_3d_hip),
           C.POINTER(STRUCTS12[(d, "d")]))
    _proto(f"create_multi_UBspline_{d}d_c_hip_conv", C.POINTER(multi_UBspline_3d_hip),
           C.POINTER(STRUCTS12[(d, "z")]))
_proto("set_multi_UBspline_1d_d_BC", None, C.POINTER(STRUCTS12[(1, "d")]), _i, _vp, BCtype_d)
_proto("set_multi_UBspline_1d_z_BC", None, C.POINTER(STRUCTS12[(1, "z")]), _i, _vp, BCtype_z)
for k in "sdcz":
    _proto(f"eval_multi_multi_UBspline_1d_{k}_hip", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _vp, _i)
_proto("eval_multi_multi_UBspline_1d_s_vgl_hip", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _vp, _vp, _vp, _i)
for k in "dcz":
    _proto(f"eval_multi_multi_UBspline_1d_{k}_vgl_hip", _i, C.POINTER(multi_UBspline_3d_hip),
           _vp, _vp, _vp, _i, _i)
for k in "sdcz":
    _proto(f"eval_multi_multi_UBspline_3d_{k}_hip", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _vp, _i)
    _proto(f"eval_multi_multi_UBspline_3d_{k}_vgh_hip", _i, C.POINTER(multi_UBspline_3d_hip),
           _vp, _vp, _vp, _vp, _i)
    _proto(f"eval_multi_multi_UBspline_3d_{k}_vgl_hip", _i, C.POINTER(multi_UBspline_3d_hip),
           _vp, _vp, _vp, _vp, _i, _i)
_proto("eval_multi_multi_UBspline_3d_s_vgl_sign_hip", _i, C.POINTER(multi_UBspline_3d_hip),
       _vp, _vp, _vp, _vp, _vp, _i, _i)
_proto("eval_multi_multi_UBspline_3d_s_sign_hip", _i, C.POINTER(multi_UBspline_3d_hip),
       _vp, _vp, _vp, _i)
_proto("create_multi_UBspline_3d_s_hip_conv", C.POINTER(multi_UBspline_3d_hip), C.POINTER(multi_UBspline_3d_d))
_proto("create_multi_UBspline_3d_c_hip_conv", C.POINTER(multi_UBspline_3d_hip), C.POINTER(multi_UBspline_3d_z))
_proto("create_multi_UBspline_3d_s_hip_store", C.POINTER(multi_UBspline_3d_hip),
       C.POINTER(multi_UBspline_3d_d), _i)
_proto("destroy_multi_UBspline_3d_hip", None, C.POINTER(multi_UBspline_3d_hip))
_proto("destroy_Bspline", None, _vp)
_proto("einspline_hip_invalidate", None, _vp)
_proto("einspline_set_tile", _i, C.POINTER(multi_UBspline_3d_hip), _i, _i)
_proto("einspline_autotune", _i, C.POINTER(multi_UBspline_3d_hip), _i, _i, C.POINTER(_i),
       C.POINTER(_i), C.POINTER(_i), C.POINTER(C.c_float))
_proto("einspline_autotune_points", _i, C.POINTER(multi_UBspline_3d_hip), _i, C.POINTER(C.c_double), _i,
       C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(C.c_float))
_proto("einspline_set_store_policy", _i, C.POINTER(multi_UBspline_3d_hip), _i)
_proto("einspline_set_point_order", _i, C.POINTER(multi_UBspline_3d_hip), _i)
_proto("einspline_set_kernel", _i, C.POINTER(multi_UBspline_3d_hip), _i)
_proto("einspline_set_point_service", _i, _i, _i)
_proto("einspline_point_service_resident", _i)
_proto("einspline_get_kernel", _i, C.POINTER(multi_UBspline_3d_hip))
_proto("einspline_get_tuning", _i, C.POINTER(multi_UBspline_3d_hip), C.POINTER(_i), C.POINTER(_i),
       C.POINTER(_i), C.POINTER(_i))
_proto("einspline_eval_v_batch", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _i, _vp, _vp)
_proto("einspline_eval_vg_batch", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _i, _vp, _vp, _vp)
_proto("einspline_eval_vgl_batch", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _i, _vp, _vp, _vp, _vp)
_proto("einspline_eval_vgh_batch", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _i, _vp, _vp, _vp, _i, _vp)
_proto("einspline_eval_vgh_batch_noalloc", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _i, _vp, _vp, _vp, _i,
       _vp)
_proto("einspline_eval_vghgh_batch", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _i, _vp, _vp, _vp, _vp, _vp)
_proto("einspline_hip_device_count", _i)
_proto("einspline_hip_last_error", C.c_char_p)
_proto("einspline_hip_malloc", _vp, _sz)
_proto("einspline_hip_free", _i, _vp)
_proto("einspline_hip_memcpy_htod", _i, _vp, _vp, _sz)
_proto("einspline_hip_memcpy_dtoh", _i, _vp, _vp, _sz)
_proto("einspline_hip_memset", _i, _vp, _i, _sz)
_proto("einspline_hip_synchronize", _i)
_proto("einspline_hip_stream_create", _vp)
_proto("einspline_hip_stream_destroy", _i, _vp)
_proto("einspline_hip_stream_synchronize", _i, _vp)
_proto("einspline_hip_event_create", _vp)
_proto("einspline_hip_event_destroy", _i, _vp)
_proto("einspline_hip_event_record", _i, _vp, _vp)
_proto("einspline_hip_event_synchronize", _i, _vp)
_proto("einspline_hip_event_elapsed_ms", C.c_float, _vp, _vp)
_proto("einspline_hip_stream_copy", _i, _vp, _vp, _sz, _vp)
_proto("einspline_hip_stream_fill", _i, _vp, C.c_uint, _sz, _i, _vp)
_proto("einspline_table_save", _i, C.POINTER(multi_UBspline_3d_hip), C.c_char_p)
_proto("einspline_build_table_hip", C.POINTER(multi_UBspline_3d_hip), _vp, _vp, _i, _vp)


# ---------------------------------------------------------------------------
# non-uniform grids and 1-D non-uniform multi-splines (include/nugrid.h,
# include/nubasis.h, include/multi_nubspline.h)

GENERAL_GRID, CENTER_GRID, LOG_GRID = 1, 2, 3


class NUgrid(C.Structure):
    _fields_ = [("code", C.c_int), ("start", C.c_double), ("end", C.c_double),
                ("points", C.POINTER(C.c_double)), ("num_points", C.c_int), ("reverse_map", C.c_void_p)]


class NUBasis(C.Structure):
    _fields_ = [("grid", C.POINTER(NUgrid)), ("xVals", C.POINTER(C.c_double)),
                ("dxInv", C.POINTER(C.c_double)), ("periodic", C.c_bool)]


def _nub(bc):
    return [("spcode", C.c_int), ("tcode", C.c_int), ("coefs", C.c_void_p), ("x_stride", C.c_ssize_t),
            ("xBC", bc), ("num_splines", C.c_int), ("x_grid", C.POINTER(NUgrid)),
            ("x_basis", C.POINTER(NUBasis))]


class multi_NUBspline_1d_d(C.Structure):
    _fields_ = _nub(BCtype_d)


class multi_NUBspline_1d_z(C.Structure):
    _fields_ = _nub(BCtype_z)


_GP, _BP = C.POINTER(NUgrid), C.POINTER(NUBasis)
_proto("create_general_grid", _GP, _vp, _i)
_proto("create_center_grid", _GP, _d, _d, _d, _i)
_proto("create_log_grid", _GP, _d, _d, _i)
_proto("destroy_grid", None, _GP)
_proto("create_NUBasis", _BP, _GP, C.c_bool)
_proto("destroy_NUBasis", None, _BP)
for _t in "sd":
    _proto(f"get_NUBasis_funcs_{_t}", _i, _BP, _d, _vp)
    _proto(f"get_NUBasis_funcs_{_t}i", None, _BP, _i, _vp)
    _proto(f"get_NUBasis_dfuncs_{_t}", _i, _BP, _d, _vp, _vp)
    _proto(f"get_NUBasis_dfuncs_{_t}i", None, _BP, _i, _vp, _vp)
    _proto(f"get_NUBasis_d2funcs_{_t}", _i, _BP, _d, _vp, _vp, _vp)
    _proto(f"get_NUBasis_d2funcs_{_t}i", None, _BP, _i, _vp, _vp, _vp)
_proto("create_multi_NUBspline_1d_d", C.POINTER(multi_NUBspline_1d_d), _GP, BCtype_d, _i)
_proto("set_multi_NUBspline_1d_d", None, C.POINTER(multi_NUBspline_1d_d), _i, _vp)
_proto("set_multi_NUBspline_1d_d_BC", None, C.POINTER(multi_NUBspline_1d_d), _i, _vp, BCtype_d)
_proto("create_multi_NUBspline_1d_z", C.POINTER(multi_NUBspline_1d_z), _GP, BCtype_z, _i)
_proto("set_multi_NUBspline_1d_z", None, C.POINTER(multi_NUBspline_1d_z), _i, _vp)
_proto("set_multi_NUBspline_1d_z_BC", None, C.POINTER(multi_NUBspline_1d_z), _i, _vp, BCtype_z)
_proto("destroy_multi_NUBspline", None, _vp)
_proto("eval_multi_NUBspline_1d_z", None, C.POINTER(multi_NUBspline_1d_z), _d, _vp)
_proto("eval_multi_NUBspline_1d_z_vg", None, C.POINTER(multi_NUBspline_1d_z), _d, _vp, _vp)
_proto("eval_multi_NUBspline_1d_z_vgl", None, C.POINTER(multi_NUBspline_1d_z), _d, _vp, _vp, _vp)
_proto("eval_multi_NUBspline_1d_z_vgh", None, C.POINTER(multi_NUBspline_1d_z), _d, _vp, _vp, _vp)
_proto("create_multi_NUBspline_1d_z_hip", _vp, C.POINTER(multi_NUBspline_1d_z))
_proto("destroy_multi_NUBspline_1d_hip", None, _vp)
_proto("einspline_eval_nub_1d_z_batch", _i, _vp, _vp, _i, _vp, _vp, _vp, _i, _vp)
_proto("einspline_table_load", C.POINTER(multi_UBspline_3d_hip), C.c_char_p)
# Slater-determinant kernels (include/determinant_update_hip.h)
for _k in "sd":
    _proto(f"einspline_update_inverse_{_k}", _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp)
    _proto(f"einspline_calc_ratios_{_k}", _i, _vp, _vp, _vp, _i, _i, _i, _i, _vp)
    _proto(f"einspline_fill_ratios_{_k}", _i, C.POINTER(multi_UBspline_3d_hip), _vp, _vp, _vp, _vp,
           _i, _i, _i, _i, _vp)


class EinsplineError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed ({code}): {lib.einspline_hip_last_error().decode()}")
        self.code = code


def check(rc, what="einspline call"):
    if rc != 0:
        raise EinsplineError(rc, what)
    return rc


def device_count() -> int:
    return lib.einspline_hip_device_count()


# ---------------------------------------------------------------------------
# device memory, streams, events


class DeviceBuffer:
    """Raw device allocation owned by Python."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib.einspline_hip_malloc(self.nbytes)
        if not self.ptr:
            raise EinsplineError(E_NOMEM, f"einspline_hip_malloc({self.nbytes})")

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib.einspline_hip_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes), "memcpy H2D")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib.einspline_hip_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes), "memcpy D2H")
        return out

    def zero(self):
        check(lib.einspline_hip_memset(self.ptr, 0, self.nbytes), "memset")

    def free(self):
        if self.ptr:
            lib.einspline_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.ptr = lib.einspline_hip_stream_create()
        if not self.ptr:
            raise EinsplineError(E_HIP, "stream create")

    def synchronize(self):
        check(lib.einspline_hip_stream_synchronize(self.ptr), "stream synchronize")

    def __del__(self):
        try:
            if self.ptr:
                lib.einspline_hip_stream_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.ptr = lib.einspline_hip_event_create()
        if not self.ptr:
            raise EinsplineError(E_HIP, "event create")

    def record(self, stream=None):
        check(lib.einspline_hip_event_record(self.ptr, stream.ptr if stream else None), "event record")

    def synchronize(self):
        check(lib.einspline_hip_event_synchronize(self.ptr), "event synchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = lib.einspline_hip_event_elapsed_ms(self.ptr, end.ptr)
        if ms < 0:
            raise EinsplineError(E_HIP, "event elapsed")
        return ms

    def __del__(self):
        try:
            if self.ptr:
                lib.einspline_hip_event_destroy(self.ptr)
        except Exception:
            pass


def synchronize():
    check(lib.einspline_hip_synchronize(), "device synchronize")


def set_point_service(enable=True, idle_us=0):
    """Serve single-point calls on 3-D tables with the resident polling
    kernel (enable) or with one launch per call; idle_us > 0 sets how long
    the kernel stays without calls.  Resident kernels leave first."""
    check(lib.einspline_set_point_service(1 if enable else 0, int(idle_us)), "set_point_service")


def point_service_resident() -> int:
    """Devices on which the single-point kernel is resident now."""
    return int(lib.einspline_point_service_resident())


# ---------------------------------------------------------------------------
# host splines (create/set run on the host; single-point evals on the GPU)


def _bc(kind, lcode, rcode, lval=0.0, rval=0.0):
    BC = KINDS[kind][1]
    if kind in "sd":
        return BC(lcode, rcode, float(np.real(lval)), float(np.real(rval)))
    return BC(lcode, rcode, float(np.real(lval)), float(np.imag(lval)),
              float(np.real(rval)), float(np.imag(rval)))


class HostSpline:
    """A multi_UBspline_{1,2,3}d_{s,d,c,z} created by libeinspline.

    M, start, end: per-axis grid (sequences of 1, 2 or 3; their length is
    the dimension); bc: per-axis (lcode, rcode[, lval, rval]) tuples,
    default periodic.
    """

    def __init__(self, kind, M, start, end, num_splines, bc=None):
        S, BC, RT, dt, nc = KINDS[kind]
        self.kind, self.dtype, self.ncomp = kind, dt, nc
        self.dim = dim = len(M)
        self.N = int(num_splines)
        bc = bc or [(PERIODIC, PERIODIC)] * dim
        grids = [Ugrid(float(s), float(e), int(m), 0.0, 0.0) for m, s, e in zip(M, start, end)]
        bcs = [_bc(kind, *b) for b in bc]
        self.ptr = getattr(lib, f"create_multi_UBspline_{dim}d_{kind}")(*grids, *bcs, self.N)
        if not self.ptr:
            raise EinsplineError(E_INVAL, f"create_multi_UBspline_{dim}d_{kind}")
        st = self.ptr.contents
        self.M = tuple(int(m) for m in M)
        self.shape = tuple(int(m) + (3 if b[0] in (PERIODIC, ANTIPERIODIC) else 2) for m, b in zip(M, bc))

    @property
    def struct(self):
        return self.ptr.contents

    def coefs(self) -> np.ndarray:
        """Zero-copy view of the host table, shape (Nx, Ny, Nz, N[, 2])."""
        n = int(np.prod(self.shape)) * self.N * self.ncomp
        RT = KINDS[self.kind][2]
        buf = (RT * n).from_address(self.struct.coefs)
        a = np.ctypeslib.as_array(buf)
        shp = self.shape + (self.N,) + ((2,) if self.ncomp == 2 else ())
        return a.reshape(shp)

    def set(self, n, data):
        data = self._data(data, 1)
        getattr(lib, f"set_multi_UBspline_{self.dim}d_{self.kind}")(self.ptr, int(n), data.ctypes.data)

    def set_block(self, first, data):
        """data: (count, M_0, .., M_{dim-1}) real or complex array."""
        count = data.shape[0]
        if self.dim != 3:
            for o in range(count):
                self.set(first + o, data[o])
            return
        data = self._data(data, count)
        getattr(lib, f"set_multi_UBspline_3d_{self.kind}_block")(self.ptr, int(first), int(count),
                                                                 data.ctypes.data)

    def _data(self, data, count):
        data = np.asarray(data)
        if self.ncomp == 2:
            cdt = np.complex64 if self.kind == "c" else np.complex128
            data = np.ascontiguousarray(data, dtype=cdt)
        else:
            data = np.ascontiguousarray(data, dtype=self.dtype)
        assert data.size == count * int(np.prod(self.M)), "data must hold Mx*My*Mz values per orbital"
        return data

    def invalidate(self):
        lib.einspline_hip_invalidate(self.ptr)

    def eval(self, mode, *x):
        """Single-point evaluation through eval_multi_UBspline_<dim>d_<kind>[_mode];
        x holds dim coordinates."""
        N, D = self.N, self.dim
        assert len(x) == D, f"{D} coordinates"
        cdt = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}[self.kind]
        nh = {3: 9, 2: 4, 1: 1}[D]
        v = np.zeros(N, cdt)
        g = np.zeros(D * N, cdt)
        h = np.zeros(nh * N if mode in ("vgh", "vghgh") else N, cdt)
        t3 = np.zeros(27 * N, cdt)
        suf = "" if mode == "v" else "_" + mode
        f = getattr(lib, f"eval_multi_UBspline_{D}d_{self.kind}{suf}")
        args = [v.ctypes.data] + ([g.ctypes.data] if mode != "v" else []) + \
               ([h.ctypes.data] if mode in ("vgl", "vgh", "vghgh") else []) + \
               ([t3.ctypes.data] if mode == "vghgh" else [])
        f(self.ptr, *[float(c) for c in x], *args)
        out = [v]
        if mode != "v":
            out.append(g.reshape(N, D))
        if mode == "vgl" or (mode == "vgh" and D == 1):
            out.append(h)
        elif mode in ("vgh", "vghgh"):
            out.append(h.reshape(N, D, D))
        if mode == "vghgh":
            out.append(t3.reshape(N, 3, 3, 3))
        return tuple(out)

    def destroy(self):
        if self.ptr:
            lib.destroy_Bspline(C.cast(self.ptr, C.c_void_p))
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# device tables and batched evaluation


class DeviceTable:
    """A multi_UBspline_3d_hip built from a HostSpline.

    storage: None (the host type), or STORE_F32/BF16/F16 from a 'd' spline
    (precision form), or 'conv' to convert d->s / z->c.
    """

    def __init__(self, host: HostSpline, storage=None):
        k, D = host.kind, host.dim
        if storage is None:
            p = getattr(lib, f"create_multi_UBspline_{D}d_{k}_hip")(host.ptr)
        elif storage == "conv":
            p = getattr(lib, {"d": f"create_multi_UBspline_{D}d_s_hip_conv",
                              "z": f"create_multi_UBspline_{D}d_c_hip_conv"}[k])(host.ptr)
        else:
            assert k == "d", "storage conversion starts from a double table"
            p = lib.create_multi_UBspline_3d_s_hip_store(host.ptr, int(storage))
        if not p:
            raise EinsplineError(E_INVAL, "create device table")
        self._adopt(p)

    @classmethod
    def build(cls, host: "HostSpline", data, storage=None, stream=None):
        """Solve the table on the GPU (einspline_build_table_hip) from samples
        shaped like set_block's, without a host solve; storage as for
        einspline_build_table_hip (None: the spline's own type)."""
        arr = host._data(data, host.N)
        buf = DeviceBuffer.from_array(arr)
        p = lib.einspline_build_table_hip(host.ptr, buf.ptr, -1 if storage is None else int(storage),
                                          stream.ptr if stream else None)
        buf.free()
        if not p:
            raise EinsplineError(E_INVAL, "einspline_build_table_hip")
        self = cls.__new__(cls)
        self._adopt(p)
        return self

    @classmethod
    def load(cls, path):
        """A table saved by save() (einspline_table_load): no solve, no conversion."""
        p = lib.einspline_table_load(os.fsencode(path))
        if not p:
            raise EinsplineError(E_INVAL, f"einspline_table_load({path})")
        self = cls.__new__(cls)
        self._adopt(p)
        return self

    def save(self, path):
        check(lib.einspline_table_save(self.ptr, os.fsencode(path)), f"einspline_table_save({path})")

    def _adopt(self, p):
        self.ptr = p
        t = p.contents
        self.N = t.num_splines
        self.tcode = t.tcode
        self.storage = t.storage
        self.complex = t.tcode in (SINGLE_COMPLEX, DOUBLE_COMPLEX)
        self.result_dtype = np.float64 if t.storage == STORE_F64 else np.float32
        self.nreal = self.N * (2 if self.complex else 1)
        self.bytes = t.bytes
        self.dim = t.dim

    def set_tile(self, lane_bytes=0, block=0):
        check(lib.einspline_set_tile(self.ptr, int(lane_bytes), int(block)), "einspline_set_tile")

    def set_store_policy(self, policy=0):
        """0 staged + streaming (default), 1 staged + ordinary, 2 direct."""
        check(lib.einspline_set_store_policy(self.ptr, int(policy)), "set_store_policy")

    def set_point_order(self, order=ORDER_GIVEN):
        """ORDER_GIVEN evaluates points as passed; ORDER_LOCALITY sorts each
        batch of >= 1024 points by cell first (same results, better reuse)."""
        check(lib.einspline_set_point_order(self.ptr, int(order)), "set_point_order")

    def set_kernel(self, kernel=KERNEL_VECTOR):
        """KERNEL_MATRIX evaluates vgh of a BF16/FP16 table (strided batches)
        on the matrix cores; KERNEL_VECTOR (default) uses the vector kernel."""
        check(lib.einspline_set_kernel(self.ptr, int(kernel)), "set_kernel")

    @property
    def kernel(self):
        k = lib.einspline_get_kernel(self.ptr)
        check(min(k, 0), "get_kernel")
        return k

    def tuning(self):
        """(lane_bytes, block, store_policy, point_order) in effect."""
        v = [_i() for _ in range(4)]
        check(lib.einspline_get_tuning(self.ptr, *[C.byref(x) for x in v]), "get_tuning")
        return tuple(x.value for x in v)

    def autotune(self, mode="vgh", sample_points=8192):
        """Pick the fastest tile and store policy for `mode` on this table
        (cached per shape).  Returns (lane_bytes, block, store_policy, ms)."""
        lb, blk, nt, ms = _i(), _i(), _i(), C.c_float()
        check(lib.einspline_autotune(self.ptr, MODES.index(mode), int(sample_points), C.byref(lb),
                                     C.byref(blk), C.byref(nt), C.byref(ms)), "einspline_autotune")
        return lb.value, blk.value, nt.value, ms.value

    def autotune_points(self, mode, points):
        """Like autotune, but timed on the caller's own host positions
        (an (n, 3) array, a sample of the points to be evaluated), so the
        point order suits their distribution.  Not cached."""
        pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        lb, blk, nt, ms = _i(), _i(), _i(), C.c_float()
        check(lib.einspline_autotune_points(self.ptr, MODES.index(mode),
                                            pts.ctypes.data_as(C.POINTER(C.c_double)), len(pts),
                                            C.byref(lb), C.byref(blk), C.byref(nt), C.byref(ms)),
              "einspline_autotune_points")
        return lb.value, blk.value, nt.value, ms.value

    def out_sizes(self, mode, packed=False):
        """Reals per point of (vals, grads, lapl/hess, third derivatives)."""
        n, D = self.nreal, self.dim
        hv = {3: (6 if packed else 9), 2: 4, 1: 1}[D]
        h = {"v": 0, "vg": 0, "vgl": n, "vgh": hv * n, "vghgh": 9 * n}[mode]
        return n, (D * n if mode != "v" else 0), h, (27 * n if mode == "vghgh" else 0)

    def alloc_outputs(self, mode, npts, packed=False):
        es = np.dtype(self.result_dtype).itemsize
        return tuple(DeviceBuffer(max(1, s * npts * es)) if s else None
                     for s in self.out_sizes(mode, packed))

    def eval_batch(self, mode, pos: DeviceBuffer, npts, outs, packed=False, stream=None):
        v, g, h, gh = (tuple(outs) + (None,) * 4)[:4]
        sp = stream.ptr if stream else None
        if mode == "v":
            rc = lib.einspline_eval_v_batch(self.ptr, pos.ptr, npts, v.ptr, sp)
        elif mode == "vg":
            rc = lib.einspline_eval_vg_batch(self.ptr, pos.ptr, npts, v.ptr, g.ptr, sp)
        elif mode == "vgl":
            rc = lib.einspline_eval_vgl_batch(self.ptr, pos.ptr, npts, v.ptr, g.ptr, h.ptr, sp)
        elif mode == "vghgh":
            rc = lib.einspline_eval_vghgh_batch(self.ptr, pos.ptr, npts, v.ptr, g.ptr, h.ptr, gh.ptr, sp)
        else:
            rc = lib.einspline_eval_vgh_batch(self.ptr, pos.ptr, npts, v.ptr, g.ptr, h.ptr,
                                              1 if packed else 0, sp)
        check(rc, f"einspline_eval_{mode}_batch")

    def download(self, mode, npts, outs, packed=False):
        """Copy batch results to host arrays shaped per point (complex for c/z)."""
        res = []
        for buf, s in zip(outs, self.out_sizes(mode, packed)):
            if not s:
                continue
            a = buf.download(self.result_dtype, (npts, s))
            if self.complex:
                a = a.view(np.complex64 if self.result_dtype == np.float32 else np.complex128)
            res.append(a)
        return tuple(res)

    def destroy(self):
        if self.ptr:
            lib.destroy_multi_UBspline_3d_hip(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class NUGrid:
    """A non-uniform grid made by the library: NUGrid.general(points),
    NUGrid.center(start, end, ratio, n) or NUGrid.log(start, end, n)."""

    def __init__(self, ptr):
        if not ptr:
            raise EinsplineError(E_INVAL, "create grid")
        self.ptr = ptr

    @classmethod
    def general(cls, points):
        p = np.ascontiguousarray(points, np.float64)
        return cls(lib.create_general_grid(p.ctypes.data, len(p)))

    @classmethod
    def center(cls, start, end, ratio, n):
        return cls(lib.create_center_grid(float(start), float(end), float(ratio), int(n)))

    @classmethod
    def log(cls, start, end, n):
        return cls(lib.create_log_grid(float(start), float(end), int(n)))

    @property
    def points(self):
        g = self.ptr.contents
        return np.ctypeslib.as_array(g.points, (g.num_points,)).copy()

    def reverse(self, x):
        g = self.ptr.contents
        f = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double)(g.reverse_map)
        return f(C.cast(self.ptr, C.c_void_p), float(x))

    def destroy(self):
        if self.ptr:
            lib.destroy_grid(self.ptr)
            self.ptr = None


class NUBSpline1D:
    """A multi_NUBspline_1d_{d,z}: kind 'd' or 'z', a NUGrid (which must
    outlive it), bc = (lcode, rcode[, lval, rval]), N orbitals."""

    def __init__(self, kind, grid: NUGrid, bc, N):
        assert kind in "dz"
        self.kind, self.grid, self.N = kind, grid, int(N)
        lc, rc, lv, rv = tuple(bc) + (0.0,) * (4 - len(bc))
        if kind == "d":
            b = BCtype_d(lc, rc, float(lv), float(rv))
            self.ptr = lib.create_multi_NUBspline_1d_d(grid.ptr, b, self.N)
        else:
            lv, rv = complex(lv), complex(rv)
            b = BCtype_z(lc, rc, lv.real, lv.imag, rv.real, rv.imag)
            self.ptr = lib.create_multi_NUBspline_1d_z(grid.ptr, b, self.N)
        if not self.ptr:
            raise EinsplineError(E_INVAL, f"create_multi_NUBspline_1d_{kind}")
        self.dtype = np.complex128 if kind == "z" else np.float64

    def set(self, n, data):
        d = np.ascontiguousarray(data, self.dtype)
        getattr(lib, f"set_multi_NUBspline_1d_{self.kind}")(self.ptr, int(n), d.ctypes.data)

    def coefs(self):
        t = self.ptr.contents
        rows = t.x_grid.contents.num_points + 2
        buf = (C.c_char * (rows * self.N * np.dtype(self.dtype).itemsize)).from_address(t.coefs)
        return np.frombuffer(buf, self.dtype).reshape(rows, self.N).copy()

    def eval(self, mode, x):
        """Single-point evaluation (z only): 'v', 'vg', 'vgl' or 'vgh'."""
        outs = [np.zeros(self.N, np.complex128) for _ in range({"v": 1, "vg": 2}.get(mode, 3))]
        suf = "" if mode == "v" else "_" + mode
        getattr(lib, f"eval_multi_NUBspline_1d_z{suf}")(self.ptr, float(x), *[o.ctypes.data for o in outs])
        return tuple(outs)

    def destroy(self):
        if self.ptr:
            lib.destroy_multi_NUBspline(self.ptr)
            self.ptr = None
