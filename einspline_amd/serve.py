"""ctypes front end of the serving runtime (include/einspline_serve.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import (DeviceBuffer, EinsplineError, E_INVAL, check, lib)


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("slots", C.c_int), ("max_batch_points", C.c_int),
                ("max_batch_requests", C.c_int), ("hess_packed", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("evals", C.c_double), ("admissions", C.c_uint64), ("evictions", C.c_uint64),
                ("deferrals", C.c_uint64), ("resident_tables", C.c_int),
                ("resident_bytes", C.c_size_t)]


class Load(C.Structure):
    _fields_ = [("rate_hz", C.c_double), ("duration_s", C.c_double), ("max_points", C.c_int),
                ("seed", C.c_uint64), ("ntables", C.c_int), ("table_ids", C.POINTER(C.c_int)),
                ("vals_d", C.POINTER(C.c_void_p)), ("grads_d", C.POINTER(C.c_void_p)),
                ("hess_d", C.POINTER(C.c_void_p)), ("box", C.c_double * 3)]


class AdmissionTable(C.Structure):
    _fields_ = [("bytes", C.c_size_t), ("state", C.c_int), ("busy", C.c_int), ("pending", C.c_int),
                ("last_use", C.c_uint64)]


ADMIT_RESIDENT, ADMIT_UPLOAD, ADMIT_DEFER = 0, 1, 2

_vp = C.c_void_p
lib.einspline_server_plan_admission.restype = C.c_int
lib.einspline_server_plan_admission.argtypes = [C.POINTER(AdmissionTable), C.c_int, C.c_int, C.c_size_t,
                                                C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
lib.einspline_server_create.restype = _vp
lib.einspline_server_create.argtypes = [C.POINTER(ServerConfig)]
lib.einspline_server_destroy.restype = None
lib.einspline_server_destroy.argtypes = [_vp]
lib.einspline_server_register_table.restype = C.c_int
lib.einspline_server_register_table.argtypes = [_vp, _vp, C.c_int, C.POINTER(C.c_int)]
lib.einspline_server_submit.restype = C.c_int
lib.einspline_server_submit.argtypes = [_vp, C.c_int, _vp, C.c_int, _vp, _vp, _vp,
                                        C.POINTER(C.c_uint64)]
lib.einspline_server_test.restype = C.c_int
lib.einspline_server_test.argtypes = [_vp, C.c_uint64, C.POINTER(C.c_double)]
lib.einspline_server_wait.restype = C.c_int
lib.einspline_server_wait.argtypes = [_vp, C.c_uint64, C.POINTER(C.c_double)]
lib.einspline_server_drain.restype = C.c_int
lib.einspline_server_drain.argtypes = [_vp]
lib.einspline_server_get_stats.restype = C.c_int
lib.einspline_server_get_stats.argtypes = [_vp, C.POINTER(ServerStats)]
lib.einspline_server_table_resident.restype = C.c_int
lib.einspline_server_table_resident.argtypes = [_vp, C.c_int]
lib.einspline_server_run_load.restype = C.c_int
lib.einspline_server_run_load.argtypes = [_vp, C.POINTER(Load), C.POINTER(C.c_double), C.c_size_t,
                                          C.POINTER(C.c_size_t), C.POINTER(C.c_double)]


def plan_admission(tables, table_id, used, budget):
    """einspline_server_plan_admission on a list of dicts with keys bytes,
    state (0 absent, 1 resident, 2 uploading), busy, pending, last_use.
    Returns (decision or negative error code, victims, drain)."""
    n = len(tables)
    arr = (AdmissionTable * n)(*[AdmissionTable(int(t["bytes"]), int(t["state"]), int(t.get("busy", 0)),
                                                int(t.get("pending", 0)), int(t.get("last_use", 0)))
                                 for t in tables])
    v, d = (C.c_int * n)(), (C.c_int * n)()
    nv, nd = C.c_int(), C.c_int()
    rc = lib.einspline_server_plan_admission(arr, n, int(table_id), int(used), int(budget), v, C.byref(nv),
                                             d, C.byref(nd))
    return rc, sorted(v[:nv.value]), sorted(d[:nd.value])


class Server:
    def __init__(self, hbm_budget=0, slots=4, max_batch_points=16384, max_batch_requests=256,
                 hess_packed=False):
        cfg = ServerConfig(int(hbm_budget), slots, max_batch_points, max_batch_requests,
                           1 if hess_packed else 0)
        self.cfg = cfg
        self.ptr = lib.einspline_server_create(C.byref(cfg))
        if not self.ptr:
            raise EinsplineError(E_INVAL, "einspline_server_create")
        self.tables = {}

    def register(self, host_spline, storage):
        tid = C.c_int()
        check(lib.einspline_server_register_table(self.ptr, C.cast(host_spline.ptr, C.c_void_p),
                                                  int(storage), C.byref(tid)), "register_table")
        self.tables[tid.value] = host_spline
        return tid.value

    def submit(self, table_id, pos, vals_d, grads_d, hess_d, ticket=True):
        pos = np.ascontiguousarray(pos, np.float64)
        t = C.c_uint64()
        check(lib.einspline_server_submit(self.ptr, table_id, pos.ctypes.data, pos.shape[0],
                                          vals_d.ptr, grads_d.ptr, hess_d.ptr,
                                          C.byref(t) if ticket else None), "submit")
        return t.value if ticket else None

    def wait(self, ticket):
        lat = C.c_double()
        check(lib.einspline_server_wait(self.ptr, ticket, C.byref(lat)), "wait")
        return lat.value

    def test(self, ticket):
        lat = C.c_double()
        rc = lib.einspline_server_test(self.ptr, ticket, C.byref(lat))
        if rc < 0:
            check(rc, "test")
        return lat.value if rc == 1 else None

    def drain(self):
        check(lib.einspline_server_drain(self.ptr), "drain")

    def stats(self) -> dict:
        s = ServerStats()
        check(lib.einspline_server_get_stats(self.ptr, C.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def resident(self, tid) -> bool:
        rc = lib.einspline_server_table_resident(self.ptr, tid)
        if rc < 0:
            check(rc, "table_resident")
        return bool(rc)

    def run_load(self, table_ids, sinks, rate_hz, duration_s, max_points=2048, seed=1,
                 box=(1.0, 1.0, 1.0), cap=2_000_000):
        n = len(table_ids)
        ids = (C.c_int * n)(*table_ids)
        vv = (C.c_void_p * n)(*[s[0].ptr for s in sinks])
        gg = (C.c_void_p * n)(*[s[1].ptr for s in sinks])
        hh = (C.c_void_p * n)(*[s[2].ptr for s in sinks])
        ld = Load(float(rate_hz), float(duration_s), int(max_points), int(seed), n, ids, vv, gg, hh,
                  (C.c_double * 3)(*box))
        lat = np.zeros(cap)
        nreq = C.c_size_t()
        wall = C.c_double()
        check(lib.einspline_server_run_load(self.ptr, C.byref(ld),
                                            lat.ctypes.data_as(C.POINTER(C.c_double)), cap,
                                            C.byref(nreq), C.byref(wall)), "run_load")
        return lat[:min(nreq.value, cap)], wall.value

    def close(self):
        if self.ptr:
            lib.einspline_server_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
