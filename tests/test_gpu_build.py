"""Tables solved on the GPU (einspline_build_table_hip) against the host
build (set_multi_UBspline_3d_* + create_multi_UBspline_3d_*_hip): the same
solve in the same order, so the device tables must agree byte for byte, for
every kind, every boundary pair, the storage conversions, several orbital
batches and non-cubic grids.  The host build is itself pinned by the oracle
(tests/test_create_host.py), so this pins the GPU build too."""
import numpy as np
import pytest

import einspline_amd as ea

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    assert ea.device_count() >= 1, "no HIP device visible: GPU tests must run on a GPU"


def coef_bytes(tab):
    t = tab.ptr.contents
    out = np.empty(t.bytes, np.uint8)
    ea.check(ea.lib.einspline_hip_memcpy_dtoh(out.ctypes.data, t.coefs, t.bytes), "dtoh")
    return out


def samples(kind, N, M, rng):
    d = rng.uniform(-1, 1, (N,) + tuple(M))
    if kind in "cz":
        d = d + 1j * rng.uniform(-1, 1, (N,) + tuple(M))
    return d


BCS = {
    "periodic": [(ea.PERIODIC, ea.PERIODIC)] * 3,
    "mixed": [(ea.NATURAL, ea.FLAT), (ea.ANTIPERIODIC, ea.ANTIPERIODIC), (ea.DERIV1, ea.DERIV2, 0.7, -1.3)],
    "deriv": [(ea.DERIV1, ea.DERIV1, 0.25, -0.5), (ea.DERIV2, ea.NATURAL, 1.5, 0.0), (ea.FLAT, ea.DERIV2, 0.0, 2.0)],
}


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
@pytest.mark.parametrize("bcs", list(BCS))
def test_gpu_build_equals_host_build(kind, bcs, rng):
    M, N = (9, 7, 6), 37
    h = ea.HostSpline(kind, M, (0.0, -1.0, 0.5), (1.0, 2.0, 1.5), N, bc=BCS[bcs])
    data = samples(kind, N, M, rng)
    h.set_block(0, data)
    host_tab = ea.DeviceTable(h)
    gpu_tab = ea.DeviceTable.build(h, data)
    t0, t1 = host_tab.ptr.contents, gpu_tab.ptr.contents
    assert (t1.row_reals, t1.tcode, t1.storage, t1.Nx, t1.Ny, t1.Nz) == \
           (t0.row_reals, t0.tcode, t0.storage, t0.Nx, t0.Ny, t0.Nz)
    np.testing.assert_array_equal(coef_bytes(gpu_tab), coef_bytes(host_tab))


@pytest.mark.parametrize("storage,host_fn", [(ea.STORE_F32, None), (ea.STORE_BF16, None), (ea.STORE_F16, None),
                                             ("conv", None)])
def test_gpu_build_storage_conversions(storage, host_fn, rng):
    kind = "z" if storage == "conv" else "d"
    M, N = (8, 8, 8), 24
    h = ea.HostSpline(kind, M, (0, 0, 0), (1, 1, 1), N)
    data = samples(kind, N, M, rng)
    h.set_block(0, data)
    host_tab = ea.DeviceTable(h, storage=storage)
    gpu_tab = ea.DeviceTable.build(h, data, storage=ea.STORE_F32 if storage == "conv" else storage)
    assert gpu_tab.tcode == host_tab.tcode and gpu_tab.storage == host_tab.storage
    np.testing.assert_array_equal(coef_bytes(gpu_tab), coef_bytes(host_tab))


def test_gpu_build_in_several_orbital_batches(rng):
    """A table whose double workspace exceeds one pass is solved in orbital
    batches; the result must not depend on the batching."""
    M, N = (40, 40, 40), 600               # 43^3 x 8 B x 600 = 381 MB: two passes of 256 MB
    h = ea.HostSpline("s", M, (0, 0, 0), (1, 1, 1), N)
    data = samples("s", N, M, rng).astype(np.float32)
    h.set_block(0, data)
    host_tab = ea.DeviceTable(h)
    gpu_tab = ea.DeviceTable.build(h, data)
    np.testing.assert_array_equal(coef_bytes(gpu_tab), coef_bytes(host_tab))


def test_gpu_build_rejects_bad_requests(rng):
    h1 = ea.HostSpline("s", (10,), (0.0,), (1.0,), 4)
    with pytest.raises(ea.EinsplineError):
        ea.DeviceTable.build(h1, np.zeros((4, 10), np.float32))          # 1-D: host build only
    h = ea.HostSpline("s", (6, 6, 6), (0, 0, 0), (1, 1, 1), 4)
    with pytest.raises(ea.EinsplineError):
        ea.DeviceTable.build(h, np.zeros((4, 6, 6, 6), np.float32), storage=ea.STORE_F64)
    assert not ea.lib.einspline_build_table_hip(h.ptr, None, -1, None)
