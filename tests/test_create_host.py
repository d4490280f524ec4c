"""Host construction (libeinspline create/set, src/multi_bspline_create.c)
checked against the oracle and the dense interpolation maps.  CPU only:
create/set never touch the GPU."""
import numpy as np
import pytest

import einspline_amd as ea
import oracle
from oracle import brute


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


BCS = {
    "periodic": [(ea.PERIODIC, ea.PERIODIC)] * 3,
    "antiperiodic": [(ea.ANTIPERIODIC, ea.ANTIPERIODIC)] * 3,
    "natural": [(ea.NATURAL, ea.NATURAL)] * 3,
    "flat": [(ea.FLAT, ea.FLAT)] * 3,
    "deriv": [(ea.DERIV1, ea.DERIV2, 0.7, -0.4), (ea.DERIV2, ea.DERIV1, 0.3, 1.1),
              (ea.DERIV1, ea.DERIV1, -0.5, 0.25)],
    "mixed": [(ea.PERIODIC, ea.PERIODIC), (ea.FLAT, ea.NATURAL), (ea.ANTIPERIODIC, ea.ANTIPERIODIC)],
}


@pytest.mark.parametrize("name", list(BCS))
def test_create_d_matches_oracle_and_dense(name, rng):
    bcs = BCS[name]
    M, start, end, N = (7, 6, 5), (3.1, 8.7, 4.5), (9.1, 12.7, 9.3), 4
    h = ea.HostSpline("d", M, start, end, N, bc=bcs)
    o = oracle.Table("d", M, start, end, N, bc=bcs)
    data = rng.uniform(-1, 1, (N,) + M)
    for n in range(N):
        h.set(n, data[n])
        o.set(n, data[n])
    c = h.coefs()
    assert c.shape == o.coefs.shape
    assert rel(c, o.coefs) < 1e-13
    lr = [(b[2], b[3]) if len(b) > 2 else (0.0, 0.0) for b in bcs]
    for n in range(N):
        ref = brute.solve_dense(data[n], [b[:2] for b in bcs], o.delta_inv, lr)
        assert rel(c[..., n], ref) < 1e-13
    st = h.struct
    assert st.num_splines == N and st.z_stride == N
    assert st.y_stride == c.shape[2] * N and st.x_stride == c.shape[1] * c.shape[2] * N
    np.testing.assert_allclose([st.x_grid.delta_inv, st.y_grid.delta_inv, st.z_grid.delta_inv],
                               o.delta_inv, rtol=0, atol=0)


def test_create_s_is_double_solve_rounded(rng):
    """The library solves float tables in double (rounded once); the reference's
    default build solves in float.  Both agree to float precision."""
    M, N = (9, 8, 10), 6
    h = ea.HostSpline("s", M, (0, 0, 0), (1, 1, 1), N)
    o = oracle.Table("s", M, (0, 0, 0), (1, 1, 1), N)
    od = oracle.Table("d", M, (0, 0, 0), (1, 1, 1), N)
    data = rng.uniform(-1, 1, (N,) + M).astype(np.float32)
    for n in range(N):
        h.set(n, data[n])
        o.set(n, data[n])
        od.set(n, data[n].astype(np.float64))
    c = h.coefs()
    np.testing.assert_array_equal(c, od.coefs.astype(np.float32))   # exactly the rounded double solve
    assert rel(c, o.coefs) < 2e-6


@pytest.mark.parametrize("kind", ["c", "z"])
def test_create_complex_matches_oracle(kind, rng):
    bcs = [(ea.PERIODIC, ea.PERIODIC), (ea.DERIV1, ea.DERIV1, 0.2 - 0.3j, 0.5 + 0.1j),
           (ea.NATURAL, ea.NATURAL)]
    M, N = (5, 6, 7), 3
    h = ea.HostSpline(kind, M, (0, 0, 0), (1, 2, 3), N, bc=bcs)
    o = oracle.Table(kind, M, (0, 0, 0), (1, 2, 3), N, bc=bcs)
    data = rng.uniform(-1, 1, (N,) + M) + 1j * rng.uniform(-1, 1, (N,) + M)
    for n in range(N):
        h.set(n, data[n])
        o.set(n, data[n])
    c = h.coefs().reshape(o.coefs.shape)
    assert rel(c, o.coefs) < (1e-13 if kind == "z" else 2e-6)


@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_block_set_equals_single_sets(kind, rng):
    M, N = (6, 7, 8), 9
    a = ea.HostSpline(kind, M, (0, 0, 0), (1, 1, 1), N)
    b = ea.HostSpline(kind, M, (0, 0, 0), (1, 1, 1), N)
    data = rng.uniform(-1, 1, (N,) + M)
    if kind in "cz":
        data = data + 1j * rng.uniform(-1, 1, (N,) + M)
    for n in range(N):
        a.set(n, data[n])
    b.set_block(0, data[:4])
    b.set_block(4, data[4:])
    np.testing.assert_array_equal(a.coefs(), b.coefs())


def test_unset_orbitals_are_zero_and_set_touches_one_column(rng):
    M, N = (5, 5, 5), 4
    h = ea.HostSpline("d", M, (0, 0, 0), (1, 1, 1), N)
    assert not h.coefs().any()
    h.set(2, rng.uniform(-1, 1, M))
    c = h.coefs()
    assert c[..., 2].any() and not c[..., [0, 1, 3]].any()


def test_invalid_arguments_return_null():
    with pytest.raises(ea.EinsplineError):
        ea.HostSpline("d", (1, 5, 5), (0, 0, 0), (1, 1, 1), 2)
    with pytest.raises(ea.EinsplineError):
        ea.HostSpline("s", (5, 5, 5), (0, 0, 0), (1, 1, 1), 0)


# ---- 1-D and 2-D tables (src/multi_bspline_create.c:88-250, :389-574,
# :767-935, :1091-1315) ----------------------------------------------------

LOW_BCS = {
    "periodic": [(ea.PERIODIC, ea.PERIODIC)] * 2,
    "antiperiodic": [(ea.ANTIPERIODIC, ea.ANTIPERIODIC)] * 2,
    "deriv": [(ea.DERIV1, ea.DERIV2, 0.7, -0.4), (ea.FLAT, ea.NATURAL)],
}


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("name", list(LOW_BCS))
@pytest.mark.parametrize("kind", ["s", "d", "c", "z"])
def test_low_dim_create_matches_oracle_and_dense(dim, name, kind, rng):
    bcs = LOW_BCS[name][:dim]
    if kind in "cz":
        bcs = [(b[0], b[1], b[2] + 0.25j, b[3] - 0.5j) if len(b) > 2 else b for b in bcs]
    M, start, end, N = (9, 7)[:dim], (1.5, -2.0)[:dim], (4.0, 3.5)[:dim], 3
    h = ea.HostSpline(kind, M, start, end, N, bc=bcs)
    o = oracle.Table(kind, M, start, end, N, bc=bcs)
    data = rng.uniform(-1, 1, (N,) + M)
    if kind in "cz":
        data = data + 1j * rng.uniform(-1, 1, (N,) + M)
    for n in range(N):
        h.set(n, data[n])
        o.set(n, data[n])
    c = h.coefs().reshape(o.coefs.shape)
    tol = 2e-6 if kind in "sc" else 1e-13
    assert rel(c, o.coefs) < tol
    if kind in "dz":
        lr = [(b[2], b[3]) if len(b) > 2 else (0.0, 0.0) for b in bcs]
        view = c.view(np.complex128) if kind == "z" else c
        for n in range(N):
            ref = brute.solve_dense(data[n], [b[:2] for b in bcs], o.delta_inv, lr)
            assert rel(view[..., n], ref) < 1e-13
    st = h.struct
    assert st.num_splines == N and st.spcode == (ea.MULTI_U1D if dim == 1 else ea.MULTI_U2D)
    if dim == 1:
        assert st.x_stride == N
    else:
        assert st.y_stride == N and st.x_stride == h.shape[1] * N


def test_1d_bc_setters_use_their_own_conditions(rng):
    """set_multi_UBspline_1d_d_BC / _1d_z_BC solve with the given BCs
    (src/multi_bspline_create.c:830-837, :1168-1197)."""
    M, N = 11, 2
    table_bc = [(ea.DERIV1, ea.DERIV1, 0.0, 0.0)]
    own = (ea.DERIV2, ea.DERIV1, 1.5, -0.5)
    for kind, bct in (("d", ea.BCtype_d), ("z", ea.BCtype_z)):
        h = ea.HostSpline(kind, (M,), (0.0,), (2.0,), N, bc=table_bc)
        data = rng.uniform(-1, 1, (N, M)) + (1j * rng.uniform(-1, 1, (N, M)) if kind == "z" else 0)
        arr = np.ascontiguousarray(data[1], dtype=np.complex128 if kind == "z" else np.float64)
        bc = ea._bc(kind, *own)
        getattr(ea.lib, f"set_multi_UBspline_1d_{kind}_BC")(h.ptr, 1, arr.ctypes.data, bc)
        o = oracle.Table(kind, (M,), (0.0,), (2.0,), N, bc=[own])
        o.set(1, data[1])
        c = h.coefs().reshape(o.coefs.shape)
        assert rel(c[:, 1], o.coefs[:, 1]) < 1e-13
        assert not c[:, 0].any()


def test_low_dim_struct_layout_and_destroy(rng):
    for dim in (1, 2):
        for kind in "sdcz":
            h = ea.HostSpline(kind, (6, 5)[:dim], (0, 0)[:dim], (1, 1)[:dim], 4)
            h.set(0, rng.uniform(-1, 1, (6, 5)[:dim]))
            assert h.coefs()[..., 0].any()
            h.destroy()
