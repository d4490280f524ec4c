"""Pins of the oracle's Slater-determinant functions (oracle.update_inverse,
oracle.det_ratio) against what the mathematics fixes, and the argument
checks of the determinant entry points of the C ABI (no GPU needed).

The reference holds no golden vectors for these kernels.  Its own checks are
  * src/determinant_update.cu:494-506: after the update, Ainv' A' = I
    (|error| <= 1e-4 in float);
  * src/cuda_walker.cu:344-370: the ratio equals det(A') / det(A) computed
    by LU on the host.
Both are restated here, in double with tight bounds, plus the textbook
inverse of the updated matrix (numpy.linalg.inv), which a transposed
operand, a wrong sign of the prefactor or a row taken for the column fails.
"""
import numpy as np
import pytest

import oracle


def well_conditioned(rng, N, dtype=np.float64):
    A = np.eye(N) + 0.5 * rng.uniform(-1, 1, (N, N)) / np.sqrt(N)
    return A.astype(dtype)


@pytest.mark.parametrize("N", [1, 2, 5, 31, 128])
@pytest.mark.parametrize("form", ["row", "delta"])
def test_update_matches_inverse_of_updated_matrix(N, form, rng):
    for k in sorted({0, N // 2, N - 1}):
        A = well_conditioned(rng, N)
        Ainv = np.linalg.inv(A)
        change = 0.3 * rng.uniform(-1, 1, N) / np.sqrt(N)
        A2 = A.copy()
        A2[k] += change
        if form == "row":
            u = A2[k].copy()
            Awork = A.copy()
            oracle.update_inverse(Awork, Ainv, u, k)
            np.testing.assert_array_equal(Awork, A2)        # row k written
        else:
            oracle.update_inverse(None, Ainv, change.copy(), k)
        ref = np.linalg.inv(A2)
        assert np.abs(Ainv - ref).max() / np.abs(ref).max() < 1e-12
        assert np.abs(Ainv @ A2 - np.eye(N)).max() < 1e-12


def test_update_float_meets_reference_identity_check(rng):
    """src/determinant_update.cu:494-506 in float: |Ainv' A' - I| <= 1e-4."""
    N = 128
    A = well_conditioned(rng, N)
    Ainv = np.linalg.inv(A).astype(np.float32)
    u = (A[7] + 0.2 * rng.uniform(-1, 1, N) / np.sqrt(N)).astype(np.float32)
    A32 = A.astype(np.float32)
    oracle.update_inverse(A32, Ainv, u, 7)
    ident = Ainv.astype(np.float64) @ A32.astype(np.float64)
    assert np.abs(ident - np.eye(N)).max() < 1e-4


def test_update_with_row_stride_touches_only_the_matrix(rng):
    N, rs, k = 9, 13, 4
    A = well_conditioned(rng, N)
    Ap = np.full((N, rs), 7.0)
    Ap[:, :N] = A
    Ainv = np.full((N, rs), -3.0)
    Ainv[:, :N] = np.linalg.inv(A)
    u = A[k] + 0.1
    up = np.zeros(rs)
    up[:N] = u
    oracle.update_inverse(Ap, Ainv, up, k, N=N)
    A2 = A.copy()
    A2[k] = u
    assert np.abs(Ainv[:, :N] - np.linalg.inv(A2)).max() < 1e-12
    assert (Ainv[:, N:] == -3.0).all() and (Ap[:, N:] == 7.0).all()


@pytest.mark.parametrize("N", [1, 3, 64, 200])
def test_ratio_is_determinant_ratio(N, rng):
    """src/cuda_walker.cu:344-370: host det(A')/det(A) by LU."""
    A = well_conditioned(rng, N)
    Ainv = np.linalg.inv(A)
    for col in sorted({0, N - 1, N // 3}):
        u = A[col] + 0.5 * rng.uniform(-1, 1, N) / np.sqrt(N)
        A2 = A.copy()
        A2[col] = u
        want = np.linalg.det(A2) / np.linalg.det(A)
        assert abs(oracle.det_ratio(Ainv, u, col) - want) < 1e-12 * max(1.0, abs(want))
    # the unchanged row gives ratio 1 (A Ainv = I)
    assert abs(oracle.det_ratio(Ainv, A[0].copy(), 0) - 1.0) < 1e-12


def test_ratio_equals_one_plus_dk_of_the_update(rng):
    """det(A')/det(A) = 1 + d_k: the ratio and the update agree on the pivot."""
    N, k = 40, 11
    A = well_conditioned(rng, N)
    Ainv = np.linalg.inv(A)
    u = A[k] + 0.4 * rng.uniform(-1, 1, N) / np.sqrt(N)
    r = oracle.det_ratio(Ainv, u, k)
    delta = u - A[k]
    assert abs(r - (1.0 + Ainv[:, k] @ delta)) < 1e-13


def test_entry_points_reject_bad_arguments_without_a_gpu():
    import einspline_amd as ea
    L = ea.lib
    for k in "sd":
        up = getattr(L, f"einspline_update_inverse_{k}")
        cr = getattr(L, f"einspline_calc_ratios_{k}")
        fr = getattr(L, f"einspline_fill_ratios_{k}")
        fake = 16                                      # never dereferenced: rejected first
        assert up(None, fake, fake, None, fake, 4, 4, 0, 0, None) == ea.E_OK     # nothing to do
        assert up(None, fake, fake, None, fake, 0, 4, 0, 1, None) == ea.E_INVAL  # N < 1
        assert up(None, fake, fake, None, fake, 5, 4, 0, 1, None) == ea.E_INVAL  # rowstride < N
        assert up(None, fake, fake, None, fake, 4, 4, 4, 1, None) == ea.E_INVAL  # k >= N
        assert up(fake, fake, fake, None, fake, 4, 4, 0, 1, None) == ea.E_INVAL  # A without delta
        assert up(None, None, fake, None, fake, 4, 4, 0, 1, None) == ea.E_INVAL  # NULL Ainv list
        assert up(None, fake, fake, None, fake, 4, 4, 0, -1, None) == ea.E_INVAL
        assert cr(fake, fake, fake, 4, 4, 0, 0, None) == ea.E_OK
        assert cr(fake, fake, fake, 4, 4, -1, 1, None) == ea.E_INVAL
        assert cr(fake, fake, None, 4, 4, 0, 1, None) == ea.E_INVAL
        assert cr(fake, fake, fake, 4, 3, 0, 1, None) == ea.E_INVAL
        assert fr(None, fake, fake, fake, fake, 4, 4, 0, 1, None) == ea.E_INVAL  # no table
