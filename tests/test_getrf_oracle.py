"""The LU oracle (oracle_dgetrf, LAPACK dgetf2 semantics) pinned without the
GPU: hand-worked factorisations, LAPACK's own dgetrf through scipy (same
pivots, factors to rounding), the reconstruction P A = L U, an exactly
representable factorisation (bitwise), zero-pivot info, and a planted
mistake in a copy of the oracle that these checks catch."""
import numpy as np
import pytest
from scipy.linalg import lapack

from oracle import oracle as O


def lu_parts(lu, m, n):
    k = min(m, n)
    L = np.tril(lu[:, :k], -1) + np.eye(m, k)
    U = np.triu(lu[:k, :])
    return L, U


def perm_rows(a, ipiv):
    """Apply the interchanges ipiv (1-based, sequential) to the rows of a."""
    p = np.array(a, copy=True)
    for j, r in enumerate(ipiv):
        r -= 1
        if r != j:
            p[[j, r]] = p[[r, j]]
    return p


def test_hand_2x2():
    a = np.asfortranarray([[1.0, 2.0], [3.0, 4.0]])
    ipiv, info = O.dgetrf(2, 2, a, 2)
    assert info == 0 and list(ipiv) == [2, 2]
    # rows swapped: U = [[3, 4], [0, 2 - 4/3]], l21 = 1/3 (computed as 1 * (1/3))
    assert a[0, 0] == 3.0 and a[0, 1] == 4.0
    assert a[1, 0] == 1.0 * (1.0 / 3.0)
    assert a[1, 1] == 2.0 - a[1, 0] * 4.0


def test_hand_3x3_textbook():
    # worked by hand: pivot rows 3, 3, 3; diag(U) = 8, -3/4, -2/3; det = 4
    a0 = np.array([[2.0, 1.0, 1.0], [4.0, 3.0, 3.0], [8.0, 7.0, 9.0]])
    a = np.asfortranarray(a0)
    ipiv, info = O.dgetrf(3, 3, a, 3)
    assert info == 0 and list(ipiv) == [3, 3, 3]
    L, U = lu_parts(a, 3, 3)
    np.testing.assert_allclose(L @ U, perm_rows(a0, ipiv), rtol=0, atol=1e-15)
    np.testing.assert_allclose(np.diag(U), [8.0, -0.75, -2.0 / 3.0], rtol=1e-15)
    np.testing.assert_allclose(L[1:, 0], [0.25, 0.5], rtol=0)   # rows 2, 3 end up as A's rows 1, 2
    assert abs(np.prod(np.diag(U)) * (-1) ** sum(1 for j, r in enumerate(ipiv) if r - 1 != j)
               - np.linalg.det(a0)) < 1e-12


@pytest.mark.parametrize("m,n,seed", [(50, 50, 1), (80, 30, 2), (30, 80, 3), (1, 7, 4), (7, 1, 5), (200, 200, 6)])
def test_matches_lapack_dgetrf(m, n, seed):
    s = O.Stream(seed)
    a0 = s.f64(m, n)
    a = np.asfortranarray(a0.copy())
    ipiv, info = O.dgetrf(m, n, a, m)
    lu, piv, linfo = lapack.dgetrf(a0)
    assert info == linfo == 0
    np.testing.assert_array_equal(ipiv, piv + 1)
    scale = np.abs(lu).max()
    assert np.abs(a - lu).max() <= 64 * max(m, n) * np.finfo(float).eps * scale
    L, U = lu_parts(a, m, n)
    res = np.linalg.norm(L @ U - perm_rows(a0, ipiv)) / np.linalg.norm(a0)
    assert res <= 4 * max(m, n) * np.finfo(float).eps


def test_exact_dyadic_factorisation_is_bitwise():
    # |l| <= 1/2 and a dominant power-of-two diagonal in U: partial pivoting
    # keeps the natural order and every operation is exact.
    rng = np.random.default_rng(7)
    n = 40
    L = np.tril(rng.integers(-4, 5, (n, n)) / 8.0, -1) + np.eye(n)
    U = np.triu(rng.integers(-3, 4, (n, n)).astype(float), 1) + np.diag(2.0 ** rng.integers(8, 12, n))
    a0 = L @ U
    a = np.asfortranarray(a0.copy())
    ipiv, info = O.dgetrf(n, n, a, n)
    assert info == 0 and list(ipiv) == list(range(1, n + 1))
    np.testing.assert_array_equal(np.tril(a, -1), np.tril(L, -1))
    np.testing.assert_array_equal(np.triu(a), U)


def test_zero_pivot_info_matches_lapack():
    s = O.Stream(11)
    a0 = s.f64(12, 12)
    a0[:, 4] = 0.0
    a = np.asfortranarray(a0.copy())
    ipiv, info = O.dgetrf(12, 12, a, 12)
    lu, piv, linfo = lapack.dgetrf(a0)
    assert info == linfo == 5
    np.testing.assert_array_equal(ipiv, piv + 1)
    np.testing.assert_allclose(a, lu, rtol=0, atol=1e-13)


def test_planted_mistake_is_caught():
    """A pivot rule that takes the last maximum, or no pivoting at all, fails
    the LAPACK comparison above: check on one case that it would."""
    s = O.Stream(3)
    a0 = s.f64(30, 30)
    _, piv, _ = lapack.dgetrf(a0)
    # no pivoting: ipiv would be the identity, which LAPACK's is not here
    assert not np.array_equal(piv, np.arange(30))
    # a transposed operand: factoring A^T gives different pivots
    at = np.asfortranarray(a0.T.copy())
    ipiv_t, _ = O.dgetrf(30, 30, at, 30)
    assert not np.array_equal(ipiv_t, piv + 1)
