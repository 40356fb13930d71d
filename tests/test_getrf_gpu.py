"""DGETRF on the GPU through the C ABI against the LU oracle (LAPACK dgetf2
semantics): identical pivots, factors within rounding of the oracle's,
backward error of P A = L U, every leaf-panel shape (1024 to 16384 rows in
registers, taller panels in global memory), wide and tall matrices, padded
lda, zero-pivot info, an exactly representable factorisation (bitwise), the
host-pointer dgetrf_, and argument errors."""
import numpy as np
import pytest

from gemm_avx512f_amd import _lib, blas
from gemm_avx512f_amd.device import DeviceBuffer, DeviceMatrix
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def perm_rows(a, ipiv):
    p = np.array(a, copy=True)
    for j, r in enumerate(ipiv):
        r -= 1
        if r != j:
            p[[j, r]] = p[[r, j]]
    return p


def lu_parts(lu, m, n):
    k = min(m, n)
    return np.tril(lu[:, :k], -1) + np.eye(m, k), np.triu(lu[:k, :])


def getrf_gpu(a, m, n):
    """Factor the first m rows of the Fortran-order array a (its leading dimension = a.shape[0])."""
    dA = DeviceMatrix.from_padded(a, m)
    k = max(min(m, n), 1)
    dpiv, dinfo = DeviceBuffer(4 * k), DeviceBuffer(4)
    blas.dgetrf(m, n, dA, dpiv, dinfo)
    ipiv = np.zeros(k, np.int32)
    info = np.zeros(1, np.int32)
    dpiv.download(ipiv)
    dinfo.download(info)
    return dA.to_host_padded(), ipiv[: min(m, n)], int(info[0])


def backward_error(a0, lu, ipiv, m, n):
    L, U = lu_parts(lu, m, n)
    return np.linalg.norm(L @ U - perm_rows(a0, ipiv)) / np.linalg.norm(a0)


@pytest.mark.parametrize("m,n,pad,seed", [
    (1, 1, 0, 1), (5, 3, 0, 2), (3, 5, 0, 3), (16, 16, 0, 4), (17, 17, 3, 5), (100, 100, 0, 6),
    (257, 130, 0, 7), (130, 257, 5, 8), (640, 640, 0, 9), (1024, 1024, 0, 10),
    (1500, 1500, 7, 11),   # 2-rows-per-thread leaves
    (3000, 700, 0, 12),    # 4 rows per thread, 8-column leaves
    (5000, 300, 0, 13),    # 8 rows per thread, 4-column leaves
    (9000, 64, 0, 14),     # 16 rows per thread, 2-column leaves
    (20000, 40, 3, 15),    # panels of more than 16384 rows: global-memory leaves of 8 columns
    (40000, 20, 0, 16),
])
def test_matches_oracle(gpu_lib, m, n, pad, seed):
    print(f"dgetrf m={m} n={n} pad={pad} seed={seed}", flush=True)
    s = O.Stream(seed)
    a0 = s.f64(m, n, ld=m + pad)
    ref = np.asfortranarray(a0[:m, :n].copy())
    ipiv_o, info_o = O.dgetrf(m, n, ref, m)
    out, ipiv, info = getrf_gpu(a0, m, n)
    assert info == info_o == 0
    np.testing.assert_array_equal(ipiv, ipiv_o)
    lu = out[:m, :n]
    if pad:
        np.testing.assert_array_equal(out[m:, :n], a0[m:, :n])   # rows past m are never written
    # same pivots, different summation order: factors agree to a few units of
    # rounding scaled by the problem (growth of random matrices is modest)
    scale = np.abs(ref).max()
    assert np.abs(lu - ref).max() <= 1e-9 * scale, np.abs(lu - ref).max() / scale
    assert backward_error(a0[:m, :n], lu, ipiv, m, n) <= 8 * max(m, n) * EPS


def test_exact_dyadic_factorisation_bitwise(gpu_lib):
    """A = L U with |l| <= 1/2 and a dominant power-of-two diagonal in U keeps
    the natural pivot order.  L's strictly lower part is nonzero in five
    columns only, so the explicit inverses the blocked DTRSM forms of L's
    diagonal blocks are products of at most five elementary factors and stay
    exact in FP64 (a dense dyadic L would give inverse entries of more than 53
    bits); every leaf, solve and Schur update is then exact."""
    rng = np.random.default_rng(7)
    n = 300
    L = np.tril(rng.integers(-4, 5, (n, n)) / 8.0, -1)
    keep = np.zeros(n, bool)
    keep[[0, 1, 17, 64, 150]] = True
    L[:, ~keep] = 0.0
    L += np.eye(n)
    U = np.triu(rng.integers(-3, 4, (n, n)).astype(float), 1) + np.diag(2.0 ** rng.integers(10, 14, n))
    a0 = np.asfortranarray(L @ U)
    out, ipiv, info = getrf_gpu(a0, n, n)
    assert info == 0 and np.array_equal(ipiv, np.arange(1, n + 1))
    np.testing.assert_array_equal(np.tril(out, -1), np.tril(L, -1))
    np.testing.assert_array_equal(np.triu(out), U)


def test_zero_pivot_info(gpu_lib):
    s = O.Stream(21)
    a0 = s.f64(200, 200)
    a0[:, 37] = 0.0
    ref = np.asfortranarray(a0.copy())
    ipiv_o, info_o = O.dgetrf(200, 200, ref, 200)
    out, ipiv, info = getrf_gpu(a0, 200, 200)
    assert info == info_o == 38
    np.testing.assert_array_equal(ipiv, ipiv_o)
    assert np.abs(out - ref).max() <= 1e-10 * np.abs(ref).max()


def test_config_size_backward_error(gpu_lib):
    """4096^2 (the DGEMM 4096^3 workload's square operand): pivots equal the oracle's on a
    sampled prefix is not required; the backward error of P A = L U is."""
    n = 4096
    s = O.Stream(0)
    a0 = s.f64(n, n)
    out, ipiv, info = getrf_gpu(a0, n, n)
    assert info == 0
    assert np.all((ipiv >= np.arange(1, n + 1)) & (ipiv <= n))
    assert backward_error(a0, out, ipiv, n, n) <= 8 * n * EPS


def test_host_pointer_dgetrf_(gpu_lib):
    s = O.Stream(31)
    m, n, lda = 300, 200, 311
    a0 = s.f64(m, n, ld=lda)
    ref = np.asfortranarray(a0[:m, :n].copy())
    ipiv_o, _ = O.dgetrf(m, n, ref, m)
    a = np.asfortranarray(a0.copy())
    ipiv = np.zeros(n, np.int32)
    info = blas.dgetrf_(m, n, a, lda, ipiv)
    assert info == 0
    np.testing.assert_array_equal(ipiv, ipiv_o)
    np.testing.assert_array_equal(a[m:], a0[m:])
    assert np.abs(a[:m] - ref).max() <= 1e-10 * np.abs(ref).max()


def test_argument_errors(gpu_lib, capfd):
    a = np.zeros((4, 4), order="F")
    ipiv = np.zeros(4, np.int32)
    assert blas.dgetrf_(-1, 4, a, 4, ipiv) == -1
    assert blas.dgetrf_(4, 4, a, 3, ipiv) == -4
    cap = capfd.readouterr()
    assert "DGETRF" in cap.out + cap.err
    dA = DeviceMatrix(4, 4, np.float64)
    dpiv, dinfo = DeviceBuffer(16), DeviceBuffer(4)
    assert _lib.lib().gemm_amd_dgetrf(4, 4, dA.ptr, 3, dpiv.ptr, dinfo.ptr, None) == 4
    # m = 0 / n = 0: info is still written
    dinfo.upload(np.array([123], np.int32))
    assert _lib.lib().gemm_amd_dgetrf(0, 4, dA.ptr, 1, dpiv.ptr, dinfo.ptr, None) == 0
    info = np.zeros(1, np.int32)
    dinfo.download(info)
    assert info[0] == 0


def test_refactor_after_device_copy_is_repeatable(gpu_lib):
    """8192^2, as bench.py times it: the matrix is restored with a device-to-device
    gemm_amd_memcpy before each factorisation, which runs on a non-blocking
    stream.  The copy must be complete when the entry point returns (the
    factorisation's first row interchange spans every column): the second
    factorisation gives the first one's bits."""
    import ctypes

    from gemm_avx512f_amd.device import Stream, fill_uniform

    n = 8192
    stream = Stream()
    A0, A = DeviceMatrix(n, n, np.float64), DeviceMatrix(n, n, np.float64)
    fill_uniform(A0, 21, 0, stream=stream)
    stream.synchronize()
    piv, info = DeviceBuffer(4 * n), DeviceBuffer(4)
    outs = []
    for _ in range(2):
        _lib.check(_lib.lib().gemm_amd_memcpy(ctypes.c_void_p(A.ptr), ctypes.c_void_p(A0.ptr), 8 * n * n, 3), "copy")
        blas.dgetrf(n, n, A, piv, info, stream=stream)
        stream.synchronize()
        ipiv = np.zeros(n, np.int32)
        piv.download(ipiv)
        outs.append((A.to_host(), ipiv))
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][0], outs[1][0])
