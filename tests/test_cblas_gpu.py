"""CBLAS (row- and column-major) and ILP64 entry points on the GPU through the
C ABI: row-major results against numpy on C-order arrays (GEMM error bound,
DTRSM residual, all 16 DTRSM variants), column-major CBLAS and every ILP64
symbol bitwise equal to the LP64 Fortran routine they wrap, CBLAS-numbered
argument errors, and the ILP64 32-bit range check (nothing written)."""
import ctypes

import numpy as np
import pytest

from gemm_avx512f_amd import _lib, blas

pytestmark = pytest.mark.gpu

ROW, COL = 101, 102
TR = {"N": 111, "T": 112, "C": 113}
UPLO = {"U": 121, "L": 122}
DIAG = {"N": 131, "U": 132}
SIDE = {"L": 141, "R": 142}


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def I64(v):
    return ctypes.byref(ctypes.c_int64(v))


def C(s):
    return ctypes.c_char_p(s.encode())


def op(x, t):
    return x if t == "N" else (x.T if t == "T" else x.conj().T)


def rowmajor_padded(rng, rows, cols, pad, dtype):
    full = rng.uniform(-1, 1, (rows, cols + pad)).astype(dtype)
    if np.issubdtype(dtype, np.complexfloating):
        full = full + 1j * rng.uniform(-1, 1, (rows, cols + pad))
        full = full.astype(dtype)
    return full, full[:, :cols]


@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("tb", "NT")
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_cblas_gemm_row_major(gpu_lib, ta, tb, dt):
    rng = np.random.default_rng(5)
    m, n, k, pad = 150, 97, 203, 3
    A_full, A = rowmajor_padded(rng, *((m, k) if ta == "N" else (k, m)), pad, dt)
    B_full, B = rowmajor_padded(rng, *((k, n) if tb == "N" else (n, k)), pad, dt)
    C_full, C0 = rowmajor_padded(rng, m, n, pad, dt)
    C0 = C0.copy()   # the call overwrites C_full, of which C0 is a view
    c_ref = 1.5 * op(A.astype(np.float64), ta) @ op(B.astype(np.float64), tb) + 0.5 * C0.astype(np.float64)
    fn = _lib.lib().cblas_dgemm if dt == np.float64 else _lib.lib().cblas_sgemm
    fn(ROW, TR[ta], TR[tb], m, n, k, 1.5, P(A_full), A_full.shape[1], P(B_full), B_full.shape[1], 0.5, P(C_full),
       C_full.shape[1])
    scale = 1.5 * np.abs(op(A, ta)).astype(np.float64) @ np.abs(op(B, tb)).astype(np.float64) + 0.5 * np.abs(C0)
    err = np.max(np.abs(C_full[:, :n] - c_ref) / scale)
    assert err <= 4 * (k + 2) * np.finfo(dt).eps, err


def test_cblas_dgemm_col_major_equals_fortran(gpu_lib):
    rng = np.random.default_rng(6)
    m, n, k = 130, 70, 90
    a = np.asfortranarray(rng.uniform(-1, 1, (k, m)))
    b = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    c0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    c1, c2 = c0.copy(order="F"), c0.copy(order="F")
    _lib.lib().cblas_dgemm(COL, TR["T"], TR["N"], m, n, k, 0.7, P(a), k, P(b), k, -1.0, P(c1), m)
    blas.dgemm_("T", "N", m, n, k, 0.7, a, k, b, k, -1.0, c2, m)
    np.testing.assert_array_equal(c1, c2)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "T"), ("T", "C")])
def test_cblas_zgemm3m_row_major(gpu_lib, ta, tb):
    rng = np.random.default_rng(8)
    m, n, k = 64, 48, 80
    A_full, A = rowmajor_padded(rng, *((m, k) if ta == "N" else (k, m)), 0, np.complex128)
    B_full, B = rowmajor_padded(rng, *((k, n) if tb == "N" else (n, k)), 0, np.complex128)
    C_full = np.ascontiguousarray(rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n)))
    alpha, beta = np.array([0.5, -1.25]), np.array([1.0, 0.25])
    ref = complex(*alpha) * op(A, ta) @ op(B, tb) + complex(*beta) * C_full
    _lib.lib().cblas_zgemm3m(ROW, TR[ta], TR[tb], m, n, k, P(alpha), P(A_full), A_full.shape[1], P(B_full),
                             B_full.shape[1], P(beta), P(C_full), n)
    assert np.max(np.abs(C_full - ref)) <= 64 * k * np.finfo(float).eps * np.max(np.abs(ref))


@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("uplo", "UL")
@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("diag", "NU")
def test_cblas_dtrsm_row_major(gpu_lib, side, uplo, ta, diag):
    rng = np.random.default_rng(9)
    m, n = 200, 75
    na = m if side == "L" else n
    a = rng.uniform(-1, 1, (na, na)) + np.diag(np.full(na, 8.0))
    tri = np.triu(a) if uplo == "U" else np.tril(a)
    if diag == "U":
        np.fill_diagonal(tri, 1.0)
    b0 = rng.uniform(-1, 1, (m, n))
    b = np.ascontiguousarray(b0.copy())
    a_c = np.ascontiguousarray(a)
    _lib.lib().cblas_dtrsm(ROW, SIDE[side], UPLO[uplo], TR[ta], DIAG[diag], m, n, 2.0, P(a_c), na, P(b), n)
    opa = tri if ta == "N" else tri.T
    lhs = opa @ b if side == "L" else b @ opa
    # componentwise residual bound of a backward-stable solve: |op(T) X - alpha B| <= c n eps |op(T)| |X|
    # (a random unit triangle is badly conditioned, so X itself can be large)
    bound = (np.abs(opa) @ np.abs(b)) if side == "L" else (np.abs(b) @ np.abs(opa))
    assert np.max(np.abs(lhs - 2.0 * b0) / bound) <= 8 * na * np.finfo(float).eps


def test_ilp64_bitwise_equal_lp64(gpu_lib):
    rng = np.random.default_rng(10)
    m, n, k = 90, 60, 110
    L = _lib.lib()
    for dt, f32, f64 in ((np.float64, blas.dgemm_, L.dgemm_64_), (np.float32, blas.sgemm_, L.sgemm_64_)):
        a = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dt))
        b = np.asfortranarray(rng.uniform(-1, 1, (k, n)).astype(dt))
        c0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
        c1, c2 = c0.copy(order="F"), c0.copy(order="F")
        ct = ctypes.c_double if dt == np.float64 else ctypes.c_float
        f32("N", "N", m, n, k, 1.25, a, m, b, k, 0.5, c1, m)
        f64(C("N"), C("N"), I64(m), I64(n), I64(k), ctypes.byref(ct(1.25)), P(a), I64(m), P(b), I64(k),
            ctypes.byref(ct(0.5)), P(c2), I64(m))
        np.testing.assert_array_equal(c1, c2)
    # complex 3M
    a = np.asfortranarray(rng.uniform(-1, 1, (m, k)) + 1j * rng.uniform(-1, 1, (m, k)))
    b = np.asfortranarray(rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n)))
    c0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n)))
    c1, c2 = c0.copy(order="F"), c0.copy(order="F")
    blas.zgemm3m_("N", "N", m, n, k, 0.5 + 0.5j, a, m, b, k, 1.0 + 0j, c1, m)
    al, be = np.array([0.5, 0.5]), np.array([1.0, 0.0])
    L.zgemm3m_64_(C("N"), C("N"), I64(m), I64(n), I64(k), P(al), P(a), I64(m), P(b), I64(k), P(be), P(c2), I64(m))
    np.testing.assert_array_equal(c1, c2)
    # DTRSM
    t = np.asfortranarray(np.tril(rng.uniform(-1, 1, (m, m))) + 6.0 * np.eye(m))
    x0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    x1, x2 = x0.copy(order="F"), x0.copy(order="F")
    blas.dtrsm_("L", "L", "N", "N", m, n, 1.0, t, m, x1, m)
    L.dtrsm_64_(C("L"), C("L"), C("N"), C("N"), I64(m), I64(n), ctypes.byref(ctypes.c_double(1.0)), P(t), I64(m),
                P(x2), I64(m))
    np.testing.assert_array_equal(x1, x2)
    # DGETRF
    g0 = np.asfortranarray(rng.uniform(-1, 1, (m, m)))
    g1, g2 = g0.copy(order="F"), g0.copy(order="F")
    piv32 = np.zeros(m, np.int32)
    assert blas.dgetrf_(m, m, g1, m, piv32) == 0
    piv64, info = np.zeros(m, np.int64), ctypes.c_int64(-7)
    L.dgetrf_64_(I64(m), I64(m), P(g2), I64(m), P(piv64), ctypes.byref(info))
    assert info.value == 0
    np.testing.assert_array_equal(g1, g2)
    np.testing.assert_array_equal(piv32.astype(np.int64), piv64)


def test_argument_and_range_errors(gpu_lib, capfd):
    L = _lib.lib()
    a = np.zeros((8, 8))
    c = np.full((8, 8), 3.0)
    L.cblas_dgemm(0, TR["N"], TR["N"], 8, 8, 8, 1.0, P(a), 8, P(a), 8, 0.0, P(c), 8)
    L.cblas_dgemm(ROW, TR["N"], TR["T"], 8, 8, 16, 1.0, P(a), 8, P(a), 16, 0.0, P(c), 8)   # lda < K
    L.cblas_dtrsm(COL, SIDE["L"], 0, TR["N"], DIAG["N"], 8, 8, 1.0, P(a), 8, P(c), 8)
    big = 1 << 31
    L.dgemm_64_(C("N"), C("N"), I64(big), I64(8), I64(8), ctypes.byref(ctypes.c_double(1.0)), P(a), I64(big),
                P(a), I64(8), ctypes.byref(ctypes.c_double(0.0)), P(c), I64(big))
    info = ctypes.c_int64(0)
    L.dgetrf_64_(I64(8), I64(big), P(a), I64(8), P(np.zeros(8, np.int64)), ctypes.byref(info))
    assert info.value == -2
    cap = capfd.readouterr()
    text = cap.out + cap.err
    assert "cblas_dgemm parameter number  1" in text
    assert "cblas_dgemm parameter number  9" in text
    assert "cblas_dtrsm parameter number  3" in text
    assert "DGEMM  parameter number  3 exceeds the 32-bit range" in text
    assert np.all(c == 3.0)   # nothing was written
