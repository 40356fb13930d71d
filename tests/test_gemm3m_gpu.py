"""cgemm3m_/zgemm3m_ on the GPU against the 4M long-double complex oracle.

Bound: the 3M imaginary part carries an extra error term proportional to
|Re A + Im A| |Re B + Im B| (Higham, "Stability of a method for multiplying
complex matrices with three real matrix multiplications", 1992), so the test
normalises by (|Ar|+|Ai|)(|Br|+|Bi|)."""
import numpy as np
import pytest

from gemm_avx512f_amd import blas
from gemm_avx512f_amd.device import DeviceMatrix
from oracle import oracle as O

pytestmark = pytest.mark.gpu

OPS = [("N", "N"), ("T", "N"), ("N", "C"), ("C", "T"), ("C", "C")]


def op(x, t):
    return x if t == "N" else (x.T if t == "T" else x.conj().T)


def case(dtype, seed, ta, tb, m, n, k, pad):
    s = O.Stream(seed)
    mk = s.c128 if dtype == np.complex128 else s.c64
    a = mk(m, k, m + pad) if ta == "N" else mk(k, m, k + pad)
    b = mk(k, n, k + pad) if tb == "N" else mk(n, k, n + pad)
    c = mk(m, n, m + pad)
    return a, b, c


def check(dtype, ta, tb, m, n, k, alpha, beta, seed, pad=0, host=False):
    a, b, c = case(dtype, seed, ta, tb, m, n, k, pad)
    ref = c.copy(order="F")
    (O.zgemm if dtype == np.complex128 else O.cgemm)(ta, tb, m, n, k, alpha, a, a.shape[0], b, b.shape[0], beta, ref,
                                                     c.shape[0])
    if host:
        out = c.copy(order="F")
        (blas.zgemm3m_ if dtype == np.complex128 else blas.cgemm3m_)(ta, tb, m, n, k, alpha, a, a.shape[0], b,
                                                                    b.shape[0], beta, out, c.shape[0])
    else:
        dA, dB = DeviceMatrix.from_padded(a, a.shape[0]), DeviceMatrix.from_padded(b, b.shape[0])
        dC = DeviceMatrix.from_padded(c, m)
        (blas.zgemm3m if dtype == np.complex128 else blas.cgemm3m)(ta, tb, m, n, k, alpha, dA, dB, beta, dC)
        out = dC.to_host_padded()
    assert np.array_equal(out[m:, :], c[m:, :])
    A, B = op(a[: (m if ta == "N" else k), :], ta), op(b[: (k if tb == "N" else n), :], tb)
    A, B = A[:m, :k], B[:k, :n]
    bound_mat = abs(alpha) * ((np.abs(A.real) + np.abs(A.imag)) @ (np.abs(B.real) + np.abs(B.imag))) \
        + abs(beta) * np.abs(c[:m, :n]) + 1e-300
    u = 2.0 ** -53 if dtype == np.complex128 else 2.0 ** -24
    err = np.max(np.abs(out[:m, :n] - ref[:m, :n]) / bound_mat)
    assert err <= 8 * (k + 4) * u, err


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64], ids=["z", "c"])
@pytest.mark.parametrize("ta,tb", OPS)
def test_gemm3m_ops(dtype, ta, tb):
    check(dtype, ta, tb, 131, 77, 95, 0.75 - 0.5j, -0.25 + 1.0j, seed=len(ta + tb) * 3 + (ta == "C"), pad=2)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64], ids=["z", "c"])
def test_gemm3m_host_dropin_and_special_scalars(dtype):
    check(dtype, "N", "N", 64, 48, 40, 1.0, 0.0, seed=5, host=True)
    check(dtype, "C", "N", 33, 17, 300, 2.0 + 0j, 0.5j, seed=6, host=True)
    check(dtype, "N", "T", 20, 30, 0, 1.0, 2.0 - 1j, seed=7)             # K = 0 -> beta*C


def test_zgemm3m_real_inputs_give_real_result():
    m, n, k = 50, 40, 60
    s = O.Stream(8)
    a = np.asfortranarray(s.f64(m, k) + 0j)
    b = np.asfortranarray(s.f64(k, n) + 0j)
    dA, dB = DeviceMatrix.from_host(a), DeviceMatrix.from_host(b)
    dC = DeviceMatrix(m, n, np.complex128)
    blas.zgemm3m("N", "N", m, n, k, 1.0, dA, dB, 0.0, dC)
    out = dC.to_host()
    assert np.all(out.imag == 0)       # T3 - T1 - T2 with Ai = Bi = 0: identical products cancel exactly
    cd = np.zeros((m, n), order="F")
    O.dgemm("N", "N", m, n, k, 1.0, np.asfortranarray(a.real), m, np.asfortranarray(b.real), k, 0.0, cd, m)
    assert np.max(np.abs(out.real - cd)) < 1e-13 * k
