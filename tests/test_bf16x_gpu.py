"""bf16-split SGEMM on the GPU: the split is bit-exact against the oracle's
restatement; bf16-representable inputs give the FP32 product of a single
bf16 GEMM; errors against the long-double oracle follow the split width
(~2^-16 for 3 products, FP32 level for 6)."""
import numpy as np
import pytest

from gemm_avx512f_amd import _lib, blas
from gemm_avx512f_amd.device import DeviceBuffer, DeviceMatrix
from oracle import oracle as O
from tests.util import gemm_case, opmat

pytestmark = pytest.mark.gpu


def pad64(x):
    return (x + 63) // 64 * 64


@pytest.mark.parametrize("nsplit", [1, 2, 3])
def test_split_bit_exact(nsplit):
    s = O.Stream(60 + nsplit)
    x = s.f32(130, 75)
    rows, k = 130, 75
    Rp, Kp = pad64(rows), pad64(k)
    dx = DeviceMatrix.from_host(x)
    dp = DeviceBuffer(nsplit * Rp * Kp * 2)
    _lib.check(_lib.lib().gemm_amd_bf16_split(dx.ptr, 1, rows, rows, k, nsplit, dp.ptr, Rp, Kp, None), "split")
    planes = dp.download(np.empty(nsplit * Rp * Kp, dtype=np.uint16)).reshape(nsplit, Rp, Kp)
    want = O.bf16_split(np.ascontiguousarray(x), nsplit).reshape(nsplit, rows, k)   # rows of x, k fastest
    assert np.array_equal(planes[:, :rows, :k], want)
    assert np.all(planes[:, rows:, :] == 0) and np.all(planes[:, :, k:] == 0)


def run(ta, tb, m, n, k, alpha, a, b, beta, c, nprod):
    dA, dB = DeviceMatrix.from_padded(a, a.shape[0]), DeviceMatrix.from_padded(b, b.shape[0])
    dC = DeviceMatrix.from_padded(c, m)
    blas.sgemm_bf16x(ta, tb, m, n, k, alpha, dA, dB, beta, dC, nprod=nprod)
    return dC.to_host_padded()


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("nprod", [3, 6])
def test_bf16x_vs_oracle(ta, tb, nprod):
    m, n, k = 277, 193, 301
    a, b, c, lda, ldb, ldc = gemm_case(np.float32, 70 + nprod, ta, tb, m, n, k, pad=2)
    out = run(ta, tb, m, n, k, 1.25, a, b, -0.5, c, nprod)
    assert np.array_equal(out[m:, :], c[m:, :])
    ref = np.asfortranarray(c.astype(np.float64))
    O.dgemm(ta, tb, m, n, k, 1.25, np.asfortranarray(a.astype(np.float64)), lda,
            np.asfortranarray(b.astype(np.float64)), ldb, -0.5, ref, ldc)
    absab = 1.25 * (np.abs(opmat(a, ta, m, k)) @ np.abs(opmat(b, tb, k, n))) + 0.5 * np.abs(c[:m, :n])
    err = np.max(np.abs(out[:m, :n] - ref[:m, :n]) / absab)
    bound = 4 * 2.0 ** -16 if nprod == 3 else 4 * (k + 2) * 2.0 ** -24
    assert err < bound, (err, bound)


def test_bf16_representable_inputs_exact_products():
    """Inputs that are bf16 values with small exponents: every split after the first is zero and the
    products are exact in FP32, so the result equals the exactly rounded FP32 product sum."""
    rng = np.random.default_rng(4)
    m, n, k = 96, 80, 64
    a = np.asfortranarray(rng.integers(-128, 128, (m, k)).astype(np.float32))   # 8-bit integers: exact in bf16
    b = np.asfortranarray(rng.integers(-128, 128, (k, n)).astype(np.float32))
    c = np.zeros((m, n), np.float32, order="F")
    want = (a.astype(np.int64) @ b.astype(np.int64)).astype(np.float64)           # |sum| < 2^24: exact in FP32
    for nprod in (3, 6):
        out = run("N", "N", m, n, k, 1.0, a, b, 0.0, c, nprod)
        assert np.array_equal(out.astype(np.float64), want), nprod


def test_error_ordering_and_checks():
    m = n = k = 512
    s = O.Stream(9)
    a, b = s.f32(m, k), s.f32(k, n)
    c = np.zeros((m, n), np.float32, order="F")
    ref = np.zeros((m, n), order="F")
    O.dgemm("N", "N", m, n, k, 1.0, np.asfortranarray(a.astype(np.float64)), m,
            np.asfortranarray(b.astype(np.float64)), k, 0.0, ref, m)
    e3 = O.err(np.asfortranarray(run("N", "N", m, n, k, 1.0, a, b, 0.0, c, 3).astype(np.float64)), ref)
    e6 = O.err(np.asfortranarray(run("N", "N", m, n, k, 1.0, a, b, 0.0, c, 6).astype(np.float64)), ref)
    assert e6["frob_rel"] < e3["frob_rel"] / 16
    assert e6["frob_rel"] < 1e-6
    L = _lib.lib()
    assert L.gemm_amd_sgemm_bf16x(b"N", b"N", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, 4, None) < 0
    assert L.gemm_amd_sgemm_bf16x(b"N", b"N", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 3, 3, None) == 13
