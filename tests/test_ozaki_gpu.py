"""Emulated DGEMM (int8 slices) on the GPU.

* the split step is byte-exact against the oracle's restatement;
* integer-valued inputs (one slice holds them exactly) give the exact product;
* on uniform and wide-dynamic-range inputs the error against the __float128
  oracle is bounded by the slicing resolution 2^(-7s) and, at s >= 7, is no
  worse than the native FP64 kernel's."""
import ctypes

import numpy as np
import pytest

from gemm_avx512f_amd import _lib, blas
from gemm_avx512f_amd.device import DeviceBuffer, DeviceMatrix
from oracle import oracle as O
from tests.util import gemm_case

pytestmark = pytest.mark.gpu


def pad128(x):
    return (x + 127) // 128 * 128


@pytest.mark.parametrize("by", ["rows", "cols"])
@pytest.mark.parametrize("ns", [4, 7, 8])
def test_split_is_byte_exact(by, ns):
    s = O.Stream(30 + ns)
    x = s.f64(150, 200, r=8)        # 150 x 200, column-major, wide dynamic range
    rows, k = (150, 200) if by == "rows" else (200, 150)
    rs, ks = (1, 150) if by == "rows" else (150, 1)
    want_sl, want_ex = O.ozaki_split(x, ns, by=by)
    dx = DeviceMatrix.from_host(x)
    Rp, Kp = pad128(rows), pad128(k)
    dsl = DeviceBuffer(ns * Rp * Kp)
    dex = DeviceBuffer(Rp * 4)
    _lib.check(_lib.lib().gemm_amd_ozaki_split(dx.ptr, rs, ks, rows, k, ns, dsl.ptr, Rp, Kp, dex.ptr, None), "split")
    sl = dsl.download(np.empty(ns * Rp * Kp, dtype=np.int8)).reshape(ns, Rp, Kp)
    ex = dex.download(np.empty(Rp, dtype=np.int32))
    assert np.array_equal(sl[:, :rows, :k], want_sl)
    assert np.all(sl[:, rows:, :] == 0) and np.all(sl[:, :, k:] == 0)
    assert np.array_equal(ex[:rows], want_ex)


def run_oz(ta, tb, m, n, k, alpha, a, b, beta, c, ns):
    dA, dB = DeviceMatrix.from_padded(a, a.shape[0]), DeviceMatrix.from_padded(b, b.shape[0])
    dC = DeviceMatrix.from_padded(c, m)
    blas.dgemm_ozaki(ta, tb, m, n, k, alpha, dA, dB, beta, dC, nslices=ns)
    return dC.to_host_padded()


def run_native(ta, tb, m, n, k, alpha, a, b, beta, c):
    dA, dB = DeviceMatrix.from_padded(a, a.shape[0]), DeviceMatrix.from_padded(b, b.shape[0])
    dC = DeviceMatrix.from_padded(c, m)
    blas.dgemm(ta, tb, m, n, k, alpha, dA, dB, beta, dC)
    return dC.to_host_padded()


def test_integer_inputs_exact():
    rng = np.random.default_rng(1)
    m, n, k = 200, 170, 300
    a = np.asfortranarray(rng.integers(-127, 128, (m, k)).astype(np.float64))
    b = np.asfortranarray(rng.integers(-127, 128, (k, n)).astype(np.float64))
    c = np.asfortranarray(rng.integers(-50, 50, (m, n)).astype(np.float64))
    want = (a.astype(np.int64) @ b.astype(np.int64)) * 2 - c.astype(np.int64)
    # with s >= 3 every slice product that is nonzero for 8-bit integers is in the t+u <= s+1 set
    for ns in (3, 7):
        out = run_oz("N", "N", m, n, k, 2.0, a, b, -1.0, c, ns)
        assert np.array_equal(out, want.astype(np.float64)), ns


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_transposes_and_edges_vs_oracle(ta, tb):
    m, n, k = 261, 197, 333
    a, b, c, lda, ldb, ldc = gemm_case(np.float64, 41, ta, tb, m, n, k, pad=3)
    out = run_oz(ta, tb, m, n, k, 1.5, a, b, 0.5, c, 8)
    assert np.array_equal(out[m:, :], c[m:, :])
    hi, lo = O.dgemm_q(ta, tb, m, n, k, 1.5, a, lda, b, ldb, 0.5, c, ldc)
    e = O.err(np.asfortranarray(out[:m, :n]), hi, lo)
    assert e["frob_rel"] < 4e-16 and e["max_rel"] < 2e-15, e


@pytest.mark.parametrize("r", [0, 8, 32])
def test_dynamic_range_error_vs_native(r):
    """A, B entries u*2^e with e in [-r, r]: error of s=7,8 emulation vs the FP64 kernel's, against __float128."""
    m = n = k = 512
    s = O.Stream(100 + r)
    a, b = s.f64(m, k, r=r), s.f64(k, n, r=r)
    c = np.zeros((m, n), order="F")
    hi, lo = O.dgemm_q("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    nat = O.err(np.asfortranarray(run_native("N", "N", m, n, k, 1.0, a, b, 0.0, c)), hi, lo)
    errs = {}
    for ns in (4, 6, 7, 8):
        errs[ns] = O.err(np.asfortranarray(run_oz("N", "N", m, n, k, 1.0, a, b, 0.0, c, ns)), hi, lo)
    # resolution bound: per-row/column slicing keeps 7s bits below each row/column maximum
    for ns, e in errs.items():
        assert e["max_rel"] < 4.0 * k * 2.0 ** (-7 * ns) * 2.0 ** (2 * r), (ns, e)
    if r == 0:
        assert errs[8]["frob_rel"] <= nat["frob_rel"] and errs[8]["max_rel"] <= nat["max_rel"]
        # 7 slices keep ~49 bits below each row maximum: an order of magnitude short of FP64
        assert errs[7]["frob_rel"] <= 16 * nat["frob_rel"]
    # more slices never hurt
    assert errs[8]["frob_rel"] <= errs[6]["frob_rel"] <= errs[4]["frob_rel"]


def test_alpha_zero_k_zero_and_argument_checks():
    m, n, k = 40, 30, 20
    a, b, c, *_ = gemm_case(np.float64, 3, "N", "N", m, n, k)
    out = run_oz("N", "N", m, n, k, 0.0, a, b, 2.0, c, 7)
    assert np.array_equal(out, 2.0 * c)
    L = _lib.lib()
    assert L.gemm_amd_dgemm_ozaki(b"N", b"N", 4, 4, 4, 1.0, None, 3, None, 4, 0.0, None, 4, 7, None) == 8
    assert L.gemm_amd_dgemm_ozaki(b"N", b"N", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, 9, None) < 0
