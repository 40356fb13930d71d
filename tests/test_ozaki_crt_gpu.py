"""Emulated DGEMM, modular (CRT) form, on the GPU.

* the residue step is byte-exact against the oracle's restatement;
* integer-valued inputs that the scaling keeps exact give the exact product
  (the CRT rebuild is exact);
* against the __float128 oracle the error is bounded by the kept bits
  (2^-bits of the row and column maxima), and with 16 moduli it is no worse
  than the native FP64 kernel's on uniform inputs."""
import numpy as np
import pytest

from gemm_avx512f_amd import _lib, blas
from gemm_avx512f_amd.device import DeviceBuffer, DeviceMatrix
from oracle import oracle as O
from tests.util import gemm_case, opmat

pytestmark = pytest.mark.gpu


def pad128(x):
    return (x + 127) // 128 * 128


def bits_for(nmod, k):
    return _lib.lib().gemm_amd_ozaki_crt_bits(nmod, k)


def test_bits_rule():
    L = _lib.lib()
    assert L.gemm_amd_ozaki_crt_bits(16, 8192) == 55
    assert L.gemm_amd_ozaki_crt_bits(15, 8192) == 51
    assert L.gemm_amd_ozaki_crt_bits(16, 16) == 56          # capped
    assert L.gemm_amd_ozaki_crt_bits(1, 100) == -1


@pytest.mark.parametrize("by", ["rows", "cols"])
@pytest.mark.parametrize("nmod", [4, 15, 16])
def test_residues_byte_exact(by, nmod):
    s = O.Stream(50 + nmod)
    x = s.f64(150, 200, r=8)
    rows, k = (150, 200) if by == "rows" else (200, 150)
    rs, ks = (1, 150) if by == "rows" else (150, 1)
    bits = min(56, 7 * nmod)
    want, want_ex = O.ozaki_crt_split(x, nmod, bits, by=by)
    dx = DeviceMatrix.from_host(x)
    Rp, Kp = pad128(rows), pad128(k)
    dres = DeviceBuffer(nmod * Rp * Kp)
    dex = DeviceBuffer(Rp * 4)
    _lib.check(_lib.lib().gemm_amd_ozaki_crt_split(dx.ptr, rs, ks, rows, k, nmod, bits, dres.ptr, Rp, Kp, dex.ptr,
                                                   None), "crt split")
    res = dres.download(np.empty(nmod * Rp * Kp, dtype=np.int8)).reshape(nmod, Rp, Kp)
    ex = dex.download(np.empty(Rp, dtype=np.int32))
    assert np.array_equal(res[:, :rows, :k], want)
    assert np.all(res[:, rows:, :] == 0) and np.all(res[:, :, k:] == 0)
    assert np.array_equal(ex[:rows], want_ex)


def run_crt(ta, tb, m, n, k, alpha, a, b, beta, c, nmod):
    dA, dB = DeviceMatrix.from_padded(a, a.shape[0]), DeviceMatrix.from_padded(b, b.shape[0])
    dC = DeviceMatrix.from_padded(c, m)
    blas.dgemm_ozaki_crt(ta, tb, m, n, k, alpha, dA, dB, beta, dC, nmoduli=nmod)
    return dC.to_host_padded()


def run_native(ta, tb, m, n, k, alpha, a, b, beta, c):
    dA, dB = DeviceMatrix.from_padded(a, a.shape[0]), DeviceMatrix.from_padded(b, b.shape[0])
    dC = DeviceMatrix.from_padded(c, m)
    blas.dgemm(ta, tb, m, n, k, alpha, dA, dB, beta, dC)
    return dC.to_host_padded()


def test_integer_inputs_exact():
    rng = np.random.default_rng(2)
    m, n, k = 300, 270, 700
    a = np.asfortranarray(rng.integers(-2 ** 20, 2 ** 20, (m, k)).astype(np.float64))
    b = np.asfortranarray(rng.integers(-2 ** 20, 2 ** 20, (k, n)).astype(np.float64))
    c = np.asfortranarray(rng.integers(-50, 50, (m, n)).astype(np.float64))
    exact = a.astype(object) @ b.astype(object)
    # |A B| < 700 * 2^40: 2*A*B - C is an integer below 2^53, exactly representable
    want = np.array(exact * 2 - c.astype(np.int64).astype(object), dtype=np.float64)
    for nmod in (8, 16):        # 8 moduli keep 26 bits per operand at K = 700, >= the 21 the inputs need
        assert bits_for(nmod, k) >= 21
        out = run_crt("N", "N", m, n, k, 2.0, a, b, -1.0, c, nmod)
        assert np.array_equal(out, want), nmod


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_transposes_and_edges_vs_oracle(ta, tb):
    m, n, k = 261, 197, 333
    a, b, c, lda, ldb, ldc = gemm_case(np.float64, 43, ta, tb, m, n, k, pad=3)
    out = run_crt(ta, tb, m, n, k, 1.5, a, b, 0.5, c, 16)
    assert np.array_equal(out[m:, :], c[m:, :])
    hi, lo = O.dgemm_q(ta, tb, m, n, k, 1.5, a, lda, b, ldb, 0.5, c, ldc)
    e = O.err(np.asfortranarray(out[:m, :n]), hi, lo)
    assert e["frob_rel"] < 4e-16 and e["max_rel"] < 2e-15, e


@pytest.mark.parametrize("nmod", [10, 13])
def test_error_bounded_by_kept_bits(nmod):
    """|C - C_exact| <= 2.1 K 2^-bits max_k|a_ik| max_k|b_kj| (+ rounding of the result)."""
    m, n, k = 200, 150, 1000
    a, b, c, lda, ldb, ldc = gemm_case(np.float64, 44 + nmod, "N", "T", m, n, k)
    bits = bits_for(nmod, k)
    out = run_crt("N", "T", m, n, k, 1.0, a, b, 0.0, c, nmod)
    hi, lo = O.dgemm_q("N", "T", m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc)
    A, B = opmat(a, "N", m, k), opmat(b, "T", k, n)
    bound = 2.1 * k * 2.0 ** -bits * np.outer(np.abs(A).max(axis=1), np.abs(B).max(axis=0)) + 4 * 2.0 ** -53 * np.abs(hi)
    assert np.all(np.abs((out[:m, :n] - hi) - lo) <= bound)


@pytest.mark.parametrize("r", [0, 8, 32])
def test_dynamic_range_error_vs_native(r):
    """A, B entries u*2^e with e in [-r, r] against __float128: 16 moduli are no worse than FP64 at r = 0."""
    m = n = k = 512
    s = O.Stream(100 + r)
    a, b = s.f64(m, k, r=r), s.f64(k, n, r=r)
    c = np.zeros((m, n), order="F")
    hi, lo = O.dgemm_q("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    nat = O.err(np.asfortranarray(run_native("N", "N", m, n, k, 1.0, a, b, 0.0, c)), hi, lo)
    errs = {nm: O.err(np.asfortranarray(run_crt("N", "N", m, n, k, 1.0, a, b, 0.0, c, nm)), hi, lo)
            for nm in (12, 14, 16)}
    for nm, e in errs.items():
        bits = bits_for(nm, k)
        assert e["max_rel"] < 4.0 * k * 2.0 ** -bits * 2.0 ** (2 * r), (nm, e)
    if r == 0:
        assert errs[16]["frob_rel"] <= nat["frob_rel"] and errs[16]["max_rel"] <= nat["max_rel"]
    assert errs[16]["frob_rel"] <= errs[14]["frob_rel"] <= errs[12]["frob_rel"]


def test_long_k_reduction_and_checks():
    """K beyond one exact int32 chunk (131008 terms) exercises the in-kernel modular fold."""
    m, n, k = 64, 48, 140000
    s = O.Stream(7)
    a, b = s.f64(m, k), s.f64(k, n)
    c = np.zeros((m, n), order="F")
    out = run_crt("N", "N", m, n, k, 1.0, a, b, 0.0, c, 16)
    rows, cols = np.arange(0, 64, 7), np.arange(0, 48, 5)          # 10 sampled entries
    hi, lo = O.dgemm_q("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, rows=rows, cols=cols)
    # sqrt(K)-sized sums of uniform products: FP64-level agreement
    assert np.max(np.abs(out[rows, cols] - hi)) <= 1e-13 * np.max(np.abs(hi))
    L = _lib.lib()
    assert L.gemm_amd_dgemm_ozaki_crt(b"N", b"N", 4, 4, 4, 1.0, None, 3, None, 4, 0.0, None, 4, 16, None) == 8
    assert L.gemm_amd_dgemm_ozaki_crt(b"N", b"N", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, 17, None) < 0
    out = run_crt("N", "N", 40, 30, 20, 0.0, a[:40, :20].copy(order="F"), b[:20, :30].copy(order="F"), 2.0,
                  np.ones((40, 30), order="F"), 16)
    assert np.all(out == 2.0)


def test_auto_counts_match_explicit():
    """nmoduli = 0 / nslices = 0 take the tuned counts for K (16 moduli, 8
    slices at K = 1500) and give the explicit calls' bits."""
    from gemm_avx512f_amd import tuning
    m, n, k = 96, 80, 1500
    a, b, c, lda, ldb, ldc = gemm_case(np.float64, 71, "N", "N", m, n, k)
    nm, ns = tuning.emulation_counts(k)
    assert (nm, ns) == (16, 8)
    assert np.array_equal(run_crt("N", "N", m, n, k, 1.0, a, b, 0.5, c, 0), run_crt("N", "N", m, n, k, 1.0, a, b, 0.5, c, nm))
    dA, dB = DeviceMatrix.from_padded(a, a.shape[0]), DeviceMatrix.from_padded(b, b.shape[0])
    outs = []
    for count in (0, ns):
        dC = DeviceMatrix.from_padded(c, m)
        blas.dgemm_ozaki("N", "N", m, n, k, 1.0, dA, dB, 0.5, dC, nslices=count)
        outs.append(dC.to_host_padded())
    assert np.array_equal(outs[0], outs[1])
