"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it:
known answers, the reference's own outputs (golden fixtures written from the
reference build), an independent implementation (numpy / scipy), closed forms
and exact integer cases.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.util import EPS, gemm_bound, gemm_case, golden_arrays, golden_index, opmat


# ------------------------------------------------------------------ RNG
def test_splitmix64_known_answers():
    # published SplitMix64 outputs for seed 0
    s = O.Stream(0)
    assert [s.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_fill_formulas_match_first_draws():
    x0 = 0xE220A8397B1DCDAF
    s = O.Stream(0)
    a = s.f64(1, 1)
    assert a[0, 0] == (x0 >> 11) * 2.0 ** -53 * 2 - 1
    s = O.Stream(0)
    f = s.f32(1, 1)
    assert f[0, 0] == np.float32((x0 >> 40) * 2.0 ** -24 * 2 - 1)


def test_fill_order_and_range():
    s = O.Stream(7)
    a = s.f64(5, 3, ld=8)
    assert np.all(a[5:, :] == 0)                       # padding untouched
    assert np.all((a[:5] >= -1) & (a[:5] < 1))
    t = O.Stream(7)                                   # column-major order: same stream as a 15 x 1 fill
    col = t.f64(15, 1)
    assert np.array_equal(a[:5, :3].reshape(-1, order="F"), col[:, 0])


def test_fill_range_exponents():
    s = O.Stream(3)
    x = s.f64(200, 50, r=8)
    e = np.frexp(x[x != 0])[1]
    assert e.min() >= -8 - 60 and e.max() <= 8
    assert (np.abs(x) < 2.0 ** 8).all()


# ------------------------------------------------------------------ GEMM vs the reference's own outputs
def _gemm_golden_cases():
    return [c for c in golden_index()["gemm"]]


@pytest.mark.parametrize("case", _gemm_golden_cases(), ids=lambda c: c["name"])
def test_oracle_gemm_matches_reference_build(case):
    dtype = np.float64 if case["dtype"] == "f64" else np.float32
    m, n, k = case["m"], case["n"], case["k"]
    if case["sampled"] and m * n * k > 2 ** 31:
        pytest.skip("full-size config checked by the sampled __float128 test below")
    a, b, c, lda, ldb, ldc = gemm_case(dtype, case["seed"], case["ta"], case["tb"], m, n, k, case["lda"],
                                       case["ldb"], case["ldc"])
    G = golden_arrays("golden_gemm.npz")
    fn = O.dgemm if dtype == np.float64 else O.sgemm
    if case["sampled"]:
        rows, cols = G[case["name"] + "_rows"], G[case["name"] + "_cols"]
        fn(case["ta"], case["tb"], m, n, k, case["alpha"], a, lda, b, ldb, case["beta"], c, ldc)
        got = c[rows, cols].astype(np.float64)
        want = G[case["name"] + "_vals"].astype(np.float64)
    else:
        c0 = c.copy(order="F")
        fn(case["ta"], case["tb"], m, n, k, case["alpha"], a, lda, b, ldb, case["beta"], c, ldc)
        ref = G[case["name"] + "_c"]
        assert np.array_equal(c[m:, :], c0[m:, :]) and np.array_equal(ref[m:, :], c0[m:, :])  # padding untouched
        got, want = c[:m, :n].astype(np.float64), ref[:m, :n].astype(np.float64)
    scale = np.abs(case["alpha"]) * k + abs(case["beta"]) + 1e-300
    assert np.max(np.abs(got - want)) <= gemm_bound(dtype, k) * scale


def test_oracle_q_matches_reference_at_config_size_sampled():
    """DGEMM 4096^3 NN/TN (alpha 1.5, beta 0.5): __float128 oracle on the sampled entries vs the reference."""
    G = golden_arrays("golden_gemm.npz")
    for case in [c for c in golden_index()["gemm"] if c["name"].startswith("config_dgemm")]:
        a, b, c, lda, ldb, ldc = gemm_case(np.float64, case["seed"], case["ta"], case["tb"], case["m"], case["n"],
                                           case["k"])
        rows, cols = G[case["name"] + "_rows"][:256], G[case["name"] + "_cols"][:256]
        hi, lo = O.dgemm_q(case["ta"], case["tb"], case["m"], case["n"], case["k"], case["alpha"], a, lda, b, ldb,
                           case["beta"], c, ldc, rows, cols)
        want = G[case["name"] + "_vals"][:256]
        assert np.max(np.abs(hi - want)) <= gemm_bound(np.float64, case["k"]) * case["k"]


# ------------------------------------------------------------------ GEMM closed forms / independent checks
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T"), ("c", "n")])
def test_oracle_gemm_vs_numpy(ta, tb):
    m, n, k = 37, 29, 41
    a, b, c, lda, ldb, ldc = gemm_case(np.float64, 99, ta, tb, m, n, k, pad=3)
    want = 0.75 * opmat(a, ta, m, k) @ opmat(b, tb, k, n) - 1.25 * c[:m, :n]
    O.dgemm(ta, tb, m, n, k, 0.75, a, lda, b, ldb, -1.25, c, ldc)
    assert np.max(np.abs(c[:m, :n] - want)) < 1e-13


def test_oracle_gemm_integer_inputs_exact():
    rng = np.random.default_rng(5)
    a = np.asfortranarray(rng.integers(-8, 9, (20, 30)).astype(np.float32))
    b = np.asfortranarray(rng.integers(-8, 9, (30, 25)).astype(np.float32))
    c = np.asfortranarray(rng.integers(-8, 9, (20, 25)).astype(np.float32))
    want = 2 * (a.astype(np.int64) @ b.astype(np.int64)) + 3 * c.astype(np.int64)
    O.sgemm("N", "N", 20, 25, 30, 2.0, a, 20, b, 30, 3.0, c, 20)
    assert np.array_equal(c.astype(np.int64), want)


def test_oracle_gemm_beta0_never_reads_c_and_transpose_identity():
    m, n, k = 9, 7, 11
    a, b, c, lda, ldb, ldc = gemm_case(np.float64, 4, "N", "N", m, n, k)
    c[:] = np.nan
    O.dgemm("N", "N", m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc)
    assert np.isfinite(c).all()
    # (A B)^T = B^T A^T with the same summation order -> bitwise
    ct = np.zeros((n, m), order="F")
    O.dgemm("T", "T", n, m, k, 1.0, b, ldb, a, lda, 0.0, ct, n)
    assert np.array_equal(ct, c.T)


def test_oracle_gemm_identity_and_quick_returns():
    n = 12
    eye = np.asfortranarray(np.eye(n))
    _, b, c, *_ = gemm_case(np.float64, 8, "N", "N", n, n, n)
    c0 = c.copy(order="F")
    O.dgemm("N", "N", n, n, n, 2.0, eye, n, b, n, 0.5, c, n)
    assert np.array_equal(c, 2.0 * b + 0.5 * c0)                    # exact: powers of two
    c1 = c0.copy(order="F")
    O.dgemm("N", "N", n, n, 0, 2.0, eye, n, b, n, 0.5, c1, n)       # K = 0 -> beta*C
    assert np.array_equal(c1, 0.5 * c0)
    c2 = c0.copy(order="F")
    O.dgemm("N", "N", n, n, n, 0.0, eye, n, b, n, -2.0, c2, n)      # alpha = 0 -> beta*C
    assert np.array_equal(c2, -2.0 * c0)


def test_oracle_q_agrees_with_long_double():
    m, n, k = 16, 12, 300
    a, b, c, lda, ldb, ldc = gemm_case(np.float64, 12, "T", "N", m, n, k)
    hi, lo = O.dgemm_q("T", "N", m, n, k, 1.5, a, lda, b, ldb, 0.5, c, ldc)
    O.dgemm("T", "N", m, n, k, 1.5, a, lda, b, ldb, 0.5, c, ldc)
    assert np.max(np.abs(hi - c)) <= 2 * EPS[np.float64] * np.max(np.abs(c))
    assert np.max(np.abs(lo)) <= EPS[np.float64] * np.max(np.abs(hi))


def test_blocked_port_matches_oracle():
    m, n, k = 130, 70, 300
    for ta, tb in [("N", "N"), ("T", "N"), ("N", "T")]:
        a, b, c, lda, ldb, ldc = gemm_case(np.float64, 77, ta, tb, m, n, k, pad=2)
        c2 = c.copy(order="F")
        O.dgemm(ta, tb, m, n, k, 1.5, a, lda, b, ldb, 0.5, c, ldc)
        O.dgemm_blocked(ta, tb, m, n, k, 1.5, a, lda, b, ldb, 0.5, c2, ldc, 4)
        assert np.max(np.abs(c - c2)) < gemm_bound(np.float64, k) * k


# ------------------------------------------------------------------ DTRSM
TRSM_VARIANTS = [(s, u, t, d) for s in "LR" for u in "UL" for t in "NT" for d in "NU"]


def _tri_system(side, uplo, m, n, seed, diag_boost):
    na = m if side == "L" else n
    s = O.Stream(seed)
    a = s.f64(na, na)
    for i in range(na):
        a[i, i] = diag_boost + abs(a[i, i])
    # poison the triangle that must not be read
    mask = np.triu(np.ones((na, na), bool), 1) if uplo == "L" else np.tril(np.ones((na, na), bool), -1)
    a[mask] = np.nan
    b = s.f64(m, n)
    return a, b


@pytest.mark.parametrize("side,uplo,transa,diag", TRSM_VARIANTS)
def test_oracle_dtrsm_residual(side, uplo, transa, diag):
    m, n = 23, 17
    a, b = _tri_system(side, uplo, m, n, 31, 8.0)
    x = b.copy(order="F")
    O.dtrsm(side, uplo, transa, diag, m, n, 1.5, a, a.shape[0], x, m)
    assert np.isfinite(x).all()
    tri = np.nan_to_num(a, nan=0.0)
    if diag == "U":
        np.fill_diagonal(tri, 1.0)
    op = tri if transa == "N" else tri.T
    lhs = op @ x if side == "L" else x @ op
    assert np.max(np.abs(lhs - 1.5 * b)) < 1e-13
    # independent library solve
    from scipy.linalg import solve_triangular
    if side == "L":
        want = solve_triangular(tri, 1.5 * b, lower=(uplo == "L"), trans=transa, unit_diagonal=(diag == "U"))
    else:
        want = solve_triangular(tri, (1.5 * b).T, lower=(uplo == "L"), trans=("T" if transa == "N" else "N"),
                                unit_diagonal=(diag == "U")).T
    assert np.max(np.abs(x - want)) < 1e-12


def test_oracle_dtrsm_diagonal_exact():
    m, n = 8, 5
    d = 2.0 ** np.arange(-3, 5)
    a = np.asfortranarray(np.diag(d))
    _, b = _tri_system("L", "L", m, n, 3, 1.0)
    x = b.copy(order="F")
    O.dtrsm("L", "L", "N", "N", m, n, 1.0, a, m, x, m)
    assert np.array_equal(x, b / d[:, None])


def test_oracle_dtrsm_kernel_matches_reference_kernels():
    """The kernel-level restatement reproduces all 62 reference solvers (sb write-back included)."""
    G = golden_arrays("golden_trsm_kernel.npz")
    from tests.golden.make_golden import trsm_kernel_inputs
    for case in golden_index()["trsm_kernel"]:
        sa, sb, C, K = trsm_kernel_inputs(case["variant"], case["M"], case["N"], case["seed"])
        O.dtrsm_kernel(case["variant"], case["M"], case["N"], sa, sb, C, case["LDB"], case["LDC"])
        scale = max(1.0, np.max(np.abs(G[case["name"] + "_c"])))
        assert np.max(np.abs(C - G[case["name"] + "_c"])) <= 1e-13 * scale, case["name"]
        assert np.max(np.abs(sa - G[case["name"] + "_sa"])) <= 1e-13 * scale, case["name"]
        assert np.max(np.abs(sb - G[case["name"] + "_sb"])) <= 1e-13 * scale, case["name"]


# ------------------------------------------------------------------ complex
def _op(x, t):
    return x if t == "N" else (x.T if t == "T" else x.conj().T)


def test_oracle_zgemm_vs_numpy_and_real_special_case():
    m, n, k = 11, 9, 13
    s = O.Stream(17)
    alpha, beta = 0.5 - 1.25j, 0.75 + 0.5j
    for ta, tb in [("N", "N"), ("C", "N"), ("T", "C"), ("N", "T")]:
        a = s.c128(m, k) if ta == "N" else s.c128(k, m)
        b = s.c128(k, n) if tb == "N" else s.c128(n, k)
        c = s.c128(m, n)
        want = alpha * (_op(a, ta) @ _op(b, tb)) + beta * c
        O.zgemm(ta, tb, m, n, k, alpha, a, a.shape[0], b, b.shape[0], beta, c, m)
        assert np.max(np.abs(c - want)) < 1e-13
        cf = s.c64(m, n)
        af, bf = a.astype(np.complex64), b.astype(np.complex64)
        wantf = alpha * (_op(af.astype(np.complex128), ta) @ _op(bf.astype(np.complex128), tb)) + beta * cf
        O.cgemm(ta, tb, m, n, k, alpha, np.asfortranarray(af), m if ta == "N" else k, np.asfortranarray(bf),
                k if tb == "N" else n, beta, cf, m)
        assert np.max(np.abs(cf - wantf)) < 1e-5
    # real-only operands and scalars -> the real GEMM, imaginary part exactly 0
    a, b = s.c128(m, k), s.c128(k, n)
    ar, br = np.asfortranarray(a.real + 0j), np.asfortranarray(b.real + 0j)
    cz = np.zeros((m, n), np.complex128, order="F")
    O.zgemm("N", "N", m, n, k, 1.5 + 0j, ar, m, br, k, 0j, cz, m)
    cd = np.zeros((m, n), order="F")
    O.dgemm("N", "N", m, n, k, 1.5, np.asfortranarray(a.real), m, np.asfortranarray(b.real), k, 0.0, cd, m)
    assert np.array_equal(cz.real, cd) and np.all(cz.imag == 0)


# ------------------------------------------------------------------ emulation splits
def test_ozaki_split_reconstructs_and_is_exact_for_short_mantissas():
    s = O.Stream(9)
    x = s.f64(30, 40, r=8)
    for ns in (2, 4, 6, 8):
        sl, ex = O.ozaki_split(x, ns, by="rows")
        assert sl.min() >= -64 and sl.max() <= 64
        w = 2.0 ** (1 - 7 * np.arange(1, ns + 1))
        recon = np.einsum("t,trk->rk", w, sl.astype(np.float64)) * (2.0 ** ex)[:, None]
        assert np.all(np.abs(recon - x) <= 2.0 ** (ex[:, None] - 7 * ns))
        amax = np.max(np.abs(x), axis=1)
        assert np.all((amax < 2.0 ** ex) & (amax >= 2.0 ** (ex - 1)))
    # round-to-nearest slicing: the residual of a random row is unbiased (a
    # truncating split would leave a residual with the sign of x)
    big = s.f64(1, 20000)
    sl, ex = O.ozaki_split(big, 3, by="rows")
    recon = np.einsum("t,trk->rk", 2.0 ** (1 - 7 * np.arange(1, 4)), sl.astype(np.float64)) * (2.0 ** ex)[:, None]
    res = (big - recon) * np.sign(big)
    assert abs(res.mean()) < 0.1 * np.abs(res).mean()
    # values with <= 6 + 7*(ns-1) significant bits below the row maximum reconstruct exactly
    y = np.asfortranarray(np.round(s.f64(8, 16) * 2 ** 20) / 2 ** 20)
    sl, ex = O.ozaki_split(y, 4, by="rows")
    recon = np.einsum("t,trk->rk", 2.0 ** (1 - 7 * np.arange(1, 5)), sl.astype(np.float64)) * (2.0 ** ex)[:, None]
    assert np.array_equal(recon, y)
    # column slicing of x equals row slicing of x^T
    slc, exc = O.ozaki_split(x, 5, by="cols")
    slr, exr = O.ozaki_split(np.asfortranarray(x.T), 5, by="rows")
    assert np.array_equal(slc, slr) and np.array_equal(exc, exr)


def test_bf16_rounding_known_values_and_split_exactness():
    cases = [(1.0, 0x3F80), (1.0 + 2 ** -8, 0x3F80), (1.0 + 3 * 2 ** -8, 0x3F82), (-2.0, 0xC000), (0.0, 0x0000),
             (-0.0, 0x8000), (65536.0 * 1.00390625, 0x4780)]
    for v, h in cases:
        assert O.lib().oracle_f32_to_bf16(v) == h, v
    s = O.Stream(2)
    x = s.f32(4096, 1)[:, 0]
    p2 = O.bf16_split(x, 2)
    hi, lo = O.bf16_to_f32(p2[0]), O.bf16_to_f32(p2[1])
    assert np.all(np.abs((hi.astype(np.float64) + lo) - x) <= np.abs(x) * 2.0 ** -16)
    p3 = O.bf16_split(x, 3)
    r3 = O.bf16_to_f32(p3[0]).astype(np.float64) + O.bf16_to_f32(p3[1]) + O.bf16_to_f32(p3[2])
    assert np.all(np.abs(r3 - x) <= np.abs(x) * 2.0 ** -24)


def test_ozaki_crt_split_residues_rebuild_the_scaled_integers():
    """Residues of the modular form: each is the symmetric residue of the row's scaled integer, and the
    Chinese remainder theorem over all 16 moduli (exact Python integers) gives that integer back."""
    from math import prod
    s = O.Stream(10)
    x = s.f64(20, 30, r=8)
    mods = O.CRT_MODULI
    M = prod(mods)
    for bits in (20, 48, 56):
        res, ex = O.ozaki_crt_split(x, 16, bits, by="rows")
        amax = np.max(np.abs(x), axis=1)
        assert np.all((amax < 2.0 ** (ex + bits)) & (amax >= 2.0 ** (ex + bits - 1)))
        for r in range(0, 20, 7):
            for p in range(0, 30, 11):
                v = float(x[r, p]) * 2.0 ** (bits - (int(ex[r]) + bits))      # x * 2^(bits - e_r)
                want = int(np.rint(v))
                assert abs(want) <= 2 ** bits
                acc = 0
                for l, m in enumerate(mods):
                    q = int(res[l, r, p])
                    assert -m // 2 <= q <= m // 2 or (m == 256 and q == -128)
                    assert (want - q) % m == 0
                    Ml = M // m
                    acc += q * Ml * pow(Ml, -1, m)
                acc %= M
                if acc > M // 2:
                    acc -= M
                assert acc == want
    rc, exc = O.ozaki_crt_split(x, 7, 40, by="cols")
    rr, exr = O.ozaki_crt_split(np.asfortranarray(x.T), 7, 40, by="rows")
    assert np.array_equal(rc, rr) and np.array_equal(exc, exr)


def test_err_metrics_known_answers():
    # hand-computed: c - r = (0, -0.5, 1); max|r| = 2.5; ||c - r|| = sqrt(1.25), ||r|| = sqrt(11.25)
    c = np.array([1.0, 2.0, 3.0])
    r = np.array([1.0, 2.5, 2.0])
    e = O.err(c, r)
    assert e["max_rel"] == pytest.approx(0.4, rel=1e-15)
    assert e["frob_rel"] == pytest.approx(1.0 / 3.0, rel=1e-15)
    assert e["max_elem_rel"] == pytest.approx(0.5, rel=1e-15)
    e32 = O.err(c.astype(np.float32), r)
    assert (e32["max_rel"], e32["frob_rel"], e32["max_elem_rel"]) == pytest.approx((0.4, 1.0 / 3.0, 0.5), rel=1e-15)
    # the low word of a double-double reference is added, not subtracted or dropped
    hi = np.array([1.0, 2.0, 4.0])
    lo = np.array([2.0 ** -60, 0.0, -(2.0 ** -55)])
    exact = O.err(hi, hi, lo)
    assert exact["max_rel"] == pytest.approx(2.0 ** -55 / 4.0, rel=1e-12)
    assert O.err(hi + 0.0, hi, None)["max_rel"] == 0.0
    # an all-zero reference reports the absolute error
    assert O.err(np.array([0.0, 0.25]), np.zeros(2))["max_rel"] == 0.25
