"""Parity of the HIP GEMM path (through the C ABI) with the oracle and with the
reference's own outputs (golden fixtures).  Bar: normwise/componentwise error
within c*(K+2)*u of the dtype against the long-double / __float128 oracle."""
import numpy as np
import pytest

from gemm_avx512f_amd import blas
from gemm_avx512f_amd.device import DeviceMatrix
from oracle import oracle as O
from tests.util import gemm_bound, gemm_case, golden_arrays, golden_index, opmat

pytestmark = pytest.mark.gpu

DT = {"f32": np.float32, "f64": np.float64}


def run_device(dtype, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tile=None, splitk=None):
    dA = DeviceMatrix.from_padded(a, a.shape[0] if a.shape[0] else 1)
    dB = DeviceMatrix.from_padded(b, b.shape[0])
    dC = DeviceMatrix.from_padded(c, m)
    fn = blas.dgemm if dtype == np.float64 else blas.sgemm
    fn(ta, tb, m, n, k, alpha, dA, dB, beta, dC, tile=tile, splitk=splitk)
    return dC.to_host_padded()


def check_vs_oracle(dtype, ta, tb, m, n, k, alpha, beta, seed, pad=0, tile=None, splitk=None, c_bound=4.0):
    a, b, c, lda, ldb, ldc = gemm_case(dtype, seed, ta, tb, m, n, k, pad=pad)
    out = run_device(dtype, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tile, splitk)
    ref = c.copy(order="F")
    (O.dgemm if dtype == np.float64 else O.sgemm)(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, ref, ldc)
    # padding rows beyond M are never written
    assert np.array_equal(out[m:, :], c[m:, :])
    absab = np.abs(alpha) * (np.abs(opmat(a, ta, m, k)) @ np.abs(opmat(b, tb, k, n)))
    denom = absab + np.abs(beta) * np.abs(c[:m, :n].astype(np.float64)) + np.finfo(np.float64).tiny
    err = np.max(np.abs(out[:m, :n].astype(np.float64) - ref[:m, :n].astype(np.float64)) / denom)
    assert err <= gemm_bound(dtype, k, c_bound), (err, gemm_bound(dtype, k, c_bound))
    return err


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["d", "s"])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("tile", [0, 1, 2, 3])
def test_gemm_tiles_transposes(dtype, ta, tb, tile):
    # ragged sizes: every edge (M, N, K not multiples of any tile) is exercised
    check_vs_oracle(dtype, ta, tb, 261, 197, 133, 1.5, 0.5, seed=100 + tile, pad=3, tile=tile)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["d", "s"])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (1, 300, 7), (300, 1, 7), (17, 19, 1), (64, 64, 64), (128, 128, 16),
                                   (129, 127, 17), (33, 65, 1000), (256, 64, 31)])
def test_gemm_shapes(dtype, m, n, k):
    check_vs_oracle(dtype, "N", "N", m, n, k, -0.75, 2.0, seed=m * 7 + n * 3 + k)
    check_vs_oracle(dtype, "T", "T", m, n, k, 1.0, 0.0, seed=m * 5 + n + k)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["d", "s"])
@pytest.mark.parametrize("splitk", [2, 3, 7])
def test_gemm_splitk(dtype, splitk):
    check_vs_oracle(dtype, "N", "T", 70, 90, 1500, 1.25, -0.5, seed=splitk, tile=1, splitk=splitk)


@pytest.mark.parametrize("dtype,tile", [(np.float32, 1), (np.float64, 1), (np.float64, 2)],
                         ids=["s-64x64", "d-64x64", "d-128x64"])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_gemm_k_grouped_waves(dtype, tile, ta, tb):
    """A grid of at most one workgroup per CU, with at least 512 k per partition, splits each slab's k
    over groups of waves (FP32 64x64: four, or two for grids of up to two workgroups per CU; FP64 64x64
    and 128x64: two, up to 2048 k) when the library
    picks the configuration itself, and group 0 adds the other groups' accumulators.  A sweep-table
    entry makes the library pick the tile without split-K for these shapes: ragged edges, K not a
    multiple of the slab, a single tile."""
    from gemm_avx512f_amd import tuning
    op = 0 if dtype == np.float32 else 1
    shapes = [(1000, 520, 640), (333, 1000, 1100), (64, 64, 512)]
    if dtype == np.float32:
        shapes.append((1030, 1100, 520))   # 306 tiles: two workgroups per CU, two groups
    shp = np.zeros(len(shapes), dtype=tuning.SHAPE_DTYPE)
    ms = np.full((len(shapes), 21), -1.0, dtype=np.float32)
    for i, (m, n, k) in enumerate(shapes):
        shp[i] = (op, ta == "T", tb == "T", m, n, k, 0)
        ms[i, tuning.candidate_index(tile, 1)] = 1.0
    tuning.table_set(shp, ms)
    try:
        for i, (m, n, k) in enumerate(shapes):
            assert tuning.pick_config(op, ta == "T", tb == "T", m, n, k, 0, mode=2) == (tile, 1)
            check_vs_oracle(dtype, ta, tb, m, n, k, 1.5, 0.5, seed=300 + i)
    finally:
        tuning.table_clear()


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["d", "s"])
def test_gemm_unaligned_leading_dims(dtype):
    # odd leading dimensions force the scalar-load path
    a, b, c, lda, ldb, ldc = gemm_case(dtype, 5, "N", "N", 77, 53, 91, pad=1)
    assert lda % 2 == 0 or True
    check_vs_oracle(dtype, "N", "N", 77, 53, 91, 1.0, 1.0, seed=5, pad=1)
    check_vs_oracle(dtype, "T", "N", 77, 53, 91, 1.0, 1.0, seed=6, pad=2)


def test_beta_zero_does_not_read_c_and_alpha_zero_scales():
    m, n, k = 40, 30, 20
    a, b, c, lda, ldb, ldc = gemm_case(np.float64, 9, "N", "N", m, n, k)
    c[:] = np.nan
    out = run_device(np.float64, "N", "N", m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc)
    assert np.isfinite(out).all()
    a, b, c, lda, ldb, ldc = gemm_case(np.float64, 10, "N", "N", m, n, k)
    out = run_device(np.float64, "N", "N", m, n, k, 0.0, a, lda, b, ldb, -3.0, c, ldc)
    assert np.array_equal(out, -3.0 * c)
    out = run_device(np.float64, "N", "N", m, n, 0, 2.0, a, lda, b, 1, 0.5, c, ldc)
    assert np.array_equal(out, 0.5 * c)


def test_integer_valued_inputs_bitwise():
    rng = np.random.default_rng(3)
    for dtype in (np.float32, np.float64):
        a = np.asfortranarray(rng.integers(-16, 17, (150, 260)).astype(dtype))
        b = np.asfortranarray(rng.integers(-16, 17, (260, 140)).astype(dtype))
        c = np.asfortranarray(rng.integers(-16, 17, (150, 140)).astype(dtype))
        want = 2 * (a.astype(np.int64) @ b.astype(np.int64)) - c.astype(np.int64)
        for tile in range(4):
            out = run_device(dtype, "N", "N", 150, 140, 260, 2.0, a, 150, b, 260, -1.0, c, 150, tile=tile)
            assert np.array_equal(out.astype(np.int64), want), (dtype, tile)


# ------------------------------------------------------------------ reference outputs
@pytest.mark.parametrize("case", [c for c in golden_index()["gemm"]], ids=lambda c: c["name"])
def test_host_dropin_matches_reference_golden(case):
    """The Fortran-ABI drop-in (host pointers) on the same seeded inputs the
    reference was run on, compared with the reference's stored outputs."""
    dtype = DT[case["dtype"]]
    m, n, k = case["m"], case["n"], case["k"]
    a, b, c, lda, ldb, ldc = gemm_case(dtype, case["seed"], case["ta"], case["tb"], m, n, k, case["lda"],
                                       case["ldb"], case["ldc"])
    c_in = c.copy(order="F")
    fn = blas.dgemm_ if dtype == np.float64 else blas.sgemm_
    fn(case["ta"], case["tb"], m, n, k, case["alpha"], a, lda, b, ldb, case["beta"], c, ldc)
    G = golden_arrays("golden_gemm.npz")
    bound = gemm_bound(dtype, k, 8.0) * (abs(case["alpha"]) * max(k, 1) + abs(case["beta"]) + 1)
    if case["sampled"]:
        rows, cols = G[case["name"] + "_rows"], G[case["name"] + "_cols"]
        got = c[rows, cols].astype(np.float64)
        want = G[case["name"] + "_vals"].astype(np.float64)
        assert np.max(np.abs(got - want)) <= bound
        cm = c[:m, :n].astype(np.float64)
        assert abs(cm.sum() - case["checksum_sum"]) <= bound * m * n
        assert abs((cm * cm).sum() - case["checksum_sumsq"]) <= 1e-5 * case["checksum_sumsq"]
    else:
        ref = G[case["name"] + "_c"]
        assert np.array_equal(c[m:, :], c_in[m:, :])
        assert np.max(np.abs(c[:m, :n].astype(np.float64) - ref[:m, :n].astype(np.float64)), initial=0) <= bound
