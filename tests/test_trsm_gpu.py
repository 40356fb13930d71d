"""DTRSM on the GPU through the C ABI: all 16 BLAS variants against the oracle
(residual and forward error), the config-size L/L/N/N solve, NaN poisoning of
the unread triangle, and the kernel-level solvers against the reference's own
outputs (all 62 dtrsm_solve_* kernels)."""
import numpy as np
import pytest

from gemm_avx512f_amd import blas
from gemm_avx512f_amd.device import DeviceBuffer, DeviceMatrix
from oracle import oracle as O
from tests.util import golden_arrays, golden_index

pytestmark = pytest.mark.gpu

VARIANTS = [(s, u, t, d) for s in "LR" for u in "UL" for t in "NT" for d in "NU"]


def tri_system(side, uplo, m, n, seed, boost, pad=0):
    na = m if side == "L" else n
    s = O.Stream(seed)
    a = s.f64(na, na, ld=na + pad)
    for i in range(na):
        a[i, i] = boost + abs(a[i, i])
    mask = np.triu(np.ones((na, na), bool), 1) if uplo == "L" else np.tril(np.ones((na, na), bool), -1)
    blk = a[:na, :na]
    blk[mask] = np.nan                      # the triangle that must never be read
    a[:na, :na] = blk
    b = s.f64(m, n, ld=m + pad)
    return a, b


def solve_gpu(side, uplo, transa, diag, m, n, alpha, a, b):
    dA = DeviceMatrix.from_padded(a, a.shape[0])
    dB = DeviceMatrix.from_padded(b, m)
    blas.dtrsm(side, uplo, transa, diag, m, n, alpha, dA, dB)
    return dB.to_host_padded()


def residual(side, uplo, transa, diag, a, x, b, alpha, m, n):
    na = m if side == "L" else n
    tri = np.nan_to_num(a[:na, :na], nan=0.0)
    if diag == "U":
        np.fill_diagonal(tri, 1.0)
    op = tri if transa == "N" else tri.T
    lhs = op @ x[:m, :n] if side == "L" else x[:m, :n] @ op
    r = np.max(np.abs(lhs - alpha * b[:m, :n]))
    scale = np.max(np.abs(op)) * np.max(np.abs(x[:m, :n])) * na + abs(alpha) * np.max(np.abs(b[:m, :n]))
    return r / scale


@pytest.mark.parametrize("shape", [(300, 170), (777, 45), (45, 777), (9, 3)])
@pytest.mark.parametrize("side,uplo,transa,diag", VARIANTS)
def test_dtrsm_variants_vs_oracle(side, uplo, transa, diag, shape):
    # 300 x 170: a full 256-row leaf plus a ragged one; 777: three leaves and a
    # partial 64-block, with few (side L) or many (side R) slabs of right-hand
    # sides; 9 x 3: a single partial block
    m, n = shape
    a, b = tri_system(side, uplo, m, n, seed=VARIANTS.index((side, uplo, transa, diag)) + 300, boost=float(max(m, n)), pad=3)
    x = solve_gpu(side, uplo, transa, diag, m, n, 1.25, a, b)
    assert np.isfinite(x[:m, :n]).all()
    assert np.array_equal(x[m:, :], b[m:, :])           # padding rows untouched
    ref = b.copy(order="F")
    O.dtrsm(side, uplo, transa, diag, m, n, 1.25, a, a.shape[0], ref, ref.shape[0])
    rel = np.max(np.abs(x[:m, :n] - ref[:m, :n])) / np.max(np.abs(ref[:m, :n]))
    assert rel < 1e-12, rel
    assert residual(side, uplo, transa, diag, a, x, b, 1.25, m, n) < 64 * max(m, n) * 2.0 ** -53


def test_dtrsm_config_size_lower_left():
    """Config: side L, uplo L, N, N, M=4096, N=512, triangle uniform[-1,1), diagonal 4096+u[0,1)."""
    m, n = 4096, 512
    s = O.Stream(4242)
    a = s.f64(m, m)
    u = O.Stream(4243).f64(m, 1)[:, 0]
    a[np.arange(m), np.arange(m)] = 4096.0 + (u + 1.0) / 2.0
    a[np.triu_indices(m, 1)] = np.nan
    b = s.f64(m, n)
    x = solve_gpu("L", "L", "N", "N", m, n, 1.0, a, b)
    assert np.isfinite(x).all()
    assert residual("L", "L", "N", "N", a, x, b, 1.0, m, n) < 16 * m * 2.0 ** -53
    # forward error vs the oracle on a sample of columns
    cols = [0, 1, 255, 511]
    ref = np.asfortranarray(b[:, cols])
    O.dtrsm("L", "L", "N", "N", m, len(cols), 1.0, a, m, ref, m)
    assert np.max(np.abs(x[:, cols] - ref)) / np.max(np.abs(ref)) < 1e-13


def test_dtrsm_alpha_zero_and_host_dropin():
    m, n = 70, 40
    a, b = tri_system("L", "U", m, n, 3, 70.0)
    x = solve_gpu("L", "U", "N", "N", m, n, 0.0, a, b)
    assert np.all(x[:m, :n] == 0)
    # Fortran-ABI host path
    a2, b2 = tri_system("R", "L", m, n, 4, 70.0)
    xh = b2.copy(order="F")
    blas.dtrsm_("R", "L", "T", "N", m, n, 0.5, a2, a2.shape[0], xh, xh.shape[0])
    ref = b2.copy(order="F")
    O.dtrsm("R", "L", "T", "N", m, n, 0.5, a2, a2.shape[0], ref, ref.shape[0])
    assert np.max(np.abs(xh - ref)) / np.max(np.abs(ref)) < 1e-12


def test_dtrsm_kernel_level_matches_reference_solvers():
    """gemm_amd_dtrsm_kernel reproduces the reference's 62 dtrsm_solve_* kernels on their own outputs."""
    from tests.golden.make_golden import trsm_kernel_inputs
    G = golden_arrays("golden_trsm_kernel.npz")
    for case in golden_index()["trsm_kernel"]:
        sa, sb, C, K = trsm_kernel_inputs(case["variant"], case["M"], case["N"], case["seed"])
        bufs = []
        for arr in (sa, sb, C):
            d = DeviceBuffer(arr.nbytes)
            d.upload(arr)
            bufs.append(d)
        blas.dtrsm_kernel(case["variant"], case["M"], case["N"], bufs[0], bufs[1], bufs[2], case["LDB"], case["LDC"])
        out = [d.download(np.empty_like(arr)) for d, arr in zip(bufs, (sa, sb, C))]
        scale = max(1.0, np.max(np.abs(G[case["name"] + "_c"])))
        for got, key in zip(out, ("_sa", "_sb", "_c")):
            assert np.max(np.abs(got - G[case["name"] + key])) <= 1e-13 * scale, (case["name"], key)
