"""gemm_amd_dtrsm_noalloc (the DTRSM entry point whose scratch is a fixed
device array, for callers that cannot allocate) against gemm_amd_dtrsm on the
same inputs, through the C ABI: two calls in a row, the second with a smaller
triangle, for side L and side R, each bitwise equal to the allocating entry
point's result (same kernels, same configurations: the allocating call runs
first and leaves its trailing-update configurations in the picker's cache),
plus the rejection of a triangle beyond the fixed maximum."""
import ctypes

import numpy as np
import pytest

from gemm_avx512f_amd import _lib
from gemm_avx512f_amd.device import DeviceMatrix
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _system(m, n, side, seed):
    na = m if side == "L" else n
    s = O.Stream(seed)
    a = s.f64(na, na, ld=na + 3)
    a[np.arange(na), np.arange(na)] = na + np.abs(a[np.arange(na), np.arange(na)])
    a[:na, :na][np.triu(np.ones((na, na), bool), 1)] = np.nan   # lower solve: the upper triangle is never read
    b = s.f64(m, n, ld=m + 1)
    return a, b


def _solve(fn, side, m, n, a, b):
    dA = DeviceMatrix.from_padded(a, a.shape[0])
    dB = DeviceMatrix.from_padded(b, m)
    rc = fn(side.encode(), b"L", b"N", b"N", m, n, 0.75, ctypes.c_void_p(dA.ptr), dA.ld, ctypes.c_void_p(dB.ptr),
            dB.ld, None)
    assert rc == 0, _lib.last_error()
    assert _lib.lib().gemm_amd_synchronize() == 0
    return dB.to_host_padded()


@pytest.mark.parametrize("side", ["L", "R"])
def test_noalloc_matches_allocating_entry_point(gpu_lib, side):
    # first call: three 256-row leaves and trailing updates of 1000 x 300 (side
    # L) / 300 x 1000 (side R); second call: a smaller triangle on the same
    # fixed scratch, whose earlier contents must not leak into the result
    shapes = [(1000, 300), (333, 77)] if side == "L" else [(300, 1000), (77, 333)]
    for i, (m, n) in enumerate(shapes):
        a, b = _system(m, n, side, seed=900 + 10 * i + (side == "R"))
        want = _solve(gpu_lib.gemm_amd_dtrsm, side, m, n, a, b)
        got = _solve(gpu_lib.gemm_amd_dtrsm_noalloc, side, m, n, a, b)
        assert np.isfinite(got[:m, :n]).all()
        assert np.array_equal(got[:m, :n], want[:m, :n]), (side, m, n)
        assert np.array_equal(got[m:, :], b[m:, :])   # padding rows untouched
        # and the solution solves the system (residual of op(A) X = alpha B)
        na = m if side == "L" else n
        tri = np.tril(np.nan_to_num(a[:na, :na]))
        x = got[:m, :n]
        lhs = tri @ x if side == "L" else x @ tri
        r = np.max(np.abs(lhs - 0.75 * b[:m, :n])) / (np.max(np.abs(tri)) * np.max(np.abs(x)) * na)
        assert r < 1e-15, r


def test_noalloc_rejects_triangle_beyond_its_scratch(gpu_lib):
    big = 32769
    rc = gpu_lib.gemm_amd_dtrsm_noalloc(b"L", b"L", b"N", b"N", big, 4, 1.0, None, big, None, big, None)
    assert rc == -1
    assert "exceeds the maximum of 32768" in _lib.last_error()
