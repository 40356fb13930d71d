"""Shared helpers for the tests: seeded inputs, oracle references, tolerances."""
import json
import os

import numpy as np

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

EPS = {np.float32: 2.0 ** -24, np.float64: 2.0 ** -53}


def golden_index():
    with open(os.path.join(GOLDEN, "golden_index.json")) as f:
        return json.load(f)


def golden_arrays(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def gemm_case(dtype, seed, ta, tb, m, n, k, lda=None, ldb=None, ldc=None, pad=0):
    """Seeded column-major inputs: one splitmix64 stream fills A, then B, then C."""
    ar, ac = (m, k) if ta in "Nn" else (k, m)
    br, bc = (k, n) if tb in "Nn" else (n, k)
    lda = lda or max(1, ar + pad)
    ldb = ldb or max(1, br + pad)
    ldc = ldc or max(1, m + pad)
    s = O.Stream(seed)
    fill = s.f64 if dtype == np.float64 else s.f32
    a = fill(ar, ac, lda)
    b = fill(br, bc, ldb)
    c = fill(m, n, ldc)
    return a, b, c, lda, ldb, ldc


def gemm_bound(dtype, k, c=4.0):
    """Normwise relative-error bound c*(K+2)*u for an FP32/FP64 GEMM result vs the exact one."""
    return c * (k + 2) * EPS[dtype]


def scaled_err(c_out, c_ref, a_abs_b_abs, beta_c_abs):
    """max |c - ref| / (|alpha| |A||B| + |beta||C|) elementwise (the BLAS componentwise measure)."""
    denom = a_abs_b_abs + beta_c_abs
    denom = np.where(denom == 0, 1.0, denom)
    return float(np.max(np.abs(c_out - c_ref) / denom))


def opmat(x, t, rows, cols):
    """op(X) as a dense float64 array from a column-major storage array."""
    return np.asarray(x[:rows, :cols], dtype=np.float64) if t in "Nn" else np.asarray(x[:cols, :rows],
                                                                                        dtype=np.float64).T
