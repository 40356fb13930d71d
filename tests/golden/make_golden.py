"""Writes the golden fixtures in this directory from the reference's own CPU
build (oracle/_ref/{SGEMM,DGEMM,libdtrsm16x4}.so, compiled by oracle/Makefile
from the reference sources).

Inputs are not stored: every case names a splitmix64 seed, and the inputs
are regenerated by the oracle's generator (A, then B, then C, column-major),
whose known-answer test pins it.  Stored are the reference's outputs: whole C
for small cases; for the full-size configs a fixed sample of entries plus
float64 checksums of the whole C.

Run:  python tests/golden/make_golden.py      (needs an AVX-512F host)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def gemm_inputs(dtype, seed, ta, tb, m, n, k, lda, ldb, ldc):
    s = O.Stream(seed)
    fill = s.f64 if dtype == np.float64 else s.f32
    ar, ac = (m, k) if ta in "Nn" else (k, m)
    br, bc = (k, n) if tb in "Nn" else (n, k)
    a = fill(ar, ac, lda)
    b = fill(br, bc, ldb)
    c = fill(m, n, ldc)
    return a, b, c


GEMM_CASES = [
    # name, dtype, seed, ta, tb, m, n, k, alpha, beta, pad (ld = rows + pad), sample
    ("s_nn_small", "f32", 11, "N", "N", 64, 48, 40, 1.0, 0.0, 0, False),
    ("s_tn_edge", "f32", 12, "T", "N", 49, 257, 385, 1.0, 0.0, 3, False),
    ("s_nt_edge", "f32", 13, "N", "T", 47, 255, 383, -0.75, 2.0, 1, False),
    ("s_tt_edge", "f32", 14, "T", "T", 48, 256, 384, 0.5, -1.0, 0, False),
    ("d_nn_edge", "f64", 21, "N", "N", 25, 161, 385, 1.5, 0.5, 2, False),
    ("d_tn_edge", "f64", 22, "T", "N", 24, 160, 384, 1.5, 0.5, 0, False),
    ("d_nt_edge", "f64", 23, "N", "T", 23, 159, 383, -2.0, 0.25, 5, False),
    ("d_tt_edge", "f64", 24, "T", "T", 97, 33, 129, 1.0, 1.0, 0, False),
    ("d_k0", "f64", 25, "N", "N", 17, 9, 0, 1.5, 0.5, 0, False),
    ("d_alpha0", "f64", 26, "N", "N", 17, 9, 5, 0.0, -3.0, 0, False),
    ("s_1x1x1", "f32", 27, "N", "N", 1, 1, 1, 2.0, 0.0, 0, False),
    # the configs themselves, sampled
    ("config_sgemm_1024_nn", "f32", 0, "N", "N", 1024, 1024, 1024, 1.0, 0.0, 0, True),
    ("config_dgemm_4096_nn", "f64", 1, "N", "N", 4096, 4096, 4096, 1.5, 0.5, 0, True),
    ("config_dgemm_4096_tn", "f64", 2, "T", "N", 4096, 4096, 4096, 1.5, 0.5, 0, True),
]

NSAMPLE = 2048


def sample_idx(m, n, seed):
    rng = np.random.default_rng(seed + 1000)
    return rng.integers(0, m, NSAMPLE).astype(np.int32), rng.integers(0, n, NSAMPLE).astype(np.int32)


def make_gemm():
    index, arrays = [], {}
    for name, dt, seed, ta, tb, m, n, k, alpha, beta, pad, sample in GEMM_CASES:
        dtype = np.float64 if dt == "f64" else np.float32
        lda = (m if ta == "N" else k) + pad or 1
        ldb = (k if tb == "N" else n) + pad or 1
        ldc = m + pad or 1
        lda, ldb, ldc = max(lda, 1), max(ldb, 1), max(ldc, 1)
        a, b, c = gemm_inputs(dtype, seed, ta, tb, m, n, k, lda, ldb, ldc)
        ok = O.ref_gemm(dtype, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        if not ok:
            raise SystemExit("reference build not available (oracle/_ref, AVX-512F host)")
        cm = np.asarray(c[:m, :n], dtype=np.float64)
        entry = dict(name=name, dtype=dt, seed=seed, ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, lda=lda,
                     ldb=ldb, ldc=ldc, sampled=sample)
        if sample:
            ri, ci = sample_idx(m, n, seed)
            arrays[name + "_rows"] = ri
            arrays[name + "_cols"] = ci
            arrays[name + "_vals"] = c[ri, ci].copy()
            entry["checksum_sum"] = float(cm.sum())
            entry["checksum_sumsq"] = float((cm * cm).sum())
        else:
            arrays[name + "_c"] = np.asarray(c).copy()   # whole padded C: padding rows must stay untouched
        index.append(entry)
        print("gemm", name, "done", flush=True)
    return index, arrays


# DTRSM kernel-level cases: every solver the reference exports
# (dtrsm_kernel_16x4_tester.c:297-365 lists them).
TRSM_L = [(m, n) for m in (1, 2, 4, 8, 16) for n in (1, 2, 4, 8, 12)]
TRSM_R = [(m, n) for m in (8, 16) for n in (4, 8, 12)]


def trsm_kernel_inputs(variant, M, N, seed):
    """sa, sb, C as the tester builds them (sizes :236-241) from a seeded stream.
    The pivots are drawn in [1, 2) and stored inverted, as the packed format
    requires (the solvers multiply by them)."""
    is_r = variant >= 2
    K = N if is_r else M
    s = O.Stream(seed)
    sa = s.f64(M * K, 1)[:, 0].copy()
    sb = s.f64(N * K, 1)[:, 0].copy()
    C = s.f64(M * N, 1)[:, 0].copy()
    if not is_r:
        for i in range(M):
            sa[i * M + i] = 1.0 / (1.5 + 0.5 * sa[i * M + i])
    else:
        # diagonal of the packed B triangle at get_b_ref(k, k)
        unroll = 4 if N > 4 else N
        for kk in range(N):
            off = (kk >> 2) * (K << 2) + kk * unroll + (kk & 3)
            sb[off] = 1.0 / (1.5 + 0.5 * sb[off])
    return sa, sb, C, K


def make_trsm():
    index, arrays = [], {}
    seed = 500
    for variant, cases in ((0, TRSM_L), (1, TRSM_L), (2, TRSM_R), (3, TRSM_R)):
        for M, N in cases:
            seed += 1
            sa, sb, C, K = trsm_kernel_inputs(variant, M, N, seed)
            ok = O.ref_dtrsm_kernel(variant, M, N, sa, sb, C, K, M)
            if not ok:
                raise SystemExit("reference dtrsm build not available")
            name = "trsm_%s_m%dn%d" % (["ln", "lt", "rn", "rt"][variant], M, N)
            arrays[name + "_sa"] = sa
            arrays[name + "_sb"] = sb
            arrays[name + "_c"] = C
            index.append(dict(name=name, variant=variant, M=M, N=N, seed=seed, LDB=K, LDC=M))
    print("trsm kernels done", flush=True)
    return index, arrays


def main():
    gi, ga = make_gemm()
    ti, ta = make_trsm()
    np.savez_compressed(os.path.join(HERE, "golden_gemm.npz"), **ga)
    np.savez_compressed(os.path.join(HERE, "golden_trsm_kernel.npz"), **ta)
    with open(os.path.join(HERE, "golden_index.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "source": "reference CPU build (oracle/_ref)",
                   "gemm": gi, "trsm_kernel": ti}, f, indent=1)


if __name__ == "__main__":
    main()
