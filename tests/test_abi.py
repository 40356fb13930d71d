"""The C-ABI library: it loads, exports every symbol the headers declare, and
its host-side argument checking behaves like reference BLAS -- all without a
GPU (argument checks run before any device work)."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

from gemm_avx512f_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M):
            names.add(m.group(1))
    names.discard("extern")
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert "dgemm_" in syms and "sgemm_" in syms
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding covers exactly the declared ABI
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_exported_dynamic_symbols_are_only_the_abi():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    text_syms = {line.split()[-1] for line in out.splitlines() if " T " in line}
    ours = {s for s in text_syms if s.endswith("_") and s[:-1].isalpha() or s.startswith("gemm_amd_")}
    assert set(declared_symbols()) <= text_syms
    assert ours <= set(declared_symbols()) | {s for s in text_syms if not s.startswith("gemm_amd_")}


@pytest.mark.parametrize("args,info", [
    (("X", "N", 4, 4, 4, 4, 4, 4), 1),
    (("N", "Q", 4, 4, 4, 4, 4, 4), 2),
    (("N", "N", -1, 4, 4, 4, 4, 4), 3),
    (("N", "N", 4, -1, 4, 4, 4, 4), 4),
    (("N", "N", 4, 4, -1, 4, 4, 4), 5),
    (("N", "N", 4, 4, 4, 3, 4, 4), 8),
    (("T", "N", 4, 4, 5, 4, 5, 4), 8),     # A stored K x M: lda >= K
    (("N", "N", 4, 4, 5, 4, 4, 4), 10),
    (("N", "T", 4, 6, 5, 4, 5, 4), 10),    # B stored N x K: ldb >= N
    (("N", "N", 4, 4, 4, 4, 4, 3), 13),
])
def test_device_entry_argument_checks(args, info, capfd):
    ta, tb, m, n, k, lda, ldb, ldc = args
    L = _lib.lib()
    for fn in (L.gemm_amd_dgemm, L.gemm_amd_sgemm):
        rc = fn(ta.encode(), tb.encode(), m, n, k, 1.0, None, lda, None, ldb, 0.0, None, ldc, None)
        assert rc == info
    err = capfd.readouterr().err
    assert "parameter number %2d had an illegal value" % info in err
    assert "DGEMM" in err and "SGEMM" in err


def test_host_entry_rejects_and_leaves_c_untouched(capfd):
    from gemm_avx512f_amd import blas
    c = np.asfortranarray(np.full((4, 4), 7.0))
    a = np.asfortranarray(np.zeros((4, 4)))
    blas.dgemm_("N", "N", 4, 4, 4, 1.0, a, 3, a, 4, 0.0, c, 4)   # lda < m
    assert np.all(c == 7.0)
    assert "DGEMM  parameter number  8" in capfd.readouterr().err


def test_quick_returns_need_no_device():
    L = _lib.lib()
    # m == 0, n == 0, and (alpha == 0 or k == 0) with beta == 1 do nothing and succeed
    assert L.gemm_amd_dgemm(b"N", b"N", 0, 5, 5, 1.0, None, 1, None, 5, 0.0, None, 1, None) == 0
    assert L.gemm_amd_dgemm(b"N", b"N", 5, 0, 5, 1.0, None, 5, None, 5, 0.0, None, 5, None) == 0
    assert L.gemm_amd_sgemm(b"N", b"N", 5, 5, 0, 1.0, None, 5, None, 1, 1.0, None, 5, None) == 0
    assert L.gemm_amd_sgemm(b"T", b"T", 5, 5, 5, 0.0, None, 5, None, 5, 1.0, None, 5, None) == 0
    c = np.asfortranarray(np.full((3, 3), 2.0))
    from gemm_avx512f_amd import blas
    blas.dgemm_("N", "N", 3, 3, 0, 1.0, c, 3, c, 1, 1.0, c, 3)
    assert np.all(c == 2.0)


def test_python_binding_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(_lib.GemmAmdError):
        _lib.lib()


def test_dgetrf_host_reports_runtime_failure(capfd):
    # a legal call that cannot run (no GPU here, or a HIP failure) leaves a
    # nonzero info (GEMM_AMD_INFO_RUNTIME_ERROR), the matrix and the pivots
    # untouched, so LAPACK callers read it as a failure; the ILP64 symbol
    # passes the same info through and leaves its pivots alone
    from gemm_avx512f_amd import blas
    if _lib.lib().gemm_amd_init() == 0:
        pytest.skip("a GPU is present: the call would succeed")
    m, n = 20000, 2
    a = np.asfortranarray(np.full((m, n), 3.0))
    ipiv = np.full(n, -7, np.int32)
    info = blas.dgetrf_(m, n, a, m, ipiv)
    assert info == -1000
    assert np.all(a == 3.0) and np.all(ipiv == -7)
    assert "DGETRF" in capfd.readouterr().err
    L = _lib.lib()
    m64, n64, lda64, info64 = (ctypes.c_int64(v) for v in (m, n, m, 0))
    piv64 = np.full(n, -7, np.int64)
    L.dgetrf_64_(ctypes.byref(m64), ctypes.byref(n64), a.ctypes.data_as(ctypes.c_void_p), ctypes.byref(lda64),
                 piv64.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info64))
    assert info64.value == -1000
    assert np.all(a == 3.0) and np.all(piv64 == -7)


def test_noalloc_dtrsm_rejects_before_device_work():
    # argument checks and the scratch limit come before any device work
    L = _lib.lib()
    assert L.gemm_amd_dtrsm_noalloc(b"X", b"L", b"N", b"N", 4, 4, 1.0, None, 4, None, 4, None) == 1
    assert L.gemm_amd_dtrsm_noalloc(b"R", b"L", b"N", b"N", 4, 32769, 1.0, None, 32769, None, 4, None) == -1
    assert "exceeds the maximum of 32768" in _lib.last_error()
    assert L.gemm_amd_dtrsm_noalloc(b"L", b"L", b"N", b"N", 0, 4, 1.0, None, 1, None, 1, None) == 0


def test_emulation_counts_table():
    # the tuned table compiled into the library: 15 moduli up to K = 512, 16
    # beyond; 8 slices everywhere; installing a table is validated
    from gemm_avx512f_amd import tuning
    assert tuning.emulation_counts(64) == (15, 8)
    assert tuning.emulation_counts(512) == (15, 8)
    assert tuning.emulation_counts(513) == (16, 8)
    assert tuning.emulation_counts(8192) == (16, 8)
    bad = (ctypes.c_int * 33)(*([1] * 33))
    ok = (ctypes.c_int * 33)(*([8] * 33))
    assert _lib.lib().gemm_amd_emulation_counts_set(bad, ok, 33) < 0
    assert _lib.lib().gemm_amd_emulation_counts(0, None, None) < 0


def test_counts_tables_from_tuning_result():
    from gemm_avx512f_amd import tuning
    tuned = {64: {"moduli": 12, "slices": 6}, 1024: {"moduli": 14, "slices": 7}, 8192: {"moduli": 16, "slices": 8}}
    mod, sl = tuning.counts_tables(tuned)
    assert len(mod) == len(sl) == 33
    assert mod[0] == mod[6] == 12 and mod[7] == mod[10] == 14 and mod[11] == mod[13] == 16
    assert mod[14] == 16 and sl[32] == 8                      # beyond the tuned range: most accurate
    assert sl[6] == 6 and sl[10] == 7
