"""ctypes front end of the CPU oracle (oracle.c) and of the reference's own
CPU build (oracle/_ref/, built by oracle/Makefile from the reference sources).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker (or the timed CPU
baseline).  The shipped library never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_vp = ctypes.c_void_p
_i = ctypes.c_int
_l = ctypes.c_long
_i64 = ctypes.c_int64
_d = ctypes.c_double
_f = ctypes.c_float
_c = ctypes.c_char

_lib = None


def build():
    """Compile liboracle.so (and _ref/ when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        sig = {
            "oracle_splitmix64_next": (ctypes.c_uint64, [ctypes.POINTER(ctypes.c_uint64)]),
            "oracle_fill_f64": (None, [ctypes.POINTER(ctypes.c_uint64), _i64, _i64, _i64, _vp]),
            "oracle_fill_f32": (None, [ctypes.POINTER(ctypes.c_uint64), _i64, _i64, _i64, _vp]),
            "oracle_fill_f64_range": (None, [ctypes.POINTER(ctypes.c_uint64), _i64, _i64, _i64, _i, _vp]),
            "oracle_dgemm": (None, [_c, _c, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i]),
            "oracle_sgemm": (None, [_c, _c, _i, _i, _i, _f, _vp, _i, _vp, _i, _f, _vp, _i]),
            "oracle_dgemm_q": (None, [_c, _c, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i, _i64, _vp, _vp, _vp,
                                      _vp]),
            "oracle_dtrsm": (None, [_c, _c, _c, _c, _i, _i, _d, _vp, _i, _vp, _i]),
            "oracle_dtrsm_kernel": (None, [_i, _l, _l, _vp, _vp, _vp, _l, _l]),
            "oracle_dgetrf": (_i, [_i, _i, _vp, _i, _vp]),
            "oracle_zgemm": (None, [_c, _c, _i, _i, _i, _vp, _vp, _i, _vp, _i, _vp, _vp, _i]),
            "oracle_cgemm": (None, [_c, _c, _i, _i, _i, _vp, _vp, _i, _vp, _i, _vp, _vp, _i]),
            "oracle_err_f64": (None, [_i64, _vp, _vp, _vp, _vp]),
            "oracle_err_f32": (None, [_i64, _vp, _vp, _vp]),
            "oracle_ozaki_split": (None, [_i, _i, _vp, _i64, _i64, _i, _vp, _vp]),
            "oracle_f32_to_bf16": (ctypes.c_uint16, [_f]),
            "oracle_bf16_to_f32": (_f, [ctypes.c_uint16]),
            "oracle_bf16_split": (None, [_i64, _vp, _i, _vp]),
            "oracle_dgemm_blocked": (None, [_c, _c, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i, _i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _ch(x):
    return x.encode() if isinstance(x, str) else x


# ------------------------------------------------------------------ seeded inputs
class Stream:
    """One splitmix64 stream; fills consume it in call order (A, then B, then C)."""

    def __init__(self, seed):
        self.state = ctypes.c_uint64(seed)

    def next(self):
        return lib().oracle_splitmix64_next(ctypes.byref(self.state))

    def f64(self, rows, cols, ld=None, r=None):
        ld = ld or max(1, rows)
        x = np.zeros((ld, max(cols, 1)), dtype=np.float64, order="F")
        if r is None:
            lib().oracle_fill_f64(ctypes.byref(self.state), rows, cols, ld, _p(x))
        else:
            lib().oracle_fill_f64_range(ctypes.byref(self.state), rows, cols, ld, r, _p(x))
        return x

    def f32(self, rows, cols, ld=None):
        ld = ld or max(1, rows)
        x = np.zeros((ld, max(cols, 1)), dtype=np.float32, order="F")
        lib().oracle_fill_f32(ctypes.byref(self.state), rows, cols, ld, _p(x))
        return x

    def c128(self, rows, cols, ld=None):
        """Complex double: (re, im) pairs drawn in storage order."""
        ld = ld or max(1, rows)
        x = self.f64(2 * ld, max(cols, 1), 2 * ld)   # fills interleaved storage column by column
        v = x.reshape(-1, order="F").view(np.complex128).reshape((ld, max(cols, 1)), order="F")
        if ld > rows:
            v[rows:, :] = 0
        return np.asfortranarray(v)

    def c64(self, rows, cols, ld=None):
        ld = ld or max(1, rows)
        x = self.f32(2 * ld, max(cols, 1), 2 * ld)
        v = x.reshape(-1, order="F").view(np.complex64).reshape((ld, max(cols, 1)), order="F")
        if ld > rows:
            v[rows:, :] = 0
        return np.asfortranarray(v)


# ------------------------------------------------------------------ oracle calls
def dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    lib().oracle_dgemm(_ch(ta), _ch(tb), m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)


def sgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    lib().oracle_sgemm(_ch(ta), _ch(tb), m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)


def dgemm_q(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, rows=None, cols=None):
    """Double-double (hi, lo) reference from __float128 sums; all entries or the sampled ones."""
    if rows is None:
        hi = np.zeros((m, n), dtype=np.float64, order="F")
        lo = np.zeros((m, n), dtype=np.float64, order="F")
        lib().oracle_dgemm_q(_ch(ta), _ch(tb), m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc, 0, None,
                             None, _p(hi), _p(lo))
    else:
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        hi = np.zeros(len(rows), dtype=np.float64)
        lo = np.zeros(len(rows), dtype=np.float64)
        lib().oracle_dgemm_q(_ch(ta), _ch(tb), m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc, len(rows),
                             _p(rows), _p(cols), _p(hi), _p(lo))
    return hi, lo


def dtrsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb):
    lib().oracle_dtrsm(_ch(side), _ch(uplo), _ch(transa), _ch(diag), m, n, alpha, _p(a), lda, _p(b), ldb)


def dtrsm_kernel(variant, M, N, sa, sb, C, LDB, LDC):
    lib().oracle_dtrsm_kernel(variant, M, N, _p(sa), _p(sb), _p(C), LDB, LDC)


def dgetrf(m, n, a, lda):
    """LU with partial pivoting in place on a Fortran-order float64 array;
    returns (ipiv int32 1-based, info)."""
    assert a.dtype == np.float64 and a.flags.f_contiguous
    ipiv = np.zeros(max(min(m, n), 1), dtype=np.int32)
    info = lib().oracle_dgetrf(m, n, _p(a), lda, _p(ipiv))
    return ipiv[: min(m, n)], info


def zgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    al = np.array([alpha.real, alpha.imag], dtype=np.float64)
    be = np.array([beta.real, beta.imag], dtype=np.float64)
    lib().oracle_zgemm(_ch(ta), _ch(tb), m, n, k, _p(al), _p(a), lda, _p(b), ldb, _p(be), _p(c), ldc)


def cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    al = np.array([alpha.real, alpha.imag], dtype=np.float32)
    be = np.array([beta.real, beta.imag], dtype=np.float32)
    lib().oracle_cgemm(_ch(ta), _ch(tb), m, n, k, _p(al), _p(a), lda, _p(b), ldb, _p(be), _p(c), ldc)


def err(c, ref_hi, ref_lo=None):
    """(max |c-r| / max |r|, ||c-r||_F / ||r||_F, max elementwise |c-r|/|r|)."""
    c = np.ascontiguousarray(c)
    out = np.zeros(3)
    if c.dtype == np.float32:
        lib().oracle_err_f32(c.size, _p(c), _p(np.ascontiguousarray(ref_hi, dtype=np.float64)), _p(out))
    else:
        rl = None if ref_lo is None else np.ascontiguousarray(ref_lo, dtype=np.float64)
        lib().oracle_err_f64(c.size, _p(np.ascontiguousarray(c, dtype=np.float64)),
                             _p(np.ascontiguousarray(ref_hi, dtype=np.float64)), _p(rl), _p(out))
    return {"max_rel": out[0], "frob_rel": out[1], "max_elem_rel": out[2]}


def ozaki_split(x, nslices, by="rows"):
    """Slices of a 2-D float64 array, per row (by='rows') or per column; returns (slices[s, r, k], exps[r])."""
    x = np.asfortranarray(x, dtype=np.float64)
    if by == "rows":
        rows, k = x.shape
        rs, ks = 1, x.shape[0]
    else:
        k, rows = x.shape
        rs, ks = x.shape[0], 1
    sl = np.zeros((nslices, rows, k), dtype=np.int8)
    ex = np.zeros(rows, dtype=np.int32)
    lib().oracle_ozaki_split(rows, k, _p(x), rs, ks, nslices, _p(sl), _p(ex))
    return sl, ex


CRT_MODULI = (256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193)


def ozaki_crt_split(x, nmod, bits, by="rows"):
    """Symmetric residues of the integer-scaled rows (or columns) of a float64 array modulo the first
    nmod moduli; returns (res[l, r, k] int8, exps[r] = e_r - bits)."""
    x = np.asfortranarray(x, dtype=np.float64)
    if by == "rows":
        rows, k = x.shape
        rs, ks = 1, x.shape[0]
    else:
        k, rows = x.shape
        rs, ks = x.shape[0], 1
    res = np.zeros((nmod, rows, k), dtype=np.int8)
    ex = np.zeros(rows, dtype=np.int32)
    mods = np.array(CRT_MODULI, dtype=np.int32)
    L = lib()
    L.oracle_ozaki_crt_split.argtypes = [_i, _i, _vp, _i64, _i64, _i, _vp, _i, _vp, _vp]
    L.oracle_ozaki_crt_split.restype = None
    L.oracle_ozaki_crt_split(rows, k, _p(x), rs, ks, nmod, _p(mods), bits, _p(res), _p(ex))
    return res, ex


def bf16_split(x, nsplit):
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    parts = np.zeros((nsplit, x.size), dtype=np.uint16)
    lib().oracle_bf16_split(x.size, _p(x), nsplit, _p(parts))
    return parts


def bf16_to_f32(h):
    h = np.asarray(h, dtype=np.uint16)
    return (h.astype(np.uint32) << 16).view(np.float32)


def dgemm_blocked(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nthreads):
    lib().oracle_dgemm_blocked(_ch(ta), _ch(tb), m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc, nthreads)


# ------------------------------------------------------------------ picker perceptron (mlp_oracle.c)
MLP_IN, MLP_HID, MLP_OUT = 34, 64, 20
MLP_NPARAMS = MLP_IN * MLP_HID + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID * MLP_OUT + MLP_OUT


def _mlp_sig():
    L = lib()
    if not hasattr(L, "_mlp_ready"):
        L.oracle_mlp_forward.argtypes = [_vp, _vp, _i, _vp]
        L.oracle_mlp_forward.restype = None
        L.oracle_mlp_loss.argtypes = [_vp, _vp, _vp, _vp, _i]
        L.oracle_mlp_loss.restype = _d
        L.oracle_mlp_grad.argtypes = [_vp, _vp, _vp, _vp, _i, _vp, _vp]
        L.oracle_mlp_grad.restype = None
        L.oracle_adam.argtypes = [_vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _i]
        L.oracle_adam.restype = None
        L._mlp_ready = True
    return L


def _d64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mlp_forward(params, x):
    p, x = _d64(params), _d64(x)
    y = np.zeros((x.shape[0], MLP_OUT))
    _mlp_sig().oracle_mlp_forward(_p(p), _p(x), x.shape[0], _p(y))
    return y


def mlp_loss(params, x, t, mask):
    p, x, t, mask = _d64(params), _d64(x), _d64(t), _d64(mask)
    return _mlp_sig().oracle_mlp_loss(_p(p), _p(x), _p(t), _p(mask), x.shape[0])


def mlp_grad(params, x, t, mask):
    p, x, t, mask = _d64(params), _d64(x), _d64(t), _d64(mask)
    g = np.zeros(MLP_NPARAMS)
    loss = ctypes.c_double()
    _mlp_sig().oracle_mlp_grad(_p(p), _p(x), _p(t), _p(mask), x.shape[0], _p(g), ctypes.byref(loss))
    return g, loss.value


def adam(params, g, m, v, lr, b1, b2, eps, step):
    """One Adam step in place on float64 arrays."""
    _mlp_sig().oracle_adam(_p(params), _p(_d64(g)), _p(m), _p(v), params.size, lr, b1, b2, eps, step)


# ------------------------------------------------------------------ reference build (oracle/_ref)
_ref_cache = {}


def ref_lib(name):
    """The reference's own CPU library (DGEMM.so, SGEMM.so, libdtrsm16x4.so) or None if not built.
    Each .so is loaded with RTLD_LOCAL so that their dgemm_/sgemm_ never collide."""
    if name in _ref_cache:
        return _ref_cache[name]
    path = os.path.join(REF_DIR, name)
    L = None
    if os.path.exists(path) and _cpu_has_avx512f():
        L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    _ref_cache[name] = L
    return L


def _cpu_has_avx512f():
    try:
        with open("/proc/cpuinfo") as f:
            return " avx512f" in f.read()
    except OSError:
        return False


def ref_gemm(dtype, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Call the reference sgemm_/dgemm_ (Fortran ABI). Returns False if it is unavailable."""
    L = ref_lib("DGEMM.so" if dtype == np.float64 else "SGEMM.so")
    if L is None:
        return False
    ct = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    fn = L.dgemm_ if dtype == np.float64 else L.sgemm_
    fn.restype = None
    bi = lambda v: ctypes.byref(ctypes.c_int(v))  # noqa: E731
    fn(ctypes.c_char_p(_ch(ta)), ctypes.c_char_p(_ch(tb)), bi(m), bi(n), bi(k), ctypes.byref(ct(alpha)), _p(a),
       bi(lda), _p(b), bi(ldb), ctypes.byref(ct(beta)), _p(c), bi(ldc))
    return True


def ref_dtrsm_kernel(variant, M, N, sa, sb, C, LDB, LDC):
    """Call the reference dtrsm_solve_{ln,lt,rn,rt}_m{M}n{N} kernel; False if unavailable."""
    L = ref_lib("libdtrsm16x4.so")
    if L is None:
        return False
    name = "dtrsm_solve_%s_m%dn%d" % (["ln", "lt", "rn", "rt"][variant], M, N)
    fn = getattr(L, name)
    fn.restype = None
    fn.argtypes = [_vp, _vp, _vp, _l, _l]
    fn(_p(sa), _p(sb), _p(C), LDB, LDC)
    return True
