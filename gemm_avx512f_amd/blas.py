"""Python mirror of the reference's BLAS-3 interface, backed by libgemm_amd.so.

Two layers, as in the C ABI:

* ``sgemm_`` / ``dgemm_`` -- the Fortran-convention drop-in exactly as the
  reference exports it (src/gemm_driver.c:116): every argument by reference,
  host arrays, column-major, C updated in place.
* ``sgemm`` / ``dgemm`` -- the same contract on device matrices
  (:class:`DeviceMatrix`), stream-ordered.

Nothing here computes: every call goes through the shared library, and a
missing library raises.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceMatrix


def _ref(x, ctype):
    return ctypes.byref(ctype(x))


def _char(c):
    return ctypes.c_char_p(c.encode() if isinstance(c, str) else bytes(c))


def _host_ptr(a, dtype):
    if a.dtype != dtype:
        raise TypeError("expected %s array, got %s" % (dtype, a.dtype))
    if not a.flags.f_contiguous and not (a.ndim == 1 and a.flags.c_contiguous):
        raise ValueError("host matrices must be Fortran-ordered (column-major)")
    return a.ctypes.data_as(ctypes.c_void_p)


def _xgemm_(fn, ctype, dtype, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    fn(_char(transa), _char(transb), _ref(m, ctypes.c_int), _ref(n, ctypes.c_int), _ref(k, ctypes.c_int),
       _ref(alpha, ctype), _host_ptr(a, dtype), _ref(lda, ctypes.c_int), _host_ptr(b, dtype),
       _ref(ldb, ctypes.c_int), _ref(beta, ctype), _host_ptr(c, dtype), _ref(ldc, ctypes.c_int))


def sgemm_(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Fortran-ABI SGEMM on host arrays (C updated in place)."""
    _xgemm_(_lib.lib().sgemm_, ctypes.c_float, np.float32, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta,
            c, ldc)


def dgemm_(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Fortran-ABI DGEMM on host arrays (C updated in place)."""
    _xgemm_(_lib.lib().dgemm_, ctypes.c_double, np.float64, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta,
            c, ldc)


def _stream(stream):
    return stream.handle if stream is not None else None


def _enc(c):
    return c.encode() if isinstance(c, str) else c


def sgemm(transa, transb, m, n, k, alpha, a: DeviceMatrix, b: DeviceMatrix, beta, c: DeviceMatrix, stream=None,
          tile=None, splitk=None):
    """C <- alpha*op(A)*op(B) + beta*C on device matrices (FP32)."""
    L = _lib.lib()
    if tile is None and splitk is None:
        rc = L.gemm_amd_sgemm(_enc(transa), _enc(transb), m, n, k, alpha, a.ptr, a.ld, b.ptr, b.ld, beta, c.ptr,
                              c.ld, _stream(stream))
    else:
        rc = L.gemm_amd_sgemm_cfg(_enc(transa), _enc(transb), m, n, k, alpha, a.ptr, a.ld, b.ptr, b.ld, beta, c.ptr,
                                  c.ld, 0 if tile is None else tile, 1 if splitk is None else splitk,
                                  _stream(stream))
    return _lib.check(rc, "sgemm")


def dgemm(transa, transb, m, n, k, alpha, a: DeviceMatrix, b: DeviceMatrix, beta, c: DeviceMatrix, stream=None,
          tile=None, splitk=None):
    """C <- alpha*op(A)*op(B) + beta*C on device matrices (FP64)."""
    L = _lib.lib()
    if tile is None and splitk is None:
        rc = L.gemm_amd_dgemm(_enc(transa), _enc(transb), m, n, k, alpha, a.ptr, a.ld, b.ptr, b.ld, beta, c.ptr,
                              c.ld, _stream(stream))
    else:
        rc = L.gemm_amd_dgemm_cfg(_enc(transa), _enc(transb), m, n, k, alpha, a.ptr, a.ld, b.ptr, b.ld, beta, c.ptr,
                                  c.ld, 0 if tile is None else tile, 1 if splitk is None else splitk,
                                  _stream(stream))
    return _lib.check(rc, "dgemm")


# ------------------------------------------------------------------ DTRSM
def dtrsm_(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb):
    """Fortran-ABI DTRSM on host arrays (B overwritten with the solution)."""
    _lib.lib().dtrsm_(_char(side), _char(uplo), _char(transa), _char(diag), _ref(m, ctypes.c_int),
                      _ref(n, ctypes.c_int), _ref(alpha, ctypes.c_double), _host_ptr(a, np.float64),
                      _ref(lda, ctypes.c_int), _host_ptr(b, np.float64), _ref(ldb, ctypes.c_int))


def dtrsm(side, uplo, transa, diag, m, n, alpha, a: DeviceMatrix, b: DeviceMatrix, stream=None):
    """B <- alpha*op(A)^-1*B (side L) or alpha*B*op(A)^-1 (side R) on device matrices."""
    rc = _lib.lib().gemm_amd_dtrsm(_enc(side), _enc(uplo), _enc(transa), _enc(diag), m, n, alpha, a.ptr, a.ld, b.ptr,
                                   b.ld, _stream(stream))
    return _lib.check(rc, "dtrsm")


TRSM_KERNEL_VARIANTS = {"ln": 0, "lt": 1, "rn": 2, "rt": 3}


def dtrsm_kernel(variant, M, N, sa, sb, c, LDB, LDC, stream=None):
    """Kernel-level solve (reference dtrsm_solve_<variant>_m<M>n<N> contract) on device buffers."""
    v = TRSM_KERNEL_VARIANTS[variant] if isinstance(variant, str) else int(variant)
    rc = _lib.lib().gemm_amd_dtrsm_kernel(v, M, N, sa.ptr, sb.ptr, c.ptr, LDB, LDC, _stream(stream))
    return _lib.check(rc, "dtrsm_kernel")


# ------------------------------------------------------------------ LU (DGETRF)
def dgetrf_(m, n, a, lda, ipiv):
    """LAPACK-ABI DGETRF on host arrays: ``a`` (Fortran order, lda x n) is
    overwritten by L and U, ``ipiv`` (int32, min(m, n)) receives the 1-based
    pivot rows.  Returns info."""
    info = ctypes.c_int(0)
    if ipiv.dtype != np.int32 or not ipiv.flags["C_CONTIGUOUS"]:
        raise TypeError("ipiv must be a contiguous int32 array")
    _lib.lib().dgetrf_(_ref(m, ctypes.c_int), _ref(n, ctypes.c_int), _host_ptr(a, np.float64),
                       _ref(lda, ctypes.c_int), ipiv.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info))
    return info.value


def dgetrf(m, n, a: DeviceMatrix, ipiv_dev, info_dev, stream=None):
    """P A = L U in place on a device matrix; ``ipiv_dev`` (min(m, n) int32) and
    ``info_dev`` (one int32) are device buffers (anything with ``.ptr``)."""
    rc = _lib.lib().gemm_amd_dgetrf(m, n, a.ptr, a.ld, ipiv_dev.ptr, info_dev.ptr, _stream(stream))
    return _lib.check(rc, "dgetrf")


# ------------------------------------------------------------------ 3M complex GEMM
def _cpair(x, ctype):
    x = complex(x)
    return (ctype * 2)(x.real, x.imag)


def zgemm3m_(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Fortran-ABI ZGEMM3M on host complex128 arrays (C updated in place)."""
    _lib.lib().zgemm3m_(_char(transa), _char(transb), _ref(m, ctypes.c_int), _ref(n, ctypes.c_int),
                        _ref(k, ctypes.c_int), _cpair(alpha, ctypes.c_double), _host_ptr(a, np.complex128),
                        _ref(lda, ctypes.c_int), _host_ptr(b, np.complex128), _ref(ldb, ctypes.c_int),
                        _cpair(beta, ctypes.c_double), _host_ptr(c, np.complex128), _ref(ldc, ctypes.c_int))


def cgemm3m_(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Fortran-ABI CGEMM3M on host complex64 arrays (C updated in place)."""
    _lib.lib().cgemm3m_(_char(transa), _char(transb), _ref(m, ctypes.c_int), _ref(n, ctypes.c_int),
                        _ref(k, ctypes.c_int), _cpair(alpha, ctypes.c_float), _host_ptr(a, np.complex64),
                        _ref(lda, ctypes.c_int), _host_ptr(b, np.complex64), _ref(ldb, ctypes.c_int),
                        _cpair(beta, ctypes.c_float), _host_ptr(c, np.complex64), _ref(ldc, ctypes.c_int))


def zgemm3m(transa, transb, m, n, k, alpha, a: DeviceMatrix, b: DeviceMatrix, beta, c: DeviceMatrix, stream=None):
    rc = _lib.lib().gemm_amd_zgemm3m(_enc(transa), _enc(transb), m, n, k, _cpair(alpha, ctypes.c_double), a.ptr, a.ld,
                                     b.ptr, b.ld, _cpair(beta, ctypes.c_double), c.ptr, c.ld, _stream(stream))
    return _lib.check(rc, "zgemm3m")


def cgemm3m(transa, transb, m, n, k, alpha, a: DeviceMatrix, b: DeviceMatrix, beta, c: DeviceMatrix, stream=None):
    rc = _lib.lib().gemm_amd_cgemm3m(_enc(transa), _enc(transb), m, n, k, _cpair(alpha, ctypes.c_float), a.ptr, a.ld,
                                     b.ptr, b.ld, _cpair(beta, ctypes.c_float), c.ptr, c.ld, _stream(stream))
    return _lib.check(rc, "cgemm3m")


# ------------------------------------------------------------------ emulated precision
def dgemm_ozaki(transa, transb, m, n, k, alpha, a: DeviceMatrix, b: DeviceMatrix, beta, c: DeviceMatrix, nslices=7,
                stream=None):
    """DGEMM emulated on int8 matrix cores with `nslices` 7-bit slices per operand (Ozaki scheme)."""
    rc = _lib.lib().gemm_amd_dgemm_ozaki(_enc(transa), _enc(transb), m, n, k, alpha, a.ptr, a.ld, b.ptr, b.ld, beta,
                                         c.ptr, c.ld, nslices, _stream(stream))
    return _lib.check(rc, "dgemm_ozaki")


def dgemm_ozaki_crt(transa, transb, m, n, k, alpha, a: DeviceMatrix, b: DeviceMatrix, beta, c: DeviceMatrix,
                    nmoduli=16, stream=None):
    """DGEMM emulated on int8 matrix cores in modular form: one int8 GEMM per modulus (2..16), CRT rebuild."""
    rc = _lib.lib().gemm_amd_dgemm_ozaki_crt(_enc(transa), _enc(transb), m, n, k, alpha, a.ptr, a.ld, b.ptr, b.ld,
                                             beta, c.ptr, c.ld, nmoduli, _stream(stream))
    return _lib.check(rc, "dgemm_ozaki_crt")


def sgemm_bf16x(transa, transb, m, n, k, alpha, a: DeviceMatrix, b: DeviceMatrix, beta, c: DeviceMatrix, nprod=3,
                stream=None):
    """SGEMM on bf16 matrix cores from exact bf16 splits: nprod=3 (bf16x3) or 6 (bf16x6)."""
    rc = _lib.lib().gemm_amd_sgemm_bf16x(_enc(transa), _enc(transb), m, n, k, alpha, a.ptr, a.ld, b.ptr, b.ld, beta,
                                         c.ptr, c.ld, nprod, _stream(stream))
    return _lib.check(rc, "sgemm_bf16x")
