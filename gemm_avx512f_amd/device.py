"""Device memory, streams and events through libgemm_amd.so (no torch needed).

Matrices are column-major, like the BLAS they serve: a host matrix is a
numpy array in Fortran order, optionally the leading ``rows`` rows of an
``ld x cols`` allocation.
"""
import ctypes

import numpy as np

from . import _lib


class DeviceBuffer:
    """An owned device allocation of ``nbytes``."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        L = _lib.lib()
        self.ptr = L.gemm_amd_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise _lib.GemmAmdError("gemm_amd_malloc(%d): %s" % (self.nbytes, _lib.last_error()))

    def free(self):
        if self.ptr:
            _lib.lib().gemm_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr, offset=0):
        arr = np.asarray(arr)
        if not (arr.flags.f_contiguous or arr.flags.c_contiguous):
            arr = np.asfortranarray(arr)
        if offset + arr.nbytes > self.nbytes:
            raise ValueError("upload exceeds buffer")
        _lib.check(_lib.lib().gemm_amd_memcpy(ctypes.c_void_p(self.ptr + offset), arr.ctypes.data_as(ctypes.c_void_p),
                                              arr.nbytes, 1), "upload")

    def download(self, out, offset=0):
        if not (out.flags.f_contiguous or out.flags.c_contiguous):
            raise ValueError("download target must be contiguous")
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download exceeds buffer")
        _lib.check(_lib.lib().gemm_amd_memcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr + offset),
                                              out.nbytes, 2), "download")
        return out


class DeviceMatrix:
    """A column-major ``rows x cols`` matrix with leading dimension ``ld`` in HBM."""

    def __init__(self, rows, cols, dtype, ld=None):
        self.rows, self.cols = int(rows), int(cols)
        self.dtype = np.dtype(dtype)
        self.ld = int(ld) if ld is not None else max(1, self.rows)
        if self.ld < max(1, self.rows):
            raise ValueError("ld < rows")
        self.buf = DeviceBuffer(self.ld * max(self.cols, 1) * self.dtype.itemsize)

    @property
    def ptr(self):
        return self.buf.ptr

    @classmethod
    def from_host(cls, a, ld=None):
        """Upload a 2-D array (any order); its padding rows (if ld > rows) are zero."""
        a = np.asarray(a)
        m = cls(a.shape[0], a.shape[1], a.dtype, ld)
        full = np.zeros((m.ld, max(m.cols, 1)), dtype=a.dtype, order="F")
        full[: m.rows, : m.cols] = a
        m.buf.upload(full)
        return m

    @classmethod
    def from_padded(cls, full, rows):
        """Upload an ``ld x cols`` Fortran-order array whose first ``rows`` rows are the matrix."""
        full = np.asfortranarray(full)
        m = cls(rows, full.shape[1], full.dtype, full.shape[0])
        m.buf.upload(full)
        return m

    def to_host_padded(self):
        full = np.empty((self.ld, max(self.cols, 1)), dtype=self.dtype, order="F")
        self.buf.download(full)
        return full

    def to_host(self):
        return np.array(self.to_host_padded()[: self.rows, : self.cols], order="F")


FILL_F64, FILL_F32, FILL_F64_RANGE = 0, 1, 2


def fill_uniform(mat, seed, first_draw=0, kind=None, range_r=0, stream=None):
    """Fill a DeviceMatrix in HBM from the seeded splitmix64 stream (see
    gemm_amd_fill_uniform).  Returns the next unused draw number, so that
    A, then B, then C can be filled from one stream."""
    if kind is None:
        kind = FILL_F64 if mat.dtype == np.float64 else FILL_F32
    _lib.check(_lib.lib().gemm_amd_fill_uniform(mat.ptr, kind, mat.rows, mat.cols, mat.ld, seed, first_draw, range_r,
                                                stream.handle if stream else None), "fill_uniform")
    return first_draw + mat.rows * mat.cols * (2 if kind == FILL_F64_RANGE else 1)


def synchronize():
    _lib.check(_lib.lib().gemm_amd_synchronize(), "synchronize")


def device_info():
    info = (ctypes.c_int64 * 4)()
    _lib.check(_lib.lib().gemm_amd_device_info(info), "device_info")
    return {"cus": info[0], "clock_khz": info[1], "lds_bytes": info[2], "l2_bytes": info[3]}


class Stream:
    def __init__(self):
        self.handle = _lib.lib().gemm_amd_stream_create()
        if not self.handle:
            raise _lib.GemmAmdError("stream_create: " + _lib.last_error())

    def synchronize(self):
        _lib.check(_lib.lib().gemm_amd_stream_synchronize(self.handle), "stream_synchronize")

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().gemm_amd_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = _lib.lib().gemm_amd_event_create()
        if not self.handle:
            raise _lib.GemmAmdError("event_create: " + _lib.last_error())

    def record(self, stream=None):
        _lib.check(_lib.lib().gemm_amd_event_record(self.handle, stream.handle if stream else None), "event_record")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        _lib.check(_lib.lib().gemm_amd_event_elapsed_ms(self.handle, later.handle, ctypes.byref(ms)), "elapsed")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().gemm_amd_event_destroy(self.handle)
        except Exception:
            pass
