"""Grouped GEMM plans and the batched 3M complex GEMM (libgemm_amd.so).

A :class:`GroupedGemm` holds many independent column-major problems
``C_p <- alpha_p*op(A_p)*op(B_p) + beta_p*C_p`` (each with its own shape and
transposes) that run as one persistent launch.  ``run()`` launches it;
``capture(stream)`` records the launch into a hipGraph once and ``replay()``
relaunches that graph without any host-side work.
"""
import ctypes

import numpy as np

from . import _lib

GROUPED_F32, GROUPED_BF16, GROUPED_F64 = 0, 1, 2
_IN_DTYPE = {GROUPED_F32: np.float32, GROUPED_BF16: np.uint16, GROUPED_F64: np.float64}
_OUT_DTYPE = {GROUPED_F32: np.float32, GROUPED_BF16: np.float32, GROUPED_F64: np.float64}


def _arr(ctype, vals):
    return (ctype * max(len(vals), 1))(*vals)


def _chars(vals):
    return ctypes.create_string_buffer(b"".join(v.encode() if isinstance(v, str) else v for v in vals) or b"N")


class GroupedGemm:
    """A plan over ``problems``: a list of dicts with keys transa, transb, m, n,
    k, alpha, a, b, beta, c (a, b, c are DeviceMatrix; their ``ld`` is used).

    dtype: GROUPED_F32, GROUPED_BF16 (operands hold bf16 bit patterns as
    uint16, C is FP32) or GROUPED_F64.  large_min: smallest M and N of a
    problem on the large tile class (None: the library's default for the
    dtype: bf16 96, FP64 128, FP32 none).  dynamic: True, workgroups claim
    tiles from an atomic counter; False, they take the static stride; None,
    the library's choice (the counter with more than two tiles per
    workgroup)."""

    def __init__(self, dtype, problems, large_min=None, dynamic=None):
        self.dtype = dtype
        self.problems = list(problems)
        for p in self.problems:
            if p["a"].dtype != _IN_DTYPE[dtype] or p["b"].dtype != _IN_DTYPE[dtype] or \
                    p["c"].dtype != _OUT_DTYPE[dtype]:
                raise TypeError("operand dtypes do not match grouped dtype %d" % dtype)
        P = self.problems
        cnt = len(P)
        L = _lib.lib()
        self._keep = [
            _chars([p["transa"] for p in P]), _chars([p["transb"] for p in P]),
            _arr(ctypes.c_int, [p["m"] for p in P]), _arr(ctypes.c_int, [p["n"] for p in P]),
            _arr(ctypes.c_int, [p["k"] for p in P]), _arr(ctypes.c_double, [p["alpha"] for p in P]),
            _arr(ctypes.c_void_p, [p["a"].ptr for p in P]), _arr(ctypes.c_int, [p["a"].ld for p in P]),
            _arr(ctypes.c_void_p, [p["b"].ptr for p in P]), _arr(ctypes.c_int, [p["b"].ld for p in P]),
            _arr(ctypes.c_double, [p["beta"] for p in P]),
            _arr(ctypes.c_void_p, [p["c"].ptr for p in P]), _arr(ctypes.c_int, [p["c"].ld for p in P]),
        ]
        if large_min is None:
            self.handle = L.gemm_amd_grouped_create(dtype, cnt, *self._keep)
        else:
            self.handle = L.gemm_amd_grouped_create_ex(dtype, cnt, *self._keep, int(large_min))
        if not self.handle:
            raise _lib.GemmAmdError("grouped_create: " + _lib.last_error())
        self.dynamic = dynamic
        if dynamic is not None:
            _lib.check(L.gemm_amd_grouped_set_schedule(self.handle, int(bool(dynamic)), None), "grouped_set_schedule")

    @property
    def tiles(self):
        return int(_lib.lib().gemm_amd_grouped_tiles(self.handle))

    @property
    def grid(self):
        return int(_lib.lib().gemm_amd_grouped_grid(self.handle))

    def timeline(self, stream=None):
        """Runs the plan once (not captured) with per-workgroup timing and
        returns (start, end, tiles) arrays: start and end in microseconds from
        the first workgroup's start (100 MHz real-time counter), tiles
        processed per workgroup.  Call before capture()."""
        from .device import DeviceBuffer, synchronize
        L = _lib.lib()
        g = self.grid
        buf = DeviceBuffer(8 * 3 * g)
        dyn = int(self.dynamic if self.dynamic is not None else self.tiles > 2 * g)
        _lib.check(L.gemm_amd_grouped_set_schedule(self.handle, dyn, ctypes.c_void_p(buf.ptr)), "grouped_set_schedule")
        try:
            self.run(stream)
            synchronize()
            out = buf.download(np.zeros(3 * g, dtype=np.uint64)).reshape(g, 3)
        finally:
            _lib.check(L.gemm_amd_grouped_set_schedule(self.handle, dyn, None), "grouped_set_schedule")
        t0 = out[:, 0].min()
        return ((out[:, 0] - t0) / 100.0, (out[:, 1] - t0) / 100.0, out[:, 2].astype(np.int64))

    def run(self, stream=None):
        return _lib.check(_lib.lib().gemm_amd_grouped_run(self.handle, stream.handle if stream else None),
                          "grouped_run")

    def capture(self, stream):
        return _lib.check(_lib.lib().gemm_amd_grouped_capture(self.handle, stream.handle), "grouped_capture")

    def replay(self, stream=None):
        return _lib.check(_lib.lib().gemm_amd_grouped_replay(self.handle, stream.handle if stream else None),
                          "grouped_replay")

    def destroy(self):
        if self.handle:
            _lib.lib().gemm_amd_grouped_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def zgemm3m_batched(problems, stream=None):
    """Run a batch of complex-double 3M GEMMs (zgemm3m_ contract per problem):
    ``problems`` as for :class:`GroupedGemm`, alpha/beta complex, matrices
    complex128 DeviceMatrix."""
    P = list(problems)
    for p in P:
        for key in ("a", "b", "c"):
            if p[key].dtype != np.complex128:
                raise TypeError("zgemm3m_batched needs complex128 matrices")
    ab = []
    be = []
    for p in P:
        al, bt = complex(p["alpha"]), complex(p["beta"])
        ab += [al.real, al.imag]
        be += [bt.real, bt.imag]
    args = [
        _chars([p["transa"] for p in P]), _chars([p["transb"] for p in P]),
        _arr(ctypes.c_int, [p["m"] for p in P]), _arr(ctypes.c_int, [p["n"] for p in P]),
        _arr(ctypes.c_int, [p["k"] for p in P]), _arr(ctypes.c_double, ab),
        _arr(ctypes.c_void_p, [p["a"].ptr for p in P]), _arr(ctypes.c_int, [p["a"].ld for p in P]),
        _arr(ctypes.c_void_p, [p["b"].ptr for p in P]), _arr(ctypes.c_int, [p["b"].ld for p in P]),
        _arr(ctypes.c_double, be),
        _arr(ctypes.c_void_p, [p["c"].ptr for p in P]), _arr(ctypes.c_int, [p["c"].ld for p in P]),
    ]
    rc = _lib.lib().gemm_amd_zgemm3m_batched(len(P), *args, stream.handle if stream else None)
    return _lib.check(rc, "zgemm3m_batched")
