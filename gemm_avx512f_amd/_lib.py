"""ctypes binding of libgemm_amd.so (the C ABI declared in include/gemm_amd.h).

The shared library is built in-tree by ``make -C gemm_avx512f_amd`` (or
``__graft_entry__.build()``).  There is no fallback: if the library is
missing or cannot be loaded, every call raises.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgemm_amd.so")

_lock = threading.Lock()
_lib = None

c_int = ctypes.c_int
c_long = ctypes.c_long
c_char = ctypes.c_char
c_float = ctypes.c_float
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t
c_vp = ctypes.c_void_p
c_ip = ctypes.POINTER(ctypes.c_int)
c_fp = ctypes.POINTER(ctypes.c_float)
c_dp = ctypes.POINTER(ctypes.c_double)
c_i64p = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); one entry per symbol of include/gemm_amd.h
SIGNATURES = {
    # Fortran-ABI drop-in (host pointers, everything by reference)
    "sgemm_": (None, [c_vp] * 13),
    "dgemm_": (None, [c_vp] * 13),
    "dtrsm_": (None, [c_vp] * 11),
    "dgetrf_": (None, [c_vp] * 6),
    # CBLAS (enum arguments as int) and ILP64 Fortran symbols
    "cblas_sgemm": (None, [c_int] * 6 + [c_float, c_vp, c_int, c_vp, c_int, c_float, c_vp, c_int]),
    "cblas_dgemm": (None, [c_int] * 6 + [c_double, c_vp, c_int, c_vp, c_int, c_double, c_vp, c_int]),
    "cblas_cgemm3m": (None, [c_int] * 6 + [c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_int]),
    "cblas_zgemm3m": (None, [c_int] * 6 + [c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_int]),
    "cblas_dtrsm": (None, [c_int] * 7 + [c_double, c_vp, c_int, c_vp, c_int]),
    "sgemm_64_": (None, [c_vp] * 13),
    "dgemm_64_": (None, [c_vp] * 13),
    "cgemm3m_64_": (None, [c_vp] * 13),
    "zgemm3m_64_": (None, [c_vp] * 13),
    "dtrsm_64_": (None, [c_vp] * 11),
    "dgetrf_64_": (None, [c_vp] * 6),
    "cgemm3m_": (None, [c_vp] * 13),
    "zgemm3m_": (None, [c_vp] * 13),
    "gemm_amd_cgemm3m": (c_int, [c_char, c_char, c_int, c_int, c_int, c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_vp,
                                 c_int, c_vp]),
    "gemm_amd_zgemm3m": (c_int, [c_char, c_char, c_int, c_int, c_int, c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_vp,
                                 c_int, c_vp]),
    # device-pointer variants
    "gemm_amd_dtrsm": (c_int, [c_char, c_char, c_char, c_char, c_int, c_int, c_double, c_vp, c_int, c_vp, c_int,
                               c_vp]),
    "gemm_amd_dtrsm_noalloc": (c_int, [c_char, c_char, c_char, c_char, c_int, c_int, c_double, c_vp, c_int, c_vp, c_int,
                               c_vp]),
    "gemm_amd_dtrsm_kernel": (c_int, [c_int, c_long, c_long, c_vp, c_vp, c_vp, c_long, c_long, c_vp]),
    "gemm_amd_dgetrf": (c_int, [c_int, c_int, c_vp, c_int, c_vp, c_vp, c_vp]),
    "gemm_amd_sgemm": (c_int, [c_char, c_char, c_int, c_int, c_int, c_float, c_vp, c_int, c_vp, c_int,
                               c_float, c_vp, c_int, c_vp]),
    "gemm_amd_dgemm": (c_int, [c_char, c_char, c_int, c_int, c_int, c_double, c_vp, c_int, c_vp, c_int,
                               c_double, c_vp, c_int, c_vp]),
    "gemm_amd_sgemm_cfg": (c_int, [c_char, c_char, c_int, c_int, c_int, c_float, c_vp, c_int, c_vp, c_int,
                                   c_float, c_vp, c_int, c_int, c_int, c_vp]),
    "gemm_amd_dgemm_cfg": (c_int, [c_char, c_char, c_int, c_int, c_int, c_double, c_vp, c_int, c_vp, c_int,
                                   c_double, c_vp, c_int, c_int, c_int, c_vp]),
    "gemm_amd_dgemm_ozaki": (c_int, [c_char, c_char, c_int, c_int, c_int, c_double, c_vp, c_int, c_vp, c_int,
                                     c_double, c_vp, c_int, c_int, c_vp]),
    "gemm_amd_ozaki_split": (c_int, [c_vp, ctypes.c_int64, ctypes.c_int64, c_int, c_int, c_int, c_vp, c_int, c_int,
                                     c_vp, c_vp]),
    "gemm_amd_dgemm_ozaki_crt": (c_int, [c_char, c_char, c_int, c_int, c_int, c_double, c_vp, c_int, c_vp, c_int,
                                         c_double, c_vp, c_int, c_int, c_vp]),
    "gemm_amd_emulation_counts": (c_int, [c_int, c_vp, c_vp]),
    "gemm_amd_emulation_counts_set": (c_int, [c_vp, c_vp, c_int]),
    "gemm_amd_ozaki_crt_bits": (c_int, [c_int, c_int]),
    "gemm_amd_ozaki_crt_split": (c_int, [c_vp, ctypes.c_int64, ctypes.c_int64, c_int, c_int, c_int, c_int, c_vp,
                                         c_int, c_int, c_vp, c_vp]),
    "gemm_amd_sgemm_bf16x": (c_int, [c_char, c_char, c_int, c_int, c_int, c_float, c_vp, c_int, c_vp, c_int,
                                     c_float, c_vp, c_int, c_int, c_vp]),
    "gemm_amd_bf16_split": (c_int, [c_vp, ctypes.c_int64, ctypes.c_int64, c_int, c_int, c_int, c_vp, c_int, c_int,
                                    c_vp]),
    "gemm_amd_grouped_create": (c_vp, [c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                       c_vp, c_vp, c_vp]),
    "gemm_amd_grouped_create_ex": (c_vp, [c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                       c_vp, c_vp, c_vp, c_int]),
    "gemm_amd_grouped_run": (c_int, [c_vp, c_vp]),
    "gemm_amd_grouped_capture": (c_int, [c_vp, c_vp]),
    "gemm_amd_grouped_replay": (c_int, [c_vp, c_vp]),
    "gemm_amd_grouped_tiles": (ctypes.c_int64, [c_vp]),
    "gemm_amd_grouped_set_schedule": (c_int, [c_vp, c_int, c_vp]),
    "gemm_amd_grouped_grid": (c_int, [c_vp]),
    "gemm_amd_grouped_destroy": (c_int, [c_vp]),
    "gemm_amd_zgemm3m_batched": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                         c_vp, c_vp, c_vp]),
    # autotuner, picker, training kernels
    "gemm_amd_tune_candidates": (c_int, [c_vp, c_vp]),
    "gemm_amd_tune_sweep": (c_int, [c_int, c_char, c_char, c_int, c_int, c_int, c_int, c_vp]),
    "gemm_amd_pick_config": (c_int, [c_int, c_char, c_char, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp]),
    "gemm_amd_picker_load": (c_int, [c_vp, c_int]),
    "gemm_amd_picker_set_mode": (c_int, [c_int]),
    "gemm_amd_picker_choose": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "gemm_amd_picker_features": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp]),
    "gemm_amd_tune_table_set": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "gemm_amd_tune_table_clear": (c_int, []),
    "gemm_amd_mlp_forward": (c_int, [c_vp, c_vp, c_int, c_vp, c_vp]),
    "gemm_amd_mlp_grad": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp]),
    "gemm_amd_mlp_adam": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_float, c_float, c_float, c_float, c_int, c_vp]),
    "gemm_amd_mlp_train": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_float,
                                   ctypes.c_uint64, c_vp]),
    # runtime
    "gemm_amd_init": (c_int, []),
    "gemm_amd_device_count": (c_int, []),
    "gemm_amd_device_info": (c_int, [c_i64p]),
    "gemm_amd_last_error": (ctypes.c_char_p, []),
    "gemm_amd_malloc": (c_vp, [c_size_t]),
    "gemm_amd_free": (c_int, [c_vp]),
    "gemm_amd_memcpy": (c_int, [c_vp, c_vp, c_size_t, c_int]),
    "gemm_amd_memset": (c_int, [c_vp, c_int, c_size_t]),
    "gemm_amd_synchronize": (c_int, []),
    "gemm_amd_stream_create": (c_vp, []),
    "gemm_amd_stream_destroy": (c_int, [c_vp]),
    "gemm_amd_stream_synchronize": (c_int, [c_vp]),
    "gemm_amd_event_create": (c_vp, []),
    "gemm_amd_event_destroy": (c_int, [c_vp]),
    "gemm_amd_event_record": (c_int, [c_vp, c_vp]),
    "gemm_amd_event_elapsed_ms": (c_int, [c_vp, c_vp, ctypes.POINTER(c_float)]),
    "gemm_amd_fill_uniform": (c_int, [c_vp, c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64,
                                      ctypes.c_uint64, c_int, c_vp]),
}


class GemmAmdError(RuntimeError):
    pass


def lib():
    """Load libgemm_amd.so once and attach the C signatures; raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise GemmAmdError(
                    f"{LIB_PATH} is not built; run `make -C {HERE}` or __graft_entry__.build()")
            handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def last_error():
    msg = lib().gemm_amd_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(rc, what):
    """Raise on a non-zero status from a device-pointer entry point."""
    if rc != 0:
        kind = "illegal argument #%d" % rc if rc > 0 else "HIP error %d" % (-rc)
        raise GemmAmdError(f"{what}: {kind}: {last_error()}")
    return rc
