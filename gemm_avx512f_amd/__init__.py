"""MI355X (gfx950) BLAS-3: a drop-in for the column-major sgemm_/dgemm_ of
wjc404/GEMM_AVX512F, running hand-written HIP kernels on one GPU.

The compute lives in libgemm_amd.so (C ABI: include/gemm_amd.h); this
package is a thin Python mirror of that interface.
"""
from ._lib import GemmAmdError, LIB_PATH, lib, last_error  # noqa: F401
from .device import DeviceBuffer, DeviceMatrix, Event, Stream, device_info, synchronize  # noqa: F401
from .blas import (cgemm3m, cgemm3m_, dgemm, dgemm_, dgemm_ozaki, dtrsm, dtrsm_, dtrsm_kernel, sgemm, sgemm_,  # noqa: F401
                   sgemm_bf16x, zgemm3m, zgemm3m_)
from .grouped import GROUPED_BF16, GROUPED_F32, GROUPED_F64, GroupedGemm, zgemm3m_batched  # noqa: F401
