// DTRSM entry point that allocates nothing: gemm_amd_dtrsm_noalloc.
//
// Same computation as gemm_amd_dtrsm (blas_trsm.cpp) and the same kernels
// (launch_dtrsm, trsm.hip).  Its scratch memory -- the inverted 64x64
// diagonal blocks and the split-K partial sums of the trailing updates --
// lives in `scratch`, a fixed array in device memory that the code object
// brings with it, instead of in library workspace that grows on demand.  No
// counterpart in the reference; it is for callers that may not allocate
// (code under graph capture, tight serving loops, harnesses that forbid
// allocation).
//
// Layout of `scratch` (NOALLOC_SCRATCH_BYTES = 64 MiB):
//   [0, inv)        inverted diagonal blocks of the left-form triangle,
//                   trsm_workspace_inv_bytes(na) bytes, rounded up to 256;
//   [inv, end)      split-K partials of the trailing GEMMs: a trailing update
//                   whose chosen configuration needs more runs unsplit.
// na (m for side L, n for side R) is limited to NOALLOC_TRSM_MAX_NA = 32768,
// whose inverses take 16 MiB.  Every kernel writes what it later reads
// (the inverse kernel writes its blocks and the zero row the panel kernel
// reads past the edge; split-K partials are written before they are summed),
// so nothing in `scratch` needs clearing between calls.
//
// All calls on one GPU share that GPU's array: calls that overlap in time on
// one GPU (two streams, or two host threads) race on it and must be ordered
// by the caller.
#include "runtime.hpp"
#include "trsm.hpp"
#include "tuner.hpp"
#include "../../include/gemm_amd.h"
#include <mutex>
#include <string>

namespace gemm_amd {

constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
constexpr int NOALLOC_TRSM_MAX_NA = GEMM_AMD_DTRSM_NOALLOC_MAX_NA;

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

// Device address of `scratch` per device, looked up once per device with that
// device current (each GPU holds its own copy of the array).  In host code the
// array's name is only a host-side handle, never a device address.
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::once_flag g_scratch_once[NOALLOC_MAX_DEVICES];
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static hipError_t g_scratch_err[NOALLOC_MAX_DEVICES];

static int scratch_address(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_status(e, "hipGetDevice");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_error("gemm_amd_dtrsm_noalloc: device index beyond the supported 64");
        return -(int)hipErrorInvalidDevice;
    }
    std::call_once(g_scratch_once[dev], [dev] {
        void *p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? p : nullptr;
    });
    if (g_scratch_err[dev] != hipSuccess) return hip_status(g_scratch_err[dev], "hipGetSymbolAddress (scratch)");
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace gemm_amd

using namespace gemm_amd;

extern "C" int gemm_amd_dtrsm_noalloc(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                                      const double *a, int lda, double *b, int ldb, void *stream) {
    const int info = trsm_check_args(side, uplo, transa, diag, m, n, lda, ldb);
    if (info) {
        xerbla("DTRSM", info);
        return info;
    }
    if (m == 0 || n == 0) return 0;
    const bool left = lsame(side, 'L');
    const int na = left ? m : n;
    if (na > NOALLOC_TRSM_MAX_NA) {
        set_error("gemm_amd_dtrsm_noalloc: the triangle's order " + std::to_string(na) + " exceeds the maximum of " +
                  std::to_string(NOALLOC_TRSM_MAX_NA) + " that its fixed scratch holds (use gemm_amd_dtrsm)");
        return -(int)hipErrorInvalidValue;
    }
    if (int st = init_status()) return st;
    void *base = nullptr;
    if (int st = scratch_address(&base)) return st;
    const size_t inv = (trsm_workspace_inv_bytes(na) + 255) & ~(size_t)255;
    double *inv_ws = static_cast<double *>(base);
    void *gemm_ws = static_cast<char *>(base) + inv;
    // the trailing updates' configurations come from the sweep table, the
    // picker's cache or the heuristic: no inference (its first use allocates)
    PickWithoutDevice no_device;
    const int e = launch_dtrsm(left, lsame(uplo, 'U'), !notrans(transa), lsame(diag, 'U'), m, n, alpha, a, lda, b,
                               ldb, inv_ws, gemm_ws, NOALLOC_SCRATCH_BYTES - inv, rt().cus, (hipStream_t)stream);
    return hip_status((hipError_t)e, "DTRSM (no-alloc)");
}
