// C-ABI LU entry points: LAPACK-convention dgetrf_ (host pointers) and the
// device-pointer, stream-ordered variant.
#include "runtime.hpp"
#include "getrf.hpp"
#include "trsm.hpp"
#include "../../include/gemm_amd.h"
#include <cstdio>
#include <string>

using namespace gemm_amd;

namespace {

// LAPACK dgetrf argument order: m, n, a, lda, ipiv, info
int getrf_check(int m, int n, int lda) {
    if (m < 0) return 1;
    if (n < 0) return 2;
    if (lda < (m > 1 ? m : 1)) return 4;
    return 0;
}

int getrf_run(int m, int n, double *a, int lda, int *ipiv, int *info, hipStream_t s) {
    const int mn = m < n ? m : n;
    int st = 0;
    WsLease lease(s);
    double *inv = (double *)lease.get(WS_TRSM_INV, trsm_workspace_inv_bytes(mn > 0 ? mn : 1), &st);
    if (st) return st;
    const size_t gbytes = getrf_gemm_workspace_bytes(m, n);
    void *gws = lease.get(WS_GEMM, gbytes, &st);
    if (st) return st;
    const int e = launch_dgetrf(m, n, a, lda, ipiv, info, inv, gws, gbytes, rt().cus, s);
    return hip_status((hipError_t)e, "DGETRF");
}

}  // namespace

extern "C" {

int gemm_amd_dgetrf(int m, int n, double *a, int lda, int *ipiv, int *info, void *stream) {
    const int arg = getrf_check(m, n, lda);
    if (arg) {
        xerbla("DGETRF", arg);
        return arg;
    }
    if (int st = init_status()) return st;
    return getrf_run(m, n, a, lda, ipiv, info, (hipStream_t)stream);
}

void dgetrf_(const int *pm, const int *pn, double *a, const int *plda, int *ipiv, int *info) {
    const int m = *pm, n = *pn, lda = *plda;
    const int arg = getrf_check(m, n, lda);
    if (arg) {
        *info = -arg;
        xerbla("DGETRF", arg);
        return;
    }
    *info = 0;
    if (m == 0 || n == 0) return;
    // every failure below leaves info = GEMM_AMD_INFO_RUNTIME_ERROR (the
    // matrix and ipiv are then not a factorisation)
    if (init_status()) {
        *info = GEMM_AMD_INFO_RUNTIME_ERROR;
        fprintf(stderr, "DGETRF: %s\n", get_error());
        return;
    }
    Runtime &R = rt();
    std::lock_guard<std::mutex> lk(R.mu);
    hipStream_t s = R.stream;
    const int mn = m < n ? m : n;
    int st = buf(BUF_A).ensure((size_t)m * n * sizeof(double));
    if (!st) st = buf(BUF_WORK).ensure((size_t)(mn + 1) * sizeof(int));
    if (st) {
        *info = GEMM_AMD_INFO_RUNTIME_ERROR;
        fprintf(stderr, "DGETRF: %s\n", get_error());
        return;
    }
    double *dA = (double *)buf(BUF_A).ptr;
    int *dpiv = (int *)buf(BUF_WORK).ptr, *dinfo = dpiv + mn;
    hipError_t e = hipMemcpy2DAsync(dA, (size_t)m * 8, a, (size_t)lda * 8, (size_t)m * 8, n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        hip_status(e, "DGETRF upload");
        *info = GEMM_AMD_INFO_RUNTIME_ERROR;
        fprintf(stderr, "DGETRF: %s\n", get_error());
        return;
    }
    if (getrf_run(m, n, dA, m, dpiv, dinfo, s)) {
        *info = GEMM_AMD_INFO_RUNTIME_ERROR;
        fprintf(stderr, "DGETRF: %s\n", get_error());
        return;
    }
    e = hipMemcpy2DAsync(a, (size_t)lda * 8, dA, (size_t)m * 8, (size_t)m * 8, n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ipiv, dpiv, (size_t)mn * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        hip_status(e, "DGETRF download");
        *info = GEMM_AMD_INFO_RUNTIME_ERROR;
        fprintf(stderr, "DGETRF: %s\n", get_error());
    }
}

}  // extern "C"
