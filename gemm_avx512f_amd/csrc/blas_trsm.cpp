// C-ABI DTRSM entry points: Fortran-convention dtrsm_ (host pointers), the
// device-pointer variant, and the kernel-level solve of the reference's
// dtrsm_16x4 family.
#include "runtime.hpp"
#include "trsm.hpp"
#include "../../include/gemm_amd.h"
#include <cstdio>

using namespace gemm_amd;

namespace gemm_amd {

// netlib dtrsm argument order: side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb
int trsm_check_args(char side, char uplo, char transa, char diag, int m, int n, int lda, int ldb) {
    const bool left = lsame(side, 'L');
    const int nrowa = left ? m : n;
    if (!left && !lsame(side, 'R')) return 1;
    if (!lsame(uplo, 'U') && !lsame(uplo, 'L')) return 2;
    if (!is_trans_char(transa)) return 3;
    if (!lsame(diag, 'U') && !lsame(diag, 'N')) return 4;
    if (m < 0) return 5;
    if (n < 0) return 6;
    if (lda < (nrowa > 1 ? nrowa : 1)) return 9;
    if (ldb < (m > 1 ? m : 1)) return 11;
    return 0;
}

}  // namespace gemm_amd

namespace {

int trsm_run(char side, char uplo, char transa, char diag, int m, int n, double alpha, const double *a, int lda,
             double *b, int ldb, hipStream_t s) {
    const bool left = lsame(side, 'L');
    const int na = left ? m : n;
    int st = 0;
    WsLease lease(s);
    double *inv = (double *)lease.get(WS_TRSM_INV, trsm_workspace_inv_bytes(na), &st);
    if (st) return st;
    // split-K partials of the widest trailing update: at most 8 partitions of an m x n (or smaller) panel
    const size_t gbytes = (size_t)8 * (size_t)m * (size_t)n * sizeof(double);
    void *gws = lease.get(WS_GEMM, gbytes, &st);
    if (st) return st;
    const int e = launch_dtrsm(left, lsame(uplo, 'U'), !notrans(transa), lsame(diag, 'U'), m, n, alpha, a, lda, b,
                               ldb, inv, gws, gbytes, rt().cus, s);
    return hip_status((hipError_t)e, "DTRSM");
}

}  // namespace

extern "C" {

int gemm_amd_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha, const double *a,
                   int lda, double *b, int ldb, void *stream) {
    const int info = trsm_check_args(side, uplo, transa, diag, m, n, lda, ldb);
    if (info) {
        xerbla("DTRSM", info);
        return info;
    }
    if (m == 0 || n == 0) return 0;
    if (int st = init_status()) return st;
    return trsm_run(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, (hipStream_t)stream);
}

void dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *pm, const int *pn,
            const double *palpha, const double *a, const int *plda, double *b, const int *pldb) {
    const int m = *pm, n = *pn, lda = *plda, ldb = *pldb;
    const int info = trsm_check_args(*side, *uplo, *transa, *diag, m, n, lda, ldb);
    if (info) {
        xerbla("DTRSM", info);
        return;
    }
    if (m == 0 || n == 0) return;
    if (init_status()) {
        fprintf(stderr, "DTRSM: %s\n", get_error());
        return;
    }
    Runtime &R = rt();
    std::lock_guard<std::mutex> lk(R.mu);
    hipStream_t s = R.stream;
    const int na = lsame(*side, 'L') ? m : n;
    int st = buf(BUF_A).ensure((size_t)na * na * sizeof(double));
    if (!st) st = buf(BUF_B).ensure((size_t)m * n * sizeof(double));
    if (st) {
        fprintf(stderr, "DTRSM: %s\n", get_error());
        return;
    }
    double *dA = (double *)buf(BUF_A).ptr, *dB = (double *)buf(BUF_B).ptr;
    hipError_t e = hipSuccess;
    if (*palpha != 0.0)
        e = hipMemcpy2DAsync(dA, (size_t)na * 8, a, (size_t)lda * 8, (size_t)na * 8, na, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(dB, (size_t)m * 8, b, (size_t)ldb * 8, (size_t)m * 8, n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        hip_status(e, "DTRSM upload");
        fprintf(stderr, "DTRSM: %s\n", get_error());
        return;
    }
    int rc = trsm_run(*side, *uplo, *transa, *diag, m, n, *palpha, dA, na, dB, m, s);
    if (rc) {
        fprintf(stderr, "DTRSM: %s\n", get_error());
        return;
    }
    e = hipMemcpy2DAsync(b, (size_t)ldb * 8, dB, (size_t)m * 8, (size_t)m * 8, n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        hip_status(e, "DTRSM download");
        fprintf(stderr, "DTRSM: %s\n", get_error());
    }
}

int gemm_amd_dtrsm_kernel(int variant, long M, long N, double *sa, double *sb, double *c, long LDB, long LDC,
                          void *stream) {
    if (variant < 0 || variant > 3 || M < 0 || N < 0 || LDC < (M > 1 ? M : 1)) {
        set_error("gemm_amd_dtrsm_kernel: variant must be 0..3 (ln, lt, rn, rt), M, N >= 0, LDC >= M");
        return -(int)hipErrorInvalidValue;
    }
    if (M == 0 || N == 0) return 0;
    if (int st = init_status()) return st;
    return hip_status((hipError_t)launch_dtrsm_kernel(variant, M, N, sa, sb, c, LDB, LDC, (hipStream_t)stream),
                      "dtrsm_kernel");
}

}  // extern "C"
