// C-ABI GEMM entry points: the Fortran-convention drop-in sgemm_/dgemm_
// (host pointers) and their device-pointer variants.
//
// Contract (src/gemm_driver.c:116-195): C <- alpha*op(A)*op(B) + beta*C with
// op(X) = X for 'N'/'n' and X^T otherwise (:18-19); beta != 1 scales C even
// when alpha == 0 (:131); alpha == 0 skips the product (:132); K == 0 leaves
// beta*C.  Added over the reference: netlib argument checks with xerbla
// messages (the reference validates nothing) and the netlib beta == 0 rule
// (C is overwritten, never read).
#include "runtime.hpp"
#include "gemm_kernels.hpp"
#include "tuner.hpp"
#include "../../include/gemm_amd.h"

using namespace gemm_amd;

namespace {

template <typename T>
int launch(bool ta, bool tb, int m, int n, int k, T alpha, const T *a, int lda, const T *b, int ldb, T beta,
           T *c, int ldc, const GemmConfig &cfg, void *work, hipStream_t s) {
    if constexpr (sizeof(T) == 8)
        return launch_dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, cfg, work, s);
    else
        return launch_sgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, cfg, work, s);
}

// Device-pointer GEMM with a chosen or given config.
template <typename T>
int gemm_device(const char *name, char ta, char tb, int m, int n, int k, T alpha, const T *a, int lda,
                const T *b, int ldb, T beta, T *c, int ldc, const GemmConfig *force, hipStream_t s) {
    const int info = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) {
        xerbla(name, info);
        return info;
    }
    if (m == 0 || n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return 0;
    if (int st = init_status()) return st;
    GemmConfig cfg = force ? *force : pick_gemm_config(sizeof(T) == 8 ? DT_F64 : DT_F32, !notrans(ta),
                                                       !notrans(tb), m, n, k, lda - (notrans(ta) ? m : k),
                                                       ldb - (notrans(tb) ? k : n), rt().cus);
    if (!force) cfg.kgroups = true;
    const size_t wbytes = gemm_workspace_bytes(m, n, k, cfg, (int)sizeof(T));
    if (!wbytes)
        return hip_status((hipError_t)launch<T>(!notrans(ta), !notrans(tb), m, n, k, alpha, a, lda, b, ldb, beta, c,
                                                ldc, cfg, nullptr, s),
                          name);
    WsLease lease(s);
    int st = 0;
    void *work = lease.get(WS_GEMM, wbytes, &st);
    if (st) return st;
    int e = launch<T>(!notrans(ta), !notrans(tb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, cfg, work, s);
    return hip_status((hipError_t)e, name);
}

// Host-pointer drop-in: stage op(A), op(B) (and C when beta != 0) into the
// library's device buffers with pitched copies, run, copy C back.
template <typename T>
void gemm_host(const char *name, const char *transa, const char *transb, const int *pm, const int *pn,
               const int *pk, const T *palpha, const T *a, const int *plda, const T *b, const int *pldb,
               const T *pbeta, T *c, const int *pldc) {
    const char ta = *transa, tb = *transb;
    const int m = *pm, n = *pn, k = *pk, lda = *plda, ldb = *pldb, ldc = *pldc;
    const T alpha = *palpha, beta = *pbeta;
    const int info = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) {
        xerbla(name, info);
        return;
    }
    if (m == 0 || n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return;
    if (init_status()) {
        fprintf(stderr, "%s: %s\n", name, get_error());
        return;
    }
    Runtime &R = rt();
    std::lock_guard<std::mutex> lk(R.mu);
    hipStream_t s = R.stream;
    const bool product = alpha != T(0) && k > 0;
    const int nrowa = notrans(ta) ? m : k, ncola = notrans(ta) ? k : m;
    const int nrowb = notrans(tb) ? k : n, ncolb = notrans(tb) ? n : k;
    int st = 0;
    if (product) {
        st = buf(BUF_A).ensure((size_t)nrowa * ncola * sizeof(T));
        if (!st) st = buf(BUF_B).ensure((size_t)nrowb * ncolb * sizeof(T));
    }
    if (!st) st = buf(BUF_C).ensure((size_t)m * n * sizeof(T));
    if (st) {
        fprintf(stderr, "%s: %s\n", name, get_error());
        return;
    }
    T *dA = (T *)buf(BUF_A).ptr, *dB = (T *)buf(BUF_B).ptr, *dC = (T *)buf(BUF_C).ptr;
    hipError_t e = hipSuccess;
    if (product) {
        e = hipMemcpy2DAsync(dA, (size_t)nrowa * sizeof(T), a, (size_t)lda * sizeof(T), (size_t)nrowa * sizeof(T),
                             ncola, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(dB, (size_t)nrowb * sizeof(T), b, (size_t)ldb * sizeof(T),
                                 (size_t)nrowb * sizeof(T), ncolb, hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess && beta != T(0))
        e = hipMemcpy2DAsync(dC, (size_t)m * sizeof(T), c, (size_t)ldc * sizeof(T), (size_t)m * sizeof(T), n,
                             hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        hip_status(e, "hipMemcpy2DAsync (upload)");
        fprintf(stderr, "%s: %s\n", name, get_error());
        return;
    }
    int rc;
    if (product) {
        // the staged copies are packed (leading dimension = rows)
        GemmConfig cfg = pick_gemm_config(sizeof(T) == 8 ? DT_F64 : DT_F32, !notrans(ta), !notrans(tb), m, n, k, 0,
                                          0, R.cus);
        cfg.kgroups = true;
        const size_t wbytes = gemm_workspace_bytes(m, n, k, cfg, (int)sizeof(T));
        if (wbytes && (rc = buf(BUF_WORK).ensure(wbytes))) {
            fprintf(stderr, "%s: %s\n", name, get_error());
            return;
        }
        rc = launch<T>(!notrans(ta), !notrans(tb), m, n, k, alpha, dA, nrowa, dB, nrowb, beta, dC, m, cfg,
                       buf(BUF_WORK).ptr, s);
    } else {
        rc = (sizeof(T) == 8) ? launch_scale_d(m, n, (double)beta, (double *)dC, m, s)
                              : launch_scale_s(m, n, (float)beta, (float *)dC, m, s);
    }
    if (rc) {
        hip_status((hipError_t)rc, "kernel launch");
        fprintf(stderr, "%s: %s\n", name, get_error());
        return;
    }
    e = hipMemcpy2DAsync(c, (size_t)ldc * sizeof(T), dC, (size_t)m * sizeof(T), (size_t)m * sizeof(T), n,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        hip_status(e, "download/synchronize");
        fprintf(stderr, "%s: %s\n", name, get_error());
    }
}

}  // namespace

extern "C" {

void sgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k, const float *alpha,
            const float *a, const int *lda, const float *b, const int *ldb, const float *beta, float *c,
            const int *ldc) {
    gemm_host<float>("SGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void dgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k, const double *alpha,
            const double *a, const int *lda, const double *b, const int *ldb, const double *beta, double *c,
            const int *ldc) {
    gemm_host<double>("DGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

int gemm_amd_sgemm(char ta, char tb, int m, int n, int k, float alpha, const float *a, int lda, const float *b,
                   int ldb, float beta, float *c, int ldc, void *stream) {
    return gemm_device<float>("SGEMM", ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nullptr,
                              (hipStream_t)stream);
}

int gemm_amd_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *a, int lda, const double *b,
                   int ldb, double beta, double *c, int ldc, void *stream) {
    return gemm_device<double>("DGEMM", ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nullptr,
                               (hipStream_t)stream);
}

int gemm_amd_sgemm_cfg(char ta, char tb, int m, int n, int k, float alpha, const float *a, int lda, const float *b,
                       int ldb, float beta, float *c, int ldc, int tile, int splitk, void *stream) {
    if (tile < 0 || tile >= TILE_COUNT || splitk < 1) {
        set_error("gemm_amd_sgemm_cfg: tile must be 0..3 and splitk >= 1");
        return -(int)hipErrorInvalidValue;
    }
    GemmConfig cfg;
    cfg.tile = tile;
    cfg.splitk = splitk;
    return gemm_device<float>("SGEMM", ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, &cfg,
                              (hipStream_t)stream);
}

int gemm_amd_dgemm_cfg(char ta, char tb, int m, int n, int k, double alpha, const double *a, int lda,
                       const double *b, int ldb, double beta, double *c, int ldc, int tile, int splitk,
                       void *stream) {
    if (tile < 0 || tile >= TILE_COUNT || splitk < 1) {
        set_error("gemm_amd_dgemm_cfg: tile must be 0..3 and splitk >= 1");
        return -(int)hipErrorInvalidValue;
    }
    GemmConfig cfg;
    cfg.tile = tile;
    cfg.splitk = splitk;
    return gemm_device<double>("DGEMM", ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, &cfg,
                               (hipStream_t)stream);
}

}  // extern "C"
