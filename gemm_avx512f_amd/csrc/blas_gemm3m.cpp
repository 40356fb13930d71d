// C-ABI 3M complex GEMM entry points: cgemm3m_/zgemm3m_ (host pointers,
// OpenBLAS-extension argument list = that of cgemm_/zgemm_) and their
// device-pointer variants.
#include "runtime.hpp"
#include "gemm3m.hpp"
#include "tuner.hpp"
#include "../../include/gemm_amd.h"
#include <cstdio>

using namespace gemm_amd;

namespace {

template <typename T>
int run3m(char ta, char tb, int m, int n, int k, const T *alpha, const T *a, int lda, const T *b, int ldb,
          const T *beta, T *c, int ldc, hipStream_t s, const GemmConfig *force = nullptr) {
    int st = 0;
    WsLease lease(s);
    void *ws = lease.get(WS_COMPLEX, gemm3m_workspace_bytes(m, n, k, sizeof(T)), &st);
    if (st) return st;
    const int nrowa = notrans(ta) ? m : k;
    const GemmConfig cfg = force ? *force
                                 : pick_gemm_config(sizeof(T) == 8 ? DT_C64 : DT_C32, !notrans(ta), !notrans(tb), m,
                                                    n, k, lda - nrowa, ldb - (notrans(tb) ? k : n), rt().cus);
    void *gws = lease.get(WS_GEMM, gemm_workspace_bytes(m, n, k, cfg, (int)sizeof(T)), &st);
    if (st) return st;
    int e;
    if constexpr (sizeof(T) == 8)
        e = launch_zgemm3m(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ws, cfg, gws, s);
    else
        e = launch_cgemm3m(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ws, cfg, gws, s);
    return hip_status((hipError_t)e, sizeof(T) == 8 ? "ZGEMM3M" : "CGEMM3M");
}

template <typename T>
int device3m(const char *name, char ta, char tb, int m, int n, int k, const T *alpha, const T *a, int lda,
             const T *b, int ldb, const T *beta, T *c, int ldc, hipStream_t s) {
    const int info = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) {
        xerbla(name, info);
        return info;
    }
    const bool a0 = alpha[0] == T(0) && alpha[1] == T(0), b1 = beta[0] == T(1) && beta[1] == T(0);
    if (m == 0 || n == 0 || ((a0 || k == 0) && b1)) return 0;
    if (int st = init_status()) return st;
    return run3m<T>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
}

template <typename T>
void host3m(const char *name, const char *transa, const char *transb, const int *pm, const int *pn, const int *pk,
            const T *alpha, const T *a, const int *plda, const T *b, const int *pldb, const T *beta, T *c,
            const int *pldc) {
    const char ta = *transa, tb = *transb;
    const int m = *pm, n = *pn, k = *pk, lda = *plda, ldb = *pldb, ldc = *pldc;
    const int info = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) {
        xerbla(name, info);
        return;
    }
    const bool a0 = alpha[0] == T(0) && alpha[1] == T(0), b1 = beta[0] == T(1) && beta[1] == T(0);
    if (m == 0 || n == 0 || ((a0 || k == 0) && b1)) return;
    if (init_status()) {
        fprintf(stderr, "%s: %s\n", name, get_error());
        return;
    }
    Runtime &R = rt();
    std::lock_guard<std::mutex> lk(R.mu);
    hipStream_t s = R.stream;
    const size_t cb = 2 * sizeof(T);   // bytes per complex element
    const int nrowa = notrans(ta) ? m : k, ncola = notrans(ta) ? k : m;
    const int nrowb = notrans(tb) ? k : n, ncolb = notrans(tb) ? n : k;
    int st = buf(BUF_A).ensure((size_t)nrowa * ncola * cb);
    if (!st) st = buf(BUF_B).ensure((size_t)nrowb * ncolb * cb);
    if (!st) st = buf(BUF_C).ensure((size_t)m * n * cb);
    if (st) {
        fprintf(stderr, "%s: %s\n", name, get_error());
        return;
    }
    T *dA = (T *)buf(BUF_A).ptr, *dB = (T *)buf(BUF_B).ptr, *dC = (T *)buf(BUF_C).ptr;
    hipError_t e = hipMemcpy2DAsync(dA, nrowa * cb, a, (size_t)lda * cb, nrowa * cb, ncola, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(dB, nrowb * cb, b, (size_t)ldb * cb, nrowb * cb, ncolb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !(beta[0] == T(0) && beta[1] == T(0)))
        e = hipMemcpy2DAsync(dC, m * cb, c, (size_t)ldc * cb, m * cb, n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        hip_status(e, "upload");
        fprintf(stderr, "%s: %s\n", name, get_error());
        return;
    }
    if (run3m<T>(ta, tb, m, n, k, alpha, dA, nrowa, dB, nrowb, beta, dC, m, s)) {
        fprintf(stderr, "%s: %s\n", name, get_error());
        return;
    }
    e = hipMemcpy2DAsync(c, (size_t)ldc * cb, dC, m * cb, m * cb, n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        hip_status(e, "download");
        fprintf(stderr, "%s: %s\n", name, get_error());
    }
}

}  // namespace

extern "C" {

void cgemm3m_(const char *transa, const char *transb, const int *m, const int *n, const int *k, const float *alpha,
              const float *a, const int *lda, const float *b, const int *ldb, const float *beta, float *c,
              const int *ldc) {
    host3m<float>("CGEMM3M", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void zgemm3m_(const char *transa, const char *transb, const int *m, const int *n, const int *k, const double *alpha,
              const double *a, const int *lda, const double *b, const int *ldb, const double *beta, double *c,
              const int *ldc) {
    host3m<double>("ZGEMM3M", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

int gemm_amd_cgemm3m(char ta, char tb, int m, int n, int k, const float *alpha, const float *a, int lda,
                     const float *b, int ldb, const float *beta, float *c, int ldc, void *stream) {
    return device3m<float>("CGEMM3M", ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, (hipStream_t)stream);
}

int gemm_amd_zgemm3m(char ta, char tb, int m, int n, int k, const double *alpha, const double *a, int lda,
                     const double *b, int ldb, const double *beta, double *c, int ldc, void *stream) {
    return device3m<double>("ZGEMM3M", ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, (hipStream_t)stream);
}

}  // extern "C"
