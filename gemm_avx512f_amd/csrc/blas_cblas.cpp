// CBLAS entry points (row- and column-major) and ILP64 Fortran entry points
// (64-bit integers, "64_" symbol suffix) over the LP64 Fortran-ABI routines.
// Both are argument mappings only: the work runs in the same kernels.
#include "runtime.hpp"
#include "../../include/gemm_amd.h"
#include <climits>
#include <cstdio>
#include <vector>

using namespace gemm_amd;

namespace {

enum {
    kRowMajor = 101, kColMajor = 102,
    kNoTrans = 111, kTrans = 112, kConjTrans = 113,
    kUpper = 121, kLower = 122,
    kNonUnit = 131, kUnit = 132,
    kLeft = 141, kRight = 142
};

char trans_char(int t) { return t == kNoTrans ? 'N' : t == kTrans ? 'T' : t == kConjTrans ? 'C' : 0; }
int imax1(int a) { return a > 1 ? a : 1; }

// CBLAS numbering of the gemm argument list: layout 1, TransA 2, TransB 3,
// M 4, N 5, K 6, lda 9, ldb 11, ldc 14.
int cblas_gemm_check(int layout, int ta, int tb, int m, int n, int k, int lda, int ldb, int ldc) {
    if (layout != kRowMajor && layout != kColMajor) return 1;
    if (!trans_char(ta)) return 2;
    if (!trans_char(tb)) return 3;
    if (m < 0) return 4;
    if (n < 0) return 5;
    if (k < 0) return 6;
    const bool na = ta == kNoTrans, nb = tb == kNoTrans;
    if (layout == kColMajor) {
        if (lda < imax1(na ? m : k)) return 9;
        if (ldb < imax1(nb ? k : n)) return 11;
        if (ldc < imax1(m)) return 14;
    } else {
        if (lda < imax1(na ? k : m)) return 9;
        if (ldb < imax1(nb ? n : k)) return 11;
        if (ldc < imax1(n)) return 14;
    }
    return 0;
}

// Row-major C = op(A) op(B) is column-major C^T = op(B)^T op(A)^T: the same
// Fortran call with the operands, their transposes and M/N exchanged.
template <typename T, typename F>
void cblas_gemm(const char *name, F fortran, int layout, int ta, int tb, int m, int n, int k, const T *alpha,
                const T *a, int lda, const T *b, int ldb, const T *beta, T *c, int ldc) {
    if (int info = cblas_gemm_check(layout, ta, tb, m, n, k, lda, ldb, ldc)) {
        xerbla(name, info);
        return;
    }
    const char cta = trans_char(ta), ctb = trans_char(tb);
    if (layout == kColMajor)
        fortran(&cta, &ctb, &m, &n, &k, alpha, a, &lda, b, &ldb, beta, c, &ldc);
    else
        fortran(&ctb, &cta, &n, &m, &k, alpha, b, &ldb, a, &lda, beta, c, &ldc);
}

bool fits32(int64_t v) { return v >= INT_MIN && v <= INT_MAX; }

// ILP64 values outside the 32-bit range the kernels index with: reported
// with the argument's Fortran position, nothing computed.
bool range_error(const char *name, int pos) {
    char msg[160];
    snprintf(msg, sizeof msg, " ** On entry to %-6s parameter number %2d exceeds the 32-bit range of this library",
             name, pos);
    fprintf(stderr, "%s\n", msg);
    set_error(msg);
    return true;
}

template <typename T, typename F>
void gemm64(const char *name, F fortran, const char *ta, const char *tb, const int64_t *m, const int64_t *n,
            const int64_t *k, const T *alpha, const T *a, const int64_t *lda, const T *b, const int64_t *ldb,
            const T *beta, T *c, const int64_t *ldc) {
    const int64_t v[6] = {*m, *n, *k, *lda, *ldb, *ldc};
    const int pos[6] = {3, 4, 5, 8, 10, 13};
    for (int i = 0; i < 6; ++i)
        if (!fits32(v[i]) && range_error(name, pos[i])) return;
    const int m32 = (int)*m, n32 = (int)*n, k32 = (int)*k, la = (int)*lda, lb = (int)*ldb, lc = (int)*ldc;
    fortran(ta, tb, &m32, &n32, &k32, alpha, a, &la, b, &lb, beta, c, &lc);
}

}  // namespace

extern "C" {

void cblas_sgemm(int layout, int transa, int transb, int m, int n, int k, float alpha, const float *a, int lda,
                 const float *b, int ldb, float beta, float *c, int ldc) {
    cblas_gemm<float>("cblas_sgemm", sgemm_, layout, transa, transb, m, n, k, &alpha, a, lda, b, ldb, &beta, c, ldc);
}

void cblas_dgemm(int layout, int transa, int transb, int m, int n, int k, double alpha, const double *a, int lda,
                 const double *b, int ldb, double beta, double *c, int ldc) {
    cblas_gemm<double>("cblas_dgemm", dgemm_, layout, transa, transb, m, n, k, &alpha, a, lda, b, ldb, &beta, c,
                       ldc);
}

void cblas_cgemm3m(int layout, int transa, int transb, int m, int n, int k, const void *alpha, const void *a,
                   int lda, const void *b, int ldb, const void *beta, void *c, int ldc) {
    cblas_gemm<float>("cblas_cgemm3m", cgemm3m_, layout, transa, transb, m, n, k, (const float *)alpha,
                      (const float *)a, lda, (const float *)b, ldb, (const float *)beta, (float *)c, ldc);
}

void cblas_zgemm3m(int layout, int transa, int transb, int m, int n, int k, const void *alpha, const void *a,
                   int lda, const void *b, int ldb, const void *beta, void *c, int ldc) {
    cblas_gemm<double>("cblas_zgemm3m", zgemm3m_, layout, transa, transb, m, n, k, (const double *)alpha,
                       (const double *)a, lda, (const double *)b, ldb, (const double *)beta, (double *)c, ldc);
}

// CBLAS numbering: layout 1, Side 2, Uplo 3, TransA 4, Diag 5, M 6, N 7, lda 10, ldb 12.
void cblas_dtrsm(int layout, int side, int uplo, int transa, int diag, int m, int n, double alpha, const double *a,
                 int lda, double *b, int ldb) {
    int info = 0;
    const int na = side == kLeft ? m : n;
    if (layout != kRowMajor && layout != kColMajor) info = 1;
    else if (side != kLeft && side != kRight) info = 2;
    else if (uplo != kUpper && uplo != kLower) info = 3;
    else if (!trans_char(transa)) info = 4;
    else if (diag != kUnit && diag != kNonUnit) info = 5;
    else if (m < 0) info = 6;
    else if (n < 0) info = 7;
    else if (lda < imax1(na)) info = 10;
    else if (ldb < imax1(layout == kColMajor ? m : n)) info = 12;
    if (info) {
        xerbla("cblas_dtrsm", info);
        return;
    }
    const char ct = trans_char(transa), cd = diag == kUnit ? 'U' : 'N';
    if (layout == kColMajor) {
        const char cs = side == kLeft ? 'L' : 'R', cu = uplo == kUpper ? 'U' : 'L';
        dtrsm_(&cs, &cu, &ct, &cd, &m, &n, &alpha, a, &lda, b, &ldb);
    } else {
        // row-major B (M x N) is column-major B^T: op(A) X = alpha B becomes
        // X^T op(A)^T = alpha B^T, and the stored A read column-major is A^T,
        // so side and triangle swap while the transpose flag stays
        const char cs = side == kLeft ? 'R' : 'L', cu = uplo == kUpper ? 'L' : 'U';
        dtrsm_(&cs, &cu, &ct, &cd, &n, &m, &alpha, a, &lda, b, &ldb);
    }
}

void sgemm_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n, const int64_t *k,
               const float *alpha, const float *a, const int64_t *lda, const float *b, const int64_t *ldb,
               const float *beta, float *c, const int64_t *ldc) {
    gemm64<float>("SGEMM", sgemm_, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void dgemm_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n, const int64_t *k,
               const double *alpha, const double *a, const int64_t *lda, const double *b, const int64_t *ldb,
               const double *beta, double *c, const int64_t *ldc) {
    gemm64<double>("DGEMM", dgemm_, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void cgemm3m_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n, const int64_t *k,
                 const float *alpha, const float *a, const int64_t *lda, const float *b, const int64_t *ldb,
                 const float *beta, float *c, const int64_t *ldc) {
    gemm64<float>("CGEMM3M", cgemm3m_, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void zgemm3m_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n, const int64_t *k,
                 const double *alpha, const double *a, const int64_t *lda, const double *b, const int64_t *ldb,
                 const double *beta, double *c, const int64_t *ldc) {
    gemm64<double>("ZGEMM3M", zgemm3m_, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void dtrsm_64_(const char *side, const char *uplo, const char *transa, const char *diag, const int64_t *m,
               const int64_t *n, const double *alpha, const double *a, const int64_t *lda, double *b,
               const int64_t *ldb) {
    const int64_t v[4] = {*m, *n, *lda, *ldb};
    const int pos[4] = {5, 6, 9, 11};
    for (int i = 0; i < 4; ++i)
        if (!fits32(v[i]) && range_error("DTRSM", pos[i])) return;
    const int m32 = (int)*m, n32 = (int)*n, la = (int)*lda, lb = (int)*ldb;
    dtrsm_(side, uplo, transa, diag, &m32, &n32, alpha, a, &la, b, &lb);
}

void dgetrf_64_(const int64_t *m, const int64_t *n, double *a, const int64_t *lda, int64_t *ipiv, int64_t *info) {
    const int64_t v[3] = {*m, *n, *lda};
    const int pos[3] = {1, 2, 4};
    for (int i = 0; i < 3; ++i)
        if (!fits32(v[i]) && range_error("DGETRF", pos[i])) {
            *info = -pos[i];
            return;
        }
    const int m32 = (int)*m, n32 = (int)*n, la = (int)*lda;
    const int mn = m32 < n32 ? m32 : n32;
    std::vector<int> piv(mn > 0 ? mn : 1);
    int info32 = 0;
    dgetrf_(&m32, &n32, a, &la, piv.data(), &info32);
    *info = info32;
    if (info32 >= 0)
        for (int i = 0; i < mn; ++i) ipiv[i] = piv[i];
}

}  // extern "C"
