/*
 * CPU oracle for the BLAS-3 hot path (sgemm_/dgemm_, dtrsm_, cgemm3m_/zgemm3m_).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and there only as the checker.
 *
 * Every routine is a plain restatement of the reference's arithmetic contract
 * (file:line citations below point into the reference repository
 * wjc404/GEMM_AVX512F) computed in long double or __float128 so that it is
 * more accurate than the FP32/FP64 results it checks.
 */
#ifndef GEMM_AMD_ORACLE_H
#define GEMM_AMD_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded inputs: one splitmix64 stream fills A, then B, then C ---- */
uint64_t oracle_splitmix64_next(uint64_t *state);
/* column-major fill of the rows x cols region of a matrix with leading dim ld */
void oracle_fill_f64(uint64_t *state, int64_t rows, int64_t cols, int64_t ld, double *x);
void oracle_fill_f32(uint64_t *state, int64_t rows, int64_t cols, int64_t ld, float *x);
/* dynamic-range fill: u * 2^e, e uniform integer in [-r, r] (u from the same stream) */
void oracle_fill_f64_range(uint64_t *state, int64_t rows, int64_t cols, int64_t ld, int r, double *x);

/* ---- GEMM: C <- alpha*op(A)*op(B) + beta*C (src/gemm_driver.c:116-195) ---- */
void oracle_dgemm(char transa, char transb, int m, int n, int k, double alpha,
                  const double *a, int lda, const double *b, int ldb,
                  double beta, double *c, int ldc);
void oracle_sgemm(char transa, char transb, int m, int n, int k, float alpha,
                  const float *a, int lda, const float *b, int ldb,
                  float beta, float *c, int ldc);
/* same product, result kept as a double-double (hi + lo) from __float128 sums;
 * only the entries listed in (rows[i], cols[i]) are computed (nsample of them)
 * or every entry when rows == 0. */
void oracle_dgemm_q(char transa, char transb, int m, int n, int k, double alpha,
                    const double *a, int lda, const double *b, int ldb,
                    double beta, const double *c, int ldc,
                    int64_t nsample, const int *rows, const int *cols,
                    double *out_hi, double *out_lo);

/* ---- DTRSM (BLAS semantics; substitution order of
 *      OpenBLAS-like_implementation/dtrsm_16x4/dtrsm_kernel_16x4_tester.c:97-199) ---- */
void oracle_dtrsm(char side, char uplo, char transa, char diag, int m, int n,
                  double alpha, const double *a, int lda, double *b, int ldb);
/* kernel-level solve contract of dtrsm_solve_{ln,lt,rn,rt}_m*n* :
 * variant 0=ln 1=lt 2=rn 3=rt; sa packed triangle with inverted diagonal,
 * sb pending update in 4-column panels, written back with the solution. */
void oracle_dtrsm_kernel(int variant, long M, long N, double *sa, double *sb,
                         double *C, long LDB, long LDC);

/* ---- complex GEMM (4M, long double), interleaved storage, ld in complex units ---- */
void oracle_zgemm(char transa, char transb, int m, int n, int k, const double *alpha,
                  const double *a, int lda, const double *b, int ldb,
                  const double *beta, double *c, int ldc);
void oracle_cgemm(char transa, char transb, int m, int n, int k, const float *alpha,
                  const float *a, int lda, const float *b, int ldb,
                  const float *beta, float *c, int ldc);

/* ---- error metrics vs a double-double reference (ref_lo may be 0) ---- */
/* out[0] = max_ij |c-r| / max_ij |r|,  out[1] = ||c-r||_F / ||r||_F,
 * out[2] = max_ij |c-r| / |r| over r != 0 (elementwise relative)       */
void oracle_err_f64(int64_t n, const double *c, const double *ref_hi, const double *ref_lo, double *out);
void oracle_err_f32(int64_t n, const float *c, const double *ref_hi, double *out);

/* ---- exactly-defined split steps of the emulated paths (restated, not measured) ---- */
/* Ozaki int8 slicing of op(A) rows (or op(B) columns): see gemm_avx512f_amd/csrc/ozaki.hip */
void oracle_ozaki_split(int rows, int k, const double *x, int64_t row_stride, int64_t k_stride,
                        int nslices, int8_t *slices, int *exps);
/* emulated DGEMM, modular form: integer scaling and symmetric residues */
void oracle_ozaki_crt_split(int rows, int k, const double *x, int64_t row_stride, int64_t k_stride, int nmod,
                            const int *moduli, int bits, int8_t *res /* nmod*rows*k */, int *exps);
/* bf16 round-to-nearest-even of an FP32 value, and the 2/3-way bf16 split */
uint16_t oracle_f32_to_bf16(float x);
float oracle_bf16_to_f32(uint16_t h);
void oracle_bf16_split(int64_t n, const float *x, int nsplit, uint16_t *parts /* nsplit*n */);

/* ---- CPU "port" baseline: cache-blocked FP64 GEMM (same contract as oracle_dgemm,
 *      FP64 accumulation), OpenMP across nthreads ---- */
void oracle_dgemm_blocked(char transa, char transb, int m, int n, int k, double alpha,
                          const double *a, int lda, const double *b, int ldb,
                          double beta, double *c, int ldc, int nthreads);

/* ---- LU with partial pivoting (LAPACK dgetf2 semantics; the reference has no
 *      LU, so this is the netlib contract): column-major m x n a overwritten by
 *      L (unit, below the diagonal) and U; ipiv 1-based; returns info (first
 *      zero pivot column, 1-based, or 0).  Pivot = first row of maximal |a|. ---- */
int oracle_dgetrf(int m, int n, double *a, int lda, int *ipiv);

/* ---- configuration picker's perceptron, in double (mlp_oracle.c) ---- */
#define ORACLE_MLP_IN 34
#define ORACLE_MLP_HID 64
#define ORACLE_MLP_OUT 20
#define ORACLE_MLP_NPARAMS (ORACLE_MLP_IN * ORACLE_MLP_HID + ORACLE_MLP_HID + ORACLE_MLP_HID * ORACLE_MLP_HID + \
                            ORACLE_MLP_HID + ORACLE_MLP_HID * ORACLE_MLP_OUT + ORACLE_MLP_OUT)
void oracle_mlp_forward(const double *p, const double *x, int batch, double *y);
double oracle_mlp_loss(const double *p, const double *x, const double *t, const double *mask, int batch);
void oracle_mlp_grad(const double *p, const double *x, const double *t, const double *mask, int batch, double *g,
                     double *loss);
void oracle_adam(double *p, const double *g, double *m, double *v, int n, double lr, double b1, double b2,
                 double eps, int step);
#ifdef __cplusplus
}
#endif
#endif
