// Host-side declarations of the 3M complex GEMM launchers.
#pragma once
#include "gemm_kernels.hpp"

namespace gemm_amd {

// Scratch: three planes of op(A), op(B) and the three real products.
size_t gemm3m_workspace_bytes(int m, int n, int k, int elem_bytes);

// C <- alpha*op(A)*op(B) + beta*C, complex interleaved, op in {N, T, C};
// alpha/beta point to (re, im) on the host.  The three real products run with
// `cfg`; gemm_ws holds gemm_workspace_bytes(m, n, k, cfg, sizeof(real)).
int launch_zgemm3m(char ta, char tb, int m, int n, int k, const double *alpha, const double *A, int64_t lda,
                   const double *B, int64_t ldb, const double *beta, double *C, int64_t ldc, void *ws, const GemmConfig &cfg,
                   void *gemm_ws, hipStream_t s);
int launch_cgemm3m(char ta, char tb, int m, int n, int k, const float *alpha, const float *A, int64_t lda,
                   const float *B, int64_t ldb, const float *beta, float *C, int64_t ldc, void *ws, const GemmConfig &cfg,
                   void *gemm_ws, hipStream_t s);

}  // namespace gemm_amd

namespace gemm_amd {
// planes padded to mp x kp (A) and kp x np (B) with zeros; t is 3 x mp*np
int launch_gemm3m_split_d(char ta, char tb, int m, int n, int k, int mp, int np, int kp, const double *A,
                          int64_t lda, const double *B, int64_t ldb, double *pa, double *pb, hipStream_t s);
int launch_gemm3m_combine_d(int m, int n, int mp, int np, const double *t, const double *alpha, const double *beta,
                            double *C, int64_t ldc, hipStream_t s);
}  // namespace gemm_amd
