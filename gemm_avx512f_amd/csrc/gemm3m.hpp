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
// One operand's split: op(X) (rows x cols, op = transpose if trans, conjugate
// if conj) into planes re, im, re+im, each prows x pcols with zeros past
// rows and cols.
struct SplitJob3m {
    int rows, cols, prows, pcols, trans, conj;
    const double *X;
    int64_t ldx;
    double *planes;
};
// One problem's combine: C <- alpha * (T1 - T2 + i (T3 - T1 - T2)) + beta * C
// from the products t (3 planes of mt x nt, ld mt).
struct CombineJob3m {
    int m, n, mt, nt;
    const double *t;
    double ar, ai, br, bi;
    double *C;
    int64_t ldc;
};
// job tables in device memory
int launch_gemm3m_split_batch_d(const SplitJob3m *jobs, int njobs, hipStream_t s);
int launch_gemm3m_combine_batch_d(const CombineJob3m *jobs, int njobs, hipStream_t s);
}  // namespace gemm_amd
