// Host-side declarations of the DTRSM launchers.
#pragma once
#include "gemm_kernels.hpp"

namespace gemm_amd {

size_t trsm_workspace_inv_bytes(int na);

// netlib DTRSM argument check: 0, or the index of the first illegal argument.
int trsm_check_args(char side, char uplo, char transa, char diag, int m, int n, int lda, int ldb);

// Blocked recursive DTRSM on device pointers (op(A) = A or A^T, triangle =
// upper/lower as stored).  inv_ws: trsm_workspace_inv_bytes; gemm_ws may be
// null (no split-K).
int launch_dtrsm(bool left, bool upper, bool trans, bool unit, int m, int n, double alpha, const double *A,
                 int64_t lda, double *B, int64_t ldb, double *inv_ws, void *gemm_ws, size_t gemm_ws_bytes, int cus,
                 hipStream_t s);

// Kernel-level solve contract of dtrsm_solve_{ln,lt,rn,rt} (variant 0..3).
int launch_dtrsm_kernel(int variant, long M, long N, double *sa, double *sb, double *C, long LDB, long LDC,
                        hipStream_t s);

}  // namespace gemm_amd
