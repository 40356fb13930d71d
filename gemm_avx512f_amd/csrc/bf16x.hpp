// Host-side declarations of the bf16-split SGEMM.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace gemm_amd {

int bf16x_pad_rows(int x);
int bf16x_pad_k(int x);
size_t bf16x_workspace_bytes(int rows, int k, int nsplit);
// Part planes [nsplit][Rp][Kp] (bf16, k contiguous) of the rows of the R x K
// view of X (element (r, k) at X[r*rs + k*ks]); Rp, Kp multiples of 64.
int launch_bf16_split(const float *X, int64_t rs, int64_t ks, int R, int K, int nsplit, uint16_t *planes, int Rp,
                      int Kp, hipStream_t s);
// C <- alpha*op(A)*op(B) + beta*C with nprod = 3 (2-way split) or 6 (3-way).
int launch_sgemm_bf16x(bool ta, bool tb, int m, int n, int k, float alpha, const float *A, int64_t lda,
                       const float *B, int64_t ldb, float beta, float *C, int64_t ldc, int nprod, void *wsA, void *wsB,
                       hipStream_t s);

}  // namespace gemm_amd
