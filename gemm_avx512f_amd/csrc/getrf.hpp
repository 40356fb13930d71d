// Host-side declarations of the LU factorisation launcher.
#pragma once
#include "gemm_kernels.hpp"

namespace gemm_amd {

// Tallest panel the register-resident leaves cover (16 rows per thread of one
// 1024-thread workgroup); taller panels are factored in global memory.
constexpr int GETRF_REGISTER_ROWS = 16384;

// Split-K workspace the Schur-complement GEMMs of an m x n factorisation may use.
size_t getrf_gemm_workspace_bytes(int m, int n);

// P A = L U in place (dgetrf contract).  ipiv (min(m,n) ints) and info (one
// int) are device pointers; inv_ws: trsm_workspace_inv_bytes(min(m, n)).
// Returns 0 or a hipError_t value.
int launch_dgetrf(int m, int n, double *A, int64_t lda, int *ipiv, int *info, double *inv_ws, void *gemm_ws,
                  size_t gemm_ws_bytes, int cus, hipStream_t s);

}  // namespace gemm_amd
