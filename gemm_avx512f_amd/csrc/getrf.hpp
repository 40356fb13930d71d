// Host-side declarations of the LU factorisation launcher.
#pragma once
#include "gemm_kernels.hpp"

namespace gemm_amd {

// Tallest matrix the register-resident leaf panels cover (16 rows per thread
// of one 1024-thread workgroup).
constexpr int GETRF_MAX_ROWS = 16384;

// Split-K workspace the Schur-complement GEMMs of an m x n factorisation may use.
size_t getrf_gemm_workspace_bytes(int m, int n);

// P A = L U in place (dgetrf contract).  ipiv (min(m,n) ints) and info (one
// int) are device pointers; inv_ws: trsm_workspace_inv_bytes(min(m, n)).
// Returns 0 or a hipError_t value (hipErrorInvalidValue when m > GETRF_MAX_ROWS).
int launch_dgetrf(int m, int n, double *A, int64_t lda, int *ipiv, int *info, double *inv_ws, void *gemm_ws,
                  size_t gemm_ws_bytes, int cus, hipStream_t s);

}  // namespace gemm_amd
