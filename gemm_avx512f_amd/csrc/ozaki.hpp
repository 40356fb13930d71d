// Host-side declarations of the emulated (int8-slice) DGEMM and the
// bf16-split SGEMM.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace gemm_amd {

int ozaki_pad_rows(int x);   // output-tile multiple (256)
int ozaki_pad_k(int x);      // 128
// FP64 partial-sum buffer for an m x n product
size_t ozaki_accum_bytes(int m, int n);
// Scratch for one operand's slices (rows x k, nslices), exponents and row maxima.
size_t ozaki_workspace_bytes(int rows, int k, int nslices);
// Fills the (t, u) slice-product list for s slices ordered by t+u; returns its length.
int ozaki_pairs(int ns, int8_t *t, int8_t *u, int8_t *last);
int launch_dgemm_ozaki(bool ta, bool tb, int m, int n, int k, double alpha, const double *A, int64_t lda,
                       const double *B, int64_t ldb, double beta, double *C, int64_t ldc, int nslices, void *wsA,
                       void *wsB, void *wsW, hipStream_t s);

}  // namespace gemm_amd

namespace gemm_amd {

// Modular (CRT) form: N pairwise coprime moduli <= 256, one int8 GEMM each.
constexpr int OZ_CRT_MAXMOD = 16;
struct CrtConstHost {
    int nmod;
    double log2M;
    int m[OZ_CRT_MAXMOD];
    double w[OZ_CRT_MAXMOD][4];   // 32-bit limbs of the CRT weights W_l
    double Mk[4];                 // 32-bit limbs of M
    double invM;
};
void ozaki_crt_constants(int nmod, CrtConstHost *out);
// Bits kept per row/column (al) for K: the largest al with K 2^(2 al) <= M/4, at most 56.
int ozaki_crt_bits(int nmod, int k);
size_t ozaki_crt_workspace_bytes(int rows, int k, int nmod);
size_t ozaki_crt_accum_bytes(int m, int n, int nmod);
int launch_dgemm_ozaki_crt(bool ta, bool tb, int m, int n, int k, double alpha, const double *A, int64_t lda,
                           const double *B, int64_t ldb, double beta, double *C, int64_t ldc, int nmod, void *wsA,
                           void *wsB, void *wsW, hipStream_t s);
int launch_ozaki_crt_split_only(const double *X, int64_t rs, int64_t ks, int R, int K, int nmod, int al, int8_t *res,
                                int Rp, int Kp, int *exps, unsigned long long *rowmax_ws, hipStream_t s);

int launch_ozaki_split_only(const double *X, int64_t rs, int64_t ks, int R, int K, int ns, int8_t *slices,
                            int Rp, int Kp, int *exps, unsigned long long *rowmax_ws, hipStream_t s);
}  // namespace gemm_amd
