// C-ABI entry points of the emulated-precision products: the int8-slice
// (Ozaki) DGEMM and the bf16-split SGEMM.
#include "runtime.hpp"
#include "ozaki.hpp"
#include "bf16x.hpp"
#include "gemm_kernels.hpp"
#include "../../include/gemm_amd.h"
#include <mutex>

using namespace gemm_amd;

namespace {

#include "emulation_counts.inc"

// Counts chosen per ceil(log2 K): the compiled-in autotuner table until a
// caller installs another (gemm_amd_emulation_counts_set).
std::mutex g_counts_mu;
int g_moduli[33], g_slices[33];
bool g_counts_init = false;

int log2_bucket(int k) {
    int b = 0;
    while (b < 32 && (1ll << b) < (long long)k) ++b;
    return b;
}

void counts_for(int k, int *nmod, int *nsl) {
    std::lock_guard<std::mutex> lk(g_counts_mu);
    if (!g_counts_init) {
        for (int i = 0; i < 33; ++i) {
            g_moduli[i] = kCrtAutoModuli[i];
            g_slices[i] = kSliceAutoCount[i];
        }
        g_counts_init = true;
    }
    const int b = log2_bucket(k);
    if (nmod) *nmod = g_moduli[b];
    if (nsl) *nsl = g_slices[b];
}

}  // namespace

extern "C" int gemm_amd_emulation_counts(int k, int *nmoduli, int *nslices) {
    if (k < 1) {
        set_error("gemm_amd_emulation_counts: k >= 1");
        return -(int)hipErrorInvalidValue;
    }
    counts_for(k, nmoduli, nslices);
    return 0;
}

extern "C" int gemm_amd_emulation_counts_set(const int *moduli, const int *slices, int n) {
    if (n != 33 || !moduli || !slices) {
        set_error("gemm_amd_emulation_counts_set: 33 entries (ceil(log2 K) = 0..32) of each");
        return -(int)hipErrorInvalidValue;
    }
    for (int i = 0; i < 33; ++i)
        if (moduli[i] < 2 || moduli[i] > OZ_CRT_MAXMOD || slices[i] < 1 || slices[i] > 8) {
            set_error("gemm_amd_emulation_counts_set: moduli 2..16 and slices 1..8");
            return -(int)hipErrorInvalidValue;
        }
    counts_for(1, nullptr, nullptr);   // initialise, then replace
    std::lock_guard<std::mutex> lk(g_counts_mu);
    for (int i = 0; i < 33; ++i) {
        g_moduli[i] = moduli[i];
        g_slices[i] = slices[i];
    }
    return 0;
}

extern "C" int gemm_amd_sgemm_bf16x(char ta, char tb, int m, int n, int k, float alpha, const float *a, int lda,
                                    const float *b, int ldb, float beta, float *c, int ldc, int nprod, void *stream) {
    const int info = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) {
        xerbla("SGEMM_BF16X", info);
        return info;
    }
    if (nprod != 3 && nprod != 6) {
        set_error("gemm_amd_sgemm_bf16x: nprod must be 3 (2-way split) or 6 (3-way split)");
        return -(int)hipErrorInvalidValue;
    }
    if (m == 0 || n == 0 || ((alpha == 0.0f || k == 0) && beta == 1.0f)) return 0;
    if (int st = init_status()) return st;
    hipStream_t s = (hipStream_t)stream;
    if (alpha == 0.0f || k == 0) return hip_status((hipError_t)launch_scale_s(m, n, beta, c, ldc, s), "SGEMM_BF16X");
    const int nsplit = nprod == 3 ? 2 : 3;
    int st = 0;
    WsLease lease(s);
    void *wa = lease.get(WS_EMU_A, bf16x_workspace_bytes(m, k, nsplit), &st);
    if (st) return st;
    void *wb = lease.get(WS_EMU_B, bf16x_workspace_bytes(n, k, nsplit), &st);
    if (st) return st;
    const int e = launch_sgemm_bf16x(!notrans(ta), !notrans(tb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nprod,
                                     wa, wb, s);
    return hip_status((hipError_t)e, "SGEMM_BF16X");
}

extern "C" int gemm_amd_bf16_split(const float *x, int64_t row_stride, int64_t k_stride, int rows, int k, int nsplit,
                                   uint16_t *planes, int rows_pad, int k_pad, void *stream) {
    if (rows < 0 || k < 0 || nsplit < 1 || nsplit > 3 || rows_pad < rows || k_pad < k || rows_pad % 64 != 0 ||
        k_pad % 64 != 0) {
        set_error("gemm_amd_bf16_split: need rows, k >= 0, nsplit in 1..3, rows_pad >= rows and k_pad >= k, "
                  "both multiples of 64");
        return -(int)hipErrorInvalidValue;
    }
    if (rows == 0 || k == 0) return 0;
    if (int st = init_status()) return st;
    return hip_status((hipError_t)launch_bf16_split(x, row_stride, k_stride, rows, k, nsplit, planes, rows_pad, k_pad,
                                                    (hipStream_t)stream),
                      "bf16_split");
}

extern "C" int gemm_amd_dgemm_ozaki(char ta, char tb, int m, int n, int k, double alpha, const double *a, int lda,
                                    const double *b, int ldb, double beta, double *c, int ldc, int nslices,
                                    void *stream) {
    const int info = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) {
        xerbla("DGEMM_OZAKI", info);
        return info;
    }
    if (nslices < 0 || nslices > 8) {
        set_error("gemm_amd_dgemm_ozaki: nslices must be 1..8, or 0 for the tuned count for this K");
        return -(int)hipErrorInvalidValue;
    }
    if (nslices == 0) counts_for(k > 0 ? k : 1, nullptr, &nslices);
    if (m == 0 || n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return 0;
    if (int st = init_status()) return st;
    hipStream_t s = (hipStream_t)stream;
    if (alpha == 0.0 || k == 0) return hip_status((hipError_t)launch_scale_d(m, n, beta, c, ldc, s), "DGEMM_OZAKI");
    int st = 0;
    WsLease lease(s);
    void *wa = lease.get(WS_EMU_A, ozaki_workspace_bytes(m, k, nslices), &st);
    if (st) return st;
    void *wb = lease.get(WS_EMU_B, ozaki_workspace_bytes(n, k, nslices), &st);
    if (st) return st;
    void *ww = lease.get(WS_EMU_W, ozaki_accum_bytes(m, n), &st);
    if (st) return st;
    const int e = launch_dgemm_ozaki(!notrans(ta), !notrans(tb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nslices,
                                     wa, wb, ww, s);
    return hip_status((hipError_t)e, "DGEMM_OZAKI");
}

extern "C" int gemm_amd_ozaki_split(const double *x, int64_t row_stride, int64_t k_stride, int rows, int k,
                                    int nslices, int8_t *slices, int rows_pad, int k_pad, int *exps, void *stream) {
    if (rows < 0 || k < 0 || nslices < 1 || nslices > 8 || rows_pad < rows || k_pad < k || rows_pad % 64 != 0 ||
        k_pad % 64 != 0) {
        set_error("gemm_amd_ozaki_split: need rows, k >= 0, nslices in 1..8, rows_pad >= rows and k_pad >= k, "
                  "both multiples of 64");
        return -(int)hipErrorInvalidValue;
    }
    if (rows == 0 || k == 0) return 0;
    if (int st = init_status()) return st;
    int st = 0;
    WsLease lease((hipStream_t)stream);
    void *rm = lease.get(WS_EMU_AUX, (size_t)rows_pad * 8, &st);
    if (st) return st;
    return hip_status((hipError_t)launch_ozaki_split_only(x, row_stride, k_stride, rows, k, nslices, slices, rows_pad,
                                                          k_pad, exps, (unsigned long long *)rm, (hipStream_t)stream),
                      "ozaki_split");
}

extern "C" int gemm_amd_dgemm_ozaki_crt(char ta, char tb, int m, int n, int k, double alpha, const double *a, int lda,
                                        const double *b, int ldb, double beta, double *c, int ldc, int nmoduli,
                                        void *stream) {
    const int info = gemm_check(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) {
        xerbla("DGEMM_OZAKI_CRT", info);
        return info;
    }
    if (nmoduli != 0 && (nmoduli < 2 || nmoduli > OZ_CRT_MAXMOD)) {
        set_error("gemm_amd_dgemm_ozaki_crt: nmoduli must be 2..16, or 0 for the tuned count for this K");
        return -(int)hipErrorInvalidValue;
    }
    if (nmoduli == 0) counts_for(k > 0 ? k : 1, &nmoduli, nullptr);
    if (m == 0 || n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return 0;
    if (ozaki_crt_bits(nmoduli, k) < 1) {
        set_error("gemm_amd_dgemm_ozaki_crt: too few moduli for this K (no bits left per operand)");
        return -(int)hipErrorInvalidValue;
    }
    if (int st = init_status()) return st;
    hipStream_t s = (hipStream_t)stream;
    if (alpha == 0.0 || k == 0)
        return hip_status((hipError_t)launch_scale_d(m, n, beta, c, ldc, s), "DGEMM_OZAKI_CRT");
    int st = 0;
    WsLease lease(s);
    void *wa = lease.get(WS_EMU_A, ozaki_crt_workspace_bytes(m, k, nmoduli), &st);
    if (st) return st;
    void *wb = lease.get(WS_EMU_B, ozaki_crt_workspace_bytes(n, k, nmoduli), &st);
    if (st) return st;
    void *ww = lease.get(WS_EMU_W, ozaki_crt_accum_bytes(m, n, nmoduli), &st);
    if (st) return st;
    const int e = launch_dgemm_ozaki_crt(!notrans(ta), !notrans(tb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                                         nmoduli, wa, wb, ww, s);
    return hip_status((hipError_t)e, "DGEMM_OZAKI_CRT");
}

extern "C" int gemm_amd_ozaki_crt_bits(int nmoduli, int k) {
    if (nmoduli < 2 || nmoduli > OZ_CRT_MAXMOD || k < 1) return -1;
    return ozaki_crt_bits(nmoduli, k);
}

extern "C" int gemm_amd_ozaki_crt_split(const double *x, int64_t row_stride, int64_t k_stride, int rows, int k,
                                        int nmoduli, int bits, int8_t *residues, int rows_pad, int k_pad, int *exps,
                                        void *stream) {
    if (rows < 0 || k < 0 || nmoduli < 2 || nmoduli > OZ_CRT_MAXMOD || bits < 1 || bits > 56 || rows_pad < rows ||
        k_pad < k || rows_pad % 64 != 0 || k_pad % 64 != 0) {
        set_error("gemm_amd_ozaki_crt_split: need rows, k >= 0, nmoduli in 2..16, bits in 1..56, rows_pad >= rows "
                  "and k_pad >= k, both multiples of 64");
        return -(int)hipErrorInvalidValue;
    }
    if (rows == 0 || k == 0) return 0;
    if (int st = init_status()) return st;
    int st = 0;
    WsLease lease((hipStream_t)stream);
    void *rm = lease.get(WS_EMU_AUX, (size_t)rows_pad * 8, &st);
    if (st) return st;
    return hip_status((hipError_t)launch_ozaki_crt_split_only(x, row_stride, k_stride, rows, k, nmoduli, bits,
                                                              residues, rows_pad, k_pad, exps,
                                                              (unsigned long long *)rm, (hipStream_t)stream),
                      "ozaki_crt_split");
}
