// DTRSM on gfx950: blocked, recursive triangular solve whose bulk runs in the
// FP64 matrix-core GEMM kernel, plus the kernel-level solve contract of the
// reference's dtrsm_solve_{ln,lt,rn,rt}_m*n* family.
//
// BLAS dtrsm: side L solves op(A) X = alpha B, side R solves X op(A) = alpha B,
// X overwrites B.  Only the uplo triangle of A is read; diag 'U' never reads
// the diagonal.  Plan (replaces dtrsm_16x4's register solves,
// OpenBLAS-like_implementation/dtrsm_16x4/dtrsm_kernel_16x4_skylakex.S):
//   1. one kernel inverts every NB x NB diagonal block of op(A) in parallel
//      (column-parallel substitution against the identity; the reference
//      likewise works from inverted pivots, :1480-1484);
//   2. B <- alpha B;
//   3. recursive halving over the triangle: solve the leading half, update the
//      trailing half with one large GEMM (the FP64 MFMA kernel), solve the
//      trailing half; a leaf (<= NB) is X_b = inv(T_bb) * B_b, a GEMM into a
//      scratch panel copied back.
#include "common.hpp"
#include "gemm_kernels.hpp"
#include "trsm.hpp"

namespace gemm_amd {

static constexpr int TRSM_NB = 64;

// Element (r, c) of op(A) from column-major storage.
__device__ __forceinline__ double opa(const double *A, int64_t lda, bool trans, int64_t r, int64_t c) {
    return trans ? A[c + r * lda] : A[r + c * lda];
}

// grid.x = diagonal block index; 256 threads = 64 columns of the identity x 4
// row groups of 16 rows.  W[b] (NB x NB, ld NB) <- inv(op(A)_bb); rows/cols past
// the matrix edge are the identity.
__global__ __launch_bounds__(256) void trsm_diag_inverse_kernel(const double *__restrict__ A, int64_t lda, int na,
                                                                bool trans, bool oplower, bool unit,
                                                                double *__restrict__ W) {
    constexpr int NB = TRSM_NB;
    __shared__ double T[NB][NB + 1];   // op(A) block, [row][col]
    __shared__ double X[NB][NB + 1];   // solution, [row][col]
    __shared__ double piv[NB];
    const int b = blockIdx.x;
    const int64_t base = (int64_t)b * NB;
    const int nb = min(NB, na - (int)base);
    const int tid = threadIdx.x, c = tid & 63, rg = tid >> 6;
    for (int idx = tid; idx < NB * NB; idx += 256) {
        const int r = idx % NB, cc = idx / NB;
        double v = 0.0;
        const bool in_tri = oplower ? (r > cc) : (r < cc);
        if (r < nb && cc < nb && in_tri) v = opa(A, lda, trans, base + r, base + cc);
        T[r][cc] = v;
        X[r][cc] = (r == cc) ? 1.0 : 0.0;
    }
    if (tid < NB) {
        double d = 1.0;
        if (tid < nb && !unit) d = opa(A, lda, trans, base + tid, base + tid);
        piv[tid] = 1.0 / d;
    }
    __syncthreads();
    if (oplower) {
        for (int i = 0; i < NB; ++i) {
            if (rg == (i >> 4)) X[i][c] *= piv[i];
            __syncthreads();
            const double xi = X[i][c];
#pragma unroll 4
            for (int r = rg * 16; r < rg * 16 + 16; ++r)
                if (r > i) X[r][c] -= T[r][i] * xi;
            __syncthreads();
        }
    } else {
        for (int i = NB - 1; i >= 0; --i) {
            if (rg == (i >> 4)) X[i][c] *= piv[i];
            __syncthreads();
            const double xi = X[i][c];
#pragma unroll 4
            for (int r = rg * 16; r < rg * 16 + 16; ++r)
                if (r < i) X[r][c] -= T[r][i] * xi;
            __syncthreads();
        }
    }
    double *out = W + (int64_t)b * NB * NB;
    for (int idx = tid; idx < NB * NB; idx += 256) {
        const int r = idx % NB, cc = idx / NB;
        out[r + (int64_t)cc * NB] = X[r][cc];
    }
}

// B (rows x cols, ldb) <- alpha * B; alpha == 0 writes zeros without reading.
__global__ __launch_bounds__(256) void trsm_scale_kernel(int rows, int cols, double alpha, double *B, int64_t ldb) {
    const int64_t total = (int64_t)rows * cols;
    for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int64_t i = idx % rows, j = idx / rows;
        double *p = B + i + j * ldb;
        *p = alpha == 0.0 ? 0.0 : alpha * *p;
    }
}

// dst (rows x cols, ldd) <- src (rows x cols, lds)
__global__ __launch_bounds__(256) void copy_block_kernel(int rows, int cols, const double *__restrict__ src,
                                                         int64_t lds, double *__restrict__ dst, int64_t ldd) {
    const int64_t total = (int64_t)rows * cols;
    for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int64_t i = idx % rows, j = idx / rows;
        dst[i + j * ldd] = src[i + j * lds];
    }
}

static int grid_for(int64_t total) {
    int64_t b = (total + 255) / 256;
    return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

namespace {

struct TrsmCtx {
    bool left, trans, oplower;
    int m, n;
    const double *A;
    int64_t lda;
    double *B;
    int64_t ldb;
    const double *W;     // inverted diagonal blocks
    double *T;           // leaf scratch (NB x n or m x NB)
    void *gemm_work;     // split-K workspace for the trailing GEMMs
    size_t gemm_work_bytes;
    int cus;
    hipStream_t s;
    int err = 0;

    // op(A)[r0.., c0..] as a GEMM operand: pointer and transpose flag
    const double *opa_ptr(int64_t r0, int64_t c0) const { return trans ? A + c0 + r0 * lda : A + r0 + c0 * lda; }

    GemmConfig cfg(int mm, int nn, int kk) const {
        GemmConfig c = pick_trsm_gemm_config(mm, nn, kk, cus);
        if (gemm_workspace_bytes(mm, nn, kk, c, 8) > gemm_work_bytes) c.splitk = 1;
        return c;
    }

    void gemm(bool ta, bool tb, int mm, int nn, int kk, double alpha, const double *a, int64_t la, const double *b,
              int64_t lb, double beta, double *c, int64_t lc) {
        if (err || mm <= 0 || nn <= 0) return;
        err = launch_dgemm(ta, tb, mm, nn, kk, alpha, a, la, b, lb, beta, c, lc, cfg(mm, nn, kk), gemm_work, s);
    }

    void leaf(int p0, int len) {
        if (err) return;
        const double *inv = W + (int64_t)(p0 / TRSM_NB) * TRSM_NB * TRSM_NB;
        if (left) {   // B[p0:p0+len, :] <- inv * B[p0:p0+len, :]
            gemm(false, false, len, n, len, 1.0, inv, TRSM_NB, B + p0, ldb, 0.0, T, len);
            if (!err) {
                hipLaunchKernelGGL(copy_block_kernel, dim3(grid_for((int64_t)len * n)), dim3(256), 0, s, len, n, T,
                                   (int64_t)len, B + p0, ldb);
                err = (int)hipGetLastError();
            }
        } else {      // B[:, p0:p0+len] <- B[:, p0:p0+len] * inv
            gemm(false, false, m, len, len, 1.0, B + (int64_t)p0 * ldb, ldb, inv, TRSM_NB, 0.0, T, m);
            if (!err) {
                hipLaunchKernelGGL(copy_block_kernel, dim3(grid_for((int64_t)m * len)), dim3(256), 0, s, m, len, T,
                                   (int64_t)m, B + (int64_t)p0 * ldb, ldb);
                err = (int)hipGetLastError();
            }
        }
    }

    // Solve the diagonal range [p0, p0+len) of op(A).
    void rec(int p0, int len) {
        if (err || len <= 0) return;
        if (len <= TRSM_NB) {
            leaf(p0, len);
            return;
        }
        int h = ((len / 2 + TRSM_NB - 1) / TRSM_NB) * TRSM_NB;   // split on a block boundary
        if (h >= len) h = len - TRSM_NB;
        const int p1 = p0 + h, len2 = len - h;
        if (left) {
            if (oplower) {           // forward: X1, then B2 -= L21 X1, then X2
                rec(p0, h);
                gemm(trans, false, len2, n, h, -1.0, opa_ptr(p1, p0), lda, B + p0, ldb, 1.0, B + p1, ldb);
                rec(p1, len2);
            } else {                 // backward: X2, then B1 -= U12 X2, then X1
                rec(p1, len2);
                gemm(trans, false, h, n, len2, -1.0, opa_ptr(p0, p1), lda, B + p1, ldb, 1.0, B + p0, ldb);
                rec(p0, h);
            }
        } else {
            if (!oplower) {          // X U = B: X1 first, then B2 -= X1 U12
                rec(p0, h);
                gemm(false, trans, m, len2, h, -1.0, B + (int64_t)p0 * ldb, ldb, opa_ptr(p0, p1), lda, 1.0,
                     B + (int64_t)p1 * ldb, ldb);
                rec(p1, len2);
            } else {                 // X L = B: X2 first, then B1 -= X2 L21
                rec(p1, len2);
                gemm(false, trans, m, h, len2, -1.0, B + (int64_t)p1 * ldb, ldb, opa_ptr(p1, p0), lda, 1.0,
                     B + (int64_t)p0 * ldb, ldb);
                rec(p0, h);
            }
        }
    }
};

}  // namespace

size_t trsm_workspace_inv_bytes(int na) {
    const int64_t nblk = (na + TRSM_NB - 1) / TRSM_NB;
    return (size_t)nblk * TRSM_NB * TRSM_NB * sizeof(double);
}

size_t trsm_workspace_tmp_bytes(bool left, int m, int n) {
    return (size_t)TRSM_NB * (size_t)(left ? n : m) * sizeof(double);
}

GemmConfig pick_trsm_gemm_config(int m, int n, int k, int cus) {
    // the same occupancy rule as plain GEMM, but never a 256-row tile: the
    // trailing updates are short-and-wide panels
    GemmConfig c;
    const long t128 = (long)((m + 127) / 128) * ((n + 127) / 128);
    const long t64 = (long)((m + 63) / 64) * ((n + 63) / 64);
    if (t128 >= cus) {
        c.tile = TILE_128x128;
    } else {
        c.tile = TILE_64x64;
        if (t64 < cus && k >= 512) {
            int want = (int)((2L * cus + t64 - 1) / t64);
            int maxk = k / 256;
            if (want > maxk) want = maxk;
            if (want > 8) want = 8;
            c.splitk = want > 1 ? want : 1;
        }
    }
    return c;
}

int launch_dtrsm(bool left, bool upper, bool trans, bool unit, int m, int n, double alpha, const double *A,
                 int64_t lda, double *B, int64_t ldb, double *inv_ws, double *tmp_ws, void *gemm_ws,
                 size_t gemm_ws_bytes, int cus, hipStream_t s) {
    if (m <= 0 || n <= 0) return 0;
    if (alpha == 0.0) {
        hipLaunchKernelGGL(trsm_scale_kernel, dim3(grid_for((int64_t)m * n)), dim3(256), 0, s, m, n, 0.0, B, ldb);
        return (int)hipGetLastError();
    }
    const int na = left ? m : n;
    const bool oplower = (!upper) != trans;
    const int nblk = (na + TRSM_NB - 1) / TRSM_NB;
    hipLaunchKernelGGL(trsm_diag_inverse_kernel, dim3(nblk), dim3(256), 0, s, A, lda, na, trans, oplower, unit,
                       inv_ws);
    if (hipError_t e = hipGetLastError()) return (int)e;
    if (alpha != 1.0) {
        hipLaunchKernelGGL(trsm_scale_kernel, dim3(grid_for((int64_t)m * n)), dim3(256), 0, s, m, n, alpha, B, ldb);
        if (hipError_t e = hipGetLastError()) return (int)e;
    }
    TrsmCtx ctx;
    ctx.left = left;
    ctx.trans = trans;
    ctx.oplower = oplower;
    ctx.m = m;
    ctx.n = n;
    ctx.A = A;
    ctx.lda = lda;
    ctx.B = B;
    ctx.ldb = ldb;
    ctx.W = inv_ws;
    ctx.T = tmp_ws;
    ctx.gemm_work = gemm_ws;
    ctx.gemm_work_bytes = gemm_ws ? gemm_ws_bytes : 0;
    ctx.cus = cus;
    ctx.s = s;
    ctx.rec(0, na);
    return ctx.err;
}

// ---------------------------------------------------------------- kernel-level solves
// Contract of dtrsm_solve_{ln,lt,rn,rt}_m{M}n{N} (dtrsm_kernel_16x4_skylakex.S:
// 1882-2536), as its tester's reference functions define it
// (dtrsm_kernel_16x4_tester.c:66-199): the pending update is added first, the
// diagonal of the packed triangle holds inverse pivots (multiply), and the
// solution is written to C and back into the packed panel.  L variants: one
// thread per column of C (columns are independent); R variants: one thread
// per row of C.
__global__ __launch_bounds__(256) void trsm_solve_kernel_l(bool forward, long M, long N, const double *__restrict__ sa,
                                                           double *__restrict__ sb, double *__restrict__ C, long LDB,
                                                           long LDC) {
    for (long col = blockIdx.x * 256l + threadIdx.x; col < N; col += (long)gridDim.x * 256) {
        const long panel = col >> 2, full = N - (N & 3);
        const long w = (col < full) ? 4 : (N & 3);
        const long cc = col & 3;
        double *sbp = sb + panel * 4 * LDB;
        double *cp = C + col * LDC;
        for (long m = 0; m < M; ++m) cp[m] += sbp[m * w + cc];
        if (forward) {
            for (long mk = 0; mk < M; ++mk) {
                const double b0 = cp[mk] * sa[mk * M + mk];
                cp[mk] = b0;
                sbp[mk * w + cc] = b0;
                for (long ml = mk + 1; ml < M; ++ml) cp[ml] -= b0 * sa[mk * M + ml];
            }
        } else {
            for (long mk = M - 1; mk >= 0; --mk) {
                const double b0 = cp[mk] * sa[mk * M + mk];
                cp[mk] = b0;
                sbp[mk * w + cc] = b0;
                for (long mu = 0; mu < mk; ++mu) cp[mu] -= b0 * sa[mk * M + mu];
            }
        }
    }
}

__device__ __forceinline__ long bref_off(long ldb, long unroll, long k, long n) {
    return (n >> 2) * (ldb << 2) + k * unroll + (n & 3);
}

__global__ __launch_bounds__(256) void trsm_solve_kernel_r(bool forward, long M, long N, double *__restrict__ sa,
                                                           const double *__restrict__ sb, double *__restrict__ C,
                                                           long LDB, long LDC) {
    const long unroll = N > 4 ? 4 : N;
    for (long m = blockIdx.x * 256l + threadIdx.x; m < M; m += (long)gridDim.x * 256) {
        for (long c = 0; c < N; ++c) C[m + c * LDC] += sa[c * M + m];
        if (forward) {
            for (long nk = 0; nk < N; ++nk) {
                const double v = C[m + nk * LDC] * sb[bref_off(LDB, unroll, nk, nk)];
                C[m + nk * LDC] = v;
                sa[nk * M + m] = v;
                for (long nr = nk + 1; nr < N; ++nr) C[m + nr * LDC] -= v * sb[bref_off(LDB, unroll, nk, nr)];
            }
        } else {
            for (long nk = N - 1; nk >= 0; --nk) {
                const double v = C[m + nk * LDC] * sb[bref_off(LDB, unroll, nk, nk)];
                C[m + nk * LDC] = v;
                sa[nk * M + m] = v;
                for (long nl = 0; nl < nk; ++nl) C[m + nl * LDC] -= v * sb[bref_off(LDB, unroll, nk, nl)];
            }
        }
    }
}

int launch_dtrsm_kernel(int variant, long M, long N, double *sa, double *sb, double *C, long LDB, long LDC,
                        hipStream_t s) {
    if (M <= 0 || N <= 0) return 0;
    if (variant <= 1) {
        const int g = (int)((N + 255) / 256);
        hipLaunchKernelGGL(trsm_solve_kernel_l, dim3(g), dim3(256), 0, s, variant == 1, M, N, sa, sb, C, LDB, LDC);
    } else {
        const int g = (int)((M + 255) / 256);
        hipLaunchKernelGGL(trsm_solve_kernel_r, dim3(g), dim3(256), 0, s, variant == 2, M, N, sa, sb, C, LDB, LDC);
    }
    return (int)hipGetLastError();
}

}  // namespace gemm_amd
