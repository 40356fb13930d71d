// DTRSM on gfx950: blocked, recursive triangular solve whose bulk runs in the
// FP64 matrix-core GEMM kernel, plus the kernel-level solve contract of the
// reference's dtrsm_solve_{ln,lt,rn,rt}_m*n* family.
//
// BLAS dtrsm: side L solves op(A) X = alpha B, side R solves X op(A) = alpha B,
// X overwrites B.  Only the uplo triangle of A is read; diag 'U' never reads
// the diagonal.  Plan (replaces dtrsm_16x4's register solves,
// OpenBLAS-like_implementation/dtrsm_16x4/dtrsm_kernel_16x4_skylakex.S):
//   1. one kernel inverts every NB x NB diagonal block in parallel
//      (column-parallel substitution against the identity; the reference
//      likewise works from inverted pivots, :1480-1484);
//   2. B <- alpha B;
//   3. recursive halving over the triangle: solve the leading half, update the
//      trailing half with one large GEMM (the FP64 matrix-core kernel), solve
//      the trailing half.  A leaf of up to NL rows is one launch of the panel
//      kernel: each workgroup keeps a slab of 4 (or 8) right-hand sides in
//      FP64 matrix-core accumulators and walks the leaf's NB-row sub-blocks,
//      X_q = inv(T_qq) B_q followed by the update of the leaf's remaining
//      rows, in place.
// Side R is solved in the same "left" form: X op(A) = B  <=>  op(A)^T X^T = B^T,
// so the panel kernel and the diagonal inverses see T = op(A)^T and B^T
// through strides; the trailing updates stay GEMMs on B as stored.
#include "common.hpp"
#include "gemm_kernels.hpp"
#include "trsm.hpp"
#include "tuner.hpp"

namespace gemm_amd {

static constexpr int TRSM_NB = 64;    // inverted diagonal block
static constexpr int TRSM_NL = 256;   // rows solved by one panel launch

// Element (r, c) of op(A) from column-major storage.
__device__ __forceinline__ double opa(const double *A, int64_t lda, bool trans, int64_t r, int64_t c) {
    return trans ? A[c + r * lda] : A[r + c * lda];
}

// grid.x = diagonal block index, 256 threads per block.  W[b] (NB x NB, ld
// NB) <- inv(op(A)_bb); rows/cols past the matrix edge are the identity.
// Block 0 also writes the NB zeros that follow the last block.
//
// The block is inverted by doubling, with all four waves busy: the inverses
// of its 16 diagonal 4x4 blocks by substitution, then for h = 4, 8, 16, 32
// every pair of neighbouring h x h inverses becomes one 2h x 2h inverse,
//     inv([[A, 0], [B, C]]) = [[inv(A), 0], [-inv(C) B inv(A), inv(C)]],
// through Y = B inv(A) and -inv(C) Y.  The 16- and 32-row products run on
// v_mfma_f64_16x16x4 (one 16x16 output block per wave), the smaller ones as
// FMA sums, one output per thread.  An upper triangle is inverted as the
// lower triangle of its index-reversed block (row r, column c at 63 - r,
// 63 - c), whose inverse is the reversed inverse.  (Substituting one column
// of the inverse per lane kept one wave issuing 2048 FMA instructions.)
__global__ __launch_bounds__(256) void trsm_diag_inverse_kernel(const double *__restrict__ A, int64_t lda, int na,
                                                                bool trans, bool oplower, bool unit,
                                                                double *__restrict__ W) {
    constexpr int NB = TRSM_NB, NT = 256, P = NB + 2;
    __shared__ __align__(16) double T[NB][P];    // the (reversed, if upper) block: strict lower part
    __shared__ __align__(16) double X[NB][P];    // its inverse (zero above the diagonal)
    __shared__ __align__(16) double Y[NB][P];    // products B inv(A) of the current step
    __shared__ double piv[NB];
    const int b = blockIdx.x;
    const int64_t base = (int64_t)b * NB;
    const int nb = min(NB, na - (int)base);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // position of op(A)'s (r, c) in the lower-triangular working block
    auto wr = [&](int r) { return oplower ? r : NB - 1 - r; };
    {
        // element e = tid + NT u: its contiguous index (row for op(A) = A, column
        // for op(A) = A^T) is the lane, so every load is coalesced
        double v[NB * NB / NT];
#pragma unroll
        for (int u = 0; u < NB * NB / NT; ++u) {
            const int o = (tid + NT * u) >> 6;
            const int r = trans ? o : lane, col = trans ? lane : o;
            const bool in_tri = oplower ? (r > col) : (r < col);
            v[u] = (r < nb && col < nb && in_tri) ? opa(A, lda, trans, base + r, base + col) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NB * NB / NT; ++u) {
            const int o = (tid + NT * u) >> 6;
            const int r = trans ? o : lane, col = trans ? lane : o;
            T[wr(r)][wr(col)] = v[u];
            X[r][col] = 0.0;
        }
    }
    if (tid < NB) {
        double d = 1.0;
        if (tid < nb && !unit) d = opa(A, lda, trans, base + tid, base + tid);
        piv[wr(tid)] = 1.0 / d;
    }
    __syncthreads();
    // 4x4 diagonal blocks: thread (q, cc) substitutes column cc of block q
    if (tid < NB) {
        const int q = tid >> 2, cc = tid & 3, o = 4 * q;
        double x[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double s = (i == cc) ? 1.0 : 0.0;
#pragma unroll
            for (int j = 0; j < i; ++j) s = fma(-T[o + i][o + j], x[j], s);
            x[i] = s * piv[o + i];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) X[o + i][o + cc] = x[i];
    }
    __syncthreads();
    // h = 4, 8: pairs (top = 2h p) of h x h blocks; one output per thread
#pragma unroll
    for (int h = 4; h <= 8; h *= 2) {
        const int outs = (NB / (2 * h)) * h * h;   // 128, 256
        const int p = tid / (h * h), i = (tid / h) % h, j = tid % h, top = 2 * h * p;
        if (tid < outs) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < h; ++k) s = fma(T[top + h + i][top + k], X[top + k][top + j], s);
            Y[top + h + i][top + j] = s;
        }
        __syncthreads();
        if (tid < outs) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < h; ++k) s = fma(X[top + h + i][top + h + k], Y[top + h + k][top + j], s);
            X[top + h + i][top + j] = -s;
        }
        __syncthreads();
    }
    // h = 16, 32 on the matrix cores: D[i][j] = sum_k P[pr + i][pk + k] Q[qk + k][qc + j]
    // (A[i][k] at lane i + 16k, B[k][j] at lane j + 16k, D[i][j] at lane
    // j + 16 (i % 4), element i / 4)
    auto block16 = [&](const double (*Pm)[P], int pr, int pk, const double (*Qm)[P], int qk, int qc, int K) {
        d4_t acc = {0.0, 0.0, 0.0, 0.0};
        const int i = lane & 15, kk = lane >> 4;
        for (int k0 = 0; k0 < K; k0 += 4)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Pm[pr + i][pk + k0 + kk], Qm[qk + k0 + kk][qc + i], acc, 0, 0,
                                                        0);
        return acc;
    };
    auto put16 = [&](double (*D)[P], int dr, int dc, d4_t acc, bool neg) {
        const int j = lane & 15, i0 = lane >> 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) D[dr + 4 * e + i0][dc + j] = neg ? -acc[e] : acc[e];
    };
    {   // h = 16: pairs at top 0 and 32, waves 0 and 1
        const int top = 32 * wave;
        if (wave < 2) put16(Y, top + 16, top, block16(T, top + 16, top, X, top, top, 16), false);
        __syncthreads();
        if (wave < 2) put16(X, top + 16, top, block16(X, top + 16, top + 16, Y, top + 16, top, 16), true);
        __syncthreads();
    }
    {   // h = 32: one pair, four 16x16 output blocks, one per wave
        const int bi = 16 * (wave >> 1), bj = 16 * (wave & 1);
        put16(Y, 32 + bi, bj, block16(T, 32 + bi, 0, X, 0, bj, 32), false);
        __syncthreads();
        put16(X, 32 + bi, bj, block16(X, 32 + bi, 32, Y, 32, bj, 32), true);
        __syncthreads();
    }
    double *out = W + (int64_t)b * NB * NB;
#pragma unroll
    for (int u = 0; u < NB * NB / NT; ++u) {
        const int e = tid + NT * u, r = e & 63, c = e >> 6;
        out[e] = X[wr(r)][wr(c)];
    }
    if (b == 0 && tid < NB) W[(int64_t)gridDim.x * NB * NB + tid] = 0.0;   // the zero operand of the panel kernel
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

// Solve T Y = C on rows [p0, p0+len) of C (len <= NL, p0 a multiple of NB),
// in place, where T(r, c) = A[r*ars + c*acs] is lower (forward) or upper
// (backward) and C(i, j) = B[i*brs + j*bcs] has ncols columns.  Rows outside
// the leaf are final and already folded into C.  W holds inv(T_bb) for every
// NB block (column-major NB x NB, identity past the edge), followed by NB
// zeros (`zero`).  grid.x = column slabs of 4*NCG; one of ars/acs and one of
// brs/bcs is 1 (the coalesced side).
//
// 16 waves; wave w keeps rows [16w, 16w+16) of the slab, NCG groups of 4
// columns, in v_mfma_f64_4x4x4_4b accumulators: the instruction's 4 blocks
// are 4 row quads of the tile sharing one B operand, so that A[r][k] sits at
// lane r+16k (r = row in the tile), B[k][j] at lane 4b+j+16k for every b, and
// D[r][j] at lane 16(r%4)+4(r/4)+j.  Sub-block q (waves 4q..4q+3): its waves
// publish C_q to LDS, compute X_q = inv(T_qq) C_q (the inverse rows straight
// from registers) and publish -X_q; the waves of the rows still to solve then
// accumulate C_w += T_wq (-X_q) (T_wq straight from registers).  Every wave's
// operands for the next sub-block are fetched while the current one computes.
// Columns past ncols are zero and stay zero.
template <int NCG>
__global__ __launch_bounds__(1024) void trsm_panel_kernel(const double *__restrict__ A, int64_t ars, int64_t acs,
                                                          const double *__restrict__ W,
                                                          const double *__restrict__ zero, double *__restrict__ B,
                                                          int64_t brs, int64_t bcs, int p0, int len, int ncols,
                                                          bool lower) {
    constexpr int NT = 1024, NB = TRSM_NB, NL = TRSM_NL, WC = 4 * NCG, SP = WC + 1;
    static_assert(NB == 64 && NL == 256, "one 16-row tile per wave, 4 waves per NB block");
    __shared__ double S[NL * SP];    // slab staging for coalesced global access
    __shared__ double Cq[NB * SP];   // C_q, published by the sub-block's waves
    __shared__ double Xq[NB * SP];   // -X_q, read by the waves still to solve
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4;
    const int ar = lane & 15;                            // A operand: tile row
    const int bj = lane & 3;                             // B operand: column in the group
    const int dr = 4 * ((lane >> 2) & 3) + (lane >> 4);  // D: tile row (column bj)
    const int j0 = blockIdx.x * WC;
    const int nq = (len + NB - 1) / NB;
    const bool brow = (brs == 1), arow = (ars == 1);

    // 1: this wave's rows are sub-block q; 2: they are still to solve; 0: idle
    auto role = [&](int q) {
        const int blk = w >> 2;
        if (16 * w >= len) return 0;
        if (blk == q) return 1;
        return (lower ? blk > q : blk < q) ? 2 : 0;
    };
    // k of the s-th instruction for the lanes of group g: four consecutive k
    // per instruction where T's rows are contiguous, sixteen consecutive k per
    // lane where its columns are
    auto kof = [&](bool rows_contig, int s) { return rows_contig ? 4 * s + g : 16 * g + s; };
    // One branch-free stream of 16 loads per wave and sub-block, so that the
    // loads for the next sub-block stay in flight behind the current one's
    // work.  Entries past the matrix edge (and idle waves) read a zero.
    auto fetch = [&](int q, double *dst) {
        const int rl = role(q);
        const int row = 16 * w + ar, kc = NB * q;
        const double *p = zero;
        int64_t st = 0;
        int kmax = NB, kmul = 0, kadd = 0;   // entry s valid iff s * kmul + kadd < kmax
        if (rl == 1) {   // inverse rows 16(w&3) + ar of block q (column-major NB x NB)
            p = W + ((int64_t)(p0 / NB) + q) * (NB * NB) + 16 * (w & 3) + ar + g * NB;
            st = 4 * NB;
        } else if (rl == 2) {   // T(row, NB q + k)
            p = A + (int64_t)(p0 + row) * ars + (int64_t)(p0 + kc) * acs;
            if (arow) {
                p += g * acs;
                st = 4 * acs;
                kmul = 4;
                kadd = g;
            } else {
                p += 16 * g;
                st = 1;
                kmul = 1;
                kadd = 16 * g;
            }
            kmax = row < len ? len - kc : 0;
        }
#pragma unroll
        for (int s = 0; s < 16; ++s) dst[s] = *(s * kmul + kadd < kmax ? p + s * st : zero);
    };

    double pc[16], pn[16];
    int q = lower ? 0 : nq - 1;
    fetch(q, pn);   // in flight together with the slab
    {
        constexpr int SL = (NL * WC + NT - 1) / NT;
        double sv[SL];
#pragma unroll
        for (int i = 0; i < SL; ++i) {
            const int e = i * NT + tid;
            const int r = brow ? e % NL : e / WC, c = brow ? e / NL : e % WC;
            sv[i] = *(r < len && j0 + c < ncols ? B + (int64_t)(p0 + r) * brs + (int64_t)(j0 + c) * bcs : zero);
        }
#pragma unroll
        for (int i = 0; i < SL; ++i) {
            const int e = i * NT + tid;
            const int r = brow ? e % NL : e / WC, c = brow ? e / NL : e % WC;
            S[r * SP + c] = sv[i];
        }
    }
    __syncthreads();
    double acc[NCG];
#pragma unroll
    for (int cg = 0; cg < NCG; ++cg) acc[cg] = S[(16 * w + dr) * SP + 4 * cg + bj];

    for (int step = 0; step < nq; ++step) {
        const int rl = role(q);
#pragma unroll
        for (int s = 0; s < 16; ++s) pc[s] = pn[s];   // this sub-block's operands (landed by now)
        const int qn = lower ? q + 1 : q - 1;
        if (step + 1 < nq) fetch(qn, pn);
        if (rl == 1) {
#pragma unroll
            for (int cg = 0; cg < NCG; ++cg) Cq[(16 * (w & 3) + dr) * SP + 4 * cg + bj] = acc[cg];
        }
        __syncthreads();
        if (rl == 1) {
#pragma unroll
            for (int cg = 0; cg < NCG; ++cg) {
                double x = 0.0;
#pragma unroll
                for (int s = 0; s < 16; ++s) x = mfma_f64_4x4x4(pc[s], Cq[kof(true, s) * SP + 4 * cg + bj], x);
                acc[cg] = x;
                Xq[(16 * (w & 3) + dr) * SP + 4 * cg + bj] = -x;   // published negated
            }
        }
        __syncthreads();
        if (rl == 2) {
#pragma unroll
            for (int cg = 0; cg < NCG; ++cg) {
#pragma unroll
                for (int s = 0; s < 16; ++s)
                    acc[cg] = mfma_f64_4x4x4(pc[s], Xq[kof(arow, s) * SP + 4 * cg + bj], acc[cg]);
            }
        }
        q = qn;
    }

#pragma unroll
    for (int cg = 0; cg < NCG; ++cg) S[(16 * w + dr) * SP + 4 * cg + bj] = acc[cg];
    __syncthreads();
    constexpr int SL = (NL * WC + NT - 1) / NT;
#pragma unroll
    for (int i = 0; i < SL; ++i) {
        const int e = i * NT + tid;
        const int r = brow ? e % NL : e / WC, c = brow ? e / NL : e % WC;
        if (e < NL * WC && r < len && j0 + c < ncols) B[(int64_t)(p0 + r) * brs + (int64_t)(j0 + c) * bcs] = S[r * SP + c];
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
    const double *W;     // inverted diagonal blocks of the left-form triangle
    void *gemm_work;     // split-K workspace for the trailing GEMMs
    size_t gemm_work_bytes;
    int cus;
    hipStream_t s;
    int err = 0;

    // op(A)[r0.., c0..] as a GEMM operand: pointer and transpose flag
    const double *opa_ptr(int64_t r0, int64_t c0) const { return trans ? A + c0 + r0 * lda : A + r0 + c0 * lda; }

    // The trailing updates take the configuration every DGEMM call gets (the
    // learned picker, or its fallback), limited to the split-K workspace.
    GemmConfig cfg(bool ta, bool tb, int mm, int nn, int kk, int64_t la, int64_t lb) const {
        GemmConfig c = pick_gemm_config(DT_F64, ta, tb, mm, nn, kk, (int)(la - (ta ? kk : mm)),
                                        (int)(lb - (tb ? nn : kk)), cus);
        if (gemm_workspace_bytes(mm, nn, kk, c, 8) > gemm_work_bytes) c.splitk = 1;
        return c;
    }

    void gemm(bool ta, bool tb, int mm, int nn, int kk, double alpha, const double *a, int64_t la, const double *b,
              int64_t lb, double beta, double *c, int64_t lc) {
        if (err || mm <= 0 || nn <= 0) return;
        err = launch_dgemm(ta, tb, mm, nn, kk, alpha, a, la, b, lb, beta, c, lc, cfg(ta, tb, mm, nn, kk, la, lb), gemm_work,
                           s);
    }

    // Leaf: rows [p0, p0+len) of the left-form system T Y = C.  Side L: T =
    // op(A), C = B.  Side R: T = op(A)^T, C = B^T (columns of B).
    void leaf(int p0, int len) {
        if (err) return;
        const bool tt = left ? trans : !trans;   // T(r, c) = tt ? A[c + r lda] : A[r + c lda]
        const int64_t ars = tt ? lda : 1, acs = tt ? 1 : lda;
        const int64_t brs = left ? 1 : ldb, bcs = left ? ldb : 1;
        const int ncols = left ? n : m;
        const bool lower = left ? oplower : !oplower;
        // 4-column slabs while they fit the grid in one round of workgroups
        const int ncg = (ncols + 3) / 4 <= cus ? 1 : 2;
        const int grid = (ncols + 4 * ncg - 1) / (4 * ncg);
        const double *zero = W + (int64_t)((left ? m : n) + TRSM_NB - 1) / TRSM_NB * (TRSM_NB * TRSM_NB);
        auto kern = ncg == 1 ? trsm_panel_kernel<1> : trsm_panel_kernel<2>;
        hipLaunchKernelGGL(kern, dim3(grid), dim3(1024), 0, s, A, ars, acs, W, zero, B, brs, bcs, p0, len, ncols,
                           lower);
        err = (int)hipGetLastError();
    }

    // Solve the diagonal range [p0, p0+len) of op(A).
    void rec(int p0, int len) {
        if (err || len <= 0) return;
        if (len <= TRSM_NL) {
            leaf(p0, len);
            return;
        }
        int h = ((len / 2 + TRSM_NL - 1) / TRSM_NL) * TRSM_NL;   // split on a leaf boundary
        if (h >= len) h = len - TRSM_NL;
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
    return ((size_t)nblk * TRSM_NB * TRSM_NB + TRSM_NB) * sizeof(double);
}

int launch_dtrsm(bool left, bool upper, bool trans, bool unit, int m, int n, double alpha, const double *A,
                 int64_t lda, double *B, int64_t ldb, double *inv_ws, void *gemm_ws, size_t gemm_ws_bytes, int cus,
                 hipStream_t s) {
    if (m <= 0 || n <= 0) return 0;
    if (alpha == 0.0) {
        hipLaunchKernelGGL(trsm_scale_kernel, dim3(grid_for((int64_t)m * n)), dim3(256), 0, s, m, n, 0.0, B, ldb);
        return (int)hipGetLastError();
    }
    const int na = left ? m : n;
    const bool oplower = (!upper) != trans;
    // Inverses of the left-form triangle's diagonal blocks: op(A) for side L,
    // op(A)^T (the other transpose, the other triangle) for side R.
    const bool inv_trans = left ? trans : !trans, inv_lower = left ? oplower : !oplower;
    const int nblk = (na + TRSM_NB - 1) / TRSM_NB;
    hipLaunchKernelGGL(trsm_diag_inverse_kernel, dim3(nblk), dim3(256), 0, s, A, lda, na, inv_trans, inv_lower, unit,
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
