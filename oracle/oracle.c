/*
 * CPU oracle for the BLAS-3 hot path.  TEST INFRASTRUCTURE ONLY: the shipped
 * library never links or calls this file; tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it as the checker.  Declarations and the
 * meaning of every argument are in oracle.h.
 */
#include "oracle.h"
#include <math.h>
#include <float.h>
#include <string.h>
#include <stdlib.h>
#include <quadmath.h>

/* ------------------------------------------------------------------ RNG */
/* SplitMix64 (Steele, Lea, Flood 2014; Vigna's public-domain constants). */
uint64_t oracle_splitmix64_next(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* FP64 uniform[-1,1): ((x>>11) * 2^-53) * 2 - 1, exact in binary64 */
static double u01_to_f64(uint64_t x) { return ldexp((double)(x >> 11), -53) * 2.0 - 1.0; }
/* FP32 uniform[-1,1): ((x>>40) * 2^-24) * 2 - 1, exact in binary32 */
static float u01_to_f32(uint64_t x) { return ldexpf((float)(x >> 40), -24) * 2.0f - 1.0f; }

void oracle_fill_f64(uint64_t *state, int64_t rows, int64_t cols, int64_t ld, double *x) {
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i)
            x[i + j * ld] = u01_to_f64(oracle_splitmix64_next(state));
}

void oracle_fill_f32(uint64_t *state, int64_t rows, int64_t cols, int64_t ld, float *x) {
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i)
            x[i + j * ld] = u01_to_f32(oracle_splitmix64_next(state));
}

/* u * 2^e: u from one draw, e = -r + (draw mod (2r+1)) from the next draw
 * (two draws per element for every r, so r = 0 gives e = 0) */
void oracle_fill_f64_range(uint64_t *state, int64_t rows, int64_t cols, int64_t ld, int r, double *x) {
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) {
            double u = u01_to_f64(oracle_splitmix64_next(state));
            uint64_t d = oracle_splitmix64_next(state);
            int e = r > 0 ? -r + (int)(d % (uint64_t)(2 * r + 1)) : 0;
            x[i + j * ld] = ldexp(u, e);
        }
}

/* ------------------------------------------------------------------ GEMM */
/* op(X) = X for 'N'/'n', else X^T  (src/gemm_driver.c:18-19) */
static int notrans(char t) { return t == 'N' || t == 'n'; }

/* C <- alpha*op(A)*op(B) + beta*C, each dot product summed in long double and
 * rounded once.  Contract of src/gemm_driver.c:116-195: beta != 1 scales C
 * (:131), alpha == 0 skips the product (:132), K == 0 leaves beta*C.  One
 * documented divergence: beta == 0 overwrites C without reading it (netlib
 * BLAS), where the reference multiplies by zero (:110). */
void oracle_dgemm(char ta, char tb, int m, int n, int k, double alpha,
                  const double *a, int lda, const double *b, int ldb,
                  double beta, double *c, int ldc) {
    const int na = notrans(ta), nb = notrans(tb);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            long double s = 0.0L;
            if (alpha != 0.0)
                for (int p = 0; p < k; ++p) {
                    long double av = na ? a[i + (int64_t)p * lda] : a[p + (int64_t)i * lda];
                    long double bv = nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb];
                    s += av * bv;
                }
            double *cij = &c[i + (int64_t)j * ldc];
            long double r = (long double)alpha * s;
            if (beta != 0.0) r += (long double)beta * (long double)(*cij);
            *cij = (double)r;
        }
}

void oracle_sgemm(char ta, char tb, int m, int n, int k, float alpha,
                  const float *a, int lda, const float *b, int ldb,
                  float beta, float *c, int ldc) {
    const int na = notrans(ta), nb = notrans(tb);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            long double s = 0.0L;
            if (alpha != 0.0f)
                for (int p = 0; p < k; ++p) {
                    long double av = na ? a[i + (int64_t)p * lda] : a[p + (int64_t)i * lda];
                    long double bv = nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb];
                    s += av * bv;
                }
            float *cij = &c[i + (int64_t)j * ldc];
            long double r = (long double)alpha * s;
            if (beta != 0.0f) r += (long double)beta * (long double)(*cij);
            *cij = (float)r;
        }
}

/* __float128 accumulation (113-bit significand; products of two binary64
 * values are exact in it), result split into a double-double. */
static void q_entry(int na, int nb, int k, double alpha, const double *a, int lda,
                    const double *b, int ldb, double beta, const double *c, int ldc,
                    int i, int j, double *hi, double *lo) {
    __float128 s = 0;
    if (alpha != 0.0)
        for (int p = 0; p < k; ++p) {
            __float128 av = na ? a[i + (int64_t)p * lda] : a[p + (int64_t)i * lda];
            __float128 bv = nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb];
            s += av * bv;
        }
    __float128 r = (__float128)alpha * s;
    if (beta != 0.0) r += (__float128)beta * (__float128)c[i + (int64_t)j * ldc];
    double h = (double)r;
    *hi = h;
    *lo = (double)(r - (__float128)h);
}

void oracle_dgemm_q(char ta, char tb, int m, int n, int k, double alpha,
                    const double *a, int lda, const double *b, int ldb,
                    double beta, const double *c, int ldc,
                    int64_t nsample, const int *rows, const int *cols,
                    double *out_hi, double *out_lo) {
    const int na = notrans(ta), nb = notrans(tb);
    if (rows) {
        #pragma omp parallel for schedule(dynamic, 16)
        for (int64_t s = 0; s < nsample; ++s)
            q_entry(na, nb, k, alpha, a, lda, b, ldb, beta, c, ldc, rows[s], cols[s],
                    &out_hi[s], &out_lo[s]);
    } else {
        #pragma omp parallel for collapse(2) schedule(dynamic, 16)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i)
                q_entry(na, nb, k, alpha, a, lda, b, ldb, beta, c, ldc, i, j,
                        &out_hi[i + (int64_t)j * m], &out_lo[i + (int64_t)j * m]);
    }
}

/* ------------------------------------------------------------------ DTRSM */
/* BLAS dtrsm: side 'L' solves op(A) X = alpha B, side 'R' solves X op(A) = alpha B;
 * X overwrites B.  The triangle of A named by uplo is the only part read; the
 * other triangle is never touched (NaN-poison safe); diag 'U' takes A_ii = 1
 * without reading it.  Substitution order follows the tester's reference
 * solvers: forward for the lower/no-trans left case (solve_lt_ref,
 * dtrsm_kernel_16x4_tester.c:160-173), backward for upper
 * (solve_ln_ref, :145-158), and column sweeps for the right side
 * (solve_rn_reference :97-119, solve_rt_reference :121-143). */
static int lsame(char x, char y) { return x == y || x == y + 32; }

void oracle_dtrsm(char side, char uplo, char transa, char diag, int m, int n,
                  double alpha, const double *a, int lda, double *b, int ldb) {
    const int left = lsame(side, 'L'), lower = lsame(uplo, 'L');
    const int nt = lsame(transa, 'N'), unit = lsame(diag, 'U');
    const int na = left ? m : n;
    /* op(A)[r][c] */
#define OPA(r, c) (nt ? (long double)a[(r) + (int64_t)(c) * lda] : (long double)a[(c) + (int64_t)(r) * lda])
    /* op(A) is lower triangular when (lower && N) or (upper && T) */
    const int oplower = (lower && nt) || (!lower && !nt);
    long double *x = (long double *)malloc(sizeof(long double) * (size_t)(na > 0 ? na : 1));
    if (left) {
        for (int j = 0; j < n; ++j) {
            for (int i = 0; i < m; ++i) x[i] = (long double)alpha * b[i + (int64_t)j * ldb];
            if (oplower) {
                for (int i = 0; i < m; ++i) {           /* forward */
                    long double s = x[i];
                    for (int p = 0; p < i; ++p) s -= OPA(i, p) * x[p];
                    x[i] = unit ? s : s / OPA(i, i);
                }
            } else {
                for (int i = m - 1; i >= 0; --i) {      /* backward */
                    long double s = x[i];
                    for (int p = i + 1; p < m; ++p) s -= OPA(i, p) * x[p];
                    x[i] = unit ? s : s / OPA(i, i);
                }
            }
            for (int i = 0; i < m; ++i) b[i + (int64_t)j * ldb] = (double)x[i];
        }
    } else {
        /* X op(A) = alpha B, row by row: x (1 x n) op(A) = alpha b_row */
        for (int i = 0; i < m; ++i) {
            for (int j = 0; j < n; ++j) x[j] = (long double)alpha * b[i + (int64_t)j * ldb];
            if (oplower) {
                /* x_j = (y_j - sum_{p>j} x_p A[p][j]) / A[j][j], j descending */
                for (int j = n - 1; j >= 0; --j) {
                    long double s = x[j];
                    for (int p = j + 1; p < n; ++p) s -= x[p] * OPA(p, j);
                    x[j] = unit ? s : s / OPA(j, j);
                }
            } else {
                for (int j = 0; j < n; ++j) {
                    long double s = x[j];
                    for (int p = 0; p < j; ++p) s -= x[p] * OPA(p, j);
                    x[j] = unit ? s : s / OPA(j, j);
                }
            }
            for (int j = 0; j < n; ++j) b[i + (int64_t)j * ldb] = (double)x[j];
        }
    }
#undef OPA
    free(x);
}

/* Kernel-level contract of dtrsm_solve_{ln,lt,rn,rt}_m{M}n{N}
 * (dtrsm_kernel_16x4_skylakex.S:1882-2536), restated from the tester's
 * reference functions (dtrsm_kernel_16x4_tester.c:66-199):
 *  - L side: sb holds 4-column panels (panel p at sb + p*4*LDB, element
 *    (row m, col c) at m*w + c with w = panel width); the panel is first added
 *    into C (update_c :66-77), then C is solved against sa, whose column mk
 *    holds the triangle and whose diagonal entry sa[mk*M+mk] is the INVERSE
 *    of the pivot (it multiplies, :163-165); the solution is written to C and
 *    back into sb.  ln = backward (:145-158), lt = forward (:160-173).
 *  - R side: sa (M x N column-major) is first added into C (update_c_a
 *    :79-89); sb is addressed by get_b_ref (:91-95) and its diagonal is the
 *    inverse pivot; the solution goes to C and back into sa.
 *    rn = forward over columns (:97-119), rt = backward (:121-143). */
static const double *bref(const double *base, unsigned long ldb, unsigned unroll,
                          unsigned long k, unsigned long n) {
    return base + (n >> 2) * (ldb << 2) + k * unroll + (n & 3u);
}

void oracle_dtrsm_kernel(int variant, long M, long N, double *sa, double *sb,
                         double *C, long LDB, long LDC) {
    if (variant == 0 || variant == 1) {
        long nrem = N, w = 4;
        double *sbp = sb, *cp = C;
        for (; nrem; nrem -= w) {
            if (nrem < 4) w = nrem;
            for (long c = 0; c < w; ++c)
                for (long m = 0; m < M; ++m) cp[m + c * LDC] += sbp[m * w + c];
            if (variant == 1) {
                for (long mk = 0; mk < M; ++mk) {
                    double a0 = sa[mk * M + mk];
                    for (long c = 0; c < w; ++c) {
                        double b0 = cp[mk + c * LDC] * a0;
                        sbp[mk * w + c] = cp[mk + c * LDC] = b0;
                        for (long ml = mk + 1; ml < M; ++ml) cp[ml + c * LDC] -= b0 * sa[mk * M + ml];
                    }
                }
            } else {
                for (long mk = M - 1; mk >= 0; --mk) {
                    double a0 = sa[mk * M + mk];
                    for (long c = 0; c < w; ++c) {
                        double b0 = cp[mk + c * LDC] * a0;
                        sbp[mk * w + c] = cp[mk + c * LDC] = b0;
                        for (long mu = 0; mu < mk; ++mu) cp[mu + c * LDC] -= b0 * sa[mk * M + mu];
                    }
                }
            }
            sbp += w * LDB;
            cp += w * LDC;
        }
    } else {
        const unsigned unroll = N > 4 ? 4 : (unsigned)N;
        for (long c = 0; c < N; ++c)
            for (long m = 0; m < M; ++m) C[m + c * LDC] += sa[c * M + m];
        if (variant == 2) {
            for (long nk = 0; nk < N; ++nk) {
                double b0 = *bref(sb, LDB, unroll, nk, nk);
                double *c_ = C + nk * LDC, *a_ = sa + nk * M;
                for (long m = 0; m < M; ++m) { c_[m] *= b0; a_[m] = c_[m]; }
                for (long nr = nk + 1; nr < N; ++nr) {
                    double b1 = *bref(sb, LDB, unroll, nk, nr);
                    double *cc = C + nr * LDC;
                    for (long m = 0; m < M; ++m) cc[m] -= a_[m] * b1;
                }
            }
        } else {
            for (long nk = N - 1; nk >= 0; --nk) {
                double b0 = *bref(sb, LDB, unroll, nk, nk);
                double *c_ = C + nk * LDC, *a_ = sa + nk * M;
                for (long m = 0; m < M; ++m) { c_[m] *= b0; a_[m] = c_[m]; }
                for (long nl = 0; nl < nk; ++nl) {
                    double b1 = *bref(sb, LDB, unroll, nk, nl);
                    double *cc = C + nl * LDC;
                    for (long m = 0; m < M; ++m) cc[m] -= a_[m] * b1;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ complex */
/* C <- alpha*op(A)*op(B) + beta*C with op in {N, T, C(conjugate transpose)},
 * all four real products per entry summed in long double (the 4M form; the
 * 3M kernels cgemm3m_kernel_16x4_skylakex.c:416-440 and
 * zgemm3m_kernel_2x8_skylakex.c:336-351 compute the same contract).
 * beta == 0 overwrites C. */
#define CGEMM_BODY(T)                                                                    \
    const int na = notrans(ta), nb = notrans(tb);                                       \
    const int ca = lsame(ta, 'C'), cb = lsame(tb, 'C');                                 \
    const long double ar = alpha[0], ai = alpha[1], br = beta[0], bi = beta[1];         \
    for (int j = 0; j < n; ++j)                                                         \
        for (int i = 0; i < m; ++i) {                                                   \
            long double sr = 0, si = 0;                                                 \
            if (ar != 0 || ai != 0)                                                     \
                for (int p = 0; p < k; ++p) {                                           \
                    int64_t ia = na ? (i + (int64_t)p * lda) : (p + (int64_t)i * lda);  \
                    int64_t ib = nb ? (p + (int64_t)j * ldb) : (j + (int64_t)p * ldb);  \
                    long double xr = a[2 * ia], xi = a[2 * ia + 1];                     \
                    long double yr = b[2 * ib], yi = b[2 * ib + 1];                     \
                    if (ca) xi = -xi;                                                   \
                    if (cb) yi = -yi;                                                   \
                    sr += xr * yr - xi * yi;                                            \
                    si += xr * yi + xi * yr;                                            \
                }                                                                       \
            T *cij = &c[2 * (i + (int64_t)j * ldc)];                                    \
            long double rr = ar * sr - ai * si, ri = ar * si + ai * sr;                 \
            if (br != 0 || bi != 0) {                                                   \
                long double cr = cij[0], ci = cij[1];                                   \
                rr += br * cr - bi * ci;                                                \
                ri += br * ci + bi * cr;                                                \
            }                                                                           \
            cij[0] = (T)rr;                                                             \
            cij[1] = (T)ri;                                                             \
        }

void oracle_zgemm(char ta, char tb, int m, int n, int k, const double *alpha,
                  const double *a, int lda, const double *b, int ldb,
                  const double *beta, double *c, int ldc) { CGEMM_BODY(double) }

void oracle_cgemm(char ta, char tb, int m, int n, int k, const float *alpha,
                  const float *a, int lda, const float *b, int ldb,
                  const float *beta, float *c, int ldc) { CGEMM_BODY(float) }

/* ------------------------------------------------------------------ errors */
void oracle_err_f64(int64_t n, const double *c, const double *rh, const double *rl, double *out) {
    long double maxd = 0, maxr = 0, sd = 0, sr = 0, maxel = 0;
    for (int64_t i = 0; i < n; ++i) {
        long double r = (long double)rh[i] + (rl ? (long double)rl[i] : 0.0L);
        long double d = fabsl((long double)c[i] - r);
        long double ar = fabsl(r);
        if (d > maxd) maxd = d;
        if (ar > maxr) maxr = ar;
        sd += d * d;
        sr += r * r;
        if (ar > 0 && d / ar > maxel) maxel = d / ar;
    }
    out[0] = maxr > 0 ? (double)(maxd / maxr) : (double)maxd;
    out[1] = sr > 0 ? (double)sqrtl(sd / sr) : (double)sqrtl(sd);
    out[2] = (double)maxel;
}

void oracle_err_f32(int64_t n, const float *c, const double *rh, double *out) {
    long double maxd = 0, maxr = 0, sd = 0, sr = 0, maxel = 0;
    for (int64_t i = 0; i < n; ++i) {
        long double r = rh[i];
        long double d = fabsl((long double)c[i] - r);
        long double ar = fabsl(r);
        if (d > maxd) maxd = d;
        if (ar > maxr) maxr = ar;
        sd += d * d;
        sr += r * r;
        if (ar > 0 && d / ar > maxel) maxel = d / ar;
    }
    out[0] = maxr > 0 ? (double)(maxd / maxr) : (double)maxd;
    out[1] = sr > 0 ? (double)sqrtl(sd / sr) : (double)sqrtl(sd);
    out[2] = (double)maxel;
}

/* ------------------------------------------------------------------ emulation splits */
/* Ozaki int8 slicing of one operand (no reference counterpart; the method is
 * the error-free splitting of Ozaki, Ogita, Oishi 2012 with int8 slices as in
 * Ootomo, Ozaki, Yokota 2024).  For each row r of x (row r = elements
 * x[r*row_stride + p*k_stride], p < k):
 *   e_r = the exponent with max_p |x| in [2^(e_r-1), 2^e_r)  (frexp), 0 for an all-zero row;
 *   y   = x * 2^-e_r  (|y| < 1, exact);
 *   slice t = 1..nslices:  v = y * (t == 1 ? 64 : 128) (exact); q_t = rint(v) in [-64,64]
 *                          (round to nearest even); y = v - q_t in [-1/2, 1/2] (exact).
 * slices[(t-1)*rows*k + r*k + p] = q_t, exps[r] = e_r.  So
 *   x = 2^e_r * sum_t q_t * 2^(1-7t) + 2^e_r * 2^(1-7 nslices) * y_final. */
void oracle_ozaki_split(int rows, int k, const double *x, int64_t row_stride, int64_t k_stride,
                        int nslices, int8_t *slices, int *exps) {
    for (int r = 0; r < rows; ++r) {
        double amax = 0;
        for (int p = 0; p < k; ++p) {
            double v = fabs(x[r * row_stride + p * k_stride]);
            if (v > amax) amax = v;
        }
        int e = 0;
        if (amax > 0) frexp(amax, &e);
        exps[r] = e;
        for (int p = 0; p < k; ++p) {
            double y = ldexp(x[r * row_stride + p * k_stride], -e);
            for (int t = 0; t < nslices; ++t) {
                double v = y * (t == 0 ? 64.0 : 128.0);
                double q = rint(v);
                slices[(int64_t)t * rows * k + (int64_t)r * k + p] = (int8_t)q;
                y = v - q;
            }
        }
    }
}

/* Modular (CRT) form of the emulated DGEMM (Ozaki scheme II; no reference
 * counterpart): row r -> integers rint(x * 2^(bits - e_r)) with e_r from frexp
 * of the row's max |x| (exact: power-of-two scaling, |value| <= 2^bits <= 2^56
 * fits int64), then the symmetric residue modulo each modulus by int64
 * remainder.  A residue of +-m/2 (m even) is stored as the byte of -m/2
 * (the same residue class). */
void oracle_ozaki_crt_split(int rows, int k, const double *x, int64_t row_stride, int64_t k_stride, int nmod,
                            const int *moduli, int bits, int8_t *res, int *exps) {
    for (int r = 0; r < rows; ++r) {
        double amax = 0;
        for (int p = 0; p < k; ++p) {
            double v = fabs(x[r * row_stride + p * k_stride]);
            if (v > amax) amax = v;
        }
        int e = 0;
        if (amax > 0) frexp(amax, &e);
        exps[r] = e - bits;
        for (int p = 0; p < k; ++p) {
            const int64_t v = (int64_t)rint(ldexp(x[r * row_stride + p * k_stride], bits - e));
            for (int l = 0; l < nmod; ++l) {
                const int64_t m = moduli[l];
                int64_t q = v % m;                /* in (-m, m) */
                if (q > m / 2) q -= m;
                if (q < -(m / 2)) q += m;
                res[(int64_t)l * rows * k + (int64_t)r * k + p] = (int8_t)(uint8_t)(q & 0xFF);
            }
        }
    }
}

uint16_t oracle_f32_to_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u); /* quiet NaN */
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;                                                      /* round to nearest even */
    return (uint16_t)(u >> 16);
}

float oracle_bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* x = p_0 + p_1 (+ p_2): p_0 = bf16(x), p_1 = bf16(x - p_0), p_2 = bf16(x - p_0 - p_1),
 * each difference exact in FP32. parts[s*n + i] = p_s(x_i). */
void oracle_bf16_split(int64_t n, const float *x, int nsplit, uint16_t *parts) {
    for (int64_t i = 0; i < n; ++i) {
        float r = x[i];
        for (int s = 0; s < nsplit; ++s) {
            uint16_t h = oracle_f32_to_bf16(r);
            parts[(int64_t)s * n + i] = h;
            r = r - oracle_bf16_to_f32(h);
        }
    }
}

/* ------------------------------------------------------------------ CPU port baseline */
/* Cache-blocked FP64 GEMM, OpenMP over column blocks of C, FP64 accumulation.
 * Same contract as oracle_dgemm; used only to time a CPU leg where the
 * reference build cannot run (no AVX-512F host). */
void oracle_dgemm_blocked(char ta, char tb, int m, int n, int k, double alpha,
                          const double *a, int lda, const double *b, int ldb,
                          double beta, double *c, int ldc, int nthreads) {
    const int na = notrans(ta), nb = notrans(tb);
    const int MB = 64, NB = 64, KB = 256;
    #pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int j0 = 0; j0 < n; j0 += NB) {
        double *ap = (double *)malloc(sizeof(double) * MB * KB);
        double *bp = (double *)malloc(sizeof(double) * KB * NB);
        int jn = (n - j0 < NB) ? n - j0 : NB;
        for (int j = j0; j < j0 + jn; ++j)
            for (int i = 0; i < m; ++i) {
                double *cij = &c[i + (int64_t)j * ldc];
                *cij = (beta == 0.0) ? 0.0 : beta * *cij;
            }
        if (alpha != 0.0)
            for (int p0 = 0; p0 < k; p0 += KB) {
                int pn = (k - p0 < KB) ? k - p0 : KB;
                for (int jj = 0; jj < jn; ++jj)
                    for (int pp = 0; pp < pn; ++pp) {
                        int p = p0 + pp, j = j0 + jj;
                        bp[pp + jj * KB] = alpha * (nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb]);
                    }
                for (int i0 = 0; i0 < m; i0 += MB) {
                    int in = (m - i0 < MB) ? m - i0 : MB;
                    for (int pp = 0; pp < pn; ++pp)
                        for (int ii = 0; ii < in; ++ii) {
                            int i = i0 + ii, p = p0 + pp;
                            ap[ii + pp * MB] = na ? a[i + (int64_t)p * lda] : a[p + (int64_t)i * lda];
                        }
                    for (int jj = 0; jj < jn; ++jj) {
                        double *cc = &c[i0 + (int64_t)(j0 + jj) * ldc];
                        for (int pp = 0; pp < pn; ++pp) {
                            double bv = bp[pp + jj * KB];
                            const double *ac = &ap[pp * MB];
                            for (int ii = 0; ii < in; ++ii) cc[ii] += ac[ii] * bv;
                        }
                    }
                }
            }
        free(ap);
        free(bp);
    }
}

/* ------------------------------------------------------------------ LU */
/* Unblocked right-looking LU with partial pivoting, step for step LAPACK's
 * dgetf2: idamax pivot (first maximum of |a|), full-row interchange, column
 * scaled by the reciprocal pivot when |pivot| >= DBL_MIN (division otherwise),
 * rank-1 update of the trailing block.  A zero pivot records info and skips
 * the scaling, as dgetf2 does. */
int oracle_dgetrf(int m, int n, double *a, int lda, int *ipiv) {
    int info = 0;
    const int mn = m < n ? m : n;
    for (int j = 0; j < mn; ++j) {
        int p = j;
        double best = fabs(a[j + (int64_t)j * lda]);
        for (int i = j + 1; i < m; ++i) {
            const double v = fabs(a[i + (int64_t)j * lda]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        ipiv[j] = p + 1;
        const double piv = a[p + (int64_t)j * lda];
        if (piv != 0.0) {
            if (p != j)
                for (int c = 0; c < n; ++c) {
                    const double t = a[j + (int64_t)c * lda];
                    a[j + (int64_t)c * lda] = a[p + (int64_t)c * lda];
                    a[p + (int64_t)c * lda] = t;
                }
            if (fabs(piv) >= DBL_MIN) {
                const double r = 1.0 / piv;
                for (int i = j + 1; i < m; ++i) a[i + (int64_t)j * lda] *= r;
            } else {
                for (int i = j + 1; i < m; ++i) a[i + (int64_t)j * lda] /= piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; ++c) {
            const double u = a[j + (int64_t)c * lda];
            if (u == 0.0) continue;
            for (int i = j + 1; i < m; ++i) a[i + (int64_t)c * lda] -= a[i + (int64_t)j * lda] * u;
        }
    }
    return info;
}
