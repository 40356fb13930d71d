// Emulated DGEMM on the int8 matrix cores (Ozaki scheme, int8 slices).
//
// No reference counterpart: the reference computes FP64 GEMM with FP64 FMAs
// (src/kernel/gemm_kernel.S:11-23).  Method: error-free splitting of Ozaki,
// Ogita and Oishi (2012) with int8 slices and exact int32 accumulation as in
// Ootomo, Ozaki and Yokota, "DGEMM on integer matrix multiplication unit"
// (IJHPCA 2024).
//
//   split:  every row i of op(A) is scaled by 2^-e_i (e_i from frexp of the
//           row's max |a|, so |a|*2^-e_i < 1) and cut into s slices with
//           round-to-nearest: y = a*2^-e_i; v = 64*y for t = 1 and v = 128*y
//           after, q_t = rint(v) in [-64, 64], y = v - q_t in [-1/2, 1/2] (all
//           exact), giving a = 2^e_i * sum_t 2^(1-7t) q_t + O(2^(e_i-7s)).
//           Rounding (not truncation) keeps the slicing error unbiased, so it
//           grows like sqrt(K) in a dot product instead of like K.
//           Columns of op(B) likewise (exponents f_j, slices r_u).
//   product: C_ij = alpha * 2^(e_i+f_j) * sum_{t+u<=s+1} 2^(2-7(t+u)) (Q_t R_u)_ij
//           + beta*C_ij.  Each Q_t R_u is an int8 GEMM accumulated exactly in
//           int32 (|q|,|r| <= 64, so exact while the number of k terms since
//           the last flush stays below 2^31/64^2); pairs with the same t+u
//           share an int32 accumulator, which is flushed into an FP64
//           accumulator scaled by 2^(2-7(t+u)).
// One kernel runs all s(s+1)/2 slice products of a 256x256 output tile back to
// back (the double-buffered LDS pipeline never drains between products) and
// applies the exponents, alpha and beta in its epilogue; FP64 stays only in
// the split, the flush and the epilogue.
#include "common.hpp"
#include "ozaki.hpp"
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <type_traits>

namespace gemm_amd {

static constexpr int OZ_ROWPAD = 768;            // rows padded to a multiple of every tile height (lcm 256, 192)
// LDS rows are KB + 16 bytes (KB = k bytes per stage, 64 or 128): the 16-byte
// fragment reads of 16 consecutive rows land on distinct 16-byte bank groups
// (80*q and 144*q mod 256 are permutations of 16*q).
static constexpr int OZ_FLUSH_K = 524160;         // 524160 * 64^2 < 2^31 - 1

int ozaki_pad_rows(int x) { return (x + OZ_ROWPAD - 1) / OZ_ROWPAD * OZ_ROWPAD; }
int ozaki_pad_k(int x) { return (x + 127) / 128 * 128; }

size_t ozaki_workspace_bytes(int rows, int k, int nslices) {
    const size_t rp = (size_t)ozaki_pad_rows(rows), kp = (size_t)ozaki_pad_k(k);
    return (size_t)nslices * rp * kp + rp * sizeof(int) + rp * sizeof(unsigned long long) + 256;
}

size_t ozaki_accum_bytes(int m, int n) { return (size_t)ozaki_pad_rows(m) * (size_t)ozaki_pad_rows(n) * 12; }

// ---------------------------------------------------------------- split
// X is viewed as R rows x K columns: element (r, k) at X[r*rs + k*ks].
__global__ __launch_bounds__(256) void oz_zero_u64_kernel(unsigned long long *p, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0ull;
}

// All 16 loads of a thread are issued before any is used (entries past the
// matrix edge read element 0 and are replaced by zero afterwards), so the
// tile costs one memory latency, not sixteen.  R, K >= 1.
__device__ __forceinline__ void oz_load_tile(const double *__restrict__ X, int64_t rs, int64_t ks, int R, int K,
                                             int r0, int k0, double (*T)[65]) {
    const int tid = threadIdx.x;
    auto coord = [&](int i, int &rr, int &kk) {
        const int idx = tid + 256 * i;
        if (rs == 1) { rr = idx & 63; kk = idx >> 6; }   // rows contiguous in memory
        else { kk = idx & 63; rr = idx >> 6; }           // k contiguous
    };
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int rr, kk;
        coord(i, rr, kk);
        const int r = r0 + rr, k = k0 + kk;
        v[i] = X[(r < R && k < K) ? (int64_t)r * rs + (int64_t)k * ks : 0];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int rr, kk;
        coord(i, rr, kk);
        T[rr][kk] = (r0 + rr < R && k0 + kk < K) ? v[i] : 0.0;
    }
}

__global__ __launch_bounds__(256) void oz_rowmax_kernel(const double *__restrict__ X, int64_t rs, int64_t ks, int R,
                                                        int K, unsigned long long *__restrict__ rowmax) {
    __shared__ double T[64][65];
    const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
    oz_load_tile(X, rs, ks, R, K, r0, k0, T);
    __syncthreads();
    const int tid = threadIdx.x, rr = tid >> 2, part = tid & 3;
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) m = fmax(m, fabs(T[rr][part * 16 + i]));
    m = fmax(m, __shfl_xor(m, 1));
    m = fmax(m, __shfl_xor(m, 2));
    // non-negative doubles order like their bit patterns
    if (part == 0 && r0 + rr < R && m > 0.0) atomicMax(&rowmax[r0 + rr], (unsigned long long)__double_as_longlong(m));
}

// grid (Kp/64, Rp/64): writes slices [t][Rp][Kp] (k contiguous; zero in the
// padding) and exps[r] for every padded row.
__global__ __launch_bounds__(256) void oz_slice_kernel(const double *__restrict__ X, int64_t rs, int64_t ks, int R,
                                                       int K, int Rp, int Kp, int ns,
                                                       const unsigned long long *__restrict__ rowmax,
                                                       int8_t *__restrict__ out, int *__restrict__ exps) {
    __shared__ double T[64][65];
    const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
    oz_load_tile(X, rs, ks, R, K, r0, k0, T);
    __syncthreads();
    const int tid = threadIdx.x, rr = tid >> 2, kq = (tid & 3) * 16;
    const double amax = __longlong_as_double((long long)rowmax[r0 + rr]);
    int e = 0;
    if (amax > 0.0) frexp(amax, &e);
    if (blockIdx.x == 0 && (tid & 3) == 0) exps[r0 + rr] = e;
    double y[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) y[i] = ldexp(T[rr][kq + i], -e);
    const int64_t plane = (int64_t)Rp * Kp;
    int8_t *dst = out + (int64_t)(r0 + rr) * Kp + k0 + kq;
    for (int t = 0; t < ns; ++t) {
        i4_t w = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const double v = y[i] * (t == 0 ? 64.0 : 128.0);
            const double q = rint(v);
            y[i] = v - q;
            w[i >> 2] |= ((int)q & 0xFF) << (8 * (i & 3));
        }
        *reinterpret_cast<i4_t *>(dst + t * plane) = w;
    }
}

// ---------------------------------------------------------------- int8 GEMM
// Slice product number pr of the triangle t+u <= s+1, ordered by d = t+u and
// then t: diagonal d holds t = 1..d-1, so pr -> (t, u, closes-its-diagonal)
// by a short scalar walk (kernel arguments are never indexed dynamically,
// which would copy them to scratch memory).
__device__ __forceinline__ void oz_pair(int pr, int &t, int &u, bool &last) {
    int d = 2, base = 0;
    while (pr >= base + (d - 1)) {
        base += d - 1;
        ++d;
    }
    t = pr - base + 1;
    u = d - t;
    last = (t == d - 1);
}

struct OzParams {
    int M, N, Kp, npairs, ns, tiles_m, tiles_n;
    int64_t planeA, planeB;       // bytes per slice plane
    const int8_t *SA, *SB;
    const int *ea, *eb;
    double *W;                    // FP64 partial sums per tile, accumulator (fragment) order
    int *Wi;                      // int32 staging of the accumulators, same order
    double alpha, beta;
    double *C;
    int64_t ldc;
};

// One workgroup = one BM x BN output tile of WM x WN waves, each wave a
// (32*TM) x (32*TN) block of v_mfma_i32_32x32x32_i8 tiles (16 int32
// accumulators per tile per lane).  When a diagonal t+u is complete its int32
// sums are scaled and added into the tile's FP64 partial sums W, stored in
// accumulator order so every flush is a coalesced 512-byte row per 64 lanes;
// after the last product a memory loop over W writes column-major C.
template <int WM, int WN, int TM, int TN, int KB>
__global__ __launch_bounds__(WM *WN * 64) void ozaki_gemm_kernel(OzParams p) {
    constexpr int KP = KB + 16;   // LDS row pitch (bytes)
    constexpr int NT = WM * WN * 64;
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    constexpr int STAGE = (BM + BN) * KP;
    constexpr int CH = KB / 16;                 // 16-byte chunks per row per stage
    constexpr int PA = BM * CH / NT, PB = BN * CH / NT;
    static_assert(BM * CH % NT == 0 && BN * CH % NT == 0, "stage must split evenly over the workgroup");
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WM) * (TM * 32), wn0 = (wave / WM) * (TN * 32);
    int tm, tn;
    tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, 8, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int nkt = p.Kp / KB;
    const int total = p.npairs * nkt;
    const int64_t wslot = ((int64_t)(tm * p.tiles_n + tn) * (WM * WN) + wave) * (TM * TN * 16 * 64);
    double *W = p.W + wslot;
    int *Wi = p.Wi + wslot;

    i4_t ra[PA], rb[PB];
    auto gload = [&](int st) {
        const int pr = st / nkt, kt = st - pr * nkt;
        int pt, pu;
        bool pl;
        oz_pair(pr, pt, pu, pl);
        const int8_t *a = p.SA + (int64_t)(pt - 1) * p.planeA + (int64_t)m0 * p.Kp + kt * KB;
        const int8_t *b = p.SB + (int64_t)(pu - 1) * p.planeB + (int64_t)n0 * p.Kp + kt * KB;
#pragma unroll
        for (int i = 0; i < PA; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            ra[i] = *reinterpret_cast<const i4_t *>(a + (int64_t)row * p.Kp + ch);
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            rb[i] = *reinterpret_cast<const i4_t *>(b + (int64_t)row * p.Kp + ch);
        }
    };
    auto sstore = [&](unsigned char *buf) {
#pragma unroll
        for (int i = 0; i < PA; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            *reinterpret_cast<i4_t *>(buf + row * KP + ch) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            *reinterpret_cast<i4_t *>(buf + (BM + row) * KP + ch) = rb[i];
        }
    };

    const int q = lane & 31, g = lane >> 5;
    const int aoff = (wm0 + q) * KP + 16 * g;
    const int boff = (BM + wn0 + q) * KP + 16 * g;
    i16_t acc[TM][TN];

    if (total > 0) {
        gload(0);
        sstore(smem);
    }
    __syncthreads();
    // Segments of the stage sequence: one per diagonal d = t+u (split further
    // if its k-term count could overflow int32).  Each segment starts from zero
    // accumulators and ends with a flush; the global-load prefetch runs on the
    // global stage index, so the pipeline does not drain between segments.
    constexpr int CHUNK = OZ_FLUSH_K / KB;
    int st = 0;
    bool first = true;
    for (int d = 2; d <= p.ns + 1; ++d) {
        const int dstages = (d - 1) * nkt;
        const double scale = ldexp(1.0, 2 - 7 * d);
        for (int c0 = 0; c0 < dstages; c0 += CHUNK) {
            const int c1 = min(dstages, c0 + CHUNK);
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
#pragma unroll
                    for (int e = 0; e < 16; ++e) acc[a][b][e] = 0;
            for (int i = c0; i < c1; ++i, ++st) {
                const unsigned char *cur = smem + (st & 1) * STAGE;
                unsigned char *nxt = smem + ((st + 1) & 1) * STAGE;
                if (st + 1 < total) gload(st + 1);
                // the fragments of k-step ks + 1 are read while the matrix
                // instructions of k-step ks run (the sched barriers keep the
                // reads from being sunk next to their uses)
                i4_t av[2][TM], bv[2][TN];
                auto frag = [&](int ks, i4_t *a, i4_t *b) {
#pragma unroll
                    for (int s = 0; s < TM; ++s) a[s] = *reinterpret_cast<const i4_t *>(cur + aoff + 32 * s * KP + 32 * ks);
#pragma unroll
                    for (int s = 0; s < TN; ++s) b[s] = *reinterpret_cast<const i4_t *>(cur + boff + 32 * s * KP + 32 * ks);
                };
                frag(0, av[0], bv[0]);
#pragma unroll
                for (int ks = 0; ks < KB / 32; ++ks) {
                    if (ks + 1 < KB / 32) frag(ks + 1, av[(ks + 1) & 1], bv[(ks + 1) & 1]);
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int a = 0; a < TM; ++a)
#pragma unroll
                        for (int b = 0; b < TN; ++b)
                            acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[ks & 1][a], bv[ks & 1][b], acc[a][b],
                                                                              0, 0, 0);
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (st + 1 < total) sstore(nxt);
                __syncthreads();
            }
            // flush: the accumulators leave through plain int32 stores (which
            // read the accumulation registers directly; converting in place
            // would pull them all into the VGPR file), then the wave folds its
            // own rows into the FP64 partial sums
            // an opaque base keeps the compiler from hoisting the 16*TM*TN store
            // addresses out of the stage loop (they would pin hundreds of VGPRs)
            int *wb = Wi + lane;
            __asm__ volatile("" : "+v"(wb));
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
#pragma unroll
                    for (int e = 0; e < 16; ++e) wb[((a * TN + b) * 16 + e) * 64] = acc[a][b][e];
            for (int idx = 0; idx < TM * TN * 16; ++idx) {
                double *w = W + idx * 64 + lane;
                const double prev = first ? 0.0 : *w;
                *w = fma((double)Wi[idx * 64 + lane], scale, prev);
            }
            first = false;
        }
    }

    // epilogue from the FP64 partial sums (a memory loop: no register arrays
    // are indexed here): C = alpha * 2^(e_i + f_j) * W + beta * C
    for (int idx = 0; idx < TM * TN * 16; ++idx) {
        const int ab = idx >> 4, e = idx & 15, a = ab / TN, b = ab % TN;
        const int i = m0 + wm0 + 32 * a + (e >> 2) * 8 + g * 4 + (e & 3);
        const int j = n0 + wn0 + 32 * b + q;
        const double sum = first ? 0.0 : W[idx * 64 + lane];
        if (i < p.M && j < p.N) {
            const double v = p.alpha * ldexp(sum, p.ea[i] + p.eb[j]);
            double *dst = p.C + i + (int64_t)j * p.ldc;
            *dst = (p.beta == 0.0) ? v : v + p.beta * *dst;
        }
    }
}

// ---------------------------------------------------------------- host
static int grid1(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

// Splits the rows of the R x K view (element (r,k) at X[r*rs + k*ks]) into ws.
static int ozaki_split(const double *X, int64_t rs, int64_t ks, int R, int K, int ns, void *ws, int8_t **slices,
                       int **exps, hipStream_t s) {
    const int Rp = ozaki_pad_rows(R), Kp = ozaki_pad_k(K);
    unsigned char *base = static_cast<unsigned char *>(ws);
    *slices = reinterpret_cast<int8_t *>(base);
    size_t off = (size_t)ns * Rp * Kp;
    off = (off + 255) & ~(size_t)255;
    unsigned long long *rowmax = reinterpret_cast<unsigned long long *>(base + off);
    off += (size_t)Rp * sizeof(unsigned long long);
    *exps = reinterpret_cast<int *>(base + off);
    hipLaunchKernelGGL(oz_zero_u64_kernel, dim3(grid1(Rp)), dim3(256), 0, s, rowmax, Rp);
    if (K > 0 && R > 0)
        hipLaunchKernelGGL(oz_rowmax_kernel, dim3((K + 63) / 64, (R + 63) / 64), dim3(256), 0, s, X, rs, ks, R, K,
                           rowmax);
    hipLaunchKernelGGL(oz_slice_kernel, dim3(Kp / 64, Rp / 64), dim3(256), 0, s, X, rs, ks, R, K, Rp, Kp, ns, rowmax,
                       *slices, *exps);
    return (int)hipGetLastError();
}

int ozaki_pairs(int ns, int8_t *t, int8_t *u, int8_t *last) {
    int n = 0;
    for (int d = 2; d <= ns + 1; ++d)
        for (int a = 1; a <= ns; ++a) {
            const int b = d - a;
            if (b < 1 || b > ns) continue;
            t[n] = (int8_t)a;
            u[n] = (int8_t)b;
            last[n] = 0;
            ++n;
        }
    // mark the last pair of every diagonal
    for (int i = 0; i < n; ++i) last[i] = (i + 1 == n) || (t[i + 1] + u[i + 1] != t[i] + u[i]);
    return n;
}

template <int WM, int WN, int TM, int TN, int KB>
static int launch_oz_variant(OzParams &p, hipStream_t s) {
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    p.tiles_m = (p.M + BM - 1) / BM;
    p.tiles_n = (p.N + BN - 1) / BN;
    constexpr size_t lds = 2ull * (BM + BN) * (KB + 16);
    auto kern = ozaki_gemm_kernel<WM, WN, TM, TN, KB>;
    if (hipError_t he = lds_limit_once<ozaki_gemm_kernel<WM, WN, TM, TN, KB>>(lds)) return (int)he;
    hipLaunchKernelGGL(kern, dim3(p.tiles_m * p.tiles_n), dim3(WM * WN * 64), lds, s, p);
    return (int)hipGetLastError();
}

int launch_dgemm_ozaki(bool ta, bool tb, int m, int n, int k, double alpha, const double *A, int64_t lda,
                       const double *B, int64_t ldb, double beta, double *C, int64_t ldc, int nslices, void *wsA,
                       void *wsB, void *wsW, hipStream_t s) {
    if (m <= 0 || n <= 0 || k <= 0) return 0;
    // op(A) rows: element (i, p) at A[i + p*lda] (N) or A[p + i*lda] (T)
    int8_t *sa, *sb;
    int *ea, *eb;
    int e = ozaki_split(A, ta ? lda : 1, ta ? 1 : lda, m, k, nslices, wsA, &sa, &ea, s);
    if (e) return e;
    // op(B) columns as rows: element (j, p) = op(B)[p][j] at B[p + j*ldb] (N) or B[j + p*ldb] (T)
    e = ozaki_split(B, tb ? 1 : ldb, tb ? ldb : 1, n, k, nslices, wsB, &sb, &eb, s);
    if (e) return e;
    OzParams p;
    p.M = m;
    p.N = n;
    p.Kp = ozaki_pad_k(k);
    p.planeA = (int64_t)ozaki_pad_rows(m) * p.Kp;
    p.planeB = (int64_t)ozaki_pad_rows(n) * p.Kp;
    p.SA = sa;
    p.SB = sb;
    p.ea = ea;
    p.eb = eb;
    p.W = static_cast<double *>(wsW);
    p.Wi = reinterpret_cast<int *>(static_cast<char *>(wsW) + (size_t)ozaki_pad_rows(m) * ozaki_pad_rows(n) * 8);
    p.alpha = alpha;
    p.beta = beta;
    p.C = C;
    p.ldc = ldc;
    p.ns = nslices;
    p.npairs = nslices * (nslices + 1) / 2;
    static const char *var = getenv("GEMM_AMD_OZ_VARIANT");
    const int v = var ? atoi(var) : 0;
    if (v == 1) return launch_oz_variant<2, 4, 4, 2, 64>(p, s);    // 256x256 tile, 8 waves of 128x64, 64-deep
    if (v == 2) return launch_oz_variant<2, 2, 3, 4, 128>(p, s);   // 192x256 tile, 4 waves of 96x128
    return launch_oz_variant<2, 4, 4, 2, 128>(p, s);               // 256x256 tile, 8 waves of 128x64
}

}  // namespace gemm_amd

namespace gemm_amd {

// Exposes the split step alone: slices (padded [t][Rp][Kp]) and exponents of
// the rows of the R x K view of X; rowmax_ws holds Rp 64-bit words.
int launch_ozaki_split_only(const double *X, int64_t rs, int64_t ks, int R, int K, int ns, int8_t *slices,
                            int Rp, int Kp, int *exps, unsigned long long *rowmax_ws, hipStream_t s) {
    // the caller's planes are Rp x Kp with Rp >= R, Kp >= K, both multiples of 64
    if (R <= 0 || K <= 0) return 0;
    hipLaunchKernelGGL(oz_zero_u64_kernel, dim3(grid1(Rp)), dim3(256), 0, s, rowmax_ws, Rp);
    hipLaunchKernelGGL(oz_rowmax_kernel, dim3((K + 63) / 64, (R + 63) / 64), dim3(256), 0, s, X, rs, ks, R, K,
                       rowmax_ws);
    hipLaunchKernelGGL(oz_slice_kernel, dim3(Kp / 64, Rp / 64), dim3(256), 0, s, X, rs, ks, R, K, Rp, Kp, ns,
                       rowmax_ws, slices, exps);
    return (int)hipGetLastError();
}

}  // namespace gemm_amd

// ============================================================================
// Emulated DGEMM on the int8 matrix cores, modular (CRT) form.
//
// Method: Ozaki, Uchino and Imamura, "Ozaki Scheme II: A GEMM-oriented
// emulation of floating-point matrix multiplication using an integer modular
// technique" (2025).  No reference counterpart.
//
//   scale:   every row i of op(A) becomes integers A'_ik = rint(a_ik 2^(al-e_i))
//            (e_i from frexp of the row's max |a|, so |A'| <= 2^al); columns
//            of op(B) likewise with f_j.  al is the largest value with
//            K 2^(2 al) <= M/4, M = m_1 ... m_N (pairwise coprime moduli
//            <= 256), so the exact integer product V = A'B' has |V| <= M/4.
//   residues: A'_ik mod m_l and B'_kj mod m_l in symmetric form fit int8.
//   products: one int8 GEMM per modulus with exact int32 accumulation
//            (|r| <= 128: exact for 131008 k terms, longer K is reduced in
//            between), reduced mod m_l to an int8 residue of V.
//   CRT:     V = sum_l r_l W_l - q M with W_l = (M/m_l) ((M/m_l)^-1 mod m_l);
//            W_l and M are split into four 32-bit limbs, so the limb sums
//            s_k = sum_l r_l w_lk (< 2^43) and s_k - q M_k are exact in FP64,
//            q = rint(s/M) is exact because |V|/M <= 1/4, and a carry pass
//            in int64 rebuilds V, rounded once or twice to FP64.
//   C_ij = alpha 2^(e_i + f_j - 2 al) V_ij + beta C_ij.
// N moduli cost N int8 GEMMs (15 or 16 give FP64-level accuracy at K = 8192)
// against s(s+1)/2 - 1 for s slices.
// ============================================================================

namespace gemm_amd {

struct CrtConst {
    int nmod, pad;
    int m[OZ_CRT_MAXMOD];
    double invm[OZ_CRT_MAXMOD];
    double w[OZ_CRT_MAXMOD][4];   // limbs of W_l, low first, each < 2^32
    double Mk[4];                 // limbs of M
    double invM;                  // 1 / M (rounded)
    double c32[OZ_CRT_MAXMOD];    // 2^32 mod m_l
};

static constexpr int OZ_CRT_CHUNK_K = 131008;   // 131008 * 128^2 < 2^31

// grid (Kp/64, Rp/64): residues [l][Rp][Kp] (zero padding) and exps[r] = e_r - al.
__global__ __launch_bounds__(256) void oz_residue_kernel(const double *__restrict__ X, int64_t rs, int64_t ks, int R,
                                                         int K, int Rp, int Kp, int pitch, int al,
                                                         const unsigned long long *__restrict__ rowmax,
                                                         const CrtConst *__restrict__ cc, int8_t *__restrict__ out,
                                                         int *__restrict__ exps) {
    __shared__ double T[64][65];
    const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
    oz_load_tile(X, rs, ks, R, K, r0, k0, T);
    __syncthreads();
    const int tid = threadIdx.x, rr = tid >> 2, kq = (tid & 3) * 16;
    const double amax = __longlong_as_double((long long)rowmax[r0 + rr]);
    int e = 0;
    if (amax > 0.0) frexp(amax, &e);
    if (blockIdx.x == 0 && (tid & 3) == 0) exps[r0 + rr] = e - al;
    // x = rint(a 2^(al-e)) (|x| <= 2^al, an exact integer) is split once as
    // x = xh 2^32 + xl with 0 <= xl < 2^32 (both exact); per modulus
    // y = xh (2^32 mod m) + xl is exact, |y| < 2^33, so rint(y/m) is the
    // exact nearest quotient (ties only at m = 256, where either residue
    // gives the same byte) and y - m rint(y/m) is the symmetric residue
    double xh[16], xl[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const double x = rint(ldexp(T[rr][kq + i], al - e));
        xh[i] = floor(x * 0x1p-32);
        xl[i] = fma(-xh[i], 0x1p32, x);
    }
    const int64_t plane = (int64_t)Rp * pitch;
    int8_t *dst = out + (int64_t)(r0 + rr) * pitch + k0 + kq;
    const int nmod = cc->nmod;
    for (int l = 0; l < nmod; ++l) {
        const double m = (double)cc->m[l], im = cc->invm[l], c = cc->c32[l];
        i4_t w = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const double y = fma(xh[i], c, xl[i]);
            const double r = fma(-rint(y * im), m, y);
            w[i >> 2] |= ((int)r & 0xFF) << (8 * (i & 3));
        }
        *reinterpret_cast<i4_t *>(dst + l * plane) = w;
    }
}

struct CrtParams {
    int M, N, Kp, tiles_m, tiles_n;
    int pitch;                    // bytes between rows of a residue plane (>= Kp)
    int64_t planeA, planeB;       // bytes per residue plane
    const int8_t *RA, *RB;
    const int *ea, *eb;
    const CrtConst *cc;
    int *Wi;                      // int32 staging of the accumulators (per wave, fragment order)
    int *Wc;                      // running residues when K spans several exact chunks, same order
    unsigned *Rv;                 // product residues, 4 per word, per wave [l][idx/4][lane]
    double alpha, beta;
    double *C;
    int64_t ldc;
};

// Symmetric residue of an int32 mod m (exact: |x| < 2^31, FP64 quotient).
__device__ __forceinline__ int crt_mod(int x, int m, double im) {
    const int q = (int)rint((double)x * im);
    int r = x - q * m;
    const int h = m >> 1;
    r = r > h ? r - m : (r < -h ? r + m : r);
    return r;
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void global_void_t;

// DIRECT staging: global_load_lds_dwordx4 moves each 16-byte chunk from HBM/L2
// straight into LDS (lane l of an instruction writes LDS base + 16 l), so no
// VGPRs hold the next stages and no LDS store instructions stall the waves;
// NS stage buffers keep NS - 1 stages in flight.  Rows are unpadded (KB bytes)
// and chunk c of row r sits at position c ^ sw(r), sw(r) = (r >> 1) & 7 for
// 128-byte rows, (r >> 2) & 3 for 64-byte rows: the 16-byte fragment reads of
// 16 consecutive rows fall on 16 distinct 16-byte groups of the 256-byte bank
// width.
__device__ __forceinline__ void wait_vmem_at_most(int n) {   // n wave-uniform; waits for at most n (rounded down)
    if (n >= 32) asm volatile("s_waitcnt vmcnt(32)" ::: "memory");
    else if (n >= 16) asm volatile("s_waitcnt vmcnt(16)" ::: "memory");
    else if (n >= 8) asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
    else if (n >= 4) asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
    else if (n >= 2) asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <int WM, int WN, int TM, int TN, int KB, bool DIRECT, int NS = 2>
__global__ __launch_bounds__(WM *WN * 64) void ozaki_crt_kernel(CrtParams p) {
    static_assert(!DIRECT || KB == 128 || KB == 64, "direct staging takes 64- or 128-byte rows");
    static_assert(DIRECT || NS == 2, "register staging is double-buffered");
    constexpr int KP = DIRECT ? KB : KB + 16;
    constexpr int NT = WM * WN * 64;
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    constexpr int STAGE = (BM + BN) * KP;
    constexpr int CH = KB / 16;
    constexpr int PA = BM * CH / NT, PB = BN * CH / NT;
    constexpr int NACC = TM * TN * 16;            // int32 accumulators per lane
    static_assert(BM * CH % NT == 0 && BN * CH % NT == 0, "stage must split evenly over the workgroup");
    extern __shared__ __align__(16) unsigned char smem[];
    // the wave index is uniform: keeping it (and what derives from it) in
    // scalar registers leaves the vector registers to the accumulators
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave % WM) * (TM * 32), wn0 = (wave / WM) * (TN * 32);
    int tm, tn;
    tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, 8, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int nkt = p.Kp / KB;
    const int nmod = p.cc->nmod;
    const int total = nmod * nkt;
    const int64_t wv = (int64_t)(tm * p.tiles_n + tn) * (WM * WN) + wave;
    int *Wi = p.Wi + wv * (NACC * 64);
    int *Wc = p.Wc + wv * (NACC * 64);
    unsigned *Rv = p.Rv + wv * ((int64_t)nmod * (NACC / 4) * 64);

    i4_t ra[PA], rb[PB];
    auto gload = [&](int st) {
        const int l = st / nkt, kt = st - l * nkt;
        const int8_t *a = p.RA + (int64_t)l * p.planeA + (int64_t)m0 * p.pitch + kt * KB;
        const int8_t *b = p.RB + (int64_t)l * p.planeB + (int64_t)n0 * p.pitch + kt * KB;
#pragma unroll
        for (int i = 0; i < PA; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            ra[i] = *reinterpret_cast<const i4_t *>(a + (int64_t)row * p.pitch + ch);
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            rb[i] = *reinterpret_cast<const i4_t *>(b + (int64_t)row * p.pitch + ch);
        }
    };
    auto sstore = [&](unsigned char *buf) {
#pragma unroll
        for (int i = 0; i < PA; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            *reinterpret_cast<i4_t *>(buf + row * KP + ch) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            *reinterpret_cast<i4_t *>(buf + (BM + row) * KP + ch) = rb[i];
        }
    };

    // direct staging: instruction j of wave w fills rows R0 = RPI (w + 8 j) ..
    // R0 + RPI - 1 of the stage (A rows first, then B rows); lane l loads
    // chunk (l % CPR) ^ sw(row) of row R0 + l / CPR
    constexpr int CPR = KB / 16;                              // chunks per row
    constexpr int RPI = 64 / CPR;                             // rows per load instruction
    constexpr int NW = NT / 64;                               // waves
    constexpr int LPW = (BM + BN) / RPI / NW;                 // load instructions per wave per stage
    constexpr int JA = BM / (RPI * NW);                       // of which fill A rows
    static_assert(BM % (RPI * NW) == 0 && NW % 2 == 0, "whole A/B instructions, even wave count");
    auto swz = [](int row) { return CPR == 8 ? (row >> 1) & 7 : (row >> 2) & 3; };
    // lane l of instruction j loads row RPI (wave + NW j) + l / CPR; with an
    // even wave count its swizzle, and so its byte offset in the row, does not
    // depend on j: one per-lane offset plus a uniform base per instruction
    const int lrow = lane / CPR;
    const int loff = lrow * p.pitch + 16 * ((lane % CPR) ^ swz(RPI * wave + lrow));
    auto dload = [&](int st) {
        const int l = st / nkt, kt = st - l * nkt;
        unsigned char *buf = smem + (st % NS) * STAGE;
        const int8_t *ua = p.RA + (int64_t)l * p.planeA + (int64_t)(m0 + RPI * wave) * p.pitch + kt * KB;
        const int8_t *ub = p.RB + (int64_t)l * p.planeB + (int64_t)(n0 + RPI * wave) * p.pitch + kt * KB;
        const int64_t js = (int64_t)(RPI * NW) * p.pitch;
#pragma unroll
        for (int j = 0; j < LPW; ++j) {
            const int8_t *u = j < JA ? ua + j * js : ub + (j - JA) * js;
            __builtin_amdgcn_global_load_lds((global_void_t *)(u + loff), (lds_void_t *)(buf + RPI * (wave + NW * j) * KP),
                                             16, 0, 0);
        }
    };

    const int q = lane & 31, g = lane >> 5;
    const int aoff = (wm0 + q) * KP + (DIRECT ? 0 : 16 * g);
    const int boff = (BM + wn0 + q) * KP + (DIRECT ? 0 : 16 * g);
    const int qx = swz(q);   // swizzle of the lane's rows (32-row blocks keep it)
    i16_t acc[TM][TN];

    if (total > 0) {
        if constexpr (DIRECT) {
            const int pre = min(NS - 1, total);
            for (int i = 0; i < pre; ++i) dload(i);
            wait_vmem_at_most(LPW * (pre - 1));   // stage 0 has landed
        } else {
            gload(0);
            sstore(smem);
        }
    }
    __syncthreads();
    constexpr int CHUNK = OZ_CRT_CHUNK_K / KB;
    int st = 0;
    for (int l = 0; l < nmod; ++l) {
        const int m = p.cc->m[l];
        const double im = p.cc->invm[l];
        const bool multi = nkt > CHUNK;   // K longer than one exact int32 chunk
        for (int c0 = 0; c0 < nkt; c0 += CHUNK) {
            const int c1 = min(nkt, c0 + CHUNK);
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
#pragma unroll
                    for (int e = 0; e < 16; ++e) acc[a][b][e] = 0;
            for (int i = c0; i < c1; ++i, ++st) {
                const unsigned char *cur = smem + (DIRECT ? st % NS : st & 1) * STAGE;
                unsigned char *nxt = smem + ((st + 1) & 1) * STAGE;
                if constexpr (DIRECT) {
                    // the buffer of stage st + NS - 1 was last read in stage st - 1
                    if (st + NS - 1 < total) dload(st + NS - 1);
                } else {
                    if (st + 1 < total) gload(st + 1);
                }
                if constexpr (DIRECT) {
                    // fragments of k-step ks + 1 are read while the matrix
                    // instructions of k-step ks run
                    i4_t av[2][TM], bv[2][TN];
                    auto frag = [&](int ks, i4_t *a, i4_t *b) {
                        const int ko = ((2 * ks + g) ^ qx) << 4;
#pragma unroll
                        for (int s = 0; s < TM; ++s)
                            a[s] = *reinterpret_cast<const i4_t *>(cur + aoff + 32 * s * KP + ko);
#pragma unroll
                        for (int s = 0; s < TN; ++s)
                            b[s] = *reinterpret_cast<const i4_t *>(cur + boff + 32 * s * KP + ko);
                    };
                    frag(0, av[0], bv[0]);
#pragma unroll
                    for (int ks = 0; ks < KB / 32; ++ks) {
                        if (ks + 1 < KB / 32) frag(ks + 1, av[(ks + 1) & 1], bv[(ks + 1) & 1]);
                        // the scheduler would otherwise sink each read next to
                        // its first use and stall the matrix pipe on LDS latency
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int a = 0; a < TM; ++a)
#pragma unroll
                            for (int b = 0; b < TN; ++b)
                                acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[ks & 1][a], bv[ks & 1][b],
                                                                                  acc[a][b], 0, 0, 0);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                } else
#pragma unroll
                for (int ks = 0; ks < KB / 32; ++ks) {
                    // byte offset of the lane's 16-byte k chunk 2 ks + g in its row
                    const int ko = DIRECT ? (((2 * ks + g) ^ qx) << 4) : 32 * ks;
                    i4_t av[TM], bv[TN];
#pragma unroll
                    for (int s = 0; s < TM; ++s)
                        av[s] = *reinterpret_cast<const i4_t *>(cur + aoff + 32 * s * KP + ko);
#pragma unroll
                    for (int s = 0; s < TN; ++s)
                        bv[s] = *reinterpret_cast<const i4_t *>(cur + boff + 32 * s * KP + ko);
#pragma unroll
                    for (int a = 0; a < TM; ++a)
#pragma unroll
                        for (int b = 0; b < TN; ++b)
                            acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[a], bv[b], acc[a][b], 0, 0, 0);
                    __builtin_amdgcn_sched_barrier(0);
                }
                if constexpr (DIRECT) {
                    // this wave's chunks of stage st + 1 have landed in LDS once
                    // at most the loads of the stages after it are outstanding
                    if (st + 1 < total) wait_vmem_at_most(LPW * (min(total - 1, st + NS - 1) - (st + 1)));
                } else {
                    if (st + 1 < total) sstore(nxt);
                }
                __syncthreads();
            }
            if (!multi) {
                // one exact chunk: residues straight from the accumulators,
                // four per word, in accumulator order
                unsigned *rv = Rv + (int64_t)l * (NACC / 4) * 64 + lane;
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int b = 0; b < TN; ++b)
#pragma unroll
                        for (int e4 = 0; e4 < 4; ++e4) {
                            unsigned word = 0;
#pragma unroll
                            for (int e = 0; e < 4; ++e)
                                word |= (unsigned)(crt_mod(acc[a][b][4 * e4 + e], m, im) & 0xFF) << (8 * e);
                            rv[((a * TN + b) * 4 + e4) * 64] = word;
                        }
                continue;
            }
            // several chunks: the accumulators leave through plain int32
            // stores (an opaque base keeps the store addresses from being
            // hoisted into VGPRs) and fold into the running residue
            int *wb = Wi + lane;
            __asm__ volatile("" : "+v"(wb));
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
#pragma unroll
                    for (int e = 0; e < 16; ++e) wb[((a * TN + b) * 16 + e) * 64] = acc[a][b][e];
            for (int idx = 0; idx < NACC; ++idx) {
                const int prev = c0 == 0 ? 0 : Wc[idx * 64 + lane];
                Wc[idx * 64 + lane] = crt_mod(prev + Wi[idx * 64 + lane], m, im);
            }
        }
        if (multi) {
            // residues of this modulus, four per word, in accumulator order
            for (int w4 = 0; w4 < NACC / 4; ++w4) {
                unsigned word = 0;
#pragma unroll
                for (int e = 0; e < 4; ++e) word |= (unsigned)(Wc[(w4 * 4 + e) * 64 + lane] & 0xFF) << (8 * e);
                Rv[((int64_t)l * (NACC / 4) + w4) * 64 + lane] = word;
            }
        }
        // the flush's stores count in the same vector-memory counter as the
        // stage loads in flight: drain it so the counted waits stay exact
        if constexpr (DIRECT) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }

    // CRT reconstruction (a memory loop over the wave's residues)
    const double two32 = 4294967296.0;
    for (int w4 = 0; w4 < NACC / 4; ++w4) {
        double s[4][4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int k = 0; k < 4; ++k) s[e][k] = 0.0;
        for (int l = 0; l < nmod; ++l) {
            const unsigned word = Rv[((int64_t)l * (NACC / 4) + w4) * 64 + lane];
            const double w0 = p.cc->w[l][0], w1 = p.cc->w[l][1], w2 = p.cc->w[l][2], w3 = p.cc->w[l][3];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double r = (double)(int)(signed char)(word >> (8 * e));
                s[e][0] = fma(r, w0, s[e][0]);
                s[e][1] = fma(r, w1, s[e][1]);
                s[e][2] = fma(r, w2, s[e][2]);
                s[e][3] = fma(r, w3, s[e][3]);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = w4 * 4 + e;
            const int ab = idx >> 4, ee = idx & 15, a = ab / TN, b = ab % TN;
            const int i = m0 + wm0 + 32 * a + (ee >> 2) * 8 + g * 4 + (ee & 3);
            const int j = n0 + wn0 + 32 * b + q;
            if (i >= p.M || j >= p.N) continue;
            const double approx = (s[e][3] * two32 + s[e][2]) * (two32 * two32) + s[e][1] * two32;
            const double qq = rint(approx * p.cc->invM);
            long long t0 = (long long)fma(-qq, p.cc->Mk[0], s[e][0]);
            long long t1 = (long long)fma(-qq, p.cc->Mk[1], s[e][1]);
            long long t2 = (long long)fma(-qq, p.cc->Mk[2], s[e][2]);
            long long t3 = (long long)fma(-qq, p.cc->Mk[3], s[e][3]);
            // carries: t0, t1, t2 into [0, 2^32) (floor division by 2^32)
            constexpr long long B32 = 4294967296ll;
            long long c = t0 >> 32;
            t0 -= c * B32;
            t1 += c;
            c = t1 >> 32;
            t1 -= c * B32;
            t2 += c;
            c = t2 >> 32;
            t2 -= c * B32;
            t3 += c;
            // V = hi 2^64 + lo with lo taken signed, so small |V| loses nothing
            long long hi = t3 * B32 + t2;
            unsigned long long ulo = ((unsigned long long)t1 << 32) | (unsigned long long)t0;
            long long lo = (long long)ulo;
            if (ulo >> 63) hi += 1;   // lo as signed = ulo - 2^64
            const double V = fma((double)hi, two32 * two32, (double)lo);
            const double v = p.alpha * ldexp(V, p.ea[i] + p.eb[j]);
            double *dst = p.C + i + (int64_t)j * p.ldc;
            *dst = (p.beta == 0.0) ? v : v + p.beta * *dst;
        }
    }
}

// Wide form of the CRT product kernel: one 256x256 output tile per workgroup
// of 4 waves, one wave per SIMD, each wave a 128x128 block of 4x4
// v_mfma_i32_32x32x32_i8 tiles (256 int32 accumulators per lane, held in the
// accumulation registers).  Against the 8-wave 128x64 form it reads 8
// fragments per 16 matrix instructions instead of 6 per 8, so LDS carries a
// third less per operation, and one wave per SIMD leaves the other half of
// the register file to the fragments and addresses.
//
// Pipeline (NS stages of KB k-bytes each, rows XOR-swizzled as above):
//   * stage st is multiplied out of LDS buffer st % NS while the direct
//     loads of stages st+1 .. st+NS-1 are in flight;
//   * a k-step's fragments are read while the previous k-step's matrix
//     instructions run (read and matrix instructions interleaved);
//   * in the last k-step of stage st, after its first matrix instructions:
//     wait for this wave's loads of stage st+1, one barrier (all waves have
//     then landed stage st+1 and finished reading stage st), issue the loads
//     of stage st+NS into buffer st % NS and read stage st+1's first
//     fragments, all while the rest of the k-step's matrix instructions run.
//     The barrier is the stage's only synchronisation point.
//   * when a modulus' last stage is done, its residues leave the
//     accumulators (exact FP64 quotient; m = 256 is the low byte) as one
//     16-byte store per 32x32 tile, four residues per word.
// K must fit one exact int32 chunk (Kp <= OZ_CRT_CHUNK_K); the caller uses
// ozaki_crt_kernel otherwise.
template <int KB, int NS, int PARTS = 1>
__global__ __launch_bounds__(256, 1) void ozaki_crt_wide_kernel(CrtParams p) {
    static_assert(KB == 128 || KB == 64, "64- or 128-byte rows");
    static_assert(PARTS == 1 || (NS == 2 && PARTS <= KB / 32), "split issue: two stages, a part per k-step");
    constexpr int NW = 4, BM = 256, BN = 256;
    constexpr int CPR = KB / 16;                  // 16-byte chunks per row
    constexpr int RPI = 64 / CPR;                 // rows per load instruction
    constexpr int STAGE = (BM + BN) * KB;         // bytes per stage
    constexpr int LPW = (BM + BN) / RPI / NW;     // load instructions per wave per stage
    constexpr int JA = BM / (RPI * NW);           // of which fill A rows
    constexpr int KS = KB / 32;                   // k-steps per stage
    static_assert(KS % 2 == 0, "the next stage's first fragments go to fragment set 0");
    extern __shared__ __align__(16) unsigned char smem[];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave & 1) * 128, wn0 = (wave >> 1) * 128;
    int tm, tn;
    tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, 8, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int nkt = p.Kp / KB;
    const int nmod = p.cc->nmod;
    const int total = nmod * nkt;
    const int64_t wv = (int64_t)(tm * p.tiles_n + tn) * NW + wave;
    // residues of this wave: [modulus][tile ab][lane] x 16 bytes
    i4_t *Rv = reinterpret_cast<i4_t *>(p.Rv) + wv * ((int64_t)nmod * 16 * 64) + lane;

    auto swz = [](int row) { return CPR == 8 ? (row >> 1) & 7 : (row >> 2) & 3; };
    // direct loads: instruction j of wave w fills stage rows RPI (w + NW j) ..
    // + RPI - 1 (A rows first, then B rows); lane l loads chunk
    // (l % CPR) ^ sw(row) of row RPI (w + NW j) + l / CPR, whose swizzle does
    // not depend on j, so one per-lane offset serves every instruction
    const int lrow = lane / CPR;
    const int loff = lrow * p.pitch + 16 * ((lane % CPR) ^ swz(RPI * wave + lrow));
    const int64_t jstride = (int64_t)(RPI * NW) * p.pitch;
    // stage sd's loads (sd < total), one instruction at a time: dma_setup
    // fixes the stage's addresses, dma_one(j) issues instruction j; dl, dkt:
    // the next stage's modulus and k-tile
    int dl = 0, dkt = 0;
    unsigned char *dbuf = nullptr;
    const int8_t *dua = nullptr, *dub = nullptr;
    auto dma_setup = [&](int sd) {
        dbuf = smem + (sd % NS) * STAGE + RPI * wave * KB;
        dua = p.RA + (int64_t)dl * p.planeA + (int64_t)(m0 + RPI * wave) * p.pitch + dkt * KB + loff;
        dub = p.RB + (int64_t)dl * p.planeB + (int64_t)(n0 + RPI * wave) * p.pitch + dkt * KB + loff;
        if (++dkt == nkt) {
            dkt = 0;
            ++dl;
        }
    };
    auto dma_one = [&](int j) {
        const int8_t *u = j < JA ? dua + j * jstride : dub + (j - JA) * jstride;
        __builtin_amdgcn_global_load_lds((global_void_t *)u, (lds_void_t *)(dbuf + RPI * NW * j * KB), 16, 0, 0);
    };

    // fragment reads: rows wm0 + 32 s + q (A) and 256 + wn0 + 32 s + q (B),
    // chunk 2 ks + g of the lane's k-step, swizzled by the row (32-row blocks
    // and the 256-row B offset keep the swizzle of q)
    const int q = lane & 31, g = lane >> 5, qx = swz(q);
    int ko[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) ko[ks] = ((2 * ks + g) ^ qx) << 4;
    const int aoff = (wm0 + q) * KB, boff = (BM + wn0 + q) * KB;
    i4_t fa[2][4], fb[2][4];
    // fragment f of k-step ks (f < 4: A block f, else B block f - 4)
    auto read_one = [&](const unsigned char *stg, int ks, int f, i4_t *A4, i4_t *B4) {
        if (f < 4) A4[f] = *reinterpret_cast<const i4_t *>(stg + aoff + 32 * f * KB + ko[ks]);
        else B4[f - 4] = *reinterpret_cast<const i4_t *>(stg + boff + 32 * (f - 4) * KB + ko[ks]);
    };
    i16_t acc[4][4];
    auto mfma = [&](int i, const i4_t *A4, const i4_t *B4) {
        acc[i >> 2][i & 3] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A4[i >> 2], B4[i & 3], acc[i >> 2][i & 3], 0, 0, 0);
    };

    if (total > 0) {
        const int pre = min(NS, total);
        for (int i = 0; i < pre; ++i) {
            dma_setup(i);
#pragma unroll
            for (int j = 0; j < LPW; ++j) dma_one(j);
        }
        wait_vmem_at_most(LPW * (pre - 1));   // this wave's part of stage 0 has landed
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
#pragma unroll
        for (int f = 0; f < 8; ++f) read_one(smem, 0, f, fa[0], fb[0]);
    }
    // One k-step: its 16 matrix instructions, each followed by at most one
    // fragment read for the next k-step and at most one direct load, in that
    // order (the scheduler may not move anything across the groups): the
    // reads and loads issue while the matrix pipe is busy instead of holding
    // it up in a block.
    int st = 0;
    for (int l = 0; l < nmod; ++l) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[a][b][e] = 0;
        for (int kt = 0; kt < nkt; ++kt, ++st) {
            const unsigned char *cur = smem + (st % NS) * STAGE;
            // stage st+NS's loads are issued in PARTS parts behind matrix
            // instructions: part 0 behind this stage's last k-step, part p
            // behind k-step p-1 of the next stage
            constexpr int PART = LPW / PARTS;
            const bool later_parts = PARTS > 1 && st >= 1 && st + 1 < total;
#pragma unroll
            for (int ks = 0; ks < KS - 1; ++ks) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    mfma(i, fa[ks & 1], fb[ks & 1]);
                    if (i < 8) read_one(cur, ks + 1, i, fa[(ks + 1) & 1], fb[(ks + 1) & 1]);
                    if (PARTS > 1 && ks < PARTS - 1 && i < PART && later_parts) dma_one((ks + 1) * PART + i);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            // last k-step: the stage hand-over runs behind its matrix instructions
            constexpr int L = KS - 1;
            mfma(0, fa[L & 1], fb[L & 1]);
            mfma(1, fa[L & 1], fb[L & 1]);
            __builtin_amdgcn_sched_barrier(0);
            if (st + 1 < total) {
                // stages st+2 .. min(st+NS-1, total-1) may stay in flight
                const int later = min(st + NS - 1, total - 1) - (st + 1);
                wait_vmem_at_most(LPW * later);
            }
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
            const bool issue = st + NS < total, next = st + 1 < total;
            if (issue) dma_setup(st + NS);
            const unsigned char *nxt = smem + ((st + 1) % NS) * STAGE;
            constexpr int NLAST = PART;   // loads behind this k-step
#pragma unroll
            for (int i = 2; i < 16; ++i) {
                mfma(i, fa[L & 1], fb[L & 1]);
                if (i - 2 < 8 && next) read_one(nxt, 0, i - 2, fa[0], fb[0]);
                if constexpr (NLAST <= 14) {
                    if (i - 2 < NLAST && issue) dma_one(i - 2);
                } else {
                    // more loads than matrix instructions: two per instruction
                    if (2 * (i - 2) < NLAST && issue) dma_one(2 * (i - 2));
                    if (2 * (i - 2) + 1 < NLAST && issue) dma_one(2 * (i - 2) + 1);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // residues of modulus l straight from the accumulators
        const int m = p.cc->m[l];
        const double md = (double)m, im = p.cc->invm[l];
#pragma unroll
        for (int ab = 0; ab < 16; ++ab) {
            const i16_t v = acc[ab >> 2][ab & 3];
            i4_t w;
#pragma unroll
            for (int e4 = 0; e4 < 4; ++e4) {
                unsigned word = 0;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int x = v[4 * e4 + e];
                    int r = x;   // m = 256: the low byte is the residue
                    if (m != 256) {
                        const double xd = (double)x;
                        r = (int)__builtin_fma(-__builtin_rint(xd * im), md, xd);   // exact: |x| < 2^31
                    }
                    word |= (unsigned)(r & 0xFF) << (8 * e);
                }
                w[e4] = (int)word;
            }
            Rv[((int64_t)l * 16 + ab) * 64] = w;
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// CRT reconstruction after ozaki_crt_wide_kernel, one thread per (wave,
// 32x32 tile, lane) of the product kernel: the thread loads that lane's
// residues of the tile under every modulus at once (16 bytes per modulus,
// 1 KB per modulus across a wave) and rebuilds its 16 elements of V.  Run as
// its own launch it has the whole chip and many waves in flight for what is
// a memory-bound pass; inside the product kernel it ran with one wave per
// SIMD and nothing to hide its loads behind.
//
// A workgroup rebuilds a 32 x 128 strip of C (wave b: columns 32b..32b+31).
// A lane's 16 elements lie in one column, so writing them straight to
// column-major C put each store instruction's 64 lanes on 32 columns (8-byte
// pieces of 64 cache lines; the launch wrote four times C's size to HBM).
// The strip passes through LDS instead and each wave then writes two
// columns' 32 contiguous rows per instruction; the exponents, alpha and beta
// are applied in that pass (one row exponent per lane).
__global__ __launch_bounds__(256) void ozaki_crt_rebuild_kernel(CrtParams p) {
    constexpr int SP = 33;   // strip pitch in doubles (column-major, odd: conflict-free row reads)
    __shared__ double S[128 * SP];
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = (int)(t & 63), ab = (int)((t >> 6) & 15);
    const int64_t wv = t >> 10;
    const int wave = (int)(wv & 3), tile = (int)(wv >> 2);
    const int tm = tile / p.tiles_n, tn = tile - tm * p.tiles_n;
    const int m0 = tm * 256 + (wave & 1) * 128, n0 = tn * 256 + (wave >> 1) * 128;
    const int a = ab >> 2, b = ab & 3, q = lane & 31, g = lane >> 5;
    const int nmod = p.cc->nmod;
    const i4_t *Rv = reinterpret_cast<const i4_t *>(p.Rv) + wv * ((int64_t)nmod * 16 * 64) + ab * 64 + lane;
    i4_t R[OZ_CRT_MAXMOD];
#pragma unroll
    for (int l = 0; l < OZ_CRT_MAXMOD; ++l) R[l] = l < nmod ? Rv[(int64_t)l * 16 * 64] : i4_t{0, 0, 0, 0};
    const double two32 = 4294967296.0;
    double *col = S + (32 * b + q) * SP;
#pragma unroll
    for (int e4 = 0; e4 < 4; ++e4) {
        double s[4][4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int k = 0; k < 4; ++k) s[e][k] = 0.0;
        // moduli past nmod have zero residues and zero weights
#pragma unroll
        for (int l = 0; l < OZ_CRT_MAXMOD; ++l) {
            const unsigned word = (unsigned)R[l][e4];
            const double w0 = p.cc->w[l][0], w1 = p.cc->w[l][1], w2 = p.cc->w[l][2], w3 = p.cc->w[l][3];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double r = (double)(int)(signed char)(word >> (8 * e));
                s[e][0] = fma(r, w0, s[e][0]);
                s[e][1] = fma(r, w1, s[e][1]);
                s[e][2] = fma(r, w2, s[e][2]);
                s[e][3] = fma(r, w3, s[e][3]);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int ee = 4 * e4 + e;
            const double approx = (s[e][3] * two32 + s[e][2]) * (two32 * two32) + s[e][1] * two32;
            const double qq = rint(approx * p.cc->invM);
            long long t0 = (long long)fma(-qq, p.cc->Mk[0], s[e][0]);
            long long t1 = (long long)fma(-qq, p.cc->Mk[1], s[e][1]);
            long long t2 = (long long)fma(-qq, p.cc->Mk[2], s[e][2]);
            long long t3 = (long long)fma(-qq, p.cc->Mk[3], s[e][3]);
            // carries: t0, t1, t2 into [0, 2^32) (floor division by 2^32)
            constexpr long long B32 = 4294967296ll;
            long long c = t0 >> 32;
            t0 -= c * B32;
            t1 += c;
            c = t1 >> 32;
            t1 -= c * B32;
            t2 += c;
            c = t2 >> 32;
            t2 -= c * B32;
            t3 += c;
            // V = hi 2^64 + lo with lo taken signed, so small |V| loses nothing
            long long hi = t3 * B32 + t2;
            const unsigned long long ulo = ((unsigned long long)t1 << 32) | (unsigned long long)t0;
            const long long lo = (long long)ulo;
            if (ulo >> 63) hi += 1;
            col[(ee >> 2) * 8 + g * 4 + (ee & 3)] = fma((double)hi, two32 * two32, (double)lo);
        }
    }
    __syncthreads();
    // write pass: wave w, round u: columns 8u + 2w + g, row q
    const int i = m0 + 32 * a + q;
    const int wq = threadIdx.x >> 6;
    if (i < p.M) {
        const int ea = p.ea[i];
#pragma unroll 4
        for (int u = 0; u < 16; ++u) {
            const int cl = 8 * u + 2 * wq + g, j = n0 + cl;
            if (j < p.N) {
                const double v = p.alpha * ldexp(S[cl * SP + q], ea + p.eb[j]);
                double *dst = p.C + i + (int64_t)j * p.ldc;
                *dst = (p.beta == 0.0) ? v : v + p.beta * *dst;
            }
        }
    }
}

// ---------------------------------------------------------------- host side of the CRT form
static const int kCrtModuli[OZ_CRT_MAXMOD] = {256, 255, 253, 251, 247, 241, 239, 233,
                                              229, 227, 223, 217, 211, 199, 197, 193};

static int inv_mod(int a, int m) {   // a^-1 mod m (gcd(a, m) = 1)
    int t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
    while (nr) {
        const int qd = r / nr;
        int tmp = t - qd * nt;
        t = nt;
        nt = tmp;
        tmp = r - qd * nr;
        r = nr;
        nr = tmp;
    }
    return t < 0 ? t + m : t;
}

void ozaki_crt_constants(int nmod, CrtConstHost *out) {
    unsigned __int128 M = 1;
    for (int l = 0; l < nmod; ++l) M *= (unsigned)kCrtModuli[l];
    out->nmod = nmod;
    out->log2M = 0.0;
    for (int l = 0; l < nmod; ++l) out->log2M += std::log2((double)kCrtModuli[l]);
    for (int l = 0; l < nmod; ++l) {
        const int m = kCrtModuli[l];
        const unsigned __int128 Ml = M / (unsigned)m;
        const int inv = inv_mod((int)(Ml % (unsigned)m), m);
        const unsigned __int128 W = Ml * (unsigned)inv;   // < M: no reduction needed
        out->m[l] = m;
        for (int k = 0; k < 4; ++k) out->w[l][k] = (double)(uint64_t)((W >> (32 * k)) & 0xFFFFFFFFu);
    }
    for (int k = 0; k < 4; ++k) out->Mk[k] = (double)(uint64_t)((M >> (32 * k)) & 0xFFFFFFFFu);
    out->invM = 1.0 / (double)M;
}

int ozaki_crt_bits(int nmod, int k) {
    CrtConstHost h;
    ozaki_crt_constants(nmod, &h);
    const int lk = (int)std::ceil(std::log2((double)(k > 1 ? k : 1)));
    int al = (int)std::floor((h.log2M - 2.0 - lk) / 2.0);
    return al > 56 ? 56 : al;   // residue arithmetic of oz_residue_kernel is exact up to 2^56
}

// Device copy of the constants per modulus count (built once, synchronously).
static const CrtConst *crt_device_constants(int nmod, int *status) {
    static std::mutex mu;
    static const CrtConst *table[OZ_CRT_MAXMOD + 1] = {};
    std::lock_guard<std::mutex> lk(mu);
    if (table[nmod]) return table[nmod];
    CrtConstHost h;
    ozaki_crt_constants(nmod, &h);
    CrtConst c{};
    c.nmod = nmod;
    for (int l = 0; l < nmod; ++l) {
        c.m[l] = h.m[l];
        c.invm[l] = 1.0 / h.m[l];
        for (int k = 0; k < 4; ++k) c.w[l][k] = h.w[l][k];
    }
    for (int k = 0; k < 4; ++k) c.Mk[k] = h.Mk[k];
    c.invM = h.invM;
    for (int l = 0; l < nmod; ++l) c.c32[l] = (double)((1ull << 32) % (unsigned long long)h.m[l]);
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(CrtConst));   // kept for the life of the process
    if (e == hipSuccess) e = hipMemcpy(d, &c, sizeof(CrtConst), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        *status = (int)e;
        return nullptr;
    }
    table[nmod] = static_cast<const CrtConst *>(d);
    return table[nmod];
}

// Row pitch of the residue planes: k padded to 128 bytes, plus 128 bytes when
// that is a multiple of 512, so that the rows a stage loads (one pitch apart)
// do not all start at the same offset modulo a large power of two.
static int crt_pitch(int Kp) { return Kp % 512 == 0 ? Kp + 128 : Kp; }

size_t ozaki_crt_workspace_bytes(int rows, int k, int nmod) {
    const size_t rp = (size_t)ozaki_pad_rows(rows), kp = (size_t)crt_pitch(ozaki_pad_k(k));
    return (size_t)nmod * rp * kp + 256 + rp * sizeof(unsigned long long) + rp * sizeof(int) + 256;
}

size_t ozaki_crt_accum_bytes(int m, int n, int nmod) {
    const size_t mn = (size_t)ozaki_pad_rows(m) * (size_t)ozaki_pad_rows(n);
    return mn * 8 + mn * (size_t)nmod + 256;
}

static int crt_split(const double *X, int64_t rs, int64_t ks, int R, int K, int al, const CrtConst *cc, int nmod,
                     void *ws, int8_t **res, int **exps, hipStream_t s) {
    const int Rp = ozaki_pad_rows(R), Kp = ozaki_pad_k(K), pitch = crt_pitch(Kp);
    unsigned char *base = static_cast<unsigned char *>(ws);
    *res = reinterpret_cast<int8_t *>(base);
    size_t off = ((size_t)nmod * Rp * pitch + 255) & ~(size_t)255;
    unsigned long long *rowmax = reinterpret_cast<unsigned long long *>(base + off);
    off += (size_t)Rp * sizeof(unsigned long long);
    *exps = reinterpret_cast<int *>(base + off);
    hipLaunchKernelGGL(oz_zero_u64_kernel, dim3(grid1(Rp)), dim3(256), 0, s, rowmax, Rp);
    if (K > 0 && R > 0)
        hipLaunchKernelGGL(oz_rowmax_kernel, dim3((K + 63) / 64, (R + 63) / 64), dim3(256), 0, s, X, rs, ks, R, K,
                           rowmax);
    hipLaunchKernelGGL(oz_residue_kernel, dim3(Kp / 64, Rp / 64), dim3(256), 0, s, X, rs, ks, R, K, Rp, Kp, pitch,
                       al, rowmax, cc, *res, *exps);
    return (int)hipGetLastError();
}

// Tile layouts of the CRT kernel share one workspace layout: the int32 staging
// and the residues are per output element (tile area), whatever the wave split.
template <int WM, int WN, int TM, int TN, int KB, bool DIRECT = false, int NS = 2>
static int launch_crt_variant(CrtParams &p, hipStream_t s) {
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    static_assert(BM == 256 && BN == 256, "rows are padded for 256-row tiles");
    p.tiles_m = (p.M + BM - 1) / BM;
    p.tiles_n = (p.N + BN - 1) / BN;
    constexpr size_t lds = (size_t)NS * (BM + BN) * (DIRECT ? KB : KB + 16);
    auto kern = ozaki_crt_kernel<WM, WN, TM, TN, KB, DIRECT, NS>;
    if (hipError_t he = lds_limit_once<ozaki_crt_kernel<WM, WN, TM, TN, KB, DIRECT, NS>>(lds)) return (int)he;
    hipLaunchKernelGGL(kern, dim3(p.tiles_m * p.tiles_n), dim3(WM * WN * 64), lds, s, p);
    return (int)hipGetLastError();
}

template <int KB, int NS, int PARTS = 1>
static int launch_crt_wide(CrtParams &p, hipStream_t s) {
    if (p.Kp > OZ_CRT_CHUNK_K) return launch_crt_variant<2, 4, 4, 2, 128, true>(p, s);   // K over one exact chunk
    p.tiles_m = (p.M + 255) / 256;
    p.tiles_n = (p.N + 255) / 256;
    constexpr size_t lds = (size_t)NS * 512 * KB;
    auto kern = ozaki_crt_wide_kernel<KB, NS, PARTS>;
    if (hipError_t he = lds_limit_once<ozaki_crt_wide_kernel<KB, NS, PARTS>>(lds)) return (int)he;
    hipLaunchKernelGGL(kern, dim3(p.tiles_m * p.tiles_n), dim3(256), lds, s, p);
    if (hipError_t he = hipGetLastError()) return (int)he;
    // 4 waves x 16 tiles x 64 lanes per 256x256 tile, 256 threads per block
    hipLaunchKernelGGL(ozaki_crt_rebuild_kernel, dim3(p.tiles_m * p.tiles_n * 16), dim3(256), 0, s, p);
    return (int)hipGetLastError();
}

int launch_dgemm_ozaki_crt(bool ta, bool tb, int m, int n, int k, double alpha, const double *A, int64_t lda,
                           const double *B, int64_t ldb, double beta, double *C, int64_t ldc, int nmod, void *wsA,
                           void *wsB, void *wsW, hipStream_t s) {
    if (m <= 0 || n <= 0 || k <= 0) return 0;
    int st = 0;
    const CrtConst *cc = crt_device_constants(nmod, &st);
    if (!cc) return st;
    const int al = ozaki_crt_bits(nmod, k);
    int8_t *ra, *rb;
    int *ea, *eb;
    int e = crt_split(A, ta ? lda : 1, ta ? 1 : lda, m, k, al, cc, nmod, wsA, &ra, &ea, s);
    if (e) return e;
    e = crt_split(B, tb ? 1 : ldb, tb ? ldb : 1, n, k, al, cc, nmod, wsB, &rb, &eb, s);
    if (e) return e;
    CrtParams p;
    p.M = m;
    p.N = n;
    p.Kp = ozaki_pad_k(k);
    p.pitch = crt_pitch(p.Kp);
    p.planeA = (int64_t)ozaki_pad_rows(m) * p.pitch;
    p.planeB = (int64_t)ozaki_pad_rows(n) * p.pitch;
    p.RA = ra;
    p.RB = rb;
    p.ea = ea;
    p.eb = eb;
    p.cc = cc;
    const size_t mn = (size_t)ozaki_pad_rows(m) * (size_t)ozaki_pad_rows(n);
    p.Wi = static_cast<int *>(wsW);
    p.Wc = p.Wi + mn;
    p.Rv = reinterpret_cast<unsigned *>(static_cast<char *>(wsW) + mn * 8);
    p.alpha = alpha;
    p.beta = beta;
    p.C = C;
    p.ldc = ldc;
    static const char *var = getenv("GEMM_AMD_CRT_VARIANT");
    const int v = var ? atoi(var) : 0;
    // Default: 256x256 tiles of 4 waves (128x128 each), two 128-byte direct
    // stages, each stage's loads issued in two halves behind matrix
    // instructions, reconstruction in its own launch (at 8192^3, 16 moduli,
    // whole call: 113.9 TFLOP/s FP64-equivalent against 103.4 for the 8-wave
    // form).  The others stay selectable for comparison (same results):
    if (v == 1) return launch_crt_variant<2, 4, 4, 2, 128>(p, s);   // 8 waves, register staging
    if (v == 2) return launch_crt_variant<2, 4, 4, 2, 64>(p, s);    // 8 waves, register staging, 64-deep stages
    if (v == 3) return launch_crt_variant<2, 4, 4, 2, 128, true>(p, s);   // 8 waves, direct, 2 stages
    if (v == 4) return launch_crt_variant<2, 4, 4, 2, 64, true, 4>(p, s);   // 8 waves, direct, 4 stages of 64 bytes
    if (v == 5) return launch_crt_wide<128, 2>(p, s);       // 4 waves, loads issued in one block
    if (v == 6) return launch_crt_wide<64, 4>(p, s);        // 4 waves, four 64-byte stages
    return launch_crt_wide<128, 2, 2>(p, s);
}

int launch_ozaki_crt_split_only(const double *X, int64_t rs, int64_t ks, int R, int K, int nmod, int al,
                                int8_t *res, int Rp, int Kp, int *exps, unsigned long long *rowmax_ws, hipStream_t s) {
    if (R <= 0 || K <= 0) return 0;
    int st = 0;
    const CrtConst *cc = crt_device_constants(nmod, &st);
    if (!cc) return st;
    hipLaunchKernelGGL(oz_zero_u64_kernel, dim3(grid1(Rp)), dim3(256), 0, s, rowmax_ws, Rp);
    hipLaunchKernelGGL(oz_rowmax_kernel, dim3((K + 63) / 64, (R + 63) / 64), dim3(256), 0, s, X, rs, ks, R, K,
                       rowmax_ws);
    hipLaunchKernelGGL(oz_residue_kernel, dim3(Kp / 64, Rp / 64), dim3(256), 0, s, X, rs, ks, R, K, Rp, Kp, Kp, al,
                       rowmax_ws, cc, res, exps);
    return (int)hipGetLastError();
}

}  // namespace gemm_amd
