// SGEMM from bf16 splits on the bf16 matrix cores ("bf16x3" / "bf16x6").
//
// No reference counterpart: the reference's SGEMM is FP32 FMA
// (src/kernel/gemm_kernel.S:11-23).  Method (Henry, Tang, Heinecke,
// "Leveraging the bfloat16 artificial intelligence datatype for higher-precision
// computations", ARITH 2019): every FP32 element is split exactly into bf16
// parts x = x0 + x1 (+ x2), x_i = bf16_rne(x - x0 - ... - x_{i-1}) (the
// differences are exact in FP32), and C = alpha * sum over the selected part
// products A_i B_j + beta*C with FP32 accumulation:
//   3 products (2-way split): A0B0 + A0B1 + A1B0        (~16 significant bits)
//   6 products (3-way split): + A0B2 + A2B0 + A1B1      (~24 bits, FP32 level)
// The products share one FP32 accumulator: the part planes are laid out so the
// GEMM runs over the K-concatenation [A_i1 | A_i2 | ...] x [B_j1; B_j2; ...],
// smallest products first.  Planes are bf16, k contiguous, rows padded to the
// 256-row tile and k to the 64-deep stage, zero filled.
#include "common.hpp"
#include "bf16x.hpp"

namespace gemm_amd {

static constexpr int BX_ROWPAD = 256;
static constexpr int BX_KPAD = 64;

int bf16x_pad_rows(int x) { return (x + BX_ROWPAD - 1) / BX_ROWPAD * BX_ROWPAD; }
int bf16x_pad_k(int x) { return (x + BX_KPAD - 1) / BX_KPAD * BX_KPAD; }

size_t bf16x_workspace_bytes(int rows, int k, int nsplit) {
    return (size_t)nsplit * bf16x_pad_rows(rows) * bf16x_pad_k(k) * 2;
}

// FP32 -> bf16, round to nearest even (NaN stays a quiet NaN)
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Rows of the R x K view of X (element (r, k) at X[r*rs + k*ks]) -> planes
// [part][Rp][Kp] of bf16; grid (Kp/64, Rp/64).
__global__ __launch_bounds__(256) void bf16_split_kernel(const float *__restrict__ X, int64_t rs, int64_t ks, int R,
                                                         int K, int Rp, int Kp, int nsplit,
                                                         uint16_t *__restrict__ out) {
    __shared__ float T[64][65];
    const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64, tid = threadIdx.x;
    for (int idx = tid; idx < 64 * 64; idx += 256) {
        int rr, kk;
        if (rs == 1) { rr = idx & 63; kk = idx >> 6; }
        else { kk = idx & 63; rr = idx >> 6; }
        const int r = r0 + rr, k = k0 + kk;
        T[rr][kk] = (r < R && k < K) ? X[(int64_t)r * rs + (int64_t)k * ks] : 0.0f;
    }
    __syncthreads();
    const int rr = tid >> 2, kq = (tid & 3) * 16;
    float rem[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) rem[i] = T[rr][kq + i];
    const int64_t plane = (int64_t)Rp * Kp;
    uint16_t *dst = out + (int64_t)(r0 + rr) * Kp + k0 + kq;
    for (int s = 0; s < nsplit; ++s) {
        i4_t w0 = {0, 0, 0, 0}, w1 = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint16_t h = f32_to_bf16_rne(rem[i]);
            rem[i] = rem[i] - bf16_to_f32(h);
            if (i < 8) w0[i >> 1] |= (int)h << (16 * (i & 1));
            else w1[(i - 8) >> 1] |= (int)h << (16 * (i & 1));
        }
        *reinterpret_cast<i4_t *>(dst + s * plane) = w0;
        *reinterpret_cast<i4_t *>(dst + s * plane + 8) = w1;
    }
}

struct BxParams {
    int M, N, Kp, nprod, tiles_m, tiles_n;
    uint64_t plist;            // product p: A part (plist >> 8p) & 3, B part (plist >> (8p+4)) & 3
    int64_t planeA, planeB;    // elements per plane
    const uint16_t *PA, *PB;
    float alpha, beta;
    float *C;
    int64_t ldc;
};

// Workgroup = BM x BN tile of WM x WN waves, each (32*TM) x (32*TN) of
// v_mfma_f32_32x32x16_bf16 tiles; stage = KB bytes (KB/2 k values) of every row.
template <int WM, int WN, int TM, int TN, int KB>
__global__ __launch_bounds__(WM *WN * 64) void bf16x_gemm_kernel(BxParams p) {
    constexpr int KP = KB + 16;
    constexpr int NT = WM * WN * 64;
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    constexpr int STAGE = (BM + BN) * KP;
    constexpr int CH = KB / 16;
    constexpr int PA = BM * CH / NT, PB = BN * CH / NT;
    static_assert(BM * CH % NT == 0 && BN * CH % NT == 0, "stage must split evenly over the workgroup");
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WM) * (TM * 32), wn0 = (wave / WM) * (TN * 32);
    int tm, tn;
    tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, 8, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int kel = KB / 2;                       // k values per stage
    const int nkt = p.Kp / kel;
    const int total = p.nprod * nkt;

    i4_t ra[PA], rb[PB];
    auto gload = [&](int st) {
        const int pr = st / nkt, kt = st - pr * nkt;
        const int ia = (p.plist >> (8 * pr)) & 3, ib = (p.plist >> (8 * pr + 4)) & 3;
        const unsigned char *a = reinterpret_cast<const unsigned char *>(p.PA + ia * p.planeA + (int64_t)m0 * p.Kp +
                                                                         kt * kel);
        const unsigned char *b = reinterpret_cast<const unsigned char *>(p.PB + ib * p.planeB + (int64_t)n0 * p.Kp +
                                                                         kt * kel);
        const int64_t pitch = (int64_t)p.Kp * 2;
#pragma unroll
        for (int i = 0; i < PA; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            ra[i] = *reinterpret_cast<const i4_t *>(a + row * pitch + ch);
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            const int v = tid + i * NT, row = v / CH, ch = (v % CH) * 16;
            rb[i] = *reinterpret_cast<const i4_t *>(b + row * pitch + ch);
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
    f16_t acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.0f;

    if (total > 0) {
        gload(0);
        sstore(smem);
    }
    __syncthreads();
    for (int st = 0; st < total; ++st) {
        const unsigned char *cur = smem + (st & 1) * STAGE;
        unsigned char *nxt = smem + ((st + 1) & 1) * STAGE;
        if (st + 1 < total) gload(st + 1);
        // 16 k values (32 bytes) per MFMA step; the fragments of step ks + 1
        // are read while the matrix instructions of step ks run (the sched
        // barriers keep the scheduler from sinking each read next to its use,
        // which stalls the matrix pipe on LDS latency)
        bf8_t av[2][TM], bv[2][TN];
        auto frag = [&](int ks, bf8_t *a, bf8_t *b) {
#pragma unroll
            for (int s = 0; s < TM; ++s) a[s] = *reinterpret_cast<const bf8_t *>(cur + aoff + 32 * s * KP + 32 * ks);
#pragma unroll
            for (int s = 0; s < TN; ++s) b[s] = *reinterpret_cast<const bf8_t *>(cur + boff + 32 * s * KP + 32 * ks);
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
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bv[ks & 1][b], av[ks & 1][a], acc[a][b], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (st + 1 < total) sstore(nxt);
        __syncthreads();
    }

    // B's fragment is the matrix instructions' first operand, so lane q holds
    // row q of each 32x32 block: one element index across the wave writes 32
    // consecutive rows of two columns (whole 128-byte runs of column-major C).
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int i = m0 + wm0 + 32 * a + q;
                const int j = n0 + wn0 + 32 * b + (e >> 2) * 8 + g * 4 + (e & 3);
                if (i < p.M && j < p.N) {
                    float *dst = p.C + i + (int64_t)j * p.ldc;
                    const float v = p.alpha * acc[a][b][e];
                    *dst = (p.beta == 0.0f) ? v : v + p.beta * *dst;
                }
            }
}

static int bx_grid(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

int launch_bf16_split(const float *X, int64_t rs, int64_t ks, int R, int K, int nsplit, uint16_t *planes, int Rp,
                      int Kp, hipStream_t s) {
    if (Rp <= 0 || Kp <= 0) return 0;
    hipLaunchKernelGGL(bf16_split_kernel, dim3(Kp / 64, Rp / 64), dim3(256), 0, s, X, rs, ks, R, K, Rp, Kp, nsplit,
                       planes);
    return (int)hipGetLastError();
}

// product lists (A part, B part), smallest first
static uint64_t plist_for(int nprod) {
    // 3: (1,0) (0,1) (0,0); 6: (2,0) (0,2) (1,1) (1,0) (0,1) (0,0)
    const int p3[3][2] = {{1, 0}, {0, 1}, {0, 0}};
    const int p6[6][2] = {{2, 0}, {0, 2}, {1, 1}, {1, 0}, {0, 1}, {0, 0}};
    uint64_t v = 0;
    for (int i = 0; i < nprod; ++i) {
        const int *pp = nprod == 3 ? p3[i] : p6[i];
        v |= (uint64_t)(pp[0] | (pp[1] << 4)) << (8 * i);
    }
    return v;
}

int launch_sgemm_bf16x(bool ta, bool tb, int m, int n, int k, float alpha, const float *A, int64_t lda,
                       const float *B, int64_t ldb, float beta, float *C, int64_t ldc, int nprod, void *wsA, void *wsB,
                       hipStream_t s) {
    if (m <= 0 || n <= 0 || k <= 0) return 0;
    const int nsplit = nprod == 3 ? 2 : 3;
    const int Mp = bf16x_pad_rows(m), Np = bf16x_pad_rows(n), Kp = bf16x_pad_k(k);
    uint16_t *pa = static_cast<uint16_t *>(wsA), *pb = static_cast<uint16_t *>(wsB);
    int e = launch_bf16_split(A, ta ? lda : 1, ta ? 1 : lda, m, k, nsplit, pa, Mp, Kp, s);
    if (e) return e;
    e = launch_bf16_split(B, tb ? 1 : ldb, tb ? ldb : 1, n, k, nsplit, pb, Np, Kp, s);
    if (e) return e;
    constexpr int WM = 2, WN = 4, TM = 4, TN = 2, KB = 128;
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    BxParams p;
    p.M = m;
    p.N = n;
    p.Kp = Kp;
    p.nprod = nprod;
    p.plist = plist_for(nprod);
    p.tiles_m = Mp / BM;
    p.tiles_n = Np / BN;
    p.planeA = (int64_t)Mp * Kp;
    p.planeB = (int64_t)Np * Kp;
    p.PA = pa;
    p.PB = pb;
    p.alpha = alpha;
    p.beta = beta;
    p.C = C;
    p.ldc = ldc;
    constexpr size_t lds = 2ull * (BM + BN) * (KB + 16);
    auto kern = bf16x_gemm_kernel<WM, WN, TM, TN, KB>;
    if (hipError_t he = lds_limit_once<bf16x_gemm_kernel<WM, WN, TM, TN, KB>>(lds)) return (int)he;
    hipLaunchKernelGGL(kern, dim3(p.tiles_m * p.tiles_n), dim3(WM * WN * 64), lds, s, p);
    return (int)hipGetLastError();
}

}  // namespace gemm_amd
