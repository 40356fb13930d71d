// FP64 / FP32 GEMM on gfx950 matrix cores.
//
// Replaces the reference's blocked CPU GEMM (src/gemm_driver.c:116-195, its
// packing in src/gemm_copy.c and the AVX-512 micro-kernels in
// src/kernel/gemm_kernel.S:818-1213) with one kernel per (dtype, transA,
// transB, tile shape):
//   * one workgroup owns a BM x BN macro-tile of C (the reference's per-thread
//     N chunk + BLOCK_DIM_M x BLOCK_DIM_N blocks become the grid);
//   * op(A) and op(B) K-slabs are staged global -> registers -> LDS, double
//     buffered, and transposed on the way in when the operand is stored
//     M/N-contiguous (the reference's pack step, src/gemm_copy.c:4-199, done
//     in-kernel); both LDS tiles are [row][k] with a 16-byte row pad so the
//     16-byte fragment reads of 8 (16) consecutive rows hit distinct banks;
//   * waves run 64x64 (or 32x32) register tiles of MFMA instructions;
//   * the epilogue applies alpha and beta once per element (beta == 0 does not
//     read C), the reference's C_MULT_BETA (src/gemm_driver.c:104-114) and
//     alpha-folding into B (src/gemm_copy.c:64) fused;
//   * split-K writes per-partition partial sums to a workspace and a second
//     pass reduces them in a fixed order (deterministic).
#include "common.hpp"
#include "gemm_kernels.hpp"
#include "gemm_tile.hpp"
#include "runtime.hpp"

namespace gemm_amd {


// ---------------------------------------------------------------- kernel
// FEPI (FP64 only): tiles wholly inside C store through store_full (no range
// checks, no 64-bit multiply per element).  It is a separate instance, taken
// for shallow K where the epilogue is a visible part of a tile, because the
// extra path changes how the compiler lays out the main loop of the others
// (the 8192^3 launch measured 2% slower with it).  Both compute the same bits.
//
// KW > 1: KW groups of WM x WN waves share each BK-deep slab, group g
// multiplying its k = g*BK/KW .. (g+1)*BK/KW - 1; at the end groups 1..KW-1
// pass their accumulators through LDS and group 0 adds them in group order
// and stores.  It is for grids of at most one workgroup per CU, where a
// 4-wave workgroup leaves each SIMD one wave to hide its latencies with.
template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WM, int WN, bool VEC, bool FEPI = false,
          int KW = 1>
// FP64 wave tiles of up to 64x32 fit two waves per SIMD without spilling;
// 64x64 ones need the whole register file of one.
__global__ __launch_bounds__(WM *WN *KW * 64, (sizeof(T) == 8 && KW == 1 && (BM / WM) * (BN / WN) <= 2048) ? 2 : 1)
void gemm_mfma_kernel(GemmParams<T> p) {
    static_assert(KW == 1 || (sizeof(T) != 2 && BK % (8 * KW) == 0), "k-grouped waves: FP32 and FP64");
    constexpr int NT = WM * WN * KW * 64;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    using Lay = SlabLayout<T, TA, TB, BM, BN, BK>;
    using LA = TileLoader<T, BM, BK, NT, TA, VEC, Lay::KMA>;   // op(A) rows = M; k-contiguous iff A is transposed
    using LB = TileLoader<T, BN, BK, NT, !TB, VEC, Lay::KMB>;  // op(B) rows = N; k-contiguous iff B is not transposed
    using WaveKW = std::conditional_t<sizeof(T) == 8,
                                      WaveF64<WTM, WTN, BK / KW, Lay::KMA ? BM + 4 : 0, Lay::KMB ? BN + 4 : 0, Lay::PITCH>,
                                      WaveF32<WTM, WTN, BK / KW, Lay::PITCH>>;
    using Wave = std::conditional_t<KW == 1, typename Lay::template Wave<WTM, WTN>, WaveKW>;
    constexpr int STAGE = Lay::STAGE;
    constexpr int BOFF = LA::SIZE;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wmn = KW == 1 ? wave : wave % (WM * WN), kg = KW == 1 ? 0 : wave / (WM * WN);
    const int wm0 = (wmn % WM) * WTM, wn0 = (wmn / WM) * WTN;

    int tm, tn;
    tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, 8, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int kbeg = blockIdx.y * p.kchunk;
    const int kend = min(p.K, kbeg + p.kchunk);
    const int mlim = p.M - m0, nlim = p.N - n0;

    const T *abase = TA ? p.A + (int64_t)m0 * p.lda + kbeg : p.A + m0 + (int64_t)kbeg * p.lda;
    const T *bbase = TB ? p.B + n0 + (int64_t)kbeg * p.ldb : p.B + (int64_t)n0 * p.ldb + kbeg;
    const int64_t astep = TA ? BK : (int64_t)BK * p.lda;
    const int64_t bstep = TB ? (int64_t)BK * p.ldb : BK;

    LA la;
    LB lb;
    Wave w;
    if constexpr (KW == 1) w.init(lane, wm0, wn0);
    else w.init(lane, wm0, wn0, kg * (BK / KW));

    const int nk = (kend - kbeg + BK - 1) / BK;
    // Interior tiles (the bulk of any large problem) stream whole aligned
    // slabs with no per-element range checks.
    const bool interior = VEC && mlim >= BM && nlim >= BN && (kend - kbeg) % BK == 0;
    if (nk > 0) {
        if (interior) {
            la.load_full(abase, p.lda, tid);
            lb.load_full(bbase, p.ldb, tid);
        } else {
            la.load(abase, p.lda, mlim, kend - kbeg, tid);
            lb.load(bbase, p.ldb, nlim, kend - kbeg, tid);
        }
        la.store(smem, tid);
        lb.store(smem + BOFF, tid);
    }
    __syncthreads();
    // Interior slabs are fetched two ahead where a second register set fits
    // (staging of at most 160 bytes per thread: all shapes but the FP64
    // 128x128 one, which would spill): while slab kt is multiplied out of
    // LDS, slab kt+1 waits in one set for its store and slab kt+2 is in
    // flight into the other.  The loads go through a buffer resource based at
    // the slab, at loop-invariant per-lane offsets, so the loop rewrites no
    // address register an outstanding load still reads (which would make the
    // compiler wait for that load); the barrier is fenced on LDS only
    // (__syncthreads would wait for the slab in flight); the loop has one
    // exit and the stores are unconditional, which keeps the wait counts
    // exact.  Same arithmetic as the one-ahead loop, so results are identical.
    // (FP32 256-row tiles that gather a rows-contiguous A keep the one-ahead
    // loop: measured 118 against 125 TFLOP/s at 8192^3 NN, 108 against 120 NT)
    constexpr bool AHEAD2 =
        (BM + BN) * BK * (int)sizeof(T) / NT <= 160 && !(sizeof(T) == 4 && !TA && BM >= 256);
    // per-lane byte offsets within a slab must fit the 32-bit buffer offset
    const int64_t aspan = (TA ? (int64_t)(BM - 1) * p.lda + BK : BM + (int64_t)(BK - 1) * p.lda) * (int64_t)sizeof(T);
    const int64_t bspan = (TB ? BN + (int64_t)(BK - 1) * p.ldb : (int64_t)(BN - 1) * p.ldb + BK) * (int64_t)sizeof(T);
    const bool spans_ok = aspan < 0x7fffffff && bspan < 0x7fffffff;
    if (interior && AHEAD2 && spans_ok) {
        LA la1;
        LB lb1;
        typename LA::Offsets oa;
        typename LB::Offsets ob;
        LA::chunk_offsets(oa, 0, p.lda, BM, tid);
        LB::chunk_offsets(ob, 0, p.ldb, BN, tid);
        // slab kt (past the end: the last slab again, into the idle set)
        auto fetch = [&](LA &xa, LB &xb, int kt) {
            const int ks = kt < nk ? kt : nk - 1;
            xa.load_slab(operand_rsrc(abase + ks * astep), oa, 0, p.lda);
            xb.load_slab(operand_rsrc(bbase + ks * bstep), ob, 0, p.ldb);
        };
        // multiply slab kt; (sa, sb) hold slab kt+1, (ia, ib) receive slab kt+2
        auto step = [&](int kt, const LA &sa, const LB &sb, LA &ia, LB &ib) {
            T *cur = smem + (kt & 1) * STAGE;
            T *nxt = smem + ((kt + 1) & 1) * STAGE;
            fetch(ia, ib, kt + 2);
            __builtin_amdgcn_sched_barrier(0);   // issue the loads before the slab's work
            w.compute(cur, cur + BOFF);
            sa.store(nxt, tid);
            sb.store(nxt + BOFF, tid);
            lds_barrier();
        };
        fetch(la1, lb1, 1);   // slab 0 is in LDS
        int kt = 0;
        for (; kt + 1 < nk; kt += 2) {
            step(kt, la1, lb1, la, lb);
            step(kt + 1, la, lb, la1, lb1);
        }
        if (kt < nk) step(kt, la1, lb1, la, lb);
    } else if (interior) {
        for (int kt = 0; kt < nk; ++kt) {
            T *cur = smem + (kt & 1) * STAGE;
            T *nxt = smem + ((kt + 1) & 1) * STAGE;
            const bool more = kt + 1 < nk;
            if (more) {
                la.load_full(abase + (kt + 1) * astep, p.lda, tid);
                lb.load_full(bbase + (kt + 1) * bstep, p.ldb, tid);
            }
            w.compute(cur, cur + BOFF);
            if (more) {
                la.store(nxt, tid);
                lb.store(nxt + BOFF, tid);
            }
            __syncthreads();
        }
    } else {
        for (int kt = 0; kt < nk; ++kt) {
            T *cur = smem + (kt & 1) * STAGE;
            T *nxt = smem + ((kt + 1) & 1) * STAGE;
            const bool more = kt + 1 < nk;
            if (more) {
                const int krem = kend - kbeg - (kt + 1) * BK;
                la.load(abase + (kt + 1) * astep, p.lda, mlim, krem, tid);
                lb.load(bbase + (kt + 1) * bstep, p.ldb, nlim, krem, tid);
            }
            w.compute(cur, cur + BOFF);
            if (more) {
                la.store(nxt, tid);
                lb.store(nxt + BOFF, tid);
            }
            __syncthreads();
        }
    }

    if constexpr (KW > 1) {
        // accumulators of one wave (FP32: 16 per sub-tile, FP64: 4 rotations)
        constexpr int PER_SUB = sizeof(T) == 8 ? 4 : 16;
        constexpr int PER_WAVE = Wave::SM * Wave::SN * PER_SUB * 64;
        static_assert((KW - 1) * WM * WN * PER_WAVE <= 2 * STAGE, "partials fit the slab buffers");
        T *red = smem;
        __syncthreads();   // every wave is done with the slabs
        if (kg > 0) {
            T *dst = red + ((kg - 1) * WM * WN + wmn) * PER_WAVE + lane;
#pragma unroll
            for (int sm = 0; sm < Wave::SM; ++sm)
#pragma unroll
                for (int sn = 0; sn < Wave::SN; ++sn)
#pragma unroll
                    for (int e = 0; e < PER_SUB; ++e) dst[((sm * Wave::SN + sn) * PER_SUB + e) * 64] = w.acc[sm][sn][e];
        }
        __syncthreads();
        if (kg > 0) return;
#pragma unroll
        for (int g = 1; g < KW; ++g) {
            const T *src = red + ((g - 1) * WM * WN + wmn) * PER_WAVE + lane;
#pragma unroll
            for (int sm = 0; sm < Wave::SM; ++sm)
#pragma unroll
                for (int sn = 0; sn < Wave::SN; ++sn)
#pragma unroll
                    for (int e = 0; e < PER_SUB; ++e) w.acc[sm][sn][e] += src[((sm * Wave::SN + sn) * PER_SUB + e) * 64];
        }
    }

    // FEPI: a tile wholly inside C stores without range checks (store_full);
    // the partials of split-K are alpha = 1 stores of the same kind
    const bool full = FEPI && mlim >= BM && nlim >= BN;
    if (p.work) {
        T *out = p.work + (int64_t)blockIdx.y * p.M * p.N;
        if constexpr (FEPI) {
            if (full) w.store_full(out + m0 + (int64_t)n0 * p.M, p.M, lane, wm0, wn0, 1.0, 0.0);
        }
        if (!full) {
            w.for_each(lane, wm0, wn0, [&](int r, int c, T v) {
                if (r < mlim && c < nlim) out[(m0 + r) + (int64_t)(n0 + c) * p.M] = v;
            });
        }
    } else {
        T *cbase = p.C + m0 + (int64_t)n0 * p.ldc;
        const T alpha = p.alpha, beta = p.beta;
        if (full) {
            if constexpr (FEPI) w.store_full(cbase, p.ldc, lane, wm0, wn0, alpha, beta);
        } else if (beta == T(0)) {
            w.for_each(lane, wm0, wn0, [&](int r, int c, T v) {
                if (r < mlim && c < nlim) cbase[r + (int64_t)c * p.ldc] = alpha * v;
            });
        } else {
            w.for_each(lane, wm0, wn0, [&](int r, int c, T v) {
                if (r < mlim && c < nlim) {
                    T *dst = cbase + r + (int64_t)c * p.ldc;
                    *dst = axpby(alpha, v, beta, dst);
                }
            });
        }
    }
}

// Sum split-K partials in partition order, then apply alpha/beta.  Grid
// (row blocks, columns): rows and columns come from the indices, not from a
// 64-bit division of a flat index.
template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int M, int N, int splits, const T *__restrict__ work,
                                                             T alpha, T beta, T *__restrict__ C, int64_t ldc) {
    const int64_t total = (int64_t)M * N;
    for (int j = blockIdx.y; j < N; j += gridDim.y) {
        const T *wj = work + (int64_t)j * M;
        T *cj = C + (int64_t)j * ldc;
        for (int i = blockIdx.x * 256 + threadIdx.x; i < M; i += gridDim.x * 256) {
            T s = wj[i];
            for (int z = 1; z < splits; ++z) s += wj[i + (int64_t)z * total];
            cj[i] = axpby(alpha, s, beta, cj + i);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(int M, int N, T beta, T *__restrict__ C, int64_t ldc) {
    const int64_t total = (int64_t)M * N;
    for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int i = (int)(idx % M);
        const int64_t j = idx / M;
        T *dst = C + i + j * ldc;
        *dst = (beta == T(0)) ? T(0) : beta * *dst;
    }
}

// ---------------------------------------------------------------- host launch
struct TileShape { int BM, BN, WM, WN, BK; };
// FP64 wave tiles hold 64 FP64 accumulators (128 VGPRs) at 64x64, so FP64
// keeps one wave per SIMD (4-wave workgroups); FP32 affords 8-wave groups.
// The 128x128 FP64 tile (one workgroup per CU) takes 32-deep slabs to halve
// the barriers per flop; smaller tiles keep 16 so two workgroups fit in LDS.
static constexpr TileShape kTilesF64[TILE_COUNT] = {
    {128, 128, 2, 2, 32}, {64, 64, 2, 2, 16}, {128, 64, 2, 2, 16}, {256, 64, 4, 1, 16}};
static constexpr TileShape kTilesF32[TILE_COUNT] = {
    {128, 128, 2, 2, 32}, {64, 64, 2, 2, 32}, {128, 64, 2, 2, 32}, {256, 128, 4, 2, 32}};

static const TileShape &tile_shape(int t, int elem_bytes) {
    return elem_bytes == 8 ? kTilesF64[t] : kTilesF32[t];
}
int tile_bm(int t, int elem_bytes) { return tile_shape(t, elem_bytes).BM; }
int tile_bn(int t, int elem_bytes) { return tile_shape(t, elem_bytes).BN; }
int tile_bk(int t, int elem_bytes) { return tile_shape(t, elem_bytes).BK; }

// K depth of each partition: ceil(K / splitk) rounded up to the 32-deep slab
// every tile shape divides (partitions start on a slab boundary).
static int k_chunk(int k, int splitk) {
    constexpr int BK = 32;
    const int s = splitk > 1 ? splitk : 1;
    const int chunk = (k + s - 1) / s;
    return (chunk + BK - 1) / BK * BK;
}

int gemm_k_partitions(int k, int splitk) {
    if (k <= 0) return 1;
    const int chunk = k_chunk(k, splitk);
    return (k + chunk - 1) / chunk;
}

size_t gemm_workspace_bytes(int m, int n, int, const GemmConfig &cfg, int elem_bytes) {
    if (cfg.splitk <= 1) return 0;
    return (size_t)cfg.splitk * (size_t)m * (size_t)n * (size_t)elem_bytes;
}

template <typename T, bool TA, bool TB, int BM, int BN, int WM, int WN, int BK, bool VEC, bool FEPI, int KW = 1>
static hipError_t launch_one_as(const GemmParams<T> &p, int splits, hipStream_t s) {
    constexpr size_t lds = 2ull * SlabLayout<T, TA, TB, BM, BN, BK>::STAGE * sizeof(T);
    auto kern = gemm_mfma_kernel<T, TA, TB, BM, BN, BK, WM, WN, VEC, FEPI, KW>;
    if (hipError_t e = lds_limit_once<gemm_mfma_kernel<T, TA, TB, BM, BN, BK, WM, WN, VEC, FEPI, KW>>(lds)) return e;
    dim3 grid(p.tiles_m * p.tiles_n, splits);
    hipLaunchKernelGGL(kern, grid, dim3(WM * WN * KW * 64), lds, s, p);
    return hipGetLastError();
}

// FP64 launches of at most GEMM_FEPI_MAX_K per partition take the
// fast-epilogue instance
constexpr int GEMM_FEPI_MAX_K = 2048;
template <typename T, bool TA, bool TB, int BM, int BN, int WM, int WN, int BK, bool VEC>
static hipError_t launch_one(const GemmParams<T> &p, int splits, hipStream_t s) {
    if constexpr (sizeof(T) == 8)
        if (p.kchunk <= GEMM_FEPI_MAX_K) return launch_one_as<T, TA, TB, BM, BN, WM, WN, BK, VEC, true>(p, splits, s);
    return launch_one_as<T, TA, TB, BM, BN, WM, WN, BK, VEC, false>(p, splits, s);
}

// Grids of at most one workgroup per CU with partitions of at least 512 k
// take the k-grouped instances where the library chose the configuration:
// FP32 64x64 with four groups (SGEMM 1024^3 27.9 -> 24.6 us; shallower
// partitions measured slower), and with two groups on 64-deep slabs for
// grids of up to two workgroups per CU (1-5% faster on 1024 x 2048 x 1024,
// 2048 x 1024 x 2048 and the like), FP64 64x64 and 128x64 with two groups on the
// fast-epilogue instance, partitions of at most 2048 k (DGEMM 1024^3 58.5 ->
// 49.5 us; four FP64 groups spill at 1024 threads).  The DTRSM and LU
// updates keep one group: their solves measured 2% slower with two.
constexpr int KGROUP_MIN_K = 512;
template <typename T>
static bool use_kgroups(const GemmParams<T> &p, int splits, bool kgroups, int per_cu = 1) {
    return kgroups && p.kchunk >= KGROUP_MIN_K && (sizeof(T) == 4 || p.kchunk <= GEMM_FEPI_MAX_K) &&
           (int64_t)p.tiles_m * p.tiles_n * splits <= (int64_t)per_cu * rt().cus;
}

template <typename T, bool TA, bool TB, bool VEC>
static hipError_t launch_tile(int tile, const GemmParams<T> &p, int splits, bool kgroups, hipStream_t s) {
    switch (tile) {
    case TILE_128x128: return launch_one<T, TA, TB, 128, 128, 2, 2, 32, VEC>(p, splits, s);
    case TILE_64x64:
        if constexpr (sizeof(T) == 4)
            if (use_kgroups(p, splits, kgroups, 2))
                if ((int64_t)p.tiles_m * p.tiles_n * splits > rt().cus)
                    // two workgroups per CU: two k-groups on 64-deep slabs
                    // (69 KB of LDS each), 16 waves per CU instead of 8
                    return launch_one_as<T, TA, TB, 64, 64, 2, 2, 64, VEC, false, 2>(p, splits, s);
        if (use_kgroups(p, splits, kgroups)) {
            // FP32: four k-groups of four waves on 128-deep slabs (16 waves
            // per CU instead of 4); FP64: two groups on 32-deep slabs
            if constexpr (sizeof(T) == 8) return launch_one_as<T, TA, TB, 64, 64, 2, 2, 32, VEC, true, 2>(p, splits, s);
            else return launch_one_as<T, TA, TB, 64, 64, 2, 2, 128, VEC, false, 4>(p, splits, s);
        }
        return launch_one<T, TA, TB, 64, 64, 2, 2, sizeof(T) == 8 ? 16 : 32, VEC>(p, splits, s);
    case TILE_128x64:
        if constexpr (sizeof(T) == 8)
            if (use_kgroups(p, splits, kgroups)) return launch_one_as<T, TA, TB, 128, 64, 2, 2, 32, VEC, true, 2>(p, splits, s);
        return launch_one<T, TA, TB, 128, 64, 2, 2, sizeof(T) == 8 ? 16 : 32, VEC>(p, splits, s);
    case TILE_LARGE:
        if constexpr (sizeof(T) == 8) return launch_one<T, TA, TB, 256, 64, 4, 1, 16, VEC>(p, splits, s);
        else return launch_one<T, TA, TB, 256, 128, 4, 2, 32, VEC>(p, splits, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
static hipError_t launch_gemm_t(bool ta, bool tb, int m, int n, int k, T alpha, const T *A, int64_t lda,
                                const T *B, int64_t ldb, T beta, T *C, int64_t ldc, const GemmConfig &cfg,
                                void *work, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (k <= 0 || alpha == T(0)) {
        if (beta == T(1)) return hipSuccess;
        const int64_t total = (int64_t)m * n;
        int blocks = (int)((total + 255) / 256);
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(scale_kernel<T>, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc);
        return hipGetLastError();
    }
    const int tile = (cfg.tile >= 0 && cfg.tile < TILE_COUNT) ? cfg.tile : TILE_128x128;
    constexpr int VW = 16 / (int)sizeof(T);
    GemmParams<T> p;
    p.M = m; p.N = n; p.K = k;
    p.alpha = alpha; p.beta = beta;
    p.A = A; p.lda = lda; p.B = B; p.ldb = ldb; p.C = C; p.ldc = ldc;
    const TileShape &ts = tile_shape(tile, (int)sizeof(T));
    p.tiles_m = (m + ts.BM - 1) / ts.BM;
    p.tiles_n = (n + ts.BN - 1) / ts.BN;
    const int splits = gemm_k_partitions(k, cfg.splitk);
    p.kchunk = k_chunk(k, cfg.splitk);
    p.work = (splits > 1) ? static_cast<T *>(work) : nullptr;
    if (splits > 1 && !work) return hipErrorInvalidValue;

    const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && (lda % VW == 0) &&
                     (ldb % VW == 0);
    hipError_t e;
    const int sel = (ta ? 2 : 0) | (tb ? 1 : 0);
    if (vec) {
        switch (sel) {
        case 0: e = launch_tile<T, false, false, true>(tile, p, splits, cfg.kgroups, s); break;
        case 1: e = launch_tile<T, false, true, true>(tile, p, splits, cfg.kgroups, s); break;
        case 2: e = launch_tile<T, true, false, true>(tile, p, splits, cfg.kgroups, s); break;
        default: e = launch_tile<T, true, true, true>(tile, p, splits, cfg.kgroups, s); break;
        }
    } else {
        switch (sel) {
        case 0: e = launch_tile<T, false, false, false>(tile, p, splits, cfg.kgroups, s); break;
        case 1: e = launch_tile<T, false, true, false>(tile, p, splits, cfg.kgroups, s); break;
        case 2: e = launch_tile<T, true, false, false>(tile, p, splits, cfg.kgroups, s); break;
        default: e = launch_tile<T, true, true, false>(tile, p, splits, cfg.kgroups, s); break;
        }
    }
    if (e != hipSuccess || splits == 1) return e;
    // one 256-row block per workgroup column, up to 65535 columns per launch
    // row (the kernel strides over the rest)
    const dim3 grid((m + 255) / 256 < 65535 ? (m + 255) / 256 : 65535, n < 65535 ? n : 65535);
    hipLaunchKernelGGL(splitk_reduce_kernel<T>, grid, dim3(256), 0, s, m, n, splits, static_cast<const T *>(work),
                       alpha, beta, C, ldc);
    return hipGetLastError();
}

int launch_dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double *A, int64_t lda,
                 const double *B, int64_t ldb, double beta, double *C, int64_t ldc, const GemmConfig &cfg,
                 void *work, hipStream_t s) {
    return (int)launch_gemm_t<double>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg, work, s);
}

int launch_sgemm(bool ta, bool tb, int m, int n, int k, float alpha, const float *A, int64_t lda,
                 const float *B, int64_t ldb, float beta, float *C, int64_t ldc, const GemmConfig &cfg,
                 void *work, hipStream_t s) {
    return (int)launch_gemm_t<float>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg, work, s);
}

int launch_scale_d(int m, int n, double beta, double *C, int64_t ldc, hipStream_t s) {
    if (m <= 0 || n <= 0) return 0;
    const int64_t total = (int64_t)m * n;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(scale_kernel<double>, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc);
    return (int)hipGetLastError();
}

int launch_scale_s(int m, int n, float beta, float *C, int64_t ldc, hipStream_t s) {
    if (m <= 0 || n <= 0) return 0;
    const int64_t total = (int64_t)m * n;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(scale_kernel<float>, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc);
    return (int)hipGetLastError();
}

}  // namespace gemm_amd
