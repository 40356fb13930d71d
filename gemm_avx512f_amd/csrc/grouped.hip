// Grouped GEMM: many independent problems of different shapes and transposes
// in ONE persistent launch (no reference counterpart: the reference's drivers
// handle one GEMM per call, src/gemm_driver.c:116; its closest analogue is the
// CAS-claimed tile queue of OpenBLAS-like_implementation/dgemm.c:464-482).
//
// The host sorts the problems by work (largest first), lays their output
// tiles out back to back (tile_base = exclusive prefix sum) and uploads one
// descriptor per problem and a tile -> problem table.  Each workgroup of the
// persistent grid claims tiles in that order from an atomic counter (or, with
// the static schedule, walks them with a fixed stride), so big tiles are
// spread over all CUs first; the launch is replayable from a captured hipGraph
// (a one-thread kernel zeroes the counter ahead of it).
// A tile finds its problem through the table and runs the same
// LDS-staged matrix-core tile as the single-problem kernel, with range checks
// at the problem's edges (sizes are arbitrary) and 16-byte vector loads for
// problems whose operands are aligned for them.
#include "common.hpp"
#include "gemm_tile.hpp"
#include "grouped.hpp"
#include <algorithm>
#include <numeric>
#include <vector>

namespace gemm_amd {

struct GroupDesc {
    int m, n, k, ta, tb, tiles_m, tiles_n, vec;   // vec: A, B 16-byte aligned with ld a multiple of 16 bytes
    int chunked;                                  // vec, and m, n, k multiples of the 16-byte chunk
    int cls;                                      // tile class: 0 = small tiles, 1 = large tiles
    int64_t lda, ldb, ldc, tile_base;
    const void *A, *B;
    void *C;
    double alpha, beta;
};

// Elements of one operand slab of ROWS rows: the larger of its layouts
// ([row][k], and k-major for FP64 and bf16).
template <typename TI, int ROWS, int BK>
constexpr size_t slab_elems() {
    const size_t rowk = (size_t)ROWS * (BK + 16 / sizeof(TI));
    const size_t kmaj = sizeof(TI) == 4 ? 0 : (size_t)BK * (ROWS + kmajor_pad<TI>());
    return rowk > kmaj ? rowk : kmaj;
}

// One BM x BN output tile of problem d; TI = operand element type, TO = C's.
// CHUNKED problems (every dimension a multiple of the 16-byte chunk, aligned
// operands) load whole chunks without branches and fetch slabs two ahead:
// while slab kt is multiplied out of LDS, slab kt+1 waits in one register set
// for its LDS store and slab kt+2 is in flight into the other, so each load
// has two slabs of matrix work to arrive in.  The others take load()'s
// range-checked element path one slab ahead.
template <typename TI, typename TO, bool TA, bool TB, int BM, int BN, int BK, int WM, int WN, bool CHUNKED>
__device__ __forceinline__ void grouped_tile(const GroupDesc &d, int tm, int tn, unsigned char *smem_raw,
                                             int tid) {   // d: in global memory
    constexpr int NT = WM * WN * 64;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    // rows-contiguous FP64 and bf16 operands are staged k-major (one 16-byte
    // store per register); FP64 reads them as the single-problem kernel does
    // (row / column pairs), bf16 with LDS transpose reads.  FP32 gathers them
    // k-wise.  The per-element arithmetic is the single-problem kernel's.
    constexpr bool BF = sizeof(TI) == 2, DP = sizeof(TI) == 8;
    constexpr bool KMA = (BF || DP) && !TA, KMB = (BF || DP) && TB;
    using LA = TileLoader<TI, BM, BK, NT, TA, true, KMA>;
    using LB = TileLoader<TI, BN, BK, NT, !TB, true, KMB>;
    constexpr int AKP = KMA ? LA::KPITCH : 0, BKP = KMB ? LB::KPITCH : 0;
    using Wave = std::conditional_t<
        BF, WaveBF16<WTM, WTN, BK, AKP, BKP>,
        std::conditional_t<DP, WaveF64<WTM, WTN, BK, AKP, BKP>, typename WaveSel<TI, WTM, WTN, BK>::type>>;
    constexpr int BOFF = LA::SIZE;               // B's slab follows A's in a stage
    constexpr int STAGE = LA::SIZE + LB::SIZE;   // at most what the launch reserves per stage
    static_assert(STAGE <= slab_elems<TI, BM, BK>() + slab_elems<TI, BN, BK>(), "stage exceeds the reserved LDS");
    TI *smem = reinterpret_cast<TI *>(smem_raw);
    const TI *A = static_cast<const TI *>(d.A);
    const TI *B = static_cast<const TI *>(d.B);
    const int lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WM) * WTM, wn0 = (wave / WM) * WTN;
    const int m0 = tm * BM, n0 = tn * BN;
    const int mlim = d.m - m0, nlim = d.n - n0;
    const TI *abase = TA ? A + (int64_t)m0 * d.lda : A + m0;
    const TI *bbase = TB ? B + n0 : B + (int64_t)n0 * d.ldb;
    const int64_t astep = TA ? BK : (int64_t)BK * d.lda;
    const int64_t bstep = TB ? (int64_t)BK * d.ldb : BK;
    Wave w;
    w.init(lane, wm0, wn0);
    const int nk = (d.k + BK - 1) / BK;
    if constexpr (CHUNKED) {
        LA la0, la1;
        LB lb0, lb1;
        const __amdgpu_buffer_rsrc_t ra = operand_rsrc(A), rb = operand_rsrc(B);
        const int64_t aoff = abase - A, boff = bbase - B;
        const int nf = d.k / BK;   // full slabs; a partial last one follows them
        if (nf > 0) {
            typename LA::Offsets oa;
            typename LB::Offsets ob;
            LA::chunk_offsets(oa, aoff, d.lda, mlim, tid);
            LB::chunk_offsets(ob, boff, d.ldb, nlim, tid);
            const int sa_step = (int)(astep * (int64_t)sizeof(TI)), sb_step = (int)(bstep * (int64_t)sizeof(TI));
            // full slab kt (clamped to the last one: the prefetch past the end
            // re-reads it into the idle register set)
            auto fetch = [&](LA &la, LB &lb, int kt) {
                const int ks = kt < nf ? kt : nf - 1;
                la.load_slab(ra, oa, ks * sa_step, d.lda);
                lb.load_slab(rb, ob, ks * sb_step, d.ldb);
            };
            // multiply slab kt; (sa, sb) hold slab kt+1, (ia, ib) receive slab kt+2
            auto step = [&](int kt, const LA &sa, const LB &sb, LA &ia, LB &ib) {
                TI *cur = smem + (kt & 1) * STAGE;
                TI *nxt = smem + ((kt + 1) & 1) * STAGE;
                fetch(ia, ib, kt + 2);
                __builtin_amdgcn_sched_barrier(0);   // issue the loads before the slab's work
                w.compute(cur, cur + BOFF);
                // unconditional (after the last slab it fills the idle stage):
                // a skipped store would leave this set's loads pending on one
                // path and cost exact wait counts everywhere after it
                sa.store(nxt, tid);
                sb.store(nxt + BOFF, tid);
                lds_barrier();   // slab kt+2 stays in flight
            };
            fetch(la0, lb0, 0);
            fetch(la1, lb1, 1);
            la0.store(smem, tid);
            lb0.store(smem + BOFF, tid);
            lds_barrier();
            // one exit, at the loop test: every path round the loop runs both
            // steps and carries exactly one set in flight (la1/lb1), so the
            // wait counts stay exact (an exit between the steps would merge a
            // path with la0/lb0 still in flight into the back edge)
            int kt = 0;
            for (; kt + 1 < nf; kt += 2) {
                step(kt, la1, lb1, la0, lb0);
                step(kt + 1, la0, lb0, la1, lb1);
            }
            if (kt < nf) step(kt, la1, lb1, la0, lb0);
        }
        if (nf < nk) {
            // partial last slab, range-checked in k, into the stage after the
            // last full one (idle since the barrier that ended that slab)
            TI *st = smem + (nf & 1) * STAGE;
            const int krem = d.k - nf * BK;
            la0.load_buf(ra, aoff + nf * astep, d.lda, mlim, krem, tid);
            lb0.load_buf(rb, boff + nf * bstep, d.ldb, nlim, krem, tid);
            la0.store(st, tid);
            lb0.store(st + BOFF, tid);
            lds_barrier();
            w.compute(st, st + BOFF);
        }
    } else {
        // range-checked element path, one slab ahead (its branches would make a
        // second register set wait with the first)
        LA la;
        LB lb;
        la.vec = lb.vec = d.vec != 0;
        if (nk > 0) {
            la.load(abase, d.lda, mlim, d.k, tid);
            lb.load(bbase, d.ldb, nlim, d.k, tid);
            la.store(smem, tid);
            lb.store(smem + BOFF, tid);
        }
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            TI *cur = smem + (kt & 1) * STAGE;
            TI *nxt = smem + ((kt + 1) & 1) * STAGE;
            const bool more = kt + 1 < nk;
            if (more) {
                const int krem = d.k - (kt + 1) * BK;
                la.load(abase + (kt + 1) * astep, d.lda, mlim, krem, tid);
                lb.load(bbase + (kt + 1) * bstep, d.ldb, nlim, krem, tid);
            }
            w.compute(cur, cur + BOFF);
            if (more) {
                la.store(nxt, tid);
                lb.store(nxt + BOFF, tid);
            }
            __syncthreads();
        }
    }
    auto cbase = as_global(static_cast<TO *>(d.C)) + m0 + (int64_t)n0 * d.ldc;
    const TO alpha = (TO)d.alpha, beta = (TO)d.beta;
    // beta is tested once per tile (its descriptor value is uniform); a tile
    // wholly inside C with beta == 0 stores without range checks
    if constexpr (DP) {
        if (mlim >= BM && nlim >= BN) {
            w.store_full(cbase, d.ldc, lane, wm0, wn0, alpha, beta);
            return;
        }
    } else {
        if (d.beta == 0.0 && mlim >= BM && nlim >= BN) {
            w.store_full(cbase, d.ldc, lane, wm0, wn0, alpha);
            return;
        }
    }
    if (d.beta == 0.0) {
        w.for_each(lane, wm0, wn0, [&](int r, int c, TO v) {
            if (r < mlim && c < nlim) cbase[r + (int64_t)c * d.ldc] = alpha * v;
        });
    } else {
        w.for_each(lane, wm0, wn0, [&](int r, int c, TO v) {
            if (r < mlim && c < nlim) {
                auto dst = cbase + r + (int64_t)c * d.ldc;
                *dst = axpby(alpha, v, beta, dst);
            }
        });
    }
}

// One tile of shape BM x BN (slab depth BK) in the variant the problem needs.
template <typename TI, typename TO, int BM, int BN, int BK, int WM, int WN>
__device__ __forceinline__ void grouped_dispatch(const GroupDesc &d, int tm, int tn, unsigned char *smem, int tid) {
    switch (d.ta * 4 + d.tb * 2 + d.chunked) {
    case 0: grouped_tile<TI, TO, false, false, BM, BN, BK, WM, WN, false>(d, tm, tn, smem, tid); break;
    case 1: grouped_tile<TI, TO, false, false, BM, BN, BK, WM, WN, true>(d, tm, tn, smem, tid); break;
    case 2: grouped_tile<TI, TO, false, true, BM, BN, BK, WM, WN, false>(d, tm, tn, smem, tid); break;
    case 3: grouped_tile<TI, TO, false, true, BM, BN, BK, WM, WN, true>(d, tm, tn, smem, tid); break;
    case 4: grouped_tile<TI, TO, true, false, BM, BN, BK, WM, WN, false>(d, tm, tn, smem, tid); break;
    case 5: grouped_tile<TI, TO, true, false, BM, BN, BK, WM, WN, true>(d, tm, tn, smem, tid); break;
    case 6: grouped_tile<TI, TO, true, true, BM, BN, BK, WM, WN, false>(d, tm, tn, smem, tid); break;
    default: grouped_tile<TI, TO, true, true, BM, BN, BK, WM, WN, true>(d, tm, tn, smem, tid); break;
    }
}

// Two tile classes in one launch (4 waves each): small (S) for problems that
// a large tile would mostly pad, large (L) for the rest; the host picks the
// class per problem.  A plan whose problems are all small launches the
// instantiation with L = S, whose registers (the small class alone) keep
// more waves resident.
//
// Schedules: static (queue == null), tile blockIdx.x + i * gridDim.x; or
// dynamic (the default with more than two tiles per workgroup): the tiles are
// dealt into 8 lists, one per XCD (workgroup b runs on XCD b mod 8), each
// problem's tile grid cut into 8 near-square blocks so that an XCD's
// workgroups share the A and B panels of their block in that XCD's L2; a
// workgroup starts on entry b / 8 of its XCD's list and claims each later
// entry from that list's counter (one atomic, issued while the current tile
// runs), taking entries from the other XCDs' lists once its own is empty.
// Tiles thus go to whichever workgroup is free first, largest problems first.
// The counters (queue[16 q], one per list) are zeroed by
// grouped_queue_reset_kernel before the launch.  timeline (null: off): per
// workgroup, the 100 MHz real-time counter when it starts and when it has
// finished its last tile, and its tile count.
template <typename TI, typename TO, typename S, typename L>
__global__ __launch_bounds__(256) void grouped_gemm_kernel(const GroupDesc *__restrict__ descs,
                                                           const int *__restrict__ tilemap, int64_t total_tiles,
                                                           int *__restrict__ queue, const int *__restrict__ xlist,
                                                           const int64_t *__restrict__ xoff,
                                                           unsigned long long *__restrict__ timeline) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int64_t s_t;
    __shared__ int s_q;
    const unsigned long long t_start = timeline ? __builtin_amdgcn_s_memrealtime() : 0;
    const int x = blockIdx.x & 7;
    // workgroups on XCD q
    auto nwg = [&](int q) { return (int64_t)((int)gridDim.x - q + 7) >> 3; };
    // thread 0: the first tile at list index i of list q, or the next one any
    // list still holds (claimed from the counters); -1 when all are empty
    auto resolve = [&](int q, int64_t i, bool own) -> void {
        int64_t tt = -1;
        int qq = q;
        for (int s = 0; s < 8; ++s) {
            qq = (q + s) & 7;
            if (s > 0 || !own) i = nwg(qq) + (int64_t)atomicAdd(queue + 16 * qq, 1);
            if (i < xoff[qq + 1] - xoff[qq]) {
                tt = xlist[xoff[qq] + i];
                break;
            }
        }
        s_t = tt;
        s_q = qq;
    };
    int64_t t = blockIdx.x;
    int q = x;
    if (queue) {
        if (threadIdx.x == 0) resolve(x, blockIdx.x >> 3, true);
        __syncthreads();
        t = s_t;
        q = s_q;
        __syncthreads();
    }
    int done = 0;
    while (queue ? t >= 0 : t < total_tiles) {
        int64_t claimed = 0;
        if (queue && threadIdx.x == 0) claimed = nwg(q) + (int64_t)atomicAdd(queue + 16 * q, 1);
        const GroupDesc &d = descs[tilemap[t]];   // read in place (uniform scalar loads), never copied to private memory
        const int local = (int)(t - d.tile_base);
        const int tm = local % d.tiles_m, tn = local / d.tiles_m;
        // the thread index passes through an empty asm each tile, so the tile
        // variants' LDS addresses are formed inside their own code paths
        // instead of being hoisted out of this loop and held live in registers
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        if constexpr (std::is_same_v<S, L>) {
            grouped_dispatch<TI, TO, S::BM, S::BN, S::BK, S::WM, S::WN>(d, tm, tn, smem, tid);
        } else {
            if (d.cls) grouped_dispatch<TI, TO, L::BM, L::BN, L::BK, L::WM, L::WN>(d, tm, tn, smem, tid);
            else grouped_dispatch<TI, TO, S::BM, S::BN, S::BK, S::WM, S::WN>(d, tm, tn, smem, tid);
        }
        ++done;
        if (queue) {
            if (threadIdx.x == 0) {
                // the claimed entry of this list if it exists, else another list's
                if (claimed < xoff[q + 1] - xoff[q]) {
                    s_t = xlist[xoff[q] + claimed];
                    s_q = q;
                } else {
                    resolve(q, claimed, true);
                }
            }
            __syncthreads();   // the LDS stages are free again and s_t is visible
            t = s_t;
            q = s_q;
            __syncthreads();   // every thread has read s_t before it is written again
        } else {
            lds_barrier();   // the next tile reuses the LDS stages
            t += gridDim.x;
        }
    }
    if (timeline && threadIdx.x == 0) {
        unsigned long long *row = timeline + 3 * (size_t)blockIdx.x;
        row[0] = t_start;
        row[1] = __builtin_amdgcn_s_memrealtime();
        row[2] = (unsigned long long)done;
    }
}

// Zeroes the dynamic schedule's 8 tile counters (a kernel, not a memset, so
// it can sit in a captured graph ahead of the grouped launch).
__global__ __launch_bounds__(64) void grouped_queue_reset_kernel(int *queue) {
    if (threadIdx.x < 8) queue[16 * threadIdx.x] = 0;
}

// ---------------------------------------------------------------- host side
struct GroupedPlan {
    int dtype = 0, count = 0;
    int64_t total_tiles = 0;
    GroupDesc *d_descs = nullptr;
    int *d_tilemap = nullptr;   // problem (descriptor index) of every tile
    int grid = 0;
    size_t lds = 0;
    const void *kernel = nullptr;
    int64_t resident = 0;   // workgroups resident at once
    int *d_queue = nullptr;   // the dynamic schedule's 8 tile counters (64 bytes apart)
    int *d_xlist = nullptr;   // the dynamic schedule's tile lists, one per XCD, back to back
    int64_t *d_xoff = nullptr;   // list q = d_xlist[d_xoff[q] .. d_xoff[q+1])
    bool dynamic = true;      // default: on when there are more than two tiles per workgroup
                              // (on the 512-problem workload the static stride left 15-17% of
                              // the workgroup time idle in the tail)
    unsigned long long *timeline = nullptr;   // caller's device buffer, 3 words per workgroup
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// Tile classes per dtype, four waves each.  Small: 64x64 tiles of 32x32
// waves, the single-problem kernel's 64x64 tile and slab depth (FP32, FP64),
// so the k order (and the result) is the same bit for bit.  Large: the
// single-problem kernel's main tiles with the same slab depths (FP32 128x128
// of 64x64 waves, FP64 128x64 of 64x32 waves; bf16 128x128 of 64x64 waves,
// 64-deep slabs): four times (twice) the matrix work per fragment read and
// per barrier.  A problem takes the large class when both of its output
// dimensions reach the plan's threshold (FP64: M that, N half of it);
// grouped_default_large_min gives the default per dtype.
struct TileCfg64x64F32 { static constexpr int BM = 64, BN = 64, BK = 32, WM = 2, WN = 2; };
struct TileCfg128F32 { static constexpr int BM = 128, BN = 128, BK = 32, WM = 2, WN = 2; };
struct TileCfg64x64F64 { static constexpr int BM = 64, BN = 64, BK = 16, WM = 2, WN = 2; };
struct TileCfg128x64F64 { static constexpr int BM = 128, BN = 64, BK = 16, WM = 2, WN = 2; };
struct TileCfg64x64BF16 { static constexpr int BM = 64, BN = 64, BK = 64, WM = 2, WN = 2; };
struct TileCfg128BF16 { static constexpr int BM = 128, BN = 128, BK = 64, WM = 2, WN = 2; };
template <typename TI, typename TO>
struct GroupedCfg;
template <> struct GroupedCfg<float, float> { using S = TileCfg64x64F32; using L = TileCfg128F32; };
template <> struct GroupedCfg<double, double> { using S = TileCfg64x64F64; using L = TileCfg128x64F64; };
template <> struct GroupedCfg<uint16_t, float> { using S = TileCfg64x64BF16; using L = TileCfg128BF16; };

template <typename TI, typename C>
constexpr size_t stage_bytes() {
    return 2ull * (slab_elems<TI, C::BM, C::BK>() + slab_elems<TI, C::BN, C::BK>()) * sizeof(TI);
}

// LDS for the classes the plan uses (a plan with small tiles only keeps more
// workgroups resident), and the resident-grid size at that LDS.
template <typename TI, typename TO>
static int prepare(GroupedPlan &P, int cus, bool any_large) {
    using G = GroupedCfg<TI, TO>;
    const size_t ls = stage_bytes<TI, typename G::S>(), ll = stage_bytes<TI, typename G::L>();
    P.lds = any_large && ll > ls ? ll : ls;
    const void *kern = any_large
                           ? reinterpret_cast<const void *>(grouped_gemm_kernel<TI, TO, typename G::S, typename G::L>)
                           : reinterpret_cast<const void *>(grouped_gemm_kernel<TI, TO, typename G::S, typename G::S>);
    P.kernel = kern;
    // the limit covers either class, so plans of both kinds can be launched in any order
    hipError_t e = hipFuncSetAttribute(P.kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(ll > ls ? ll : ls));
    // persistent grid: as many workgroups as can be resident at once
    int per_cu = 0;
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, P.lds);
    P.resident = (int64_t)cus * (per_cu > 0 ? per_cu : 1);
    return (int)e;
}

template <typename TI, typename TO>
static void tile_dims(int cls, int *bm, int *bn) {
    using G = GroupedCfg<TI, TO>;
    *bm = cls ? G::L::BM : G::S::BM;
    *bn = cls ? G::L::BN : G::S::BN;
}
static void tile_dims_of(int dtype, int cls, int *bm, int *bn) {
    if (dtype == GROUPED_F64) tile_dims<double, double>(cls, bm, bn);
    else if (dtype == GROUPED_F32) tile_dims<float, float>(cls, bm, bn);
    else tile_dims<uint16_t, float>(cls, bm, bn);
}

void *grouped_plan_create(int dtype, int count, const char *ta, const char *tb, const int *m, const int *n,
                          const int *k, const double *alpha, const void *const *A, const int *lda,
                          const void *const *B, const int *ldb, const double *beta, void *const *Cp, const int *ldc,
                          int cus, int large_min, int *status) {
    *status = 0;
    GroupedPlan *P = new GroupedPlan();
    P->dtype = dtype;
    if (large_min <= 0) large_min = grouped_default_large_min(dtype);
    const int eb = dtype == GROUPED_F64 ? 8 : (dtype == GROUPED_F32 ? 4 : 2);   // operand element bytes
    const int ebc = dtype == GROUPED_F64 ? 8 : 4;                                // C element bytes
    // Work items: an FP64 problem on the large tile class (128x64) whose last
    // tile row would hold at most a small tile's rows runs as two items, the
    // rows of whole large tiles and the remaining rows on small tiles (the
    // classes compute the same bits), so that no large tile row is mostly
    // padding.  (For bf16 and FP32, whose small tiles are also narrower and
    // much slower per element than their 128x128 ones, it did not pay.)
    struct Item {
        int p, r0, rows, cls;
    };
    std::vector<Item> items;
    items.reserve(count);
    {
        int BML, BNL, BMS, BNS;
        tile_dims_of(dtype, 1, &BML, &BNL);
        tile_dims_of(dtype, 0, &BMS, &BNS);
        for (int p = 0; p < count; ++p) {
            const int cls = m[p] >= large_min && n[p] >= (dtype == GROUPED_F64 ? large_min / 2 : large_min);
            const int rem = m[p] % BML;
            if (dtype == GROUPED_F64 && cls && m[p] > BML && rem > 0 && rem <= BMS) {
                items.push_back({p, 0, m[p] - rem, 1});
                items.push_back({p, m[p] - rem, rem, 0});
            } else {
                items.push_back({p, 0, m[p], cls});
            }
        }
    }
    const int nitems = (int)items.size();
    P->count = nitems;
    // largest items first: their tiles are spread over all CUs before the small ones
    std::stable_sort(items.begin(), items.end(), [&](const Item &x, const Item &y) {
        return (double)x.rows * n[x.p] * k[x.p] > (double)y.rows * n[y.p] * k[y.p];
    });
    std::vector<GroupDesc> h(nitems);
    int64_t base = 0;
    bool any_large = false;
    for (int i = 0; i < nitems; ++i) {
        const int p = items[i].p;
        GroupDesc &d = h[i];
        d.m = items[i].rows;
        d.n = n[p];
        d.k = k[p];
        d.ta = !(ta[p] == 'N' || ta[p] == 'n');
        d.tb = !(tb[p] == 'N' || tb[p] == 'n');
        d.cls = items[i].cls;
        any_large = any_large || d.cls;
        // the item's rows of op(A) and C
        const int64_t r0 = items[i].r0;
        const void *Ai = static_cast<const char *>(A[p]) + (d.ta ? r0 * lda[p] : r0) * eb;
        void *Ci = static_cast<char *>(Cp[p]) + r0 * ebc;
        int BM, BN;
        tile_dims_of(dtype, d.cls, &BM, &BN);
        d.tiles_m = (d.m + BM - 1) / BM;
        d.tiles_n = (d.n + BN - 1) / BN;
        {
            const int vw = 16 / eb;
            d.vec = ((uintptr_t)Ai % 16 == 0) && ((uintptr_t)B[p] % 16 == 0) && (lda[p] % vw == 0) &&
                    (ldb[p] % vw == 0);
            // extents in bytes, for the 32-bit buffer offsets of the chunked path
            const int64_t ar = d.ta ? k[p] : d.m, ac = d.ta ? d.m : k[p];
            const int64_t br = d.tb ? n[p] : k[p], bc = d.tb ? k[p] : n[p];
            const int64_t ea = ((int64_t)lda[p] * (ac - 1) + ar) * eb, ebb = ((int64_t)ldb[p] * (bc - 1) + br) * eb;
            d.chunked = d.vec && d.m % vw == 0 && n[p] % vw == 0 && k[p] % vw == 0 && ea < (1ll << 31) - 64 &&
                        ebb < (1ll << 31) - 64;
        }
        d.lda = lda[p];
        d.ldb = ldb[p];
        d.ldc = ldc[p];
        d.A = Ai;
        d.B = B[p];
        d.C = Ci;
        d.alpha = alpha[p];
        d.beta = beta[p];
        d.tile_base = base;
        base += (int64_t)d.tiles_m * d.tiles_n;
    }
    count = nitems;   // descriptors from here on
    P->total_tiles = base;
    // tile -> problem, one word per tile: a tile finds its descriptor with one
    // load instead of a binary search of dependent loads over the table
    std::vector<int> tilemap((size_t)(base > 0 ? base : 1), 0);
    for (int i = 0; i < count; ++i)
        std::fill(tilemap.begin() + h[i].tile_base, tilemap.begin() + h[i].tile_base +
                  (int64_t)h[i].tiles_m * h[i].tiles_n, i);
    hipError_t e = hipMalloc(&P->d_descs, sizeof(GroupDesc) * (count > 0 ? count : 1));
    if (e == hipSuccess) e = hipMalloc(&P->d_tilemap, sizeof(int) * tilemap.size());
    if (e == hipSuccess)
        e = hipMemcpy(P->d_tilemap, tilemap.data(), sizeof(int) * tilemap.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&P->d_queue, 8 * 16 * sizeof(int));
    // the dynamic schedule's per-XCD lists: each problem's tile grid in 8
    // near-square blocks (problems of fewer than 16 tiles dealt tile by tile),
    // the blocks' XCDs rotated from problem to problem so the lists even out
    if (e == hipSuccess) {
        std::vector<std::vector<int>> lists(8);
        int rot = 0;
        for (int i = 0; i < count; ++i) {
            const GroupDesc &d = h[i];
            const int tm = d.tiles_m, tn = d.tiles_n, nt = tm * tn;
            if (nt < 16) {
                for (int j = 0; j < nt; ++j) lists[(rot + j) & 7].push_back((int)(d.tile_base + j));
                rot = (rot + nt) & 7;
                continue;
            }
            int pm = 1, pn = 8;
            double bestr = 1e30;
            for (int a = 1; a <= 8; a *= 2) {   // blocks of (tm / a) x (tn / (8 / a)) tiles, nearest to square
                const int b = 8 / a;
                if (a > tm || b > tn) continue;
                const double r = (double)tm / a / ((double)tn / b);
                const double dev = r > 1 ? r : 1 / r;
                if (dev < bestr) {
                    bestr = dev;
                    pm = a;
                    pn = b;
                }
            }
            if (bestr == 1e30) {   // thin grids (fewer than 8 tiles in both directions' splits): deal tiles
                for (int j = 0; j < nt; ++j) lists[(rot + j) & 7].push_back((int)(d.tile_base + j));
                rot = (rot + nt) & 7;
                continue;
            }
            for (int bj = 0; bj < pn; ++bj)
                for (int bi = 0; bi < pm; ++bi) {
                    const int q = (bi * pn + bj + rot) & 7;
                    const int m0 = bi * tm / pm, m1 = (bi + 1) * tm / pm;
                    const int n0 = bj * tn / pn, n1 = (bj + 1) * tn / pn;
                    for (int c = n0; c < n1; ++c)
                        for (int r = m0; r < m1; ++r) lists[q].push_back((int)(d.tile_base + r + (int64_t)c * tm));
                }
            rot = (rot + 1) & 7;
        }
        std::vector<int> flat;
        std::vector<int64_t> off(9, 0);
        for (int q = 0; q < 8; ++q) {
            off[q] = (int64_t)flat.size();
            flat.insert(flat.end(), lists[q].begin(), lists[q].end());
        }
        off[8] = (int64_t)flat.size();
        if (flat.empty()) flat.push_back(0);
        e = hipMalloc(&P->d_xlist, sizeof(int) * flat.size());
        if (e == hipSuccess) e = hipMemcpy(P->d_xlist, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&P->d_xoff, sizeof(int64_t) * 9);
        if (e == hipSuccess) e = hipMemcpy(P->d_xoff, off.data(), sizeof(int64_t) * 9, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && count > 0)
        e = hipMemcpy(P->d_descs, h.data(), sizeof(GroupDesc) * count, hipMemcpyHostToDevice);
    int st = (int)e;
    if (!st) {
        if (dtype == GROUPED_F32) st = prepare<float, float>(*P, cus, any_large);
        else if (dtype == GROUPED_F64) st = prepare<double, double>(*P, cus, any_large);
        else st = prepare<uint16_t, float>(*P, cus, any_large);
    }
    int64_t g = P->resident;
    if (g > P->total_tiles) g = P->total_tiles;
    P->grid = (int)(g > 0 ? g : 1);
    // the dynamic schedule pays where workgroups run several tiles of unequal
    // cost; with about one tile each its atomics only add latency
    P->dynamic = P->total_tiles > 2 * (int64_t)P->grid;
    if (st) {
        grouped_plan_destroy(P);
        *status = st;
        return nullptr;
    }
    return P;
}

int grouped_plan_launch(void *plan, hipStream_t s) {
    GroupedPlan *P = static_cast<GroupedPlan *>(plan);
    if (P->total_tiles == 0) return 0;
    int *queue = P->dynamic ? P->d_queue : nullptr;
    if (queue) {
        hipLaunchKernelGGL(grouped_queue_reset_kernel, dim3(1), dim3(64), 0, s, queue);
        if (hipError_t e = hipGetLastError()) return (int)e;
    }
    void *args[] = {&P->d_descs, &P->d_tilemap, &P->total_tiles, &queue, &P->d_xlist, &P->d_xoff, &P->timeline};
    return (int)hipLaunchKernel(P->kernel, dim3(P->grid), dim3(256), args, P->lds, s);
}

int grouped_plan_configure(void *plan, int dynamic, unsigned long long *timeline) {
    GroupedPlan *P = static_cast<GroupedPlan *>(plan);
    if (P->exec) return (int)hipErrorInvalidValue;   // the captured launch keeps its arguments
    P->dynamic = dynamic != 0;
    P->timeline = timeline;
    return 0;
}

int grouped_plan_grid(void *plan) { return static_cast<GroupedPlan *>(plan)->grid; }

int grouped_plan_capture(void *plan, hipStream_t s) {
    GroupedPlan *P = static_cast<GroupedPlan *>(plan);
    if (P->exec) return 0;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return (int)e;
    int le = grouped_plan_launch(plan, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (le) {
        if (g) (void)hipGraphDestroy(g);
        return le;
    }
    if (e != hipSuccess) return (int)e;
    P->graph = g;
    e = hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0);
    return (int)e;
}

int grouped_plan_replay(void *plan, hipStream_t s) {
    GroupedPlan *P = static_cast<GroupedPlan *>(plan);
    if (!P->exec) return (int)hipErrorInvalidValue;
    return (int)hipGraphLaunch(P->exec, s);
}

int grouped_plan_destroy(void *plan) {
    GroupedPlan *P = static_cast<GroupedPlan *>(plan);
    if (!P) return 0;
    hipError_t e = hipSuccess;
    if (P->exec) e = hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->d_descs) (void)hipFree(P->d_descs);
    if (P->d_tilemap) (void)hipFree(P->d_tilemap);
    if (P->d_queue) (void)hipFree(P->d_queue);
    if (P->d_xlist) (void)hipFree(P->d_xlist);
    if (P->d_xoff) (void)hipFree(P->d_xoff);
    delete P;
    return (int)e;
}

int64_t grouped_plan_tiles(void *plan) { return static_cast<GroupedPlan *>(plan)->total_tiles; }

}  // namespace gemm_amd
