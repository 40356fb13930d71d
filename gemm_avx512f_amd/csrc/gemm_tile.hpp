// Device-side building blocks of the FP32/FP64 matrix-core GEMM tiles, shared by
// the single-problem kernel (gemm_mfma.hip) and the grouped kernel (grouped.hip).
#pragma once
#include "common.hpp"
#include <type_traits>

namespace gemm_amd {

template <typename T>
struct GemmParams {
    int M, N, K;
    T alpha, beta;
    const T *A;
    int64_t lda;
    const T *B;
    int64_t ldb;
    T *C;
    int64_t ldc;
    int tiles_m, tiles_n;
    int kchunk;   // K extent per split-K partition (multiple of BK)
    T *work;      // split-K partials, [split][N][M], or nullptr
};

// ---------------------------------------------------------------- tile loader
// Moves a ROWS x BK slab of an operand (row = M index of op(A) or N index of
// op(B)) into LDS laid out [row][PITCH].  KCONTIG: the operand is stored with
// k contiguous (element (r,k) at base[k + r*ld]); otherwise rows contiguous
// (element (r,k) at base[r + k*ld]).  Elements outside (rlim, klim) read 0.
//
// Rows-contiguous FP32 operands are gathered k-wise: a thread reads the 4
// consecutive k of one row with scalar loads (neighbouring lanes read
// neighbouring rows, so every load is one coalesced 256-byte segment per
// wave) and writes them as one 16-byte LDS vector.  Transposing in the other
// direction (a 16-byte global load of 4 rows, scattered into 4 LDS rows) puts
// 32 lanes of a wave on 2 LDS banks with a 36-float row pitch.  FP64 keeps
// the transposing store: its 4-way conflicts cost less than the doubled load
// instruction count of a k-pair gather (measured 57.1 vs 53.7 TFLOP/s at
// 8192^3 NN).
//
// KMAJOR (rows-contiguous FP64 and bf16 operands): the slab is stored
// [k][row] with a row pitch of KPITCH = ROWS + 4 doubles (ROWS + 8 bf16)
// instead, so the 16-byte register of neighbouring rows lands as one
// conflict-free LDS store (the transposing store of the [row][k] layout puts
// 8 lanes on each FP64 bank pair, and writes bf16 two bytes at a time).  The
// bf16 wave reads such a slab with LDS transpose reads (WaveBF16).
// Row pitch padding (elements) of a k-major slab.
template <typename T>
constexpr int kmajor_pad() { return sizeof(T) == 2 ? 8 : 4; }

template <typename T, int ROWS, int BK, int NT, bool KCONTIG, bool VEC, bool KMAJOR = false>
struct TileLoader {
    using V = typename Vec16<T>::t;
    static constexpr int VW = Vec16<T>::N;
    static constexpr int PITCH = BK + VW;
    static constexpr int KPITCH = ROWS + kmajor_pad<T>();
    static constexpr int SIZE = KMAJOR ? BK * KPITCH : ROWS * PITCH;   // elements per stage
    static constexpr int NVEC = ROWS * BK / VW;
    static constexpr bool KGATHER = !KCONTIG && sizeof(T) == 4;
    static_assert(!KMAJOR || (!KCONTIG && sizeof(T) != 4), "k-major slabs are for rows-contiguous FP64 / bf16 operands");
    static_assert(NVEC % NT == 0, "tile must split evenly over the workgroup");
    static constexpr int PER = NVEC / NT;
    V reg[PER];
    bool vec = true;   // run-time veto of the vector path (VEC must also be set)

    __device__ __forceinline__ static void coords(int v, int &r, int &k) {
        if (KCONTIG) {
            r = v / (BK / VW);
            k = (v % (BK / VW)) * VW;
        } else if (KGATHER) {
            r = v % ROWS;
            k = (v / ROWS) * VW;
        } else {
            k = v / (ROWS / VW);
            r = (v % (ROWS / VW)) * VW;
        }
    }

    // Interior slab: every element in range and 16-byte aligned.
    __device__ __forceinline__ void load_full(const T *__restrict__ base, int64_t ld, int tid) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int r, k;
            coords(tid + i * NT, r, k);
            const T *src = KCONTIG ? base + (int64_t)r * ld + k : base + r + (int64_t)k * ld;
            if (KGATHER) {
#pragma unroll
                for (int e = 0; e < VW; ++e) reg[i][e] = src[(int64_t)e * ld];
            } else {
                reg[i] = *reinterpret_cast<const V *>(src);
            }
        }
    }

    // Range-checked slab (edges, unaligned operands).  The operand is read
    // through a global-address-space pointer: pointers taken from a
    // descriptor table would otherwise compile to flat loads.
    __device__ __forceinline__ void load(const T *__restrict__ base, int64_t ld, int rlim, int klim,
                                         int tid) {
        using GP = const __attribute__((address_space(1))) T *;
        const GP gbase = (GP)base;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int r, k;
            coords(tid + i * NT, r, k);
            const GP src = KCONTIG ? gbase + (int64_t)r * ld + k : gbase + r + (int64_t)k * ld;
            if (KGATHER) {
                V t;
#pragma unroll
                for (int e = 0; e < VW; ++e) t[e] = (r < rlim && k + e < klim) ? src[(int64_t)e * ld] : T(0);
                reg[i] = t;
                continue;
            }
            const bool full = KCONTIG ? (r < rlim && k + VW <= klim) : (r + VW <= rlim && k < klim);
            if (VEC && vec && full) {
                reg[i] = *reinterpret_cast<const __attribute__((address_space(1))) V *>(src);
            } else {
                V t;
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    const bool ok = KCONTIG ? (r < rlim && k + e < klim) : (r + e < rlim && k < klim);
                    t[e] = ok ? src[e] : T(0);
                }
                reg[i] = t;
            }
        }
    }

    // Whole-chunk loads through a buffer resource over the operand
    // (num_records 2^31 - 1 bytes), for operands whose 16-byte chunks lie
    // either wholly inside or wholly outside (rlim, klim) (the chunked
    // dimension a multiple of VW, rows 16-byte aligned) and whose extent is
    // below 2 GiB.  Every chunk is one unconditional load at a 32-bit byte
    // offset; an out-of-range chunk is given offset 2^31, past num_records,
    // which the buffer unit returns as zeros.
    //
    // load_buf: the slab at element offset `off`, range-checked in rows and k.
    __device__ __forceinline__ void load_buf(__amdgpu_buffer_rsrc_t rs, int64_t off, int64_t ld, int rlim, int klim,
                                             int tid) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int r, k;
            coords(tid + i * NT, r, k);
            const bool ok = r < rlim && k < klim;
            const int64_t e0 = off + (KCONTIG ? (int64_t)r * ld + k : r + (int64_t)k * ld);
            load_chunk(rs, i, ok ? (int)(e0 * (int64_t)sizeof(T)) : (int)0x80000000u, 0, ld);
        }
    }

    // For the k-loop over full slabs: the per-chunk byte offsets of the first
    // slab (rows past rlim at 2^31) are computed once per tile, and each slab
    // adds a wave-uniform byte offset `soff` (the instruction's scalar
    // offset), so the loop rewrites no address registers.  (A rewritten
    // address register that is still the source of an outstanding load makes
    // the compiler wait for that load.)
    struct Offsets {
        int v[PER];
    };
    __device__ __forceinline__ static void chunk_offsets(Offsets &o, int64_t off, int64_t ld, int rlim, int tid) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int r, k;
            coords(tid + i * NT, r, k);
            const int64_t e0 = off + (KCONTIG ? (int64_t)r * ld + k : r + (int64_t)k * ld);
            o.v[i] = r < rlim ? (int)(e0 * (int64_t)sizeof(T)) : (int)0x80000000u;
        }
    }
    __device__ __forceinline__ void load_slab(__amdgpu_buffer_rsrc_t rs, const Offsets &o, int soff, int64_t ld) {
#pragma unroll
        for (int i = 0; i < PER; ++i) load_chunk(rs, i, o.v[i], soff, ld);
    }

    __device__ __forceinline__ void load_chunk(__amdgpu_buffer_rsrc_t rs, int i, int voff, int soff, int64_t ld) {
        if constexpr (KGATHER) {
#pragma unroll
            for (int e = 0; e < VW; ++e)
                reg[i][e] = __builtin_bit_cast(
                    T, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff + (int)(e * ld * (int64_t)sizeof(T)), 0));
        } else {
            reg[i] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0));
        }
    }

    __device__ __forceinline__ void store(T *lds, int tid) const {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int r, k;
            coords(tid + i * NT, r, k);
            if (KMAJOR) {
                *reinterpret_cast<V *>(lds + k * KPITCH + r) = reg[i];
            } else if (KCONTIG || KGATHER) {
                *reinterpret_cast<V *>(lds + r * PITCH + k) = reg[i];
            } else {
#pragma unroll
                for (int e = 0; e < VW; ++e) lds[(r + e) * PITCH + k] = reg[i][e];
            }
        }
    }
};

// Buffer resource over memory at p for the raw buffer loads above: stride 0,
// num_records 2^31 - 1 bytes, the gfx9 data-format word (32-bit elements).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t operand_rsrc(const void *p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, 0x7fffffff, 0x00020000);
}

// ---------------------------------------------------------------- wave tiles
// FP64: v_mfma_f64_4x4x4_4b.  A 16x16 sub-tile is four instructions: all use
// the same A fragment (rows 4b..4b+3 of the slab in block b) and rotation r
// feeds block b the B columns 4((b+r)%4)..+3, so the four rotations cover the
// 16 column groups x 4 row groups once.  Each rotation's B fragment is read
// straight from LDS (the row offset of lane q is its column 4((b+r)%4)+j):
// VALU work between the matrix instructions (e.g. DPP lane rotations)
// measurably starves the FP64 matrix pipe on gfx950, LDS reads do not.  Lane
// group g = lane/16 owns k = g*BK/4 .. (g+1)*BK/4-1 of each LDS slab, read two
// at a time (any k assignment is valid as long as A and B use the same one).
//
// APK > 0: the A slab is k-major with row pitch APK, and A sub-tiles come in
// row pairs: lane q reads rows 2q, 2q+1 of a 32-row band at one k (one
// 16-byte read) and feeds the first to sub-tile 2p (the band's even rows) and
// the second to sub-tile 2p+1 (its odd rows), both against the same B
// fragment; the two k of a k pair are two such reads.  BPK > 0 does the same
// for a k-major B slab: column pairs of a 32-column band feed the band's even
// and odd sub-tiles (the rotation picks the pair).
// RP is the row pitch of a [row][k] slab and koff the first k of the wave's
// part (a slab deeper than BK shared by several waves along k).
template <int WTM, int WTN, int BK, int APK = 0, int BPK = 0, int RP = BK + 2>
struct WaveF64 {
    static constexpr int SM = WTM / 16, SN = WTN / 16;
    static constexpr int PITCH = RP;
    static_assert(BK % 8 == 0, "FP64 slab depth must be a multiple of 8");
    static_assert(APK == 0 || SM % 2 == 0, "row pairs need an even number of A sub-tiles");
    static_assert(BPK == 0 || SN % 2 == 0, "column pairs need an even number of B sub-tiles");
    double acc[SM][SN][4];
    int aoff[SM], boff[SN][4];

    __device__ __forceinline__ void init(int lane, int wm0, int wn0, int koff = 0) {
        const int q = lane & 15, g = lane >> 4, b = q >> 2, j = q & 3;
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
            aoff[sm] = APK ? (koff + (BK / 4) * g) * APK + wm0 + 32 * (sm >> 1) + 2 * q
                           : (wm0 + sm * 16 + q) * PITCH + koff + (BK / 4) * g;
#pragma unroll
        for (int sn = 0; sn < SN; ++sn)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                boff[sn][r] = BPK ? (koff + (BK / 4) * g) * BPK + wn0 + 32 * (sn >> 1) + 2 * (4 * ((b + r) & 3) + j)
                                  : (wn0 + sn * 16 + 4 * ((b + r) & 3) + j) * PITCH + koff + (BK / 4) * g;
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
#pragma unroll
            for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[sm][sn][r] = 0.0;
    }

    __device__ __forceinline__ void compute(const double *As, const double *Bs) {
#pragma unroll
        for (int h = 0; h < BK / 8; ++h) {
            d2_t av[SM], bv[SN][4];
            if (APK) {
                // av[2p + e] = rows (2q, 2q+1) of band p at k = 2h + e
#pragma unroll
                for (int p = 0; p < SM / 2; ++p)
#pragma unroll
                    for (int e = 0; e < 2; ++e)
                        av[2 * p + e] = *reinterpret_cast<const d2_t *>(As + aoff[2 * p] + (2 * h + e) * APK);
            } else {
#pragma unroll
                for (int sm = 0; sm < SM; ++sm) av[sm] = *reinterpret_cast<const d2_t *>(As + aoff[sm] + 2 * h);
            }
            if (BPK) {
                // bv[2p + e][r] = columns of band p, rotation r, at k = 2h + e
#pragma unroll
                for (int p = 0; p < SN / 2; ++p)
#pragma unroll
                    for (int e = 0; e < 2; ++e)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            bv[2 * p + e][r] = *reinterpret_cast<const d2_t *>(Bs + boff[2 * p][r] + (2 * h + e) * BPK);
            } else {
#pragma unroll
                for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                    for (int r = 0; r < 4; ++r) bv[sn][r] = *reinterpret_cast<const d2_t *>(Bs + boff[sn][r] + 2 * h);
            }
#pragma unroll
            for (int e = 0; e < 2; ++e)
#pragma unroll
                for (int sm = 0; sm < SM; ++sm)
#pragma unroll
                    for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            // k-major: sub-tile sm = 2p + hf takes element hf of the band's read at k = 2h + e
                            const double a = APK ? av[2 * (sm >> 1) + e][sm & 1] : av[sm][e];
                            const double bb = BPK ? bv[2 * (sn >> 1) + e][r][sn & 1] : bv[sn][r][e];
                            acc[sm][sn][r] = mfma_f64_4x4x4(a, bb, acc[sm][sn][r]);
                        }
            // keep each phase's fragment reads next to its matrix ops: hoisting
            // every phase's reads to the top would need 4x the fragment registers
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    // visit(row, col, value) for every accumulator element, tile-relative
    template <typename F>
    __device__ __forceinline__ void for_each(int lane, int wm0, int wn0, F &&visit) const {
        const int i = lane >> 4, b = (lane >> 2) & 3, j = lane & 3;
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
#pragma unroll
            for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    visit(APK ? wm0 + 32 * (sm >> 1) + 2 * (4 * b + i) + (sm & 1) : wm0 + sm * 16 + 4 * b + i,
                          BPK ? wn0 + 32 * (sn >> 1) + 2 * (4 * ((b + r) & 3) + j) + (sn & 1)
                              : wn0 + sn * 16 + 4 * ((b + r) & 3) + j,
                          acc[sm][sn][r]);
    }

    // C = alpha * acc (beta == 0: C is not read) or axpby(alpha, acc, beta, C)
    // for a tile wholly inside C, c at the tile's origin: the map of for_each
    // with one column pointer per (sub-tile column, rotation), formed once,
    // and the rows as small offsets from it -- no range checks and no 64-bit
    // multiply per element.
    template <typename P>
    __device__ __forceinline__ void store_full(P c, int64_t ldc, int lane, int wm0, int wn0, double alpha,
                                               double beta) const {
        const int i = lane >> 4, b = (lane >> 2) & 3, j = lane & 3;
        int roff[SM];
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
            roff[sm] = APK ? wm0 + 32 * (sm >> 1) + 2 * (4 * b + i) + (sm & 1) : wm0 + sm * 16 + 4 * b + i;
        auto col = [&](int sn, int r) {
            return BPK ? wn0 + 32 * (sn >> 1) + 2 * (4 * ((b + r) & 3) + j) + (sn & 1)
                       : wn0 + sn * 16 + 4 * ((b + r) & 3) + j;
        };
        if (beta == 0.0) {
#pragma unroll
            for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const P cp = c + (int64_t)col(sn, r) * ldc;
#pragma unroll
                    for (int sm = 0; sm < SM; ++sm) cp[roff[sm]] = alpha * acc[sm][sn][r];
                }
        } else {
#pragma unroll
            for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const P cp = c + (int64_t)col(sn, r) * ldc;
#pragma unroll
                    for (int sm = 0; sm < SM; ++sm) cp[roff[sm]] = axpby(alpha, acc[sm][sn][r], beta, cp + roff[sm]);
                }
        }
    }
};

// FP32: v_mfma_f32_32x32x2_f32.  Lane group g = lane/32 owns k = g*BK/2 ..
// (g+1)*BK/2 - 1 of the wave's BK-deep part of each slab, read 4 at a time.
// KP is the slab's row pitch and koff the first k of the wave's part (a slab
// deeper than BK shared by several waves along k).
template <int WTM, int WTN, int BK, int KP = BK + 4>
struct WaveF32 {
    static constexpr int SM = WTM / 32, SN = WTN / 32;
    static constexpr int PITCH = KP;
    static_assert(BK % 8 == 0, "FP32 slab depth must be a multiple of 8");
    f16_t acc[SM][SN];
    int aoff[SM], boff[SN];

    __device__ __forceinline__ void init(int lane, int wm0, int wn0, int koff = 0) {
        const int q = lane & 31, g = lane >> 5;
#pragma unroll
        for (int sm = 0; sm < SM; ++sm) aoff[sm] = (wm0 + sm * 32 + q) * PITCH + koff + (BK / 2) * g;
#pragma unroll
        for (int sn = 0; sn < SN; ++sn) boff[sn] = (wn0 + sn * 32 + q) * PITCH + koff + (BK / 2) * g;
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
#pragma unroll
            for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[sm][sn][e] = 0.0f;
    }

    __device__ __forceinline__ void compute(const float *As, const float *Bs) {
#pragma unroll
        for (int h = 0; h < BK / 8; ++h) {
            f4_t av[SM], bv[SN];
#pragma unroll
            for (int sm = 0; sm < SM; ++sm) av[sm] = *reinterpret_cast<const f4_t *>(As + aoff[sm] + 4 * h);
#pragma unroll
            for (int sn = 0; sn < SN; ++sn) bv[sn] = *reinterpret_cast<const f4_t *>(Bs + boff[sn] + 4 * h);
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int sm = 0; sm < SM; ++sm)
#pragma unroll
                    for (int sn = 0; sn < SN; ++sn)
                        acc[sm][sn] = mfma_f32_32x32x2(bv[sn][e], av[sm][e], acc[sm][sn]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    // B's fragment is the matrix instruction's first operand, so each
    // accumulator block is the transposed tile: lane l holds row l % 32 of C,
    // and one element index across the wave is 32 consecutive rows of two
    // columns (whole 128-byte runs of column-major C for the epilogue).
    template <typename F>
    __device__ __forceinline__ void for_each(int lane, int wm0, int wn0, F &&visit) const {
        const int i = lane & 31, g = lane >> 5;
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
#pragma unroll
            for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                for (int e = 0; e < 16; ++e)
                    visit(wm0 + sm * 32 + i, wn0 + sn * 32 + (e >> 2) * 8 + g * 4 + (e & 3), acc[sm][sn][e]);
    }

    // C = alpha * acc for a tile wholly inside C (no range checks), c at the
    // tile's (0, 0): a lane's elements lie on its row at column offsets that
    // are constants times ldc, so each store is one address add
    template <typename P>
    __device__ __forceinline__ void store_full(P c, int64_t ldc, int lane, int wm0, int wn0, float alpha) const {
        store_rows(acc, c, ldc, lane, wm0, wn0, alpha);
    }
    template <typename A, typename P>
    __device__ __forceinline__ static void store_rows(const A &acc, P c, int64_t ldc, int lane, int wm0, int wn0,
                                                      float alpha) {
        const int i = lane & 31, g = lane >> 5;
#pragma unroll
        for (int sm = 0; sm < SM; ++sm) {
            const P row = c + (wm0 + sm * 32 + i) + (int64_t)(wn0 + g * 4) * ldc;
#pragma unroll
            for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                for (int e = 0; e < 16; ++e)
                    row[(int64_t)(sn * 32 + (e >> 2) * 8 + (e & 3)) * ldc] = alpha * acc[sm][sn][e];
        }
    }
};

// BF16 inputs, FP32 accumulation: v_mfma_f32_32x32x16_bf16.  Operands are bf16
// bit patterns (uint16_t) in LDS rows of BK + 8 elements; lane group
// g = lane/32 supplies k = 8g..8g+7 of every 16-deep MFMA step.
//
// AKP / BKP > 0: that operand's slab is k-major ([k][row], row pitch AKP / BKP
// elements) and is read with ds_read_b64_tr_b16.  That instruction works on
// groups of 16 lanes: lane i of a group receives, as its element e (0..3),
// element i % 4 of the 8-byte word addressed by lane 4e + i / 4 of the group
// (checked on gfx950 with an index-valued LDS probe).  So if lane j of the
// group addresses (k0 + j / 4, r0 + 4 (j % 4)), lane i receives rows r0 + i at
// k0 .. k0+3: two such reads (k0 and k0 + 4) give a lane its 8 consecutive k.
template <int WTM, int WTN, int BK, int AKP = 0, int BKP = 0>
struct WaveBF16 {
    static constexpr int SM = WTM / 32, SN = WTN / 32;
    static constexpr int PITCH = BK + 8;
    static_assert(BK % 16 == 0, "bf16 slab depth must be a multiple of 16");
    static_assert(AKP % 4 == 0 && BKP % 4 == 0, "transpose reads need 8-byte aligned rows");
    f16_t acc[SM][SN];
    int aoff[SM], boff[SN];

    // offset of the lane's fragment in a 32-row block starting at row `row0`
    template <int KP>
    __device__ __forceinline__ static int frag_off(int lane, int row0) {
        if constexpr (KP > 0) {
            const int i = lane & 15, grp = lane >> 4;
            return (8 * (grp >> 1) + (i >> 2)) * KP + row0 + 16 * (grp & 1) + 4 * (i & 3);
        } else {
            return (row0 + (lane & 31)) * PITCH + 8 * (lane >> 5);
        }
    }
    // the lane's 8 k values of 16-deep k-step s
    template <int KP>
    __device__ __forceinline__ static bf8_t frag(const uint16_t *S, int off, int s) {
        if constexpr (KP > 0) {
            typedef short s4_t __attribute__((ext_vector_type(4)));
            typedef short s8_t __attribute__((ext_vector_type(8)));
            using LP = __attribute__((address_space(3))) s4_t *;
            const uint16_t *p = S + off + 16 * s * KP;
            const s4_t lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((LP)(p));
            const s4_t hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((LP)(p + 4 * KP));
            const s8_t v = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
            return __builtin_bit_cast(bf8_t, v);
        } else {
            return *reinterpret_cast<const bf8_t *>(S + off + 16 * s);
        }
    }

    __device__ __forceinline__ void init(int lane, int wm0, int wn0) {
#pragma unroll
        for (int sm = 0; sm < SM; ++sm) aoff[sm] = frag_off<AKP>(lane, wm0 + sm * 32);
#pragma unroll
        for (int sn = 0; sn < SN; ++sn) boff[sn] = frag_off<BKP>(lane, wn0 + sn * 32);
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
#pragma unroll
            for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[sm][sn][e] = 0.0f;
    }

    // The fragments of k-step s + 1 are read while the matrix instructions of
    // step s run (the sched barriers keep the scheduler from sinking each read
    // next to its use, which would leave the matrix pipe waiting on LDS).
    __device__ __forceinline__ void compute(const uint16_t *As, const uint16_t *Bs) {
        bf8_t av[2][SM], bv[2][SN];
        auto load = [&](int s, bf8_t *a, bf8_t *b) {
#pragma unroll
            for (int sm = 0; sm < SM; ++sm) a[sm] = frag<AKP>(As, aoff[sm], s);
#pragma unroll
            for (int sn = 0; sn < SN; ++sn) b[sn] = frag<BKP>(Bs, boff[sn], s);
        };
        load(0, av[0], bv[0]);
#pragma unroll
        for (int s = 0; s < BK / 16; ++s) {
            if (s + 1 < BK / 16) load(s + 1, av[(s + 1) & 1], bv[(s + 1) & 1]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int sm = 0; sm < SM; ++sm)
#pragma unroll
                for (int sn = 0; sn < SN; ++sn)
                    acc[sm][sn] =
                        __builtin_amdgcn_mfma_f32_32x32x16_bf16(bv[s & 1][sn], av[s & 1][sm], acc[sm][sn], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    // The matrix instructions take B's fragment as their first operand, so each
    // accumulator block is the transposed tile: lane l holds row l % 32 of C
    // (and 16 of its columns), and one element index across the wave is 32
    // consecutive rows of two columns, whole 128-byte runs of column-major C.
    template <typename F>
    __device__ __forceinline__ void for_each(int lane, int wm0, int wn0, F &&visit) const {
        const int i = lane & 31, g = lane >> 5;
#pragma unroll
        for (int sm = 0; sm < SM; ++sm)
#pragma unroll
            for (int sn = 0; sn < SN; ++sn)
#pragma unroll
                for (int e = 0; e < 16; ++e)
                    visit(wm0 + sm * 32 + i, wn0 + sn * 32 + (e >> 2) * 8 + g * 4 + (e & 3), acc[sm][sn][e]);
    }

    // as WaveF32::store_full (the same accumulator map)
    template <typename P>
    __device__ __forceinline__ void store_full(P c, int64_t ldc, int lane, int wm0, int wn0, float alpha) const {
        WaveF32<WTM, WTN, BK>::store_rows(acc, c, ldc, lane, wm0, wn0, alpha);
    }
};

template <typename T, int WTM, int WTN, int BK> struct WaveSel;
template <int WTM, int WTN, int BK> struct WaveSel<double, WTM, WTN, BK> { using type = WaveF64<WTM, WTN, BK>; };
template <int WTM, int WTN, int BK> struct WaveSel<float, WTM, WTN, BK> { using type = WaveF32<WTM, WTN, BK>; };
template <int WTM, int WTN, int BK> struct WaveSel<uint16_t, WTM, WTN, BK> { using type = WaveBF16<WTM, WTN, BK>; };

// LDS slab layout of the single-problem kernel: rows-contiguous FP64
// operands (A not transposed, B transposed) go k-major and are read in row /
// column pairs; everything else is [row][k].  STAGE = elements of one A + B
// stage.
template <typename T, bool TA, bool TB, int BM, int BN, int BK>
struct SlabLayout {
    static constexpr bool KMA = sizeof(T) == 8 && !TA;
    static constexpr bool KMB = sizeof(T) == 8 && TB;
    static constexpr int PITCH = BK + 16 / (int)sizeof(T);
    static constexpr int SA = KMA ? BK * (BM + 4) : BM * PITCH;
    static constexpr int SB = KMB ? BK * (BN + 4) : BN * PITCH;
    static constexpr int STAGE = SA + SB;
    template <int WTM, int WTN>
    using Wave = std::conditional_t<sizeof(T) == 8, WaveF64<WTM, WTN, BK, KMA ? BM + 4 : 0, KMB ? BN + 4 : 0>,
                                    typename WaveSel<T, WTM, WTN, BK>::type>;
};

}  // namespace gemm_amd
