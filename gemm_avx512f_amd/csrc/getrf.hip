// Blocked LU factorisation with partial pivoting (LAPACK dgetrf contract) on
// device pointers: P A = L U, column-major, ipiv 1-based.
//
// The factorisation is recursive over columns (left half, then the U12 solve
// and the Schur-complement GEMM, then the right half), so all but the leaf
// columns run in the DTRSM and FP64 matrix-core GEMM kernels of this library.
// A leaf is a panel of up to 32 columns and every row below its diagonal,
// factored by one workgroup of up to 1024 threads with the panel held in
// registers (rows strided over threads); panels of more than 16384 rows stay
// in global memory instead (getrf_leaf_global_kernel), so m is not limited.
// After each leaf a chip-wide kernel applies
// the leaf's row interchanges to every other column of the matrix, so the
// state after every leaf is that of LAPACK's unblocked dgetf2 at the same
// column and no deferred row-swap passes are needed.
#include "getrf.hpp"
#include "trsm.hpp"
#include "tuner.hpp"
#include <hip/hip_runtime.h>
#include <float.h>

namespace gemm_amd {

namespace {

constexpr int LU_THREADS = 1024;
constexpr int LU_WAVES = LU_THREADS / 64;

// Pivot reductions in two passes: the largest magnitude first, then the
// smallest row among the lanes that hold it.  That is the order (larger |a|,
// then smaller row) in which idamax takes the first maximum, so the pivots
// are those of one combined comparison; the passes take a plain maximum and
// a plain minimum per step instead of a compare-and-select of two values.
// Magnitudes are never NaN here (a NaN never becomes a candidate), and
// lanes without a candidate hold -1 and row 0x7fffffff.

// max(x, the value of the lane that the DPP pattern CTRL names), within 16-lane rows
template <int CTRL>
__device__ __forceinline__ double dpp_max(double x) {
    const long long bits = __double_as_longlong(x);
    const int lo = (int)bits, hi = (int)(bits >> 32);
    const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
    const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
    return fmax(x, __longlong_as_double((long long)(((unsigned long long)(unsigned)ohi << 32) | (unsigned)olo)));
}
template <int CTRL>
__device__ __forceinline__ int dpp_min(int x) {
    return min(x, __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false));
}
// every lane of a 16-lane row: exchanges with lane ^ 1 and lane ^ 2 (quad
// permutes), then with the mirrored lane in the half row and in the row
__device__ __forceinline__ double row_max(double x) {
    return dpp_max<0x140>(dpp_max<0x141>(dpp_max<0x4E>(dpp_max<0xB1>(x))));
}
__device__ __forceinline__ int row_min(int x) {
    return dpp_min<0x140>(dpp_min<0x141>(dpp_min<0x4E>(dpp_min<0xB1>(x))));
}
__device__ __forceinline__ double lane_double(double x, int l) {
    const long long bits = __double_as_longlong(x);
    return __longlong_as_double(
        (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(bits >> 32), l) << 32) |
                    (unsigned)__builtin_amdgcn_readlane((int)bits, l)));
}

// Every lane of a 16-lane row ends with the row's best candidate.
__device__ __forceinline__ void pivot_row_reduce(double &best, int &brow) {
    const double m = row_max(best);
    brow = row_min(best == m ? brow : 0x7fffffff);
    best = m;
}

// The wave's best candidate (every lane gets it, as wave-uniform values).
__device__ __forceinline__ void pivot_wave_reduce(double &best, int &brow) {
    const double rm = row_max(best);
    const double m = fmax(fmax(lane_double(rm, 0), lane_double(rm, 16)), fmax(lane_double(rm, 32), lane_double(rm, 48)));
    const int rr = row_min(best == m ? brow : 0x7fffffff);
    brow = min(min(__builtin_amdgcn_readlane(rr, 0), __builtin_amdgcn_readlane(rr, 16)),
               min(__builtin_amdgcn_readlane(rr, 32), __builtin_amdgcn_readlane(rr, 48)));
    best = m;
}

// Factor columns [c0, c0 + w) of A over rows [c0, m) (T rows per thread,
// row c0 + i in register row i / NT of thread i % NT, where NT = blockDim.x is
// a multiple of 64 between 64 and 1024 sized to the panel; at most W columns;
// FULL: w == W); getrf_swap_kernel then applies the w interchanges to the
// other columns.  ipiv[c] receives the 1-based pivot row
// of column c; info the first (1-based) column whose pivot is exactly zero,
// as dgetf2 reports it.
//
// One barrier per column.  Before it, every wave publishes its best pivot
// candidate together with that row's panel entries (the lane that owns the
// row writes them), and thread j publishes row c0+j.  After it, every wave
// reduces the candidates and takes the pivot row from the winning wave's
// slot, so no second exchange is needed.  Slots are double buffered by
// column parity: column j+2 rewrites column j's slots only after every wave
// has passed column j+1's barrier, that is, after it finished reading them.
// A column step costs the length of one wave's dependent instruction chain
// (every wave runs the whole step), so the step is kept short: conditions
// that hold for every row but the first few (t > 0 rows lie below any
// diagonal of the leaf) are spelled out, a full-width leaf tests no column
// bounds, LDS reads are issued together before their values are used, and
// the workgroup has no more waves than the panel's rows need.
template <int T, int W, bool FULL>
__global__ __launch_bounds__(LU_THREADS) void getrf_leaf_kernel(double *__restrict__ A, int64_t lda, int m, int c0,
                                                                int w, int *__restrict__ ipiv,
                                                                int *__restrict__ info) {
    static_assert(W <= 64, "rows c0 .. c0+W-1 are the first register rows of threads 0 .. W-1 (NT >= 64)");
    __shared__ double s_mag[2][LU_WAVES];
    __shared__ int s_row[2][LU_WAVES];
    __shared__ double s_cand[2][LU_WAVES][W];   // each wave's best candidate row (panel entries)
    __shared__ double s_crow[2][W];             // row c0+j's panel entries (move to the pivot row)
    __shared__ int s_ipiv[W];                    // pivots of the leaf, stored to ipiv after the loop (no global
                                                 // store may be outstanding at a column's barrier)

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int NT = blockDim.x, nwaves = NT >> 6;
    int zero_col = -1;   // thread 0: first column of this leaf with an exactly zero pivot
    auto incol = [&](int k) { return FULL ? k < W : k < w; };

    double v[T][W];
    bool live[T];   // the register row lies inside the matrix
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int r = c0 + tid + t * NT;
        live[t] = r < m;
#pragma unroll
        for (int k = 0; k < W; ++k) v[t][k] = (live[t] && incol(k)) ? A[r + (int64_t)(c0 + k) * lda] : 0.0;
    }

    // Unrolled over the panel width so that the column index is a constant:
    // the register panel is then addressed directly, without select chains.
#pragma clang loop unroll(full)
    for (int j = 0; j < W; ++j) {
        if (!FULL && j >= w) continue;   // w is uniform: every thread skips the same columns
        const int cj = c0 + j, b = j & 1;
        // 1. row cj (thread j's first register row); pivot candidates over the
        // thread's rows >= cj: rows ascend with t, so a strict comparison keeps
        // the first maximum of |a| (a NaN magnitude never wins)
        if (tid == j)
#pragma unroll
            for (int k = 0; k < W; ++k) s_crow[b][k] = v[0][k];
        double best = -1.0;
        int brow = 0x7fffffff;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int r = c0 + tid + t * NT;
            const double mag = fabs(v[t][j]);
            if ((t > 0 || r >= cj) && live[t] && mag > best) {
                best = mag;
                brow = r;
            }
        }
        pivot_wave_reduce(best, brow);
        // the winning row's owner publishes its entries (tested row by row:
        // a test on the register row's index would let the compiler index
        // the register panel dynamically)
#pragma unroll
        for (int t = 0; t < T; ++t)
            if (c0 + tid + t * NT == brow)
#pragma unroll
                for (int k = 0; k < W; ++k) s_cand[b][wave][k] = v[t][k];
        if (lane == 0) {
            s_mag[b][wave] = best;
            s_row[b][wave] = brow;
        }
        __syncthreads();
        // 2. every wave reduces the 16 candidates (lanes 0-15, one 16-lane row)
        static_assert(LU_WAVES == 16, "the partials fit one 16-lane row");
        const int wrow = lane < nwaves ? s_row[b][lane] : 0x7fffffff;   // lane w: wave w's candidate row
        double bm = lane < nwaves ? s_mag[b][lane] : -1.0;
        int br = wrow;
        pivot_row_reduce(bm, br);
        br = __builtin_amdgcn_readlane(br, 0);
        // no candidate compared larger (NaN column): keep the diagonal row
        const bool found = br < m;
        const int p = found ? br : cj;
        // the winning wave is the lane whose candidate row is the pivot row (rows are distinct)
        const unsigned long long won = __builtin_amdgcn_ballot_w64(found && wrow == p);
        const double *prow = found ? s_cand[b][__builtin_ctzll(won | (1ull << 63))] : s_crow[b];
        if (tid == 0) s_ipiv[j] = p + 1;
        // 3. swap, scale the column below the diagonal, rank-1 update of the panel
        const unsigned long long pbits = __builtin_bit_cast(unsigned long long, prow[j]);
        const double piv = __builtin_bit_cast(
            double, ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(pbits >> 32)) << 32) |
                        (unsigned)__builtin_amdgcn_readfirstlane((int)pbits));
        if (tid == 0 && piv == 0.0 && zero_col < 0) zero_col = cj;
        const bool recip = fabs(piv) >= DBL_MIN;
        const double rpiv = 1.0 / piv;
        // the pivot row's entries right of the diagonal, read once for all of
        // the thread's register rows (read inside each row's branch they would
        // be one LDS round trip per row)
        double pk[W];
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (T > 1 && k > j && incol(k)) pk[k] = prow[k];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int r = c0 + tid + t * NT;
            // (all reads first, then an empty asm on each value: it keeps the
            // copy register moves -- as a block copy from LDS the register
            // row would be given memory -- without waiting on each read)
            if (t == 0 && r == cj) {
                double x[W];
#pragma unroll
                for (int k = 0; k < W; ++k) x[k] = prow[k];
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    asm volatile("" : "+v"(x[k]));
                    v[0][k] = x[k];
                }
            } else if (r == p) {
                double x[W];
#pragma unroll
                for (int k = 0; k < W; ++k) x[k] = s_crow[b][k];
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    asm volatile("" : "+v"(x[k]));
                    v[t][k] = x[k];
                }
            }
            if ((t > 0 || r > cj) && live[t] && piv != 0.0) {
                const double l = recip ? v[t][j] * rpiv : v[t][j] / piv;
                v[t][j] = l;
#pragma clang loop unroll(full)
                for (int k = j + 1; k < W; ++k)
                    if (incol(k)) v[t][k] = fma(-l, T > 1 ? pk[k] : prow[k], v[t][k]);
            }
        }
    }
    __syncthreads();   // s_ipiv is complete

#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int r = c0 + tid + t * NT;
        if (live[t])
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (incol(k)) A[r + (int64_t)(c0 + k) * lda] = v[t][k];
    }
    if (tid < w) ipiv[c0 + tid] = s_ipiv[tid];
    if (tid == 0) {
        // leaves run in column order on one stream: the first leaf resets info
        if (c0 == 0) *info = zero_col >= 0 ? zero_col + 1 : 0;
        else if (zero_col >= 0 && *info == 0) *info = zero_col + 1;
    }
}

// Panels taller than the register leaves hold (more than 16384 rows below the
// diagonal): the same column steps with the panel left in global memory.  Row
// r of the panel belongs to thread (r - c0) mod 1024 for the whole leaf, so
// every panel entry is read and written by one thread only (its own program
// order keeps it coherent); what the other threads need of a column step --
// row cj and the pivot row -- passes through LDS, double buffered by column
// parity as in the register leaf.  ipiv and info as getrf_leaf_kernel.
template <int W>
__global__ __launch_bounds__(1024) void getrf_leaf_global_kernel(double *__restrict__ A, int64_t lda, int m, int c0,
                                                                 int w, int *__restrict__ ipiv,
                                                                 int *__restrict__ info) {
    constexpr int NT = 1024, NW = NT / 64;
    __shared__ double s_mag[2][NW];
    __shared__ int s_row[2][NW];
    __shared__ double s_prow[2][W];   // the pivot row's panel entries (become row c0+j)
    __shared__ double s_crow[2][W];   // row c0+j's panel entries (move to the pivot row)
    __shared__ int s_ipiv[W];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int zero_col = -1;
    for (int j = 0; j < w; ++j) {
        const int cj = c0 + j, b = j & 1;
        const double *colj = A + (int64_t)cj * lda;
        // 1. row cj (thread j owns it); pivot candidates over the thread's rows >= cj
        if (tid == j)
#pragma unroll
            for (int k = 0; k < W; ++k) s_crow[b][k] = k < w ? A[cj + (int64_t)(c0 + k) * lda] : 0.0;
        double best = -1.0;
        int brow = 0x7fffffff;
        for (int r = c0 + tid; r < m; r += NT) {
            if (r < cj) continue;
            const double mag = fabs(colj[r]);
            if (mag > best) {   // rows ascend: a strict comparison keeps the first maximum
                best = mag;
                brow = r;
            }
        }
        pivot_wave_reduce(best, brow);
        if (lane == 0) {
            s_mag[b][wave] = best;
            s_row[b][wave] = brow;
        }
        __syncthreads();
        static_assert(NW == 16, "the partials fill one 16-lane row");
        double bm = lane < NW ? s_mag[b][lane] : -1.0;
        int br = lane < NW ? s_row[b][lane] : 0x7fffffff;
        pivot_row_reduce(bm, br);
        br = __builtin_amdgcn_readlane(br, 0);
        const int p = br >= m ? cj : br;   // no candidate compared larger (NaN column): keep the diagonal row
        const int powner = (p - c0) % NT;
        if (tid == 0) s_ipiv[j] = p + 1;
        if (tid == powner)
#pragma unroll
            for (int k = 0; k < W; ++k) s_prow[b][k] = k < w ? A[p + (int64_t)(c0 + k) * lda] : 0.0;
        __syncthreads();
        // 2. swap rows cj and p (each by its owner), scale, rank-1 update of the later panel columns
        const double piv = s_prow[b][j];
        if (tid == 0 && piv == 0.0 && zero_col < 0) zero_col = cj;
        if (p != cj) {
            if (tid == j)
                for (int k = 0; k < w; ++k) A[cj + (int64_t)(c0 + k) * lda] = s_prow[b][k];
            if (tid == powner)
                for (int k = 0; k < w; ++k) A[p + (int64_t)(c0 + k) * lda] = s_crow[b][k];
        }
        if (piv != 0.0) {
            const bool recip = fabs(piv) >= DBL_MIN;
            const double rpiv = 1.0 / piv;
            for (int r = c0 + tid; r < m; r += NT) {
                if (r <= cj) continue;
                double *row = A + r;
                const double l = recip ? row[(int64_t)cj * lda] * rpiv : row[(int64_t)cj * lda] / piv;
                row[(int64_t)cj * lda] = l;
                for (int k = j + 1; k < w; ++k) row[(int64_t)(c0 + k) * lda] = fma(-l, s_prow[b][k], row[(int64_t)(c0 + k) * lda]);
            }
        }
    }
    __syncthreads();   // s_ipiv is complete
    if (tid < w) ipiv[c0 + tid] = s_ipiv[tid];
    if (tid == 0) {
        if (c0 == 0) *info = zero_col >= 0 ? zero_col + 1 : 0;
        else if (zero_col >= 0 && *info == 0) *info = zero_col + 1;
    }
}

// Applies the interchanges ipiv[c0 .. c0+w) (1-based rows, in order) to every
// column in [0, ncols) outside [c0, c0+w).  Each workgroup first reduces the
// w swaps to their net effect on the rows they touch (the value that ends at
// row dst started at row src; at most 2w rows move), then 32 or 64 lanes (one
// per move) move one column: every lane of a column loads its source before
// any lane of that column stores (one instruction each, one wave per column
// group).
constexpr int SWAP_THREADS = 256;
// lanes that move one column: one per row move (at most 2w), 32 or 64
constexpr int swap_lanes_per_column(int w) { return 2 * w > 32 ? 64 : 32; }

template <int W>
__global__ __launch_bounds__(SWAP_THREADS) void getrf_swap_kernel(double *__restrict__ A, int64_t lda, int ncols,
                                                                  int c0, int w, const int *__restrict__ ipiv) {
    static_assert(2 * W <= 64, "one wave holds the slots");
    __shared__ int s_dst[2 * W], s_src[2 * W];
    __shared__ int s_npairs;
    const int tid = threadIdx.x;
    if (tid < 64) {
        // Slots 0..w-1 are rows c0..c0+w-1; pivot rows below the block get
        // slots w.. in order of first appearance.  Lane i tracks which
        // original row's value currently sits in slot i.
        int prow[W], pslot[W];
        int next = w;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            prow[j] = j < w ? ipiv[c0 + j] - 1 : c0;
            int sl = prow[j] - c0;
            if (j < w && sl >= w) {
                sl = -1;
#pragma unroll
                for (int q = 0; q < W; ++q)
                    if (q < j && sl < 0 && prow[q] == prow[j]) sl = pslot[q];
                if (sl < 0) sl = next++;
            }
            pslot[j] = sl;
        }
        const int lane = tid;
        int row = lane < w ? c0 + lane : -1;
#pragma unroll
        for (int j = 0; j < W; ++j)
            if (j < w && pslot[j] == lane && lane >= w) row = prow[j];
        int held = row;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (j < w) {
                const int va = __shfl(held, j), vb = __shfl(held, pslot[j]);
                if (lane == j) held = vb;
                if (lane == pslot[j]) held = va;
            }
        }
        const bool moved = row >= 0 && held != row;
        const unsigned long long mask = __ballot(moved);
        if (moved) {
            const int idx = __popcll(mask & ((1ull << lane) - 1ull));
            s_dst[idx] = row;
            s_src[idx] = held;
        }
        if (lane == 0) s_npairs = __popcll(mask);
    }
    __syncthreads();
    constexpr int LPC = swap_lanes_per_column(W);
    const int np = s_npairs, pair = tid % LPC;
    const int c = blockIdx.x * (SWAP_THREADS / LPC) + tid / LPC;
    if (pair >= np || c >= ncols || (c >= c0 && c < c0 + w)) return;
    double *col = A + (int64_t)c * lda;
    const double x = col[s_src[pair]];
    col[s_dst[pair]] = x;
}

// X <- L^-1 X for a unit lower n1 x n1 triangle L (n1 <= NMAX) and an
// n1 x n2 block X, by forward substitution: one thread per column of X, the
// column in registers, L staged in LDS and read as broadcasts.
template <int NMAX>
__global__ __launch_bounds__(64) void getrf_trsm_small_kernel(const double *__restrict__ L, int64_t lda,
                                                              double *__restrict__ X, int n1, int n2) {
    __shared__ double s_l[NMAX * NMAX];   // column-major, leading dimension NMAX
    const int tid = threadIdx.x;
    for (int e = tid; e < n1 * n1; e += 64) {
        const int i = e % n1, k = e / n1;
        s_l[i + k * NMAX] = L[i + (int64_t)k * lda];
    }
    __syncthreads();
    const int c = blockIdx.x * 64 + tid;
    if (c >= n2) return;
    double *col = X + (int64_t)c * lda;
    double x[NMAX];
#pragma unroll
    for (int i = 0; i < NMAX; ++i) x[i] = i < n1 ? col[i] : 0.0;
#pragma unroll
    for (int k = 0; k < NMAX - 1; ++k) {
#pragma unroll
        for (int i = k + 1; i < NMAX; ++i)
            if (i < n1) x[i] = fma(-s_l[i + k * NMAX], x[k], x[i]);
    }
#pragma unroll
    for (int i = 0; i < NMAX; ++i)
        if (i < n1) col[i] = x[i];
}

// T > 0: the register leaf (T rows per thread); T == 0: the global-memory leaf.
template <int T, int W>
int leaf(double *A, int64_t lda, int m, int ncols, int c0, int w, int *ipiv, int *info, hipStream_t s) {
    if constexpr (T > 0) {
        // as many waves as T rows per thread need (at least one)
        const int per = (m - c0 + T - 1) / T;
        const int nt = per <= 64 ? 64 : per >= LU_THREADS ? LU_THREADS : (per + 63) / 64 * 64;
        if (w == W)
            hipLaunchKernelGGL((getrf_leaf_kernel<T, W, true>), dim3(1), dim3(nt), 0, s, A, lda, m, c0, w, ipiv, info);
        else
            hipLaunchKernelGGL((getrf_leaf_kernel<T, W, false>), dim3(1), dim3(nt), 0, s, A, lda, m, c0, w, ipiv, info);
    } else
        hipLaunchKernelGGL((getrf_leaf_global_kernel<W>), dim3(1), dim3(1024), 0, s, A, lda, m, c0, w, ipiv, info);
    if (w > 0 && ncols > w) {
        constexpr int cols_per_block = SWAP_THREADS / swap_lanes_per_column(W);
        const int grid = (ncols + cols_per_block - 1) / cols_per_block;
        hipLaunchKernelGGL((getrf_swap_kernel<W>), dim3(grid), dim3(SWAP_THREADS), 0, s, A, lda, ncols, c0, w, ipiv);
    }
    return (int)hipGetLastError();
}

struct Lu {
    double *A;
    int64_t lda;
    int m, n;
    int *ipiv, *info;
    double *inv_ws;
    void *gemm_ws;
    size_t gemm_ws_bytes;
    int cus;
    hipStream_t s;
    int err = 0;

    double *at(int r, int c) const { return A + r + (int64_t)c * lda; }

    // widest leaf the register budget allows for a panel of `rows` rows; past
    // GETRF_REGISTER_ROWS the panel stays in global memory, 8 columns a leaf
    static int leaf_width(int rows) {
        return rows <= LU_THREADS ? 32 : rows <= 2 * LU_THREADS ? 16 : rows <= 4 * LU_THREADS ? 8
               : rows <= 8 * LU_THREADS ? 4 : rows <= GETRF_REGISTER_ROWS ? 2 : 8;
    }

    void run_leaf(int c0, int w) {
        if (err) return;
        const int rows = m - c0;
        if (rows <= LU_THREADS) err = leaf<1, 32>(A, lda, m, n, c0, w, ipiv, info, s);
        else if (rows <= 2 * LU_THREADS) err = leaf<2, 16>(A, lda, m, n, c0, w, ipiv, info, s);
        else if (rows <= 4 * LU_THREADS) err = leaf<4, 8>(A, lda, m, n, c0, w, ipiv, info, s);
        else if (rows <= 8 * LU_THREADS) err = leaf<8, 4>(A, lda, m, n, c0, w, ipiv, info, s);
        else if (rows <= GETRF_REGISTER_ROWS) err = leaf<16, 2>(A, lda, m, n, c0, w, ipiv, info, s);
        else err = leaf<0, 8>(A, lda, m, n, c0, w, ipiv, info, s);
    }

    void gemm(int mm, int nn, int kk, const double *a, const double *b, double *c) {
        if (err || mm <= 0 || nn <= 0 || kk <= 0) return;
        GemmConfig cfg = pick_gemm_config(DT_F64, false, false, mm, nn, kk, (int)(lda - mm), (int)(lda - kk), cus);
        if (gemm_workspace_bytes(mm, nn, kk, cfg, 8) > gemm_ws_bytes) cfg.splitk = 1;
        err = launch_dgemm(false, false, mm, nn, kk, -1.0, a, lda, b, lda, 1.0, c, lda, cfg, gemm_ws, s);
    }

    // L11^-1 (unit lower, n1 x n1 at (c0, c0)) applied to the n1 x nn block at (c0, c1)
    void trsm(int c0, int n1, int c1, int nn) {
        if (err || nn <= 0) return;
        if (n1 <= 16) {
            hipLaunchKernelGGL(getrf_trsm_small_kernel<16>, dim3((nn + 63) / 64), dim3(64), 0, s, at(c0, c0), lda,
                               at(c0, c1), n1, nn);
            err = (int)hipGetLastError();
            return;
        }
        if (n1 <= 32) {   // two substitutions around one GEMM beat the blocked solve's inverse pass here
            const int h = n1 / 2;
            trsm(c0, h, c1, nn);
            gemm(n1 - h, nn, h, at(c0 + h, c0), at(c0, c1), at(c0 + h, c1));
            trsm(c0 + h, n1 - h, c1, nn);
            return;
        }
        err = launch_dtrsm(true, false, false, true, n1, nn, 1.0, at(c0, c0), lda, at(c0, c1), lda, inv_ws, gemm_ws,
                           gemm_ws_bytes, cus, s);
    }

    // Factor columns [c0, c0 + nw) over rows [c0, m); columns right of the
    // range are updated by the caller.
    void rec(int c0, int nw) {
        if (err || nw <= 0) return;
        if (nw <= leaf_width(m - c0)) {
            run_leaf(c0, nw);
            return;
        }
        const int n1 = nw / 2, n2 = nw - n1;
        rec(c0, n1);
        trsm(c0, n1, c0 + n1, n2);
        gemm(m - c0 - n1, n2, n1, at(c0 + n1, c0), at(c0, c0 + n1), at(c0 + n1, c0 + n1));
        rec(c0 + n1, n2);
    }
};

}  // namespace

size_t getrf_gemm_workspace_bytes(int m, int n) {
    // split-K partials of the widest Schur update, at most 8 partitions of an
    // m x min(m, n)/2 block (the recursion never updates a wider one), capped
    // at 512 MiB: a GEMM whose partials do not fit runs unsplit (large
    // updates fill the chip without splitting K)
    const int mn = m < n ? m : n;
    const size_t want = (size_t)8 * (size_t)(m > 0 ? m : 1) * (size_t)((mn + 1) / 2 + 1) * sizeof(double);
    const size_t cap = (size_t)512 << 20;
    return want < cap ? want : cap;
}

int launch_dgetrf(int m, int n, double *A, int64_t lda, int *ipiv, int *info, double *inv_ws, void *gemm_ws,
                  size_t gemm_ws_bytes, int cus, hipStream_t s) {
    Lu lu{A, lda, m, n, ipiv, info, inv_ws, gemm_ws, gemm_ws_bytes, cus, s};
    const int mn = m < n ? m : n;
    if (mn == 0) lu.run_leaf(0, 0);   // writes info = 0 and nothing else
    lu.rec(0, mn);
    // columns right of a wide matrix's square part: U12 = L11^-1 A12
    if (n > m) lu.trsm(0, m, m, n - m);
    return lu.err;
}

}  // namespace gemm_amd
