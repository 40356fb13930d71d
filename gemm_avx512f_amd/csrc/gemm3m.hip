// Complex GEMM by the 3M method (cgemm3m_/zgemm3m_) on gfx950.
//
// The reference's 3M kernels compute one real product T of packed real
// planes and accumulate C += (alpha_r + i alpha_i) * T into interleaved
// complex C (cgemm3m_kernel_16x4_skylakex.c:41-58, 416-440;
// zgemm3m_kernel_2x8_skylakex.c:336-351); OpenBLAS's gemm3m driver (not in the
// reference) calls them three times on the planes Ar, Ai, Ar+Ai.  Here:
//   1. one pass splits op(A) and op(B) (transpose / conjugate applied) into the
//      real planes Xr, Xi, Xr+Xi, column-major without padding;
//   2. three real GEMMs on the FP32/FP64 matrix-core kernel:
//      T1 = Ar Br, T2 = Ai Bi, T3 = (Ar+Ai)(Br+Bi);
//   3. one pass combines Re = T1 - T2, Im = T3 - T1 - T2 and applies the
//      complex alpha and beta (beta == 0 does not read C).
#include "common.hpp"
#include "gemm_kernels.hpp"
#include "gemm3m.hpp"

namespace gemm_amd {

// planes[0..2] of op(X) (rows x cols): re, im, re+im, each prows x pcols
// (ld prows) with zeros past rows and cols.  op = N, T or C (conjugate
// transpose).  X is interleaved complex with leading dim ldx.
// The element loops are shared by the one-problem kernels and the batched
// ones (one launch over a job table, blockIdx.y = job); XP/PP/TP/CP are the
// pointer types (global-address-space ones for pointers read from a table).
template <typename T, typename XP, typename PP>
__device__ __forceinline__ void split_range(int rows, int cols, int prows, int pcols, XP X, int64_t ldx, bool trans,
                                            bool conj, PP planes) {
    const int64_t total = (int64_t)prows * pcols;
    for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int64_t r = idx % prows, c = idx / prows;
        T re = T(0), im = T(0);
        if (r < rows && c < cols) {
            const int64_t src = trans ? (c + r * ldx) : (r + c * ldx);
            re = X[2 * src];
            im = X[2 * src + 1];
            if (conj) im = -im;
        }
        planes[idx] = re;
        planes[idx + total] = im;
        planes[idx + 2 * total] = re + im;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void split_planes_kernel(int rows, int cols, int prows, int pcols,
                                                           const T *__restrict__ X, int64_t ldx, bool trans, bool conj,
                                                           T *__restrict__ planes) {
    split_range<T>(rows, cols, prows, pcols, X, ldx, trans, conj, planes);
}

// t: the three real products, each mt x nt (ld mt, mt >= m, nt >= n).
template <typename T, typename TP, typename CP>
__device__ __forceinline__ void combine_range(int m, int n, int mt, int nt, TP t, T ar, T ai, T br, T bi, CP C,
                                              int64_t ldc) {
    const int64_t total = (int64_t)m * n, plane = (int64_t)mt * nt;
    for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int64_t i = idx % m, j = idx / m, ti = i + j * mt;
        const T t1 = t[ti], t2 = t[ti + plane], t3 = t[ti + 2 * plane];
        const T re = t1 - t2, im = t3 - t1 - t2;
        const auto c = C + 2 * (i + j * ldc);
        T cr = ar * re - ai * im, ci = ar * im + ai * re;
        if (br != T(0) || bi != T(0)) {
            const T xr = c[0], xi = c[1];
            cr += br * xr - bi * xi;
            ci += br * xi + bi * xr;
        }
        c[0] = cr;
        c[1] = ci;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void combine3m_kernel(int m, int n, int mt, int nt, const T *__restrict__ t, T ar,
                                                        T ai, T br, T bi, T *__restrict__ C, int64_t ldc) {
    combine_range<T>(m, n, mt, nt, t, ar, ai, br, bi, C, ldc);
}

__global__ __launch_bounds__(256) void split_planes_batch_kernel(const SplitJob3m *__restrict__ jobs) {
    const SplitJob3m &j = jobs[blockIdx.y];
    split_range<double>(j.rows, j.cols, j.prows, j.pcols, as_global(j.X), j.ldx, j.trans != 0, j.conj != 0,
                        as_global(j.planes));
}

__global__ __launch_bounds__(256) void combine3m_batch_kernel(const CombineJob3m *__restrict__ jobs) {
    const CombineJob3m &j = jobs[blockIdx.y];
    combine_range<double>(j.m, j.n, j.mt, j.nt, as_global(j.t), j.ar, j.ai, j.br, j.bi, as_global(j.C), j.ldc);
}

template <typename T>
__global__ __launch_bounds__(256) void cscale_kernel(int m, int n, T br, T bi, T *__restrict__ C, int64_t ldc) {
    const int64_t total = (int64_t)m * n;
    for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int64_t i = idx % m, j = idx / m;
        T *c = C + 2 * (i + j * ldc);
        if (br == T(0) && bi == T(0)) {
            c[0] = T(0);
            c[1] = T(0);
        } else {
            const T xr = c[0], xi = c[1];
            c[0] = br * xr - bi * xi;
            c[1] = br * xi + bi * xr;
        }
    }
}

static int grid_of(int64_t total) {
    int64_t b = (total + 255) / 256;
    return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

size_t gemm3m_workspace_bytes(int m, int n, int k, int elem_bytes) {
    return (size_t)elem_bytes * 3 * ((size_t)m * k + (size_t)k * n + (size_t)m * n);
}

template <typename T>
static int gemm3m_t(char ta, char tb, int m, int n, int k, const T *alpha, const T *A, int64_t lda, const T *B,
                    int64_t ldb, const T *beta, T *C, int64_t ldc, void *ws, const GemmConfig &cfg, void *gemm_ws,
                    hipStream_t s) {
    if (m <= 0 || n <= 0) return 0;
    const T ar = alpha[0], ai = alpha[1], br = beta[0], bi = beta[1];
    if (k <= 0 || (ar == T(0) && ai == T(0))) {
        if (br == T(1) && bi == T(0)) return 0;
        hipLaunchKernelGGL(cscale_kernel<T>, dim3(grid_of((int64_t)m * n)), dim3(256), 0, s, m, n, br, bi, C, ldc);
        return (int)hipGetLastError();
    }
    T *pa = static_cast<T *>(ws);
    T *pb = pa + 3 * (size_t)m * k;
    T *pt = pb + 3 * (size_t)k * n;
    const bool tA = !(ta == 'N' || ta == 'n'), cA = (ta == 'C' || ta == 'c');
    const bool tB = !(tb == 'N' || tb == 'n'), cB = (tb == 'C' || tb == 'c');
    hipLaunchKernelGGL(split_planes_kernel<T>, dim3(grid_of((int64_t)m * k)), dim3(256), 0, s, m, k, m, k, A, lda, tA,
                       cA, pa);
    hipLaunchKernelGGL(split_planes_kernel<T>, dim3(grid_of((int64_t)k * n)), dim3(256), 0, s, k, n, k, n, B, ldb, tB,
                       cB, pb);
    if (hipError_t e = hipGetLastError()) return (int)e;
    const size_t mk = (size_t)m * k, kn = (size_t)k * n, mn = (size_t)m * n;
    for (int p = 0; p < 3; ++p) {
        int e;
        if constexpr (sizeof(T) == 8)
            e = launch_dgemm(false, false, m, n, k, 1.0, pa + p * mk, m, pb + p * kn, k, 0.0, pt + p * mn, m, cfg,
                             gemm_ws, s);
        else
            e = launch_sgemm(false, false, m, n, k, 1.0f, pa + p * mk, m, pb + p * kn, k, 0.0f, pt + p * mn, m, cfg,
                             gemm_ws, s);
        if (e) return e;
    }
    hipLaunchKernelGGL(combine3m_kernel<T>, dim3(grid_of((int64_t)m * n)), dim3(256), 0, s, m, n, m, n, pt, ar, ai, br,
                       bi, C, ldc);
    return (int)hipGetLastError();
}

int launch_zgemm3m(char ta, char tb, int m, int n, int k, const double *alpha, const double *A, int64_t lda,
                   const double *B, int64_t ldb, const double *beta, double *C, int64_t ldc, void *ws, const GemmConfig &cfg,
                   void *gemm_ws, hipStream_t s) {
    return gemm3m_t<double>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ws, cfg, gemm_ws, s);
}

int launch_cgemm3m(char ta, char tb, int m, int n, int k, const float *alpha, const float *A, int64_t lda,
                   const float *B, int64_t ldb, const float *beta, float *C, int64_t ldc, void *ws, const GemmConfig &cfg,
                   void *gemm_ws, hipStream_t s) {
    return gemm3m_t<float>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ws, cfg, gemm_ws, s);
}

// Pieces of the 3M method for callers that batch the three real products
// themselves (the grouped zgemm3m batch): one launch splits every listed
// operand into its zero-padded planes, one launch combines every problem's
// products (job tables in device memory; 64 workgroups per job).
int launch_gemm3m_split_batch_d(const SplitJob3m *jobs, int njobs, hipStream_t s) {
    for (int j0 = 0; j0 < njobs; j0 += 65535) {
        const int nj = njobs - j0 < 65535 ? njobs - j0 : 65535;
        hipLaunchKernelGGL(split_planes_batch_kernel, dim3(64, nj), dim3(256), 0, s, jobs + j0);
    }
    return (int)hipGetLastError();
}

int launch_gemm3m_combine_batch_d(const CombineJob3m *jobs, int njobs, hipStream_t s) {
    for (int j0 = 0; j0 < njobs; j0 += 65535) {
        const int nj = njobs - j0 < 65535 ? njobs - j0 : 65535;
        hipLaunchKernelGGL(combine3m_batch_kernel, dim3(64, nj), dim3(256), 0, s, jobs + j0);
    }
    return (int)hipGetLastError();
}

}  // namespace gemm_amd
