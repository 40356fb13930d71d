// Seeded synthetic inputs generated in HBM: the splitmix64 stream used for
// every benchmark and parity input, made random-access.  Draw number t
// (t = 0, 1, ...) of the stream seeded with s is mix(s + (t+1)*gamma), so an
// element's value depends only on its position in the column-major fill
// order; a whole matrix fills in one pass at HBM bandwidth.
#include "common.hpp"
#include "../../include/gemm_amd.h"
#include "runtime.hpp"

namespace gemm_amd {

__device__ __forceinline__ uint64_t splitmix_draw(uint64_t seed, uint64_t t) {
    uint64_t z = seed + (t + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// kind 0: FP64 ((x>>11)*2^-53)*2-1; kind 1: FP32 ((x>>40)*2^-24)*2-1;
// kind 2: FP64 u*2^e with u from draw 2t and e = -r + draw(2t+1) mod (2r+1).
template <int KIND>
__global__ __launch_bounds__(256) void fill_kernel(void *out, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                                                   uint64_t first, int r) {
    const int64_t total = rows * cols;
    for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int64_t i = idx % rows, j = idx / rows;
        if (KIND == 0) {
            const uint64_t x = splitmix_draw(seed, first + idx);
            ((double *)out)[i + j * ld] = (double)(x >> 11) * 0x1.0p-53 * 2.0 - 1.0;
        } else if (KIND == 1) {
            const uint64_t x = splitmix_draw(seed, first + idx);
            ((float *)out)[i + j * ld] = (float)(x >> 40) * 0x1.0p-24f * 2.0f - 1.0f;
        } else {
            const uint64_t x = splitmix_draw(seed, first + 2 * idx);
            double u = (double)(x >> 11) * 0x1.0p-53 * 2.0 - 1.0;
            int e = 0;
            if (r > 0) e = -r + (int)(splitmix_draw(seed, first + 2 * idx + 1) % (uint64_t)(2 * r + 1));
            ((double *)out)[i + j * ld] = ldexp(u, e);
        }
    }
}

}  // namespace gemm_amd

using namespace gemm_amd;

extern "C" int gemm_amd_fill_uniform(void *dptr, int kind, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                                     uint64_t first_draw, int range_r, void *stream) {
    if (kind < 0 || kind > 2 || rows < 0 || cols < 0 || ld < (rows > 1 ? rows : 1) || range_r < 0) {
        set_error("gemm_amd_fill_uniform: bad arguments");
        return -(int)hipErrorInvalidValue;
    }
    if (rows == 0 || cols == 0) return 0;
    if (int st = init_status()) return st;
    const int64_t total = rows * cols;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipStream_t s = (hipStream_t)stream;
    if (kind == 0)
        hipLaunchKernelGGL(fill_kernel<0>, dim3(blocks), dim3(256), 0, s, dptr, rows, cols, ld, seed, first_draw, 0);
    else if (kind == 1)
        hipLaunchKernelGGL(fill_kernel<1>, dim3(blocks), dim3(256), 0, s, dptr, rows, cols, ld, seed, first_draw, 0);
    else
        hipLaunchKernelGGL(fill_kernel<2>, dim3(blocks), dim3(256), 0, s, dptr, rows, cols, ld, seed, first_draw,
                           range_r);
    return hip_status(hipGetLastError(), "fill_kernel");
}
