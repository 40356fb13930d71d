// Shared device-side definitions for the gfx950 BLAS-3 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef double d2_t __attribute__((ext_vector_type(2)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef float f16_t __attribute__((ext_vector_type(16)));
typedef int i4_t __attribute__((ext_vector_type(4)));
typedef int i16_t __attribute__((ext_vector_type(16)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

template <typename T> struct Vec16;  // 16-byte vector of T
template <> struct Vec16<double> { using t = d2_t; static constexpr int N = 2; };
template <> struct Vec16<float> { using t = f4_t; static constexpr int N = 4; };
typedef unsigned short us8_t __attribute__((ext_vector_type(8)));
template <> struct Vec16<uint16_t> { using t = us8_t; static constexpr int N = 8; };   // bf16 bit patterns

// Lane layouts of the gfx950 matrix instructions used here (each one was
// checked on hardware by one-hot / random-operand probes):
//  v_mfma_f64_4x4x4_4b_f64: 4 independent blocks b of D_b(4x4) += A_b(4x4) B_b(4x4);
//     A_b[i][k] at lane 4b+i+16k, B_b[k][j] at lane 4b+j+16k, D_b[i][j] at lane 16i+4b+j.
//     (the 16x16x4 f64 form issues at ~60% of the FP64 peak on gfx950, this one at ~92%)
//  v_mfma_f64_16x16x4_f64: A[i][k] at lane i+16k, B[k][j] at lane j+16k,
//     D[i][j] at lane j+16*(i%4), element i/4.
//  v_mfma_f32_32x32x2_f32: A[i][k] lane i+32k, B[k][j] lane j+32k,
//     D[i][j] lane j+32*((i/4)%2), element (i/8)*4 + i%4.
//  v_mfma_f32_32x32x16_bf16 / v_mfma_i32_32x32x32_i8: lane l supplies 8 (bf16) or
//     16 (i8) consecutive-k values of row i = l%32 (A) / column j = l%32 (B) for
//     k-group l/32; D as for 32x32x2.
__device__ __forceinline__ double mfma_f64_4x4x4(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f16_t mfma_f32_32x32x2(float a, float b, f16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// The GEMM epilogue alpha*v + beta*c, spelled with one explicit fused
// multiply-add so that every kernel rounds it identically (beta == 0 does not
// read c).
// c points at C's element in any address space.
template <typename T, typename P>
__device__ __forceinline__ T axpby(T alpha, T v, T beta, P c) {
    if (beta == T(0)) return alpha * v;
    const T cv = *c;
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(beta, cv, alpha * v);
    else return __builtin_fma(beta, cv, alpha * v);
}

// The same pointer in the global address space.  Accesses through pointers
// read from memory (a descriptor table) otherwise compile to flat
// instructions, which count against both the vector-memory and the LDS
// counters and so turn the counted waits of a load pipeline into full drains.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *as_global(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

// Workgroup barrier that orders LDS only.  __syncthreads() fences every
// address space, so it drains the wave's outstanding global loads as well
// (s_waitcnt vmcnt(0) before s_barrier); fenced on LDS alone, the barrier
// waits for the wave's LDS accesses (lgkmcnt) and leaves loads that are
// prefetching later slabs in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Map a linear workgroup id to a tile id so that each of the 8 XCDs (which
// receive workgroups round-robin: wg i runs on XCD i % 8) owns a contiguous
// run of tiles, then order tiles in groups of `group` tile-rows so that
// neighbouring tiles on one XCD share A row-panels and B column-panels in
// that XCD's L2.
__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int group,
                                            int &tm, int &tn) {
    const int ntiles = tiles_m * tiles_n;
    int t = bid;
    if ((ntiles & 7) == 0) t = (bid & 7) * (ntiles >> 3) + (bid >> 3);
    const int per_group = group * tiles_n;
    const int gid = t / per_group;
    const int first_m = gid * group;
    int gsize = tiles_m - first_m;
    if (gsize > group) gsize = group;
    const int r = t - gid * per_group;
    tm = first_m + r % gsize;
    tn = r / gsize;
}

// Raises the dynamic-LDS limit of the kernel KERN to `bytes`: one
// hipFuncSetAttribute call per process, made by whichever thread launches the
// kernel first; every caller gets that call's status.
template <auto KERN>
inline hipError_t lds_limit_once(size_t bytes) {
    static std::once_flag once;
    static hipError_t status = hipSuccess;
    std::call_once(once, [bytes] {
        status = hipFuncSetAttribute(reinterpret_cast<const void *>(KERN), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)bytes);
    });
    return status;
}
