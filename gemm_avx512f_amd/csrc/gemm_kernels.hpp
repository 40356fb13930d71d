// Host-side declarations of the GEMM kernel launchers (gfx950 MFMA kernels).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gemm_amd {

// Macro-tile shapes the FP32/FP64 kernels are instantiated for (BM x BN).
enum TileId : int {
    TILE_128x128 = 0,   // 4 waves (2x2), 64x64 per wave
    TILE_64x64 = 1,     // 4 waves (2x2), 32x32 per wave
    TILE_128x64 = 2,    // 4 waves (2x2), 64x32 per wave
    TILE_LARGE = 3,     // FP64: 256x64, 4 waves (4x1); FP32: 256x128, 8 waves (4x2); 64x64 per wave
    TILE_COUNT = 4
};

struct GemmConfig {
    int tile = TILE_128x128;
    int splitk = 1;     // number of K partitions (>1: partial sums + reduction pass)
    // 64x64 (and FP64 128x64) tiles may split each slab's k over groups of
    // waves when the grid has at most one workgroup per CU (a different k order:
    // set only where the library chooses the configuration, so an explicit
    // tile and split-K always name one kernel)
    bool kgroups = false;
};

int tile_bm(int tile, int elem_bytes);
int tile_bn(int tile, int elem_bytes);
int tile_bk(int tile, int elem_bytes);

// Number of K partitions a split-K request of `splitk` runs as: each
// partition is a whole number of 32-deep slabs, so for shallow K there may be
// fewer than asked.
int gemm_k_partitions(int k, int splitk);

// Workspace bytes a config needs (split-K partials), 0 if none.
size_t gemm_workspace_bytes(int m, int n, int k, const GemmConfig &cfg, int elem_bytes);

// C <- alpha*op(A)*op(B) + beta*C on device pointers, column-major.
// Returns 0 or a hipError_t value.  beta == 0 never reads C.
int launch_dgemm(bool ta, bool tb, int m, int n, int k, double alpha,
                 const double *A, int64_t lda, const double *B, int64_t ldb,
                 double beta, double *C, int64_t ldc, const GemmConfig &cfg,
                 void *work, hipStream_t s);
int launch_sgemm(bool ta, bool tb, int m, int n, int k, float alpha,
                 const float *A, int64_t lda, const float *B, int64_t ldb,
                 float beta, float *C, int64_t ldc, const GemmConfig &cfg,
                 void *work, hipStream_t s);

// C <- beta*C (beta == 0 writes zeros without reading) on an m x n column-major block.
int launch_scale_d(int m, int n, double beta, double *C, int64_t ldc, hipStream_t s);
int launch_scale_s(int m, int n, float beta, float *C, int64_t ldc, hipStream_t s);

}  // namespace gemm_amd
