// Host-side declarations of the grouped (many-problem) GEMM.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gemm_amd {

enum GroupedDtype : int { GROUPED_F32 = 0, GROUPED_BF16 = 1, GROUPED_F64 = 2 };

// Default smallest output dimension of a problem on the large tile class,
// per dtype, from the grouped autotuner on the 512-problem workload (M, N, K
// log-uniform over [32, 2048]): bf16 96; FP64 128; FP32 none (its 128x128
// tiles gather rows-contiguous operands k-wise and measured slower than
// keeping every problem on small tiles at higher occupancy).
constexpr int GROUPED_NO_LARGE = 1 << 30;
inline int grouped_default_large_min(int dtype) {
    return dtype == GROUPED_BF16 ? 96 : dtype == GROUPED_F64 ? 128 : GROUPED_NO_LARGE;
}

// Builds the device problem table (problems sorted by work, tiles laid out
// back to back) and prepares the persistent kernel; returns a plan or null
// (status = hipError_t).  large_min: smallest output dimension of a problem
// on the large tile class (<= 0: grouped_default_large_min(dtype)).  Pointers are device pointers; bf16 operands are
// uint16 bit patterns with FP32 C.
void *grouped_plan_create(int dtype, int count, const char *ta, const char *tb, const int *m, const int *n,
                          const int *k, const double *alpha, const void *const *A, const int *lda,
                          const void *const *B, const int *ldb, const double *beta, void *const *C, const int *ldc,
                          int cus, int large_min, int *status);
int grouped_plan_launch(void *plan, hipStream_t s);    // one persistent launch
int grouped_plan_capture(void *plan, hipStream_t s);   // capture that launch into a hipGraph (once)
int grouped_plan_replay(void *plan, hipStream_t s);    // hipGraphLaunch of the captured graph
int grouped_plan_destroy(void *plan);
int64_t grouped_plan_tiles(void *plan);
// Before capture: dynamic != 0 claims tiles from an atomic counter instead of
// the static stride; timeline (device, 3 x grid words, or null) receives each
// workgroup's start and end time (100 MHz counter) and tile count.
int grouped_plan_configure(void *plan, int dynamic, unsigned long long *timeline);
int grouped_plan_grid(void *plan);   // workgroups of the persistent launch

}  // namespace gemm_amd
