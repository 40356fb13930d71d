// Host-side declarations of the grouped (many-problem) GEMM.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gemm_amd {

enum GroupedDtype : int { GROUPED_F32 = 0, GROUPED_BF16 = 1, GROUPED_F64 = 2 };

// Builds the device problem table (problems sorted by work, tiles laid out
// back to back) and prepares the persistent kernel; returns a plan or null
// (status = hipError_t).  Pointers are device pointers; bf16 operands are
// uint16 bit patterns with FP32 C.
void *grouped_plan_create(int dtype, int count, const char *ta, const char *tb, const int *m, const int *n,
                          const int *k, const double *alpha, const void *const *A, const int *lda,
                          const void *const *B, const int *ldb, const double *beta, void *const *C, const int *ldc,
                          int cus, int *status);
int grouped_plan_launch(void *plan, hipStream_t s);    // one persistent launch
int grouped_plan_capture(void *plan, hipStream_t s);   // capture that launch into a hipGraph (once)
int grouped_plan_replay(void *plan, hipStream_t s);    // hipGraphLaunch of the captured graph
int grouped_plan_destroy(void *plan);
int64_t grouped_plan_tiles(void *plan);

}  // namespace gemm_amd
