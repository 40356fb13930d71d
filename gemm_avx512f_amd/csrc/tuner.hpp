// Kernel-configuration picking: the runtime replacement of the reference's
// compile-time tile macros (sgemm_tune.h:6-15, dgemm_tune.h:6-15,
// src/gemm_set_parameters.h:66-82).
//
// Three sources, in order of precedence:
//   1. a forced config from the environment (GEMM_AMD_FORCE_TILE / _SPLITK);
//   2. the exhaustive-sweep table: exact shapes whose 20 candidate configs
//      were all timed (entries installed by gemm_amd_tune_table_set);
//   3. the learned picker (a small perceptron run by a HIP inference kernel;
//      predictions are cached per shape) when weights are loaded and the mode
//      allows it, else the occupancy heuristic.
#pragma once
#include "gemm_kernels.hpp"

namespace gemm_amd {

enum DType : int { DT_F32 = 0, DT_F64 = 1, DT_C32 = 2, DT_C64 = 3, DT_BF16 = 4 };

// Tuning operations (the picker's op one-hot): SGEMM, DGEMM, CGEMM3M.
enum TuneOp : int { OP_SGEMM = 0, OP_DGEMM = 1, OP_CGEMM3M = 2, OP_COUNT = 3 };
int dtype_to_op(int dtype);   // -1 for dtypes the picker does not cover

// The candidate set: every tile shape x split-K in {1, 2, 4, 8, 16}.
constexpr int TUNE_NSPLIT = 5;
constexpr int TUNE_NCFG = TILE_COUNT * TUNE_NSPLIT;
GemmConfig tune_candidate(int i);
int tune_index(const GemmConfig &c);   // -1 if c is not a candidate
// A candidate is valid when each K partition is at least 64 deep and the
// split-K partials stay within 1 GiB.
bool tune_valid(int op, int m, int n, int k, const GemmConfig &c);

enum PickMode : int { PICK_HEURISTIC = 0, PICK_MLP = 1 };

// lda_pad, ldb_pad: leading dimension minus the rows of op(A), op(B) storage.
GemmConfig pick_gemm_config(int dtype, bool ta, bool tb, int m, int n, int k, int lda_pad, int ldb_pad, int cus);

// Whether the GEMM kernels take their 16-byte vector-load path for this
// problem (both leading dimensions a multiple of 16 bytes; operands from
// hipMalloc are aligned).  The CGEMM3M products read packed planes
// (leading dimensions M and K).
bool gemm_vec_path(int op, bool ta, bool tb, int m, int n, int k, int lda_pad, int ldb_pad);

// While one lives on a thread, pick_gemm_config on that thread does no device
// work: a shape that neither the sweep table nor the picker's cache holds gets
// the heuristic's config instead of a network inference (whose first use
// allocates the picker's device buffers).  For callers that must not allocate.
class PickWithoutDevice {
public:
    PickWithoutDevice();
    ~PickWithoutDevice();
    PickWithoutDevice(const PickWithoutDevice &) = delete;
    PickWithoutDevice &operator=(const PickWithoutDevice &) = delete;

private:
    bool prev_;
};
GemmConfig heuristic_gemm_config(int dtype, int m, int n, int k, int cus);

// Learned picker control (thread-safe).
int picker_load(const float *params, int nparams);   // 0 or a hipError_t value
bool picker_loaded();
void picker_set_mode(int mode);
int picker_mode();
// Batched HIP inference: choice[i] = index of the predicted-fastest valid candidate.
int picker_choose(int count, const int *op, const char *ta, const char *tb, const int *m, const int *n, const int *k,
                  const int *lda_pad, int cus, int *choice);

// Exhaustive-sweep table (shapes as swept: ldb = rows of op(B)).
void table_set(int op, bool ta, bool tb, int m, int n, int k, int lda_pad, int choice);
bool table_get(int op, bool ta, bool tb, int m, int n, int k, int lda_pad, GemmConfig *out);
void table_clear();

}  // namespace gemm_amd
