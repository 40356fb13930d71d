// Kernel-configuration picking: the runtime replacement of the reference's
// compile-time tile macros (sgemm_tune.h:6-15, dgemm_tune.h:6-15,
// src/gemm_set_parameters.h:66-82).
#pragma once
#include "gemm_kernels.hpp"

namespace gemm_amd {

enum DType : int { DT_F32 = 0, DT_F64 = 1, DT_C32 = 2, DT_C64 = 3, DT_BF16 = 4 };

// Chooses tile shape and split-K for one problem.  Order of precedence: a
// forced config from the environment (GEMM_AMD_FORCE_TILE / _SPLITK), the
// learned picker when one is loaded, then the built-in occupancy heuristic.
GemmConfig pick_gemm_config(int dtype, bool ta, bool tb, int m, int n, int k, int cus);

// The occupancy heuristic alone (the picker's fallback and the baseline it is
// compared against).
GemmConfig heuristic_gemm_config(int dtype, int m, int n, int k, int cus);

}  // namespace gemm_amd
