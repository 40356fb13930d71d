// Config picking for the GEMM kernels.
#include "tuner.hpp"
#include <cstdlib>

namespace gemm_amd {

static int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Occupancy heuristic: the largest macro-tile that still gives every CU work;
// when even 64x64 tiles leave CUs idle and K is deep, split K so that about
// two workgroups per CU exist, keeping at least 256 of K per partition.
// Measured on MI355X at 4096^3..8192^3: FP64 is fastest on 128x64 tiles (two
// waves per SIMD), FP32 on 64x64 tiles (four waves per SIMD).
GemmConfig heuristic_gemm_config(int dtype, int m, int n, int k, int cus) {
    GemmConfig c;
    const int eb = (dtype == DT_F64 || dtype == DT_C64) ? 8 : 4;
    const long t128x64 = (long)ceil_div(m, 128) * ceil_div(n, 64);
    const long t64 = (long)ceil_div(m, 64) * ceil_div(n, 64);
    if (eb == 8 && t128x64 >= cus) {
        c.tile = TILE_128x64;
    } else if (t64 >= cus) {
        c.tile = TILE_64x64;
    } else {
        c.tile = TILE_64x64;
        int want = ceil_div(2L * cus, t64);
        int maxk = k / 256;
        if (want > maxk) want = maxk;
        if (want > 16) want = 16;
        c.splitk = want > 1 ? want : 1;
    }
    return c;
}

GemmConfig pick_gemm_config(int dtype, bool, bool, int m, int n, int k, int cus) {
    static const char *ft = getenv("GEMM_AMD_FORCE_TILE");
    static const char *fs = getenv("GEMM_AMD_FORCE_SPLITK");
    GemmConfig c = heuristic_gemm_config(dtype, m, n, k, cus);
    if (ft) {
        int t = atoi(ft);
        if (t >= 0 && t < TILE_COUNT) c.tile = t;
    }
    if (fs) {
        int s = atoi(fs);
        if (s >= 1) c.splitk = s;
    }
    return c;
}

}  // namespace gemm_amd
