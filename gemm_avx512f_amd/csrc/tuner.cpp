// Config picking for the GEMM kernels: occupancy heuristic, exhaustive-sweep
// table and the learned picker (HIP inference of the perceptron in mlp.hip).
#include "tuner.hpp"
#include "mlp.hpp"
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace gemm_amd {

// Trained picker weights compiled into the library (count 0 = none).
#include "picker_weights.inc"

static int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

int dtype_to_op(int dtype) {
    switch (dtype) {
    case DT_F32: return OP_SGEMM;
    case DT_F64: return OP_DGEMM;
    case DT_C32: return OP_CGEMM3M;
    default: return -1;
    }
}

GemmConfig tune_candidate(int i) {
    GemmConfig c;
    c.tile = i / TUNE_NSPLIT;
    c.splitk = 1 << (i % TUNE_NSPLIT);
    return c;
}

int tune_index(const GemmConfig &c) {
    if (c.tile < 0 || c.tile >= TILE_COUNT || c.splitk < 1 || (c.splitk & (c.splitk - 1))) return -1;
    const int l = __builtin_ctz((unsigned)c.splitk);
    return l < TUNE_NSPLIT ? c.tile * TUNE_NSPLIT + l : -1;
}

bool tune_valid(int op, int m, int n, int k, const GemmConfig &c) {
    if (c.splitk == 1) return true;
    const double eb = op == OP_DGEMM ? 8.0 : 4.0;
    return k >= 64 * c.splitk && (double)m * n * c.splitk * eb <= (double)(1u << 30);
}

// Occupancy heuristic: the largest macro-tile that still gives every CU work;
// when even 64x64 tiles leave CUs idle and K is deep, split K so that about
// two workgroups per CU exist, keeping at least 256 of K per partition.
// Measured on MI355X at 4096^3..8192^3: FP64 is fastest on 128x64 tiles (two
// waves per SIMD), FP32 on 256x128 then 128x128 tiles (64x64 per wave keeps
// the LDS read traffic per matrix instruction low).  The complex 3M
// products use 128x128 tiles when those fill the chip and split K at most 8
// ways.
GemmConfig heuristic_gemm_config(int dtype, int m, int n, int k, int cus) {
    GemmConfig c;
    const long t128 = (long)ceil_div(m, 128) * ceil_div(n, 128);
    const long t128x64 = (long)ceil_div(m, 128) * ceil_div(n, 64);
    const long t64 = (long)ceil_div(m, 64) * ceil_div(n, 64);
    if (dtype == DT_C32 || dtype == DT_C64) {
        // the three real products use the real dtype's tile rule below
        const bool f64 = dtype == DT_C64;
        const long t256x128c = (long)ceil_div(m, 256) * ceil_div(n, 128);
        if (f64 && t128x64 >= cus) {
            c.tile = TILE_128x64;
            return c;
        }
        if (!f64 && t256x128c >= cus) {
            c.tile = TILE_LARGE;
            return c;
        }
        if (t128 >= cus) return c;   // TILE_128x128
        c.tile = TILE_64x64;
        if (t64 < cus && k >= 512) {
            int want = ceil_div(2L * cus, t64);
            int maxk = k / 256;
            if (want > maxk) want = maxk;
            if (want > 8) want = 8;
            c.splitk = want > 1 ? want : 1;
        }
        return c;
    }
    const bool fp64 = dtype == DT_F64;
    const long t256x128 = (long)ceil_div(m, 256) * ceil_div(n, 128);
    if (fp64 && t128x64 >= cus) {
        c.tile = TILE_128x64;
    } else if (!fp64 && t256x128 >= cus) {
        c.tile = TILE_LARGE;   // FP32: 256x128 tiles of eight 64x64 waves, once they fill the chip
    } else if (!fp64 && t128 >= cus) {
        c.tile = TILE_128x128;
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

// ---------------------------------------------------------------- picker state
namespace {

struct Key {
    int op, ta, tb, m, n, k, pad, vec;
    bool operator==(const Key &o) const {
        return op == o.op && ta == o.ta && tb == o.tb && m == o.m && n == o.n && k == o.k && pad == o.pad &&
               vec == o.vec;
    }
};
struct KeyHash {
    size_t operator()(const Key &x) const {
        uint64_t h = 1469598103934665603ull;
        for (int v : {x.op, x.ta, x.tb, x.m, x.n, x.k, x.pad, x.vec}) h = (h ^ (uint32_t)v) * 1099511628211ull;
        return (size_t)h;
    }
};

Key make_key(int op, bool ta, bool tb, int m, int n, int k, int pad, bool vec) {
    return Key{op, ta ? 1 : 0, tb ? 1 : 0, m, n, k, ((pad % 64) + 64) % 64, vec ? 1 : 0};
}

struct Picker {
    std::mutex mu;
    bool loaded = false, defaults_tried = false;
    int mode = -1;   // -1: not decided yet
    float *d_params = nullptr, *d_x = nullptr, *d_y = nullptr, *d_ws = nullptr;
    int cap = 0;
    hipStream_t stream = nullptr;
    std::unordered_map<Key, GemmConfig, KeyHash> table, cache;
};

Picker &picker() {
    static Picker *p = new Picker();   // never destroyed: may be used during process teardown
    return *p;
}

int ensure_capacity(Picker &P, int batch) {
    if (batch <= P.cap) return 0;
    if (P.d_x) (void)hipFree(P.d_x);
    if (P.d_y) (void)hipFree(P.d_y);
    if (P.d_ws) (void)hipFree(P.d_ws);
    P.d_x = P.d_y = P.d_ws = nullptr;
    P.cap = 0;
    int cap = batch < 64 ? 64 : batch;
    hipError_t e = hipMalloc(&P.d_x, sizeof(float) * cap * MLP_IN);
    if (e == hipSuccess) e = hipMalloc(&P.d_y, sizeof(float) * cap * MLP_OUT);
    if (e == hipSuccess) e = hipMalloc(&P.d_ws, sizeof(float) * mlp_forward_ws_floats(cap));
    if (e != hipSuccess) return (int)e;
    P.cap = cap;
    return 0;
}

int load_locked(Picker &P, const float *params, int n) {
    if (n != MLP_NPARAMS) return (int)hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    if (!P.stream) e = hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking);
    if (e == hipSuccess && !P.d_params) e = hipMalloc(&P.d_params, sizeof(float) * MLP_NPARAMS);
    if (e == hipSuccess) e = hipMemcpy(P.d_params, params, sizeof(float) * MLP_NPARAMS, hipMemcpyHostToDevice);
    if (e != hipSuccess) return (int)e;
    if (int r = ensure_capacity(P, 64)) return r;
    P.loaded = true;
    P.cache.clear();
    return 0;
}

void decide_defaults_locked(Picker &P) {
    if (P.defaults_tried) return;
    P.defaults_tried = true;
    if (!P.loaded && kPickerWeightsCount == MLP_NPARAMS) load_locked(P, kPickerWeights, kPickerWeightsCount);
    if (P.mode < 0) {
        const char *env = getenv("GEMM_AMD_PICKER");
        if (env && !strcmp(env, "heuristic")) P.mode = PICK_HEURISTIC;
        else if (env && !strcmp(env, "mlp")) P.mode = PICK_MLP;
        else P.mode = P.loaded ? PICK_MLP : PICK_HEURISTIC;
    }
}

// Runs the network on `count` feature rows; fills out[count * MLP_OUT].
int infer_locked(Picker &P, int count, const float *feat, float *out) {
    if (int r = ensure_capacity(P, count)) return r;
    // the inference runs on the picker's own stream; relaxed capture mode keeps
    // it legal while a caller's stream is being captured into a graph
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    hipError_t e = hipMemcpyAsync(P.d_x, feat, sizeof(float) * count * MLP_IN, hipMemcpyHostToDevice, P.stream);
    int r = e == hipSuccess ? mlp_forward(P.d_params, P.d_x, count, P.d_y, P.d_ws, P.stream) : (int)e;
    if (!r) e = hipMemcpyAsync(out, P.d_y, sizeof(float) * count * MLP_OUT, hipMemcpyDeviceToHost, P.stream);
    if (!r && e == hipSuccess) e = hipStreamSynchronize(P.stream);
    if (!r) r = (int)e;
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    return r;
}

// The valid candidate with the lowest predicted cost; a split-K choice gives
// way to the same tile without split-K when the network predicts that one
// within PICK_TIE (log2 units, ~1.4%): no reduce pass, no workspace, and the
// network's resolution is a few percent.
constexpr float PICK_TIE = 0.02f;
int argmin_valid(int op, int m, int n, int k, const float *y) {
    int best = -1;
    for (int c = 0; c < TUNE_NCFG; ++c) {
        if (!tune_valid(op, m, n, k, tune_candidate(c))) continue;
        if (best < 0 || y[c] < y[best]) best = c;
    }
    if (best < 0) return 0;
    const int one = (best / TUNE_NSPLIT) * TUNE_NSPLIT;   // c = tile * TUNE_NSPLIT + log2(split-K)
    if (one != best && tune_valid(op, m, n, k, tune_candidate(one)) && y[one] - y[best] <= PICK_TIE) best = one;
    return best;
}

}  // namespace

int picker_load(const float *params, int nparams) {
    Picker &P = picker();
    std::lock_guard<std::mutex> lk(P.mu);
    P.defaults_tried = true;   // explicit weights replace the compiled-in ones
    return load_locked(P, params, nparams);
}

bool picker_loaded() {
    Picker &P = picker();
    std::lock_guard<std::mutex> lk(P.mu);
    decide_defaults_locked(P);
    return P.loaded;
}

void picker_set_mode(int mode) {
    Picker &P = picker();
    std::lock_guard<std::mutex> lk(P.mu);
    decide_defaults_locked(P);
    P.mode = mode;
}

int picker_mode() {
    Picker &P = picker();
    std::lock_guard<std::mutex> lk(P.mu);
    decide_defaults_locked(P);
    return P.mode;
}

int picker_choose(int count, const int *op, const char *ta, const char *tb, const int *m, const int *n, const int *k,
                  const int *lda_pad, int cus, int *choice) {
    Picker &P = picker();
    std::lock_guard<std::mutex> lk(P.mu);
    decide_defaults_locked(P);
    if (!P.loaded) return (int)hipErrorNotInitialized;
    if (count <= 0) return 0;
    std::vector<float> feat((size_t)count * MLP_IN), y((size_t)count * MLP_OUT);
    for (int i = 0; i < count; ++i) {
        const bool a = !(ta[i] == 'N' || ta[i] == 'n'), b = !(tb[i] == 'N' || tb[i] == 'n');
        mlp_features(op[i], a, b, m[i], n[i], k[i], lda_pad[i], gemm_vec_path(op[i], a, b, m[i], n[i], k[i],
                     lda_pad[i], 0), cus, &feat[(size_t)i * MLP_IN]);
    }
    if (int r = infer_locked(P, count, feat.data(), y.data())) return r;
    for (int i = 0; i < count; ++i) choice[i] = argmin_valid(op[i], m[i], n[i], k[i], &y[(size_t)i * MLP_OUT]);
    return 0;
}

void table_set(int op, bool ta, bool tb, int m, int n, int k, int lda_pad, int choice) {
    Picker &P = picker();
    std::lock_guard<std::mutex> lk(P.mu);
    P.table[make_key(op, ta, tb, m, n, k, lda_pad, gemm_vec_path(op, ta, tb, m, n, k, lda_pad, 0))] =
        tune_candidate(choice);
}

bool table_get(int op, bool ta, bool tb, int m, int n, int k, int lda_pad, GemmConfig *out) {
    Picker &P = picker();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.table.find(make_key(op, ta, tb, m, n, k, lda_pad, gemm_vec_path(op, ta, tb, m, n, k, lda_pad, 0)));
    if (it == P.table.end()) return false;
    *out = it->second;
    return true;
}

void table_clear() {
    Picker &P = picker();
    std::lock_guard<std::mutex> lk(P.mu);
    P.table.clear();
}

static thread_local bool t_pick_without_device = false;

PickWithoutDevice::PickWithoutDevice() : prev_(t_pick_without_device) { t_pick_without_device = true; }
PickWithoutDevice::~PickWithoutDevice() { t_pick_without_device = prev_; }

bool gemm_vec_path(int op, bool ta, bool tb, int m, int n, int k, int lda_pad, int ldb_pad) {
    if (op == OP_CGEMM3M) return m % 4 == 0 && k % 4 == 0;
    const int vw = op == OP_DGEMM ? 2 : 4;   // elements per 16 bytes
    const long lda = (long)(ta ? k : m) + lda_pad, ldb = (long)(tb ? n : k) + ldb_pad;
    return lda % vw == 0 && ldb % vw == 0;
}

GemmConfig pick_gemm_config(int dtype, bool ta, bool tb, int m, int n, int k, int lda_pad, int ldb_pad, int cus) {
    static const char *ft = getenv("GEMM_AMD_FORCE_TILE");
    static const char *fs = getenv("GEMM_AMD_FORCE_SPLITK");
    GemmConfig c = heuristic_gemm_config(dtype, m, n, k, cus);
    if (ft || fs) {
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
    const int op = dtype_to_op(dtype);
    if (op < 0) return c;
    Picker &P = picker();
    std::lock_guard<std::mutex> lk(P.mu);
    const bool vec = gemm_vec_path(op, ta, tb, m, n, k, lda_pad, ldb_pad);
    const Key key = make_key(op, ta, tb, m, n, k, lda_pad, vec);
    auto it = P.table.find(key);
    if (it != P.table.end()) return it->second;
    if (t_pick_without_device) {
        // only what is already known: the picker's cached choice, else the heuristic
        if (!P.defaults_tried || P.mode != PICK_MLP || !P.loaded) return c;
        auto ct = P.cache.find(key);
        return ct != P.cache.end() ? ct->second : c;
    }
    decide_defaults_locked(P);
    if (P.mode != PICK_MLP || !P.loaded) return c;
    auto ct = P.cache.find(key);
    if (ct != P.cache.end()) return ct->second;
    float feat[MLP_IN], y[MLP_OUT];
    mlp_features(op, ta, tb, m, n, k, lda_pad, vec, cus, feat);
    if (infer_locked(P, 1, feat, y)) return c;   // inference failure: keep the heuristic
    const GemmConfig g = tune_candidate(argmin_valid(op, m, n, k, y));
    P.cache[key] = g;
    return g;
}

}  // namespace gemm_amd
