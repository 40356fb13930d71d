// C-ABI of the autotuner (exhaustive sweep over the candidate configs), the
// sweep table, the learned picker and the perceptron's training kernels.
#include "runtime.hpp"
#include "tuner.hpp"
#include "mlp.hpp"
#include "gemm3m.hpp"
#include "../../include/gemm_amd.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace gemm_amd;

namespace {

bool op_ok(int op) { return op >= 0 && op < OP_COUNT; }

// One run of `op` with config `cfg` on the sweep buffers.
int run_config(int op, bool ta, bool tb, int m, int n, int k, void *A, int lda, void *B, int ldb, void *C,
               const GemmConfig &cfg, void *work, void *ws3m, hipStream_t s) {
    if (op == OP_SGEMM)
        return launch_sgemm(ta, tb, m, n, k, 1.0f, (const float *)A, lda, (const float *)B, ldb, 0.0f, (float *)C, m,
                            cfg, work, s);
    if (op == OP_DGEMM)
        return launch_dgemm(ta, tb, m, n, k, 1.0, (const double *)A, lda, (const double *)B, ldb, 0.0, (double *)C,
                            m, cfg, work, s);
    const float one[2] = {1.0f, 0.0f}, zero[2] = {0.0f, 0.0f};
    return launch_cgemm3m(ta ? 'T' : 'N', tb ? 'T' : 'N', m, n, k, one, (const float *)A, lda, (const float *)B, ldb,
                          zero, (float *)C, m, ws3m, cfg, work, s);
}

// Milliseconds per run: one warm-up run, then the best of 2-3 batches of runs
// (batches are long enough, up to 32 runs, to hide launch jitter).
int time_config(int op, bool ta, bool tb, int m, int n, int k, void *A, int lda, void *B, int ldb, void *C,
                const GemmConfig &cfg, void *work, void *ws3m, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                float *ms) {
    hipError_t e = hipEventRecord(e0, s);
    if (e != hipSuccess) return (int)e;
    int r = run_config(op, ta, tb, m, n, k, A, lda, B, ldb, C, cfg, work, ws3m, s);
    if (r) return r;
    if ((e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess) return (int)e;
    float warm = 0.0f;
    if ((e = hipEventElapsedTime(&warm, e0, e1)) != hipSuccess) return (int)e;
    int reps = warm > 0.0f ? (int)std::ceil(0.3f / warm) : 32;
    reps = std::max(1, std::min(32, reps));
    const int batches = warm > 2.0f ? 2 : 3;
    float best = INFINITY;
    for (int b = 0; b < batches; ++b) {
        if ((e = hipEventRecord(e0, s)) != hipSuccess) return (int)e;
        for (int i = 0; i < reps; ++i)
            if ((r = run_config(op, ta, tb, m, n, k, A, lda, B, ldb, C, cfg, work, ws3m, s))) return r;
        if ((e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess) return (int)e;
        float t = 0.0f;
        if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) return (int)e;
        best = std::min(best, t / reps);
    }
    *ms = best;
    return 0;
}

}  // namespace

extern "C" {

int gemm_amd_tune_candidates(int *tile, int *splitk) {
    for (int i = 0; i < TUNE_NCFG; ++i) {
        const GemmConfig c = tune_candidate(i);
        if (tile) tile[i] = c.tile;
        if (splitk) splitk[i] = c.splitk;
    }
    return TUNE_NCFG;
}

int gemm_amd_tune_sweep(int op, char transa, char transb, int m, int n, int k, int lda_pad, float *ms) {
    if (!op_ok(op) || m < 1 || n < 1 || k < 1 || lda_pad < 0 || !ms) {
        set_error("gemm_amd_tune_sweep: op 0..2, m, n, k >= 1, lda_pad >= 0");
        return -(int)hipErrorInvalidValue;
    }
    if (int st = init_status()) return st;
    const bool ta = !notrans(transa), tb = !notrans(transb);
    const int nrowa = ta ? k : m, ncola = ta ? m : k, nrowb = tb ? n : k, ncolb = tb ? k : n;
    const int lda = nrowa + lda_pad, ldb = nrowb;
    const int esz = op == OP_DGEMM ? 8 : 4, cw = op == OP_CGEMM3M ? 2 : 1;   // reals per element
    int st = 0;
    hipStream_t s = rt().stream;
    WsLease lease(s);
    void *A = lease.get(WS_TUNE_A, (size_t)esz * cw * lda * ncola, &st);
    void *B = st ? nullptr : lease.get(WS_TUNE_B, (size_t)esz * cw * ldb * ncolb, &st);
    void *C = st ? nullptr : lease.get(WS_TUNE_C, (size_t)esz * cw * m * n, &st);
    void *ws3m = nullptr;
    if (!st && op == OP_CGEMM3M) ws3m = lease.get(WS_COMPLEX, gemm3m_workspace_bytes(m, n, k, 4), &st);
    if (st) return st;
    const int kind = op == OP_DGEMM ? 0 : 1;
    int r = gemm_amd_fill_uniform(A, kind, (int64_t)cw * lda, ncola, (int64_t)cw * lda, 11, 0, 0, s);
    if (!r) r = gemm_amd_fill_uniform(B, kind, (int64_t)cw * ldb, ncolb, (int64_t)cw * ldb, 12, 0, 0, s);
    if (r) return r;
    const GemmConfig heur = heuristic_gemm_config(op == OP_SGEMM ? DT_F32 : op == OP_DGEMM ? DT_F64 : DT_C32, m, n, k,
                                                  rt().cus);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he != hipSuccess) return hip_status(he, "tune events");
    for (int i = 0; i <= TUNE_NCFG && !r; ++i) {
        const GemmConfig cfg = i < TUNE_NCFG ? tune_candidate(i) : heur;
        ms[i] = -1.0f;
        if (i < TUNE_NCFG && !tune_valid(op, m, n, k, cfg)) continue;
        void *work = lease.get(WS_GEMM, gemm_workspace_bytes(m, n, k, cfg, esz), &st);
        if (st) {
            r = st;
            break;
        }
        r = time_config(op, ta, tb, m, n, k, A, lda, B, ldb, C, cfg, work, ws3m, s, e0, e1, &ms[i]);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return r > 0 ? hip_status((hipError_t)r, "tune sweep") : r;   // r < 0: already a status
}

int gemm_amd_pick_config(int op, char transa, char transb, int m, int n, int k, int lda_pad, int mode, int *tile,
                         int *splitk) {
    if (!op_ok(op) || mode < 0 || mode > 2 || !tile || !splitk) {
        set_error("gemm_amd_pick_config: op 0..2, mode 0 (heuristic), 1 (picker) or 2 (library order)");
        return -(int)hipErrorInvalidValue;
    }
    if (int st = init_status()) return st;
    const int dt = op == OP_SGEMM ? DT_F32 : op == OP_DGEMM ? DT_F64 : DT_C32;
    const bool ta = !notrans(transa), tb = !notrans(transb);
    GemmConfig c;
    if (mode == 0) {
        c = heuristic_gemm_config(dt, m, n, k, rt().cus);
    } else if (mode == 1) {
        int choice = 0;
        const char t1 = ta ? 'T' : 'N', t2 = tb ? 'T' : 'N';
        int r = picker_choose(1, &op, &t1, &t2, &m, &n, &k, &lda_pad, rt().cus, &choice);
        if (r) return hip_status((hipError_t)r, "picker (no weights loaded?)");
        c = tune_candidate(choice);
    } else {
        c = pick_gemm_config(dt, ta, tb, m, n, k, lda_pad, 0, rt().cus);   // ldb as in the sweep: rows of op(B)
    }
    *tile = c.tile;
    *splitk = c.splitk;
    return 0;
}

int gemm_amd_picker_load(const float *params, int nparams) {
    if (!params || nparams != MLP_NPARAMS) {
        set_error("gemm_amd_picker_load: expects the 6100 parameters of the 9-64-64-20 network");
        return -(int)hipErrorInvalidValue;
    }
    if (int st = init_status()) return st;
    return hip_status((hipError_t)picker_load(params, nparams), "picker load");
}

int gemm_amd_picker_set_mode(int mode) {
    if (mode != PICK_HEURISTIC && mode != PICK_MLP) {
        set_error("gemm_amd_picker_set_mode: 0 (heuristic) or 1 (learned picker)");
        return -(int)hipErrorInvalidValue;
    }
    if (int st = init_status()) return st;
    picker_set_mode(mode);
    return 0;
}

int gemm_amd_picker_features(int count, const int *op, const char *transa, const char *transb, const int *m,
                             const int *n, const int *k, const int *lda_pad, int cus, float *out) {
    if (count < 0 || cus < 1 || (count > 0 && !out)) {
        set_error("gemm_amd_picker_features: count >= 0, cus >= 1 and an output array");
        return -(int)hipErrorInvalidValue;
    }
    for (int i = 0; i < count; ++i) {
        if (!op_ok(op[i]) || m[i] < 1 || n[i] < 1 || k[i] < 1 || lda_pad[i] < 0) {
            set_error("gemm_amd_picker_features: op 0..2, m, n, k >= 1 and lda_pad >= 0 for every shape");
            return -(int)hipErrorInvalidValue;
        }
        const bool ta = !notrans(transa[i]), tb = !notrans(transb[i]);
        mlp_features(op[i], ta, tb, m[i], n[i], k[i], lda_pad[i],
                     gemm_vec_path(op[i], ta, tb, m[i], n[i], k[i], lda_pad[i], 0), cus, out + (size_t)i * MLP_IN);
    }
    return 0;
}

int gemm_amd_picker_choose(int count, const int *op, const char *transa, const char *transb, const int *m,
                           const int *n, const int *k, const int *lda_pad, int *choice) {
    if (count < 0) return -(int)hipErrorInvalidValue;
    for (int i = 0; i < count; ++i)
        if (!op_ok(op[i]) || m[i] < 1 || n[i] < 1 || k[i] < 1) {
            set_error("gemm_amd_picker_choose: op 0..2 and m, n, k >= 1 for every shape");
            return -(int)hipErrorInvalidValue;
        }
    if (int st = init_status()) return st;
    return hip_status((hipError_t)picker_choose(count, op, transa, transb, m, n, k, lda_pad, rt().cus, choice),
                      "picker (no weights loaded?)");
}

int gemm_amd_tune_table_set(int count, const int *op, const char *transa, const char *transb, const int *m,
                            const int *n, const int *k, const int *lda_pad, const int *choice) {
    for (int i = 0; i < count; ++i)
        if (!op_ok(op[i]) || choice[i] < 0 || choice[i] >= TUNE_NCFG) {
            set_error("gemm_amd_tune_table_set: op 0..2 and choice 0..19");
            return -(int)hipErrorInvalidValue;
        }
    for (int i = 0; i < count; ++i)
        table_set(op[i], !notrans(transa[i]), !notrans(transb[i]), m[i], n[i], k[i], lda_pad[i], choice[i]);
    return 0;
}

int gemm_amd_tune_table_clear(void) {
    table_clear();
    return 0;
}

// ---------------------------------------------------------------- training kernels
int gemm_amd_mlp_forward(const float *params, const float *x, int batch, float *y, void *stream) {
    if (batch < 0) return -(int)hipErrorInvalidValue;
    if (int st = init_status()) return st;
    int st = 0;
    WsLease lease((hipStream_t)stream);
    float *ws = (float *)lease.get(WS_MLP, sizeof(float) * mlp_forward_ws_floats(batch) + 256, &st);
    if (st) return st;
    return hip_status((hipError_t)mlp_forward(params, x, batch, y, ws, (hipStream_t)stream), "mlp forward");
}

int gemm_amd_mlp_grad(const float *params, const float *x, const float *t, const float *mask, int batch, float *grad,
                      float *loss, void *stream) {
    if (batch < 1) return -(int)hipErrorInvalidValue;
    if (int st = init_status()) return st;
    int st = 0;
    WsLease lease((hipStream_t)stream);
    float *ws = (float *)lease.get(WS_MLP, sizeof(float) * mlp_grad_ws_floats(batch) + 256, &st);
    if (st) return st;
    return hip_status((hipError_t)mlp_grad(params, x, t, mask, batch, grad, loss, ws, (hipStream_t)stream),
                      "mlp grad");
}

int gemm_amd_mlp_adam(float *params, const float *grad, float *m, float *v, int n, float lr, float beta1, float beta2,
                      float eps, int step, void *stream) {
    if (n < 1 || step < 1) return -(int)hipErrorInvalidValue;
    if (int st = init_status()) return st;
    return hip_status((hipError_t)mlp_adam(params, grad, m, v, n, lr, beta1, beta2, eps, step, (hipStream_t)stream),
                      "mlp adam");
}

int gemm_amd_mlp_train(float *params, float *adam_m, float *adam_v, const float *x, const float *t, const float *mask,
                       int nsamples, int batch, int epochs, float lr, uint64_t seed, float *epoch_loss) {
    if (nsamples < 1 || batch < 1 || epochs < 1 || lr <= 0.0f) {
        set_error("gemm_amd_mlp_train: nsamples, batch, epochs >= 1 and lr > 0");
        return -(int)hipErrorInvalidValue;
    }
    if (int st = init_status()) return st;
    if (batch > nsamples) batch = nsamples;
    const int nb = nsamples / batch;   // whole batches per epoch (the remainder rotates through epochs)
    hipStream_t s = rt().stream;
    int st = 0;
    const size_t f_x = (size_t)batch * MLP_IN, f_t = (size_t)batch * MLP_OUT;
    const size_t bytes = sizeof(float) * (mlp_grad_ws_floats(batch) + f_x + 2 * f_t + MLP_NPARAMS + nb) +
                         sizeof(int) * (size_t)nsamples + 1024;
    WsLease lease(s);
    char *base = (char *)lease.get(WS_MLP, bytes, &st);
    if (st) return st;
    float *ws = (float *)base;
    float *bx = ws + mlp_grad_ws_floats(batch), *bt = bx + f_x, *bm = bt + f_t, *grad = bm + f_t;
    float *losses = grad + MLP_NPARAMS;
    int *d_idx = (int *)(losses + nb);
    std::vector<int> idx(nsamples);
    std::vector<float> hl(nb);
    uint64_t state = seed;
    auto next = [&]() {   // splitmix64
        uint64_t z = (state += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    int step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = 0; i < nsamples; ++i) idx[i] = i;
        for (int i = nsamples - 1; i > 0; --i) std::swap(idx[i], idx[(int)(next() % (uint64_t)(i + 1))]);
        hipError_t e = hipMemcpyAsync(d_idx, idx.data(), sizeof(int) * nsamples, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_status(e, "mlp train");
        // cosine decay of the step size over the run
        const float lr_ep = lr * 0.5f * (1.0f + (float)std::cos(M_PI * ep / epochs));
        for (int b = 0; b < nb; ++b) {
            const int *bi = d_idx + (size_t)b * batch;
            int r = mlp_gather(x, bi, batch, MLP_IN, bx, s);
            if (!r) r = mlp_gather(t, bi, batch, MLP_OUT, bt, s);
            if (!r) r = mlp_gather(mask, bi, batch, MLP_OUT, bm, s);
            if (!r) r = mlp_grad(params, bx, bt, bm, batch, grad, losses + b, ws, s);
            if (!r) r = mlp_adam(params, grad, adam_m, adam_v, MLP_NPARAMS, lr_ep, 0.9f, 0.999f, 1e-8f, ++step, s);
            if (r) return hip_status((hipError_t)r, "mlp train");
        }
        e = hipMemcpyAsync(hl.data(), losses, sizeof(float) * nb, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_status(e, "mlp train");
        double sum = 0.0;
        for (float v : hl) sum += v;
        if (epoch_loss) epoch_loss[ep] = (float)(sum / nb);
    }
    return 0;
}

}  // extern "C"
