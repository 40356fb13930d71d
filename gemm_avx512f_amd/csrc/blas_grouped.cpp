// C-ABI of the grouped GEMM (plans, direct launch, hipGraph capture/replay)
// and of the batched 3M complex GEMM built on it.
#include "runtime.hpp"
#include "grouped.hpp"
#include "gemm3m.hpp"
#include "../../include/gemm_amd.h"
#include <cstdio>
#include <string>
#include <vector>

using namespace gemm_amd;

namespace {

int check_all(const char *name, int count, const char *ta, const char *tb, const int *m, const int *n, const int *k,
              const int *lda, const int *ldb, const int *ldc) {
    for (int p = 0; p < count; ++p) {
        const int info = gemm_check(ta[p], tb[p], m[p], n[p], k[p], lda[p], ldb[p], ldc[p]);
        if (info) {
            char msg[200];
            snprintf(msg, sizeof msg, " ** On entry to %s problem %d: parameter number %d had an illegal value", name,
                     p, info);
            fprintf(stderr, "%s\n", msg);
            set_error(msg);
            return info;
        }
    }
    return 0;
}

}  // namespace

extern "C" {

void *gemm_amd_grouped_create_ex(int dtype, int count, const char *transa, const char *transb, const int *m,
                                 const int *n, const int *k, const double *alpha, const void *const *a,
                                 const int *lda, const void *const *b, const int *ldb, const double *beta,
                                 void *const *c, const int *ldc, int large_min) {
    if (dtype < 0 || dtype > 2 || count < 0 || large_min < 0) {
        set_error("gemm_amd_grouped_create: dtype must be 0 (FP32), 1 (BF16 in, FP32 out) or 2 (FP64); count >= 0; "
                  "large_min >= 0");
        return nullptr;
    }
    if (check_all("GROUPED_GEMM", count, transa, transb, m, n, k, lda, ldb, ldc)) return nullptr;
    if (init_status()) return nullptr;
    int st = 0;
    void *plan = grouped_plan_create(dtype, count, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                                     rt().cus, large_min, &st);
    if (!plan) hip_status((hipError_t)st, "grouped plan");
    return plan;
}

void *gemm_amd_grouped_create(int dtype, int count, const char *transa, const char *transb, const int *m,
                              const int *n, const int *k, const double *alpha, const void *const *a, const int *lda,
                              const void *const *b, const int *ldb, const double *beta, void *const *c,
                              const int *ldc) {
    return gemm_amd_grouped_create_ex(dtype, count, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, 0);
}

int gemm_amd_grouped_run(void *plan, void *stream) {
    if (!plan) return -(int)hipErrorInvalidValue;
    return hip_status((hipError_t)grouped_plan_launch(plan, (hipStream_t)stream), "grouped launch");
}

int gemm_amd_grouped_capture(void *plan, void *stream) {
    if (!plan || !stream) {
        set_error("gemm_amd_grouped_capture: needs a plan and a non-null stream");
        return -(int)hipErrorInvalidValue;
    }
    return hip_status((hipError_t)grouped_plan_capture(plan, (hipStream_t)stream), "grouped capture");
}

int gemm_amd_grouped_replay(void *plan, void *stream) {
    if (!plan) return -(int)hipErrorInvalidValue;
    return hip_status((hipError_t)grouped_plan_replay(plan, (hipStream_t)stream), "grouped replay");
}

int64_t gemm_amd_grouped_tiles(void *plan) { return plan ? grouped_plan_tiles(plan) : -1; }

int gemm_amd_grouped_set_schedule(void *plan, int dynamic, unsigned long long *timeline) {
    if (!plan) return -(int)hipErrorInvalidValue;
    if (grouped_plan_configure(plan, dynamic, timeline)) {
        set_error("gemm_amd_grouped_set_schedule: the plan is already captured (its launch keeps its arguments)");
        return -(int)hipErrorInvalidValue;
    }
    return 0;
}

int gemm_amd_grouped_grid(void *plan) { return plan ? grouped_plan_grid(plan) : -1; }

int gemm_amd_grouped_destroy(void *plan) {
    if (!plan) return 0;
    hipError_t e = hipDeviceSynchronize();   // the descriptor table may still be in use
    if (e != hipSuccess) return hip_status(e, "grouped destroy");
    return hip_status((hipError_t)grouped_plan_destroy(plan), "grouped destroy");
}

int gemm_amd_zgemm3m_batched(int count, const char *transa, const char *transb, const int *m, const int *n,
                             const int *k, const double *alpha, const double *const *a, const int *lda,
                             const double *const *b, const int *ldb, const double *beta, double *const *c,
                             const int *ldc, void *stream) {
    if (count < 0) return -(int)hipErrorInvalidValue;
    if (int info = check_all("ZGEMM3M_BATCHED", count, transa, transb, m, n, k, lda, ldb, ldc)) return info;
    if (count == 0) return 0;
    if (int st = init_status()) return st;
    hipStream_t s = (hipStream_t)stream;
    // scratch: per problem 3 planes of op(A), op(B) and the three real products,
    // each dimension padded to even with zeros, so that every real product has
    // whole 16-byte chunks and aligned planes (the grouped kernel's chunked
    // path).  The padded k adds exact zero products; padded rows and columns
    // of the products are not combined.
    auto even = [](int x) { return x + (x & 1); };
    std::vector<size_t> off(count + 1, 0);
    for (int p = 0; p < count; ++p) {
        const size_t mp = even(m[p]), np = even(n[p]), kp = even(k[p]);
        off[p + 1] = off[p] + 3 * (mp * kp + kp * np + mp * np);
    }
    // then the split and combine job tables
    const size_t tab = (off[count] * sizeof(double) + 255) / 256 * 256;
    const size_t tabs = 2 * (size_t)count * sizeof(SplitJob3m), tabc = (size_t)count * sizeof(CombineJob3m);
    int st = 0;
    WsLease lease(s);
    char *wsb = (char *)lease.get(WS_COMPLEX, tab + tabs + tabc + 256, &st);
    if (st) return st;
    double *ws = (double *)wsb;
    SplitJob3m *d_split = (SplitJob3m *)(wsb + tab);
    CombineJob3m *d_comb = (CombineJob3m *)(wsb + tab + tabs);
    std::vector<SplitJob3m> sj(2 * (size_t)count);
    std::vector<CombineJob3m> cj(count);
    std::vector<char> ta3, tb3;
    std::vector<int> m3, n3, k3, la3, lb3, lc3;
    std::vector<double> al3, be3;
    std::vector<const void *> A3, B3;
    std::vector<void *> C3;
    for (int p = 0; p < count; ++p) {
        const int mp = even(m[p]), np = even(n[p]), kp = even(k[p]);
        double *pa = ws + off[p];
        double *pb = pa + 3 * (size_t)mp * kp;
        double *pt = pb + 3 * (size_t)kp * np;
        const bool tA = !(transa[p] == 'N' || transa[p] == 'n'), cA = (transa[p] == 'C' || transa[p] == 'c');
        const bool tB = !(transb[p] == 'N' || transb[p] == 'n'), cB = (transb[p] == 'C' || transb[p] == 'c');
        sj[2 * p] = {m[p], k[p], mp, kp, tA, cA, a[p], lda[p], pa};
        sj[2 * p + 1] = {k[p], n[p], kp, np, tB, cB, b[p], ldb[p], pb};
        cj[p] = {m[p], n[p], mp, np, pt, alpha[2 * p], alpha[2 * p + 1], beta[2 * p], beta[2 * p + 1], c[p], ldc[p]};
        for (int q = 0; q < 3; ++q) {
            ta3.push_back('N');
            tb3.push_back('N');
            m3.push_back(mp);
            n3.push_back(np);
            k3.push_back(kp);
            la3.push_back(mp > 1 ? mp : 1);
            lb3.push_back(kp > 1 ? kp : 1);
            lc3.push_back(mp > 1 ? mp : 1);
            al3.push_back(1.0);
            be3.push_back(0.0);
            A3.push_back(pa + q * (size_t)mp * kp);
            B3.push_back(pb + q * (size_t)kp * np);
            C3.push_back(pt + q * (size_t)mp * np);
        }
    }
    void *plan = grouped_plan_create(GROUPED_F64, 3 * count, ta3.data(), tb3.data(), m3.data(), n3.data(), k3.data(),
                                     al3.data(), A3.data(), la3.data(), B3.data(), lb3.data(), be3.data(), C3.data(),
                                     lc3.data(), rt().cus, 0, &st);
    if (!plan) return hip_status((hipError_t)st, "ZGEMM3M_BATCHED plan");
    // the host tables stay alive until the synchronisation below
    int e = (int)hipMemcpyAsync(d_split, sj.data(), tabs, hipMemcpyHostToDevice, s);
    if (!e) e = (int)hipMemcpyAsync(d_comb, cj.data(), tabc, hipMemcpyHostToDevice, s);
    if (!e) e = launch_gemm3m_split_batch_d(d_split, 2 * count, s);
    if (!e) e = grouped_plan_launch(plan, s);
    if (!e) e = launch_gemm3m_combine_batch_d(d_comb, count, s);
    hipError_t se = hipStreamSynchronize(s);   // the plan's and the jobs' tables must outlive the launches
    grouped_plan_destroy(plan);
    if (e) return hip_status((hipError_t)e, "ZGEMM3M_BATCHED");
    return hip_status(se, "ZGEMM3M_BATCHED");
}

}  // extern "C"
