// C-ABI of the grouped GEMM (plans, direct launch, hipGraph capture/replay)
// and of the batched 3M complex GEMM built on it.
#include "runtime.hpp"
#include "grouped.hpp"
#include "gemm3m.hpp"
#include "../../include/gemm_amd.h"
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
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

// Plans of recent zgemm3m batches: the grouped plan of the 3·count real
// products and the split / combine job tables, keyed by every argument and by
// the workspace address the planes and products live at.  A batch repeated
// with the same arguments (shapes, pointers, scalars) launches without
// building a plan or uploading tables.  An entry is freed when the last call
// holding it has synchronised and it has left the cache.
struct Batch3m {
    std::vector<uint64_t> key;
    void *plan = nullptr;
    SplitJob3m *d_split = nullptr;
    CombineJob3m *d_comb = nullptr;
    ~Batch3m() {
        if (plan) grouped_plan_destroy(plan);
        if (d_split) (void)hipFree(d_split);
        if (d_comb) (void)hipFree(d_comb);
    }
};
constexpr size_t BATCH3M_CACHED = 4;
std::mutex g_b3_mu;
// least recently used first; never destroyed (no device frees after the HIP
// runtime has shut down at process exit)
std::vector<std::shared_ptr<Batch3m>> &g_b3 = *new std::vector<std::shared_ptr<Batch3m>>();

template <typename T>
uint64_t key_bits(T v) {
    uint64_t u = 0;
    static_assert(sizeof(T) <= sizeof(u), "key words are 64-bit");
    std::memcpy(&u, &v, sizeof(T));
    return u;
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
    const size_t tab = (off[count] * sizeof(double) + 255) / 256 * 256;
    const size_t tabs = 2 * (size_t)count * sizeof(SplitJob3m), tabc = (size_t)count * sizeof(CombineJob3m);
    int st = 0;
    WsLease lease(s);
    char *wsb = (char *)lease.get(WS_COMPLEX, tab + 256, &st);
    if (st) return st;
    double *ws = (double *)wsb;
    std::vector<uint64_t> key;
    key.reserve(2 + 16 * (size_t)count);
    key.push_back((uint64_t)count);
    key.push_back(key_bits(wsb));
    for (int p = 0; p < count; ++p) {
        for (uint64_t v : {(uint64_t)(unsigned char)transa[p], (uint64_t)(unsigned char)transb[p], (uint64_t)m[p],
                           (uint64_t)n[p], (uint64_t)k[p], key_bits(alpha[2 * p]), key_bits(alpha[2 * p + 1]),
                           key_bits(a[p]), (uint64_t)lda[p], key_bits(b[p]), (uint64_t)ldb[p], key_bits(beta[2 * p]),
                           key_bits(beta[2 * p + 1]), key_bits(c[p]), (uint64_t)ldc[p]})
            key.push_back(v);
    }
    std::shared_ptr<Batch3m> bp;
    {
        std::lock_guard<std::mutex> lk(g_b3_mu);
        for (size_t i = 0; i < g_b3.size(); ++i)
            if (g_b3[i]->key == key) {
                bp = g_b3[i];
                g_b3.erase(g_b3.begin() + (long)i);
                g_b3.push_back(bp);   // most recently used
                break;
            }
    }
    if (bp) {
        int e = launch_gemm3m_split_batch_d(bp->d_split, 2 * count, s);
        if (!e) e = grouped_plan_launch(bp->plan, s);
        if (!e) e = launch_gemm3m_combine_batch_d(bp->d_comb, count, s);
        const hipError_t se = hipStreamSynchronize(s);
        if (e) return hip_status((hipError_t)e, "ZGEMM3M_BATCHED");
        return hip_status(se, "ZGEMM3M_BATCHED");
    }
    bp = std::make_shared<Batch3m>();
    bp->key = std::move(key);
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
    bp->plan = grouped_plan_create(GROUPED_F64, 3 * count, ta3.data(), tb3.data(), m3.data(), n3.data(), k3.data(),
                                   al3.data(), A3.data(), la3.data(), B3.data(), lb3.data(), be3.data(), C3.data(),
                                   lc3.data(), rt().cus, 0, &st);
    if (!bp->plan) return hip_status((hipError_t)st, "ZGEMM3M_BATCHED plan");
    hipError_t he = hipMalloc(&bp->d_split, tabs);
    if (he == hipSuccess) he = hipMalloc(&bp->d_comb, tabc);
    if (he == hipSuccess) he = hipMemcpy(bp->d_split, sj.data(), tabs, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(bp->d_comb, cj.data(), tabc, hipMemcpyHostToDevice);
    if (he != hipSuccess) return hip_status(he, "ZGEMM3M_BATCHED tables");
    int e = launch_gemm3m_split_batch_d(bp->d_split, 2 * count, s);
    if (!e) e = grouped_plan_launch(bp->plan, s);
    if (!e) e = launch_gemm3m_combine_batch_d(bp->d_comb, count, s);
    const hipError_t se = hipStreamSynchronize(s);
    if (!e && se == hipSuccess) {
        std::shared_ptr<Batch3m> evicted;   // freed outside the lock
        std::lock_guard<std::mutex> lk(g_b3_mu);
        g_b3.push_back(bp);
        if (g_b3.size() > BATCH3M_CACHED) {
            evicted = g_b3.front();
            g_b3.erase(g_b3.begin());
        }
    }
    if (e) return hip_status((hipError_t)e, "ZGEMM3M_BATCHED");
    return hip_status(se, "ZGEMM3M_BATCHED");
}

}  // extern "C"
