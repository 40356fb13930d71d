// Host runtime of libgemm_amd.so: device state, buffer cache, errors, and
// the plumbing entry points (memory, copies, streams, events).
#include "runtime.hpp"
#include "../../include/gemm_amd.h"
#include <cstdio>
#include <cstring>
#include <vector>

namespace gemm_amd {

static int g_init_status = 1;   // 1 = not yet initialised
static Runtime g_rt;
static std::once_flag g_once;

static void do_init() {
    hipError_t e = hipGetDevice(&g_rt.device);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, g_rt.device);
        if (e == hipSuccess) g_rt.cus = prop.multiProcessorCount;
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking);
    g_init_status = (e == hipSuccess) ? 0 : -(int)e;
    if (e != hipSuccess) {
        std::string m = "gemm_amd: HIP initialisation failed: ";
        m += hipGetErrorString(e);
        set_error(m);
    }
}

Runtime &rt() {
    std::call_once(g_once, do_init);
    return g_rt;
}

int init_status() {
    rt();
    return g_init_status;
}

static thread_local std::string t_err;
void set_error(const std::string &msg) { t_err = msg; }
const char *get_error() { return t_err.c_str(); }

int hip_status(hipError_t err, const char *where) {
    if (err == hipSuccess) return 0;
    std::string m = "gemm_amd: ";
    m += where;
    m += ": ";
    m += hipGetErrorString(err);
    set_error(m);
    return -(int)err;
}

int DevBuf::ensure(size_t need) {
    if (need <= bytes && ptr) return 0;
    if (ptr) {
        hipError_t e = hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        if (e != hipSuccess) return hip_status(e, "hipFree");
    }
    size_t alloc = need < 256 ? 256 : need;
    alloc = (alloc + 0xFFFFF) & ~(size_t)0xFFFFF;   // round up to 1 MiB
    hipError_t e = hipMalloc(&ptr, alloc);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_status(e, "hipMalloc (buffer cache)");
    }
    bytes = alloc;
    return 0;
}

static DevBuf g_bufs[BUF_COUNT];
DevBuf &buf(BufId id) { return g_bufs[id]; }

struct WsSet {
    DevBuf ws[WS_COUNT];
    hipEvent_t done = nullptr;     // recorded on the last holder's stream when it let the set go
    hipStream_t last = nullptr;    // that stream
    bool used = false;             // `done` has been recorded
    bool busy = false;             // held by a lease
    bool captured = false;         // owned by a captured graph: never reused, never freed
    std::vector<void *> retired;   // captured sets: buffers outgrown during the capture
};

static std::mutex g_pool_mu;
static std::vector<WsSet *> g_pool;   // sets are never destroyed (process lifetime)

static bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive;
}

WsLease::WsLease(hipStream_t s) : s_(s) {
    const bool capturing = stream_capturing(s);
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (!capturing) {
            for (WsSet *w : g_pool)   // this stream's own free set first
                if (!w->busy && !w->captured && w->used && w->last == s) {
                    set_ = w;
                    break;
                }
            if (!set_)
                for (WsSet *w : g_pool)
                    if (!w->busy && !w->captured) {
                        set_ = w;
                        break;
                    }
        }
        if (!set_) {
            set_ = new WsSet();
            set_->captured = capturing;
            g_pool.push_back(set_);
        }
        set_->busy = true;
    }
    if (set_->used && set_->last != s) {
        // kernels of the previous holder (another stream) may still use the buffers
        status_ = hip_status(hipStreamWaitEvent(s, set_->done, 0), "hipStreamWaitEvent (workspace hand-over)");
    }
}

WsLease::~WsLease() {
    if (!set_) return;
    if (!set_->captured) {
        hipError_t e = hipSuccess;
        if (!set_->done) e = hipEventCreateWithFlags(&set_->done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(set_->done, s_);
        if (e == hipSuccess) {
            set_->last = s_;
            set_->used = true;
        } else {
            // without a recorded event the next holder could not order itself
            // after this one: drain the stream instead
            (void)hipStreamSynchronize(s_);
            set_->used = false;
        }
    }
    std::lock_guard<std::mutex> lk(g_pool_mu);
    set_->busy = false;
}

void *WsLease::get(int slot, size_t bytes, int *status) {
    int st = status_;
    DevBuf &w = set_->ws[slot];
    if (st == 0 && (bytes > w.bytes || !w.ptr)) {
        if (set_->captured) {
            // the captured kernels keep the old buffer: retire it, never free it
            hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
            (void)hipThreadExchangeStreamCaptureMode(&mode);
            if (w.ptr) set_->retired.push_back(w.ptr);
            w.ptr = nullptr;
            w.bytes = 0;
            st = w.ensure(bytes);
            (void)hipThreadExchangeStreamCaptureMode(&mode);
        } else {
            // this stream waits for every earlier holder of the set, and this
            // lease's own earlier kernels run on it: once it drains, nothing
            // queued can still use the old buffer
            if (w.ptr) st = hip_status(hipStreamSynchronize(s_), "hipStreamSynchronize (workspace growth)");
            if (st == 0) st = w.ensure(bytes);
        }
    }
    if (status) *status = st;
    return st == 0 ? w.ptr : nullptr;
}

void xerbla(const char *name, int info) {
    char msg[160];
    snprintf(msg, sizeof msg, " ** On entry to %-6s parameter number %2d had an illegal value", name, info);
    fprintf(stderr, "%s\n", msg);
    set_error(msg);
}

int gemm_check(char ta, char tb, int m, int n, int k, int lda, int ldb, int ldc) {
    const int nrowa = notrans(ta) ? m : k;
    const int nrowb = notrans(tb) ? k : n;
    if (!is_trans_char(ta)) return 1;
    if (!is_trans_char(tb)) return 2;
    if (m < 0) return 3;
    if (n < 0) return 4;
    if (k < 0) return 5;
    if (lda < (nrowa > 1 ? nrowa : 1)) return 8;
    if (ldb < (nrowb > 1 ? nrowb : 1)) return 10;
    if (ldc < (m > 1 ? m : 1)) return 13;
    return 0;
}

}  // namespace gemm_amd

using namespace gemm_amd;

extern "C" {

int gemm_amd_init(void) { return init_status(); }

int gemm_amd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gemm_amd_device_info(int64_t *info) {
    if (int st = init_status()) return st;
    hipDeviceProp_t p;
    hipError_t e = hipGetDeviceProperties(&p, rt().device);
    if (e != hipSuccess) return hip_status(e, "hipGetDeviceProperties");
    info[0] = p.multiProcessorCount;
    info[1] = p.clockRate;
    info[2] = (int64_t)p.sharedMemPerBlock;
    info[3] = p.l2CacheSize;
    return 0;
}

const char *gemm_amd_last_error(void) { return get_error(); }

void *gemm_amd_malloc(size_t bytes) {
    if (init_status()) return nullptr;
    void *p = nullptr;
    if (hip_status(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc")) return nullptr;
    return p;
}

int gemm_amd_free(void *p) { return hip_status(hipFree(p), "hipFree"); }

int gemm_amd_memcpy(void *dst, const void *src, size_t bytes, int kind) {
    if (int st = init_status()) return st;
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                                                                     : hipMemcpyDeviceToDevice;
    hipError_t e = hipMemcpy(dst, src, bytes, k);
    // a device-to-device hipMemcpy may return before the copy has run; the
    // entry point is synchronous, so wait for it (kernels on non-blocking
    // streams do not wait for the null stream)
    if (e == hipSuccess && k == hipMemcpyDeviceToDevice) e = hipStreamSynchronize(nullptr);
    return hip_status(e, "hipMemcpy");
}

int gemm_amd_memset(void *p, int value, size_t bytes) {
    if (int st = init_status()) return st;
    hipError_t e = hipMemset(p, value, bytes);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);   // synchronous, like gemm_amd_memcpy
    return hip_status(e, "hipMemset");
}

int gemm_amd_synchronize(void) {
    if (int st = init_status()) return st;
    return hip_status(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

void *gemm_amd_stream_create(void) {
    if (init_status()) return nullptr;
    hipStream_t s = nullptr;
    if (hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate")) return nullptr;
    return (void *)s;
}

int gemm_amd_stream_destroy(void *s) { return hip_status(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy"); }

int gemm_amd_stream_synchronize(void *s) {
    return hip_status(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize");
}

void *gemm_amd_event_create(void) {
    if (init_status()) return nullptr;
    hipEvent_t e = nullptr;
    if (hip_status(hipEventCreate(&e), "hipEventCreate")) return nullptr;
    return (void *)e;
}

int gemm_amd_event_destroy(void *e) { return hip_status(hipEventDestroy((hipEvent_t)e), "hipEventDestroy"); }

int gemm_amd_event_record(void *e, void *s) {
    return hip_status(hipEventRecord((hipEvent_t)e, (hipStream_t)s), "hipEventRecord");
}

int gemm_amd_event_elapsed_ms(void *a, void *b, float *ms) {
    hipError_t e = hipEventSynchronize((hipEvent_t)b);
    if (e != hipSuccess) return hip_status(e, "hipEventSynchronize");
    return hip_status(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
}

}  // extern "C"
