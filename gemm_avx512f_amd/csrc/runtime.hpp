// Host runtime shared by the C-ABI entry points: device state, buffer cache,
// error reporting and argument checking.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <mutex>

namespace gemm_amd {

struct Runtime {
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;   // the library's stream for host-pointer calls
    std::mutex mu;                  // serialises host-pointer calls and the buffer cache
};

Runtime &rt();                       // initialises on first use (throws nothing; check init_status)
int init_status();                   // 0 or negative hipError_t of the first init

// Thread-local last-error message.
void set_error(const std::string &msg);
const char *get_error();

// Returns 0, or -(int)err after recording the message.
int hip_status(hipError_t err, const char *where);

// A growable device buffer owned by the library (never shrinks).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);         // 0 or negative HIP error
};

// Named scratch buffers: host-pointer staging (A, B, C) and kernel workspaces.
enum BufId { BUF_A = 0, BUF_B, BUF_C, BUF_WORK, BUF_WORK2, BUF_WORK3, BUF_COUNT };
DevBuf &buf(BufId id);                     // for host-pointer calls (under rt().mu)
// Workspaces of the device-pointer calls.  A routine holds several slots at
// once (e.g. DTRSM's inverse blocks while its trailing GEMMs use the split-K
// slot).
enum WsSlot { WS_GEMM = 0, WS_TRSM_INV, WS_EMU_A, WS_EMU_B, WS_EMU_AUX, WS_EMU_W, WS_COMPLEX, WS_GROUPED,
              WS_TUNE_A, WS_TUNE_B, WS_TUNE_C, WS_MLP,
              WS_COUNT };

struct WsSet;

// One entry point's claim on a set of workspace slots, from its first
// workspace request until every kernel that uses them is enqueued (construct
// it before the first request, let it go out of scope after the last launch).
//
// Sets are pooled.  A set is held by one lease at a time, so calls on
// different streams (from different host threads) never share a buffer; a
// free set whose last holder ran on this stream is preferred, so a stream
// keeps reusing its own buffers.  Taking a set last used on another stream
// makes this stream wait (hipStreamWaitEvent) for the event recorded when
// that holder let it go, so kernels of the two holders never overlap on the
// same memory.  Growing a slot synchronises this stream (which by then waits
// for every earlier holder too) before the old buffer is freed, so no queued
// kernel can still read it.  A stream under graph capture gets a set of its
// own that is never reused or freed: the captured graph keeps its buffers.
class WsLease {
public:
    explicit WsLease(hipStream_t s);
    ~WsLease();
    WsLease(const WsLease &) = delete;
    WsLease &operator=(const WsLease &) = delete;
    // Device memory of at least `bytes` for `slot`, valid for this lease's
    // kernels; nullptr with *status < 0 on failure.
    void *get(int slot, size_t bytes, int *status);

private:
    WsSet *set_ = nullptr;
    hipStream_t s_ = nullptr;
    int status_ = 0;
};

// netlib xerbla message: " ** On entry to NAME parameter number I had an illegal value"
void xerbla(const char *name, int info);

inline bool lsame(char a, char b) { return a == b || a == (char)(b + 32); }
inline bool is_trans_char(char t) { return lsame(t, 'N') || lsame(t, 'T') || lsame(t, 'C'); }
inline bool notrans(char t) { return t == 'N' || t == 'n'; }

// Checks the GEMM argument list exactly as reference BLAS does (info = index of
// the first illegal argument in the Fortran argument list, 0 if all legal).
int gemm_check(char ta, char tb, int m, int n, int k, int lda, int ldb, int ldc);

}  // namespace gemm_amd
