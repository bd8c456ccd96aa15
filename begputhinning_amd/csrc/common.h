// common.h -- shared host/device definitions of the MI355X thinning library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bgt.h"

namespace bgt {

// ---- errors ----------------------------------------------------------------
void set_error(int code, const std::string &msg);
std::string last_error_message();  // the calling thread's last message
int fail(int code, const char *fmt, ...);

#define BGT_HIP(expr)                                                                \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return ::bgt::fail(BGT_ERR_HIP, "%s failed: %s (%s:%d)", #expr,         \
                               hipGetErrorString(e_), __FILE__, __LINE__);           \
    } while (0)

#define BGT_CHECK(expr)                                                              \
    do {                                                                             \
        int s_ = (expr);                                                             \
        if (s_ != BGT_OK) return s_;                                                 \
    } while (0)

// ---- device memory -------------------------------------------------------
// Every device allocation of the library goes through here, so that tests can
// make the next ones fail (bgt_debug_fail_device_allocs).
hipError_t device_malloc(void **p, size_t bytes);

// Grow-only per-thread workspace used by the host-pointer drop-in entry points
// so that repeated calls do not hipMalloc/hipFree every time.
// A thread's workspaces and pinned poll buffers are released when the thread
// exits (server workers, caller thread pools); the process's main thread
// keeps them until the process ends.
//
// An outgrown buffer is retired, not freed: hipFree waits for the whole
// device, which stalls every other thread's work on it for milliseconds under
// load.  Each growth at least doubles the buffer, so the retired buffers
// together hold less than the current one.
struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
    int device = -1;
    std::vector<std::pair<void *, int>> retired;  // (buffer, device)
    ~Workspace();
};

// The size to allocate for a grow-only buffer of `have` bytes that must hold
// `need`: a quarter more than asked, and at least twice what it had.
inline size_t grow_bytes(size_t need, size_t have)
{
    const size_t want = need + need / 4 + 4096;
    return have && want < 2 * have ? 2 * have : want;
}
// Pinned host memory of one thread (hipHostMalloc), released at thread exit.
// Mapped and coherent (pinned_mapped): kernels write it directly.
struct PinnedBuffer {
    void *ptr = nullptr;
    void *dptr = nullptr;  // its device address
    ~PinnedBuffer();
};
// Gives b `bytes` of mapped, coherent pinned memory, unless it has some.
int pinned_mapped(PinnedBuffer &b, size_t bytes);
// Allocates mapped, coherent pinned memory: *host, and its device address.
int host_alloc_mapped(size_t bytes, void **host, void **dev);
// Copies n ints of device memory to mapped pinned memory, given by its device
// address, with a kernel on stream s (one workgroup per 256 ints, at most
// 256).  Under load from other
// threads a device-to-host hipMemcpyAsync can block in the runtime for
// milliseconds before it is even queued; a kernel launch does not.
int copy_ints_to_host(const int *src, int *dst, int n, hipStream_t s);
bool on_main_thread();
int workspace_reserve(Workspace &ws, size_t bytes);
Workspace &thread_workspace(int slot);

// ---- pixel types ---------------------------------------------------------
inline size_t dtype_size(int dt)
{
    switch (dt) {
    case BGT_U8: case BGT_I8: case BGT_BOOL: return 1;
    case BGT_U16: case BGT_I16: return 2;
    case BGT_U32: case BGT_I32: case BGT_F32: return 4;
    case BGT_F64: return 8;
    default: return 0;
    }
}

inline size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
inline int pow2_ceil(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace bgt
