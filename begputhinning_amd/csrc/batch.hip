// batch.hip -- many images and volumes of uneven size thinned together.
//
// All items of a batch live in one device allocation.  A pass launch covers
// the bands of every 2D item (one launch) and the bricks of every 3D item (one
// launch per 3D line-width class, as the kernel is specialised on the register
// words per x-line); items that have converged make their tiles exit at once,
// and every item keeps its own pass counter, change flags and iteration count
// on the device.  The host only polls one pair of live-item counters after each
// chunk of passes, and the chunk can be replayed as a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "batch.h"
#include "common.h"
#include "thin2d.h"
#include "thin2d_kernels.h"
#include "thin3d.h"
#include "thin3d_kernels.h"

namespace bgt {

struct IO2D {
    const uint8_t *in;
    size_t in_pitch;
    uint8_t *out;
    size_t out_pitch;
};

struct IO3D {
    const uint8_t *in;
    uint8_t *out;
};

namespace {

__device__ __forceinline__ int find_item(const long *off, int n, long g)
{
    int lo = 0, hi = n - 1;  // largest i with off[i] <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t nz4b(uint32_t x)
{
    uint32_t t = x | (x >> 4);
    t |= t >> 2;
    t |= t >> 1;
    t &= 0x01010101u;
    return (t | (t >> 7) | (t >> 14) | (t >> 21)) & 0xfu;
}

__device__ __forceinline__ uint32_t spread4b(uint32_t b)
{
    return ((b & 0xfu) * 0x00204081u) & 0x01010101u;
}

// 32 bytes -> 32 bits (p 16-byte aligned)
__device__ __forceinline__ uint32_t pack32(const uint8_t *p)
{
    const uint4 a = reinterpret_cast<const uint4 *>(p)[0];
    const uint4 b = reinterpret_cast<const uint4 *>(p)[1];
    return nz4b(a.x) | nz4b(a.y) << 4 | nz4b(a.z) << 8 | nz4b(a.w) << 12 | nz4b(b.x) << 16 |
           nz4b(b.y) << 20 | nz4b(b.z) << 24 | nz4b(b.w) << 28;
}

__device__ __forceinline__ void unpack32(uint32_t v, uint8_t *p)
{
    uint4 *q = reinterpret_cast<uint4 *>(p);
    q[0] = make_uint4(spread4b(v), spread4b(v >> 4), spread4b(v >> 8), spread4b(v >> 12));
    q[1] = make_uint4(spread4b(v >> 16), spread4b(v >> 20), spread4b(v >> 24), spread4b(v >> 28));
}

__device__ __forceinline__ uint32_t pack_bytes(const uint8_t *row, int x0, int w)
{
    if (x0 + 32 <= w && ((((uintptr_t)(row + x0)) & 15) == 0)) return pack32(row + x0);
    uint32_t v = 0;
    for (int k = 0; k < 32 && x0 + k < w; ++k)
        if (row[x0 + k]) v |= 1u << k;
    return v;
}

__device__ __forceinline__ void unpack_bytes(uint32_t v, uint8_t *row, int x0, int w)
{
    if (x0 + 32 <= w && ((((uintptr_t)(row + x0)) & 15) == 0)) {
        unpack32(v, row + x0);
        return;
    }
    for (int k = 0; k < 32 && x0 + k < w; ++k) row[x0 + k] = (uint8_t)((v >> k) & 1u);
}

// one thread per 32-bit half of a 64-bit row word
__global__ void batch_pack2d(const Item2D *__restrict__ items, const IO2D *__restrict__ io,
                             const long *__restrict__ off, int n, long total)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int i = find_item(off, n, g);
    const Item2D it = items[i];
    const long local = g - off[i];  // half-word index
    const long wi = local >> 1;
    const int half = (int)(local & 1);
    const int y = (int)(wi >> it.lgL), c = (int)(wi & (it.L - 1));
    const uint8_t *row = io[i].in + (size_t)y * io[i].in_pitch;
    const uint32_t v = pack_bytes(row, c * 64 + half * 32, it.w);
    reinterpret_cast<uint32_t *>(it.buf[0])[local] = v;
}

__global__ void batch_unpack2d(const Item2D *__restrict__ items, const IO2D *__restrict__ io,
                               const long *__restrict__ off, int n, long total)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int i = find_item(off, n, g);
    const Item2D it = items[i];
    const long local = g - off[i];
    const long wi = local >> 1;
    const int half = (int)(local & 1);
    const int y = (int)(wi >> it.lgL), c = (int)(wi & (it.L - 1));
    const int x0 = c * 64 + half * 32;
    if (x0 >= it.w) return;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(it.st[ST_PARITY] ? it.buf[1] : it.buf[0]);
    unpack_bytes(src[local], io[i].out + (size_t)y * io[i].out_pitch, x0, it.w);
}

__global__ void batch_pack3d(const Item3D *__restrict__ items, const IO3D *__restrict__ io,
                             const long *__restrict__ off, int n, long total)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int i = find_item(off, n, g);
    const Item3D it = items[i];
    const long local = g - off[i];
    const long line = local / it.nh;
    const int j = (int)(local - line * it.nh);
    it.buf[0][local] = pack_bytes(io[i].in + (size_t)line * it.w, j * 32, it.w);
}

__global__ void batch_unpack3d(const Item3D *__restrict__ items, const IO3D *__restrict__ io,
                               const long *__restrict__ off, int n, long total)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int i = find_item(off, n, g);
    const Item3D it = items[i];
    const long local = g - off[i];
    const long line = local / it.nh;
    const int j = (int)(local - line * it.nh);
    if (j * 32 >= it.w) return;
    const uint32_t *src = it.st[ST_PARITY] ? it.buf[1] : it.buf[0];
    unpack_bytes(src[local], io[i].out + (size_t)line * it.w, j * 32, it.w);
}

// zero every item state and the launch counters; live[0]/live[1] = item counts
__global__ void batch_init(int *states, long nstate, unsigned *ctrs, int nctr, int *live, int n2,
                           int n3)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nstate) states[g] = 0;
    if (g < nctr) ctrs[g] = 0;
    if (g == 0) {
        live[0] = n2;
        live[1] = n3;
    }
}

struct Carver {
    char *base;
    size_t used = 0;
    template <typename T>
    T *take(size_t bytes)
    {
        T *r = reinterpret_cast<T *>(base ? base + used : nullptr);
        used += round_up(bytes ? bytes : 1, 256);
        return r;
    }
};

}  // namespace

struct Batch {
    Config2D c2;
    Config3D c3;
    struct In2 {
        IO2D io;
        int w, h;
    };
    struct In3 {
        IO3D io;
        int w, h, d;
    };
    std::vector<In2> in2;
    std::vector<In3> in3;

    // prepared state
    bool ready = false;
    void *mem = nullptr;
    size_t bytes = 0;
    std::vector<int> order3;  // 3D items sorted by class: index into in3
    struct Class3 {
        int NH;
        int first, count;     // range in the sorted 3D item array
        Tile3D *tiles;
        int ntiles;
    };
    std::vector<Class3> classes;
    Item2D *d_items2 = nullptr;
    IO2D *d_io2 = nullptr;
    long *d_off2 = nullptr;
    Tile2D *d_tiles2 = nullptr;
    int ntiles2 = 0;
    long words2 = 0;
    Item3D *d_items3 = nullptr;
    IO3D *d_io3 = nullptr;
    long *d_off3 = nullptr;
    long words3 = 0;
    int *d_states = nullptr;
    long nstate = 0;
    unsigned *d_ctr = nullptr;
    int *d_live = nullptr;
    int *h_live = nullptr;  // pinned
    std::vector<int> iters;
    // optional per-launch kernel timing (direct launches only)
    bool timing = false;
    std::vector<hipEvent_t> events;
    size_t ev_used = 0;
    std::vector<std::pair<size_t, int>> ev_marks;  // (first event index, kind 0=2D 1=3D)
    double ms2d = 0, ms3d = 0;
    int n2d = 0, n3d = 0;
    // captured chunk of passes
    hipGraphExec_t graph = nullptr;
    hipStream_t graph_stream = nullptr;
    int graph_chunk = 0;

    ~Batch() { release(); }
    void release()
    {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        events.clear();
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        if (mem) (void)hipFree(mem);
        mem = nullptr;
        if (h_live) (void)hipHostFree(h_live);
        h_live = nullptr;
        ready = false;
    }
};

static int prepare(Batch &b)
{
    if (b.ready) return BGT_OK;
    if (b.in2.empty() && b.in3.empty()) return fail(BGT_ERR_INVALID, "empty batch");
    BGT_CHECK(check_config2d(b.c2));
    const int n2 = (int)b.in2.size(), n3 = (int)b.in3.size();
    std::vector<Plan2D> p2(n2);
    std::vector<Plan3D> p3(n3);
    for (int i = 0; i < n2; ++i) BGT_CHECK(plan2d(b.in2[i].w, b.in2[i].h, b.c2, p2[i]));
    for (int i = 0; i < n3; ++i) BGT_CHECK(plan3d(b.in3[i].w, b.in3[i].h, b.in3[i].d, b.c3, p3[i]));

    // 3D items grouped by kernel class
    b.order3.resize(n3);
    std::iota(b.order3.begin(), b.order3.end(), 0);
    std::stable_sort(b.order3.begin(), b.order3.end(),
                     [&](int a, int c) { return p3[a].NH < p3[c].NH; });
    b.classes.clear();
    for (int k = 0; k < n3; ++k) {
        const int NH = p3[b.order3[k]].NH;
        if (b.classes.empty() || b.classes.back().NH != NH) b.classes.push_back({NH, k, 0, nullptr, 0});
        b.classes.back().count++;
    }

    // 2D tiles: largest images first so that the launch tail holds small work
    std::vector<int> ord2(n2);
    std::iota(ord2.begin(), ord2.end(), 0);
    std::stable_sort(ord2.begin(), ord2.end(),
                     [&](int a, int c) { return p2[a].state_words > p2[c].state_words; });
    std::vector<Tile2D> t2;
    for (int i : ord2)
        for (int k = 0; k < p2[i].nbands; ++k) t2.push_back({i, k});
    std::vector<std::vector<Tile3D>> t3(b.classes.size());
    for (size_t c = 0; c < b.classes.size(); ++c)
        for (int k = 0; k < b.classes[c].count; ++k) {
            const int it = b.order3[b.classes[c].first + k];
            for (int br = 0; br < p3[it].nbricks; ++br) t3[c].push_back({k, br});
        }

    // layout (two sweeps: size, then carve)
    for (int sweep = 0; sweep < 2; ++sweep) {
        Carver cv{sweep ? (char *)b.mem : nullptr};
        b.d_items2 = cv.take<Item2D>(sizeof(Item2D) * n2);
        b.d_io2 = cv.take<IO2D>(sizeof(IO2D) * n2);
        b.d_off2 = cv.take<long>(sizeof(long) * (n2 + 1));
        b.d_tiles2 = cv.take<Tile2D>(sizeof(Tile2D) * t2.size());
        b.d_items3 = cv.take<Item3D>(sizeof(Item3D) * n3);
        b.d_io3 = cv.take<IO3D>(sizeof(IO3D) * n3);
        b.d_off3 = cv.take<long>(sizeof(long) * (n3 + 1));
        for (size_t c = 0; c < b.classes.size(); ++c)
            b.classes[c].tiles = cv.take<Tile3D>(sizeof(Tile3D) * t3[c].size());
        b.nstate = (long)(n2 + n3) * ST_INTS;
        b.d_states = cv.take<int>(sizeof(int) * b.nstate);
        b.d_ctr = cv.take<unsigned>(sizeof(unsigned) * (1 + b.classes.size()));
        b.d_live = cv.take<int>(sizeof(int) * 2);

        std::vector<Item2D> h2(n2);
        std::vector<long> off2(n2 + 1, 0);
        for (int i = 0; i < n2; ++i) {
            Item2D &v = h2[i];
            v.buf[0] = cv.take<uint64_t>(p2[i].state_words * sizeof(uint64_t));
            v.buf[1] = cv.take<uint64_t>(p2[i].state_words * sizeof(uint64_t));
            v.chg = cv.take<uint8_t>((size_t)2 * p2[i].nbands);
            v.st = b.d_states ? b.d_states + (size_t)i * ST_INTS : nullptr;
            v.w = p2[i].w; v.h = p2[i].h; v.L = p2[i].L; v.lgL = p2[i].lgL;
            v.nbands = p2[i].nbands;
            v.band_rows = p2[i].band_rows;
            off2[i + 1] = off2[i] + (long)p2[i].state_words * 2;  // 32-bit halves
        }
        std::vector<Item3D> h3(n3);
        std::vector<long> off3(n3 + 1, 0);
        std::vector<IO3D> io3(n3);
        for (int k = 0; k < n3; ++k) {
            const int i = b.order3[k];
            Item3D &v = h3[k];
            v.buf[0] = cv.take<uint32_t>(p3[i].state_words * sizeof(uint32_t));
            v.buf[1] = cv.take<uint32_t>(p3[i].state_words * sizeof(uint32_t));
            v.chg = cv.take<uint8_t>((size_t)2 * p3[i].nbricks);
            v.st = b.d_states ? b.d_states + (size_t)(n2 + k) * ST_INTS : nullptr;
            v.w = p3[i].w; v.h = p3[i].h; v.d = p3[i].d; v.nh = p3[i].nh;
            v.iy = p3[i].iy; v.iz = p3[i].iz;
            v.nby = p3[i].nby; v.nbz = p3[i].nbz; v.nbricks = p3[i].nbricks;
            off3[k + 1] = off3[k] + (long)p3[i].state_words;
            io3[k] = b.in3[i].io;
        }
        if (sweep == 0) {
            b.bytes = cv.used;
            hipError_t e = hipMalloc(&b.mem, b.bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                b.mem = nullptr;
                return fail(BGT_ERR_NOMEM, "batch allocation of %zu bytes failed", b.bytes);
            }
            continue;
        }
        std::vector<IO2D> io2(n2);
        for (int i = 0; i < n2; ++i) io2[i] = b.in2[i].io;
        b.words2 = off2[n2];
        b.words3 = off3[n3];
        b.ntiles2 = (int)t2.size();
        BGT_HIP(hipMemcpy(b.d_items2, h2.data(), sizeof(Item2D) * n2, hipMemcpyHostToDevice));
        BGT_HIP(hipMemcpy(b.d_io2, io2.data(), sizeof(IO2D) * n2, hipMemcpyHostToDevice));
        BGT_HIP(hipMemcpy(b.d_off2, off2.data(), sizeof(long) * (n2 + 1), hipMemcpyHostToDevice));
        if (!t2.empty())
            BGT_HIP(hipMemcpy(b.d_tiles2, t2.data(), sizeof(Tile2D) * t2.size(), hipMemcpyHostToDevice));
        BGT_HIP(hipMemcpy(b.d_items3, h3.data(), sizeof(Item3D) * n3, hipMemcpyHostToDevice));
        BGT_HIP(hipMemcpy(b.d_io3, io3.data(), sizeof(IO3D) * n3, hipMemcpyHostToDevice));
        BGT_HIP(hipMemcpy(b.d_off3, off3.data(), sizeof(long) * (n3 + 1), hipMemcpyHostToDevice));
        for (size_t c = 0; c < b.classes.size(); ++c) {
            b.classes[c].ntiles = (int)t3[c].size();
            BGT_HIP(hipMemcpy(b.classes[c].tiles, t3[c].data(), sizeof(Tile3D) * t3[c].size(),
                              hipMemcpyHostToDevice));
        }
    }
    if (!b.h_live && hipHostMalloc((void **)&b.h_live, 64, hipHostMallocDefault) != hipSuccess)
        return fail(BGT_ERR_NOMEM, "pinned allocation failed");
    b.iters.assign(n2 + n3, 0);
    b.ready = true;
    return BGT_OK;
}

static int timed_mark(Batch &b, hipStream_t s, int kind, bool begin)
{
    if (!b.timing) return BGT_OK;
    if (b.ev_used == b.events.size()) {
        for (int i = 0; i < 64; ++i) {
            hipEvent_t e;
            BGT_HIP(hipEventCreate(&e));
            b.events.push_back(e);
        }
    }
    if (begin) b.ev_marks.push_back({b.ev_used, kind});
    BGT_HIP(hipEventRecord(b.events[b.ev_used++], s));
    return BGT_OK;
}

static int launch_one_pass(Batch &b, hipStream_t s)
{
    const int n2 = (int)b.in2.size();
    if (n2) {
        BGT_CHECK(timed_mark(b, s, 0, true));
        BGT_CHECK(launch_pass2d(b.c2, b.ntiles2, b.d_items2, b.d_tiles2, n2, b.d_ctr, b.d_live, s));
        BGT_CHECK(timed_mark(b, s, 0, false));
    }
    for (size_t c = 0; c < b.classes.size(); ++c) {
        const Batch::Class3 &k = b.classes[c];
        BGT_CHECK(timed_mark(b, s, 1, true));
        BGT_CHECK(launch_pass3d_nh(b.c3, k.NH, k.ntiles, b.d_items3 + k.first, k.tiles, k.count,
                                   b.d_ctr + 1 + c, b.d_live + 1, s));
        BGT_CHECK(timed_mark(b, s, 1, false));
    }
    return BGT_OK;
}

static int collect_timing(Batch &b)
{
    b.ms2d = b.ms3d = 0;
    b.n2d = b.n3d = 0;
    for (auto &m : b.ev_marks) {
        float ms = 0;
        BGT_HIP(hipEventElapsedTime(&ms, b.events[m.first], b.events[m.first + 1]));
        if (m.second == 0) { b.ms2d += ms; b.n2d++; }
        else { b.ms3d += ms; b.n3d++; }
    }
    b.ev_marks.clear();
    b.ev_used = 0;
    return BGT_OK;
}

int batch_run(Batch &b, hipStream_t s, bool use_graph)
{
    BGT_CHECK(prepare(b));
    const int n2 = (int)b.in2.size(), n3 = (int)b.in3.size();
    const long init_n = std::max<long>(b.nstate, 1 + (long)b.classes.size());
    hipLaunchKernelGGL(batch_init, dim3((unsigned)((init_n + 255) / 256)), dim3(256), 0, s,
                       b.d_states, b.nstate, b.d_ctr, (int)(1 + b.classes.size()), b.d_live, n2, n3);
    if (b.words2)
        hipLaunchKernelGGL(batch_pack2d, dim3((unsigned)((b.words2 + 255) / 256)), dim3(256), 0, s,
                           b.d_items2, b.d_io2, b.d_off2, n2, b.words2);
    if (b.words3)
        hipLaunchKernelGGL(batch_pack3d, dim3((unsigned)((b.words3 + 255) / 256)), dim3(256), 0, s,
                           b.d_items3, b.d_io3, b.d_off3, n3, b.words3);
    BGT_HIP(hipGetLastError());

    const int chunk = 4;
    if (b.timing) use_graph = false;
    b.ev_marks.clear();
    b.ev_used = 0;
    if (use_graph && (!b.graph || b.graph_stream != s || b.graph_chunk != chunk)) {
        if (b.graph) (void)hipGraphExecDestroy(b.graph);
        b.graph = nullptr;
        hipStream_t cap = nullptr;
        BGT_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        BGT_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        int rc = BGT_OK;
        for (int i = 0; i < chunk && rc == BGT_OK; ++i) rc = launch_one_pass(b, cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (rc != BGT_OK) return rc;
        if (e != hipSuccess) return fail(BGT_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&b.graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(BGT_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
        b.graph_stream = s;
        b.graph_chunk = chunk;
    }
    long max_vox = 0;
    for (auto &v : b.in2) max_vox = std::max<long>(max_vox, (long)v.w * v.h);
    for (auto &v : b.in3) max_vox = std::max<long>(max_vox, (long)v.w * v.h * v.d);
    const long max_pass = max_vox + 2;
    long passes = 0;
    for (;;) {
        if (use_graph) {
            BGT_HIP(hipGraphLaunch(b.graph, s));
        } else {
            for (int i = 0; i < chunk; ++i) BGT_CHECK(launch_one_pass(b, s));
        }
        passes += chunk;
        BGT_HIP(hipMemcpyAsync(b.h_live, b.d_live, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        BGT_HIP(hipStreamSynchronize(s));
        if (b.h_live[0] <= 0 && b.h_live[1] <= 0) break;
        if (passes > max_pass) return fail(BGT_ERR_NOT_CONVERGED, "batch did not converge");
    }
    if (b.words2)
        hipLaunchKernelGGL(batch_unpack2d, dim3((unsigned)((b.words2 + 255) / 256)), dim3(256), 0, s,
                           b.d_items2, b.d_io2, b.d_off2, n2, b.words2);
    if (b.words3)
        hipLaunchKernelGGL(batch_unpack3d, dim3((unsigned)((b.words3 + 255) / 256)), dim3(256), 0, s,
                           b.d_items3, b.d_io3, b.d_off3, n3, b.words3);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(s));
    if (b.timing) BGT_CHECK(collect_timing(b));
    return BGT_OK;
}

void batch_set_timing(Batch &b, bool on) { b.timing = on; }

void batch_kernel_time(const Batch &b, double *ms2d, int *n2d, double *ms3d, int *n3d)
{
    *ms2d = b.ms2d; *n2d = b.n2d; *ms3d = b.ms3d; *n3d = b.n3d;
}

int batch_iterations(Batch &b, int *out)
{
    if (!b.ready) return fail(BGT_ERR_INVALID, "batch has not run");
    const int n2 = (int)b.in2.size(), n3 = (int)b.in3.size();
    std::vector<int> st((size_t)(n2 + n3) * ST_INTS);
    BGT_HIP(hipMemcpy(st.data(), b.d_states, st.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < n2; ++i) out[i] = st[(size_t)i * ST_INTS + ST_ITERS];
    for (int k = 0; k < n3; ++k) out[n2 + b.order3[k]] = st[(size_t)(n2 + k) * ST_INTS + ST_ITERS];
    return BGT_OK;
}

Batch *batch_new() { return new Batch(); }
void batch_delete(Batch *b) { delete b; }

int batch_add2d(Batch &b, const uint8_t *in, size_t in_pitch, int w, int h, uint8_t *out,
                size_t out_pitch)
{
    if (!in || !out) return fail(BGT_ERR_INVALID, "null device pointer");
    if (w <= 0 || h <= 0 || in_pitch < (size_t)w || out_pitch < (size_t)w)
        return fail(BGT_ERR_INVALID, "bad 2D item (%d x %d)", w, h);
    if (w > 4096) return fail(BGT_ERR_UNSUPPORTED, "2D width %d exceeds 4096", w);
    b.release();
    b.in2.push_back({{in, in_pitch, out, out_pitch}, w, h});
    return BGT_OK;
}

int batch_add3d(Batch &b, const uint8_t *in, int w, int h, int d, uint8_t *out)
{
    if (!in || !out) return fail(BGT_ERR_INVALID, "null device pointer");
    if (w <= 0 || h <= 0 || d <= 0) return fail(BGT_ERR_INVALID, "bad 3D item");
    b.release();
    b.in3.push_back({{in, out}, w, h, d});
    return BGT_OK;
}

int batch_set_config(Batch &b, const Config2D &c2, const Config3D &c3)
{
    BGT_CHECK(check_config2d(c2));
    b.release();
    b.c2 = c2;
    b.c3 = c3;
    return BGT_OK;
}

size_t batch_device_bytes(Batch &b)
{
    return b.ready ? b.bytes : 0;
}

int batch_size(const Batch &b) { return (int)(b.in2.size() + b.in3.size()); }

}  // namespace bgt
