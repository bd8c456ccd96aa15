// thin2d.hip -- 2D thinning on gfx950: pack/unpack kernels, the fused pass
// kernel and the host drivers (per-image drop-in and device-resident).
//
// Reference path replaced: 2D/LibThinning2D.cu:39-130 (host driver),
// 2D/CUDA_ThinningKernelMultiBlock2D.cuh:68-170 (driver kernel),
// 2D/CUDA_SubiterationKernels2D.cuh (rules), 2D/CUDA_DeviceHelperFunctions.cuh
// (pack/unpack, lane shifts, reductions).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.h"
#include "thin2d.h"
#include "pack.h"
#include "queue2d.h"
#include "thin2d_kernels.h"

namespace bgt {

// ---------------------------------------------------------------------------
// pixel thresholds of the reference
// ---------------------------------------------------------------------------

// 2D rule: (unsigned)v > 0 with the GPU's saturating float->unsigned
// conversion (2D/CUDA_DeviceHelperFunctions.cuh:101): integers v != 0,
// floating point v >= 1 (NaN and negatives are background).
template <typename T>
__device__ __forceinline__ bool fg2d(T v)
{
    return v != 0;
}
template <>
__device__ __forceinline__ bool fg2d<float>(float v)
{
    return v >= 1.0f;
}
template <>
__device__ __forceinline__ bool fg2d<double>(double v)
{
    return v >= 1.0;
}

template <typename T>
__device__ __forceinline__ T one_value()
{
    return (T)1;
}

// ---------------------------------------------------------------------------
// pack / unpack
// ---------------------------------------------------------------------------

// uint8 rows (pitch bytes) -> packed words; one thread per 64-bit word
// Per-image pack: one thread per 64-pixel run with its own aligned loads (the
// wave-coalesced pack_runs_wave_u8 pays off on whole batches, but its LDS
// round trip costs more than it saves on a single latency-bound frame).
__global__ __launch_bounds__(256) void pack2d_u8(const uint8_t *__restrict__ in, size_t pitch,
                                                 int w, int h, int L, uint64_t *__restrict__ out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)L * h) return;
    const int y = (int)(idx / L), x0 = (int)(idx - (long)y * L) * 64;  // x0 < w
    out[idx] = lane_order(pack_run_u8(in + (size_t)y * pitch + x0, min(64, w - x0)));
}

template <typename T>
__global__ void pack2d_t(const T *__restrict__ in, int w, int h, int L, uint64_t *__restrict__ out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)L * h) return;
    const int y = (int)(idx / L), c = (int)(idx % L);
    const T *row = in + (size_t)y * w;
    uint64_t v = 0;
    for (int k = 0; k < 64; ++k) {
        int x = c * 64 + k;
        if (x < w && fg2d<T>(row[x])) v |= 1ull << lane_bit(k);
    }
    out[idx] = v;
}

// packed words (from the buffer the item's final parity selects) -> uint8 0/1,
// one thread per 16-byte chunk of the output
__global__ __launch_bounds__(256) void unpack2d_u8(const Item2D *__restrict__ item,
                                                   uint8_t *__restrict__ out, size_t pitch,
                                                   uint32_t uchunks, int flat)
{
    const Item2D it = *item;
    const uint64_t *buf = it.st[ST_PARITY] ? it.buf[1] : it.buf[0];
    unpack_image_u8(buf, (uint32_t)it.L, it.w, it.h, out, pitch, uchunks, flat,
                    blockIdx.x * 256u + threadIdx.x);
}

// launch of unpack2d_u8 for a w x h image with row pitch `pitch`
void launch_unpack_item2d(const Item2D *d_item, uint8_t *out, int w, int h, size_t pitch,
                          hipStream_t s)
{
    int flat = 0;
    uint32_t uchunks = 0;
    long threads = 0;
    unpack_image_grid(w, h, pitch, flat, uchunks, threads);
    hipLaunchKernelGGL(unpack2d_u8, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s,
                       d_item, out, pitch, uchunks, flat);
}

template <typename T>
__global__ void unpack2d_t(const Item2D *__restrict__ item, T *__restrict__ out)
{
    const Item2D it = *item;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)it.L * it.h) return;
    const int y = (int)(idx / it.L), c = (int)(idx % it.L);
    const uint64_t v = (it.st[ST_PARITY] ? it.buf[1] : it.buf[0])[idx];
    T *row = out + (size_t)y * it.w;
    for (int k = 0; k < 64; ++k) {
        int x = c * 64 + k;
        if (x < it.w) row[x] = ((v >> lane_bit(k)) & 1) ? one_value<T>() : (T)0;
    }
}

// ---------------------------------------------------------------------------
// item set-up (kernel arguments only: no host staging, no memset node)
// ---------------------------------------------------------------------------
__global__ void init_item2d(Item2D *item, Item2D val, unsigned *ctr)
{
    const int t = threadIdx.x;
    if (t == 0) *item = val;
    if (t < ST_INTS) val.st[t] = 0;
    if (t == 0 && ctr) *ctr = 0;
    for (int k = t; k < 2 * val.nbands; k += blockDim.x) val.chg[k] = 0;  // flags only get set
}

// ---------------------------------------------------------------------------
// the fused pass kernel
// ---------------------------------------------------------------------------

// Per-image pass kernel: kp workgroups per band (part = blockIdx % kp), the
// last workgroup of the launch advances the item.  With kp > 1 a band is
// split into parts that run in windows of 4 or 8 register rows per segment
// (the finest whose parts number at most kp), each with full 4K-row halos, so
// the band's result is the same while a pass's latency, bound by the rows one
// wave works through per subiteration, drops; where no such window fits,
// part 0 runs the whole band.
template <int R, int V>
__global__ __launch_bounds__(512) void thin2d_pass(const Item2D *__restrict__ items, int K,
                                                    unsigned *__restrict__ ctr, int kp)
{
    extern __shared__ uint64_t edges[];
    __shared__ TileInfo info;
    __shared__ int s_last;

    const int tid = threadIdx.x;
    const Item2D it = items[0];
    const int band = blockIdx.x / kp, part = blockIdx.x % kp;
    bool work = !it.st[ST_DONE];
    int p = 0, par = 0;
    if (work) {
        p = it.st[ST_PASS];
        par = p & 1;
        if (p > 0) {
            const uint8_t *c = it.chg + (par ^ 1) * it.nbands;
            work = c[band] || (band > 0 && c[band - 1]) || (band + 1 < it.nbands && c[band + 1]);
        }
    }
    int row0 = band * it.band_rows, nrows = it.band_rows, win = R;
    if (work && kp > 1) {
        const int seg = (int)(blockDim.x >> 6) * (64 / lanes_for(it.L, V));  // segments per window
        const int in4 = seg * 4 - 8 * K, in8 = seg * 8 - 8 * K;  // interior rows of a part
        const int np4 = in4 >= 1 ? (it.band_rows + in4 - 1) / in4 : kp + 1;
        const int np8 = in8 >= 1 ? (it.band_rows + in8 - 1) / in8 : kp + 1;
        int inner = 0, np = 1;
        if (np4 <= kp) {
            win = 4;
            inner = in4;
            np = np4;
        } else if (R > 8 && np8 <= kp) {
            win = 8;
            inner = in8;
            np = np8;
        }
        if (win != R) {
            row0 += part * inner;
            nrows = min(inner, it.band_rows - part * inner);
        }
        work = part < np && row0 < it.h;
    }
    if (work) {
        int run = 0;
        if (win == 4)
            band_work<4, V>(it, row0, nrows, false, K, par, edges, &info);
        else if (win == 8 && R > 8)
            band_work<8, V>(it, row0, nrows, false, K, par, edges, &info);
        else
            run = band_work<R, V>(it, row0, nrows, it.nbands == 1, K, par, edges, &info);
        if (tid == 0) {
            // flags are cleared before the pass (init, then each pass's last
            // workgroup), so the parts of a band only set them
            if (info.mask) it.chg[par * it.nbands + band] = 1;
            if (info.mask) atomicOr((unsigned *)&it.st[ST_MASK], info.mask);
            if (it.nbands == 1) it.st[ST_RUN] = run + 1;
        }
    }
    if (tid == 0) {
        __threadfence();
        s_last = atomicAdd(ctr, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    // every workgroup of this pass has read the previous pass's flags: clear
    // them for the next pass to set
    if (!it.st[ST_DONE])
        for (int k = tid; k < it.nbands; k += blockDim.x) it.chg[(par ^ 1) * it.nbands + k] = 0;
    if (tid == 0) {
        __threadfence();
        finalize_state(it.st, K, it.st[ST_PASS], nullptr);
        __threadfence();
        atomicExch(ctr, 0u);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

int band_rows_for(int L, int V, const Config2D &cfg)
{
    const int G = 64 / lanes_for(L, V);
    return cfg.waves * G * cfg.rows - 8 * cfg.iters_per_pass;
}

int check_config2d(const Config2D &cfg)
{
    if (cfg.waves < 1 || cfg.waves > 8 || (cfg.rows != 8 && cfg.rows != 16) ||
        cfg.iters_per_pass < 1 || cfg.iters_per_pass > 16 || cfg.blocks_per_cu < -1 ||
        cfg.blocks_per_cu > 32 ||
        (cfg.lane_words != 0 && cfg.lane_words != 1 && cfg.lane_words != 2 && cfg.lane_words != 4) ||
        cfg.fused_io < 0 || cfg.fused_io > 1 || cfg.persistent < 0 || cfg.persistent > 1)
        return fail(BGT_ERR_INVALID,
                    "bad 2D config (waves=%d rows=%d k=%d blocks/CU=%d lane words=%d fused io=%d)",
                    cfg.waves, cfg.rows, cfg.iters_per_pass, cfg.blocks_per_cu, cfg.lane_words,
                    cfg.fused_io);
    return BGT_OK;
}

int lane_words2d(int w, const Config2D &cfg)
{
    const int L = (w + 63) / 64;
    const int need = L <= 64 ? 1 : L <= 128 ? 2 : L <= 256 ? 4 : 0;
    if (!need) return 0;
    return std::max(need, cfg.lane_words > 0 ? cfg.lane_words : 1);
}

// the per-image pass kernel for `rows` register rows and V words per lane
static const void *image_pass_fn(int rows, int V)
{
    if (V == 4) return rows == 8 ? (const void *)thin2d_pass<8, 4> : (const void *)thin2d_pass<16, 4>;
    if (V == 2) return rows == 8 ? (const void *)thin2d_pass<8, 2> : (const void *)thin2d_pass<16, 2>;
    return rows == 8 ? (const void *)thin2d_pass<8, 1> : (const void *)thin2d_pass<16, 1>;
}

// LDS of a pass workgroup: the edge rows, [2 buffers][top/bottom][lanes][V words]
size_t pass_lds_bytes2d(const Config2D &cfg, int V)
{
    return (size_t)2 * 2 * cfg.waves * 64 * V * sizeof(uint64_t);
}

static int launch_image_pass2d(const Config2D &cfg, int V, int nbands, int kp, const Item2D *item,
                               unsigned *ctr, hipStream_t s)
{
    const void *fn = image_pass_fn(cfg.rows, V);
    const dim3 grid(nbands * kp), blk(cfg.waves * 64);
    int K = cfg.iters_per_pass;
    void *args[] = {(void *)&item, (void *)&K, (void *)&ctr, (void *)&kp};
    BGT_HIP(hipLaunchKernel(fn, grid, blk, args, pass_lds_bytes2d(cfg, V), s));
    return BGT_OK;
}

// Workgroups per band of the per-image pass: a multi-band image spreads its
// bands over the CUs (at most 16 parts a band); a single-band image runs to
// convergence in one workgroup.
static int image_parts(int nbands)
{
    static const int cus = [] {
        int dev = 0, c = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            c > 0)
            return c;
        return 256;
    }();
    if (nbands <= 1) return 1;
    return std::max(1, std::min(16, cus / nbands));
}

int plan2d(int w, int h, const Config2D &cfg, Plan2D &pl, int V)
{
    if (w <= 0 || h <= 0) return fail(BGT_ERR_INVALID, "2D size must be >= 1 (got %d x %d)", w, h);
    if (w > kMaxWidth2D)
        return fail(BGT_ERR_UNSUPPORTED,
                    "2D width %d exceeds %d (one wave64 of lanes of four 64-bit words)", w, kMaxWidth2D);
    BGT_CHECK(check_config2d(cfg));
    const int need = lane_words2d(w, cfg);
    if (V == 0) V = need;
    if (V != 1 && V != 2 && V != 4) return fail(BGT_ERR_INVALID, "2D lane words %d not 1, 2 or 4", V);
    if (lanes_for((w + 63) / 64, V) > 64)
        return fail(BGT_ERR_UNSUPPORTED, "2D width %d needs more than %d words per lane", w, V);
    pl.w = w;
    pl.h = h;
    pl.L = (w + 63) / 64;  // words per row; a wave holds floor(64 / ceil(L / V)) row segments
    pl.V = V;
    pl.band_rows = band_rows_for(pl.L, V, cfg);
    if (pl.band_rows < 4 * cfg.iters_per_pass)
        return fail(BGT_ERR_INVALID, "2D band of %d rows is thinner than its halo %d",
                    pl.band_rows, 4 * cfg.iters_per_pass);
    pl.nbands = (h + pl.band_rows - 1) / pl.band_rows;
    pl.state_words = (size_t)pl.L * h;
    return BGT_OK;
}

size_t workspace_bytes2d(const Plan2D &pl)
{
    size_t b = 0;
    b += round_up(sizeof(Item2D), 256);
    b += round_up(ST_INTS * sizeof(int), 256);
    b += round_up(sizeof(unsigned), 256);
    b += 2 * round_up(pl.state_words * sizeof(uint64_t), 256);
    b += round_up((size_t)2 * pl.nbands, 256);
    return b;
}

struct Carve {
    char *p;
    template <typename T>
    T *take(size_t bytes)
    {
        T *r = reinterpret_cast<T *>(p);
        p += round_up(bytes, 256);
        return r;
    }
};

int setup_item2d(const Plan2D &pl, void *ws, hipStream_t s, Item2D **d_item, unsigned **d_ctr,
                 Item2D *host_item)
{
    Carve cv{reinterpret_cast<char *>(ws)};
    Item2D *item = cv.take<Item2D>(sizeof(Item2D));
    int *st = cv.take<int>(ST_INTS * sizeof(int));
    unsigned *ctr = cv.take<unsigned>(sizeof(unsigned));
    Item2D v{};
    v.buf[0] = cv.take<uint64_t>(pl.state_words * sizeof(uint64_t));
    v.buf[1] = cv.take<uint64_t>(pl.state_words * sizeof(uint64_t));
    v.chg = cv.take<uint8_t>((size_t)2 * pl.nbands);
    v.st = st;
    v.w = pl.w; v.h = pl.h; v.L = pl.L;
    v.nbands = pl.nbands;
    v.band_rows = pl.band_rows;
    v.tile0 = 0;
    hipLaunchKernelGGL(init_item2d, dim3(1), dim3(64), 0, s, item, v, ctr);
    BGT_HIP(hipGetLastError());
    *d_item = item;
    *d_ctr = ctr;
    if (host_item) *host_item = v;
    return BGT_OK;
}

__global__ __launch_bounds__(256) void copy_ints(const int *__restrict__ src, int *__restrict__ dst, int n)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) dst[i] = src[i];
}

int copy_ints_to_host(const int *src, int *dst, int n, hipStream_t s)
{
    if (n <= 0) return BGT_OK;
    hipLaunchKernelGGL(copy_ints, dim3((unsigned)std::min(256, (n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

int host_alloc_mapped(size_t bytes, void **host, void **dev)
{
    *host = *dev = nullptr;
    void *h = nullptr, *d = nullptr;
    if (hipHostMalloc(&h, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        (void)hipGetLastError();
        return fail(BGT_ERR_NOMEM, "pinned allocation of %zu bytes failed", bytes);
    }
    if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(h);
        return fail(BGT_ERR_HIP, "pinned memory has no device address");
    }
    *host = h;
    *dev = d;
    return BGT_OK;
}

int pinned_mapped(PinnedBuffer &b, size_t bytes)
{
    if (b.ptr) return BGT_OK;
    return host_alloc_mapped(bytes, &b.ptr, &b.dptr);
}

// Per-thread pinned words for the convergence poll.
static PinnedBuffer *poll_buffer()
{
    thread_local PinnedBuffer b;  // released when the thread exits
    return pinned_mapped(b, 64) == BGT_OK ? &b : nullptr;
}

int run_passes2d(const Config2D &cfg, const Plan2D &pl, Item2D *d_item, unsigned *d_ctr,
                 const int *d_state, hipStream_t s, int *iterations, int *host_poll)
{
    // host_poll (callers that cannot allocate) takes a runtime copy
    PinnedBuffer *pb = host_poll ? nullptr : poll_buffer();
    int *hp = host_poll ? host_poll : pb ? (int *)pb->ptr : nullptr;
    if (!hp) return fail(BGT_ERR_NOMEM, "pinned poll buffer allocation failed");
    const long max_pass = (long)pl.w * pl.h / cfg.iters_per_pass + 2;  // each changing iteration deletes >= 1 px
    long launched = 0;
    // a single-band image converges in its first launch; a multi-band one
    // takes a few passes of K iterations, so the first poll waits for four
    int chunk = pl.nbands == 1 ? 1 : 4;
    const int kp = image_parts(pl.nbands);
    for (;;) {
        for (int i = 0; i < chunk; ++i)
            BGT_CHECK(launch_image_pass2d(cfg, pl.V, pl.nbands, kp, d_item, d_ctr, s));
        launched += chunk;
        if (pb)
            BGT_CHECK(copy_ints_to_host(d_state, (int *)pb->dptr, ST_INTS, s));
        else
            BGT_HIP(hipMemcpyAsync(hp, d_state, ST_INTS * sizeof(int), hipMemcpyDeviceToHost, s));
        BGT_HIP(hipStreamSynchronize(s));
        if (hp[ST_DONE]) break;
        if (launched > max_pass) return fail(BGT_ERR_NOT_CONVERGED, "2D thinning did not converge");
        chunk = std::min(chunk * 2, 16);
    }
    if (iterations) *iterations = hp[ST_ITERS];
    return BGT_OK;
}

int thin2d_device_u8(const uint8_t *d_in, size_t in_pitch, int w, int h, uint8_t *d_out,
                     size_t out_pitch, hipStream_t s, int *iterations, const Config2D &cfg)
{
    if (!d_in || !d_out) return fail(BGT_ERR_INVALID, "null device pointer");
    if (in_pitch < (size_t)w || out_pitch < (size_t)w)
        return fail(BGT_ERR_INVALID, "pitch smaller than width");
    Plan2D pl;
    BGT_CHECK(plan2d(w, h, cfg, pl));
    Workspace &ws = thread_workspace(0);
    BGT_CHECK(workspace_reserve(ws, workspace_bytes2d(pl)));
    Item2D *d_item;
    unsigned *d_ctr;
    Item2D hv;
    BGT_CHECK(setup_item2d(pl, ws.ptr, s, &d_item, &d_ctr, &hv));
    const long nw = (long)pl.L * h;
    hipLaunchKernelGGL(pack2d_u8, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, d_in,
                       in_pitch, w, h, pl.L, hv.buf[0]);
    BGT_HIP(hipGetLastError());
    BGT_CHECK(run_passes2d(cfg, pl, d_item, d_ctr, hv.st, s, iterations));
    launch_unpack_item2d(d_item, d_out, w, h, out_pitch, s);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(s));
    return BGT_OK;
}

// ---------------------------------------------------------------------------
// device-resident thinning without allocation
// ---------------------------------------------------------------------------

// Engine memory of the allocation-free entry point: one array per GPU, held
// until the process ends.  Host code never uses its name as an address (on
// the host it is a stand-in): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

static int scratch_address(void **out)
{
    static std::mutex mu;
    static void *ptrs[64] = {};
    int dev = 0;
    BGT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(BGT_ERR_UNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(mu);
    if (!ptrs[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(BGT_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        ptrs[dev] = p;
    }
    *out = ptrs[dev];
    return BGT_OK;
}

size_t thin2d_noalloc_max_bytes() { return kScratchBytes; }

int thin2d_device_u8_noalloc(const uint8_t *d_in, size_t in_pitch, int w, int h, uint8_t *d_out,
                             size_t out_pitch, hipStream_t s, int *iterations, const Config2D &cfg)
{
    if (!d_in || !d_out) return fail(BGT_ERR_INVALID, "null device pointer");
    if (in_pitch < (size_t)w || out_pitch < (size_t)w)
        return fail(BGT_ERR_INVALID, "pitch smaller than width");
    Plan2D pl;
    BGT_CHECK(plan2d(w, h, cfg, pl));
    const size_t need = workspace_bytes2d(pl);
    if (need > kScratchBytes)
        return fail(BGT_ERR_UNSUPPORTED,
                    "%d x %d image needs %zu bytes of engine memory; the allocation-free entry "
                    "point holds %zu",
                    w, h, need, kScratchBytes);
    void *ws = nullptr;
    BGT_CHECK(scratch_address(&ws));
    Item2D *d_item;
    unsigned *d_ctr;
    Item2D hv;
    // the init kernel clears the item state, pass counter and band flags that
    // an earlier call left in the array; the first pass writes every row of
    // the other state buffer
    BGT_CHECK(setup_item2d(pl, ws, s, &d_item, &d_ctr, &hv));
    const long nw = (long)pl.L * h;
    hipLaunchKernelGGL(pack2d_u8, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, d_in,
                       in_pitch, w, h, pl.L, hv.buf[0]);
    BGT_HIP(hipGetLastError());
    int poll[ST_INTS];  // pageable: the poll copies synchronise anyway
    BGT_CHECK(run_passes2d(cfg, pl, d_item, d_ctr, hv.st, s, iterations, poll));
    launch_unpack_item2d(d_item, d_out, w, h, out_pitch, s);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(s));
    return BGT_OK;
}

// OutT is T (drop-in Thinning2D) or uint8_t (0/1 mask for the tracer)
template <typename T, typename OutT = T>
static int thin2d_host_t(int w, int h, const T *in, OutT *out, const Config2D &cfg,
                         const OutT **dev_out = nullptr)
{
    Plan2D pl;
    BGT_CHECK(plan2d(w, h, cfg, pl));
    const size_t img_bytes = (size_t)w * h * sizeof(T);
    Workspace &ws = thread_workspace(0);
    Workspace &io = thread_workspace(1);
    BGT_CHECK(workspace_reserve(ws, workspace_bytes2d(pl)));
    BGT_CHECK(workspace_reserve(io, img_bytes));
    hipStream_t s = nullptr;
    BGT_HIP(hipMemcpyAsync(io.ptr, in, img_bytes, hipMemcpyHostToDevice, s));
    Item2D *d_item;
    unsigned *d_ctr;
    Item2D hv;
    BGT_CHECK(setup_item2d(pl, ws.ptr, s, &d_item, &d_ctr, &hv));
    const long nw = (long)pl.L * h;
    const dim3 g((unsigned)((nw + 255) / 256)), b(256);
    if (sizeof(T) == 1 && std::is_same<T, uint8_t>::value)
        hipLaunchKernelGGL(pack2d_u8, g, b, 0, s, (const uint8_t *)io.ptr, (size_t)w, w, h, pl.L,
                           hv.buf[0]);
    else
        hipLaunchKernelGGL(pack2d_t<T>, g, b, 0, s, (const T *)io.ptr, w, h, pl.L, hv.buf[0]);
    BGT_HIP(hipGetLastError());
    BGT_CHECK(run_passes2d(cfg, pl, d_item, d_ctr, hv.st, s, nullptr));
    if (std::is_same<OutT, uint8_t>::value)
        launch_unpack_item2d(d_item, (uint8_t *)io.ptr, w, h, (size_t)w, s);
    else
        hipLaunchKernelGGL(unpack2d_t<OutT>, g, b, 0, s, d_item, (OutT *)io.ptr);
    BGT_HIP(hipGetLastError());
    if (dev_out) {  // the result stays in this thread's I/O workspace
        *dev_out = (const OutT *)io.ptr;
        BGT_HIP(hipStreamSynchronize(s));
        return BGT_OK;
    }
    BGT_HIP(hipMemcpyAsync(out, io.ptr, (size_t)w * h * sizeof(OutT), hipMemcpyDeviceToHost, s));
    BGT_HIP(hipStreamSynchronize(s));
    return BGT_OK;
}

int thin2d_host(int dtype, int w, int h, const void *in, void *out, const Config2D &cfg)
{
    if (!in || !out) return fail(BGT_ERR_INVALID, "null host pointer");
    switch (dtype) {
    case BGT_U8: return thin2d_host_t<uint8_t>(w, h, (const uint8_t *)in, (uint8_t *)out, cfg);
    case BGT_I8: return thin2d_host_t<int8_t>(w, h, (const int8_t *)in, (int8_t *)out, cfg);
    case BGT_U16: return thin2d_host_t<uint16_t>(w, h, (const uint16_t *)in, (uint16_t *)out, cfg);
    case BGT_I16: return thin2d_host_t<int16_t>(w, h, (const int16_t *)in, (int16_t *)out, cfg);
    case BGT_U32: return thin2d_host_t<uint32_t>(w, h, (const uint32_t *)in, (uint32_t *)out, cfg);
    case BGT_I32: return thin2d_host_t<int32_t>(w, h, (const int32_t *)in, (int32_t *)out, cfg);
    case BGT_F32: return thin2d_host_t<float>(w, h, (const float *)in, (float *)out, cfg);
    case BGT_F64: return thin2d_host_t<double>(w, h, (const double *)in, (double *)out, cfg);
    case BGT_BOOL: return thin2d_host_t<bool>(w, h, (const bool *)in, (bool *)out, cfg);
    default: return fail(BGT_ERR_INVALID, "unknown pixel type %d", dtype);
    }
}

int thin2d_host_keep(int dtype, int w, int h, const void *in, const uint8_t **d_skel,
                     const Config2D &cfg)
{
    if (!in || !d_skel) return fail(BGT_ERR_INVALID, "null pointer");
    switch (dtype) {
    case BGT_U8: return thin2d_host_t<uint8_t, uint8_t>(w, h, (const uint8_t *)in, nullptr, cfg, d_skel);
    case BGT_I8: return thin2d_host_t<int8_t, uint8_t>(w, h, (const int8_t *)in, nullptr, cfg, d_skel);
    case BGT_U16: return thin2d_host_t<uint16_t, uint8_t>(w, h, (const uint16_t *)in, nullptr, cfg, d_skel);
    case BGT_I16: return thin2d_host_t<int16_t, uint8_t>(w, h, (const int16_t *)in, nullptr, cfg, d_skel);
    case BGT_U32: return thin2d_host_t<uint32_t, uint8_t>(w, h, (const uint32_t *)in, nullptr, cfg, d_skel);
    case BGT_I32: return thin2d_host_t<int32_t, uint8_t>(w, h, (const int32_t *)in, nullptr, cfg, d_skel);
    case BGT_F32: return thin2d_host_t<float, uint8_t>(w, h, (const float *)in, nullptr, cfg, d_skel);
    case BGT_F64: return thin2d_host_t<double, uint8_t>(w, h, (const double *)in, nullptr, cfg, d_skel);
    case BGT_BOOL: return thin2d_host_t<bool, uint8_t>(w, h, (const bool *)in, nullptr, cfg, d_skel);
    default: return fail(BGT_ERR_INVALID, "unknown pixel type %d", dtype);
    }
}

int thin2d_host_mask(int dtype, int w, int h, const void *in, uint8_t *out, const Config2D &cfg)
{
    if (!in || !out) return fail(BGT_ERR_INVALID, "null host pointer");
    switch (dtype) {
    case BGT_U8: return thin2d_host_t<uint8_t, uint8_t>(w, h, (const uint8_t *)in, out, cfg);
    case BGT_I8: return thin2d_host_t<int8_t, uint8_t>(w, h, (const int8_t *)in, out, cfg);
    case BGT_U16: return thin2d_host_t<uint16_t, uint8_t>(w, h, (const uint16_t *)in, out, cfg);
    case BGT_I16: return thin2d_host_t<int16_t, uint8_t>(w, h, (const int16_t *)in, out, cfg);
    case BGT_U32: return thin2d_host_t<uint32_t, uint8_t>(w, h, (const uint32_t *)in, out, cfg);
    case BGT_I32: return thin2d_host_t<int32_t, uint8_t>(w, h, (const int32_t *)in, out, cfg);
    case BGT_F32: return thin2d_host_t<float, uint8_t>(w, h, (const float *)in, out, cfg);
    case BGT_F64: return thin2d_host_t<double, uint8_t>(w, h, (const double *)in, out, cfg);
    case BGT_BOOL: return thin2d_host_t<bool, uint8_t>(w, h, (const bool *)in, out, cfg);
    default: return fail(BGT_ERR_INVALID, "unknown pixel type %d", dtype);
    }
}

}  // namespace bgt
