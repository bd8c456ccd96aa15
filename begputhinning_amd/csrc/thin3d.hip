// thin3d.hip -- 3D thinning on gfx950: pack/unpack kernels, the fused pass
// kernel and the host drivers (per-volume drop-in and device-resident).
//
// Reference path replaced: 3D/LibThinning3D.cu:34-90 (host driver),
// 3D/CUDA_ThinningKernel3D_Linear.cuh:690-891 (pre-test and driver kernel),
// 3D/CUDA_SubiterationKernels3D.cuh (rules), 3D/CUDA_DeviceHelperFunctions.cuh:
// 146-194 (pack/unpack).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "common.h"
#include "thin3d.h"
#include "thin3d_kernels.h"
#include "sparse3d.h"

namespace bgt {

// 3D rule of Thinning3D<T>: foreground iff v > 0 in T
// (3D/CUDA_DeviceHelperFunctions.cuh:162).
template <typename T>
__device__ __forceinline__ bool fg3d(T v)
{
    return v > (T)0;
}
template <>
__device__ __forceinline__ bool fg3d<bool>(bool v)
{
    return v;
}

// ---------------------------------------------------------------------------
// pack / unpack: one thread per 32-bit word of an x-line
// ---------------------------------------------------------------------------
__global__ void pack3d_u8(const uint8_t *__restrict__ in, int w, int h, int d, int nh,
                          uint32_t *__restrict__ out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)nh * h * d;
    if (idx >= total) return;
    const long line = idx / nh;
    const int j = (int)(idx % nh);
    const uint8_t *row = in + (size_t)line * w;
    const int x0 = j * 32;
    uint32_t v = 0;
    if (x0 + 32 <= w && ((((uintptr_t)(row + x0)) & 15) == 0)) {
        const uint4 *p = reinterpret_cast<const uint4 *>(row + x0);
        uint4 a = p[0], b = p[1];
        auto nz4 = [](uint32_t x) {
            uint32_t t = x | (x >> 4);
            t |= t >> 2;
            t |= t >> 1;
            t &= 0x01010101u;
            return (t | (t >> 7) | (t >> 14) | (t >> 21)) & 0xfu;
        };
        v = nz4(a.x) | nz4(a.y) << 4 | nz4(a.z) << 8 | nz4(a.w) << 12 | nz4(b.x) << 16 |
            nz4(b.y) << 20 | nz4(b.z) << 24 | nz4(b.w) << 28;
    } else {
        for (int k = 0; k < 32; ++k) {
            const int x = x0 + k;
            if (x < w && row[x]) v |= 1u << k;
        }
    }
    out[idx] = v;
}

template <typename T>
__global__ void pack3d_t(const T *__restrict__ in, int w, int h, int d, int nh,
                         uint32_t *__restrict__ out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)nh * h * d;
    if (idx >= total) return;
    const long line = idx / nh;
    const int j = (int)(idx % nh);
    const T *row = in + (size_t)line * w;
    uint32_t v = 0;
    for (int k = 0; k < 32; ++k) {
        const int x = j * 32 + k;
        if (x < w && fg3d<T>(row[x])) v |= 1u << k;
    }
    out[idx] = v;
}

template <typename T>
__global__ void unpack3d(const Item3D *__restrict__ item, T *__restrict__ out)
{
    const Item3D it = *item;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)it.nh * it.h * it.d;
    if (idx >= total) return;
    const long line = idx / it.nh;
    const int j = (int)(idx % it.nh);
    const int x0 = j * 32;
    if (x0 >= it.w) return;
    const uint32_t v = (it.st[ST_PARITY] ? it.buf[1] : it.buf[0])[idx];
    T *row = out + (size_t)line * it.w;
    if (std::is_same<T, uint8_t>::value && x0 + 32 <= it.w &&
        ((((uintptr_t)(row + x0)) & 15) == 0)) {
        auto spread4 = [](uint32_t b) { return ((b & 0xfu) * 0x00204081u) & 0x01010101u; };
        uint4 *p = reinterpret_cast<uint4 *>(row + x0);
        p[0] = make_uint4(spread4(v), spread4(v >> 4), spread4(v >> 8), spread4(v >> 12));
        p[1] = make_uint4(spread4(v >> 16), spread4(v >> 20), spread4(v >> 24), spread4(v >> 28));
    } else {
        for (int k = 0; k < 32 && x0 + k < it.w; ++k) row[x0 + k] = ((v >> k) & 1) ? (T)1 : (T)0;
    }
}

__global__ void init_item3d(Item3D *item, Item3D val, unsigned *ctr)
{
    const int t = threadIdx.x;
    if (t == 0) *item = val;
    if (t < ST_INTS) val.st[t] = 0;
    if (t == 0 && ctr) *ctr = 0;
}

// ---------------------------------------------------------------------------
// per-volume pass kernel
// ---------------------------------------------------------------------------

// Per-volume pass kernel: one workgroup per brick, the last workgroup of the
// launch advances the item.
template <int NH, int R>
__global__ __launch_bounds__(512) void thin3d_pass(const Item3D *__restrict__ items, int K,
                                                    unsigned *__restrict__ ctr)
{
    extern __shared__ uint32_t edges3[];
    __shared__ TileInfo info;
    __shared__ int s_last;

    const int tid = threadIdx.x;
    const Item3D it = items[0];
    const int brick = blockIdx.x;
    const int nbx = it.nbricks / (it.nby * it.nbz);
    const int bx = brick % nbx, by = (brick / nbx) % it.nby, bz = brick / (nbx * it.nby);
    bool work = !it.st[ST_DONE];
    int p = 0, par = 0;
    if (work) {
        p = it.st[ST_PASS];
        par = p & 1;
        if (p > 0) {
            const uint8_t *c = it.chg + (par ^ 1) * it.nbricks;
            bool act = false;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int x = bx + dx, y = by + dy, z = bz + dz;
                        if (x < 0 || y < 0 || z < 0 || x >= nbx || y >= it.nby || z >= it.nbz) continue;
                        act = act || c[(z * it.nby + y) * nbx + x];
                    }
            if (!act) {
                if (tid == 0) it.chg[par * it.nbricks + brick] = 0;
                work = false;
            }
        }
    }
    if (work) {
        brick_work<NH, R>(it, brick, K, par, edges3, &info);
        if (tid == 0) {
            it.chg[par * it.nbricks + brick] = info.mask != 0;
            if (info.mask) atomicOr((unsigned *)&it.st[ST_MASK], info.mask);
        }
    }
    if (tid == 0) {
        __threadfence();
        s_last = atomicAdd(ctr, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last && tid == 0) {
        __threadfence();
        finalize_state(it.st, K, it.st[ST_PASS], nullptr);
        __threadfence();
        atomicExch(ctr, 0u);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

int rows_for_nh(int nh)
{
    switch (nh) {
    case 2: return 8;
    case 4: return 6;
    case 6: return 4;
    default: return 3;
    }
}

int plan3d(int w, int h, int d, const Config3D &cfg, Plan3D &pl)
{
    if (w <= 0 || h <= 0 || d <= 0)
        return fail(BGT_ERR_INVALID, "3D size must be >= 1 (got %d x %d x %d)", w, h, d);
    if (cfg.waves < 1 || cfg.waves > 8 || cfg.iters_per_pass < 1 || cfg.iters_per_pass > 5)
        return fail(BGT_ERR_INVALID, "bad 3D config (waves=%d k=%d)", cfg.waves, cfg.iters_per_pass);
    pl.w = w; pl.h = h; pl.d = d;
    pl.nh = 2 * ((w + 63) / 64);
    const int H = 6 * cfg.iters_per_pass;
    if (pl.nh <= 8) {
        pl.NH = pl.nh;
        pl.nbx = 1;
    } else {
        pl.NH = 8;  // x chunks of 6 interior words with one halo word each side
        pl.nbx = (pl.nh + 5) / 6;
    }
    pl.R = rows_for_nh(pl.NH);
    pl.iy = 64 - 2 * H;
    pl.iz = cfg.waves * pl.R - 2 * H;
    if (pl.iy < H || pl.iz < H)
        return fail(BGT_ERR_INVALID, "3D brick interior (%d x %d) smaller than its halo %d", pl.iy,
                    pl.iz, H);
    pl.nby = (h + pl.iy - 1) / pl.iy;
    pl.nbz = (d + pl.iz - 1) / pl.iz;
    pl.nbricks = pl.nbx * pl.nby * pl.nbz;
    pl.state_words = (size_t)pl.nh * h * d;
    return BGT_OK;
}

size_t workspace_bytes3d(const Plan3D &pl)
{
    size_t b = 0;
    b += round_up(sizeof(Item3D), 256);
    b += round_up(ST_INTS * sizeof(int), 256);
    b += round_up(sizeof(unsigned), 256);
    b += 2 * round_up(pl.state_words * sizeof(uint32_t), 256);
    b += round_up((size_t)2 * pl.nbricks, 256);
    return b;
}

namespace {
struct Carve3 {
    char *p;
    template <typename T>
    T *take(size_t bytes)
    {
        T *r = reinterpret_cast<T *>(p);
        p += round_up(bytes, 256);
        return r;
    }
};
}  // namespace

int setup_item3d(const Plan3D &pl, void *ws, hipStream_t s, Item3D **d_item, unsigned **d_ctr,
                 Item3D *host_item)
{
    Carve3 cv{reinterpret_cast<char *>(ws)};
    Item3D *item = cv.take<Item3D>(sizeof(Item3D));
    int *st = cv.take<int>(ST_INTS * sizeof(int));
    unsigned *ctr = cv.take<unsigned>(sizeof(unsigned));
    Item3D v{};
    v.buf[0] = cv.take<uint32_t>(pl.state_words * sizeof(uint32_t));
    v.buf[1] = cv.take<uint32_t>(pl.state_words * sizeof(uint32_t));
    v.chg = cv.take<uint8_t>((size_t)2 * pl.nbricks);
    v.st = st;
    v.w = pl.w; v.h = pl.h; v.d = pl.d; v.nh = pl.nh;
    v.iy = pl.iy; v.iz = pl.iz;
    v.nby = pl.nby; v.nbz = pl.nbz; v.nbricks = pl.nbricks;
    v.nhc = pl.nbx > 1 ? pl.NH - 2 : pl.nh;
    v.tile0 = 0;
    v.occ0 = 0;
    hipLaunchKernelGGL(init_item3d, dim3(1), dim3(64), 0, s, item, v, ctr);
    BGT_HIP(hipGetLastError());
    *d_item = item;
    *d_ctr = ctr;
    if (host_item) *host_item = v;
    return BGT_OK;
}

size_t lds_bytes3d(const Config3D &cfg, int NH)
{
    // z-edge exchange buffers + per-lane protection masks
    const int R = rows_for_nh(NH);
    return ((size_t)2 * 2 * cfg.waves * NH * 64 + (size_t)(R * NH + 1) * cfg.waves * 64) *
           sizeof(uint32_t);
}

int lds_limit()
{
    static thread_local int max_lds = -1;
    if (max_lds < 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess)
            max_lds = 65536;
    }
    return max_lds;
}

static int launch_volume_pass3d(const Config3D &cfg, const Plan3D &pl, const Item3D *item,
                                unsigned *ctr, hipStream_t s)
{
    const size_t lds = lds_bytes3d(cfg, pl.NH);
    if (lds + 64 > (size_t)lds_limit())
        return fail(BGT_ERR_UNSUPPORTED, "3D pass needs %zu B of LDS, device allows %d", lds, lds_limit());
    const dim3 g(pl.nbricks), b(cfg.waves * 64);
    const int K = cfg.iters_per_pass;
    switch (pl.NH) {
    case 2: hipLaunchKernelGGL((thin3d_pass<2, 8>), g, b, lds, s, item, K, ctr); break;
    case 4: hipLaunchKernelGGL((thin3d_pass<4, 6>), g, b, lds, s, item, K, ctr); break;
    case 6: hipLaunchKernelGGL((thin3d_pass<6, 4>), g, b, lds, s, item, K, ctr); break;
    case 8: hipLaunchKernelGGL((thin3d_pass<8, 3>), g, b, lds, s, item, K, ctr); break;
    default: return fail(BGT_ERR_UNSUPPORTED, "no 3D kernel for %d words per line", pl.NH);
    }
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

static int *poll_buffer3()
{
    thread_local int *p = nullptr;
    if (!p) {
        if (hipHostMalloc((void **)&p, 64, hipHostMallocDefault) != hipSuccess) p = nullptr;
    }
    return p;
}

int run_passes3d(const Config3D &cfg, const Plan3D &pl, Item3D *d_item, unsigned *d_ctr,
                 const int *d_state, hipStream_t s, int *iterations)
{
    int *hp = poll_buffer3();
    if (!hp) return fail(BGT_ERR_NOMEM, "pinned poll buffer allocation failed");
    const long max_pass = (long)pl.w * pl.h * pl.d / cfg.iters_per_pass + 2;
    long launched = 0;
    int chunk = 2;
    for (;;) {
        for (int i = 0; i < chunk; ++i)
            BGT_CHECK(launch_volume_pass3d(cfg, pl, d_item, d_ctr, s));
        launched += chunk;
        BGT_HIP(hipMemcpyAsync(hp, d_state, ST_INTS * sizeof(int), hipMemcpyDeviceToHost, s));
        BGT_HIP(hipStreamSynchronize(s));
        if (hp[ST_DONE]) break;
        if (launched > max_pass) return fail(BGT_ERR_NOT_CONVERGED, "3D thinning did not converge");
        chunk = std::min(chunk * 2, 16);
    }
    if (iterations) *iterations = hp[ST_ITERS];
    return BGT_OK;
}

// ---------------------------------------------------------------------------
// per-volume path on the non-empty-word list engine
// ---------------------------------------------------------------------------

__global__ void pack3d_both(const uint8_t *__restrict__ in, int w, long lines, int nh,
                            uint32_t *__restrict__ b0, uint32_t *__restrict__ b1)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= lines * nh) return;
    const long line = idx / nh;
    const int j = (int)(idx % nh);
    const uint8_t *row = in + (size_t)line * w;
    const int x0 = j * 32;
    uint32_t v = 0;
    for (int k = 0; k < 32 && x0 + k < w; ++k)
        if (row[x0 + k]) v |= 1u << k;
    b0[idx] = v;
    b1[idx] = v;
}

template <typename T>
__global__ void pack3d_both_t(const T *__restrict__ in, int w, long lines, int nh,
                              uint32_t *__restrict__ b0, uint32_t *__restrict__ b1)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= lines * nh) return;
    const long line = idx / nh;
    const int j = (int)(idx % nh);
    const T *row = in + (size_t)line * w;
    uint32_t v = 0;
    for (int k = 0; k < 32 && j * 32 + k < w; ++k)
        if (fg3d<T>(row[j * 32 + k])) v |= 1u << k;
    b0[idx] = v;
    b1[idx] = v;
}

template <typename T>
__global__ void unpack3d_sparse(const uint32_t *__restrict__ b0, int w, long lines, int nh,
                                T *__restrict__ out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= lines * nh) return;
    const long line = idx / nh;
    const int j = (int)(idx % nh);
    const uint32_t v = b0[idx];
    T *row = out + (size_t)line * w;
    for (int k = 0; k < 32 && j * 32 + k < w; ++k) row[j * 32 + k] = ((v >> k) & 1) ? (T)1 : (T)0;
}

__global__ void init_sparse3d(SItem3D *item, SItem3D val, int *ctl)
{
    // ctl: [0] changed, [1] iteration, [2] live, [3] count A, [4] count B
    const int t = threadIdx.x;
    if (t == 0) *item = val;
    if (t < ST_INTS) val.st[t] = 0;
    if (t < 5) ctl[t] = t == 2 ? 1 : 0;
}

struct SparseVol {
    SItem3D *d_item;
    SItem3D host;
    SList list[2];
    uint32_t *prot;
    int *ctl;
    long words;
    int capacity;
};

static size_t sparse_bytes(long words)
{
    return round_up(sizeof(SItem3D), 256) + round_up(ST_INTS * sizeof(int), 256) +
           round_up(8 * sizeof(int), 256) + 2 * round_up(words * sizeof(uint32_t), 256) +
           2 * (round_up(words * sizeof(int), 256) + round_up(words * sizeof(uint32_t), 256)) +
           round_up(words * sizeof(uint32_t), 256) + round_up(sizeof(long) * 2, 256);
}

static int sparse_setup(int w, int h, int d, void *ws, hipStream_t s, SparseVol &sv)
{
    const int nh = 2 * ((w + 63) / 64);
    const long words = (long)nh * h * d;
    if (words > 0x7fffffffL) return fail(BGT_ERR_UNSUPPORTED, "3D volume too large");
    char *p = (char *)ws;
    auto take = [&](size_t bytes) {
        char *r = p;
        p += round_up(bytes, 256);
        return r;
    };
    sv.d_item = (SItem3D *)take(sizeof(SItem3D));
    int *st = (int *)take(ST_INTS * sizeof(int));
    sv.ctl = (int *)take(8 * sizeof(int));
    SItem3D v{};
    v.buf[0] = (uint32_t *)take(words * sizeof(uint32_t));
    v.buf[1] = (uint32_t *)take(words * sizeof(uint32_t));
    v.st = st;
    v.w = w; v.h = h; v.d = d; v.nh = nh;
    for (int k = 0; k < 2; ++k) {
        sv.list[k].item = (int *)take(words * sizeof(int));
        sv.list[k].word = (uint32_t *)take(words * sizeof(uint32_t));
        sv.list[k].count = sv.ctl + 3 + k;
    }
    sv.prot = (uint32_t *)take(words * sizeof(uint32_t));
    long *off = (long *)take(sizeof(long) * 2);
    (void)off;
    sv.host = v;
    sv.words = words;
    sv.capacity = (int)words;
    hipLaunchKernelGGL(init_sparse3d, dim3(1), dim3(64), 0, s, sv.d_item, v, sv.ctl);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

__global__ void set_offsets(long *off, long total)
{
    off[0] = 0;
    off[1] = total;
}

static int sparse_run(SparseVol &sv, hipStream_t s, int *iterations)
{
    // offsets for the list builder (one item)
    long *off = (long *)((char *)sv.prot + round_up(sv.words * sizeof(uint32_t), 256));
    hipLaunchKernelGGL(set_offsets, dim3(1), dim3(1), 0, s, off, sv.words);
    BGT_CHECK(sparse3d_build_list(sv.d_item, off, 1, sv.words, sv.list[0], s));
    int *hp = poll_buffer3();
    if (!hp) return fail(BGT_ERR_NOMEM, "pinned poll buffer allocation failed");
    const long max_iter = (long)sv.host.w * sv.host.h * sv.host.d + 2;
    long done_iters = 0;
    int cur = 0, chunk = 2;
    for (;;) {
        for (int i = 0; i < chunk; ++i) {
            BGT_CHECK(sparse3d_iteration(sv.d_item, 1, sv.list[cur], sv.list[cur ^ 1], sv.capacity,
                                         sv.prot, sv.ctl + 0, sv.ctl + 1, sv.ctl + 2, s));
            cur ^= 1;
        }
        done_iters += chunk;
        BGT_HIP(hipMemcpyAsync(hp, sv.host.st, ST_INTS * sizeof(int), hipMemcpyDeviceToHost, s));
        BGT_HIP(hipStreamSynchronize(s));
        if (hp[ST_DONE]) break;
        if (done_iters > max_iter) return fail(BGT_ERR_NOT_CONVERGED, "3D thinning did not converge");
        chunk = std::min(chunk * 2, 16);
    }
    if (iterations) *iterations = hp[ST_ITERS];
    return BGT_OK;
}

int thin3d_device_u8(const uint8_t *d_in, int w, int h, int d, uint8_t *d_out, hipStream_t s,
                     int *iterations, const Config3D &cfg)
{
    (void)cfg;
    if (!d_in || !d_out) return fail(BGT_ERR_INVALID, "null device pointer");
    if (w <= 0 || h <= 0 || d <= 0)
        return fail(BGT_ERR_INVALID, "3D size must be >= 1 (got %d x %d x %d)", w, h, d);
    const long words = (long)(2 * ((w + 63) / 64)) * h * d;
    Workspace &ws = thread_workspace(2);
    BGT_CHECK(workspace_reserve(ws, sparse_bytes(words)));
    SparseVol sv;
    BGT_CHECK(sparse_setup(w, h, d, ws.ptr, s, sv));
    const long lines = (long)h * d;
    const dim3 g((unsigned)((words + 255) / 256)), b(256);
    hipLaunchKernelGGL(pack3d_both, g, b, 0, s, d_in, w, lines, sv.host.nh, sv.host.buf[0],
                       sv.host.buf[1]);
    BGT_HIP(hipGetLastError());
    BGT_CHECK(sparse_run(sv, s, iterations));
    hipLaunchKernelGGL(unpack3d_sparse<uint8_t>, g, b, 0, s, sv.host.buf[0], w, lines, sv.host.nh,
                       d_out);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(s));
    return BGT_OK;
}

template <typename T>
static int thin3d_host_t(int w, int h, int d, const T *in, T *out, const Config3D &cfg)
{
    (void)cfg;
    if (w <= 0 || h <= 0 || d <= 0)
        return fail(BGT_ERR_INVALID, "3D size must be >= 1 (got %d x %d x %d)", w, h, d);
    const size_t bytes = (size_t)w * h * d * sizeof(T);
    const long words = (long)(2 * ((w + 63) / 64)) * h * d;
    Workspace &ws = thread_workspace(2);
    Workspace &io = thread_workspace(3);
    BGT_CHECK(workspace_reserve(ws, sparse_bytes(words)));
    BGT_CHECK(workspace_reserve(io, bytes));
    hipStream_t s = nullptr;
    BGT_HIP(hipMemcpyAsync(io.ptr, in, bytes, hipMemcpyHostToDevice, s));
    SparseVol sv;
    BGT_CHECK(sparse_setup(w, h, d, ws.ptr, s, sv));
    const long lines = (long)h * d;
    const dim3 g((unsigned)((words + 255) / 256)), b(256);
    hipLaunchKernelGGL(pack3d_both_t<T>, g, b, 0, s, (const T *)io.ptr, w, lines, sv.host.nh,
                       sv.host.buf[0], sv.host.buf[1]);
    BGT_HIP(hipGetLastError());
    BGT_CHECK(sparse_run(sv, s, nullptr));
    hipLaunchKernelGGL(unpack3d_sparse<T>, g, b, 0, s, sv.host.buf[0], w, lines, sv.host.nh,
                       (T *)io.ptr);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipMemcpyAsync(out, io.ptr, bytes, hipMemcpyDeviceToHost, s));
    BGT_HIP(hipStreamSynchronize(s));
    return BGT_OK;
}

int thin3d_host(int dtype, int w, int h, int d, const void *in, void *out, const Config3D &cfg)
{
    if (!in || !out) return fail(BGT_ERR_INVALID, "null host pointer");
    switch (dtype) {
    case BGT_U8: return thin3d_host_t<uint8_t>(w, h, d, (const uint8_t *)in, (uint8_t *)out, cfg);
    case BGT_I8: return thin3d_host_t<int8_t>(w, h, d, (const int8_t *)in, (int8_t *)out, cfg);
    case BGT_U16: return thin3d_host_t<uint16_t>(w, h, d, (const uint16_t *)in, (uint16_t *)out, cfg);
    case BGT_I16: return thin3d_host_t<int16_t>(w, h, d, (const int16_t *)in, (int16_t *)out, cfg);
    case BGT_U32: return thin3d_host_t<uint32_t>(w, h, d, (const uint32_t *)in, (uint32_t *)out, cfg);
    case BGT_I32: return thin3d_host_t<int32_t>(w, h, d, (const int32_t *)in, (int32_t *)out, cfg);
    case BGT_F32: return thin3d_host_t<float>(w, h, d, (const float *)in, (float *)out, cfg);
    case BGT_F64: return thin3d_host_t<double>(w, h, d, (const double *)in, (double *)out, cfg);
    case BGT_BOOL: return thin3d_host_t<bool>(w, h, d, (const bool *)in, (bool *)out, cfg);
    default: return fail(BGT_ERR_INVALID, "unknown pixel type %d", dtype);
    }
}

// 3D CenterlineExtraction conversion: foreground iff int(v) > 0
// (3D/CenterlineExtraction.cpp:44).  Floating values outside int's range
// convert to INT_MIN on x86 and are background here.
template <typename T>
static void thr_int_pos(const T *in, size_t n, uint8_t *m)
{
    for (size_t i = 0; i < n; ++i) {
        const T v = in[i];
        if (std::is_floating_point<T>::value) {
            const double dv = (double)v;
            m[i] = (dv >= 1.0 && dv < 2147483648.0) ? 1 : 0;
        } else {
            m[i] = (int)v > 0;
        }
    }
}

int threshold_int_positive(int dtype, const void *in, size_t n, uint8_t *mask)
{
    switch (dtype) {
    case BGT_U8: thr_int_pos((const uint8_t *)in, n, mask); break;
    case BGT_I8: thr_int_pos((const int8_t *)in, n, mask); break;
    case BGT_U16: thr_int_pos((const uint16_t *)in, n, mask); break;
    case BGT_I16: thr_int_pos((const int16_t *)in, n, mask); break;
    case BGT_U32: thr_int_pos((const uint32_t *)in, n, mask); break;
    case BGT_I32: thr_int_pos((const int32_t *)in, n, mask); break;
    case BGT_F32: thr_int_pos((const float *)in, n, mask); break;
    case BGT_F64: thr_int_pos((const double *)in, n, mask); break;
    case BGT_BOOL: thr_int_pos((const uint8_t *)in, n, mask); break;
    default: return fail(BGT_ERR_INVALID, "unknown pixel type %d", dtype);
    }
    return BGT_OK;
}

}  // namespace bgt
