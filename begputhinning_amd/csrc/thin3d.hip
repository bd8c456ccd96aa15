// thin3d.hip -- 3D thinning on gfx950, per-volume drivers: pack/unpack
// kernels and the host loop over the non-empty-word list engine (sparse3d.hip)
// for the drop-in Thinning3D and the device-resident entry point.
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
#include "pack.h"
#include "sparse3d.h"
#include "tile3d.h"

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
// per-volume path on the non-empty-word list engine
// ---------------------------------------------------------------------------

// Thinning3D<T> input -> uint8 0/1 (v > 0 in T), and back (0/1 in T).
// 3D CenterlineExtraction's foreground: int(v) > 0 (3D/CenterlineExtraction.cpp:44).
// Floating values outside int's range convert to INT_MIN on x86 and are
// background here.
template <typename T>
__device__ __forceinline__ bool int_pos(T v)
{
    if (std::is_floating_point<T>::value) {
        const double dv = (double)v;
        return dv >= 1.0 && dv < 2147483648.0;
    }
    return (int)v > 0;
}
template <typename T, bool IntPos>
__global__ void mask3d_from(const T *__restrict__ in, uint8_t *__restrict__ m, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) m[t] = (IntPos ? int_pos<T>(in[t]) : fg3d<T>(in[t])) ? 1 : 0;
}

template <typename T>
__global__ void mask3d_to(const uint8_t *__restrict__ m, T *__restrict__ out, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = m[t] ? (T)1 : (T)0;
}

__global__ void init_one3d(SItem3D *item, int2 *geo, SIO *io, SItem3D v, SIO f)
{
    if (threadIdx.x == 0) {
        *item = v;
        *geo = int2{v.P, v.plane};
        *io = f;
    }
}

int check_config3d(const Config3D &c)
{
    if (c.blocks_per_cu < 0 || c.blocks_per_cu > 32 || c.max_chunk < 1 || c.max_chunk > 64 ||
        c.tail_words < 0 || c.engine < 0 || c.engine > 2)
        return fail(BGT_ERR_INVALID,
                    "bad 3D config (blocks_per_cu=%d max_chunk=%d tail_words=%d engine=%d)",
                    c.blocks_per_cu, c.max_chunk, c.tail_words, c.engine);
    return BGT_OK;
}

// Per-thread pinned words for the convergence poll (mapped: a kernel writes
// them).
static PinnedBuffer *poll_buffer3()
{
    thread_local PinnedBuffer b;  // released when the thread exits
    return pinned_mapped(b, 256) == BGT_OK ? &b : nullptr;
}

// Engine for one volume in a workspace: [item|geo|io] [states+ctl] [changed]
// [state 0] [state 1] [list 0 item|word|block] [list 1 ...] [active ...]
// [prot] [first] [block stamps].
struct OneVolume {
    Sparse3D e;
    SItem3D item;
    Tile3D t;            // the tile engine that takes over (engine 1)
    bool tiles = false;
};

__global__ void init_one3d_titem(TileItem3D *titem, TileItem3D tv)
{
    if (threadIdx.x == 0) *titem = tv;
}

static size_t one_volume_bytes(long nwords, long data, long nblocks)
{
    return 3 * 256 + round_up((ST_INTS + SC_INTS) * sizeof(int), 256) + 256 +
           2 * round_up(nwords * sizeof(uint32_t), 256) +
           11 * round_up(data * sizeof(uint32_t), 256) + 2 * round_up(nblocks * sizeof(int), 256);
}

static int one_volume_setup(int w, int h, int d, const uint8_t *in, uint8_t *out, Workspace &ws,
                            hipStream_t s, OneVolume &ov, bool with_tiles)
{
    Layout3D lay;
    BGT_CHECK(sparse3d_layout(&w, &h, &d, 1, lay));
    const long nwords = lay.nwords, data = lay.data_words;
    std::vector<TileItem3D> ti;
    const long ntiles = with_tiles ? tile3d_items(lay, ti) : 0;
    uint32_t *no_state[2] = {nullptr, nullptr};  // sizing only: the word engine's arrays are shared
    Tile3D sizing;
    const size_t tile_bytes =
        with_tiles ? 256 + tile3d_carve(sizing, nullptr, 1, nwords, ntiles, lay.nblocks, no_state) : 0;
    BGT_CHECK(workspace_reserve(ws, one_volume_bytes(nwords, data, lay.nblocks) + tile_bytes));
    char *p = (char *)ws.ptr;
    auto take = [&](size_t bytes) {
        char *r = p;
        p += round_up(bytes ? bytes : 1, 256);
        return r;
    };
    Sparse3D &e = ov.e;
    e = Sparse3D{};
    e.nitems = 1;
    SItem3D *d_item = (SItem3D *)take(sizeof(SItem3D));
    int2 *d_geo = (int2 *)take(sizeof(int2));
    SIO *d_io = (SIO *)take(sizeof(SIO));
    e.items = d_item;
    e.geo = d_geo;
    e.io = d_io;
    e.blk = nullptr;  // every block belongs to item 0
    e.nblk = (int)lay.blk.size();
    e.ublk = nullptr;
    e.nublk = (int)lay.ublk.size();
    e.states = (int *)take((ST_INTS + SC_INTS) * sizeof(int));
    e.ctl = e.states + ST_INTS;  // polled together with the item state
    e.changed = (int *)take(sizeof(int));
    e.state[0] = (uint32_t *)take(nwords * sizeof(uint32_t));
    e.state[1] = (uint32_t *)take(nwords * sizeof(uint32_t));
    e.nwords = nwords;
    for (int k = 0; k < 2; ++k) {
        e.list[k].item = (int *)take(data * sizeof(int));
        e.list[k].word = (uint32_t *)take(data * sizeof(uint32_t));
        e.list[k].blk = (int *)take(data * sizeof(int));
        e.list[k].count = e.ctl + SC_COUNT0 + k;
    }
    e.act.item = (int *)take(data * sizeof(int));
    e.act.word = (uint32_t *)take(data * sizeof(uint32_t));
    e.act.blk = (int *)take(data * sizeof(int));
    e.act.count = nullptr;  // in ctl, by iteration parity
    e.prot = (uint32_t *)take(data * sizeof(uint32_t));
    e.first = (uint32_t *)take(data * sizeof(uint32_t));
    e.capacity = (int)data;
    e.nblocks = (int)lay.nblocks;
    e.bstamp = (int *)take(lay.nblocks * sizeof(int));
    e.dstamp = (int *)take(lay.nblocks * sizeof(int));
    ov.item = lay.items[0];
    hipLaunchKernelGGL(init_one3d, dim3(1), dim3(64), 0, s, d_item, d_geo, d_io, lay.items[0],
                       SIO{in, out, 0, 0});
    ov.tiles = with_tiles;
    if (with_tiles) {
        TileItem3D *d_titem = (TileItem3D *)take(sizeof(TileItem3D));
        Tile3D &t = ov.t;
        t = Tile3D{};
        tile3d_carve(t, p, 1, nwords, ntiles, lay.nblocks, e.state);
        t.items = d_titem;
        t.sitems = d_item;
        t.io = d_io;
        t.blk = nullptr;
        t.nblk = e.nblk;
        t.ublk = nullptr;
        t.nublk = e.nublk;
        hipLaunchKernelGGL(init_one3d_titem, dim3(1), dim3(64), 0, s, d_titem, ti[0]);
    }
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

// The tile engine's iterations from t on (it has taken over from the word
// engine), polling item 0 after chunks of 2, 4, ... up to max_chunk.  The
// launch of iteration t closes the item when iteration t - 1 changed nothing.
static int tiles_to_convergence(const Tile3D &e, int t, long max_iter, const Config3D &cfg,
                                hipStream_t s, int *hp, int *hd)
{
    const int grid = tile3d_grid(cfg.blocks_per_cu);
    int chunk = std::min(2, cfg.max_chunk);
    for (;;) {
        for (int i = 0; i < chunk; ++i) BGT_CHECK(tile3d_iteration(e, t++, grid, s));
        BGT_CHECK(copy_ints_to_host(e.states, hd, ST_INTS, s));
        BGT_HIP(hipStreamSynchronize(s));
        if (hp[ST_DONE]) return BGT_OK;
        if (t > max_iter) return fail(BGT_ERR_NOT_CONVERGED, "3D thinning did not converge");
        chunk = std::min(chunk * 2, cfg.max_chunk);
    }
}

// Iterate one volume to convergence, polling the item state every chunk of
// iterations (2, 4, ... up to max_chunk).
static int one_volume_run(OneVolume &ov, const Config3D &cfg, hipStream_t s, int *iterations)
{
    const Sparse3D &e = ov.e;
    BGT_CHECK(sparse3d_begin(e, true, s));
    PinnedBuffer *pb = poll_buffer3();
    if (!pb) return fail(BGT_ERR_NOMEM, "pinned poll buffer allocation failed");
    int *hp = (int *)pb->ptr, *hd = (int *)pb->dptr;
    const long max_iter = (long)ov.item.w * ov.item.h * ov.item.d + 2;
    const int grid = sparse3d_grid(cfg.blocks_per_cu);
    long done_iters = 0;
    int cur = 0, chunk = std::min(2, cfg.max_chunk), bound = e.capacity, abound = e.capacity;
    for (;;) {
        for (int i = 0; i < chunk; ++i) {
            BGT_CHECK(sparse3d_iteration(e, cur, bound, abound, grid, s));
            cur ^= 1;
        }
        done_iters += chunk;
        BGT_CHECK(copy_ints_to_host(e.states, hd, ST_INTS + SC_INTS, s));
        BGT_HIP(hipStreamSynchronize(s));
        if (hp[ST_DONE]) break;
        if (done_iters > max_iter) return fail(BGT_ERR_NOT_CONVERGED, "3D thinning did not converge");
        bound = hp[ST_INTS + SC_COUNT0 + cur];
        abound = hp[ST_INTS + SC_ACOUNT0 + (hp[ST_INTS + SC_ITER] & 1)];
        if (ov.tiles && cfg.tail_words > 0 && abound <= cfg.tail_words && !(done_iters & 1)) {
            // few active words left: the tile engine, one launch per iteration
            BGT_CHECK(tile3d_from_words(ov.t, e, (int)done_iters, bound, s));
            BGT_CHECK(tiles_to_convergence(ov.t, (int)done_iters, max_iter, cfg, s, hp, hd));
            break;
        }
        if (!ov.tiles && cfg.tail_words > 0 && abound <= cfg.tail_words) {
            // few active words left: the remaining iterations in one launch
            BGT_CHECK(sparse3d_tail(e, s));
            BGT_CHECK(copy_ints_to_host(e.states, hd, ST_INTS + SC_INTS, s));
            BGT_HIP(hipStreamSynchronize(s));
            if (!hp[ST_DONE]) return fail(BGT_ERR_NOT_CONVERGED, "3D tail did not converge");
            break;
        }
        chunk = std::min(chunk * 2, cfg.max_chunk);
    }
    if (iterations) *iterations = hp[ST_ITERS];
    return sparse3d_end(e, s);
}

// ---------------------------------------------------------------------------
// per-volume path on the tile engine
// ---------------------------------------------------------------------------

// One volume in a workspace: [sitem|titem|io] [engine: states, tiles, lists,
// state arrays].
struct OneVolumeTiles {
    Tile3D e;
    SItem3D item;
};

__global__ void init_one3d_tiles(SItem3D *item, TileItem3D *titem, SIO *io, SItem3D v, TileItem3D tv,
                                 SIO f)
{
    if (threadIdx.x == 0) {
        *item = v;
        *titem = tv;
        *io = f;
    }
}

static int one_volume_tiles_setup(int w, int h, int d, const uint8_t *in, uint8_t *out,
                                  Workspace &ws, hipStream_t s, OneVolumeTiles &ov)
{
    Layout3D lay;
    BGT_CHECK(sparse3d_layout(&w, &h, &d, 1, lay));
    std::vector<TileItem3D> ti;
    const long ntiles = tile3d_items(lay, ti);
    Tile3D &e = ov.e;
    e = Tile3D{};
    const size_t head = 3 * 256;
    const size_t need = head + tile3d_carve(e, nullptr, 1, lay.nwords, ntiles, lay.nblocks);
    BGT_CHECK(workspace_reserve(ws, need));
    char *p = (char *)ws.ptr;
    SItem3D *d_item = (SItem3D *)p;
    TileItem3D *d_titem = (TileItem3D *)(p + 256);
    SIO *d_io = (SIO *)(p + 512);
    tile3d_carve(e, p + head, 1, lay.nwords, ntiles, lay.nblocks);
    e.items = d_titem;
    e.sitems = d_item;
    e.io = d_io;
    e.blk = nullptr;  // every block belongs to item 0
    e.nblk = (int)lay.blk.size();
    e.ublk = nullptr;
    e.nublk = (int)lay.ublk.size();
    ov.item = lay.items[0];
    hipLaunchKernelGGL(init_one3d_tiles, dim3(1), dim3(64), 0, s, d_item, d_titem, d_io, lay.items[0],
                       ti[0], SIO{in, out, 0, 0});
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

// Iterate one volume to convergence, polling the item state after chunks of
// 2, 4, ... up to max_chunk iterations.  The launch of iteration t closes the
// item when iteration t - 1 changed nothing, so a poll after iterations
// 0 .. t - 1 sees every convergence up to iteration t - 2.
static int one_volume_tiles_run(OneVolumeTiles &ov, const Config3D &cfg, hipStream_t s,
                                int *iterations)
{
    const Tile3D &e = ov.e;
    BGT_CHECK(tile3d_begin(e, true, s));
    PinnedBuffer *pb = poll_buffer3();
    if (!pb) return fail(BGT_ERR_NOMEM, "pinned poll buffer allocation failed");
    int *hp = (int *)pb->ptr, *hd = (int *)pb->dptr;
    const long max_iter = (long)ov.item.w * ov.item.h * ov.item.d + 3;
    const int grid = tile3d_grid(cfg.blocks_per_cu);
    int t = 0, chunk = std::min(2, cfg.max_chunk);
    for (;;) {
        for (int i = 0; i < chunk; ++i) BGT_CHECK(tile3d_iteration(e, t++, grid, s));
        BGT_CHECK(copy_ints_to_host(e.states, hd, ST_INTS, s));
        BGT_HIP(hipStreamSynchronize(s));
        if (hp[ST_DONE]) break;
        if (t > max_iter) return fail(BGT_ERR_NOT_CONVERGED, "3D thinning did not converge");
        chunk = std::min(chunk * 2, cfg.max_chunk);
    }
    if (iterations) *iterations = hp[ST_ITERS];
    return tile3d_end(e, s);
}

static int thin3d_one(const uint8_t *in, int w, int h, int d, uint8_t *out, hipStream_t s,
                      int *iterations, const Config3D &cfg)
{
    if (cfg.engine == 2) {
        OneVolumeTiles ov;
        BGT_CHECK(one_volume_tiles_setup(w, h, d, in, out, thread_workspace(2), s, ov));
        return one_volume_tiles_run(ov, cfg, s, iterations);
    }
    OneVolume ov;
    BGT_CHECK(one_volume_setup(w, h, d, in, out, thread_workspace(2), s, ov, cfg.engine == 1));
    return one_volume_run(ov, cfg, s, iterations);
}

int thin3d_device_u8(const uint8_t *d_in, int w, int h, int d, uint8_t *d_out, hipStream_t s,
                     int *iterations, const Config3D &cfg)
{
    BGT_CHECK(check_config3d(cfg));
    if (!d_in || !d_out) return fail(BGT_ERR_INVALID, "null device pointer");
    if (w <= 0 || h <= 0 || d <= 0)
        return fail(BGT_ERR_INVALID, "3D size must be >= 1 (got %d x %d x %d)", w, h, d);
    BGT_CHECK(thin3d_one(d_in, w, h, d, d_out, s, iterations, cfg));
    BGT_HIP(hipStreamSynchronize(s));
    return BGT_OK;
}

template <typename T, bool IntPos = false>
static int thin3d_host_t(int w, int h, int d, const T *in, T *out, const Config3D &cfg,
                         const uint8_t **dev_mask = nullptr)
{
    BGT_CHECK(check_config3d(cfg));
    if (w <= 0 || h <= 0 || d <= 0)
        return fail(BGT_ERR_INVALID, "3D size must be >= 1 (got %d x %d x %d)", w, h, d);
    const long n = (long)w * h * d;
    const size_t bytes = (size_t)n * sizeof(T);
    // uint8 and bool inputs are their own mask (nonzero == v > 0)
    const bool direct = std::is_same<T, uint8_t>::value || std::is_same<T, bool>::value;
    Workspace &io = thread_workspace(3);
    BGT_CHECK(workspace_reserve(io, round_up(bytes, 256) + (direct ? 0 : (size_t)n)));
    hipStream_t s = nullptr;
    BGT_HIP(hipMemcpyAsync(io.ptr, in, bytes, hipMemcpyHostToDevice, s));
    uint8_t *mask = direct ? (uint8_t *)io.ptr : (uint8_t *)io.ptr + round_up(bytes, 256);
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    if (!direct) hipLaunchKernelGGL((mask3d_from<T, IntPos>), g, b, 0, s, (const T *)io.ptr, mask, n);
    BGT_HIP(hipGetLastError());
    BGT_CHECK(thin3d_one(mask, w, h, d, mask, s, nullptr, cfg));
    if (dev_mask) {  // the skeleton stays in this thread's I/O workspace
        *dev_mask = mask;
        BGT_HIP(hipStreamSynchronize(s));
        return BGT_OK;
    }
    if (!direct) hipLaunchKernelGGL(mask3d_to<T>, g, b, 0, s, mask, (T *)io.ptr, n);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipMemcpyAsync(out, io.ptr, bytes, hipMemcpyDeviceToHost, s));
    BGT_HIP(hipStreamSynchronize(s));
    return BGT_OK;
}

int thin3d_keep_int_positive(int dtype, int w, int h, int d, const void *in,
                             const uint8_t **d_skel, const Config3D &cfg)
{
    if (!in || !d_skel) return fail(BGT_ERR_INVALID, "null pointer");
    switch (dtype) {
    case BGT_U8: return thin3d_host_t<uint8_t, true>(w, h, d, (const uint8_t *)in, nullptr, cfg, d_skel);
    case BGT_I8: return thin3d_host_t<int8_t, true>(w, h, d, (const int8_t *)in, nullptr, cfg, d_skel);
    case BGT_U16: return thin3d_host_t<uint16_t, true>(w, h, d, (const uint16_t *)in, nullptr, cfg, d_skel);
    case BGT_I16: return thin3d_host_t<int16_t, true>(w, h, d, (const int16_t *)in, nullptr, cfg, d_skel);
    case BGT_U32: return thin3d_host_t<uint32_t, true>(w, h, d, (const uint32_t *)in, nullptr, cfg, d_skel);
    case BGT_I32: return thin3d_host_t<int32_t, true>(w, h, d, (const int32_t *)in, nullptr, cfg, d_skel);
    case BGT_F32: return thin3d_host_t<float, true>(w, h, d, (const float *)in, nullptr, cfg, d_skel);
    case BGT_F64: return thin3d_host_t<double, true>(w, h, d, (const double *)in, nullptr, cfg, d_skel);
    case BGT_BOOL: return thin3d_host_t<uint8_t, true>(w, h, d, (const uint8_t *)in, nullptr, cfg, d_skel);
    default: return fail(BGT_ERR_INVALID, "unknown pixel type %d", dtype);
    }
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

}  // namespace bgt
