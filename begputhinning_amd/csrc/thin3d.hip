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

// one thread per 64 voxels (two state words) of an x-line
__global__ void pack3d_both(const uint8_t *__restrict__ in, int w, long lines, int nh,
                            uint32_t *__restrict__ b0, uint32_t *__restrict__ b1)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nq = nh >> 1;
    if (idx >= lines * nq) return;
    const long line = idx / nq;
    const int x0 = (int)(idx - line * nq) * 64;
    const uint64_t v = x0 < w ? pack_run_u8(in + (size_t)line * w + x0, min(64, w - x0)) : 0ull;
    reinterpret_cast<uint64_t *>(b0)[idx] = v;
    reinterpret_cast<uint64_t *>(b1)[idx] = v;
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

// uint8 output: one thread per 64 voxels of an x-line
__global__ void unpack3d_u8(const uint32_t *__restrict__ b0, int w, long lines, int nh,
                            uint8_t *__restrict__ out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nq = nh >> 1;
    if (idx >= lines * nq) return;
    const long line = idx / nq;
    const int x0 = (int)(idx - line * nq) * 64;
    if (x0 >= w) return;
    unpack_run_u8(reinterpret_cast<const uint64_t *>(b0)[idx], out + (size_t)line * w + x0,
                  min(64, w - x0));
}

__global__ void init_sparse3d(SItem3D *item, SItem3D val, int *ctl)
{
    // ctl: [0] changed, [1] iteration, [2] live, [3] count A, [4] count B
    const int t = threadIdx.x;
    if (t == 0) *item = val;
    if (t < ST_INTS) val.st[t] = 0;
    if (t < 5) ctl[t] = t == 2 ? 1 : 0;
}

int check_config3d(const Config3D &c)
{
    if (c.blocks_per_cu < 0 || c.blocks_per_cu > 32 || c.max_chunk < 2 || c.max_chunk > 64 ||
        (c.max_chunk & 1))
        return fail(BGT_ERR_INVALID, "bad 3D config (blocks_per_cu=%d max_chunk=%d)",
                    c.blocks_per_cu, c.max_chunk);
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

static int sparse_run(SparseVol &sv, const Config3D &cfg, hipStream_t s, int *iterations)
{
    // offsets for the list builder (one item)
    long *off = (long *)((char *)sv.prot + round_up(sv.words * sizeof(uint32_t), 256));
    hipLaunchKernelGGL(set_offsets, dim3(1), dim3(1), 0, s, off, sv.words);
    BGT_CHECK(sparse3d_build_list(sv.d_item, off, 1, sv.words, sv.list[0], s));
    int *hp = poll_buffer3();
    if (!hp) return fail(BGT_ERR_NOMEM, "pinned poll buffer allocation failed");
    const long max_iter = (long)sv.host.w * sv.host.h * sv.host.d + 2;
    long done_iters = 0;
    const int grid = sparse3d_grid(cfg.blocks_per_cu);
    int cur = 0, chunk = 2;
    for (;;) {
        for (int i = 0; i < chunk; ++i) {
            BGT_CHECK(sparse3d_iteration(sv.d_item, 1, sv.list[cur], sv.list[cur ^ 1], sv.capacity,
                                         sv.prot, sv.ctl + 0, sv.ctl + 1, sv.ctl + 2, grid, s));
            cur ^= 1;
        }
        done_iters += chunk;
        BGT_HIP(hipMemcpyAsync(hp, sv.host.st, ST_INTS * sizeof(int), hipMemcpyDeviceToHost, s));
        BGT_HIP(hipStreamSynchronize(s));
        if (hp[ST_DONE]) break;
        if (done_iters > max_iter) return fail(BGT_ERR_NOT_CONVERGED, "3D thinning did not converge");
        chunk = std::min(chunk * 2, cfg.max_chunk);
    }
    if (iterations) *iterations = hp[ST_ITERS];
    return BGT_OK;
}

int thin3d_device_u8(const uint8_t *d_in, int w, int h, int d, uint8_t *d_out, hipStream_t s,
                     int *iterations, const Config3D &cfg)
{
    BGT_CHECK(check_config3d(cfg));
    if (!d_in || !d_out) return fail(BGT_ERR_INVALID, "null device pointer");
    if (w <= 0 || h <= 0 || d <= 0)
        return fail(BGT_ERR_INVALID, "3D size must be >= 1 (got %d x %d x %d)", w, h, d);
    const long words = (long)(2 * ((w + 63) / 64)) * h * d;
    Workspace &ws = thread_workspace(2);
    BGT_CHECK(workspace_reserve(ws, sparse_bytes(words)));
    SparseVol sv;
    BGT_CHECK(sparse_setup(w, h, d, ws.ptr, s, sv));
    const long lines = (long)h * d;
    const dim3 g((unsigned)((words / 2 + 255) / 256)), b(256);
    hipLaunchKernelGGL(pack3d_both, g, b, 0, s, d_in, w, lines, sv.host.nh, sv.host.buf[0],
                       sv.host.buf[1]);
    BGT_HIP(hipGetLastError());
    BGT_CHECK(sparse_run(sv, cfg, s, iterations));
    hipLaunchKernelGGL(unpack3d_u8, g, b, 0, s, sv.host.buf[0], w, lines, sv.host.nh, d_out);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(s));
    return BGT_OK;
}

template <typename T>
static int thin3d_host_t(int w, int h, int d, const T *in, T *out, const Config3D &cfg)
{
    BGT_CHECK(check_config3d(cfg));
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
    BGT_CHECK(sparse_run(sv, cfg, s, nullptr));
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
