// sparse3d.hip -- 3D thinning over the list of non-empty words.
//
// Volumes on this path are sparse (a few percent foreground), so the work is
// organised around the words (32 voxels of an x-line) that hold foreground:
// a compacted list of (item, word) entries, kept in raster order inside each
// block of the compaction so neighbouring lanes touch neighbouring words.
// A subiteration is one launch over the list; every entry gathers its 3x3x3
// word neighbourhood from the previous subiteration's buffer (the state is a
// few tens of MB and stays in the caches), evaluates the rule and writes the
// other buffer.  Empty words are zero in both buffers and can never change
// (subiterations only clear bits), so the list only ever shrinks: after each
// iteration, entries whose word became empty or whose item converged are
// compacted away.  No tile, halo or neighbour bookkeeping is needed, and each
// subiteration is synchronous over the whole volume, as in the reference's
// semantics (3D/CUDA_ThinningKernel3D_Linear.cuh:739-782).
//
// An iteration is six launches, U (with the protection pre-test of all six
// directions, 3D/CUDA_ThinningKernel3D_Linear.cuh:690-702), N, E, D, S, W,
// alternating buffers 0 -> 1 -> 0 ..., so it ends in buffer 0, followed by a
// one-workgroup finalize (per-item convergence) and the list compaction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "sparse3d.h"
#include "thin3d_kernels.h"

namespace bgt {

namespace {

__global__ void sparse3d_reset_count(int *count) { *count = 0; }

__device__ __forceinline__ uint32_t fetch(const uint32_t *buf, const SItem3D &it, int y, int z,
                                          int j)
{
    if ((unsigned)y >= (unsigned)it.h || (unsigned)z >= (unsigned)it.d || (unsigned)j >= (unsigned)it.nh)
        return 0u;
    return buf[((size_t)z * it.h + y) * it.nh + j];
}

template <int D>
__global__ __launch_bounds__(256) void sparse3d_step(const SItem3D *__restrict__ items,
                                                     const SList list, uint32_t *__restrict__ prot,
                                                     int *__restrict__ changed)
{
    const int n = *list.count;
    const int stride = gridDim.x * blockDim.x;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const int ii = list.item[k];
        const uint32_t w = list.word[k];
        const SItem3D it = items[ii];
        const uint32_t *src = (D & 1) ? it.buf[1] : it.buf[0];
        uint32_t *dst = (D & 1) ? it.buf[0] : it.buf[1];
        const int j = (int)(w % (uint32_t)it.nh);
        const uint32_t line = w / (uint32_t)it.nh;
        const int y = (int)(line % (uint32_t)it.h), z = (int)(line / (uint32_t)it.h);
        uint32_t nb[27];
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const uint32_t a = fetch(src, it, y + dy - 1, z + dz - 1, j - 1);
                const uint32_t b = fetch(src, it, y + dy - 1, z + dz - 1, j);
                const uint32_t c = fetch(src, it, y + dy - 1, z + dz - 1, j + 1);
                const int q = dz * 9 + dy * 3;
                nb[q + 0] = __builtin_amdgcn_alignbit(b, a, 31);  // x-1
                nb[q + 1] = b;
                nb[q + 2] = __builtin_amdgcn_alignbit(c, b, 1);   // x+1
            }
        const uint32_t v = nb[13];
        uint32_t del;
        if (D == 0) {
            const uint32_t cu = rule3<0>(nb);
            const uint32_t any = cu | rule3<1>(nb) | rule3<2>(nb) | rule3<3>(nb) | rule3<4>(nb) |
                                 rule3<5>(nb);
            prot[k] = v & ~any;
            del = cu & v;
        } else {
            del = rule3<D>(nb) & v & ~prot[k];
        }
        dst[w] = v ^ del;
        if (del) changed[ii] = 1;
    }
}

// Per-item convergence after an iteration: an item whose iteration deleted
// nothing is done; its iteration count includes that last iteration.
__global__ void sparse3d_finalize(const SItem3D *__restrict__ items, int nitems, int *changed,
                                  int *iteration, int *live)
{
    const int itn = *iteration;
    for (int i = threadIdx.x; i < nitems; i += blockDim.x) {
        int *st = items[i].st;
        if (st[ST_DONE]) continue;
        if (!changed[i]) {
            st[ST_DONE] = 1;
            st[ST_ITERS] = itn + 1;
            st[ST_PARITY] = 0;
            atomicSub(live, 1);
        }
        changed[i] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) *iteration = itn + 1;
}

// Keep the entries whose word is non-empty and whose item is not done.  Each
// block keeps the order of its 256 entries; blocks take output ranges with one
// atomic each.
__global__ __launch_bounds__(256) void sparse3d_compact(const SItem3D *__restrict__ items,
                                                        const SList in, SList out)
{
    __shared__ int s_scan[256];
    __shared__ int s_base;
    const int n = *in.count;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return;
    int keep = 0, ii = 0;
    uint32_t w = 0;
    if (k < n) {
        ii = in.item[k];
        w = in.word[k];
        const SItem3D it = items[ii];
        const bool live = !it.st[ST_DONE];
        keep = live && it.buf[0][w] != 0u;
        // A word emptied by the last subiteration (W) still holds its bits from
        // before that step in buffer 1; once off the list nothing rewrites it,
        // and the next iteration's odd steps read buffer 1.
        if (live && !keep) it.buf[1][w] = 0u;
    }
    s_scan[threadIdx.x] = keep;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = threadIdx.x >= off ? s_scan[threadIdx.x - off] : 0;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    if (threadIdx.x == 255) s_base = s_scan[255] ? atomicAdd(out.count, s_scan[255]) : 0;
    __syncthreads();
    if (keep) {
        const int pos = s_base + s_scan[threadIdx.x] - 1;
        out.item[pos] = ii;
        out.word[pos] = w;
    }
}

// Initial list: every non-empty word of every item (grid over all words).
__global__ __launch_bounds__(256) void sparse3d_build(const SItem3D *__restrict__ items,
                                                      const long *__restrict__ off, int nitems,
                                                      long total, SList out)
{
    __shared__ int s_scan[256];
    __shared__ int s_base;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int keep = 0, ii = 0;
    uint32_t w = 0;
    if (g < total) {
        int lo = 0, hi = nitems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= g) lo = mid;
            else hi = mid - 1;
        }
        ii = lo;
        w = (uint32_t)(g - off[ii]);
        keep = items[ii].buf[0][w] != 0u;
    }
    s_scan[threadIdx.x] = keep;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int v = threadIdx.x >= o ? s_scan[threadIdx.x - o] : 0;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    if (threadIdx.x == 255) s_base = s_scan[255] ? atomicAdd(out.count, s_scan[255]) : 0;
    __syncthreads();
    if (keep) {
        const int pos = s_base + s_scan[threadIdx.x] - 1;
        out.item[pos] = ii;
        out.word[pos] = w;
    }
}

}  // namespace

int sparse3d_grid(int blocks_per_cu)
{
    static int cus = 0, occ = 0;
    if (!cus) {
        int dev = 0, c = 256, per_cu = 4;
        if (hipGetDevice(&dev) == hipSuccess) {
            (void)hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)sparse3d_step<0>,
                                                             256, 0) != hipSuccess || per_cu < 1)
                per_cu = 4;
        }
        occ = std::max(1, per_cu);
        cus = std::max(1, c);
    }
    return cus * (blocks_per_cu > 0 ? blocks_per_cu : occ);
}

int sparse3d_build_list(const SItem3D *items, const long *off, int nitems, long total_words,
                        SList list, hipStream_t s)
{
    if (total_words <= 0) return BGT_OK;
    hipLaunchKernelGGL(sparse3d_build, dim3((unsigned)((total_words + 255) / 256)), dim3(256), 0, s,
                       items, off, nitems, total_words, list);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

int sparse3d_iteration(const SItem3D *items, int nitems, SList cur, SList next, int max_entries,
                       uint32_t *prot, int *changed, int *iteration, int *live, int grid,
                       hipStream_t s)
{
    const dim3 g(grid), b(256);
    hipLaunchKernelGGL(sparse3d_step<0>, g, b, 0, s, items, cur, prot, changed);
    hipLaunchKernelGGL(sparse3d_step<1>, g, b, 0, s, items, cur, prot, changed);
    hipLaunchKernelGGL(sparse3d_step<2>, g, b, 0, s, items, cur, prot, changed);
    hipLaunchKernelGGL(sparse3d_step<3>, g, b, 0, s, items, cur, prot, changed);
    hipLaunchKernelGGL(sparse3d_step<4>, g, b, 0, s, items, cur, prot, changed);
    hipLaunchKernelGGL(sparse3d_step<5>, g, b, 0, s, items, cur, prot, changed);
    hipLaunchKernelGGL(sparse3d_finalize, dim3(1), dim3(256), 0, s, items, nitems, changed,
                       iteration, live);
    hipLaunchKernelGGL(sparse3d_reset_count, dim3(1), dim3(1), 0, s, next.count);
    hipLaunchKernelGGL(sparse3d_compact, dim3((unsigned)((max_entries + 255) / 256)), b, 0, s,
                       items, cur, next);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

}  // namespace bgt
