// seggraph.hip -- device pre-pass of the segment tracer: classification and
// stream compaction of skeleton points, so that the host traces only them.
//
// Reference: the tracer's first pass, 2D/ExtractSegments.cpp:32-50 and
// 3D/ExtractSegments.cpp:154-171, scans the whole padded frame for skeleton
// pixels off its border, in raster order, and codes each by its number of
// skeleton neighbours (8 or 26, border pixels included).  Its second pass
// follows segments through those codes.  Here the first pass runs on the
// device over the skeleton as it leaves the thinning kernels, and yields a
// compact graph in raster order: each point's raster index in the padded
// frame, its code, and the ranks of its neighbours that are points themselves
// (in the reference's neighbour order), so that the host's second pass needs
// no frame-sized array.
//
// The skeleton is the unpadded w x h x d uint8 buffer; the padded frame
// (pw x ph x pd, zero outside the skeleton) is virtual: it only decides which
// pixels lie off the border and the raster index.
//
// Kernels, one thread per 64-pixel chunk of a row (chunks in raster order):
//   seg_count  -- the chunk's point bits, their count, and each workgroup's sum;
//   seg_scan   -- one workgroup: exclusive sums of the workgroup sums, total;
//   seg_base   -- each chunk's first rank;
//   seg_emit   -- per point: packed coordinates, code, neighbour ranks.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "common.h"
#include "seggraph.h"

namespace bgt {

namespace {

struct Frame {
    const uint8_t *skel;  // w x h x d, rows of w bytes
    int w, h, d;          // skeleton
    int pw, ph, pd;       // padded frame (>= w, h, d)
    int cpr;              // chunks per row: ceil(w / 64)
    int three_d;
};

__device__ __forceinline__ bool off_border(const Frame &f, int x, int y, int z)
{
    return x >= 1 && x <= f.pw - 2 && y >= 1 && y <= f.ph - 2 &&
           (!f.three_d || (z >= 1 && z <= f.pd - 2));
}

// skeleton byte at (x, y, z), zero outside the skeleton's extent
__device__ __forceinline__ bool set_at(const Frame &f, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= f.w || y >= f.h || z >= f.d) return false;
    return f.skel[((size_t)z * f.h + y) * f.w + x] != 0;
}

__global__ __launch_bounds__(256) void seg_count(Frame f, long nchunks, uint64_t *bits,
                                                 int *cnt, int *wsum)
{
    __shared__ int s_sum[4];
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    int n = 0;
    if (c < nchunks) {
        const long row = c / f.cpr;
        const int x0 = (int)(c - row * f.cpr) * 64;
        const int y = (int)(row % f.h), z = (int)(row / f.h);
        const uint8_t *p = f.skel + (size_t)row * f.w + x0;
        const int nx = min(64, f.w - x0);
        uint64_t m = 0;
        const bool row_ok = y >= 1 && y <= f.ph - 2 && (!f.three_d || (z >= 1 && z <= f.pd - 2));
        if (row_ok)
            for (int k = 0; k < nx; ++k) {
                const int x = x0 + k;
                if (p[k] && x >= 1 && x <= f.pw - 2) m |= 1ull << k;
            }
        bits[c] = m;
        n = __popcll(m);
        cnt[c] = n;
    }
    // workgroup sum
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) wsum[blockIdx.x] = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
}

// one workgroup of 1024 threads: exclusive sums of wsum[0..nw) in place, and
// the total in *total
__global__ __launch_bounds__(1024) void seg_scan(int *wsum, int nw, long *total)
{
    __shared__ long s_part[1024];
    __shared__ long s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int b = 0; b < nw; b += 1024) {
        const int i = b + (int)threadIdx.x;
        const long v = i < nw ? wsum[i] : 0;
        s_part[threadIdx.x] = v;
        __syncthreads();
        // inclusive Hillis-Steele scan over the 1024 values
        for (int o = 1; o < 1024; o <<= 1) {
            const long a = threadIdx.x >= (unsigned)o ? s_part[threadIdx.x - o] : 0;
            __syncthreads();
            s_part[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nw) wsum[i] = (int)(s_carry + s_part[threadIdx.x] - v);
        __syncthreads();
        if (threadIdx.x == 1023) s_carry += s_part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s_carry;
}

// first rank of every chunk: its workgroup's base plus the counts before it
__global__ __launch_bounds__(256) void seg_base(long nchunks, const int *cnt, const int *wbase,
                                                int *base)
{
    __shared__ int s_w[4];
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    const int v = c < nchunks ? cnt[c] : 0;
    // inclusive scan within the wave, then across the four waves
    int s = v;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(s, o);
        if (lane >= o) s += t;
    }
    if (lane == 63) s_w[wv] = s;
    __syncthreads();
    int before = 0;
    for (int k = 0; k < wv; ++k) before += s_w[k];
    if (c < nchunks) base[c] = wbase[blockIdx.x] + before + s - v;
}

__device__ __forceinline__ int rank_of(const Frame &f, const uint64_t *bits, const int *base,
                                       int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= f.w || y >= f.h || z >= f.d) return -1;
    const long c = ((long)z * f.h + y) * f.cpr + (x >> 6);
    const uint64_t m = bits[c];
    const int k = x & 63;
    if (!((m >> k) & 1ull)) return -1;
    return base[c] + __popcll(m & ((1ull << k) - 1ull));
}

// one thread per chunk: every point of the chunk
__global__ __launch_bounds__(256) void seg_emit(Frame f, long nchunks, const uint64_t *bits,
                                                const int *base, int64_t *idx, int8_t *code,
                                                int *nbr)
{
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= nchunks) return;
    uint64_t m = bits[c];
    if (!m) return;
    const long row = c / f.cpr;
    const int x0 = (int)(c - row * f.cpr) * 64;
    const int y = (int)(row % f.h), z = (int)(row / f.h);
    const int nd = f.three_d ? 26 : 8;
    int r = base[c];
    while (m) {
        const int k = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int x = x0 + k;
        idx[r] = (int64_t)x | (int64_t)y << 16 | (int64_t)z << 32;  // packed coordinates
        int nn = 0, j = 0;
        // the reference's neighbour order: raster order of the offsets
        for (int dz = f.three_d ? -1 : 0; dz <= (f.three_d ? 1 : 0); ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dz && !dy && !dx) continue;
                    const int qx = x + dx, qy = y + dy, qz = z + dz;
                    nn += set_at(f, qx, qy, qz);
                    nbr[(size_t)r * nd + j] =
                        off_border(f, qx, qy, qz) ? rank_of(f, bits, base, qx, qy, qz) : -1;
                    ++j;
                }
        if (nn == 0) nn = 1;
        code[r] = (int8_t)(nn == 2 ? 2 : -nn);
        ++r;
    }
}

// Grow-only pinned host staging of one thread's pre-pass results (device to
// host copies into pageable memory go through the runtime's own staging),
// released at thread exit.
struct PinnedStage {
    char *ptr = nullptr;
    size_t bytes = 0;
    ~PinnedStage()
    {
        if (ptr && !on_main_thread()) (void)hipHostFree(ptr);
    }
    int reserve(size_t n)
    {
        if (ptr && bytes >= n) return BGT_OK;
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
        const size_t want = n + n / 4 + 4096;
        if (hipHostMalloc((void **)&ptr, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            ptr = nullptr;
            return fail(BGT_ERR_NOMEM, "pinned staging of %zu bytes failed", want);
        }
        bytes = want;
        return BGT_OK;
    }
};

PinnedStage &thread_stage()
{
    thread_local PinnedStage st;
    return st;
}

}  // namespace

int segment_graph_device(const uint8_t *d_skel, int w, int h, int d, int pw, int ph, int pd,
                         bool three_d, hipStream_t s, SegGraph &g)
{
    g = SegGraph{};
    g.nd = three_d ? 26 : 8;
    if (w <= 0 || h <= 0 || d <= 0) return BGT_OK;
    if (pw < w || ph < h || pd < d) return fail(BGT_ERR_INVALID, "padded frame smaller than skeleton");
    if (pw > 32767 || ph > 32767 || pd > 32767)
        return fail(BGT_ERR_UNSUPPORTED, "coordinates are int16 (sides <= 32767)");
    Frame f{d_skel, w, h, d, pw, ph, pd, (w + 63) / 64, three_d ? 1 : 0};
    const long nchunks = (long)d * h * f.cpr;
    const long nwg = (nchunks + 255) / 256;
    if (nwg > 0x7fffffffL) return fail(BGT_ERR_UNSUPPORTED, "skeleton too large for the pre-pass");
    // scratch: bits, counts, bases, workgroup sums, total
    const size_t need = round_up(sizeof(uint64_t) * nchunks, 256) + 2 * round_up(sizeof(int) * nchunks, 256) +
                        round_up(sizeof(int) * nwg, 256) + 256;
    Workspace &ws = thread_workspace(5);
    BGT_CHECK(workspace_reserve(ws, need));
    char *p = (char *)ws.ptr;
    uint64_t *bits = (uint64_t *)p;
    p += round_up(sizeof(uint64_t) * nchunks, 256);
    int *cnt = (int *)p;
    p += round_up(sizeof(int) * nchunks, 256);
    int *base = (int *)p;
    p += round_up(sizeof(int) * nchunks, 256);
    int *wsum = (int *)p;
    p += round_up(sizeof(int) * nwg, 256);
    long *d_total = (long *)p;
    hipLaunchKernelGGL(seg_count, dim3((unsigned)nwg), dim3(256), 0, s, f, nchunks, bits, cnt, wsum);
    hipLaunchKernelGGL(seg_scan, dim3(1), dim3(1024), 0, s, wsum, (int)nwg, d_total);
    hipLaunchKernelGGL(seg_base, dim3((unsigned)nwg), dim3(256), 0, s, nchunks, cnt, wsum, base);
    BGT_HIP(hipGetLastError());
    PinnedStage &st = thread_stage();
    BGT_CHECK(st.reserve(4096));
    BGT_HIP(hipMemcpyAsync(st.ptr, d_total, sizeof(long), hipMemcpyDeviceToHost, s));
    BGT_HIP(hipStreamSynchronize(s));
    long total = 0;
    std::memcpy(&total, st.ptr, sizeof(long));
    if (total > 0x7fffffffL / g.nd) return fail(BGT_ERR_UNSUPPORTED, "too many skeleton points");
    g.npts = (int)total;
    if (!g.npts) return BGT_OK;
    const size_t out = round_up(sizeof(int64_t) * total, 256) + round_up((size_t)total, 256) +
                       round_up(sizeof(int) * total * g.nd, 256);
    Workspace &wo = thread_workspace(6);
    BGT_CHECK(workspace_reserve(wo, out));
    char *q = (char *)wo.ptr;
    int64_t *d_idx = (int64_t *)q;
    q += round_up(sizeof(int64_t) * total, 256);
    int8_t *d_code = (int8_t *)q;
    q += round_up((size_t)total, 256);
    int *d_nbr = (int *)q;
    hipLaunchKernelGGL(seg_emit, dim3((unsigned)nwg), dim3(256), 0, s, f, nchunks, bits, base,
                       d_idx, d_code, d_nbr);
    BGT_HIP(hipGetLastError());
    // the three arrays lie one after another in the workspace: one copy
    BGT_CHECK(st.reserve(out));
    BGT_HIP(hipMemcpyAsync(st.ptr, wo.ptr, out, hipMemcpyDeviceToHost, s));
    BGT_HIP(hipStreamSynchronize(s));
    // the host trace reads the arrays at random: from ordinary pageable
    // memory (this thread's copy, kept between calls) it runs faster
    thread_local std::vector<char> hv;
    if (hv.size() < out) hv.resize(out);
    std::memcpy(hv.data(), st.ptr, out);
    const char *hq = hv.data();
    g.idx = (const int64_t *)hq;
    g.code = (const int8_t *)(hq + ((const char *)d_code - (const char *)d_idx));
    g.nbr = (const int *)(hq + ((const char *)d_nbr - (const char *)d_idx));
    return BGT_OK;
}

}  // namespace bgt
