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
// pass kernel
// ---------------------------------------------------------------------------

__device__ __forceinline__ void finalize_state(int *st, int K, int *live)
{
    if (st[ST_DONE]) return;
    const int p = st[ST_PASS];
    const unsigned mask =
        __hip_atomic_load((unsigned *)&st[ST_MASK], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int m = __popc(mask);
    if (m < K) {
        st[ST_ITERS] = p * K + m + 1;
        st[ST_PARITY] = (p + 1) & 1;
        st[ST_DONE] = 1;
        if (live) atomicSub(live, 1);
    } else {
        st[ST_PASS] = p + 1;
    }
    st[ST_MASK] = 0;
}

// One subiteration of direction D (0..5 = U, N, E, D, S, W) over the R x NH
// register words of a lane.  The Up step also evaluates the protection mask P
// = voxels deletable in no direction (3D/CUDA_ThinningKernel3D_Linear.cuh:
// 690-702); later steps delete only voxels outside P.  Bit r of rows_changed is
// set when z-row r lost a voxel in one of the interior words [xa, xa + nx).
template <int D, int NH, int R>
__device__ __forceinline__ void step3d(uint32_t (&s)[R][NH], uint32_t *P,
                                       const uint32_t *edges_up, const uint32_t *edges_dn,
                                       uint32_t up_ok, uint32_t dn_ok, int lane, int xa, int nx,
                                       uint32_t &rows_changed)
{
    uint32_t pv[NH];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint32_t zm[NH], zp[NH];
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            if (r == 0) zm[j] = edges_up[j * 64 + lane] & up_ok;
            else zm[j] = pv[j];
            if (r + 1 < R) zp[j] = s[r + 1 < R ? r + 1 : 0][j];
            else zp[j] = edges_dn[j * 64 + lane] & dn_ok;
        }
        uint32_t nw[NH];
        uint32_t racc = 0;
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            uint32_t nb[27];
            gather27<NH>(zm, s[r], zp, j, nb);
            const uint32_t v = s[r][j];
            uint32_t del;
            if (D == 0) {
                const uint32_t cu = rule3<0>(nb);
                const uint32_t any = cu | rule3<1>(nb) | rule3<2>(nb) | rule3<3>(nb) |
                                     rule3<4>(nb) | rule3<5>(nb);
                P[r * NH + j] = v & ~any;
                del = cu & v;
            } else {
                del = rule3<D>(nb) & v & ~P[r * NH + j];
            }
            nw[j] = v ^ del;
            if ((unsigned)(j - xa) < (unsigned)nx) racc |= del;
            __builtin_amdgcn_sched_barrier(0);
        }
        rows_changed |= (racc != 0u ? 1u : 0u) << r;
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            pv[j] = s[r][j];
            s[r][j] = nw[j];
        }
    }
}

template <int NH, int R>
__global__ __launch_bounds__(512) void thin3d_pass(const Item3D *__restrict__ items,
                                                    const Tile3D *__restrict__ tiles, int nitems,
                                                    int K, unsigned *__restrict__ ctr,
                                                    int *__restrict__ live)
{
    extern __shared__ uint32_t edges3[];  // [buf 2][top/bot 2][NW][NH][64]
    __shared__ unsigned s_mask;
    __shared__ int s_last;

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int NW = blockDim.x >> 6;
    const Tile3D tl = tiles ? tiles[blockIdx.x] : Tile3D{0, (int)blockIdx.x};
    const Item3D it = items[tl.item];
    const int brick = tl.brick;
    // brick index = (bz * nby + by) * nbx + bx
    const int nbx = it.nbricks / (it.nby * it.nbz);
    const int bx = brick % nbx, by = (brick / nbx) % it.nby, bz = brick / (nbx * it.nby);

    if (tid == 0) s_mask = 0;
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
        const int H = 6 * K;
        const int xhalo = nbx > 1 ? 1 : 0;
        const int nhc = NH - 2 * xhalo;            // interior words per brick
        const int j_mem0 = bx * nhc - xhalo;       // memory word of register word 0
        const int y = by * it.iy - H + lane;
        const int z_reg0 = bz * it.iz - H + wv * R;
        const bool y_ok = y >= 0 && y < it.h;
        const uint32_t ymask = (lane >= H && lane < H + it.iy) ? 0xffffffffu : 0u;
        const int ra = max(0, H - wv * R), rb = min(R, H + it.iz - wv * R);
        const uint32_t zint = rb > ra ? (uint32_t)(((1ull << rb) - 1ull) & ~((1ull << ra) - 1ull)) : 0u;
        const int xa = xhalo, nx = min(nhc, it.nh - bx * nhc);
        const uint32_t *src = par ? it.buf[1] : it.buf[0];
        uint32_t *dst = par ? it.buf[0] : it.buf[1];

        uint32_t s[R][NH];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int z = z_reg0 + r;
            const bool ok = y_ok && z >= 0 && z < it.d;
            const size_t base = ((size_t)z * it.h + y) * it.nh;
#pragma unroll
            for (int j = 0; j < NH; ++j) {
                const int jm = j_mem0 + j;
                s[r][j] = (ok && jm >= 0 && jm < it.nh) ? src[base + jm] : 0u;
            }
        }
        // protection masks of this lane after the edge buffers: R*NH words per
        // lane at a stride of R*NH+1 words (odd, so the lanes hit distinct banks)
        uint32_t *P = edges3 + 2 * 2 * NW * NH * 64 + tid * (R * NH + 1);
        const int plane = NH * 64;  // words per (buf, tb, wave) block
        auto eidx = [&](int b, int tb, int wave) { return ((b * 2 + tb) * NW + wave) * plane; };
        const int wup = wv > 0 ? wv - 1 : wv, wdn = wv + 1 < NW ? wv + 1 : wv;
        const uint32_t up_ok = wv > 0 ? 0xffffffffu : 0u, dn_ok = wv + 1 < NW ? 0xffffffffu : 0u;
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            edges3[eidx(0, 0, wv) + j * 64 + lane] = s[0][j];
            edges3[eidx(0, 1, wv) + j * 64 + lane] = s[R - 1][j];
        }
        __syncthreads();

        unsigned bits = 0;
        int sidx = 0;
#pragma unroll 1
        for (int itn = 0; itn < K; ++itn) {
            uint32_t rows_changed = 0;
#define BGT_STEP3(D)                                                                          \
    {                                                                                         \
        const int rb_ = sidx & 1, wb_ = (sidx + 1) & 1;                                       \
        step3d<D, NH, R>(s, P, edges3 + eidx(rb_, 1, wup), edges3 + eidx(rb_, 0, wdn), up_ok, \
                         dn_ok, lane, xa, nx, rows_changed);                                  \
        _Pragma("unroll") for (int j = 0; j < NH; ++j)                                        \
        {                                                                                     \
            edges3[eidx(wb_, 0, wv) + j * 64 + lane] = s[0][j];                               \
            edges3[eidx(wb_, 1, wv) + j * 64 + lane] = s[R - 1][j];                           \
        }                                                                                     \
        __syncthreads();                                                                      \
        ++sidx;                                                                               \
    }
            BGT_STEP3(0) BGT_STEP3(1) BGT_STEP3(2) BGT_STEP3(3) BGT_STEP3(4) BGT_STEP3(5)
#undef BGT_STEP3
            if (ymask && (rows_changed & zint)) bits |= 1u << itn;
        }

        if (ymask && y_ok) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int z = z_reg0 + r;
                if (((zint >> r) & 1u) && z >= 0 && z < it.d) {
                    const size_t base = ((size_t)z * it.h + y) * it.nh;
#pragma unroll
                    for (int j = 0; j < NH; ++j) {
                        const int jm = j_mem0 + j;
                        if ((unsigned)(j - xa) < (unsigned)nx) dst[base + jm] = s[r][j];
                    }
                }
            }
        }
        if (bits) atomicOr(&s_mask, bits);
        __syncthreads();
        if (tid == 0) {
            it.chg[par * it.nbricks + brick] = s_mask != 0;
            if (s_mask) atomicOr((unsigned *)&it.st[ST_MASK], s_mask);
        }
    }

    if (tid == 0) {
        __threadfence();
        s_last = atomicAdd(ctr, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        for (int i = tid; i < nitems; i += blockDim.x) finalize_state(items[i].st, K, live);
        __threadfence();
        if (tid == 0) atomicExch(ctr, 0u);
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

int launch_pass3d_nh(const Config3D &cfg, int NH, int nblocks, const Item3D *items,
                     const Tile3D *tiles, int nitems, unsigned *ctr, int *live, hipStream_t s)
{
    const size_t lds = lds_bytes3d(cfg, NH);
    const dim3 g(nblocks), b(cfg.waves * 64);
    const int K = cfg.iters_per_pass;
    static thread_local int max_lds = -1;
    if (max_lds < 0) {
        int dev = 0;
        BGT_HIP(hipGetDevice(&dev));
        BGT_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    }
    if (lds + 64 > (size_t)max_lds)
        return fail(BGT_ERR_UNSUPPORTED, "3D pass needs %zu B of LDS, device allows %d", lds, max_lds);
    switch (NH) {
    case 2: hipLaunchKernelGGL((thin3d_pass<2, 8>), g, b, lds, s, items, tiles, nitems, K, ctr, live); break;
    case 4: hipLaunchKernelGGL((thin3d_pass<4, 6>), g, b, lds, s, items, tiles, nitems, K, ctr, live); break;
    case 6: hipLaunchKernelGGL((thin3d_pass<6, 4>), g, b, lds, s, items, tiles, nitems, K, ctr, live); break;
    case 8: hipLaunchKernelGGL((thin3d_pass<8, 3>), g, b, lds, s, items, tiles, nitems, K, ctr, live); break;
    default: return fail(BGT_ERR_UNSUPPORTED, "no 3D kernel for %d words per line", NH);
    }
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

int launch_pass3d(const Config3D &cfg, const Plan3D &pl, int nblocks, const Item3D *items,
                  const Tile3D *tiles, int nitems, unsigned *ctr, hipStream_t s)
{
    return launch_pass3d_nh(cfg, pl.NH, nblocks, items, tiles, nitems, ctr, nullptr, s);
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
            BGT_CHECK(launch_pass3d(cfg, pl, pl.nbricks, d_item, nullptr, 1, d_ctr, s));
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

int thin3d_device_u8(const uint8_t *d_in, int w, int h, int d, uint8_t *d_out, hipStream_t s,
                     int *iterations, const Config3D &cfg)
{
    if (!d_in || !d_out) return fail(BGT_ERR_INVALID, "null device pointer");
    Plan3D pl;
    BGT_CHECK(plan3d(w, h, d, cfg, pl));
    Workspace &ws = thread_workspace(2);
    BGT_CHECK(workspace_reserve(ws, workspace_bytes3d(pl)));
    Item3D *d_item;
    unsigned *d_ctr;
    Item3D hv;
    BGT_CHECK(setup_item3d(pl, ws.ptr, s, &d_item, &d_ctr, &hv));
    const long nw = (long)pl.state_words;
    const dim3 g((unsigned)((nw + 255) / 256)), b(256);
    hipLaunchKernelGGL(pack3d_u8, g, b, 0, s, d_in, w, h, d, pl.nh, hv.buf[0]);
    BGT_HIP(hipGetLastError());
    BGT_CHECK(run_passes3d(cfg, pl, d_item, d_ctr, hv.st, s, iterations));
    hipLaunchKernelGGL(unpack3d<uint8_t>, g, b, 0, s, d_item, d_out);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(s));
    return BGT_OK;
}

template <typename T>
static int thin3d_host_t(int w, int h, int d, const T *in, T *out, const Config3D &cfg)
{
    Plan3D pl;
    BGT_CHECK(plan3d(w, h, d, cfg, pl));
    const size_t bytes = (size_t)w * h * d * sizeof(T);
    Workspace &ws = thread_workspace(2);
    Workspace &io = thread_workspace(3);
    BGT_CHECK(workspace_reserve(ws, workspace_bytes3d(pl)));
    BGT_CHECK(workspace_reserve(io, bytes));
    hipStream_t s = nullptr;
    BGT_HIP(hipMemcpyAsync(io.ptr, in, bytes, hipMemcpyHostToDevice, s));
    Item3D *d_item;
    unsigned *d_ctr;
    Item3D hv;
    BGT_CHECK(setup_item3d(pl, ws.ptr, s, &d_item, &d_ctr, &hv));
    const long nw = (long)pl.state_words;
    const dim3 g((unsigned)((nw + 255) / 256)), b(256);
    if (std::is_same<T, uint8_t>::value)
        hipLaunchKernelGGL(pack3d_u8, g, b, 0, s, (const uint8_t *)io.ptr, w, h, d, pl.nh, hv.buf[0]);
    else
        hipLaunchKernelGGL(pack3d_t<T>, g, b, 0, s, (const T *)io.ptr, w, h, d, pl.nh, hv.buf[0]);
    BGT_HIP(hipGetLastError());
    BGT_CHECK(run_passes3d(cfg, pl, d_item, d_ctr, hv.st, s, nullptr));
    hipLaunchKernelGGL(unpack3d<T>, g, b, 0, s, d_item, (T *)io.ptr);
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
