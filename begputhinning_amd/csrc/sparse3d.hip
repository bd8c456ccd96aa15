// sparse3d.hip -- 3D thinning over the list of non-empty words.
//
// Volumes on this path are sparse (a few percent foreground), so the work is
// organised around the words (32 voxels of an x-line) that hold foreground:
// a compacted list of (item, word) entries, in raster order inside each
// workgroup of the compaction so neighbouring lanes touch neighbouring words.
// A subiteration is one launch over the list; every entry gathers its 3x3x3
// word neighbourhood from the previous subiteration's state array (the state
// of a batch is a few tens of MB and stays in the caches), evaluates the rule
// and writes the other array.  Empty words are zero in both arrays and can
// never change (subiterations only clear bits), so the list only ever
// shrinks: after each iteration, entries whose word became empty or whose item
// converged are compacted away.  No tile, halo or neighbour bookkeeping is
// needed, and each subiteration is synchronous over the whole volume, as in
// the reference's semantics (3D/CUDA_ThinningKernel3D_Linear.cuh:739-782).
//
// An iteration is six launches, U (with the protection pre-test of all six
// directions, 3D/CUDA_ThinningKernel3D_Linear.cuh:690-702), N, E, D, S, W,
// alternating arrays 0 -> 1 -> 0 ..., so it ends in array 0, and a compaction
// launch that builds the next list and records per-item convergence.  W also
// activates the blocks around the blocks its changes mark dirty.  The
// subiterations run over the active list only (the words near the previous
// iteration's changes, see sparse3d.h); the other words of the list cannot
// change and hold the same value in both arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "pack.h"
#include "sparse3d.h"
#include "thin3d_kernels.h"

namespace bgt {

namespace {

const int kCompactPer = 1024;  // list entries per compaction workgroup (4 per thread)
const int kPackRuns = kPackRuns3D;

// Appends the flagged entries (item, word[j], blk[j]) of a workgroup to
// `out`, up to N per thread, keeping their order, with one atomic per
// workgroup.  Every thread of the block must call it (flags may be 0); two
// calls in a row need a barrier between them.
template <int N>
__device__ __forceinline__ void append_n(const bool (&f)[N], int item, const uint32_t (&word)[N],
                                         const int (&blk)[N], SList out)
{
    __shared__ int s_cnt[16];
    __shared__ int s_base;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    const uint64_t lt = (1ull << lane) - 1ull;
    int c = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) c += f[j];
    // exclusive prefix over the wave of the per-thread counts (0..N), bit by bit
    int pre = 0, tot = 0;
#pragma unroll
    for (int bit = 0; (1 << bit) <= N; ++bit) {
        const uint64_t m = __ballot((c >> bit) & 1);
        pre += __popcll(m & lt) << bit;
        tot += __popcll(m) << bit;
    }
    if (lane == 0) s_cnt[wv] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int k = 0; k < nw; ++k) sum += s_cnt[k];
        s_base = sum ? atomicAdd(out.count, sum) : 0;
    }
    __syncthreads();
    int pos = s_base + pre;
    for (int k = 0; k < wv; ++k) pos += s_cnt[k];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (f[j]) {
            out.item[pos] = item;
            out.word[pos] = word[j];
            out.blk[pos] = blk[j];
            ++pos;
        }
    }
}

// Appends the entries (item, word, block) of a workgroup flagged for list
// `A` (fa) and for list `B` (fb) -- up to 4 consecutive entries per thread --
// keeping their order, with one atomic per list per workgroup.  Every thread
// of the block must call it.
__device__ __forceinline__ void append4_both(const bool (&fa)[4], const bool (&fb)[4],
                                             const int (&item)[4], const uint32_t (&word)[4],
                                             const int (&blk)[4], SList A, SList B)
{
    __shared__ int s_cnt[2][16];
    __shared__ int s_base[2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    const uint64_t lt = (1ull << lane) - 1ull;
    int ca = 0, cb = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ca += fa[j];
        cb += fb[j];
    }
    // exclusive prefix over the wave of the per-thread counts (0..4), bit by bit
    int pa = 0, pb = 0, ta = 0, tb = 0;
#pragma unroll
    for (int bit = 0; bit < 3; ++bit) {
        const uint64_t ma = __ballot((ca >> bit) & 1), mb = __ballot((cb >> bit) & 1);
        pa += __popcll(ma & lt) << bit;
        ta += __popcll(ma) << bit;
        pb += __popcll(mb & lt) << bit;
        tb += __popcll(mb) << bit;
    }
    if (lane == 0) {
        s_cnt[0][wv] = ta;
        s_cnt[1][wv] = tb;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        int sum = 0;
        for (int k = 0; k < nw; ++k) sum += s_cnt[threadIdx.x][k];
        s_base[threadIdx.x] = sum ? atomicAdd(threadIdx.x ? B.count : A.count, sum) : 0;
    }
    __syncthreads();
    pa += s_base[0];
    pb += s_base[1];
    for (int k = 0; k < wv; ++k) {
        pa += s_cnt[0][k];
        pb += s_cnt[1][k];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (fa[j]) {
            A.item[pa] = item[j];
            A.word[pa] = word[j];
            A.blk[pa] = blk[j];
            ++pa;
        }
        if (fb[j]) {
            B.item[pb] = item[j];
            B.word[pb] = word[j];
            B.blk[pb] = blk[j];
            ++pb;
        }
    }
}

__global__ void sparse3d_init(int *states, int *changed, int *ctl, int nitems, int *bstamp,
                              int *dstamp, int nblocks)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nitems * ST_INTS) states[t] = 0;
    if (t < nitems) changed[t] = 0;
    if (t < SC_INTS) ctl[t] = t == SC_LIVE ? nitems : 0;
    if (t < nblocks) {
        bstamp[t] = 0;
        dstamp[t] = 0;
    }
}

__global__ void sparse3d_zero(uint32_t *a, uint32_t *b, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        a[t] = 0u;
        b[t] = 0u;
    }
}

// Each thread packs kPackRuns runs of 64 voxels (two data words) of x-lines:
// writes both state arrays and lists the non-empty words (list 0, also the
// first iteration's active list).
__global__ __launch_bounds__(256) void sparse3d_pack(const SItem3D *__restrict__ items,
                                                     const SIO *__restrict__ io,
                                                     const int *__restrict__ blk, uint32_t *s0,
                                                     uint32_t *s1, SList out)
{
    const int i = blk ? blk[blockIdx.x] : 0;
    const SItem3D it = items[i];
    const SIO f = io[i];
    const uint32_t nq = (uint32_t)it.nh >> 1;
    const uint32_t nruns = (uint32_t)it.h * (uint32_t)it.d * nq;
    __shared__ uint32_t s_masks[4][320];
    bool fl[2 * kPackRuns];
    uint32_t wd[2 * kPackRuns];
    int bk[2 * kPackRuns];
#pragma unroll
    for (int k = 0; k < kPackRuns; ++k) {
        // runs k*256 .. k*256 + 255 of the workgroup: whole waves of consecutive runs
        const uint32_t t = ((uint32_t)(blockIdx.x - f.blk0) * kPackRuns + k) * 256u + threadIdx.x;
        const bool valid = t < nruns;
        const uint32_t line = valid ? t / nq : 0;
        const int x0 = valid ? (int)(t - line * nq) * 64 : 0;  // < w: nq = ceil(w / 64)
        // the caller's lines follow each other in memory, and so do the runs
        const uint64_t v = pack_runs_wave_u8(f.in + (size_t)line * it.w + x0, min(64, it.w - x0),
                                             valid, s_masks[threadIdx.x >> 6]);
        uint32_t lo = 0, hi = 0, g = 0;
        int b0 = 0;
        if (valid) {
            lo = (uint32_t)v;
            hi = (uint32_t)(v >> 32);
            const uint32_t y = line % (uint32_t)it.h, z = line / (uint32_t)it.h;
            g = (uint32_t)it.base + ((z + 1) * (uint32_t)(it.h + 2) + y + 1) * (uint32_t)it.P + 1 +
                (uint32_t)x0 / 32;
            b0 = it.bbase + ((int)(z >> 3) * it.nby + (int)(y >> 3)) * it.nh + x0 / 32;
            s0[g] = lo;
            s0[g + 1] = hi;
            s1[g] = lo;
            s1[g + 1] = hi;
        }
        fl[2 * k] = lo != 0;
        fl[2 * k + 1] = hi != 0;
        wd[2 * k] = g;
        wd[2 * k + 1] = g + 1;
        bk[2 * k] = b0;
        bk[2 * k + 1] = b0 + 1;
    }
    // one list atomic per workgroup: same-address atomics are what bounds a
    // pack of many small workgroups
    append_n<2 * kPackRuns>(fl, i, wd, bk, out);
}

// Pixels of one z-plane of an item in state array 0.
struct PlaneBits {
    const uint32_t *s0;
    uint32_t g0;  // word 0 of line y = 0
    uint32_t P;
    __device__ uint32_t bits16(uint32_t y, uint32_t x) const
    {
        const uint32_t j = g0 + y * P + (x >> 5);
        const uint64_t v = ((uint64_t)s0[j] | ((uint64_t)s0[j + 1] << 32)) >> (x & 31);
        return (uint32_t)v & 0xffffu;
    }
    __device__ uint32_t bit(uint32_t y, uint32_t x) const
    {
        return (s0[g0 + y * P + (x >> 5)] >> (x & 31)) & 1u;
    }
    // 64 voxels from x of line y (three words; the words after a line's data
    // belong to the next line's margin or data: the caller masks them)
    __device__ uint64_t bits64(uint32_t y, uint32_t x) const
    {
        const uint32_t j = g0 + y * P + (x >> 5), s = x & 31;
        const uint64_t a = (uint64_t)s0[j] | ((uint64_t)s0[j + 1] << 32);
        return s ? (a >> s) | ((uint64_t)s0[j + 2] << (64 - s)) : a;
    }
};

// One thread per 64-byte block of a z-plane of the caller's buffer (whole
// waves per plane: the stores are exchanged within the wave).
__global__ __launch_bounds__(256) void sparse3d_unpack(const SItem3D *__restrict__ items,
                                                       const SIO *__restrict__ io,
                                                       const int *__restrict__ ublk,
                                                       const uint32_t *__restrict__ s0)
{
    const int i = ublk ? ublk[blockIdx.x] : 0;
    const SItem3D it = items[i];
    const SIO f = io[i];
    const uint32_t plane = (uint32_t)it.w * (uint32_t)it.h;
    const uint32_t cpp = unpack_wave_blocks(plane);  // whole waves per plane
    const uint32_t t = (uint32_t)(blockIdx.x - f.ublk0) * 256u + threadIdx.x;
    const uint32_t z = t / cpp, c = t - z * cpp;
    if (z >= (uint32_t)it.d) return;
    const PlaneBits src{s0, (uint32_t)it.base + ((z + 1) * (uint32_t)(it.h + 2) + 1) * (uint32_t)it.P + 1,
                        (uint32_t)it.P};
    unpack_block64_wave_u8(f.out + (size_t)z * plane, plane, (uint32_t)it.w, c, src);
}

// Block b of item `it` as (bx, by, bz).
__device__ __forceinline__ void block_xyz(const SItem3D &it, int b, int &bx, int &by, int &bz)
{
    const int l = b - it.bbase, byz = l / it.nh;
    bx = l - byz * it.nh;
    bz = byz / it.nby;
    by = byz - bz * it.nby;
}

// The 3x3x3 word neighbourhood of interior word g, each word shifted so that
// bit k holds the neighbour of voxel k (raster (dz, dy, dx) order).
__device__ __forceinline__ void gather27(const uint32_t *__restrict__ src, uint32_t g, int2 gp,
                                         uint32_t (&nb)[27])
{
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
            const uint32_t *r = src + (g + (uint32_t)((dz - 1) * gp.y + (dy - 1) * gp.x));
            const uint32_t a = r[-1], b = r[0], c = r[1];
            const int q = dz * 9 + dy * 3;
            nb[q + 0] = __builtin_amdgcn_alignbit(b, a, 31);  // x-1
            nb[q + 1] = b;
            nb[q + 2] = __builtin_amdgcn_alignbit(c, b, 1);   // x+1
        }
}

// Subiteration D on word g: src -> dst.  U (D = 0) also returns the
// protection mask and the word's value at the start of the iteration; the
// other steps take the protection mask.  Returns the deleted voxels.
template <int D>
__device__ __forceinline__ uint32_t sub_word(const uint32_t *__restrict__ src, uint32_t *dst,
                                             uint32_t g, int2 gp, uint32_t &prot, uint32_t &first,
                                             uint32_t &nv)
{
    uint32_t nb[27];
    gather27(src, g, gp, nb);
    const uint32_t v = nb[13];
    uint32_t del;
    if (D == 0) {
        const uint32_t cu = rule3<0>(nb);
        const uint32_t any = cu | rule3<1>(nb) | rule3<2>(nb) | rule3<3>(nb) | rule3<4>(nb) |
                             rule3<5>(nb);
        prot = v & ~any;
        first = v;
        del = cu & v;
    } else {
        del = rule3<D>(nb) & v & ~prot;
    }
    nv = v ^ del;
    dst[g] = nv;
    return del;
}

// Subiteration D (0..5 = U, N, E, D, S, W) over the active list (list 0 in
// the first iteration): src -> dst.  The last one marks the blocks of the
// words the iteration changed dirty and activates the 27 blocks around each
// (the first entry to mark a block in this iteration stamps its
// neighbourhood).
template <int D>
__global__ __launch_bounds__(256) void sparse3d_step(const uint32_t *__restrict__ src,
                                                     uint32_t *__restrict__ dst,
                                                     const SItem3D *__restrict__ items,
                                                     const int2 *__restrict__ geo, SList act,
                                                     SList list0, uint32_t *__restrict__ prot,
                                                     uint32_t *__restrict__ first,
                                                     int *__restrict__ changed,
                                                     const int *__restrict__ ctl, int *dstamp,
                                                     int *bstamp)
{
    // The list is cut into one contiguous range per XCD (workgroups are dealt
    // to the 8 XCDs round-robin; the grid is a multiple of 8 or smaller than
    // 8), so that neighbouring entries, which share most of their 3x3x3
    // neighbourhoods, are gathered through the same L2.
    const int itn = ctl[SC_ITER];
    const SList list = itn == 0 ? list0 : act;
    const int n = itn == 0 ? ctl[SC_COUNT0] : ctl[SC_ACOUNT0 + (itn & 1)];
    const int nx = gridDim.x >= 8 ? 8 : 1;
    const int xcd = blockIdx.x % nx, local = blockIdx.x / nx, per = gridDim.x / nx;
    const int lo = (int)((long)n * xcd / nx), hi = (int)((long)n * (xcd + 1) / nx);
    for (int k = lo + local * 256 + threadIdx.x; k < hi; k += per * 256) {
        const int ii = list.item[k];
        const uint32_t g = list.word[k];
        uint32_t pk = D == 0 ? 0u : prot[k], fk = 0u, nv;
        const uint32_t del = sub_word<D>(src, dst, g, geo[ii], pk, fk, nv);
        if (D == 0) {
            prot[k] = pk;
            first[k] = fk;
        }
        if (del) changed[ii] = 1;
        if (D == 5 && nv != first[k]) {
            const int b = list.blk[k], stamp = itn + 1;
            if (dstamp[b] != stamp && atomicMax(&dstamp[b], stamp) < stamp) {
                const SItem3D it = items[ii];
                int bx, by, bz;
                block_xyz(it, b, bx, by, bz);
                for (int dz = -1; dz <= 1; ++dz) {
                    if (bz + dz < 0 || bz + dz >= it.nbz) continue;
                    for (int dy = -1; dy <= 1; ++dy) {
                        if (by + dy < 0 || by + dy >= it.nby) continue;
                        const int row = b + (dz * it.nby + dy) * it.nh;
                        for (int dx = -1; dx <= 1; ++dx)
                            if (bx + dx >= 0 && bx + dx < it.nh) bstamp[row + dx] = stamp;
                    }
                }
            }
        }
    }
}

// Keep the entries whose word is non-empty and whose item changed in this
// iteration (an item whose iteration deleted nothing has converged); those in
// a block activated in this iteration also form the next active list.  A word
// emptied by the last subiteration (W) still holds its bits from before that
// step in array 1; once off the list nothing would rewrite it, and the next
// iteration's odd steps read array 1, so it is cleared here.
//
// The last workgroup to finish then advances the items: each live item whose
// iteration deleted nothing is done, with that iteration counted.  It reads
// only what earlier launches wrote, and it clears `changed` and the input
// list's count only after every other workgroup has finished reading them
// (their block-counter increments follow their reads), so no fence is needed.
__global__ __launch_bounds__(256) void sparse3d_compact(const uint32_t *__restrict__ s0,
                                                        uint32_t *__restrict__ s1,
                                                        int *__restrict__ changed, const SList in,
                                                        SList out, SList act,
                                                        const int *__restrict__ bstamp,
                                                        int *__restrict__ states,
                                                        int *__restrict__ ctl, int nitems)
{
    __shared__ int s_last;
    const int n = *in.count;
    const int itn = ctl[SC_ITER];
    act.count = ctl + SC_ACOUNT0 + ((itn + 1) & 1);
    // Each thread takes 4 consecutive entries.  Workgroups past the list take
    // no part (the grid is sized by an upper bound of n); the others count
    // themselves in to elect the last one.
    const int active = n > 0 ? (n + kCompactPer - 1) / kCompactPer : 1;
    if ((int)blockIdx.x >= active) return;
    {
        const int k0 = ((int)blockIdx.x * (int)blockDim.x + (int)threadIdx.x) * 4;
        int ii[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
        uint32_t g[4] = {0u, 0u, 0u, 0u};
        bool keep[4], ak[4];
        if (k0 + 4 <= n) {  // lists are 256-byte aligned: k0 is 16-byte aligned
            const int4 vi = *reinterpret_cast<const int4 *>(in.item + k0);
            const uint4 vw = *reinterpret_cast<const uint4 *>(in.word + k0);
            const int4 vb = *reinterpret_cast<const int4 *>(in.blk + k0);
            ii[0] = vi.x; ii[1] = vi.y; ii[2] = vi.z; ii[3] = vi.w;
            g[0] = vw.x; g[1] = vw.y; g[2] = vw.z; g[3] = vw.w;
            b[0] = vb.x; b[1] = vb.y; b[2] = vb.z; b[3] = vb.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (k0 + j < n) {
                    ii[j] = in.item[k0 + j];
                    g[j] = in.word[k0 + j];
                    b[j] = in.blk[k0 + j];
                }
        }
        int live[4], bs[4];
        uint32_t v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool ok = k0 + j < n;
            live[j] = ok ? changed[ii[j]] : 0;
            v[j] = ok ? s0[g[j]] : 0u;
            bs[j] = ok ? bstamp[b[j]] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            keep[j] = live[j] != 0 && v[j] != 0u;
            if (live[j] != 0 && v[j] == 0u) s1[g[j]] = 0u;
            ak[j] = keep[j] && bs[j] == itn + 1;
        }
        append4_both(keep, ak, ii, g, b, out, act);
    }
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(ctl + SC_BLOCKS, 1) == active - 1;
    __syncthreads();
    if (!s_last) return;
    for (int i = threadIdx.x; i < nitems; i += blockDim.x) {
        int *st = states + (size_t)i * ST_INTS;
        if (!st[ST_DONE] && !changed[i]) {
            st[ST_DONE] = 1;
            st[ST_ITERS] = itn + 1;
            st[ST_PARITY] = 0;
            atomicSub(ctl + SC_LIVE, 1);
        }
        changed[i] = 0;
    }
    if (threadIdx.x == 0) {
        ctl[SC_ITER] = itn + 1;
        ctl[SC_BLOCKS] = 0;
        *in.count = 0;  // the next iteration compacts into this list
        ctl[SC_ACOUNT0 + (itn & 1)] = 0;  // and builds its active list here
    }
}


// The remaining iterations of item blockIdx.x (see sparse3d_tail).  All
// branches on shared values are taken after a barrier, so every thread
// reaches every barrier; the loop ends when an iteration changes nothing,
// which each iteration that does change something brings closer (it deletes
// a voxel).
__global__ __launch_bounds__(1024) void sparse3d_tail_kernel(
    const SItem3D *__restrict__ items, const int2 *__restrict__ geo, const SList act,
    const SList l0, const SList l1, uint32_t *s0, uint32_t *s1, uint32_t *prot, uint32_t *first,
    int *states, int *ctl, int *bstamp, int *dstamp)
{
    __shared__ int s_na, s_nd, s_nab, s_nn, s_changed;
    const int i = blockIdx.x, tid = threadIdx.x, NT = blockDim.x;
    int *st = states + (size_t)i * ST_INTS;
    if (st[ST_DONE]) return;  // no other workgroup writes item i's state
    const SItem3D it = items[i];
    const int2 gp = geo[i];
    // per-item scratch: active words (two generations), dirty and active blocks
    uint32_t *aw[2] = {l0.word + it.dbase, l1.word + it.dbase};
    int *ab[2] = {l0.blk + it.dbase, l1.blk + it.dbase};
    int *dirty = l0.item + it.dbase, *actb = l1.item + it.dbase;
    uint32_t *pr = prot + it.dbase, *fi = first + it.dbase;
    const int itn0 = ctl[SC_ITER];
    if (tid == 0) {
        s_na = 0;
        s_nd = 0;
        s_nab = 0;
        s_nn = 0;
        s_changed = 0;
    }
    __syncthreads();
    // the item's entries of the current active list (the host launches this
    // after iteration 0, whose active list is list 0, the scratch here)
    {
        const int ng = ctl[SC_ACOUNT0 + (itn0 & 1)];
        for (int k = tid; k < ng; k += NT)
            if (act.item[k] == i) {
                const int p = atomicAdd(&s_na, 1);
                aw[0][p] = act.word[k];
                ab[0][p] = act.blk[k];
            }
    }
    __syncthreads();
    int cur = 0;
    for (int t = itn0;; ++t) {
        const int na = s_na, stamp = t + 1;
        const uint32_t *w = aw[cur];
#define BGT_TAIL_SUB(D, SRC, DST)                                                          \
    for (int k = tid; k < na; k += NT) {                                                   \
        uint32_t pk = D == 0 ? 0u : pr[k], fk = 0u, nv;                                     \
        const uint32_t del = sub_word<D>(SRC, DST, w[k], gp, pk, fk, nv);                   \
        if (D == 0) {                                                                      \
            pr[k] = pk;                                                                    \
            fi[k] = fk;                                                                    \
        }                                                                                  \
        if (del) s_changed = 1;                                                            \
        if (D == 5 && nv != fi[k]) {                                                       \
            const int b = ab[cur][k];                                                      \
            if (atomicMax(&dstamp[b], stamp) < stamp) dirty[atomicAdd(&s_nd, 1)] = b;      \
        }                                                                                  \
    }                                                                                      \
    __syncthreads();
        BGT_TAIL_SUB(0, s0, s1)
        BGT_TAIL_SUB(1, s1, s0)
        BGT_TAIL_SUB(2, s0, s1)
        BGT_TAIL_SUB(3, s1, s0)
        BGT_TAIL_SUB(4, s0, s1)
        BGT_TAIL_SUB(5, s1, s0)
#undef BGT_TAIL_SUB
        if (!s_changed) {
            if (tid == 0) {
                st[ST_ITERS] = t + 1;
                st[ST_PARITY] = 0;
                st[ST_DONE] = 1;
                atomicSub(ctl + SC_LIVE, 1);
            }
            return;
        }
        // words W emptied still hold their bits from before W in array 1
        for (int k = tid; k < na; k += NT)
            if (s0[w[k]] == 0u) s1[w[k]] = 0u;
        // activate the 27 blocks around every dirty block
        const int nd = s_nd;
        for (int x = tid; x < nd * 27; x += NT) {
            const int b = dirty[x / 27], j = x % 27;
            const int dz = j / 9 - 1, dy = (j / 3) % 3 - 1, dx = j % 3 - 1;
            int bx, by, bz;
            block_xyz(it, b, bx, by, bz);
            if (bx + dx < 0 || bx + dx >= it.nh || by + dy < 0 || by + dy >= it.nby ||
                bz + dz < 0 || bz + dz >= it.nbz)
                continue;
            const int nb = b + (dz * it.nby + dy) * it.nh + dx;
            if (atomicMax(&bstamp[nb], stamp) < stamp) actb[atomicAdd(&s_nab, 1)] = nb;
        }
        __syncthreads();
        // the next active words: the non-empty words of the active blocks
        const int nab = s_nab;
        for (int x = tid; x < nab * 64; x += NT) {
            const int b = actb[x >> 6], q = x & 63;
            int bx, by, bz;
            block_xyz(it, b, bx, by, bz);
            const int y = by * 8 + (q & 7), z = bz * 8 + (q >> 3);
            if (y >= it.h || z >= it.d) continue;
            const uint32_t g = (uint32_t)it.base +
                               ((uint32_t)(z + 1) * (uint32_t)(it.h + 2) + (uint32_t)(y + 1)) *
                                   (uint32_t)it.P +
                               1u + (uint32_t)bx;
            if (s0[g] != 0u) {
                const int p = atomicAdd(&s_nn, 1);
                aw[cur ^ 1][p] = g;
                ab[cur ^ 1][p] = b;
            }
        }
        __syncthreads();
        if (tid == 0) {
            s_na = s_nn;
            s_nn = 0;
            s_nd = 0;
            s_nab = 0;
            s_changed = 0;
        }
        __syncthreads();
        cur ^= 1;
    }
}

}  // namespace

int sparse3d_tail(const Sparse3D &e, hipStream_t s)
{
    if (e.nitems > 0)
        hipLaunchKernelGGL(sparse3d_tail_kernel, dim3(e.nitems), dim3(1024), 0, s, e.items, e.geo,
                           e.act, e.list[0], e.list[1], e.state[0], e.state[1], e.prot, e.first,
                           e.states, e.ctl, e.bstamp, e.dstamp);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

int sparse3d_layout(const int *w, const int *h, const int *d, int n, Layout3D &lay)
{
    lay = Layout3D{};
    lay.items.assign(n, SItem3D{});
    lay.geo.assign(n, int2{0, 0});
    lay.blk0.assign(n, 0);
    lay.ublk0.assign(n, 0);
    long base = 0, data = 0, blocks = 0;
    for (int i = 0; i < n; ++i) {
        if (w[i] <= 0 || h[i] <= 0 || d[i] <= 0)
            return fail(BGT_ERR_INVALID, "3D size must be >= 1 (got %d x %d x %d)", w[i], h[i], d[i]);
        if ((long)w[i] * h[i] >= 0x7fffffffL)
            return fail(BGT_ERR_UNSUPPORTED, "3D plane of %d x %d voxels exceeds 2^31", w[i], h[i]);
        SItem3D &it = lay.items[i];
        it.w = w[i];
        it.h = h[i];
        it.d = d[i];
        it.nh = 2 * ((w[i] + 63) / 64);
        // + 1: the trailing zero word is the +x neighbour of the last word of
        // the last margin line, which the corner interior word gathers
        const long P = it.nh + 1, plane = P * (h[i] + 2), total = plane * (d[i] + 2) + 1;
        if (base + total >= 0x7fffffffL || plane >= 0x7fffffffL)
            return fail(BGT_ERR_UNSUPPORTED, "3D state of %ld words exceeds 2^31", base + total);
        it.P = (int)P;
        it.plane = (int)plane;
        it.base = base;
        lay.geo[i] = int2{it.P, it.plane};
        it.nby = (h[i] + 7) / 8;
        it.nbz = (d[i] + 7) / 8;
        it.bbase = (int)std::min<long>(blocks, 0x7fffffffL);
        it.dbase = (int)std::min<long>(data, 0x7fffffffL);
        blocks += (long)it.nh * it.nby * it.nbz;
        base += total;
        data += (long)h[i] * d[i] * it.nh;
        const long threads = (long)h[i] * d[i] * (it.nh / 2);
        lay.blk0[i] = (int)lay.blk.size();
        for (long k = 0; k < (threads + 256 * kPackRuns - 1) / (256 * kPackRuns); ++k)
            lay.blk.push_back(i);
        const long uthreads = (long)d[i] * unpack_wave_blocks((uint32_t)(w[i] * h[i]));
        lay.ublk0[i] = (int)lay.ublk.size();
        for (long k = 0; k < (uthreads + 255) / 256; ++k) lay.ublk.push_back(i);
    }
    if (data >= 0x7fffffffL) return fail(BGT_ERR_UNSUPPORTED, "3D data exceeds 2^31 words");
    if (blocks >= 0x7fffffffL) return fail(BGT_ERR_UNSUPPORTED, "3D blocks exceed 2^31");
    if (lay.blk.size() >= 0x7fffffffUL || lay.ublk.size() >= 0x7fffffffUL)
        return fail(BGT_ERR_UNSUPPORTED, "3D grid too large");
    lay.nwords = base;
    lay.data_words = data;
    lay.nblocks = blocks;
    return BGT_OK;
}

int sparse3d_grid(int blocks_per_cu)
{
    // device values, read once (a function-local static is initialised by
    // exactly one thread; concurrent first calls wait for it)
    static const int2 dev_cus_occ = [] {
        int dev = 0, c = 256, per_cu = 4;
        if (hipGetDevice(&dev) == hipSuccess) {
            (void)hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)sparse3d_step<0>,
                                                             256, 0) != hipSuccess || per_cu < 1)
                per_cu = 4;
        }
        return int2{std::max(1, c), std::max(1, per_cu)};
    }();
    return dev_cus_occ.x * (blocks_per_cu > 0 ? blocks_per_cu : dev_cus_occ.y);
}

int sparse3d_begin(const Sparse3D &e, bool zero, hipStream_t s)
{
    const int ninit = std::max({e.nitems * ST_INTS, (int)SC_INTS, e.nblocks});
    hipLaunchKernelGGL(sparse3d_init, dim3((ninit + 255) / 256), dim3(256), 0, s, e.states,
                       e.changed, e.ctl, e.nitems, e.bstamp, e.dstamp, e.nblocks);
    if (zero && e.nwords > 0)
        hipLaunchKernelGGL(sparse3d_zero, dim3((unsigned)((e.nwords + 255) / 256)), dim3(256), 0, s,
                           e.state[0], e.state[1], e.nwords);
    if (e.nblk > 0)
        hipLaunchKernelGGL(sparse3d_pack, dim3(e.nblk), dim3(256), 0, s, e.items, e.io, e.blk,
                           e.state[0], e.state[1], e.list[0]);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

int sparse3d_iteration(const Sparse3D &e, int cur, int bound, int abound, int grid, hipStream_t s)
{
    const SList L = e.list[cur], N = e.list[cur ^ 1], A = e.act;
    const int need = std::max(1, (bound + kCompactPer - 1) / kCompactPer);
    int steps = std::min(grid, std::max(1, (abound + 255) / 256));
    if (steps >= 8) steps &= ~7;  // whole XCD ranges
    const dim3 g(steps), b(256);
    uint32_t *s0 = e.state[0], *s1 = e.state[1];
#define BGT_STEP(D, SRC, DST)                                                                    \
    hipLaunchKernelGGL(sparse3d_step<D>, g, b, 0, s, SRC, DST, e.items, e.geo, A, e.list[0],     \
                       e.prot, e.first, e.changed, e.ctl, e.dstamp, e.bstamp)
    BGT_STEP(0, s0, s1);
    BGT_STEP(1, s1, s0);
    BGT_STEP(2, s0, s1);
    BGT_STEP(3, s1, s0);
    BGT_STEP(4, s0, s1);
    BGT_STEP(5, s1, s0);
#undef BGT_STEP
    hipLaunchKernelGGL(sparse3d_compact, dim3(need), b, 0, s, s0, s1, e.changed, L, N, A, e.bstamp,
                       e.states, e.ctl, e.nitems);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

int sparse3d_end(const Sparse3D &e, hipStream_t s)
{
    if (e.nublk > 0)
        hipLaunchKernelGGL(sparse3d_unpack, dim3(e.nublk), dim3(256), 0, s, e.items, e.io, e.ublk,
                           e.state[0]);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

}  // namespace bgt
