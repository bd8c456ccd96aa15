// tile3d.hip -- 3D thinning, one launch per iteration over queued tiles
// (see tile3d.h for the layout, the tiles and why the result is exact).
//
// Reference path replaced: 3D/CUDA_ThinningKernel3D_Linear.cuh:704-891 (the
// driver kernel: pre-test and six subiterations per iteration on a 64^3 tile
// staged in shared memory, with overlapping tiles whose halo is too thin for
// the iterations it runs) and its host loop 3D/LibThinning3D.cu:64-74 (one
// blocking device-to-host count per launch).  Here a tile's halo is exactly
// the iteration's dependence radius, the tile runs one iteration per launch,
// and only tiles near the previous iteration's changes are queued.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "pack.h"
#include "thin3d_kernels.h"
#include "tile3d.h"

namespace bgt {

namespace {

constexpr int kHalo = 6;  // voxels a voxel's value depends on, per iteration
constexpr int RX = kTileWordsX + 2;            // region words along x
constexpr int RY = kTileLines + 2 * kHalo;     // region lines
constexpr int RZ = kTilePlanes + 2 * kHalo;    // region planes
constexpr int NROWS = RY * RZ;                 // region x-rows
constexpr int NR = RX * NROWS;                 // region words
constexpr int NINT = kTileWordsX * kTileLines * kTilePlanes;  // interior words

// Region rows (ly, lz) at distance >= j from the region's y/z boundary,
// i.e. ly in [j, RY - 1 - j] and lz in [j, RZ - 1 - j].
constexpr int rows_at_least(int j) { return (RY - 2 * j) * (RZ - 2 * j); }
constexpr int imin(int a, int b) { return a < b ? a : b; }
constexpr int row_level(int ly, int lz)
{
    return imin(imin(imin(ly, RY - 1 - ly), imin(lz, RZ - 1 - lz)), kHalo);
}

// The region rows ordered by level, highest first (raster order within a
// level): the interior rows (level 6) come first, and the rows step j
// evaluates (level >= j) are a prefix.
struct RowOrder {
    uint16_t row[NROWS];
};
constexpr RowOrder make_row_order()
{
    RowOrder o{};
    int k = 0;
    for (int lev = kHalo; lev >= 0; --lev)
        for (int lz = 0; lz < RZ; ++lz)
            for (int ly = 0; ly < RY; ++ly)
                if (row_level(ly, lz) == lev) o.row[k++] = (uint16_t)(lz * RY + ly);
    return o;
}
__constant__ RowOrder kRowOrder = make_row_order();

static_assert(rows_at_least(kHalo) == kTileLines * kTilePlanes, "interior rows come first");
static_assert(kTileWordsX == 2, "a pack run (64 voxels) is one tile wide");
static_assert(kTileLines == 16 && kTilePlanes == 16, "a tile's interior is 2 x 2 x 2 activity blocks");

// The 3x3x3 word neighbourhood of LDS word r of a region, each word shifted
// so that bit k holds the neighbour of voxel k (raster (dz, dy, dx) order).
// The x-halo words at a region row's ends read the adjacent row's words as
// their outer neighbours: those bits are more than 6 voxels from the
// interior and never reach it.
__device__ __forceinline__ void gather27_lds(const uint32_t *src, int r, uint32_t (&nb)[27])
{
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
            const uint32_t *row = src + r + (dz - 1) * RX * RY + (dy - 1) * RX;
            const uint32_t a = row[-1], b = row[0], c = row[1];
            const int q = dz * 9 + dy * 3;
            nb[q + 0] = __builtin_amdgcn_alignbit(b, a, 31);  // x-1
            nb[q + 1] = b;
            nb[q + 2] = __builtin_amdgcn_alignbit(c, b, 1);   // x+1
        }
}

// Step D (0..5 = U, N, E, D, S, W) over the first n entries: src -> dst.  U
// also records the protection mask (the voxels no direction's rule deletes at
// the start of the iteration) of the first `nfull` entries, those the later
// steps evaluate (level >= 2); the other steps keep those.
template <int D, int NT>
__device__ __forceinline__ void region_step(const uint32_t *src, uint32_t *dst,
                                            const uint16_t *ent, uint32_t *prot, int n, int nfull)
{
    for (int e = threadIdx.x; e < n; e += NT) {
        const int r = ent[e] + 1;  // LDS arrays carry one pad word in front
        uint32_t nb[27];
        gather27_lds(src, r, nb);
        const uint32_t v = nb[13];
        uint32_t del;
        if (D == 0) {
            const uint32_t cu = rule3<0>(nb);
            if (e < nfull) {
                const uint32_t any = cu | rule3<1>(nb) | rule3<2>(nb) | rule3<3>(nb) | rule3<4>(nb) |
                                     rule3<5>(nb);
                prot[e] = v & ~any;
            }
            del = cu & v;
        } else {
            del = rule3<D>(nb) & v & ~prot[e];
        }
        dst[r] = v ^ del;
    }
}

// Iteration t (see tile3d_iteration).  Workgroup 0 first does the
// bookkeeping; then the workgroups take the queued tiles k = blockIdx.x,
// blockIdx.x + gridDim.x, ...
//
// Region blocks.  The region's words fall into 4 x 4 x 4 activity blocks:
// local block (lbx, lby, lbz) = (lx, (ly + 2) / 8, (lz + 2) / 8) is global
// block (2 tx - 1 + lbx, 2 ty - 1 + lby, 2 tz - 1 + lbz); the interior is
// local blocks 1..2 on every axis.  Region block (lbx, lby, lbz) belongs to
// tile (tx, ty, tz) + floor((lb - 1) / 2) on each axis.
template <int NT>
__global__ __launch_bounds__(NT) void tile3d_iter(Tile3D e, const uint32_t *__restrict__ R,
                                                  uint32_t *__restrict__ W, int t)
{
    constexpr int NW = NT / 64;
    constexpr int ROUNDS = (NR + NT - 1) / NT;
    constexpr int WROUNDS = (NINT + NT - 1) / NT;
    constexpr int NL1 = RX * rows_at_least(1);  // positions of level >= 1
    static_assert(NT >= 64 && ROUNDS <= 64, "one thread per region block, one lane per round");
    __shared__ uint32_t bufA[NR + 2], bufB[NR + 2];
    __shared__ uint16_t ent[NL1];
    __shared__ uint32_t prot[NL1];
    __shared__ int s_wc[ROUNDS][NW];
    __shared__ int s_cge[kHalo + 1];
    __shared__ uint8_t s_act[64];   // region blocks active in this iteration
    __shared__ uint8_t s_qt[27];    // neighbour tiles to queue
    __shared__ unsigned s_dirty;    // interior blocks that changed (bit per block)
    __shared__ int s_any;           // interior non-empty at the start

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    const int nitems = e.nitems;

    if (blockIdx.x == 0) {
        // items whose iteration t - 1 deleted nothing have converged: t
        // iterations, the last one changing nothing
        if (t >= 1) {
            const int *chp = e.changed + (size_t)((t + 2) % 3) * nitems;
            for (int i = tid; i < nitems; i += NT) {
                int *st = e.states + (size_t)i * ST_INTS;
                if (!st[ST_DONE] && !chp[i]) {
                    st[ST_DONE] = 1;
                    st[ST_ITERS] = t;
                    st[ST_PARITY] = 0;
                    atomicSub(&e.ctl[TC_LIVE], 1);
                }
            }
        }
        // slot (t + 1) % 3 was last read by this bookkeeping in iteration t - 1,
        // list (t + 2) % 3 last by iteration t - 1's tiles
        int *chn = e.changed + (size_t)((t + 1) % 3) * nitems;
        for (int i = tid; i < nitems; i += NT) chn[i] = 0;
        if (tid == 0) e.ctl[TC_COUNT0 + (t + 2) % 3] = 0;
    }

    const int n = e.ctl[TC_COUNT0 + t % 3];
    if ((int)blockIdx.x >= n) return;
    const int *list = e.list[t % 3];
    int *nlist = e.list[(t + 1) % 3];
    int *ncnt = &e.ctl[TC_COUNT0 + (t + 1) % 3];
    int *chg = e.changed + (size_t)(t % 3) * nitems;
    // block stamps read in this iteration, and those written for the next
    // (other workgroups still read this iteration's while stamps are written)
    const int *bact = e.bstamp + (size_t)(t & 1) * e.nblocks;
    int *bnext = e.bstamp + (size_t)((t + 1) & 1) * e.nblocks;
    uint32_t *const bufs[2] = {bufA, bufB};

    // This thread's region words are the same in every tile: staging word
    // r = q NT + tid in raster order (lx | ly << 2 | lz << 7, bit 31: past
    // the region) and compaction position p = q NT + tid in level order
    // (word | region block << 12 | level >= 1 << 18).
    uint32_t geo[ROUNDS], pos_w[ROUNDS];
#pragma unroll
    for (int q = 0; q < ROUNDS; ++q) {
        const int r = q * NT + tid;
        const int lx = r % RX, row = r / RX, ly = row % RY, lz = row / RY;
        geo[q] = r < NR ? (uint32_t)(lx | ly << 2 | lz << 7) : 0x80000000u;
        const int p = r;
        uint32_t pw = 0u;
        if (p < NL1) {
            const int prow = kRowOrder.row[p / RX], px = p % RX;
            const int ply = prow % RY, plz = prow / RY;
            pw = (uint32_t)(prow * RX + px) | (uint32_t)((((plz + 2) >> 3) * 4 + ((ply + 2) >> 3)) * 4 + px) << 12 |
                 1u << 18;
        }
        pos_w[q] = pw;
    }
    // interior words written back: lx | ly << 2 | lz << 7
    uint32_t igeo[WROUNDS];
#pragma unroll
    for (int q = 0; q < WROUNDS; ++q) {
        const int i = q * NT + tid;
        const int row = kRowOrder.row[(i < NINT ? i : 0) / kTileWordsX];
        igeo[q] = i < NINT ? (uint32_t)((1 + i % kTileWordsX) | (row % RY) << 2 | (row / RY) << 7)
                           : 0x80000000u;
    }

    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int tile = list[k];
        int lo = 0, hi = nitems - 1;  // the item holding the tile: the last with tbase <= tile
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (e.items[mid].tbase <= tile) lo = mid;
            else hi = mid - 1;
        }
        const int item = lo;
        const TileItem3D it = e.items[item];
        const int ltile = tile - it.tbase, tx = ltile % it.ntx, tyz = ltile / it.ntx;
        const int ty = tyz % it.nty, tz = tyz / it.nty;
        const int gx0 = tx * kTileWordsX - 1, y0 = ty * kTileLines - kHalo,
                  z0 = tz * kTilePlanes - kHalo;
        // word (gx0, y0, z0) of the padded layout; region word (lx, ly, lz)
        // is at org + lz plane + ly P + lx (never formed for words outside)
        const int org = (int)it.base + ((z0 + 1) * (it.h + 2) + (y0 + 1)) * it.P + 1 + gx0;

        // ---- stage the region (words outside the volume read as zero) ----
        uint32_t v[ROUNDS];
#pragma unroll
        for (int q = 0; q < ROUNDS; ++q) {
            const int lx = geo[q] & 3, ly = (geo[q] >> 2) & 31, lz = (geo[q] >> 7) & 31;
            v[q] = 0u;
            if (!(geo[q] >> 31) && (unsigned)(gx0 + lx) < (unsigned)it.nh &&
                (unsigned)(y0 + ly) < (unsigned)it.h && (unsigned)(z0 + lz) < (unsigned)it.d)
                v[q] = R[org + lz * it.plane + ly * it.P + lx];
        }
        if (tid < 64) {
            const int lbx = tid & 3, lby = (tid >> 2) & 3, lbz = tid >> 4;
            const int bx = gx0 + lbx, by = 2 * ty - 1 + lby, bz = 2 * tz - 1 + lbz;
            bool a = false;
            if (bx >= 0 && bx < it.nh && by >= 0 && by < it.nby && bz >= 0 && bz < it.nbz)
                a = t == 0 || bact[it.bbase + (bz * it.nby + by) * it.nh + bx] == t + 1;
            s_act[tid] = a;
        }
        if (tid < 27) s_qt[tid] = 0;
#pragma unroll
        for (int q = 0; q < ROUNDS; ++q) {
            const int r = q * NT + tid;
            if (r < NR) {
                bufA[r + 1] = v[q];
                bufB[r + 1] = v[q];
            }
        }
        if (tid == 0) {
            bufA[0] = bufB[0] = 0u;
            bufA[NR + 1] = bufB[NR + 1] = 0u;
            s_dirty = 0u;
            s_any = 0;
        }
        __syncthreads();

        // ---- the non-empty words of active blocks at level >= 1, in level
        // order: one ballot per round, each wave's base from the counts -----
        uint32_t fl = 0u;  // bit q: this thread's position of round q is an entry
        uint64_t m[ROUNDS];
#pragma unroll
        for (int q = 0; q < ROUNDS; ++q) {
            const uint32_t pw = pos_w[q];
            const bool f = (pw >> 18) && bufA[(pw & 0xfffu) + 1] != 0u && s_act[(pw >> 12) & 63u];
            fl |= (uint32_t)f << q;
            m[q] = __ballot(f);
            if (lane == 0) s_wc[q][wv] = __popcll(m[q]);
        }
        __syncthreads();
        // lane q of every wave sums round q's counts; an exclusive scan over
        // the lanes gives each round's base, then the waves before this one
        int tot = 0, mine = 0;
        if (lane < ROUNDS) {
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const int c = s_wc[lane][w];
                tot += c;
                mine += w < wv ? c : 0;
            }
        }
        int scan = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(scan, o, 64);
            if (lane >= o) scan += u;
        }
        const int rbase = scan - tot + mine;  // lane q: first entry of this wave in round q
#pragma unroll
        for (int q = 0; q < ROUNDS; ++q) {
            const int base = __shfl(rbase, q, 64);
            const int pos = base + __popcll(m[q] & lt);
            if ((fl >> q) & 1u) ent[pos] = (uint16_t)(pos_w[q] & 0xfffu);
            // entries of level >= j end where the rows of level < j begin
            const int p = q * NT + tid;
#pragma unroll
            for (int j = 1; j <= kHalo; ++j)
                if (p == RX * rows_at_least(j)) s_cge[j] = pos;
        }
        __syncthreads();

        // ---- the iteration: P and U, then N, E, D, S, W -------------------
        region_step<0, NT>(bufs[0], bufs[1], ent, prot, s_cge[1], s_cge[2]);
        __syncthreads();
        region_step<1, NT>(bufs[1], bufs[0], ent, prot, s_cge[2], 0);
        __syncthreads();
        region_step<2, NT>(bufs[0], bufs[1], ent, prot, s_cge[3], 0);
        __syncthreads();
        region_step<3, NT>(bufs[1], bufs[0], ent, prot, s_cge[4], 0);
        __syncthreads();
        region_step<4, NT>(bufs[0], bufs[1], ent, prot, s_cge[5], 0);
        __syncthreads();
        region_step<5, NT>(bufs[1], bufs[0], ent, prot, s_cge[6], 0);
        __syncthreads();

        // ---- write the interior: every word non-zero in either array ------
        int wg[WROUNDS];
        uint32_t wr[WROUNDS], wo[WROUNDS];
        bool wok[WROUNDS];
#pragma unroll
        for (int q = 0; q < WROUNDS; ++q) {
            const int lx = igeo[q] & 3, ly = (igeo[q] >> 2) & 31, lz = (igeo[q] >> 7) & 31;
            wok[q] = !(igeo[q] >> 31) && y0 + ly < it.h && z0 + lz < it.d;
            wg[q] = org + lz * it.plane + ly * it.P + lx;
            wr[q] = wok[q] ? R[wg[q]] : 0u;
            wo[q] = wok[q] ? W[wg[q]] : 0u;
        }
        unsigned dirty = 0u;
        bool any = false;
#pragma unroll
        for (int q = 0; q < WROUNDS; ++q) {
            if (!wok[q]) continue;
            const int lx = igeo[q] & 3, ly = (igeo[q] >> 2) & 31, lz = (igeo[q] >> 7) & 31;
            const uint32_t nv = bufA[(lz * RY + ly) * RX + lx + 1];
            if (wr[q] | wo[q]) W[wg[q]] = nv;
            any |= wr[q] != 0u;
            if (nv != wr[q]) dirty |= 1u << (((((lz + 2) >> 3) - 1) * 2 + (((ly + 2) >> 3) - 1)) * 2 + lx - 1);
        }
        const uint64_t anyw = __ballot(any);
        if (dirty) atomicOr(&s_dirty, dirty);
        if (lane == 0 && anyw) s_any = 1;
        __syncthreads();
        const unsigned dm = s_dirty;
        if (tid == 0) {
            // empty at the start: this pass cleared the other array's words,
            // so both are empty and stay so
            e.occ[tile] = s_any != 0;
            if (dm) chg[item] = 1;
        }
        if (dm) {
            // activate the region blocks within one block of a changed
            // interior block, and mark their tiles
            if (tid < 64) {
                const int lbx = tid & 3, lby = (tid >> 2) & 3, lbz = tid >> 4;
                bool a = false;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const int dx = (b & 1) + 1, dy = ((b >> 1) & 1) + 1, dz = (b >> 2) + 1;
                    a |= ((dm >> b) & 1u) && abs(lbx - dx) <= 1 && abs(lby - dy) <= 1 && abs(lbz - dz) <= 1;
                }
                const int bx = gx0 + lbx, by = 2 * ty - 1 + lby, bz = 2 * tz - 1 + lbz;
                if (a && bx >= 0 && bx < it.nh && by >= 0 && by < it.nby && bz >= 0 && bz < it.nbz) {
                    atomicMax(&bnext[it.bbase + (bz * it.nby + by) * it.nh + bx], t + 2);
                    s_qt[(((lbz + 1) >> 1) * 3 + ((lby + 1) >> 1)) * 3 + ((lbx + 1) >> 1)] = 1;
                }
            }
            __syncthreads();
            // queue the marked tiles that hold voxels for iteration t + 1
            if (tid < 27 && s_qt[tid]) {
                const int nx = tx + tid % 3 - 1, ny = ty + (tid / 3) % 3 - 1, nz = tz + tid / 9 - 1;
                const int nt = it.tbase + (nz * it.nty + ny) * it.ntx + nx;
                if ((tid == 13 || e.occ[nt]) && atomicMax(&e.stamp[nt], t + 2) < t + 2)
                    nlist[atomicAdd(ncnt, 1)] = nt;
            }
        }
        __syncthreads();  // LDS is reused by the next tile
    }
}

__global__ void tile3d_init(Tile3D e)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < (long)e.nitems * ST_INTS) e.states[g] = 0;
    if (g < 3L * e.nitems) e.changed[g] = 0;
    if (g < TC_INTS) e.ctl[g] = g == TC_LIVE ? e.nitems : 0;
    if (g < e.ntiles) {
        e.occ[g] = 0;
        e.stamp[g] = 0;
    }
    if (g < 2L * e.nblocks) e.bstamp[g] = 0;
}

__global__ void tile3d_zero(uint32_t *a, uint32_t *b, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        a[t] = 0u;
        b[t] = 0u;
    }
}

// Each thread packs kPackRuns3D runs of 64 voxels (two data words, one tile
// wide) into both state arrays and marks the runs' tiles occupied.
__global__ __launch_bounds__(256) void tile3d_pack(const SItem3D *__restrict__ items,
                                                   const TileItem3D *__restrict__ titems,
                                                   const SIO *__restrict__ io,
                                                   const int *__restrict__ blk, uint32_t *s0,
                                                   uint32_t *s1, uint8_t *occ)
{
    const int i = blk ? blk[blockIdx.x] : 0;
    const SItem3D it = items[i];
    const TileItem3D ti = titems[i];
    const SIO f = io[i];
    const uint32_t nq = (uint32_t)it.nh >> 1;
    const uint32_t nruns = (uint32_t)it.h * (uint32_t)it.d * nq;
    __shared__ uint32_t s_masks[4][320];
#pragma unroll
    for (int k = 0; k < kPackRuns3D; ++k) {
        const uint32_t t = ((uint32_t)(blockIdx.x - f.blk0) * kPackRuns3D + k) * 256u + threadIdx.x;
        const bool valid = t < nruns;
        const uint32_t line = valid ? t / nq : 0;
        const uint32_t q = valid ? t - line * nq : 0;
        const int x0 = (int)q * 64;  // < w
        const uint64_t v = pack_runs_wave_u8(f.in + (size_t)line * it.w + x0, min(64, it.w - x0),
                                             valid, s_masks[threadIdx.x >> 6]);
        if (!valid) continue;
        const uint32_t y = line % (uint32_t)it.h, z = line / (uint32_t)it.h;
        const uint32_t g = (uint32_t)it.base + ((z + 1) * (uint32_t)(it.h + 2) + y + 1) * (uint32_t)it.P +
                           1 + q * 2;
        s0[g] = (uint32_t)v;
        s0[g + 1] = (uint32_t)(v >> 32);
        s1[g] = (uint32_t)v;
        s1[g + 1] = (uint32_t)(v >> 32);
        if (v)
            occ[ti.tbase + ((int)(z / kTilePlanes) * ti.nty + (int)(y / kTileLines)) * ti.ntx + (int)q] = 1;
    }
}

// iteration 0's queue: every occupied tile
__global__ void tile3d_build(Tile3D e)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= e.ntiles || !e.occ[t]) return;
    e.stamp[t] = 1;
    e.list[0][atomicAdd(&e.ctl[TC_COUNT0], 1)] = t;
}

constexpr int kIterThreads = 256;

// ---- taking over a run of the word-list engine ------------------------------

// item states; items still live changed in iteration k - 1 (the word engine
// closes an item in the iteration that changes nothing)
__global__ void tile3d_take_states(Tile3D e, const int *__restrict__ wstates, int k)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e.nitems) return;
    const int *ws = wstates + (size_t)i * ST_INTS;
    int *st = e.states + (size_t)i * ST_INTS;
    st[ST_DONE] = ws[ST_DONE];
    st[ST_ITERS] = ws[ST_ITERS];
    st[ST_PARITY] = 0;
    if (ws[ST_DONE]) atomicSub(&e.ctl[TC_LIVE], 1);
    else e.changed[(size_t)((k + 2) % 3) * e.nitems + i] = 1;
}

__device__ __forceinline__ int item_of_block(const Tile3D &e, int b)
{
    int lo = 0, hi = e.nitems - 1;  // the last item with bbase <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (e.items[mid].bbase <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// the tile holding activity block b (1 word x 8 lines x 8 planes) of item it
__device__ __forceinline__ int tile_of_block(const TileItem3D &it, int b)
{
    const int l = b - it.bbase, bx = l % it.nh, byz = l / it.nh;
    const int by = byz % it.nby, bz = byz / it.nby;
    return it.tbase + ((bz / (kTilePlanes / 8)) * it.nty + by / (kTileLines / 8)) * it.ntx +
           bx / kTileWordsX;
}

// tile occupancy from the word engine's non-empty words
__global__ void tile3d_take_occ(Tile3D e, const int *__restrict__ blk, const int *__restrict__ count)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *count) return;
    const int b = blk[k];
    e.occ[tile_of_block(e.items[item_of_block(e, b)], b)] = 1;
}

// the word engine's blocks active in iteration k, and the occupied tiles
// holding them as iteration k's queue
__global__ void tile3d_take_blocks(Tile3D e, const int *__restrict__ wbstamp, int k)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= e.nblocks || wbstamp[b] != k) return;
    e.bstamp[(size_t)(k & 1) * e.nblocks + b] = k + 1;
    const int tile = tile_of_block(e.items[item_of_block(e, b)], b);
    if (e.occ[tile] && atomicMax(&e.stamp[tile], k + 1) < k + 1)
        e.list[k % 3][atomicAdd(&e.ctl[TC_COUNT0 + k % 3], 1)] = tile;
}

}  // namespace

long tile3d_items(const Layout3D &lay, std::vector<TileItem3D> &out)
{
    out.assign(lay.items.size(), TileItem3D{});
    long tiles = 0;
    for (size_t i = 0; i < lay.items.size(); ++i) {
        const SItem3D &s = lay.items[i];
        TileItem3D &t = out[i];
        t.base = s.base;
        t.w = s.w; t.h = s.h; t.d = s.d; t.nh = s.nh;
        t.P = s.P; t.plane = s.plane;
        t.ntx = s.nh / kTileWordsX;
        t.nty = (s.h + kTileLines - 1) / kTileLines;
        t.ntz = (s.d + kTilePlanes - 1) / kTilePlanes;
        t.tbase = (int)std::min<long>(tiles, 0x7fffffffL);
        tiles += (long)t.ntx * t.nty * t.ntz;
        t.nby = s.nby;
        t.nbz = s.nbz;
        t.bbase = s.bbase;
    }
    return tiles;
}

size_t tile3d_carve(Tile3D &e, char *base, int nitems, long nwords, long ntiles, long nblocks,
                    uint32_t *const *shared_state)
{
    size_t used = 0;
    auto take = [&](size_t bytes) {
        char *r = base ? base + used : nullptr;
        used += round_up(bytes ? bytes : 1, 256);
        return r;
    };
    e.nitems = nitems;
    for (int k = 0; k < 2; ++k)
        e.state[k] = shared_state ? shared_state[k] : (uint32_t *)take(sizeof(uint32_t) * nwords);
    e.nwords = nwords;
    e.ntiles = (int)ntiles;
    e.nblocks = (int)nblocks;
    e.bstamp = (int *)take(sizeof(int) * 2 * nblocks);
    e.occ = (uint8_t *)take((size_t)ntiles);
    e.stamp = (int *)take(sizeof(int) * ntiles);
    for (int k = 0; k < 3; ++k) e.list[k] = (int *)take(sizeof(int) * ntiles);
    e.changed = (int *)take(sizeof(int) * 3 * (size_t)nitems);
    e.states = (int *)take(sizeof(int) * ST_INTS * (size_t)nitems);
    e.ctl = (int *)take(sizeof(int) * TC_INTS);
    return used;
}

int tile3d_grid(int blocks_per_cu)
{
    static const int2 dev_cus_occ = [] {
        int dev = 0, c = 256, per_cu = 2;
        if (hipGetDevice(&dev) == hipSuccess) {
            (void)hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu, (const void *)tile3d_iter<kIterThreads>, kIterThreads, 0) != hipSuccess ||
                per_cu < 1)
                per_cu = 2;
        }
        return int2{std::max(1, c), std::max(1, per_cu)};
    }();
    return dev_cus_occ.x * (blocks_per_cu > 0 ? std::min(blocks_per_cu, dev_cus_occ.y) : dev_cus_occ.y);
}

int tile3d_begin(const Tile3D &e, bool zero, hipStream_t s)
{
    const long ninit = std::max<long>({(long)e.nitems * ST_INTS, 3L * e.nitems, (long)TC_INTS,
                                       (long)e.ntiles, 2L * e.nblocks, 1L});
    hipLaunchKernelGGL(tile3d_init, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, s, e);
    if (zero && e.nwords > 0)
        hipLaunchKernelGGL(tile3d_zero, dim3((unsigned)((e.nwords + 255) / 256)), dim3(256), 0, s,
                           e.state[0], e.state[1], e.nwords);
    if (e.nblk > 0)
        hipLaunchKernelGGL(tile3d_pack, dim3(e.nblk), dim3(256), 0, s, e.sitems, e.items, e.io, e.blk,
                           e.state[0], e.state[1], e.occ);
    if (e.ntiles > 0)
        hipLaunchKernelGGL(tile3d_build, dim3((unsigned)((e.ntiles + 255) / 256)), dim3(256), 0, s, e);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

int tile3d_iteration(const Tile3D &e, int t, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(tile3d_iter<kIterThreads>, dim3(std::max(1, grid)), dim3(kIterThreads), 0, s, e,
                       e.state[t & 1], e.state[(t + 1) & 1], t);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

int tile3d_from_words(const Tile3D &e, const Sparse3D &w, int k, int bound, hipStream_t s)
{
    if (k < 1 || (k & 1)) return fail(BGT_ERR_INVALID, "tile engine takes over after an even iteration");
    if (e.state[0] != w.state[0] || e.state[1] != w.state[1] || e.nitems != w.nitems ||
        e.nblocks != w.nblocks)
        return fail(BGT_ERR_INVALID, "tile engine does not share the word engine's layout");
    const long ninit = std::max<long>({(long)e.nitems * ST_INTS, 3L * e.nitems, (long)TC_INTS,
                                       (long)e.ntiles, 2L * e.nblocks, 1L});
    hipLaunchKernelGGL(tile3d_init, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, s, e);
    hipLaunchKernelGGL(tile3d_take_states, dim3((unsigned)((e.nitems + 255) / 256)), dim3(256), 0, s, e,
                       w.states, k);
    if (bound > 0)
        hipLaunchKernelGGL(tile3d_take_occ, dim3((unsigned)((bound + 255) / 256)), dim3(256), 0, s, e,
                           w.list[0].blk, w.list[0].count);
    if (e.nblocks > 0)
        hipLaunchKernelGGL(tile3d_take_blocks, dim3((unsigned)((e.nblocks + 255) / 256)), dim3(256), 0,
                           s, e, w.bstamp, k);
    BGT_HIP(hipGetLastError());
    return BGT_OK;
}

int tile3d_end(const Tile3D &e, hipStream_t s)
{
    Sparse3D v;  // the word-list engine's unpack reads array 0 through the same layout
    v.nitems = e.nitems;
    v.items = e.sitems;
    v.io = e.io;
    v.ublk = e.ublk;
    v.nublk = e.nublk;
    v.state[0] = e.state[0];
    return sparse3d_end(v, s);
}

}  // namespace bgt
