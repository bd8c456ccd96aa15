// batch.hip -- many images and volumes of uneven size thinned together.
//
// All items of a batch live in one device allocation.  2D items and 3D items
// form two independent pipelines on two streams, each fed by its own host
// thread; chunks of launches can be replayed as captured hipGraphs.
//
// 2D: each pass runs from a work queue of active bands.  A band is in the
// queue of pass P+1 iff it changed in pass P, or it is non-empty and a
// neighbouring band changed in pass P; empty bands are never queued.  Bands
// never queued for a pass hold identical data in both ping-pong buffers
// (packing writes both), so skipping them is exact.  The queue kernel runs a
// fixed, co-resident grid: workgroup b takes queued band b and the bands past
// the grid come from an atomic counter (or, with few bands queued, each band
// is split over several workgroups).  Its last workgroup does the pass's
// bookkeeping: it closes the items of pass P-2 (converged when one iteration
// of a pass deleted nothing), clears the idle queue counters and reports the
// pass's queued-band count to the host through host-mapped memory.  No launch
// runs between passes.
//
// 3D: the non-empty-word list engine (sparse3d.hip) over all volumes at once;
// the host polls its control words once per chunk of iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "batch.h"
#include "common.h"
#include "thin2d.h"
#include "thin2d_kernels.h"
#include "thin3d.h"
#include "thin3d_kernels.h"
#include "pack.h"
#include "queue2d.h"
#include "sparse3d.h"
#include "tile3d.h"

namespace bgt {

// Caller buffers of one item and the first block of its pack/unpack grid.
// The pack/unpack kernels give every item whole 256-thread blocks; a table
// maps each block to its item.
struct IO2D {
    const uint8_t *in;
    size_t in_pitch;
    uint8_t *out;
    size_t out_pitch;
    int blk0;           // first block of the item in the pack grid
    int ublk0;          // first block of the item in the unpack grid
    uint32_t uchunks;   // unpack chunks per segment
    int flat;           // rows contiguous: the whole image is one segment
};

namespace {

// One thread per 64-bit row word; writes both ping-pong buffers and marks the
// band of every non-empty word occupied.
__global__ __launch_bounds__(256) void batch_pack2d(const Item2D *__restrict__ items,
                                                    const IO2D *__restrict__ io,
                                                    const int *__restrict__ blk_item, uint8_t *occ)
{
    __shared__ uint32_t s_masks[4][320];
    const int i = blk_item[blockIdx.x];
    const Item2D it = items[i];
    const IO2D f = io[i];
    const int t = (blockIdx.x - f.blk0) * 256 + threadIdx.x;
    const bool valid = t < it.h * it.L;
    const int y = valid ? t / it.L : 0, x0 = valid ? (t - y * it.L) * 64 : 0;  // x0 < w
    const int n = min(64, it.w - x0);
    const uint8_t *p = f.in + (size_t)y * f.in_pitch + x0;
    uint64_t v;
    if (f.in_pitch == (size_t)it.w)  // contiguous rows (the same for the whole workgroup)
        v = pack_runs_wave_u8(p, n, valid, s_masks[threadIdx.x >> 6]);
    else
        v = valid ? pack_run_u8(p, n) : 0ull;
    if (!valid) return;
    v = lane_order(v);
    it.buf[0][t] = v;
    it.buf[1][t] = v;
    if (v) occ[it.tile0 + y / it.band_rows] = 1;
}

// One thread per 16-byte chunk of the caller's image (or of one of its rows).
__global__ __launch_bounds__(256) void batch_unpack2d(const Item2D *__restrict__ items,
                                                      const IO2D *__restrict__ io,
                                                      const int *__restrict__ ublk)
{
    const int i = ublk[blockIdx.x];
    const Item2D it = items[i];
    const IO2D f = io[i];
    const uint32_t t = (uint32_t)(blockIdx.x - f.ublk0) * 256u + threadIdx.x;
    const uint64_t *buf = it.st[ST_PARITY] ? it.buf[1] : it.buf[0];
    unpack_image_u8(buf, (uint32_t)it.L, it.w, it.h, f.out, f.out_pitch, f.uchunks, f.flat, t);
}

// zero states, iteration masks, tile flags and queue control words
__global__ void batch_init(int *states, long nstate, unsigned *hist, long nhist, uint8_t *occ,
                           int *stamp, long ntiles, int *ctl, int nctl)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nstate) states[g] = 0;
    if (g < nhist) hist[g] = 0u;
    if (g < ntiles) {
        occ[g] = 0;
        stamp[g] = 0;
    }
    if (g < nctl) ctl[g] = 0;
}

// queue of pass 0 = every occupied tile
__global__ void batch_build_list(Queue q)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= q.ntiles || !q.occ[t]) return;
    const int k = atomicAdd(&q.cnt[0], 1);
    q.list[0][k] = t;
}

// queue of pass 0 = every tile (the passes pack the images themselves, so no
// occupancy is known before it)
__global__ void batch_list_all(Queue q)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < q.ntiles) q.list[0][t] = t;
    if (t == 0) q.cnt[0] = q.ntiles;
}

__device__ __forceinline__ void enqueue(const Queue &q, int u, int next_pass, int *nlist,
                                        int *ncnt, bool force)
{
    if (!force && !q.occ[u]) return;
    if (atomicMax(&q.stamp[u], next_pass) < next_pass) nlist[atomicAdd(ncnt, 1)] = u;
}

// close item i from its mask of pass P (the first pass whose mask misses an
// iteration is the one in which it converged)
__device__ __forceinline__ void close_item(const Close2D &c, int i, int K, int P)
{
    int *st = c.states + (size_t)i * ST_INTS;
    if (st[ST_DONE]) return;
    const int base = st[ST_BASE];  // the item's first pass
    if (P < base) return;
    const int m = __popc(c.hist[(size_t)(P & 3) * c.nitems + i]);
    if (m < K) {
        st[ST_ITERS] = (P - base) * K + m + 1;  // counting the iteration that deleted nothing
        st[ST_PARITY] = (P + 1) & 1;
        st[ST_DONE] = 1;
        if (c.closed)
            __hip_atomic_store(&c.closed[i], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}


// the item skips pass P when an iteration of its pass P - 1 deleted nothing
__device__ __forceinline__ bool converged_before(const Close2D &c, const unsigned *hprev, int i,
                                                 int K, int P)
{
    return P >= 1 + c.states[(size_t)i * ST_INTS + ST_BASE] && __popc(hprev[i]) < K;
}

// tail passes: the most workgroups a band is split over
constexpr int kParts = 16;

// One pass P of the 2D batch over the queued bands (the body of a pass
// launch, or of one pass of the persistent chain).  Workgroups
// 0 .. nwg - 1 take bands; workgroup nwg does the pass's bookkeeping (closes
// the items of pass P - 2, clears the idle queue counters and history slot)
// and takes no band.
//
// IO = 2 (V = 1 only): pass 0 of a batch whose items all start there, which
// packs the images from the caller's bytes (`io`) itself and queues every
// neighbour of a changed band (band occupancy is only being found in that
// pass): no pack launch runs.  A kernel of its own, so that the other passes
// carry none of that code.  IO = 0: `io` is null.
template <int R, int V, int IO>
__device__ __forceinline__ void queue_pass_body(const Item2D *__restrict__ items,
                                                const Tile2D *__restrict__ tiles, const Queue &q,
                                                const Close2D &c, int K, int P, int nwg,
                                                const ItemIO2D *__restrict__ io, uint64_t *edges,
                                                TileInfo &info)
{
    const int tid = threadIdx.x;
    if ((int)blockIdx.x == nwg) {
        if (P >= 2)
            for (int i = tid; i < c.nitems; i += blockDim.x) close_item(c, i, K, P - 2);
        unsigned *h = c.hist + (size_t)((P + 1) & 3) * c.nitems;
        for (int i = tid; i < c.nitems; i += blockDim.x) h[i] = 0u;
        if (tid == 0) {
            q.cnt[(P + 2) % 3] = 0;
            q.wctr[(P + 2) % 3] = 0;
            // the host follows the passes through these: a pass with no queued
            // band means every item converged
            if (c.report)
                __hip_atomic_store(&c.report[P & 15],
                                   ((unsigned long long)(P + 1) << 32) | (unsigned)q.cnt[P % 3],
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        return;
    }
    const int cur = P % 3;
    const int *list = q.list[cur];
    int *nlist = q.list[(P + 1) % 3];
    int *ncnt = &q.cnt[(P + 1) % 3];
    const int n = q.cnt[cur];
    const unsigned *hprev = c.hist + (size_t)((P + 3) & 3) * c.nitems;  // pass P - 1
    unsigned *hcur = c.hist + (size_t)(P & 3) * c.nitems;
    // Fewer bands than half the CUs (the tail of a batch): each band is split
    // over kp workgroups, at most one per CU, whose windows hold 4 or 8
    // register rows per segment.  A pass's latency is bound by the rows one
    // wave works through per subiteration, so it drops while the CUs would
    // otherwise idle; with fewer idle CUs the parts' halo rows cost more than
    // they save.  Each part carries full halos, so the band's result is the
    // same.
    const int kp = n > 0 ? min(min(kParts, c.ncu / n), nwg / n) : 0;
    if (kp >= 2) {
        const int e = blockIdx.x / kp, part = blockIdx.x % kp;
        if (e >= n) return;
        const int tile = list[e];
        const Tile2D tl = tiles[tile];
        if (converged_before(c, hprev, tl.item, K, P)) return;
        const Item2D it = items[tl.item];
        const ItemIO2D *fio = io ? io + tl.item : nullptr;
        const int row0 = tl.band * it.band_rows;
        const int seg = (int)(blockDim.x >> 6) * (64 / lanes_for(it.L, V));  // segments per window
        const int in4 = seg * 4 - 8 * K, in8 = seg * 8 - 8 * K;  // interior rows of a part
        const int np4 = in4 >= 1 ? (it.band_rows + in4 - 1) / in4 : kParts + 1;
        const int np8 = in8 >= 1 ? (it.band_rows + in8 - 1) / in8 : kParts + 1;
        if (np4 <= kp) {
            if (part >= np4 || row0 + part * in4 >= it.h) return;
            band_work<4, V, IO>(it, row0 + part * in4, min(in4, it.band_rows - part * in4), false,
                                   K, P & 1, edges, &info, 0x7fffffff, fio);
        } else if (R > 8 && np8 <= kp) {
            if (part >= np8 || row0 + part * in8 >= it.h) return;
            band_work<8, V, IO>(it, row0 + part * in8, min(in8, it.band_rows - part * in8), false,
                                   K, P & 1, edges, &info, 0x7fffffff, fio);
        } else {  // windows too short for parts: one workgroup per band
            if (part != 0) return;
            band_work<R, V, IO>(it, row0, it.band_rows, false, K, P & 1, edges, &info, 0x7fffffff,
                                   fio);
        }
        if (tid == 0) {
            // parts only set the flag: a band emptied in a tail pass stays
            // marked, which can only cost a pass over an empty band
            if (info.occ) q.occ[tile] = 1;
            if (info.mask) atomicOr(&hcur[tl.item], info.mask);
        }
        // a part that changed queues its band and the band's neighbours
        if (info.mask && tid < 3) {
            const int b = tl.band - 1 + tid;
            if (b >= 0 && b < it.nbands)
                enqueue(q, it.tile0 + b, P + 1, nlist, ncnt, tid == 1 || IO == 2);
        }
        return;
    }
    // Workgroup b takes band b first, and only the bands past the grid go
    // through the work counter: same-address atomics, one per grab, cost about
    // as much as a short pass when every workgroup makes one.  Workgroups past
    // the queue leave at once.
    if ((int)blockIdx.x >= n) return;
    for (int k = 0;; ++k) {
        int t = blockIdx.x;
        if (k > 0) {
            if (n <= nwg) break;
            if (tid == 0) info.t = nwg + atomicAdd(&q.wctr[cur], 1);
            __syncthreads();
            t = info.t;
            __syncthreads();
            if (t >= n) break;
        }
        const int tile = list[t];
        const Tile2D tl = tiles[tile];
        // converged in pass P - 1: an iteration of that pass deleted nothing
        if (converged_before(c, hprev, tl.item, K, P)) continue;
        const Item2D it = items[tl.item];
        // a single-band item also runs at most K iterations per pass, so that no
        // pass waits on one item iterating to convergence while CUs idle
        band_work<R, V, IO>(it, tl.band * it.band_rows, it.band_rows, it.nbands == 1, K, P & 1,
                               edges, &info, K, io ? io + tl.item : nullptr);
        if (tid == 0) {
            q.occ[tile] = info.occ != 0;
            if (info.mask) atomicOr(&hcur[tl.item], info.mask);
            // all K iterations deleted something: not converged, go on next pass
            if (it.nbands == 1 && info.mask == (1u << K) - 1u)
                enqueue(q, tile, P + 1, nlist, ncnt, true);
        }
        if (info.mask && it.nbands > 1 && tid < 3) {
            const int b = tl.band - 1 + tid;
            if (b >= 0 && b < it.nbands)
                enqueue(q, it.tile0 + b, P + 1, nlist, ncnt, tid == 1 || IO == 2);
        }
        __syncthreads();
    }
}

// One pass P = (*pbase if pbase else 0) + p: gridDim.x - 1 band workgroups
// and the bookkeeping one.
template <int R, int V, int IO>
__global__ __launch_bounds__(512) void thin2d_queue_pass(const Item2D *__restrict__ items,
                                                          const Tile2D *__restrict__ tiles,
                                                          Queue q, Close2D c, int K,
                                                          const int *pbase, int p,
                                                          const ItemIO2D *__restrict__ io)
{
    extern __shared__ uint64_t edges[];
    __shared__ TileInfo info;
    queue_pass_body<R, V, IO>(items, tiles, q, c, K, (pbase ? *pbase : 0) + p, (int)gridDim.x - 1,
                              io, edges, info);
}

// after the last pass Pend - 1: close the items of passes Pend - 2 and Pend - 1
// (Pend from *dpend when given: the persistent chain counts its passes)
__global__ void batch_close2d(Close2D c, int K, int pend, const int *dpend)
{
    if (dpend) pend = *dpend;
    for (int P = max(0, pend - 2); P < pend; ++P)
        for (int i = threadIdx.x; i < c.nitems; i += blockDim.x) close_item(c, i, K, P);
}

// Device-wide barrier of the persistent 2D chain: each workgroup arrives once
// per pass on a monotonic counter and waits until it reaches `target`, with
// agent-scope release and acquire fences around it, so that the bands written
// in a pass are visible to every workgroup (and XCD) in the next.  A wait
// longer than kBarrierTicks of the constant-rate clock (or another
// workgroup's timeout) sets *err and returns false: every workgroup then
// leaves, and the run reports the error.
constexpr uint64_t kBarrierTicks = 100000000ull;  // about 1 s at the 100 MHz clock

__device__ __forceinline__ bool chain_barrier(unsigned *bar, unsigned target, int *err, int *s_ok)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(bar, 1u);
        const uint64_t t0 = wall_clock64();
        int ok = 1;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                wall_clock64() - t0 > kBarrierTicks) {
                atomicExch(err, 1);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

// The whole 2D chain in one launch: every pass of the band queue, each
// followed by the device-wide barrier, until a pass leaves the next queue
// empty (or after max_passes).  ctl[0] = barrier counter, ctl[1] = passes
// run, ctl[2] = barrier timeout.  The grid must be co-resident (one
// workgroup per slot the pass kernel's occupancy allows), plus the
// bookkeeping workgroup.
template <int R, int V>
__global__ __launch_bounds__(512) void thin2d_chain(const Item2D *__restrict__ items,
                                                     const Tile2D *__restrict__ tiles, Queue q,
                                                     Close2D c, int K, int max_passes, int *ctl)
{
    extern __shared__ uint64_t edges[];
    __shared__ TileInfo info;
    __shared__ int s_ok;
    const int nwg = (int)gridDim.x - 1;
    int P = 0;
    for (; P < max_passes; ++P) {
        queue_pass_body<R, V, 0>(items, tiles, q, c, K, P, nwg, nullptr, edges, info);
        if (!chain_barrier((unsigned *)&ctl[0], gridDim.x * (unsigned)(P + 1), &ctl[2], &s_ok)) return;
        // the queue of pass P + 1 is complete and stays so until pass P + 2
        if (__hip_atomic_load(&q.cnt[(P + 1) % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            ++P;
            break;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl[1] = P;
}

// graph replays: the next chunk's passes start `n` later
__global__ void batch_advance(int *pbase, int n)
{
    if (threadIdx.x == 0) *pbase += n;
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

// Workgroups of a work-queue launch: as many as fit on the device (at most
// `per_cu_cap` per CU when > 0), and no more than there are tiles.  Fewer
// than fit can be faster: the queue kernel is VALU-bound at two workgroups per
// CU, and the CU slots left free let a concurrent 3D chain run.
int occupancy_grid(const void *kernel, int block, size_t lds, int ntiles, int per_cu_cap)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return std::max(1, std::min(ntiles, 1024));
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (per_cu_cap > 0) per_cu = std::min(per_cu, per_cu_cap);
    if (cus < 1) cus = 256;
    return std::max(1, std::min(ntiles, cus * per_cu));
}

}  // namespace

struct Batch {
    Config2D c2;
    Config3D c3;
    struct In2 {
        IO2D io;
        int w, h;
    };
    struct In3 {
        SIO io;
        int w, h, d;
    };
    std::vector<In2> in2;
    std::vector<In3> in3;
    std::vector<int> order;  // the items in the order added: 2D index, or ~3D index

    bool ready = false;
    bool ran = false;  // a run completed since the layout was made
    void *mem = nullptr;
    size_t bytes = 0;
    bool owns_mem = true;       // false when `mem` is the caller's arena
    void *arena = nullptr;      // optional caller-owned device memory
    size_t arena_bytes = 0;

    // 2D pipeline: band work queue
    Item2D *d_items2 = nullptr;
    IO2D *d_io2 = nullptr;
    ItemIO2D *d_fio2 = nullptr;  // caller buffers for passes with fused pack/unpack
    int *d_blk2 = nullptr;   // pack block -> item
    int nblk2 = 0;
    int *d_ublk2 = nullptr;  // unpack block -> item
    int nublk2 = 0;
    Tile2D *d_tiles2 = nullptr;
    Queue q2{};
    Close2D close2{};
    int grid2 = 0;          // band workgroups of a pass (one more does the bookkeeping)
    int V2 = 1;             // words per lane of every 2D item's bands
    // pass 0 packs the images (Config2D::fused_io, one word per lane only;
    // wider batches keep the pack launch)
    bool fpack2 = false;
    int *d_states2 = nullptr;
    int *d_gpass2 = nullptr;  // pass number base of graph replays
    int pass2 = 0;            // 2D passes queued in this run
    hipGraphExec_t graph2 = nullptr;
    hipStream_t graph2_stream = nullptr;

    // 3D pipeline: the non-empty word list engine and/or the tile engine
    // (Config3D::engine)
    Tile3D t3;
    int titer3 = 0;        // tile engine: next iteration to queue
    bool on_tiles3 = false;  // this run's 3D chain is on the tile engine
    Sparse3D e3;
    bool zeroed3 = false;  // state margins cleared since allocation
    int bound3 = 0;        // upper bound of the current 3D list's count
    int abound3 = 0;       // upper bound of the current 3D active list's count
    int iter3 = 0;         // 3D iterations completed at the last poll
    int cur3 = 0;          // current 3D list
    std::vector<std::pair<long, hipGraphExec_t>> graphs3;  // by launch bounds
    hipStream_t graph3_stream = nullptr;

    // control words cleared by one init kernel
    int *d_ctl = nullptr;
    int nctl = 0;
    uint8_t *d_occ = nullptr;
    int *d_stamp = nullptr;
    long ntiles2 = 0;

    int *h_live = nullptr;  // pinned, two slots of 32: [8..] 3D control words; [64..66] 2D chain words
    int *dh_live = nullptr;  // its device address (kernels write the polled words)
    unsigned long long *h_rep = nullptr;  // pinned, device-mapped pass reports (Close2D)
    char *h_stage = nullptr;   // pinned staging of the tables uploaded by prepare
    size_t h_stage_bytes = 0;
    // outgrown pinned buffers, freed with the batch: hipHostFree waits for the
    // device, which stalls the other threads' work on it
    std::vector<void *> h_retired;
    int *h_states = nullptr;   // pinned copy of the item states after a run
    int *dh_states = nullptr;  // its device address
    size_t h_states_ints = 0;
    hipEvent_t ev_poll3[2] = {nullptr, nullptr};
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    // optional per-launch kernel timing (direct launches only)
    bool timing = false;
    std::mutex mark_mu;  // the two chains' host threads share the timing events
    std::vector<hipEvent_t> events;
    size_t ev_used = 0;
    struct Mark {
        size_t begin, end;  // event indices
        int kind;           // 0 = 2D pass, 1 = 3D iteration
    };
    std::vector<Mark> ev_marks;
    double ms2d = 0, ms3d = 0;
    int n2d = 0, n3d = 0;

    ~Batch()
    {
        release();
        if (side) (void)hipStreamDestroy(side);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        for (int k = 0; k < 2; ++k) {
            if (ev_poll3[k]) (void)hipEventDestroy(ev_poll3[k]);
        }
        if (h_live) (void)hipHostFree(h_live);
        if (h_rep) (void)hipHostFree(h_rep);
        if (h_stage) (void)hipHostFree(h_stage);
        if (h_states) (void)hipHostFree(h_states);
        for (void *p : h_retired) (void)hipHostFree(p);
    }
    void release()
    {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        events.clear();
        for (auto &g : graphs3) (void)hipGraphExecDestroy(g.second);
        graphs3.clear();
        if (graph2) (void)hipGraphExecDestroy(graph2);
        graph2 = nullptr;
        zeroed3 = false;
        if (mem && owns_mem) (void)hipFree(mem);
        mem = nullptr;
        ready = false;
        ran = false;
    }
};

// Workgroups per CU of the batch's 2D passes: the configured count, or when
// automatic two beside volumes (the 3D chain's kernels take the other slots)
// and as many as fit in a batch of images alone.
static int blocks2(const Batch &b)
{
    if (b.c2.blocks_per_cu >= 0) return b.c2.blocks_per_cu;
    return b.in3.empty() ? 0 : 2;
}

// Lays the batch out in device memory (its own allocation, or the caller's
// arena); with size_only, only computes b.bytes (the sizing sweep sets the
// device pointers from a null base, so a batch that is not ready is laid out
// again by its next run).
static int prepare(Batch &b, hipStream_t s, bool size_only = false)
{
    // host tables, uploaded together through pinned memory on the run stream
    struct Staged {
        void *dst;
        size_t off, n;
    };
    std::vector<Staged> staged;
    std::vector<char> host_stage;
    auto stage = [&](void *dst, const void *src, size_t n) {
        if (!n) return;
        staged.push_back({dst, host_stage.size(), n});
        host_stage.insert(host_stage.end(), (const char *)src, (const char *)src + n);
    };
    if (b.ready) return BGT_OK;  // b.bytes is that of the current layout
    if (b.in2.empty() && b.in3.empty()) return fail(BGT_ERR_INVALID, "empty batch");
    BGT_CHECK(check_config2d(b.c2));
    const int n2 = (int)b.in2.size(), n3 = (int)b.in3.size();
    // one pass kernel runs every item's bands, so all take the words per lane
    // of the widest
    b.V2 = 1;
    for (int i = 0; i < n2; ++i) {
        const int v = lane_words2d(b.in2[i].w, b.c2);
        if (!v) return fail(BGT_ERR_UNSUPPORTED, "2D width %d exceeds %d", b.in2[i].w, kMaxWidth2D);
        b.V2 = std::max(b.V2, v);
    }
    b.fpack2 = b.V2 == 1 && b.c2.fused_io != 0;
    std::vector<Plan2D> p2(n2);
    for (int i = 0; i < n2; ++i) BGT_CHECK(plan2d(b.in2[i].w, b.in2[i].h, b.c2, p2[i], b.V2));
    BGT_CHECK(check_config3d(b.c3));
    std::vector<int> w3v(n3), h3v(n3), d3v(n3);
    for (int i = 0; i < n3; ++i) {
        w3v[i] = b.in3[i].w;
        h3v[i] = b.in3[i].h;
        d3v[i] = b.in3[i].d;
    }
    Layout3D lay3;
    if (n3) BGT_CHECK(sparse3d_layout(w3v.data(), h3v.data(), d3v.data(), n3, lay3));
    const long nwords3 = lay3.nwords, data3 = lay3.data_words;
    // engine 0: word lists (one-launch tail); 1: word lists, then tiles; 2: tiles
    const bool words3 = n3 > 0 && b.c3.engine <= 1;
    const bool tiles3 = n3 > 0 && b.c3.engine >= 1;
    std::vector<TileItem3D> titems3;
    const long ntiles3 = tiles3 ? tile3d_items(lay3, titems3) : 0;
    if (ntiles3 >= 0x7fffffffL) return fail(BGT_ERR_UNSUPPORTED, "3D tiles exceed 2^31");

    std::vector<Tile2D> t2;
    std::vector<int> tile0(n2);
    for (int i = 0; i < n2; ++i) {
        tile0[i] = (int)t2.size();
        for (int k = 0; k < p2[i].nbands; ++k) t2.push_back({i, k});
    }
    b.ntiles2 = (long)t2.size();
    b.nctl = 16;
    // pack grid: whole blocks per item, one thread per 64 pixels; unpack
    // grid: one thread per 16-byte chunk of the image (rows contiguous) or of
    // each row (pitched)
    std::vector<int> blk2, blk0_2(n2), ublk2, ublk0_2(n2);
    std::vector<IO2D> io2(n2);
    for (int i = 0; i < n2; ++i) {
        blk0_2[i] = (int)blk2.size();
        const long nt = (long)p2[i].state_words;
        for (long k = 0; k < (nt + 255) / 256; ++k) blk2.push_back(i);
        IO2D &f = io2[i];
        f = b.in2[i].io;
        f.blk0 = blk0_2[i];
        long ut = 0;
        unpack_image_grid(b.in2[i].w, b.in2[i].h, f.out_pitch, f.flat, f.uchunks, ut);
        f.ublk0 = (int)ublk2.size();
        for (long k = 0; k < (ut + 255) / 256; ++k) ublk2.push_back(i);
    }
    b.nblk2 = (int)blk2.size();
    b.nublk2 = (int)ublk2.size();

    for (int sweep = 0; sweep < 2; ++sweep) {
        Carver cv{sweep ? (char *)b.mem : nullptr};
        b.d_ctl = cv.take<int>(sizeof(int) * b.nctl);
        // 2D
        b.d_items2 = cv.take<Item2D>(sizeof(Item2D) * n2);
        b.d_io2 = cv.take<IO2D>(sizeof(IO2D) * n2);
        b.d_fio2 = cv.take<ItemIO2D>(sizeof(ItemIO2D) * n2);
        b.d_blk2 = cv.take<int>(sizeof(int) * blk2.size());
        b.d_ublk2 = cv.take<int>(sizeof(int) * ublk2.size());
        b.d_tiles2 = cv.take<Tile2D>(sizeof(Tile2D) * t2.size());
        b.d_occ = cv.take<uint8_t>((size_t)b.ntiles2);
        b.d_stamp = cv.take<int>(sizeof(int) * b.ntiles2);
        int *lists2 = cv.take<int>(sizeof(int) * 3 * b.ntiles2);
        b.d_states2 = cv.take<int>(sizeof(int) * (size_t)n2 * ST_INTS);
        unsigned *hist2 = cv.take<unsigned>(sizeof(unsigned) * 4 * (size_t)n2);
        // 3D
        Sparse3D &e = b.e3;
        e = Sparse3D{};
        e.nitems = n3;
        SItem3D *d_items3 = cv.take<SItem3D>(sizeof(SItem3D) * n3);
        int2 *d_geo3 = cv.take<int2>(sizeof(int2) * n3);
        SIO *d_io3 = cv.take<SIO>(sizeof(SIO) * n3);
        int *d_blk3 = cv.take<int>(sizeof(int) * lay3.blk.size());
        int *d_ublk3 = cv.take<int>(sizeof(int) * lay3.ublk.size());
        if (words3) {
            e.items = d_items3;
            e.geo = d_geo3;
            e.io = d_io3;
            e.blk = d_blk3;
            e.nblk = (int)lay3.blk.size();
            e.ublk = d_ublk3;
            e.nublk = (int)lay3.ublk.size();
            e.state[0] = cv.take<uint32_t>(sizeof(uint32_t) * nwords3);
            e.state[1] = cv.take<uint32_t>(sizeof(uint32_t) * nwords3);
            e.nwords = nwords3;
            e.ctl = cv.take<int>(sizeof(int) * SC_INTS);
            for (int k = 0; k < 2; ++k) {
                e.list[k].item = cv.take<int>(sizeof(int) * data3);
                e.list[k].word = cv.take<uint32_t>(sizeof(uint32_t) * data3);
                e.list[k].blk = cv.take<int>(sizeof(int) * data3);
                e.list[k].count = e.ctl ? e.ctl + SC_COUNT0 + k : nullptr;
            }
            e.act.item = cv.take<int>(sizeof(int) * data3);
            e.act.word = cv.take<uint32_t>(sizeof(uint32_t) * data3);
            e.act.blk = cv.take<int>(sizeof(int) * data3);
            e.act.count = nullptr;  // in ctl, by iteration parity
            e.prot = cv.take<uint32_t>(sizeof(uint32_t) * data3);
            e.first = cv.take<uint32_t>(sizeof(uint32_t) * data3);
            e.capacity = (int)data3;
            e.nblocks = (int)lay3.nblocks;
            e.bstamp = cv.take<int>(sizeof(int) * lay3.nblocks);
            e.dstamp = cv.take<int>(sizeof(int) * lay3.nblocks);
            e.states = cv.take<int>(sizeof(int) * (size_t)n3 * ST_INTS);
            e.changed = cv.take<int>(sizeof(int) * n3);
        }
        TileItem3D *d_titems3 = nullptr;
        b.t3 = Tile3D{};
        if (tiles3) {
            // the tile engine takes over the word engine's state arrays
            // (engine 1) or has its own (engine 2)
            d_titems3 = cv.take<TileItem3D>(sizeof(TileItem3D) * n3);
            Tile3D &t = b.t3;
            uint32_t *const *shared = words3 ? e.state : nullptr;
            char *tb = cv.take<char>(tile3d_carve(t, nullptr, n3, nwords3, ntiles3, lay3.nblocks, shared));
            if (tb) tile3d_carve(t, tb, n3, nwords3, ntiles3, lay3.nblocks, shared);
            t.items = d_titems3;
            t.sitems = d_items3;
            t.io = d_io3;
            t.blk = d_blk3;
            t.nblk = (int)lay3.blk.size();
            t.ublk = d_ublk3;
            t.nublk = (int)lay3.ublk.size();
        }
        // 2D control words: [0..2] queue counts, [3..5] next slots, [6] pass base
        int *ctl = b.d_ctl;
        for (int k = 0; k < 3; ++k) b.q2.list[k] = lists2 ? lists2 + k * b.ntiles2 : nullptr;
        b.q2.cnt = ctl;
        b.q2.wctr = ctl ? ctl + 3 : nullptr;
        b.q2.stamp = b.d_stamp;
        b.q2.occ = b.d_occ;
        b.q2.ntiles = (int)b.ntiles2;
        b.d_gpass2 = ctl ? ctl + 6 : nullptr;
        b.close2.hist = hist2;
        b.close2.states = b.d_states2;
        b.close2.nitems = n2;
        b.close2.closed = nullptr;

        std::vector<Item2D> h2(n2);
        for (int i = 0; i < n2; ++i) {
            Item2D &v = h2[i];
            v.buf[0] = cv.take<uint64_t>(p2[i].state_words * sizeof(uint64_t));
            v.buf[1] = cv.take<uint64_t>(p2[i].state_words * sizeof(uint64_t));
            v.chg = nullptr;
            v.st = b.d_states2 ? b.d_states2 + (size_t)i * ST_INTS : nullptr;
            v.w = p2[i].w; v.h = p2[i].h; v.L = p2[i].L;
            v.nbands = p2[i].nbands;
            v.band_rows = p2[i].band_rows;
            v.tile0 = tile0[i];
        }
        std::vector<SIO> io3(n3);
        for (int i = 0; i < n3; ++i) {
            io3[i] = b.in3[i].io;
            io3[i].blk0 = lay3.blk0[i];
            io3[i].ublk0 = lay3.ublk0[i];
        }
        if (sweep == 0) {
            b.bytes = cv.used;
            if (size_only) return BGT_OK;
            if (b.arena) {
                if (b.bytes > b.arena_bytes)
                    return fail(BGT_ERR_NOMEM, "batch needs %zu bytes, arena holds %zu", b.bytes,
                                b.arena_bytes);
                b.mem = b.arena;
                b.owns_mem = false;
                continue;
            }
            hipError_t e = device_malloc(&b.mem, b.bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                b.mem = nullptr;
                return fail(BGT_ERR_NOMEM, "batch allocation of %zu bytes failed", b.bytes);
            }
            b.owns_mem = true;
            continue;
        }
        if (n2) {
            stage(b.d_items2, h2.data(), sizeof(Item2D) * n2);
            stage(b.d_io2, io2.data(), sizeof(IO2D) * n2);
            std::vector<ItemIO2D> fio2(n2);
            for (int i = 0; i < n2; ++i)
                fio2[i] = ItemIO2D{io2[i].in, io2[i].in_pitch, io2[i].out, io2[i].out_pitch};
            stage(b.d_fio2, fio2.data(), sizeof(ItemIO2D) * n2);
            stage(b.d_blk2, blk2.data(), sizeof(int) * blk2.size());
            stage(b.d_ublk2, ublk2.data(), sizeof(int) * ublk2.size());
            stage(b.d_tiles2, t2.data(), sizeof(Tile2D) * t2.size());
        }
        if (n3) {
            stage(d_items3, lay3.items.data(), sizeof(SItem3D) * n3);
            stage(d_geo3, lay3.geo.data(), sizeof(int2) * n3);
            stage(d_io3, io3.data(), sizeof(SIO) * n3);
            stage(d_blk3, lay3.blk.data(), sizeof(int) * lay3.blk.size());
            stage(d_ublk3, lay3.ublk.data(), sizeof(int) * lay3.ublk.size());
            if (tiles3) stage(d_titems3, titems3.data(), sizeof(TileItem3D) * n3);
        }
    }
    if (host_stage.size() > b.h_stage_bytes) {
        const size_t want = grow_bytes(host_stage.size(), b.h_stage_bytes);
        if (b.h_stage) b.h_retired.push_back(b.h_stage);
        b.h_stage = nullptr;
        b.h_stage_bytes = 0;
        if (hipHostMalloc((void **)&b.h_stage, want, hipHostMallocDefault) != hipSuccess)
            return fail(BGT_ERR_NOMEM, "pinned staging of %zu bytes failed", want);
        b.h_stage_bytes = want;
    }
    if (!host_stage.empty()) std::memcpy(b.h_stage, host_stage.data(), host_stage.size());
    for (const Staged &t : staged)
        BGT_HIP(hipMemcpyAsync(t.dst, b.h_stage + t.off, t.n, hipMemcpyHostToDevice, s));
    if (n2)
        b.grid2 = queue_pass_grid2d(b.c2, b.V2, (int)std::min<long>(b.ntiles2 * kParts, 1L << 30),
                                    blocks2(b));
    if (!b.h_live) BGT_CHECK(host_alloc_mapped(96 * sizeof(int), (void **)&b.h_live, (void **)&b.dh_live));
    if (!b.h_rep) {
        if (hipHostMalloc((void **)&b.h_rep, 16 * sizeof(unsigned long long),
                          hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
            b.h_rep = nullptr;
            return fail(BGT_ERR_NOMEM, "pinned allocation failed");
        }
    }
    void *rep_dev = nullptr;
    BGT_HIP(hipHostGetDevicePointer(&rep_dev, b.h_rep, 0));
    b.close2.report = (unsigned long long *)rep_dev;
    {
        int dev = 0, cus = 0;
        BGT_HIP(hipGetDevice(&dev));
        BGT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        b.close2.ncu = std::max(1, cus);
    }
    if (!b.ev_fork) BGT_HIP(hipEventCreateWithFlags(&b.ev_fork, hipEventDisableTiming));
    if (!b.ev_join) BGT_HIP(hipEventCreateWithFlags(&b.ev_join, hipEventDisableTiming));
    for (int k = 0; k < 2; ++k) {
        if (!b.ev_poll3[k]) BGT_HIP(hipEventCreateWithFlags(&b.ev_poll3[k], hipEventDisableTiming));
    }
    b.ready = true;
    return BGT_OK;
}

// Records the event opening (begin) or closing a timed launch on stream s;
// `open` carries the opening event's index from the one call to the other.
static int timed_mark(Batch &b, hipStream_t s, int kind, bool begin, size_t *open)
{
    if (!b.timing) return BGT_OK;
    std::lock_guard<std::mutex> lock(b.mark_mu);
    if (b.ev_used == b.events.size()) {
        for (int i = 0; i < 64; ++i) {
            hipEvent_t e;
            BGT_HIP(hipEventCreate(&e));
            b.events.push_back(e);
        }
    }
    const size_t k = b.ev_used++;
    BGT_HIP(hipEventRecord(b.events[k], s));
    if (begin)
        *open = k;
    else
        b.ev_marks.push_back({*open, k, kind});
    return BGT_OK;
}

// the queue pass kernel for `rows` register rows and V words per lane
// (the pass-0 kernel that packs from the caller's bytes when `first`; V = 1 only)
static const void *queue_pass_fn(int rows, int V, bool first = false)
{
    if (V == 4)
        return rows == 8 ? (const void *)thin2d_queue_pass<8, 4, 0> : (const void *)thin2d_queue_pass<16, 4, 0>;
    if (V == 2)
        return rows == 8 ? (const void *)thin2d_queue_pass<8, 2, 0> : (const void *)thin2d_queue_pass<16, 2, 0>;
    if (first)
        return rows == 8 ? (const void *)thin2d_queue_pass<8, 1, 2> : (const void *)thin2d_queue_pass<16, 1, 2>;
    return rows == 8 ? (const void *)thin2d_queue_pass<8, 1, 0> : (const void *)thin2d_queue_pass<16, 1, 0>;
}

static const void *chain_fn(int rows, int V)
{
    if (V == 4) return rows == 8 ? (const void *)thin2d_chain<8, 4> : (const void *)thin2d_chain<16, 4>;
    if (V == 2) return rows == 8 ? (const void *)thin2d_chain<8, 2> : (const void *)thin2d_chain<16, 2>;
    return rows == 8 ? (const void *)thin2d_chain<8, 1> : (const void *)thin2d_chain<16, 1>;
}

// the persistent 2D chain (Config2D::persistent): one launch of thin2d_chain
// on a co-resident grid (its band workgroups and the bookkeeping one)
static int launch_chain2d(Batch &b, hipStream_t s, long max_passes)
{
    const int grid = std::max(1, occupancy_grid(chain_fn(b.c2.rows, b.V2), b.c2.waves * 64,
                                                 pass_lds_bytes2d(b.c2, b.V2),
                                                 (int)std::min<long>(b.ntiles2 * kParts, 1L << 30),
                                                 blocks2(b)) - 1);
    int K = b.c2.iters_per_pass;
    int mp = (int)std::min<long>(max_passes, 0x7fffffffL);
    int *ctl = b.d_ctl + 8;
    void *args[] = {(void *)&b.d_items2, (void *)&b.d_tiles2, (void *)&b.q2, (void *)&b.close2,
                    (void *)&K, (void *)&mp, (void *)&ctl};
    size_t mk = 0;
    BGT_CHECK(timed_mark(b, s, 0, true, &mk));
    BGT_HIP(hipLaunchKernel(chain_fn(b.c2.rows, b.V2), dim3(grid + 1), dim3(b.c2.waves * 64), args,
                            pass_lds_bytes2d(b.c2, b.V2), s));
    BGT_CHECK(timed_mark(b, s, 0, false, &mk));
    return BGT_OK;
}

int launch_queue_pass2d(const Config2D &c2, int V, int grid, const Item2D *items,
                        const Tile2D *tiles, const Queue &q, const Close2D &c, const int *pbase,
                        int p, hipStream_t s, const ItemIO2D *io, bool first)
{
    const dim3 g(grid + 1), blk(c2.waves * 64);
    int K = c2.iters_per_pass;
    Queue qq = q;
    Close2D cc = c;
    void *args[] = {(void *)&items, (void *)&tiles, (void *)&qq,    (void *)&cc,
                    (void *)&K,     (void *)&pbase, (void *)&p,     (void *)&io};
    if (first && (!io || V != 1))
        return fail(BGT_ERR_INVALID, "the packing 2D pass needs caller buffers and one word per lane");
    BGT_HIP(hipLaunchKernel(queue_pass_fn(c2.rows, V, first), g, blk, args, pass_lds_bytes2d(c2, V), s));
    return BGT_OK;
}

int queue_pass_grid2d(const Config2D &c2, int V, int ntiles, int per_cu_cap)
{
    return occupancy_grid(queue_pass_fn(c2.rows, V), c2.waves * 64, pass_lds_bytes2d(c2, V), ntiles,
                          per_cu_cap);
}

// one 2D pass, number (pbase ? *pbase : 0) + p: one launch, bookkeeping included
static int launch_pass2d(Batch &b, hipStream_t s, const int *pbase, int p)
{
    size_t mk = 0;
    BGT_CHECK(timed_mark(b, s, 0, true, &mk));
    // fused pack: pass 0 (always a direct launch) reads the caller's bytes
    const bool first = b.fpack2 && !pbase && p == 0;
    BGT_CHECK(launch_queue_pass2d(b.c2, b.V2, b.grid2, b.d_items2, b.d_tiles2, b.q2, b.close2, pbase, p, s,
                                  first ? b.d_fio2 : nullptr, first));
    BGT_CHECK(timed_mark(b, s, 0, false, &mk));
    return BGT_OK;
}

// one 3D iteration from list `cur` into list cur^1; `bound` and `abound` cap
// the list and active list counts
static int launch_iter3d(Batch &b, int cur, int bound, int abound, hipStream_t s)
{
    size_t mk = 0;
    BGT_CHECK(timed_mark(b, s, 1, true, &mk));
    BGT_CHECK(sparse3d_iteration(b.e3, cur, bound, abound, sparse3d_grid(b.c3.blocks_per_cu), s));
    BGT_CHECK(timed_mark(b, s, 1, false, &mk));
    return BGT_OK;
}

// A chunk is an even number of passes / iterations so that it starts and ends
// on the same queue parity and can be captured once and replayed.
static const int kChunk = 4;
// 2D passes queued ahead of the oldest pass that has not reported yet
static const int kAhead2 = 4;

// 3D launch sizes follow the list-count bound rounded up to a power of two
// of workgroups, so that a few captured graphs cover a whole run.
static int bound_blocks3(int bound) { return pow2_ceil(std::max(1, (bound + 255) / 256)); }

static int capture(hipGraphExec_t &graph, const std::function<int(hipStream_t)> &body)
{
    graph = nullptr;
    hipStream_t cap = nullptr;
    BGT_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    BGT_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    const int rc = body(cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &g);
    (void)hipStreamDestroy(cap);
    if (rc != BGT_OK) return rc;
    if (e != hipSuccess) return fail(BGT_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        graph = nullptr;
        return fail(BGT_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    }
    return BGT_OK;
}

static int launch_chunk(Batch &b, bool three_d, bool use_graph, hipStream_t s)
{
    const int bound3 = bound_blocks3(b.bound3) * 256, abound3 = bound_blocks3(b.abound3) * 256;
    const long key3 = ((long)bound3 << 32) | (long)abound3;
    const int p0 = b.pass2;
    auto body = [&](hipStream_t st) -> int {
        for (int i = 0; i < kChunk; ++i) {
            if (three_d)
                BGT_CHECK(launch_iter3d(b, i & 1, bound3, abound3, st));
            else if (use_graph)  // pass numbers from the device, advanced per replay
                BGT_CHECK(launch_pass2d(b, st, b.d_gpass2, i));
            else
                BGT_CHECK(launch_pass2d(b, st, nullptr, p0 + i));
        }
        if (!three_d && use_graph) {
            hipLaunchKernelGGL(batch_advance, dim3(1), dim3(64), 0, st, b.d_gpass2, kChunk);
            BGT_HIP(hipGetLastError());
        }
        return BGT_OK;
    };
    if (!three_d) b.pass2 += kChunk;
    if (!use_graph) return body(s);
    if (!three_d && p0 == 0 && b.fpack2) {
        // the chunk holding pass 0 (its own kernel) runs as direct launches;
        // replays of the captured chunks then number their passes from kChunk
        for (int i = 0; i < kChunk; ++i) BGT_CHECK(launch_pass2d(b, s, nullptr, i));
        hipLaunchKernelGGL(batch_advance, dim3(1), dim3(64), 0, s, b.d_gpass2, kChunk);
        BGT_HIP(hipGetLastError());
        return BGT_OK;
    }
    if (!three_d) {
        if (!b.graph2 || b.graph2_stream != s) {
            if (b.graph2) (void)hipGraphExecDestroy(b.graph2);
            BGT_CHECK(capture(b.graph2, body));
            b.graph2_stream = s;
        }
        BGT_HIP(hipGraphLaunch(b.graph2, s));
        return BGT_OK;
    }
    if (b.graph3_stream != s) {
        for (auto &g : b.graphs3) (void)hipGraphExecDestroy(g.second);
        b.graphs3.clear();
        b.graph3_stream = s;
    }
    hipGraphExec_t graph = nullptr;
    for (auto &g : b.graphs3)
        if (g.first == key3) graph = g.second;
    if (!graph) {
        BGT_CHECK(capture(graph, body));
        b.graphs3.push_back({key3, graph});
    }
    BGT_HIP(hipGraphLaunch(graph, s));
    return BGT_OK;
}

static int collect_timing(Batch &b)
{
    b.ms2d = b.ms3d = 0;
    b.n2d = b.n3d = 0;
    for (auto &m : b.ev_marks) {
        float ms = 0;
        BGT_HIP(hipEventElapsedTime(&ms, b.events[m.begin], b.events[m.end]));
        if (m.kind == 0) { b.ms2d += ms; b.n2d++; }
        else { b.ms3d += ms; b.n3d++; }
    }
    b.ev_marks.clear();
    b.ev_used = 0;
    return BGT_OK;
}

int batch_run(Batch &b, hipStream_t s, bool use_graph)
{
    BGT_CHECK(prepare(b, s));
    const int n2 = (int)b.in2.size(), n3 = (int)b.in3.size();
    if (b.timing) use_graph = false;
    b.ev_marks.clear();
    b.ev_used = 0;

    // set-up on the caller's stream
    const long init_n = std::max<long>({(long)n2 * ST_INTS, (long)n3 * ST_INTS, 4L * n2,
                                        b.ntiles2, (long)b.nctl, (long)n3, 1L});
    hipLaunchKernelGGL(batch_init, dim3((unsigned)((init_n + 255) / 256)), dim3(256), 0, s,
                       b.d_states2, (long)n2 * ST_INTS, b.close2.hist, 4L * n2, b.d_occ,
                       b.d_stamp, b.ntiles2, b.d_ctl, b.nctl);
    b.pass2 = 0;
    BGT_HIP(hipGetLastError());

    // With both kinds of items, the 3D pipeline (its pack included) runs on a
    // side stream forked from and joined into `s`; a batch of one kind uses
    // `s` alone (no extra stream, so a caller with one stream per worker keeps
    // one hardware queue per worker).  The fork follows the 2D pack: the 2D
    // chain is the longer one, and its pack runs faster without the 3D pack
    // beside it on HBM.
    const bool fork = n2 > 0 && n3 > 0;
    if (fork && !b.side) BGT_HIP(hipStreamCreateWithFlags(&b.side, hipStreamNonBlocking));
    hipStream_t s3 = fork ? b.side : s;
    const bool fpack2 = b.fpack2;
    if (b.nblk2 && !fpack2)
        hipLaunchKernelGGL(batch_pack2d, dim3(b.nblk2), dim3(256), 0, s, b.d_items2, b.d_io2,
                           b.d_blk2, b.d_occ);
    if (fork) {
        BGT_HIP(hipEventRecord(b.ev_fork, s));
        BGT_HIP(hipStreamWaitEvent(b.side, b.ev_fork, 0));
    }
    if (n2)
        hipLaunchKernelGGL(fpack2 ? batch_list_all : batch_build_list,
                           dim3((unsigned)((b.ntiles2 + 255) / 256)), dim3(256), 0, s, b.q2);
    BGT_HIP(hipGetLastError());
    b.on_tiles3 = n3 > 0 && b.c3.engine == 2;
    if (b.on_tiles3) {
        BGT_CHECK(tile3d_begin(b.t3, !b.zeroed3, s3));
        b.zeroed3 = true;
        b.titer3 = 0;
    } else if (n3) {
        BGT_CHECK(sparse3d_begin(b.e3, !b.zeroed3, s3));
        b.zeroed3 = true;
        b.bound3 = b.abound3 = b.e3.capacity;
        b.iter3 = 0;
    }
    long max_vox = 0;
    for (auto &v : b.in2) max_vox = std::max<long>(max_vox, (long)v.w * v.h);
    for (auto &v : b.in3) max_vox = std::max<long>(max_vox, (long)v.w * v.h * v.d);
    const long max_chunks = max_vox + 3;  // an iteration that deletes nothing ends an item
    // The 2D chain follows the passes' reports: it keeps at most kAhead2
    // passes queued past the oldest unreported one, and stops at the first
    // pass that finds no queued band (every later pass finds none either and
    // leaves at once).
    auto drive2 = [&]() -> int {
        for (int k = 0; k < 16; ++k) __atomic_store_n(&b.h_rep[k], 0ull, __ATOMIC_RELAXED);
        long chunks = 0;
        for (int next = 0, idle = 0;;) {
            if (b.pass2 - next < kAhead2) {
                if (++chunks > max_chunks)
                    return fail(BGT_ERR_NOT_CONVERGED, "batch did not converge");
                BGT_CHECK(launch_chunk(b, false, use_graph, s));
                continue;
            }
            const unsigned long long v = __atomic_load_n(&b.h_rep[next & 15], __ATOMIC_ACQUIRE);
            if ((long)(v >> 32) == (long)next + 1) {
                if ((unsigned)v == 0u) return BGT_OK;
                ++next;
                idle = 0;
                continue;
            }
            // not reported yet: every 64 looks, make sure the stream still runs
            if (++idle % 64 == 0) {
                const hipError_t q = hipStreamQuery(s);
                if (q == hipSuccess) {  // all queued passes ran: the report must be there
                    const unsigned long long w =
                        __atomic_load_n(&b.h_rep[next & 15], __ATOMIC_ACQUIRE);
                    if ((long)(w >> 32) != (long)next + 1)
                        return fail(BGT_ERR_HIP, "2D pass %d did not report", next);
                    continue;
                }
                if (q != hipErrorNotReady) return fail(BGT_ERR_HIP, "2D chain: %s", hipGetErrorString(q));
                (void)hipGetLastError();  // "not ready" is not an error of a later call
            }
            std::this_thread::yield();
        }
    };
    // The 3D chain keeps two chunks in flight: chunk k+1 is queued before the
    // poll after chunk k is read, so the GPU does not wait for the host.  A
    // chunk queued after the last volume converged finds empty lists.
    struct Chain {
        bool live = false;
        bool final = false;          // its last launch is queued (the 3D tail)
        int head = 0, inflight = 0;  // oldest pending poll slot, pending polls
        bool tiles[2] = {false, false};  // a slot's poll holds the tile engine's words
        long chunks = 0;
    } c3;
    c3.live = n3 > 0;
    auto enqueue3 = [&](Chain &c) -> int {
        const int slot = (c.head + c.inflight) & 1;
        c.tiles[slot] = b.on_tiles3;
        if (b.on_tiles3) {
            // a chunk of iterations, one launch each; iteration t closes the
            // items whose iteration t - 1 changed nothing
            const int grid = tile3d_grid(b.c3.blocks_per_cu);
            size_t mk = 0;
            BGT_CHECK(timed_mark(b, s3, 1, true, &mk));
            for (int i = 0; i < kChunk; ++i) BGT_CHECK(tile3d_iteration(b.t3, b.titer3++, grid, s3));
            BGT_CHECK(timed_mark(b, s3, 1, false, &mk));
            BGT_CHECK(copy_ints_to_host(b.t3.ctl, b.dh_live + slot * 32 + 8, TC_INTS, s3));
            BGT_HIP(hipEventRecord(b.ev_poll3[slot], s3));
            ++c.inflight;
            if (++c.chunks > max_chunks) return fail(BGT_ERR_NOT_CONVERGED, "batch did not converge");
            return BGT_OK;
        }
        if (b.c3.engine == 0 && b.c3.tail_words > 0 && b.iter3 >= 1 &&
            (long)b.abound3 <= (long)b.c3.tail_words * n3) {
            // few active words per volume left: the remaining iterations in one
            // launch, one workgroup per volume
            size_t mk = 0;
            BGT_CHECK(timed_mark(b, s3, 1, true, &mk));
            BGT_CHECK(sparse3d_tail(b.e3, s3));
            BGT_CHECK(timed_mark(b, s3, 1, false, &mk));
            c.final = true;
        } else {
            BGT_CHECK(launch_chunk(b, true, use_graph, s3));
        }
        BGT_CHECK(copy_ints_to_host(b.e3.ctl, b.dh_live + slot * 32 + 8, SC_INTS, s3));
        BGT_HIP(hipEventRecord(b.ev_poll3[slot], s3));
        ++c.inflight;
        if (++c.chunks > max_chunks) return fail(BGT_ERR_NOT_CONVERGED, "batch did not converge");
        return BGT_OK;
    };
    // reads the oldest pending poll (its event has completed)
    auto consume3 = [&](Chain &c) {
        const int *hp = b.h_live + c.head * 32;
        if (c.tiles[c.head]) {
            if (hp[8 + TC_LIVE] <= 0) c.live = false;
            c.head ^= 1;
            --c.inflight;
            return;
        }
        if (b.on_tiles3) {
            // a word-list chunk queued before the tile engine took over: the
            // tile engine's polls follow it
            c.head ^= 1;
            --c.inflight;
            return;
        }
        // chunks hold an even number of iterations: list 0 is current
        b.bound3 = hp[8 + SC_COUNT0];
        b.abound3 = hp[8 + SC_ACOUNT0 + (hp[8 + SC_ITER] & 1)];
        b.iter3 = hp[8 + SC_ITER];
        if (hp[8 + SC_LIVE] <= 0) c.live = false;
        c.head ^= 1;
        --c.inflight;
    };
    // With both kinds of items each chain has a host thread of its own:
    // queueing a 3D chunk is ~32 launches, ~150 us of host time, long enough to
    // drain the short tail passes queued on the 2D chain if one thread fed
    // both.
    auto drive3 = [&](Chain &c) -> int {
        while (c.live) {
            // engine 1: once few words are active, the tile engine takes over,
            // queued behind the word-list chunks in flight: each runs kChunk
            // iterations (iterations past convergence find empty lists), so
            // the takeover iteration is known, and the list count last read
            // bounds the later ones
            const bool to_tiles = b.c3.engine == 1 && !b.on_tiles3 && b.c3.tail_words > 0 &&
                                  b.iter3 >= 1 && (long)b.abound3 <= (long)b.c3.tail_words * n3;
            if (to_tiles) {
                const int k = b.iter3 + c.inflight * kChunk;
                BGT_CHECK(tile3d_from_words(b.t3, b.e3, k, b.bound3, s3));
                b.on_tiles3 = true;
                b.titer3 = k;
                continue;
            }
            if (c.inflight < 2 && !c.final && !to_tiles) {
                BGT_CHECK(enqueue3(c));
                continue;
            }
            if (c.final && c.inflight == 0)
                return fail(BGT_ERR_NOT_CONVERGED, "3D tail did not converge");
            BGT_HIP(hipEventSynchronize(b.ev_poll3[c.head]));
            consume3(c);
        }
        return BGT_OK;
    };
    // each chain's closing launches follow its last chunk on its own stream,
    // so the 2D unpack overlaps the end of the 3D chain
    auto run2 = [&]() -> int {
        if (!n2) return BGT_OK;
        const int *dpend = nullptr;
        if (b.c2.persistent) {
            BGT_CHECK(launch_chain2d(b, s, max_vox + 3));
            dpend = b.d_ctl + 9;
        } else {
            BGT_CHECK(drive2());
        }
        hipLaunchKernelGGL(batch_close2d, dim3(1), dim3(256), 0, s, b.close2,
                           b.c2.iters_per_pass, b.pass2, dpend);
        if (b.nublk2)
            hipLaunchKernelGGL(batch_unpack2d, dim3(b.nublk2), dim3(256), 0, s, b.d_items2,
                               b.d_io2, b.d_ublk2);
        BGT_HIP(hipGetLastError());
        return BGT_OK;
    };
    auto run3 = [&]() -> int {
        if (!n3) return BGT_OK;
        BGT_CHECK(drive3(c3));
        BGT_CHECK(b.on_tiles3 ? tile3d_end(b.t3, s3) : sparse3d_end(b.e3, s3));
        if (fork) BGT_HIP(hipEventRecord(b.ev_join, s3));
        return BGT_OK;
    };
    if (fork) {
        int dev = 0;
        BGT_HIP(hipGetDevice(&dev));
        int rc3 = BGT_OK;
        std::string err3;
        std::thread t3([&] {
            const hipError_t e = hipSetDevice(dev);
            rc3 = e != hipSuccess ? fail(BGT_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e))
                                  : run3();
            if (rc3 != BGT_OK) err3 = last_error_message();
        });
        const int rc2 = run2();
        t3.join();
        BGT_CHECK(rc2);
        if (rc3 != BGT_OK) {
            set_error(rc3, err3);
            return rc3;
        }
        BGT_HIP(hipStreamWaitEvent(s, b.ev_join, 0));
    } else {
        BGT_CHECK(run2());  // at most one of the two has items
        BGT_CHECK(run3());
    }
    const size_t nst = (size_t)(n2 + n3) * ST_INTS;
    if (nst > b.h_states_ints) {
        const size_t want = grow_bytes(nst * sizeof(int), b.h_states_ints * sizeof(int)) / sizeof(int);
        if (b.h_states) b.h_retired.push_back(b.h_states);
        b.h_states = b.dh_states = nullptr;
        b.h_states_ints = 0;
        BGT_CHECK(host_alloc_mapped(want * sizeof(int), (void **)&b.h_states, (void **)&b.dh_states));
        b.h_states_ints = want;
    }
    if (n2)
        BGT_CHECK(copy_ints_to_host(b.d_states2, b.dh_states, n2 * ST_INTS, s));
    if (n3)
        BGT_CHECK(copy_ints_to_host(b.on_tiles3 ? b.t3.states : b.e3.states, b.dh_states + (size_t)n2 * ST_INTS,
                                    n3 * ST_INTS, s));
    if (n2 && b.c2.persistent)
        BGT_CHECK(copy_ints_to_host(b.d_ctl + 8, b.dh_live + 64, 3, s));
    BGT_HIP(hipStreamSynchronize(s));
    if (n2 && b.c2.persistent && b.h_live[66])
        return fail(BGT_ERR_HIP, "persistent 2D chain: a device-wide barrier waited past its limit");
    b.ran = true;
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
    if (!b.ready || !b.ran) return fail(BGT_ERR_INVALID, "batch has not run");
    // the state copy holds the 2D items, then the 3D items
    const size_t n2 = b.in2.size();
    for (size_t k = 0; k < b.order.size(); ++k) {
        const int o = b.order[k];
        const size_t i = o >= 0 ? (size_t)o : n2 + (size_t)~o;
        out[k] = b.h_states[i * ST_INTS + ST_ITERS];
    }
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
    if (w > kMaxWidth2D) return fail(BGT_ERR_UNSUPPORTED, "2D width %d exceeds %d", w, kMaxWidth2D);
    b.release();
    b.order.push_back((int)b.in2.size());
    b.in2.push_back({{in, in_pitch, out, out_pitch}, w, h});
    return BGT_OK;
}

int batch_add3d(Batch &b, const uint8_t *in, int w, int h, int d, uint8_t *out)
{
    if (!in || !out) return fail(BGT_ERR_INVALID, "null device pointer");
    if (w <= 0 || h <= 0 || d <= 0) return fail(BGT_ERR_INVALID, "bad 3D item");
    b.release();
    b.order.push_back(~(int)b.in3.size());
    b.in3.push_back({{in, out}, w, h, d});
    return BGT_OK;
}

int batch_set_config(Batch &b, const Config2D &c2, const Config3D &c3)
{
    BGT_CHECK(check_config2d(c2));
    BGT_CHECK(check_config3d(c3));
    b.release();
    b.c2 = c2;
    b.c3 = c3;
    return BGT_OK;
}

// Launch-shape autotuner: runs the batch under candidate 2D shapes (waves,
// register rows, iterations per pass) with the current 3D shape, then the
// best of those under candidate workgroups per CU, then candidate 3D step
// grids with the best 2D shape, and keeps the fastest.
// Every shape gives the same results, so only time is compared; shapes that
// do not fit an item (a band thinner than its halo) are skipped.  Each
// candidate is run once to lay the batch out and then `reps` times, taking
// the median wall time of a whole run.  The batch's current shape is timed
// first and is replaced only by a shape at least 5% faster, so that the
// choice does not flip between near-equal shapes from run to run.
int batch_autotune(Batch &b, hipStream_t s, int reps, Config2D *best2, Config3D *best3,
                   double *best_ms, int *tried)
{
    if (b.in2.empty() && b.in3.empty()) return fail(BGT_ERR_INVALID, "empty batch");
    reps = std::max(1, reps);
    static const int k2[][3] = {{4, 16, 4}, {4, 16, 3}, {4, 16, 2}, {4, 16, 6}, {4, 8, 2},
                                {4, 8, 3},  {4, 8, 4},  {2, 16, 3}, {2, 16, 4}, {4, 16, 8},
                                {8, 8, 3},  {8, 8, 4},  {8, 8, 6},  {8, 16, 3}, {8, 16, 6}};
    static const int kb2[] = {1, 2, 3, 4, 0};
    static const int k3[] = {0, 2, 4, 8, 16};
    Config2D c2 = b.c2;
    Config3D c3 = b.c3;
    double best = 1e30;
    int n = 0;
    auto trial = [&](const Config2D &t2, const Config3D &t3) {
        if (n && t2.waves == c2.waves && t2.rows == c2.rows && t2.iters_per_pass == c2.iters_per_pass &&
            t2.blocks_per_cu == c2.blocks_per_cu && t2.lane_words == c2.lane_words &&
            t2.fused_io == c2.fused_io && t2.persistent == c2.persistent &&
            t3.blocks_per_cu == c3.blocks_per_cu)
            return;  // the current shape, already timed
        if (batch_set_config(b, t2, t3) != BGT_OK) return;
        if (batch_run(b, s, false) != BGT_OK) return;  // lays out; shape may not fit
        // the median of the runs: a shape's one lucky run must not win it
        std::vector<double> ts;
        for (int r = 0; r < reps; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            if (batch_run(b, s, false) != BGT_OK) return;
            const auto t1 = std::chrono::steady_clock::now();
            ts.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
        }
        std::sort(ts.begin(), ts.end());
        const double t = ts.empty() ? 1e30 : ts[ts.size() / 2];
        const bool first = n == 0;
        ++n;
        if (first || t < best * 0.95) {
            best = t;
            c2 = t2;
            c3 = t3;
        }
    };
    const Config3D start3 = b.c3;
    Config2D cur2 = b.c2;
    cur2.blocks_per_cu = blocks2(b);  // an automatic count, as it resolves here
    trial(cur2, start3);  // the current shape first
    if (!b.in2.empty()) {
        for (const auto &k : k2) {
            Config2D t2 = cur2;
            t2.waves = k[0];
            t2.rows = k[1];
            t2.iters_per_pass = k[2];
            trial(t2, start3);
        }
        const Config2D shape2 = c2;
        for (int bpc : kb2) {
            Config2D t2 = shape2;
            t2.blocks_per_cu = bpc;
            trial(t2, start3);
        }
    }
    if (!b.in3.empty()) {
        const Config2D base2 = c2;
        for (int bpc : k3) {
            Config3D t3 = start3;
            t3.blocks_per_cu = bpc;
            trial(base2, t3);
        }
    }
    if (!n) return fail(BGT_ERR_INVALID, "no launch shape fits this batch");
    BGT_CHECK(batch_set_config(b, c2, c3));
    *best2 = c2;
    *best3 = c3;
    *best_ms = best;
    if (tried) *tried = n;
    return BGT_OK;
}

size_t batch_device_bytes(Batch &b) { return b.ready ? b.bytes : 0; }

int batch_plan_bytes(Batch &b, size_t *bytes)
{
    BGT_CHECK(prepare(b, nullptr, true));
    *bytes = b.bytes;
    return BGT_OK;
}

void batch_set_arena(Batch &b, void *mem, size_t bytes)
{
    b.release();
    b.arena = mem;
    b.arena_bytes = mem ? bytes : 0;
}

void batch_clear(Batch &b)
{
    b.release();
    b.in2.clear();
    b.in3.clear();
    b.order.clear();
}

int batch_size(const Batch &b) { return (int)(b.in2.size() + b.in3.size()); }

}  // namespace bgt
