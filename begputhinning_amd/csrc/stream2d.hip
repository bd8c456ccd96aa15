// stream2d.hip -- the continuously running 2D band queue (see stream2d.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "pack.h"
#include "queue2d.h"
#include "stream2d.h"
#include "thin2d_kernels.h"

namespace bgt {

namespace {

// passes queued ahead of the oldest one that has not started yet: an image
// submitted now joins at most this many passes later
constexpr int kAhead = 2;

// every slot empty (done, so that the pass kernel's bookkeeping skips it),
// tile t = band t % mb of slot t / mb, queue counters zero
__global__ void stream2d_init(int *states, int nslots, unsigned *hist, Tile2D *tiles, uint8_t *occ,
                              int *stamp, int ntiles, int mb, int *ctl, int nctl)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nslots * ST_INTS) states[g] = (g % ST_INTS) == ST_DONE ? 1 : 0;
    if (g < 4 * nslots) hist[g] = 0u;
    if (g < ntiles) {
        tiles[g] = Tile2D{g / mb, g % mb};
        occ[g] = 0;
        stamp[g] = 0;
    }
    if (g < nctl) ctl[g] = 0;
}

// a slot takes an image whose first pass is P: its item, a fresh state and
// empty band flags
__global__ void stream2d_admit(Item2D *items, int slot, Item2D val, int *states, uint8_t *occ,
                               int *stamp, int mb, int P)
{
    const int t = threadIdx.x;
    if (t == 0) items[slot] = val;
    if (t < ST_INTS) states[(size_t)slot * ST_INTS + t] = t == ST_BASE ? P : 0;
    for (int b = t; b < mb; b += blockDim.x) {
        occ[slot * mb + b] = 0;
        stamp[slot * mb + b] = 0;
    }
}

// uint8 rows -> both state buffers; marks the bands holding pixels
__global__ __launch_bounds__(256) void stream2d_pack(const Item2D *__restrict__ items, int slot,
                                                     const uint8_t *__restrict__ in, size_t pitch,
                                                     uint8_t *occ)
{
    const Item2D it = items[slot];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)it.L * it.h) return;
    const int y = (int)(t / it.L), x0 = (int)(t - (long)y * it.L) * 64;
    const uint64_t v = lane_order(pack_run_u8(in + (size_t)y * pitch + x0, min(64, it.w - x0)));
    it.buf[0][t] = v;
    it.buf[1][t] = v;
    if (v) occ[it.tile0 + y / it.band_rows] = 1;
}

// the slot's bands that hold pixels join pass P's queue
__global__ void stream2d_enqueue(Queue q, const Item2D *__restrict__ items, int slot, int P)
{
    const Item2D it = items[slot];
    for (int b = threadIdx.x; b < it.nbands; b += blockDim.x) {
        const int tile = it.tile0 + b;
        if (!q.occ[tile]) continue;
        q.stamp[tile] = P;
        q.list[P % 3][atomicAdd(&q.cnt[P % 3], 1)] = tile;
    }
}

}  // namespace

struct Stream2D {
    struct Req {
        const uint8_t *in;
        size_t in_pitch;
        uint8_t *out;
        size_t out_pitch;
        int w, h;
        void *tag;
    };
    enum { FREE = 0, RUNNING = 1, UNPACKING = 2 };
    struct Slot {
        int state = FREE;
        Req r{};
        hipEvent_t ev = nullptr;
    };

    Stream2DConfig cfg;
    Config2D c2;
    int V = 1;  // words per lane of every slot's bands (those of the largest image)
    Stream2DDone done;
    int device = 0;
    int mb = 0;                 // bands per slot
    size_t slot_words = 0;      // 64-bit words of each state buffer of a slot
    hipStream_t s = nullptr;    // passes and admissions
    hipStream_t su = nullptr;   // unpacks
    hipEvent_t ev_pass = nullptr;  // the passes queued before an unpack
    void *mem = nullptr;
    size_t bytes = 0;
    Item2D *d_items = nullptr;
    Tile2D *d_tiles = nullptr;
    int *d_states = nullptr;
    uint64_t *d_bufs = nullptr;  // [slots][2][slot_words]
    uint8_t *d_chg = nullptr;    // unused by the queue passes; Item2D.chg of every slot
    Queue q{};
    Close2D c{};
    int grid = 0;
    unsigned long long *h_rep = nullptr;  // host-mapped pass reports
    unsigned *h_closed = nullptr;         // host-mapped per-slot closed flags
    int *h_states = nullptr;              // pinned state copies after unpack (mapped)
    int *dh_states = nullptr;             // their device address
    std::vector<Slot> slots;

    std::mutex m;
    std::condition_variable cv;
    std::deque<Req> pending;
    bool stop = false;
    std::thread driver;
    std::atomic<uint64_t> passes{0}, admitted{0};
    int pass_next = 0, reported = 0, live = 0;

    void run();
    int admit(int k, const Req &r);
    void fail_all(int code, const char *why);
};

bool stream2d_fits(const Stream2D *e, int w, int h)
{
    Plan2D pl;
    if (w <= 0 || h <= 0 || plan2d(w, h, e->c2, pl, e->V) != BGT_OK) return false;
    return pl.nbands <= e->mb && pl.state_words <= e->slot_words;
}

int Stream2D::admit(int k, const Req &r)
{
    Plan2D pl;
    BGT_CHECK(plan2d(r.w, r.h, c2, pl, V));
    Item2D v{};
    v.buf[0] = d_bufs + (size_t)k * 2 * slot_words;
    v.buf[1] = v.buf[0] + slot_words;
    v.chg = d_chg;
    v.st = d_states + (size_t)k * ST_INTS;
    v.w = pl.w; v.h = pl.h; v.L = pl.L;
    v.nbands = pl.nbands;
    v.band_rows = pl.band_rows;
    v.tile0 = k * mb;
    __atomic_store_n(&h_closed[k], 0u, __ATOMIC_RELEASE);
    const int P = pass_next;
    hipLaunchKernelGGL(stream2d_admit, dim3(1), dim3(256), 0, s, d_items, k, v, d_states, q.occ,
                       q.stamp, mb, P);
    const long nw = (long)pl.L * pl.h;
    hipLaunchKernelGGL(stream2d_pack, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, d_items, k,
                       r.in, r.in_pitch, q.occ);
    hipLaunchKernelGGL(stream2d_enqueue, dim3(1), dim3(256), 0, s, q, d_items, k, P);
    BGT_HIP(hipGetLastError());
    slots[k].state = RUNNING;
    slots[k].r = r;
    ++live;
    ++admitted;
    return BGT_OK;
}

void Stream2D::fail_all(int code, const char *why)
{
    std::deque<Req> left;
    {
        std::lock_guard<std::mutex> lk(m);
        left.swap(pending);
    }
    for (const Req &r : left) {
        set_error(code, why);
        done(r.tag, code, 0);
    }
}

void Stream2D::run()
{
    (void)hipSetDevice(device);
    for (;;) {
        bool progress = false;
        // ---- admit waiting images into free slots, ahead of the next pass
        std::vector<std::pair<int, Req>> take;
        bool stopping;
        {
            std::unique_lock<std::mutex> lk(m);
            if (!stop && live == 0 && pending.empty() && reported == pass_next)
                cv.wait(lk, [&] { return stop || !pending.empty(); });
            stopping = stop;
            if (!stopping) {
                for (int k = 0; k < (int)slots.size() && !pending.empty(); ++k)
                    if (slots[k].state == FREE) {
                        take.push_back({k, pending.front()});
                        pending.pop_front();
                        slots[k].state = RUNNING;  // reserved
                    }
            }
        }
        if (stopping) {
            // images not admitted yet fail; the admitted ones run to the end
            fail_all(BGT_ERR_INVALID, "serving queue stopped before the request ran");
            if (live == 0 && reported == pass_next) break;
        }
        for (auto &kr : take) {
            slots[kr.first].state = FREE;
            const int rc = admit(kr.first, kr.second);
            if (rc != BGT_OK) done(kr.second.tag, rc, 0);
            progress = true;
        }
        // ---- keep the passes flowing while images are in the queue
        while (live > 0 && pass_next - reported < kAhead) {
            const int rc = launch_queue_pass2d(c2, V, grid, d_items, d_tiles, q, c, nullptr, pass_next, s);
            if (rc != BGT_OK) {
                // the queue cannot go on: every image in it or waiting fails
                const std::string why = last_error_message();
                (void)hipStreamSynchronize(s);
                (void)hipStreamSynchronize(su);
                for (Slot &sl : slots)
                    if (sl.state != FREE) {
                        sl.state = FREE;
                        set_error(rc, why);
                        done(sl.r.tag, rc, 0);
                    }
                live = 0;
                fail_all(rc, why.c_str());
                {
                    std::lock_guard<std::mutex> lk(m);
                    stop = true;
                }
                return;
            }
            ++pass_next;
            ++passes;
            progress = true;
        }
        // ---- passes that have started
        while (reported < pass_next) {
            const unsigned long long v = __atomic_load_n(&h_rep[reported & 15], __ATOMIC_ACQUIRE);
            if ((long)(v >> 32) != (long)reported + 1) break;
            ++reported;
            progress = true;
        }
        // ---- closed images: unpack on the side stream; done when it ends.
        // The pass that closed an image may still be running when its flag
        // is seen, and its state words (the parity of the final buffer) are
        // only known to be visible once it has ended: the side stream first
        // waits for the passes queued so far.
        bool waited = false;
        for (int k = 0; k < (int)slots.size(); ++k) {
            Slot &sl = slots[k];
            if (sl.state == RUNNING && __atomic_load_n(&h_closed[k], __ATOMIC_ACQUIRE)) {
                if (!waited) {
                    (void)hipEventRecord(ev_pass, s);
                    (void)hipStreamWaitEvent(su, ev_pass, 0);
                    waited = true;
                }
                launch_unpack_item2d(d_items + k, sl.r.out, sl.r.w, sl.r.h, sl.r.out_pitch, su);
                (void)copy_ints_to_host(d_states + (size_t)k * ST_INTS, dh_states + (size_t)k * ST_INTS, ST_INTS,
                                        su);
                (void)hipEventRecord(sl.ev, su);
                sl.state = UNPACKING;
                progress = true;
            } else if (sl.state == UNPACKING) {
                const hipError_t q2 = hipEventQuery(sl.ev);
                if (q2 == hipErrorNotReady) continue;
                const int rc = q2 == hipSuccess ? BGT_OK : fail(BGT_ERR_HIP, "serving unpack: %s",
                                                                 hipGetErrorString(q2));
                const int iters = h_states[(size_t)k * ST_INTS + ST_ITERS];
                sl.state = FREE;
                --live;
                done(sl.r.tag, rc, iters);
                progress = true;
            }
        }
        if (!progress) std::this_thread::yield();
    }
}

int stream2d_create(const Stream2DConfig &cfg, const Config2D &c2, Stream2DDone done,
                    Stream2D **out)
{
    *out = nullptr;
    if (cfg.slots < 1 || cfg.max_w < 1 || cfg.max_h < 1) return fail(BGT_ERR_INVALID, "bad serving queue");
    BGT_CHECK(check_config2d(c2));
    std::unique_ptr<Stream2D> e(new Stream2D());
    e->cfg = cfg;
    e->c2 = c2;
    e->done = std::move(done);
    BGT_HIP(hipGetDevice(&e->device));
    Plan2D big;
    BGT_CHECK(plan2d(cfg.max_w, cfg.max_h, c2, big));
    e->V = big.V;
    // the fewest interior rows per band any admitted width gets (widest rows)
    e->mb = (cfg.max_h + big.band_rows - 1) / big.band_rows;
    e->slot_words = big.state_words;
    const int ns = cfg.slots, nt = ns * e->mb;
    // device memory: items, tiles, flags, lists, counters, states, history,
    // band-flag stand-in, state buffers
    size_t off = 0;
    auto place = [&](size_t n) {
        const size_t at = off;
        off += round_up(n ? n : 1, 256);
        return at;
    };
    const size_t o_items = place(sizeof(Item2D) * ns), o_tiles = place(sizeof(Tile2D) * nt),
                 o_occ = place(nt), o_stamp = place(sizeof(int) * nt),
                 o_lists = place(sizeof(int) * 3 * (size_t)nt), o_ctl = place(sizeof(int) * 16),
                 o_states = place(sizeof(int) * ST_INTS * (size_t)ns),
                 o_hist = place(sizeof(unsigned) * 4 * (size_t)ns), o_chg = place(2),
                 o_bufs = place(sizeof(uint64_t) * 2 * e->slot_words * (size_t)ns);
    e->bytes = off;
    if (device_malloc(&e->mem, e->bytes) != hipSuccess) {
        (void)hipGetLastError();
        e->mem = nullptr;
        return fail(BGT_ERR_NOMEM, "serving queue of %zu bytes failed", e->bytes);
    }
    char *b = (char *)e->mem;
    e->d_items = (Item2D *)(b + o_items);
    e->d_tiles = (Tile2D *)(b + o_tiles);
    int *ctl = (int *)(b + o_ctl);
    e->d_states = (int *)(b + o_states);
    e->d_chg = (uint8_t *)(b + o_chg);
    e->d_bufs = (uint64_t *)(b + o_bufs);
    int *lists = (int *)(b + o_lists);
    for (int k = 0; k < 3; ++k) e->q.list[k] = lists + (size_t)k * nt;
    e->q.cnt = ctl;
    e->q.wctr = ctl + 3;
    e->q.stamp = (int *)(b + o_stamp);
    e->q.occ = (uint8_t *)(b + o_occ);
    e->q.ntiles = nt;
    e->c.hist = (unsigned *)(b + o_hist);
    e->c.states = e->d_states;
    e->c.nitems = ns;
    int cus = 0;
    BGT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
    e->c.ncu = std::max(1, cus);
    if (hipHostMalloc((void **)&e->h_rep, 16 * sizeof(unsigned long long),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc((void **)&e->h_closed, ns * sizeof(unsigned),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        host_alloc_mapped((size_t)ns * ST_INTS * sizeof(int), (void **)&e->h_states, (void **)&e->dh_states) !=
            BGT_OK) {
        (void)hipGetLastError();
        if (e->h_rep) (void)hipHostFree(e->h_rep);
        if (e->h_closed) (void)hipHostFree(e->h_closed);
        (void)hipFree(e->mem);
        return fail(BGT_ERR_NOMEM, "serving queue host buffers failed");
    }
    for (int k = 0; k < 16; ++k) e->h_rep[k] = 0ull;
    for (int k = 0; k < ns; ++k) e->h_closed[k] = 0u;
    void *dp = nullptr;
    BGT_HIP(hipHostGetDevicePointer(&dp, e->h_rep, 0));
    e->c.report = (unsigned long long *)dp;
    BGT_HIP(hipHostGetDevicePointer(&dp, e->h_closed, 0));
    e->c.closed = (unsigned *)dp;
    BGT_HIP(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking));
    BGT_HIP(hipStreamCreateWithFlags(&e->su, hipStreamNonBlocking));
    BGT_HIP(hipEventCreateWithFlags(&e->ev_pass, hipEventDisableTiming));
    e->slots.resize(ns);
    for (auto &sl : e->slots) BGT_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    e->grid = queue_pass_grid2d(c2, e->V, nt * 16, cfg.blocks_per_cu);
    const int ninit = std::max({ns * ST_INTS, 4 * ns, nt, 16});
    hipLaunchKernelGGL(stream2d_init, dim3((ninit + 255) / 256), dim3(256), 0, e->s, e->d_states, ns,
                       e->c.hist, e->d_tiles, e->q.occ, e->q.stamp, nt, e->mb, ctl, 16);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(e->s));
    Stream2D *raw = e.get();
    e->driver = std::thread([raw] { raw->run(); });
    *out = e.release();
    return BGT_OK;
}

void stream2d_destroy(Stream2D *e)
{
    if (!e) return;
    {
        std::lock_guard<std::mutex> lk(e->m);
        e->stop = true;
    }
    e->cv.notify_all();
    if (e->driver.joinable()) e->driver.join();  // running images finish first
    (void)hipStreamSynchronize(e->s);
    (void)hipStreamSynchronize(e->su);
    for (auto &sl : e->slots)
        if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (e->ev_pass) (void)hipEventDestroy(e->ev_pass);
    if (e->s) (void)hipStreamDestroy(e->s);
    if (e->su) (void)hipStreamDestroy(e->su);
    if (e->mem) (void)hipFree(e->mem);
    if (e->h_rep) (void)hipHostFree(e->h_rep);
    if (e->h_closed) (void)hipHostFree(e->h_closed);
    if (e->h_states) (void)hipHostFree(e->h_states);
    delete e;
}

int stream2d_submit(Stream2D *e, const uint8_t *in, size_t in_pitch, int w, int h, uint8_t *out,
                    size_t out_pitch, void *tag)
{
    if (!in || !out) return fail(BGT_ERR_INVALID, "null device pointer");
    if (in_pitch < (size_t)w || out_pitch < (size_t)w) return fail(BGT_ERR_INVALID, "pitch smaller than width");
    if (!stream2d_fits(e, w, h))
        return fail(BGT_ERR_UNSUPPORTED, "%d x %d image does not fit a serving queue slot", w, h);
    {
        std::lock_guard<std::mutex> lk(e->m);
        if (e->stop) return fail(BGT_ERR_INVALID, "serving queue stopped");
        e->pending.push_back(Stream2D::Req{in, in_pitch, out, out_pitch, w, h, tag});
    }
    e->cv.notify_all();
    return BGT_OK;
}

size_t stream2d_device_bytes(const Stream2D *e) { return e->bytes; }

void stream2d_stats(const Stream2D *e, uint64_t *passes, uint64_t *admitted)
{
    *passes = e->passes.load();
    *admitted = e->admitted.load();
}

}  // namespace bgt
