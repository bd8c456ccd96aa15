// server.cpp -- request loop with continuous batching over the batch engine.
//
// Requests (device-resident images and volumes, caller-owned input and output
// buffers) are queued per class (2D, 3D).  Worker threads -- each with its
// own stream, batch engine and grow-only device arena, so no allocation or
// free happens per batch -- take every queued request of their class that
// fits the HBM budget (the engine memory of all admitted requests stays below
// it), thin them as one batch and release their memory.  A request arriving
// while a batch runs is taken by the next idle worker at once, so its work
// overlaps the running batch on the device; frames never wait behind volumes
// because the classes have their own workers.  Each result equals the
// single-item result (the batch engine's per-item convergence), whatever the
// mix it ran in.
//
// No counterpart in the reference (it thins one image per blocking call,
// 2D/LibThinning2D.cu:39-130).
#include <hip/hip_runtime.h>
#include <sys/resource.h>
#include <sys/syscall.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.h"
#include "common.h"
#include "seggraph.h"
#include "segments.h"
#include "server.h"
#include "stream2d.h"
#include "thin2d.h"
#include "thin3d.h"

namespace bgt {

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    int cls = 0;  // 0 = 2D, 1 = 3D
    const uint8_t *in = nullptr;
    size_t in_pitch = 0;
    uint8_t *out = nullptr;
    size_t out_pitch = 0;
    int w = 0, h = 0, d = 1;
    size_t bytes = 0;  // admission footprint (engine memory of the item alone)
    double t_submit = 0, t_done = 0;
    int iterations = 0;
    int status = BGT_OK;
    std::string error;  // message of a failed batch (set on the worker thread)
    bool done = false;
    // segment tracing of served skeletons (server_enable_segments)
    bool traced = false;      // segments ready (or tracing failed or not wanted)
    int trace_status = BGT_OK;
    std::string trace_error;
    bgt_segments segs{};
    double t_segs = 0;
};

void free_segments(bgt_segments &g)
{
    std::free(g.x); std::free(g.y); std::free(g.z);
    std::free(g.seg_start); std::free(g.is_end);
    std::memset(&g, 0, sizeof(g));
}

}  // namespace

struct Server {
    ServerConfig cfg;
    Config2D c2;
    Config3D c3;
    int device = 0;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue[2];
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    size_t in_use = 0;
    uint64_t next_id = 1;
    bool stop = false;
    std::vector<std::thread> threads;
    // frames that fit its slots join the continuously running 2D queue at the
    // next pass boundary; other frames and all volumes go to the workers
    Stream2D *q2d = nullptr;
    // statistics
    uint64_t batches = 0, completed = 0;
    size_t peak_in_use = 0;
    // tracer threads: each finished request's skeleton to its segment list
    // (the device pre-pass on the thread's stream, then the host trace)
    int tracing = 0;
    bool tstop = false;  // set once no request can finish any more
    std::deque<Request *> tq;
    std::condition_variable cv_trace;
    std::vector<std::thread> tracers;
    const uint8_t *warm = nullptr;  // warm-up skeletons while tracers start
    int tracers_warm = 0;

    void worker(int cls);
    void tracer();
    void finish(const std::vector<Request *> &take, const std::vector<int> &iters, int rc);
    void completed_locked(Request *r);
};

// With tracing on, a request whose skeleton is done (successfully) goes to
// the tracers; otherwise it has nothing more to wait for.  Holds m.
void Server::completed_locked(Request *r)
{
    if (tracing && r->status == BGT_OK) {
        tq.push_back(r);
        cv_trace.notify_one();
    } else {
        r->traced = true;
    }
}

// The warm-up skeletons: a 16 x 16 frame (a line of 11 pixels with a branch
// of 5 meeting it) and, after it, an 8 x 8 x 8 volume (a line of 6 voxels
// with a branch of 3).
const int kWarm2 = 16 * 16, kWarm3 = 8 * 8 * 8;

std::vector<uint8_t> warm_skeletons()
{
    std::vector<uint8_t> v(kWarm2 + kWarm3, 0);
    for (int x = 2; x <= 12; ++x) v[8 * 16 + x] = 1;
    for (int y = 3; y <= 7; ++y) v[y * 16 + 7] = 1;
    uint8_t *c = v.data() + kWarm2;
    for (int x = 1; x <= 6; ++x) c[(4 * 8 + 4) * 8 + x] = 1;
    for (int z = 1; z <= 3; ++z) c[(z * 8 + 4) * 8 + 3] = 1;
    return v;
}

void Server::tracer()
{
    // below the engine's host threads, which must react to every pass: a
    // trace waits rather than a pass
    (void)setpriority(PRIO_PROCESS, (id_t)syscall(SYS_gettid), 10);
    (void)hipSetDevice(device);
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    // Before the first request: this thread's workspaces (for up to 2^17
    // points of a 256^3 volume, which also covers a 1024 x 1024 frame's) and
    // the pre-pass kernels' first launches, which otherwise cost the first
    // requests milliseconds.
    // Failures here are left to the requests to report.
    if (s) {
        (void)segment_reserve(256, 256, 256, true, 1L << 17);
        for (int k = 0; warm && k < 2; ++k) {
            const bool three = k == 1;
            const int e = three ? 8 : 16;
            SegGraph gr;
            bgt_segments g{};
            if (segment_graph_device(warm + (three ? kWarm2 : 0), e, e, three ? e : 1, e, e, three ? e : 1,
                                     three, s, gr) == BGT_OK)
                (void)extract_segments_graph(gr, e, e, three, &g);
            free_segments(g);
        }
    }
    {
        std::lock_guard<std::mutex> lk(m);
        ++tracers_warm;
    }
    cv_done.notify_all();
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(m);
            cv_trace.wait(lk, [&] { return tstop || !tq.empty(); });
            if (tq.empty()) break;  // stopping, nothing left
            r = tq.front();
            tq.pop_front();
        }
        bgt_segments g{};
        int rc;
        if (!s) {
            rc = fail(BGT_ERR_HIP, "tracer stream creation failed");
        } else if (r->out_pitch != (size_t)r->w && r->cls == 0) {
            rc = fail(BGT_ERR_UNSUPPORTED, "tracing a served image needs contiguous output rows");
        } else {
            SegGraph gr;
            rc = segment_graph_device(r->out, r->w, r->h, r->d, r->w, r->h, r->d, r->cls == 1, s, gr);
            if (rc == BGT_OK) rc = extract_segments_graph(gr, r->w, r->h, r->cls == 1, &g);
        }
        const std::string err = rc == BGT_OK ? std::string() : last_error_message();
        {
            std::lock_guard<std::mutex> lk(m);
            r->trace_status = rc;
            r->trace_error = err;
            r->segs = g;
            r->t_segs = now_ms();
            r->traced = true;
        }
        cv_done.notify_all();
    }
    if (s) (void)hipStreamDestroy(s);
}

// Marks a batch's requests done (with the status and, on failure, the
// worker thread's error message) and releases their memory.
void Server::finish(const std::vector<Request *> &take, const std::vector<int> &iters, int rc)
{
    const double t = now_ms();
    const std::string err = rc == BGT_OK ? std::string() : std::string(bgt_last_error());
    {
        std::lock_guard<std::mutex> lk(m);
        for (size_t k = 0; k < take.size(); ++k) {
            Request *r = take[k];
            r->status = rc;
            r->error = err;
            r->iterations = iters[k];
            r->t_done = t;
            r->done = true;
            in_use -= r->bytes;
            completed_locked(r);
        }
        completed += take.size();
        ++batches;
    }
    cv_done.notify_all();
    cv_work.notify_all();
}

void Server::worker(int cls)
{
    (void)hipSetDevice(device);
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    std::unique_ptr<Batch, void (*)(Batch *)> b(batch_new(), batch_delete);
    (void)batch_set_config(*b, c2, c3);
    void *arena = nullptr;
    size_t arena_bytes = 0;
    std::vector<void *> retired;
    std::vector<Request *> take;
    std::vector<int> iters;
    for (;;) {
        take.clear();
        {
            std::unique_lock<std::mutex> lk(m);
            cv_work.wait(lk, [&] {
                return stop || (!queue[cls].empty() && in_use + queue[cls].front()->bytes <= cfg.budget);
            });
            if (stop) break;
            while (!queue[cls].empty() && (int)take.size() < cfg.max_items &&
                   in_use + queue[cls].front()->bytes <= cfg.budget) {
                Request *r = queue[cls].front();
                queue[cls].pop_front();
                in_use += r->bytes;
                take.push_back(r);
            }
            peak_in_use = std::max(peak_in_use, in_use);
        }
        int rc = s ? BGT_OK : fail(BGT_ERR_HIP, "worker stream creation failed");
        iters.assign(take.size(), 0);
        if (rc == BGT_OK && take.size() == 1) {
            // a batch of one takes the per-item path (no batch layout, no
            // table uploads); its result is the same
            Request *r = take[0];
            rc = r->cls == 0 ? thin2d_device_u8(r->in, r->in_pitch, r->w, r->h, r->out,
                                                r->out_pitch, s, &iters[0], c2)
                             : thin3d_device_u8(r->in, r->w, r->h, r->d, r->out, s, &iters[0], c3);
            finish(take, iters, rc);
            continue;
        }
        batch_clear(*b);
        for (Request *r : take) {
            if (rc != BGT_OK) break;
            rc = r->cls == 0 ? batch_add2d(*b, r->in, r->in_pitch, r->w, r->h, r->out, r->out_pitch)
                             : batch_add3d(*b, r->in, r->w, r->h, r->d, r->out);
        }
        size_t need = 0;
        if (rc == BGT_OK) rc = batch_plan_bytes(*b, &need);
        if (rc == BGT_OK && need > arena_bytes) {
            batch_set_arena(*b, nullptr, 0);
            // the outgrown arena is freed when the worker leaves: hipFree
            // waits for the whole device, stalling the other workers
            if (arena) retired.push_back(arena);
            const size_t grow = grow_bytes(need, arena_bytes);
            arena = nullptr;
            arena_bytes = 0;
            if (device_malloc(&arena, grow) == hipSuccess) {
                arena_bytes = grow;
            } else {
                (void)hipGetLastError();
                arena = nullptr;
                rc = fail(BGT_ERR_NOMEM, "server arena of %zu bytes failed", grow);
            }
        }
        if (rc == BGT_OK) {
            batch_set_arena(*b, arena, arena_bytes);
            rc = batch_run(*b, s, false);
        }
        if (rc == BGT_OK) rc = batch_iterations(*b, iters.data());
        finish(take, iters, rc);
    }
    batch_set_arena(*b, nullptr, 0);
    b.reset();
    if (arena) (void)hipFree(arena);
    for (void *p : retired) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
}

int server_create(const ServerConfig &cfg, const Config2D &c2, const Config3D &c3, Server **out)
{
    if (cfg.budget == 0 || cfg.workers2d < 0 || cfg.workers3d < 0 ||
        cfg.workers2d + cfg.workers3d < 1 || cfg.max_items < 1)
        return fail(BGT_ERR_INVALID, "bad server configuration");
    BGT_CHECK(check_config2d(c2));
    BGT_CHECK(check_config3d(c3));
    std::unique_ptr<Server> sv(new Server());
    sv->cfg = cfg;
    sv->c2 = c2;
    sv->c3 = c3;
    BGT_HIP(hipGetDevice(&sv->device));
    if (cfg.workers2d > 0 && cfg.stream_slots > 0) {
        // slots for frames up to 2048 x 2048, within half of the budget
        Stream2DConfig qc;
        qc.max_w = qc.max_h = 2048;
        Plan2D big;
        BGT_CHECK(plan2d(qc.max_w, qc.max_h, c2, big));
        const size_t slot_bytes = 2 * big.state_words * sizeof(uint64_t) + 4096;
        qc.slots = (int)std::min<size_t>((size_t)std::min(cfg.stream_slots, 64), cfg.budget / 2 / slot_bytes);
        if (qc.slots >= 1) {
            Server *raw = sv.get();
            BGT_CHECK(stream2d_create(qc, c2,
                                      [raw](void *tag, int status, int iters) {
                                          Request *r = (Request *)tag;
                                          const std::string err =
                                              status == BGT_OK ? std::string() : last_error_message();
                                          {
                                              std::lock_guard<std::mutex> lk(raw->m);
                                              r->status = status;
                                              r->error = err;
                                              r->iterations = iters;
                                              r->t_done = now_ms();
                                              r->done = true;
                                              ++raw->completed;
                                              raw->completed_locked(r);
                                          }
                                          raw->cv_done.notify_all();
                                      },
                                      &sv->q2d));
            sv->in_use = stream2d_device_bytes(sv->q2d);  // held for the server's lifetime
            sv->peak_in_use = sv->in_use;
        }
    }
    for (int k = 0; k < cfg.workers2d; ++k) sv->threads.emplace_back(&Server::worker, sv.get(), 0);
    for (int k = 0; k < cfg.workers3d; ++k) sv->threads.emplace_back(&Server::worker, sv.get(), 1);
    *out = sv.release();
    return BGT_OK;
}

void server_destroy(Server *sv)
{
    if (!sv) return;
    {
        // requests not yet taken by a worker fail; batches in flight finish
        std::lock_guard<std::mutex> lk(sv->m);
        sv->stop = true;
        for (auto &q : sv->queue)
            for (Request *r : q) {
                r->status = BGT_ERR_INVALID;
                r->error = "server destroyed before the request ran";
                r->done = true;
                r->traced = true;
            }
        sv->queue[0].clear();
        sv->queue[1].clear();
    }
    sv->cv_work.notify_all();
    sv->cv_done.notify_all();
    stream2d_destroy(sv->q2d);  // its running frames finish, waiting ones fail
    for (auto &t : sv->threads) t.join();
    {
        std::lock_guard<std::mutex> lk(sv->m);
        sv->tstop = true;
    }
    sv->cv_trace.notify_all();  // the tracers finish the queued skeletons, then leave
    for (auto &t : sv->tracers) t.join();
    for (auto &kv : sv->reqs) free_segments(kv.second->segs);
    delete sv;
}

int server_enable_segments(Server *sv, int threads)
{
    if (!sv || threads < 1 || threads > 64) return fail(BGT_ERR_INVALID, "bad tracer count");
    std::unique_lock<std::mutex> lk(sv->m);
    if (sv->tracing) return fail(BGT_ERR_INVALID, "segment tracing is already on");
    if (!sv->reqs.empty()) return fail(BGT_ERR_INVALID, "turn segment tracing on before submitting");
    // the warm-up skeletons on the server's device (without them the tracers
    // only reserve their workspaces)
    int cur = -1;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(sv->device);
    void *warm = nullptr;
    const std::vector<uint8_t> pattern = warm_skeletons();
    if (device_malloc(&warm, pattern.size()) != hipSuccess ||
        hipMemcpy(warm, pattern.data(), pattern.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        if (warm) (void)hipFree(warm);
        warm = nullptr;
    }
    if (cur >= 0) (void)hipSetDevice(cur);
    sv->warm = (const uint8_t *)warm;
    sv->tracing = threads;
    for (int k = 0; k < threads; ++k) sv->tracers.emplace_back(&Server::tracer, sv);
    sv->cv_done.wait(lk, [&] { return sv->tracers_warm == threads; });
    sv->warm = nullptr;
    lk.unlock();
    if (warm) (void)hipFree(warm);
    return BGT_OK;
}

int server_wait_segments(Server *sv, uint64_t ticket, int *iterations, double *latency_ms,
                         double *segments_ms, bgt_segments *out)
{
    if (out) std::memset(out, 0, sizeof(*out));
    std::unique_lock<std::mutex> lk(sv->m);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return fail(BGT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
    Request *r = it->second.get();
    sv->cv_done.wait(lk, [&] { return r->done && r->traced; });
    int rc = r->status;
    std::string err = r->error;
    if (rc == BGT_OK && r->trace_status != BGT_OK) {
        rc = r->trace_status;
        err = r->trace_error;
    }
    if (iterations) *iterations = r->iterations;
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    if (segments_ms) *segments_ms = r->t_segs - r->t_submit;
    if (out && rc == BGT_OK) {
        *out = r->segs;
        std::memset(&r->segs, 0, sizeof(r->segs));
    }
    free_segments(r->segs);
    sv->reqs.erase(it);
    if (rc != BGT_OK) set_error(rc, "request failed: " + err);
    return rc;
}

static int item_bytes(const Request &r, const Config2D &c2, const Config3D &c3, size_t *bytes)
{
    std::unique_ptr<Batch, void (*)(Batch *)> b(batch_new(), batch_delete);
    BGT_CHECK(batch_set_config(*b, c2, c3));
    BGT_CHECK(r.cls == 0 ? batch_add2d(*b, r.in, r.in_pitch, r.w, r.h, r.out, r.out_pitch)
                         : batch_add3d(*b, r.in, r.w, r.h, r.d, r.out));
    return batch_plan_bytes(*b, bytes);
}

int server_submit(Server *sv, int cls, const uint8_t *in, size_t in_pitch, uint8_t *out,
                  size_t out_pitch, int w, int h, int d, uint64_t *ticket)
{
    if (!sv || !ticket) return fail(BGT_ERR_INVALID, "null server or ticket");
    if ((cls == 0 ? sv->cfg.workers2d : sv->cfg.workers3d) == 0)
        return fail(BGT_ERR_INVALID, "the server has no %s workers", cls == 0 ? "2D" : "3D");
    const double t_submit = now_ms();  // latency includes admission
    std::unique_ptr<Request> r(new Request());
    r->cls = cls;
    r->in = in;
    r->in_pitch = in_pitch;
    r->out = out;
    r->out_pitch = out_pitch;
    r->w = w;
    r->h = h;
    r->d = d;
    if (cls == 0 && sv->q2d && stream2d_fits(sv->q2d, w, h)) {
        // the frame joins the running queue (its slot is already allocated)
        Request *rp = r.get();
        {
            std::lock_guard<std::mutex> lk(sv->m);
            r->id = sv->next_id++;
            r->t_submit = t_submit;
            *ticket = r->id;
            sv->reqs.emplace(r->id, std::move(r));
        }
        const int rc = stream2d_submit(sv->q2d, in, in_pitch, w, h, out, out_pitch, rp);
        if (rc != BGT_OK) {
            std::lock_guard<std::mutex> lk(sv->m);
            sv->reqs.erase(*ticket);
        }
        return rc;
    }
    BGT_CHECK(item_bytes(*r, sv->c2, sv->c3, &r->bytes));
    if (r->bytes > sv->cfg.budget)
        return fail(BGT_ERR_ADMISSION, "request needs %zu bytes of device memory, budget is %zu",
                    r->bytes, sv->cfg.budget);
    {
        std::lock_guard<std::mutex> lk(sv->m);
        r->id = sv->next_id++;
        r->t_submit = t_submit;
        *ticket = r->id;
        sv->queue[cls].push_back(r.get());
        sv->reqs.emplace(r->id, std::move(r));
    }
    sv->cv_work.notify_all();
    return BGT_OK;
}

int server_wait(Server *sv, uint64_t ticket, int *iterations, double *latency_ms)
{
    std::unique_lock<std::mutex> lk(sv->m);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return fail(BGT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
    Request *r = it->second.get();
    // with tracing on, the record lives until its tracer is done with it
    sv->cv_done.wait(lk, [&] { return r->done && r->traced; });
    const int rc = r->status;
    const std::string err = r->error;
    if (iterations) *iterations = r->iterations;
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    free_segments(r->segs);
    sv->reqs.erase(it);
    if (rc != BGT_OK) set_error(rc, "request failed: " + err);
    return rc;
}

int server_poll(Server *sv, uint64_t ticket, int *done)
{
    std::lock_guard<std::mutex> lk(sv->m);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return fail(BGT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
    *done = it->second->done ? 1 : 0;
    return BGT_OK;
}

void server_stats(Server *sv, uint64_t *batches, uint64_t *completed, size_t *peak_bytes)
{
    std::lock_guard<std::mutex> lk(sv->m);
    *batches = sv->batches;
    *completed = sv->completed;
    *peak_bytes = sv->peak_in_use;
}

}  // namespace bgt
