// capi.cpp -- the extern "C" entry points declared in include/bgt.h, error
// reporting and the per-thread device workspace.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <sys/syscall.h>
#include <unistd.h>
#include <vector>

#include "batch.h"
#include "common.h"
#include "seggraph.h"
#include "segments.h"
#include "server.h"
#include "thin2d.h"
#include "thin3d.h"

namespace bgt {

static thread_local std::string g_error;
static thread_local int g_error_code = BGT_OK;

void set_error(int code, const std::string &msg)
{
    g_error_code = code;
    g_error = msg;
}

std::string last_error_message() { return g_error; }

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    set_error(code, buf);
    return code;
}

int workspace_reserve(Workspace &ws, size_t bytes)
{
    int dev = 0;
    BGT_HIP(hipGetDevice(&dev));
    if (ws.ptr && ws.bytes >= bytes && ws.device == dev) return BGT_OK;
    const size_t want = grow_bytes(bytes, ws.device == dev ? ws.bytes : 0);
    if (ws.ptr) {  // work queued before may still use it: freed with the workspace
        ws.retired.emplace_back(ws.ptr, ws.device);
        ws.ptr = nullptr;
        ws.bytes = 0;
    }
    void *p = nullptr;
    hipError_t e = device_malloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(BGT_ERR_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    ws.ptr = p;
    ws.bytes = want;
    ws.device = dev;
    return BGT_OK;
}

// device allocations left to fail (bgt_debug_fail_device_allocs)
static std::atomic<int> g_fail_allocs{0};

hipError_t device_malloc(void **p, size_t bytes)
{
    for (int n = g_fail_allocs.load(); n > 0; n = g_fail_allocs.load())
        if (g_fail_allocs.compare_exchange_weak(n, n - 1)) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
    return hipMalloc(p, bytes);
}

bool on_main_thread() { return (long)syscall(SYS_gettid) == (long)getpid(); }

Workspace::~Workspace()
{
    // at process exit the runtime may be going away: the main thread's memory
    // is returned with the process
    if (on_main_thread()) return;
    if (ptr) retired.emplace_back(ptr, device);
    ptr = nullptr;
    int cur = -1;
    (void)hipGetDevice(&cur);
    for (const auto &r : retired) {
        if (r.second != cur) (void)hipSetDevice(r.second);
        (void)hipFree(r.first);
        if (r.second != cur && cur >= 0) (void)hipSetDevice(cur);
    }
    retired.clear();
}

PinnedBuffer::~PinnedBuffer()
{
    if (ptr && !on_main_thread()) (void)hipHostFree(ptr);
    ptr = nullptr;
}

Workspace &thread_workspace(int slot)
{
    // 0, 1: 2D state and I/O; 2, 3: 3D state and I/O; 4: a host array traced
    // on the device; 5, 6, 7: segment pre-pass
    thread_local Workspace ws[8];
    return ws[slot & 7];
}

// Process-wide launch shapes: read from the environment once (invalid values
// are reported on stderr and the defaults kept), replaced by
// bgt_set_config2d/3d, and handed to each entry point as a copy taken under
// the lock, so a concurrent set_config never changes a running call's shape.
static std::mutex g_config_mu;

static Config2D &config2d_storage()
{
    static Config2D c = [] {
        Config2D v;
        if (const char *e = std::getenv("BGT_2D_WAVES")) v.waves = std::atoi(e);
        if (const char *e = std::getenv("BGT_2D_ROWS")) v.rows = std::atoi(e);
        if (const char *e = std::getenv("BGT_2D_K")) v.iters_per_pass = std::atoi(e);
        if (const char *e = std::getenv("BGT_2D_BLOCKS")) v.blocks_per_cu = std::atoi(e);
        if (const char *e = std::getenv("BGT_2D_LANE_WORDS")) v.lane_words = std::atoi(e);
        if (const char *e = std::getenv("BGT_2D_FUSED_IO")) v.fused_io = std::atoi(e);
        if (const char *e = std::getenv("BGT_2D_PERSISTENT")) v.persistent = std::atoi(e);
        if (check_config2d(v) != BGT_OK) {
            std::fprintf(stderr, "libbgt: ignoring BGT_2D_* environment: %s\n", g_error.c_str());
            v = Config2D{};
        }
        return v;
    }();
    return c;
}

Config2D default_config2d()
{
    std::lock_guard<std::mutex> lk(g_config_mu);
    return config2d_storage();
}

static Config3D &config3d_storage()
{
    static Config3D c = [] {
        Config3D v;
        if (const char *e = std::getenv("BGT_3D_BLOCKS")) v.blocks_per_cu = std::atoi(e);
        if (const char *e = std::getenv("BGT_3D_CHUNK")) v.max_chunk = std::atoi(e);
        if (const char *e = std::getenv("BGT_3D_TAIL")) v.tail_words = std::atoi(e);
        if (const char *e = std::getenv("BGT_3D_ENGINE")) v.engine = std::atoi(e);
        if (check_config3d(v) != BGT_OK) {
            std::fprintf(stderr, "libbgt: ignoring BGT_3D_* environment: %s\n", g_error.c_str());
            v = Config3D{};
        }
        return v;
    }();
    return c;
}

Config3D default_config3d()
{
    std::lock_guard<std::mutex> lk(g_config_mu);
    return config3d_storage();
}

}  // namespace bgt

using namespace bgt;

extern "C" {

const char *bgt_last_error(void) { return g_error.c_str(); }

int bgt_version(void) { return 1; }

int bgt_set_config2d(int waves, int rows, int iters_per_pass, int blocks_per_cu)
{
    Config2D c = default_config2d();  // the words per lane are kept
    c.waves = waves;
    c.rows = rows;
    c.iters_per_pass = iters_per_pass;
    c.blocks_per_cu = blocks_per_cu;
    BGT_CHECK(check_config2d(c));
    std::lock_guard<std::mutex> lk(g_config_mu);
    config2d_storage() = c;
    return BGT_OK;
}

int bgt_set_lane_words2d(int words)
{
    Config2D c = default_config2d();
    c.lane_words = words;
    BGT_CHECK(check_config2d(c));
    std::lock_guard<std::mutex> lk(g_config_mu);
    config2d_storage() = c;
    return BGT_OK;
}

int bgt_debug_fail_device_allocs(int n)
{
    if (n < 0) return fail(BGT_ERR_INVALID, "negative allocation count %d", n);
    g_fail_allocs.store(n);
    return BGT_OK;
}

int bgt_set_persistent2d(int on)
{
    Config2D c = default_config2d();
    c.persistent = on;
    BGT_CHECK(check_config2d(c));
    std::lock_guard<std::mutex> lk(g_config_mu);
    config2d_storage() = c;
    return BGT_OK;
}

int bgt_set_fused_io2d(int on)
{
    Config2D c = default_config2d();
    c.fused_io = on;
    BGT_CHECK(check_config2d(c));
    std::lock_guard<std::mutex> lk(g_config_mu);
    config2d_storage() = c;
    return BGT_OK;
}

int bgt_set_config3d(int blocks_per_cu, int max_chunk, int tail_words)
{
    Config3D c = default_config3d();  // the engine choice is kept
    c.blocks_per_cu = blocks_per_cu;
    c.max_chunk = max_chunk;
    c.tail_words = tail_words;
    BGT_CHECK(check_config3d(c));
    std::lock_guard<std::mutex> lk(g_config_mu);
    config3d_storage() = c;
    return BGT_OK;
}

int bgt_set_engine3d(int engine)
{
    Config3D c = default_config3d();
    c.engine = engine;
    BGT_CHECK(check_config3d(c));
    std::lock_guard<std::mutex> lk(g_config_mu);
    config3d_storage() = c;
    return BGT_OK;
}

int bgt_thinning2d(int dtype, int width, int height, const void *in, void *out)
{
    return thin2d_host(dtype, width, height, in, out, default_config2d());
}

int bgt_thin2d_device(const uint8_t *d_in, size_t in_pitch, int width, int height, uint8_t *d_out,
                      size_t out_pitch, void *stream, int *iterations)
{
    return thin2d_device_u8(d_in, in_pitch, width, height, d_out, out_pitch,
                            (hipStream_t)stream, iterations, default_config2d());
}

int bgt_thin2d_device_noalloc(const uint8_t *d_in, size_t in_pitch, int width, int height,
                              uint8_t *d_out, size_t out_pitch, void *stream, int *iterations)
{
    return thin2d_device_u8_noalloc(d_in, in_pitch, width, height, d_out, out_pitch,
                                    (hipStream_t)stream, iterations, default_config2d());
}

size_t bgt_thin2d_noalloc_max_bytes(void) { return thin2d_noalloc_max_bytes(); }

int bgt_thinning3d(int dtype, int width, int height, int depth, const void *in, void *out)
{
    return thin3d_host(dtype, width, height, depth, in, out, default_config3d());
}

int bgt_thin3d_device(const uint8_t *d_in, int width, int height, int depth, uint8_t *d_out,
                      void *stream, int *iterations)
{
    return thin3d_device_u8(d_in, width, height, depth, d_out, (hipStream_t)stream, iterations,
                            default_config3d());
}

// Host arrays of at least this many pixels are traced through the device
// when one is present (copy in, pre-pass, host walk of the compact graph):
// 2-7x faster than the frame scan from about a megapixel.  Smaller arrays,
// and every array on a host without a device, take the host's frame scan.
// Both give the same segment lists.
static constexpr long kDeviceTraceMin = 1L << 19;

static bool have_device()
{
    static const bool yes = [] {
        int n = 0;
        const bool ok = hipGetDeviceCount(&n) == hipSuccess && n > 0;
        (void)hipGetLastError();
        return ok;
    }();
    return yes;
}

static int extract_host(int dtype, const void *data, int w, int h, int d, bool three_d, bgt_segments *out)
{
    if (data && out && w > 0 && h > 0 && d > 0 && w <= 32767 && h <= 32767 && d <= 32767 &&
        dtype_size(dtype) && (long)w * h * d >= kDeviceTraceMin && have_device()) {
        std::memset(out, 0, sizeof(*out));
        SegGraph g;
        BGT_CHECK(segment_graph_host(dtype, data, w, h, d, three_d, nullptr, g));
        return extract_segments_graph(g, w, h, three_d, out);
    }
    return extract_segments(dtype, data, w, h, d, three_d, out);
}

int bgt_extract_segments2d(int dtype, const void *img, int width, int height, bgt_segments *out)
{
    return extract_host(dtype, img, width, height, 1, false, out);
}

int bgt_extract_segments3d(int dtype, const void *vol, int width, int height, int depth,
                           bgt_segments *out)
{
    return extract_host(dtype, vol, width, height, depth, true, out);
}

static int extract_device(const void *d_skel, int w, int h, int d, int pw, int ph, int pd,
                          bool three_d, void *stream, bgt_segments *out)
{
    if (!d_skel || !out) return fail(BGT_ERR_INVALID, "null pointer");
    std::memset(out, 0, sizeof(*out));
    if (w <= 0 || h <= 0 || d <= 0) return fail(BGT_ERR_INVALID, "bad size");
    if (pw < w || ph < h || pd < d) return fail(BGT_ERR_INVALID, "padded frame smaller than skeleton");
    if (pw > 32767 || ph > 32767 || pd > 32767)
        return fail(BGT_ERR_UNSUPPORTED, "coordinates are int16 (sides <= 32767)");
    SegGraph g;
    BGT_CHECK(segment_graph_device((const uint8_t *)d_skel, w, h, d, pw, ph, pd, three_d,
                                   (hipStream_t)stream, g));
    return extract_segments_graph(g, pw, ph, three_d, out);
}

int bgt_extract_segments2d_device(const void *d_skel, int width, int height, int padded_width,
                                  int padded_height, void *stream, bgt_segments *out)
{
    return extract_device(d_skel, width, height, 1, padded_width, padded_height, 1, false, stream,
                          out);
}

int bgt_extract_segments3d_device(const void *d_skel, int width, int height, int depth,
                                  int padded_width, int padded_height, int padded_depth,
                                  void *stream, bgt_segments *out)
{
    return extract_device(d_skel, width, height, depth, padded_width, padded_height, padded_depth,
                          true, stream, out);
}

void bgt_segments_free(bgt_segments *s)
{
    if (!s) return;
    std::free(s->x); std::free(s->y); std::free(s->z);
    std::free(s->seg_start); std::free(s->is_end);
    std::memset(s, 0, sizeof(*s));
}

void bgt_free(void *p) { std::free(p); }

int bgt_mex_columns(const float *x, const float *y, const float *z, const uint32_t *out_idx,
                    uint32_t nsegs, double *out)
{
    return mex_columns(x, y, z, out_idx, nsegs, out);
}

struct bgt_batch_s {
    bgt::Batch *b;
};

int bgt_batch_create(bgt_batch *out)
{
    if (!out) return fail(BGT_ERR_INVALID, "null pointer");
    *out = nullptr;
    bgt_batch h = new bgt_batch_s{batch_new()};
    const int rc = batch_set_config(*h->b, default_config2d(), default_config3d());
    if (rc != BGT_OK) {  // no handle escapes a failed create
        batch_delete(h->b);
        delete h;
        return rc;
    }
    *out = h;
    return BGT_OK;
}

int bgt_batch_use_current_config(bgt_batch b)
{
    if (!b) return fail(BGT_ERR_INVALID, "null batch");
    return batch_set_config(*b->b, default_config2d(), default_config3d());
}

int bgt_batch_add2d(bgt_batch b, const uint8_t *d_in, size_t in_pitch, int width, int height,
                    uint8_t *d_out, size_t out_pitch)
{
    if (!b) return fail(BGT_ERR_INVALID, "null batch");
    return batch_add2d(*b->b, d_in, in_pitch, width, height, d_out, out_pitch);
}

int bgt_batch_add3d(bgt_batch b, const uint8_t *d_in, int width, int height, int depth,
                    uint8_t *d_out)
{
    if (!b) return fail(BGT_ERR_INVALID, "null batch");
    return batch_add3d(*b->b, d_in, width, height, depth, d_out);
}

int bgt_batch_run(bgt_batch b, void *stream, int use_graph)
{
    if (!b) return fail(BGT_ERR_INVALID, "null batch");
    return batch_run(*b->b, (hipStream_t)stream, use_graph != 0);
}

int bgt_batch_iterations(bgt_batch b, int *iterations)
{
    if (!b || !iterations) return fail(BGT_ERR_INVALID, "null pointer");
    return batch_iterations(*b->b, iterations);
}

int bgt_batch_device_bytes(bgt_batch b, size_t *bytes)
{
    if (!b || !bytes) return fail(BGT_ERR_INVALID, "null pointer");
    *bytes = batch_device_bytes(*b->b);
    return BGT_OK;
}

int bgt_batch_set_timing(bgt_batch b, int on)
{
    if (!b) return fail(BGT_ERR_INVALID, "null batch");
    batch_set_timing(*b->b, on != 0);
    return BGT_OK;
}

int bgt_batch_kernel_time(bgt_batch b, double *ms2d, int *launches2d, double *ms3d,
                          int *launches3d)
{
    if (!b || !ms2d || !launches2d || !ms3d || !launches3d) return fail(BGT_ERR_INVALID, "null pointer");
    batch_kernel_time(*b->b, ms2d, launches2d, ms3d, launches3d);
    return BGT_OK;
}

int bgt_batch_autotune(bgt_batch b, void *stream, int reps, int *waves, int *rows,
                       int *iters_per_pass, int *blocks_per_cu_2d, int *blocks_per_cu_3d,
                       double *best_ms)
{
    if (!b) return fail(BGT_ERR_INVALID, "null batch");
    Config2D c2;
    Config3D c3;
    double ms = 0;
    BGT_CHECK(batch_autotune(*b->b, (hipStream_t)stream, reps, &c2, &c3, &ms, nullptr));
    if (waves) *waves = c2.waves;
    if (rows) *rows = c2.rows;
    if (iters_per_pass) *iters_per_pass = c2.iters_per_pass;
    if (blocks_per_cu_2d) *blocks_per_cu_2d = c2.blocks_per_cu;
    if (blocks_per_cu_3d) *blocks_per_cu_3d = c3.blocks_per_cu;
    if (best_ms) *best_ms = ms;
    return BGT_OK;
}

int bgt_batch_destroy(bgt_batch b)
{
    if (!b) return BGT_OK;
    batch_delete(b->b);
    delete b;
    return BGT_OK;
}

struct bgt_server_s {
    Server *s;
};

int bgt_server_create(bgt_server *out, size_t hbm_budget, int workers2d, int workers3d,
                      int max_items)
{
    if (!out) return fail(BGT_ERR_INVALID, "null pointer");
    ServerConfig cfg;
    cfg.budget = hbm_budget;
    cfg.workers2d = workers2d;
    cfg.workers3d = workers3d;
    cfg.max_items = max_items;
    // BGT_SERVE_QUEUE=0: frames are batched by the 2D workers only
    const char *q = std::getenv("BGT_SERVE_QUEUE");
    cfg.stream_slots = (q && std::atoi(q) == 0) ? 0 : max_items;
    Server *s = nullptr;
    BGT_CHECK(server_create(cfg, default_config2d(), default_config3d(), &s));
    *out = new bgt_server_s{s};
    return BGT_OK;
}

int bgt_server_submit2d(bgt_server sv, const uint8_t *d_in, size_t in_pitch, int width, int height,
                        uint8_t *d_out, size_t out_pitch, uint64_t *ticket)
{
    if (!sv) return fail(BGT_ERR_INVALID, "null server");
    return server_submit(sv->s, 0, d_in, in_pitch, d_out, out_pitch, width, height, 1, ticket);
}

int bgt_server_submit3d(bgt_server sv, const uint8_t *d_in, int width, int height, int depth,
                        uint8_t *d_out, uint64_t *ticket)
{
    if (!sv) return fail(BGT_ERR_INVALID, "null server");
    return server_submit(sv->s, 1, d_in, 0, d_out, 0, width, height, depth, ticket);
}

int bgt_server_wait(bgt_server sv, uint64_t ticket, int *iterations, double *latency_ms)
{
    if (!sv) return fail(BGT_ERR_INVALID, "null server");
    return server_wait(sv->s, ticket, iterations, latency_ms);
}

int bgt_server_enable_segments(bgt_server sv, int threads)
{
    if (!sv) return fail(BGT_ERR_INVALID, "null server");
    return server_enable_segments(sv->s, threads);
}

int bgt_server_wait_segments(bgt_server sv, uint64_t ticket, int *iterations, double *latency_ms,
                             double *segments_ms, bgt_segments *out)
{
    if (!sv) return fail(BGT_ERR_INVALID, "null server");
    return server_wait_segments(sv->s, ticket, iterations, latency_ms, segments_ms, out);
}

int bgt_server_poll(bgt_server sv, uint64_t ticket, int *done)
{
    if (!sv || !done) return fail(BGT_ERR_INVALID, "null pointer");
    return server_poll(sv->s, ticket, done);
}

int bgt_server_stats(bgt_server sv, uint64_t *batches, uint64_t *completed, size_t *peak_bytes)
{
    if (!sv || !batches || !completed || !peak_bytes) return fail(BGT_ERR_INVALID, "null pointer");
    server_stats(sv->s, batches, completed, peak_bytes);
    return BGT_OK;
}

int bgt_server_destroy(bgt_server sv)
{
    if (!sv) return BGT_OK;
    server_destroy(sv->s);
    delete sv;
    return BGT_OK;
}

// 2D: pad to pw = ceil32(w), ph = max(1, (h+63)/128)*128 + 64
// (2D/CenterlineExtraction.cpp:27-28); thinning does not depend on the zero
// padding, tracing does (it skips the padded frame's border).
int bgt_centerline2d(int dtype, const void *img, int width, int height, int max_segment_length,
                     int smoothing_span, float **out_x, float **out_y, uint32_t **out_idx,
                     uint32_t *nsegs)
{
    if (!img || !out_x || !out_y || !out_idx || !nsegs) return fail(BGT_ERR_INVALID, "null pointer");
    if (width <= 0 || height <= 0) return fail(BGT_ERR_INVALID, "bad size");
    const int pw = (width + 31) / 32 * 32;
    const int q = (height + 63) / 128;
    const int ph = (q > 1 ? q : 1) * 128 + 64;
    if (pw > 32767 || ph > 32767) return fail(BGT_ERR_UNSUPPORTED, "padded size exceeds int16");
    // thin on the device, then classify and compact the skeleton points there:
    // the host traces only those, not the padded frame
    const uint8_t *d_skel = nullptr;
    BGT_CHECK(thin2d_host_keep(dtype, width, height, img, &d_skel, default_config2d()));
    SegGraph g;
    BGT_CHECK(segment_graph_device(d_skel, width, height, 1, pw, ph, 1, false, nullptr, g));
    bgt_segments s;
    BGT_CHECK(extract_segments_graph(g, pw, ph, false, &s));
    int r = centerline_from_segments(s, false, max_segment_length, smoothing_span, out_x, out_y,
                                     nullptr, out_idx, nsegs);
    bgt_segments_free(&s);
    return r;
}

// 3D: pad each side to a multiple of 32 with zeros (3D/CenterlineExtraction.cpp:27-47,
// whose padding the reference leaves uninitialised), foreground iff int(v) > 0.
int bgt_centerline3d(int dtype, const void *vol, int width, int height, int depth,
                     int max_segment_length, int smoothing_span, float **out_x, float **out_y,
                     float **out_z, uint32_t **out_idx, uint32_t *nsegs)
{
    if (!vol || !out_x || !out_y || !out_z || !out_idx || !nsegs)
        return fail(BGT_ERR_INVALID, "null pointer");
    if (width <= 0 || height <= 0 || depth <= 0) return fail(BGT_ERR_INVALID, "bad size");
    const int pw = (width + 31) / 32 * 32, ph = (height + 31) / 32 * 32, pd = (depth + 31) / 32 * 32;
    if (pw > 32767 || ph > 32767 || pd > 32767)
        return fail(BGT_ERR_UNSUPPORTED, "padded size exceeds int16");
    // threshold and thin on the device, then classify and compact the skeleton
    // points there: the host traces only those, not the padded volume
    const uint8_t *d_skel = nullptr;
    BGT_CHECK(thin3d_keep_int_positive(dtype, width, height, depth, vol, &d_skel,
                                       default_config3d()));
    SegGraph g;
    BGT_CHECK(segment_graph_device(d_skel, width, height, depth, pw, ph, pd, true, nullptr, g));
    bgt_segments s;
    BGT_CHECK(extract_segments_graph(g, pw, ph, true, &s));
    int r = centerline_from_segments(s, true, max_segment_length, smoothing_span, out_x, out_y,
                                     out_z, out_idx, nsegs);
    bgt_segments_free(&s);
    return r;
}

}  // extern "C"
