// segments.cpp -- host segment tracer (ExtractSegmentsV) and the centerline
// post-processing (prune, smooth, flatten) of the product library.
//
// Reference: 2D/ExtractSegments.cpp:16-127, 3D/ExtractSegments.cpp:135-251,
// 2D/CenterlineExtraction.cpp:51-131, 3D/CenterlineExtraction.cpp:55-144.
//
// The tracer's visiting order, neighbour order and node bookkeeping decide the
// output list, so they follow the reference step for step: skeleton pixels in
// raster order, codes -1 (isolated) / 2 (path) / -n (end or junction with n
// neighbours) kept in a signed byte, segments started from every negative code
// and followed through code-2 pixels.  What differs is only how the work is
// organised: the first pass scans 64-bit words and skips empty ones, and the
// list and coordinates are kept in flat arrays.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"
#include "seggraph.h"
#include "segments.h"

namespace bgt {

namespace {

struct Flat {
    std::vector<int16_t> x, y, z;
    std::vector<uint32_t> start{0};
    std::vector<uint8_t> is_end;
};

// mask: 0/1 bytes, dims (w, h, d) with d == 1 for 2D; nd = 8 or 26
void trace(const uint8_t *mask, int w, int h, int d, bool three_d, Flat &out)
{
    const long sl = (long)w * h;
    const size_t n = (size_t)sl * d;
    long delta[26];
    int nd = 0;
    if (!three_d) {
        const long t[8] = {-(long)w - 1, -(long)w, -(long)w + 1, -1, 1, (long)w - 1, (long)w, (long)w + 1};
        for (long v : t) delta[nd++] = v;
    } else {
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    if (dz || dy || dx) delta[nd++] = dz * sl + dy * (long)w + dx;
    }
    std::vector<int8_t> code(n, 0);
    std::vector<long> pts;
    // pass 1: interior skeleton pixels in raster order
    const int z0 = three_d ? 1 : 0, z1 = three_d ? d - 1 : 1;
    for (int z = z0; z < z1; ++z)
        for (int y = 1; y < h - 1; ++y) {
            const uint8_t *row = mask + (size_t)z * sl + (size_t)y * w;
            int x = 1;
            while (x < w - 1) {
                // skip 8 zero bytes at a time
                if (x + 8 <= w - 1) {
                    uint64_t v;
                    std::memcpy(&v, row + x, 8);
                    if (!v) { x += 8; continue; }
                }
                if (row[x]) {
                    const long p = (long)z * sl + (long)y * w + x;
                    int nn = 0;
                    for (int k = 0; k < nd; ++k) nn += mask[p + delta[k]] != 0;
                    if (nn == 0) nn = 1;
                    if (nn != 2) nn = -nn;
                    code[p] = (int8_t)nn;
                    pts.push_back(p);
                }
                ++x;
            }
        }
    std::vector<int8_t> orig(code);

    auto emit = [&](long q) {
        const int zz = (int)(q / sl);
        const long r = q - (long)zz * sl;
        out.x.push_back((int16_t)(r % w));
        out.y.push_back((int16_t)(r / w));
        if (three_d) out.z.push_back((int16_t)zz);
    };
    auto close = [&](bool end) {
        out.is_end.push_back(end ? 1 : 0);
        out.start.push_back((uint32_t)out.x.size());
    };

    // pass 2: follow segments from every end point / junction
    for (long idx : pts) {
        const int8_t nodes = code[idx];
        if (nodes >= 0) continue;
        code[idx] = 0;
        int ni = -1;
        for (int8_t si = 0; si > nodes; --si) {
            bool end = std::abs((int)orig[idx]) < 2;
            emit(idx);
            ++ni;
            while (ni < nd && code[idx + delta[ni]] == 0) ++ni;
            if (ni >= nd) { close(end); continue; }
            long t = idx + delta[ni];
            while (code[t] == 2) {
                code[t] = 0;
                emit(t);
                for (int k = 0; k < nd; ++k)
                    if (code[t + delta[k]] != 0) { t += delta[k]; break; }
            }
            emit(t);
            if (std::abs((int)orig[t]) < 2) end = true;
            code[t] = (int8_t)(code[t] + 1);
            close(end);
        }
    }
}

// The second pass of trace() over the device pre-pass's compact graph
// (seggraph.hip): the same visiting order, neighbour order and code
// bookkeeping, with a point's rank in place of its raster index and a
// neighbour rank of -1 (not a point) in place of a zero code.
//
// Only nodes start segments, so the pass visits the node list.  A walk enters
// a path point only from a neighbour that is not a path point, i.e. at an end
// of its chain, and the chain is then untouched: its points change code only
// when a walk passes through them, and a walk that enters a chain passes
// through all of it (from each point the only live neighbour is the next one,
// the previous being cleared).  So the walk emits the whole chain from the
// entered end and goes on from the far end e: to e's neighbour off the chain
// if that is live (a node: nodes never reach code 2), else it stops at e, as
// the step-by-step walk does.  For a single point, the neighbour off the chain
// is the one of its two the walk did not come from.  Of a chain's points only
// the ends neighbour points off it, so only their codes are cleared; no other
// point of it is looked at again.  A path point that is not such an end can
// only be reached by stepping through a chain, which never happens, and is
// reported as an inconsistent graph.
//
// A node's neighbours that are points come as one list (node_ent), so the
// scan for the next live neighbour skips the neighbours that are not points
// without looking at them.  The segments are written straight into the
// output arrays, allocated for the most the walk can emit: a node with code
// -n starts at most n segments, each of two points besides the chain it
// passes through, and every chain is passed through at most once, so there
// are at most S = sum over nodes of n segments and 2 S + npts points.
template <bool D3>
int trace_graph(const SegGraph &g, bgt_segments *out)
{
    // offsets of the nd neighbours in the packed coordinates, in the
    // reference's order (a point's neighbours lie inside the frame, so no
    // field borrows from another)
    int64_t delta[26];
    {
        int j = 0;
        for (int dz = D3 ? -1 : 0; dz <= (D3 ? 1 : 0); ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    if (dz || dy || dx) delta[j++] = (int64_t)dx + (int64_t)dy * 65536 + (int64_t)dz * 4294967296LL;
    }
    const int8_t *orig = g.code;
    size_t max_seg = 0;
    for (int k = 0; k < g.nnodes; ++k) max_seg += (size_t)-(int)orig[g.node_rank[k]];
    const size_t max_pts = 2 * max_seg + (size_t)g.npts;
    if (max_pts > 0xffffffffu) return fail(BGT_ERR_UNSUPPORTED, "segment list beyond 2^32 points");
    int16_t *ox = (int16_t *)std::malloc(sizeof(int16_t) * (max_pts + 1));
    int16_t *oy = (int16_t *)std::malloc(sizeof(int16_t) * (max_pts + 1));
    int16_t *oz = D3 ? (int16_t *)std::malloc(sizeof(int16_t) * (max_pts + 1)) : nullptr;
    uint32_t *ost = (uint32_t *)std::malloc(sizeof(uint32_t) * (max_seg + 1));
    uint8_t *oend = (uint8_t *)std::malloc(max_seg + 1);
    out->x = ox;
    out->y = oy;
    out->z = oz;
    out->seg_start = ost;
    out->is_end = oend;
    if (!ox || !oy || (D3 && !oz) || !ost || !oend) return fail(BGT_ERR_NOMEM, "segment list allocation failed");
    // this thread's working copy of the codes, kept between calls
    thread_local std::vector<int8_t> code;
    code.assign(g.code, g.code + g.npts);
    size_t np = 0, ns = 0;
    ost[0] = 0;
    auto emit = [&](int64_t v) {
        const uint64_t q = (uint64_t)v;  // x | y << 16 | z << 32
        ox[np] = (int16_t)(q & 0xffffu);
        oy[np] = (int16_t)((q >> 16) & 0xffffu);
        if (D3) oz[np] = (int16_t)(q >> 32);
        ++np;
    };
    auto close = [&](bool end) {
        oend[ns] = end ? 1 : 0;
        ost[++ns] = (uint32_t)np;
    };
    auto live = [&](int r) { return r >= 0 && code[r] != 0; };
    for (int k = 0; k < g.nnodes; ++k) {
        const int idx = g.node_rank[k];
        const int8_t nodes = code[idx];
        if (nodes >= 0) continue;
        code[idx] = 0;
        const int64_t *ent = g.node_ent + g.node_off[k];
        const int ne = g.node_off[k + 1] - g.node_off[k];
        const int64_t here = g.node_xyz[k];
        int ei = -1;
        for (int8_t si = 0; si > nodes; --si) {
            bool end = std::abs((int)orig[idx]) < 2;
            emit(here);
            ++ei;
            while (ei < ne && !code[(uint32_t)ent[ei]]) ++ei;
            if (ei >= ne) {
                close(end);
                continue;
            }
            int t = (int)(uint32_t)ent[ei];
            int64_t txyz = here + delta[(ent[ei] >> 32) & 31];
            if (code[t] == 2) {
                const int ch = (int)(ent[ei] >> 37) - 1;
                if (ch < 0 || ch >= g.nchains)
                    return fail(BGT_ERR_HIP, "segment graph: path point %d on no chain", t);
                const SegChain c = g.chains[ch];
                if (t != c.e0 && t != c.e1)
                    return fail(BGT_ERR_HIP, "segment graph: chain %d entered inside", ch);
                const int64_t *cp = g.chain_xyz + c.off;
                int e, x;
                int64_t exyz, xxyz;
                if (c.len == 1) {
                    emit(cp[0]);
                    e = t;
                    exyz = cp[0];
                    const bool from0 = idx == c.x0;
                    x = from0 ? c.x1 : c.x0;
                    xxyz = from0 ? c.x1xyz : c.x0xyz;
                } else if (t == c.e0) {
                    for (int q = 0; q < c.len; ++q) emit(cp[q]);
                    e = c.e1;
                    exyz = cp[c.len - 1];
                    x = c.x1;
                    xxyz = c.x1xyz;
                } else {
                    for (int q = c.len - 1; q >= 0; --q) emit(cp[q]);
                    e = c.e0;
                    exyz = cp[0];
                    x = c.x0;
                    xxyz = c.x0xyz;
                }
                code[c.e0] = 0;
                code[c.e1] = 0;
                if (live(x)) {
                    t = x;
                    txyz = xxyz;
                } else {
                    t = e;
                    txyz = exyz;
                }
                if (code[t] == 2) return fail(BGT_ERR_HIP, "segment graph: chain %d leaves to a path point", ch);
            }
            emit(txyz);
            if (std::abs((int)orig[t]) < 2) end = true;
            code[t] = (int8_t)(code[t] + 1);
            close(end);
        }
    }
    out->nseg = (int)ns;
    out->npts = (int)np;
    return BGT_OK;
}

template <typename T>
void to_mask_nonzero(const T *in, size_t n, uint8_t *m)
{
    for (size_t i = 0; i < n; ++i) m[i] = in[i] != 0;
}

int mask_from(int dtype, const void *in, size_t n, std::vector<uint8_t> &m)
{
    m.resize(n);
    switch (dtype) {
    case BGT_U8: case BGT_BOOL: to_mask_nonzero((const uint8_t *)in, n, m.data()); break;
    case BGT_I8: to_mask_nonzero((const int8_t *)in, n, m.data()); break;
    case BGT_U16: to_mask_nonzero((const uint16_t *)in, n, m.data()); break;
    case BGT_I16: to_mask_nonzero((const int16_t *)in, n, m.data()); break;
    case BGT_U32: to_mask_nonzero((const uint32_t *)in, n, m.data()); break;
    case BGT_I32: to_mask_nonzero((const int32_t *)in, n, m.data()); break;
    case BGT_F32: to_mask_nonzero((const float *)in, n, m.data()); break;
    case BGT_F64: to_mask_nonzero((const double *)in, n, m.data()); break;
    default: return fail(BGT_ERR_INVALID, "unknown pixel type %d", dtype);
    }
    return BGT_OK;
}

template <typename V>
V *dup(const std::vector<V> &v)
{
    V *p = (V *)std::malloc(sizeof(V) * (v.size() ? v.size() : 1));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(V) * v.size());
    return p;
}

int export_segments(const Flat &f, bool three_d, bgt_segments *out)
{
    out->nseg = (int)f.is_end.size();
    out->npts = (int)f.x.size();
    out->x = dup(f.x);
    out->y = dup(f.y);
    out->z = three_d ? dup(f.z) : nullptr;
    out->seg_start = dup(f.start);
    out->is_end = dup(f.is_end);
    if (!out->x || !out->y || (three_d && !out->z) || !out->seg_start || !out->is_end)
        return fail(BGT_ERR_NOMEM, "segment list allocation failed");
    return BGT_OK;
}

}  // namespace

int extract_segments_mask(const uint8_t *mask, int w, int h, int d, bool three_d,
                          bgt_segments *out)
{
    std::memset(out, 0, sizeof(*out));
    Flat f;
    if (w > 0 && h > 0 && d > 0) trace(mask, w, h, d, three_d, f);
    return export_segments(f, three_d, out);
}

int extract_segments_graph(const SegGraph &g, int pw, int ph, bool three_d, bgt_segments *out)
{
    std::memset(out, 0, sizeof(*out));
    (void)pw;
    (void)ph;
    const int r = three_d ? trace_graph<true>(g, out) : trace_graph<false>(g, out);
    if (r != BGT_OK) {
        std::free(out->x);
        std::free(out->y);
        std::free(out->z);
        std::free(out->seg_start);
        std::free(out->is_end);
        std::memset(out, 0, sizeof(*out));
    }
    return r;
}

int extract_segments(int dtype, const void *data, int w, int h, int d, bool three_d,
                     bgt_segments *out)
{
    if (!data || !out) return fail(BGT_ERR_INVALID, "null pointer");
    if (w <= 0 || h <= 0 || d <= 0) return fail(BGT_ERR_INVALID, "bad size");
    if (w > 32767 || h > 32767 || d > 32767)
        return fail(BGT_ERR_UNSUPPORTED, "coordinates are int16 (sides <= 32767)");
    std::vector<uint8_t> m;
    BGT_CHECK(mask_from(dtype, data, (size_t)w * h * d, m));
    return extract_segments_mask(m.data(), w, h, d, three_d, out);
}

// ---- prune, smooth, flatten (2D/CenterlineExtraction.cpp:51-131) ----------

namespace {

// A few host threads kept for the process's lifetime, so that a call's
// parallel section costs a wake-up rather than thread creation.  run(n, fn)
// calls fn(i) for i in [0, n) on the pool and the calling thread, and returns
// when all have finished.  Calls are serialised (one parallel section at a
// time).
class HostPool {
  public:
    static HostPool &get()
    {
        static HostPool p;
        return p;
    }
    int threads() const { return (int)workers_.size() + 1; }
    void run(int n, const std::function<void(int)> &fn)
    {
        std::lock_guard<std::mutex> one(call_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            busy_ = (int)workers_.size();
            ++gen_;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return busy_ == 0; });
        fn_ = nullptr;
    }
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }

  private:
    HostPool()
    {
        const unsigned hc = std::thread::hardware_concurrency();
        const int nw = (int)std::min<unsigned>(7u, hc > 1 ? hc - 1 : 0u);
        for (int k = 0; k < nw; ++k) workers_.emplace_back([this] { loop(); });
    }
    void drain()
    {
        for (int i; (i = next_.fetch_add(1)) < n_;) (*fn_)(i);
    }
    void loop()
    {
        unsigned seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
            }
            drain();
            std::lock_guard<std::mutex> lk(mu_);
            if (--busy_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_, call_mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *fn_ = nullptr;
    std::atomic<int> next_{0};
    int n_ = 0, busy_ = 0;
    unsigned gen_ = 0;
    bool stop_ = false;
};

// The moving average of one segment (2D/CenterlineExtraction.cpp:60-104):
// a window that grows from the first point, slides, then shrinks into the
// last point; float32 running sums in the reference's order of additions and
// subtractions.  Writes n values of each coordinate from X, Y (, Z).
void smooth_segment(const int16_t *px, const int16_t *py, const int16_t *pz, int n, int span_cap,
                    float *X, float *Y, float *Z)
{
    const int span = std::min(span_cap, (n - 1) / 2);
    float ax = px[0], ay = py[0], az = pz ? pz[0] : 0.0f;
    int cnt = 1, o = 0;
    X[o] = (float)px[0]; Y[o] = (float)py[0]; if (Z) Z[o] = (float)pz[0];
    ++o;
    for (int j = 1; j <= span; ++j) {
        ax += px[2 * j - 1]; ay += py[2 * j - 1]; if (pz) az += pz[2 * j - 1];
        ax += px[2 * j];     ay += py[2 * j];     if (pz) az += pz[2 * j];
        cnt += 2;
        X[o] = ax / (float)cnt; Y[o] = ay / (float)cnt; if (Z) Z[o] = az / (float)cnt;
        ++o;
    }
    for (int j = span + 1; j < n - span; ++j) {
        ax += px[j + span]; ay += py[j + span]; if (pz) az += pz[j + span];
        ax -= px[j - span - 1]; ay -= py[j - span - 1]; if (pz) az -= pz[j - span - 1];
        X[o] = ax / (float)cnt; Y[o] = ay / (float)cnt; if (Z) Z[o] = az / (float)cnt;
        ++o;
    }
    for (int j = n - span, dd = span; j < n; ++j, --dd) {
        ax -= px[j - dd - 1]; ay -= py[j - dd - 1]; if (pz) az -= pz[j - dd - 1];
        ax -= px[j - dd];     ay -= py[j - dd];     if (pz) az -= pz[j - dd];
        cnt -= 2;
        X[o] = ax / (float)cnt; Y[o] = ay / (float)cnt; if (Z) Z[o] = az / (float)cnt;
        ++o;
    }
}

// points at or above which the segments are smoothed on the host pool
constexpr size_t kParallelPoints = size_t(1) << 14;

}  // namespace

int centerline_from_segments(const bgt_segments &s, bool three_d, int max_segment_length,
                             int smoothing_span, float **ox, float **oy, float **oz,
                             uint32_t **oidx, uint32_t *nsegs)
{
    const size_t lim = (size_t)(long)max_segment_length;  // negative -> huge, as in the reference
    const int span_cap = smoothing_span < 0 ? 0 : smoothing_span;
    // prune (2D/CenterlineExtraction.cpp:51-57): short end segments go
    std::vector<int> keep;
    size_t npts = 0;
    for (int i = 0; i < s.nseg; ++i) {
        const size_t sz = s.seg_start[i + 1] - s.seg_start[i];
        if (sz <= lim && s.is_end[i]) continue;  // pruned
        keep.push_back(i);
        npts += sz;
    }
    float *X = (float *)std::malloc(sizeof(float) * (npts ? npts : 1));
    float *Y = (float *)std::malloc(sizeof(float) * (npts ? npts : 1));
    float *Z = three_d ? (float *)std::malloc(sizeof(float) * (npts ? npts : 1)) : nullptr;
    uint32_t *I = (uint32_t *)std::malloc(sizeof(uint32_t) * (keep.size() + 1));
    if (!X || !Y || (three_d && !Z) || !I) {
        std::free(X); std::free(Y); std::free(Z); std::free(I);
        return fail(BGT_ERR_NOMEM, "centerline output allocation failed");
    }
    // flatten (:107-131): each kept segment's points start where the previous
    // one's end, so the segments are independent and smoothed in parallel
    size_t o = 0;
    for (size_t k = 0; k < keep.size(); ++k) {
        I[k] = (uint32_t)o;
        o += s.seg_start[keep[k] + 1] - s.seg_start[keep[k]];
    }
    I[keep.size()] = (uint32_t)npts;
    auto one = [&](size_t k) {
        const int sidx = keep[k];
        const int st = (int)s.seg_start[sidx];
        const int n = (int)(s.seg_start[sidx + 1] - s.seg_start[sidx]);
        smooth_segment(s.x + st, s.y + st, three_d ? s.z + st : nullptr, n, span_cap, X + I[k],
                       Y + I[k], Z ? Z + I[k] : nullptr);
    };
    HostPool &pool = HostPool::get();
    if (npts >= kParallelPoints && pool.threads() > 1) {
        // chunks of about npts / (4 threads) points, whole segments each
        const int nchunks = pool.threads() * 4;
        std::vector<size_t> bounds(1, 0);
        const size_t per = npts / (size_t)nchunks + 1;
        for (size_t k = 0; k < keep.size(); ++k)
            if (I[k + 1] >= per * bounds.size()) bounds.push_back(k + 1);
        if (bounds.back() != keep.size()) bounds.push_back(keep.size());
        pool.run((int)bounds.size() - 1, [&](int c) {
            for (size_t k = bounds[c]; k < bounds[c + 1]; ++k) one(k);
        });
    } else {
        for (size_t k = 0; k < keep.size(); ++k) one(k);
    }
    *ox = X; *oy = Y;
    if (oz) *oz = Z;
    *oidx = I;
    *nsegs = (uint32_t)keep.size();
    return BGT_OK;
}

int mex_columns(const float *x, const float *y, const float *z, const uint32_t *idx,
                uint32_t nsegs, double *out)
{
    if (!x || !y || !idx || (!out && idx[nsegs] > 0)) return fail(BGT_ERR_INVALID, "null pointer");
    const uint32_t nc = z ? 3 : 2;
    for (uint32_t i = 0; i < nsegs; ++i) {
        if (idx[i + 1] < idx[i]) return fail(BGT_ERR_INVALID, "segment starts are not ascending");
        const uint32_t n = idx[i + 1] - idx[i];
        double *blk = out + (size_t)idx[i] * nc;  // this segment's n x nc matrix, column-major
        for (uint32_t t = 0; t < n; ++t) {
            blk[t] = (double)y[idx[i] + t];
            blk[n + t] = (double)x[idx[i] + t];
            if (z) blk[2 * n + t] = (double)z[idx[i] + t];
        }
    }
    return BGT_OK;
}

}  // namespace bgt
