// Host code of the thinning path under AddressSanitizer and
// UndefinedBehaviorSanitizer: the library's tracer, centerline
// post-processing and MEX column layout (begputhinning_amd/csrc/segments.cpp)
// and the oracle (oracle/oracle.c), compiled together with
// -fsanitize=address,undefined into one program that runs without a GPU.
// The thinning itself comes from the oracle; every result of the library's
// host code is compared with the oracle's on seeded images and volumes:
//   - ExtractSegmentsV of the skeleton and of raw (unthinned) masks, also
//     through the typed entry (float pixels; negative int16 pixels, which the
//     tracer takes as set: 2D/ExtractSegments.cpp:36 tests == 0);
//   - CenterlineExtraction's prune/smooth/flatten of the padded frame's
//     segments, float bits included, for several (max_segment_length,
//     smoothing_span), against or2d/or3d_centerline;
//   - the MEX [y x (z)] columns, against a direct column-major copy;
//   - the tracer's second pass over the device pre-pass's graph (codes,
//     nodes with their neighbour ranks, chains of path points, built here on
//     the host), on the same masks and on dense noise.
// Any sanitizer report aborts the program (-fno-sanitize-recover).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../begputhinning_amd/csrc/seggraph.h"
#include "../../begputhinning_amd/csrc/segments.h"
#include "../../oracle/oracle.h"

// the library's error channel, as the tracer sees it (capi.cpp is not linked)
namespace bgt {
int fail(int code, const char *fmt, ...) { (void)fmt; return code; }
}  // namespace bgt

static int failures = 0;
#define EXPECT(cond, ...)                                     \
    do {                                                      \
        if (!(cond)) {                                        \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);                \
            std::fprintf(stderr, "\n");                       \
            ++failures;                                       \
        }                                                     \
    } while (0)

static bool same_segments(const bgt_segments &a, const or_segments &b, bool three_d)
{
    if (a.nseg != b.nseg || a.npts != b.npts) return false;
    const size_t n = (size_t)a.npts;
    if (n && (std::memcmp(a.x, b.x, n * 2) || std::memcmp(a.y, b.y, n * 2))) return false;
    if (three_d && n && std::memcmp(a.z, b.z, n * 2)) return false;
    if (std::memcmp(a.seg_start, b.seg_start, (size_t)(a.nseg + 1) * 4)) return false;
    return !a.nseg || !std::memcmp(a.is_end, b.is_end, (size_t)a.nseg);
}

static std::vector<uint8_t> random_mask(std::mt19937_64 &g, size_t n, double p)
{
    std::bernoulli_distribution b(p);
    std::vector<uint8_t> m(n);
    for (auto &v : m) v = b(g) ? 1 : 0;
    return m;
}

// blobs (discs / balls) on zero, so that thinning leaves branching skeletons
static std::vector<uint8_t> blobs(std::mt19937_64 &g, int w, int h, int d, int count)
{
    std::vector<uint8_t> m((size_t)w * h * d, 0);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (int k = 0; k < count; ++k) {
        const double cx = u(g) * w, cy = u(g) * h, cz = u(g) * d, r = 1.0 + u(g) * 6.0;
        for (int z = 0; z < d; ++z)
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x) {
                    const double dx = x - cx, dy = y - cy, dz = d > 1 ? z - cz : 0.0;
                    if (dx * dx + dy * dy + dz * dz <= r * r) m[((size_t)z * h + y) * w + x] = 1;
                }
    }
    return m;
}

static void check_segments(const std::vector<uint8_t> &m, int w, int h, int d, bool three_d,
                           const char *what)
{
    bgt_segments s{};
    or_segments o{};
    const int rc = bgt::extract_segments_mask(m.data(), w, h, d, three_d, &s);
    EXPECT(rc == BGT_OK, "%s: extract rc %d", what, rc);
    if (three_d)
        or3d_extract(m.data(), w, h, d, &o);
    else
        or2d_extract(m.data(), w, h, &o);
    EXPECT(same_segments(s, o, three_d), "%s %dx%dx%d: segments differ", what, w, h, d);
    or_segments_free(&o);
    // the typed entry: float, int16 and bool pixels of the same mask
    {
        std::vector<float> f(m.begin(), m.end());
        bgt_segments t{};
        EXPECT(bgt::extract_segments(BGT_F32, f.data(), w, h, d, three_d, &t) == BGT_OK, "%s f32", what);
        EXPECT(t.nseg == s.nseg && t.npts == s.npts, "%s: float pixels trace differently", what);
        std::free(t.x); std::free(t.y); std::free(t.z); std::free(t.seg_start); std::free(t.is_end);
        std::vector<int16_t> i16(m.size());
        for (size_t k = 0; k < m.size(); ++k) i16[k] = m[k] ? -3 : 0;  // the tracer: any nonzero
        bgt_segments u{};
        EXPECT(bgt::extract_segments(BGT_I16, i16.data(), w, h, d, three_d, &u) == BGT_OK, "%s i16", what);
        EXPECT(u.nseg == s.nseg && u.npts == s.npts &&
                   (!u.npts || !std::memcmp(u.x, s.x, (size_t)u.npts * 2)),
               "%s: negative pixels trace differently", what);
        std::free(u.x); std::free(u.y); std::free(u.z); std::free(u.seg_start); std::free(u.is_end);
    }
    std::free(s.x); std::free(s.y); std::free(s.z); std::free(s.seg_start); std::free(s.is_end);
}

// The device pre-pass's graph (begputhinning_amd/csrc/seggraph.hip) built on
// the host for a mask whose frame is the mask itself: the set voxels off the
// frame's border in raster order with their codes (neighbours counted over
// the whole mask); the chains of path points, each walked here from its end
// of lower rank (the device ranks them by pointer jumping instead); and the
// nodes (negative codes) with their packed coordinates and their neighbours
// that are points (rank, slot, chain + 1).
struct HostGraph {
    std::vector<int64_t> idx, node_xyz, chain_xyz, node_ent;
    std::vector<int8_t> code;
    std::vector<int> nbr, chain, node_rank, node_off;
    std::vector<bgt::SegChain> chains;
    bgt::SegGraph g;
};

static void host_graph(const std::vector<uint8_t> &m, int w, int h, int d, bool three_d, HostGraph &G)
{
    const int nd = three_d ? 26 : 8;
    std::vector<int> rank(m.size(), -1);
    int n = 0;
    auto inner = [&](int x, int y, int z) {
        return x >= 1 && x < w - 1 && y >= 1 && y < h - 1 && (!three_d || (z >= 1 && z < d - 1));
    };
    for (int z = 0; z < d; ++z)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                if (m[((size_t)z * h + y) * w + x] && inner(x, y, z)) rank[((size_t)z * h + y) * w + x] = n++;
    G.idx.assign(n, 0);
    G.code.assign(n, 0);
    G.nbr.assign((size_t)n * nd, -1);
    for (int z = 0; z < d; ++z)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const int r = rank[((size_t)z * h + y) * w + x];
                if (r < 0) continue;
                G.idx[r] = (int64_t)x | (int64_t)y << 16 | (int64_t)z << 32;
                int nn = 0, j = 0;
                for (int dz = three_d ? -1 : 0; dz <= (three_d ? 1 : 0); ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            if (!dz && !dy && !dx) continue;
                            const size_t q = ((size_t)(z + dz) * h + (y + dy)) * w + (x + dx);
                            nn += m[q] != 0;
                            G.nbr[(size_t)r * nd + j++] = rank[q];
                        }
                if (nn == 0) nn = 1;
                G.code[r] = (int8_t)(nn == 2 ? 2 : -nn);
            }
    // the (at most two) point neighbours of a path point, in neighbour order
    auto slots = [&](int p, int *s) {
        int k = 0;
        s[0] = s[1] = -1;
        for (int j = 0; j < nd && k < 2; ++j)
            if (G.nbr[(size_t)p * nd + j] >= 0) s[k++] = G.nbr[(size_t)p * nd + j];
    };
    G.chain.assign(n, -1);
    G.chains.clear();
    G.chain_xyz.clear();
    for (int p = 0; p < n; ++p) {
        if (G.code[p] != 2 || G.chain[p] >= 0) continue;
        int s[2];
        slots(p, s);
        int on_path = 0;
        for (int v : s) on_path += v >= 0 && G.code[v] == 2;
        if (on_path == 2) continue;  // not an end: reached from one, or on a closed run
        bgt::SegChain c{};
        c.off = (int)G.chain_xyz.size();
        c.e0 = p;
        const int id = (int)G.chains.size();
        int off_chain[2] = {-1, -1}, k = 0;
        for (int v : s)
            if (v >= 0 && G.code[v] != 2) off_chain[k++] = v;
        int prev = -1, cur = p;
        for (;;) {
            G.chain[cur] = id;
            G.chain_xyz.push_back(G.idx[cur]);
            int t[2], next = -1;
            slots(cur, t);
            for (int v : t)
                if (v >= 0 && v != prev && G.code[v] == 2 && G.chain[v] < 0) { next = v; break; }
            if (next < 0) break;
            prev = cur;
            cur = next;
        }
        c.e1 = cur;
        c.len = (int)G.chain_xyz.size() - c.off;
        c.x0 = off_chain[0];
        if (c.len == 1) {
            c.x1 = off_chain[1];
        } else {
            int t[2];
            slots(cur, t);
            c.x1 = -1;
            for (int v : t)
                if (v >= 0 && G.code[v] != 2) { c.x1 = v; break; }
        }
        c.x0xyz = c.x0 >= 0 ? G.idx[c.x0] : 0;
        c.x1xyz = c.x1 >= 0 ? G.idx[c.x1] : 0;
        G.chains.push_back(c);
    }
    G.node_rank.clear();
    G.node_xyz.clear();
    G.node_off.assign(1, 0);
    G.node_ent.clear();
    for (int r = 0; r < n; ++r)
        if (G.code[r] < 0) {
            G.node_rank.push_back(r);
            G.node_xyz.push_back(G.idx[r]);
            for (int j = 0; j < nd; ++j) {
                const int v = G.nbr[(size_t)r * nd + j];
                if (v >= 0)
                    G.node_ent.push_back((int64_t)(uint32_t)v | (int64_t)j << 32 | (int64_t)(G.chain[v] + 1) << 37);
            }
            G.node_off.push_back((int)G.node_ent.size());
        }
    G.g = bgt::SegGraph{};
    G.g.npts = n;
    G.g.nd = nd;
    G.g.code = G.code.data();
    G.g.nnodes = (int)G.node_rank.size();
    G.g.node_rank = G.node_rank.data();
    G.g.node_xyz = G.node_xyz.data();
    G.g.node_off = G.node_off.data();
    G.g.node_ent = G.node_ent.data();
    G.g.nchains = (int)G.chains.size();
    G.g.chains = G.chains.data();
    G.g.chain_xyz = G.chain_xyz.data();
}

// the host's second pass over the pre-pass graph against the oracle's tracer
static void check_graph(const std::vector<uint8_t> &m, int w, int h, int d, bool three_d, const char *what)
{
    if (w < 1 || h < 1 || d < 1) return;
    HostGraph G;
    host_graph(m, w, h, d, three_d, G);
    or_segments o{};
    if (three_d)
        or3d_extract(m.data(), w, h, d, &o);
    else
        or2d_extract(m.data(), w, h, &o);
    bgt_segments s{};
    EXPECT(bgt::extract_segments_graph(G.g, w, h, three_d, &s) == BGT_OK, "%s: graph trace", what);
    EXPECT(same_segments(s, o, three_d), "%s %dx%dx%d: graph trace differs", what, w, h, d);
    std::free(s.x); std::free(s.y); std::free(s.z); std::free(s.seg_start); std::free(s.is_end);
    or_segments_free(&o);
}

static void check_centerline(const std::vector<uint8_t> &img, int w, int h, int d, bool three_d,
                             int maxlen, int span)
{
    // the padded frame of CenterlineExtraction, thinned by the oracle
    int pw, ph, pd;
    if (three_d) {
        pw = (w + 31) / 32 * 32; ph = (h + 31) / 32 * 32; pd = (d + 31) / 32 * 32;
    } else {
        const int q = (h + 63) / 128;
        pw = (w + 31) / 32 * 32; ph = (q > 1 ? q : 1) * 128 + 64; pd = 1;
    }
    std::vector<uint8_t> buf((size_t)pw * ph * pd, 0);
    for (int z = 0; z < d; ++z)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                buf[((size_t)z * ph + y) * pw + x] = img[((size_t)z * h + y) * w + x] != 0;
    if (three_d)
        or3d_thin(buf.data(), pw, ph, pd, 0);
    else
        or2d_thin(buf.data(), pw, ph, 0);
    bgt_segments s{};
    EXPECT(bgt::extract_segments_mask(buf.data(), pw, ph, pd, three_d, &s) == BGT_OK, "trace");
    float *x = nullptr, *y = nullptr, *z = nullptr;
    uint32_t *idx = nullptr, nsegs = 0;
    const int rc = bgt::centerline_from_segments(s, three_d, maxlen, span, &x, &y, three_d ? &z : nullptr,
                                                 &idx, &nsegs);
    EXPECT(rc == BGT_OK, "centerline rc %d", rc);
    or_centerline o{};
    if (three_d)
        or3d_centerline(img.data(), w, h, d, maxlen, span, &o);
    else
        or2d_centerline(img.data(), w, h, maxlen, span, &o);
    EXPECT((int)nsegs == o.nseg, "centerline %dx%dx%d (%d,%d): %u segments, oracle %d", w, h, d, maxlen,
           span, nsegs, o.nseg);
    if ((int)nsegs == o.nseg) {
        const uint32_t npts = idx[nsegs];
        EXPECT((int)npts == o.npts && !std::memcmp(idx, o.idx, (nsegs + 1) * 4), "centerline idx");
        if ((int)npts == o.npts) {
            EXPECT(!std::memcmp(x, o.x, npts * 4) && !std::memcmp(y, o.y, npts * 4), "centerline x/y bits");
            if (three_d) EXPECT(!std::memcmp(z, o.z, npts * 4), "centerline z bits");
            // MEX columns: per segment an n x 2 (n x 3) double matrix, column-major [y x (z)]
            std::vector<double> cols((size_t)npts * (three_d ? 3 : 2) + 1, -1.0);
            EXPECT(bgt::mex_columns(x, y, z, idx, nsegs, cols.data()) == BGT_OK, "mex columns");
            size_t at = 0;
            for (uint32_t k = 0; k < nsegs; ++k) {
                const uint32_t a = idx[k], n = idx[k + 1] - a;
                for (uint32_t i = 0; i < n; ++i) {
                    EXPECT(cols[at + i] == (double)y[a + i] && cols[at + n + i] == (double)x[a + i],
                           "mex column of segment %u", k);
                    if (three_d) EXPECT(cols[at + 2 * n + i] == (double)z[a + i], "mex z column");
                }
                at += (size_t)n * (three_d ? 3 : 2);
            }
            EXPECT(cols[at] == -1.0, "mex columns wrote past their end");
        }
    }
    std::free(x); std::free(y); std::free(z); std::free(idx);
    or_centerline_free(&o);
    std::free(s.x); std::free(s.y); std::free(s.z); std::free(s.seg_start); std::free(s.is_end);
}

int main(int argc, char **argv)
{
    const int rounds = argc > 1 ? std::atoi(argv[1]) : 24;
    std::mt19937_64 g(0xBE6E5A11);
    std::uniform_int_distribution<int> side2(1, 180), side3(1, 22);
    const int params[][2] = {{10, 3}, {0, 0}, {-1, 5}, {1000, 2}, {3, 1}};
    for (int r = 0; r < rounds; ++r) {
        // 2D: raw masks and thinned blob images
        const int w = side2(g), h = side2(g);
        auto raw = random_mask(g, (size_t)w * h, 0.3);
        check_segments(raw, w, h, 1, false, "2D raw");
        check_graph(raw, w, h, 1, false, "2D raw");
        auto img = blobs(g, w, h, 1, 1 + r % 9);
        auto skel = img;
        or2d_thin(skel.data(), w, h, 0);
        check_segments(skel, w, h, 1, false, "2D skeleton");
        check_graph(skel, w, h, 1, false, "2D skeleton");
        check_centerline(img, w, h, 1, false, params[r % 5][0], params[r % 5][1]);
        // 3D
        const int w3 = side3(g), h3 = side3(g), d3 = side3(g);
        auto raw3 = random_mask(g, (size_t)w3 * h3 * d3, 0.25);
        check_segments(raw3, w3, h3, d3, true, "3D raw");
        check_graph(raw3, w3, h3, d3, true, "3D raw");
        auto vol = blobs(g, w3, h3, d3, 1 + r % 5);
        auto skel3 = vol;
        or3d_thin(skel3.data(), w3, h3, d3, 0);
        check_segments(skel3, w3, h3, d3, true, "3D skeleton");
        check_graph(skel3, w3, h3, d3, true, "3D skeleton");
        check_centerline(vol, w3, h3, d3, true, params[(r + 2) % 5][0], params[(r + 2) % 5][1]);
    }
    // dense noise at many densities: junctions of every degree, nodes next to
    // nodes, chains closing on themselves or on one node, points next to the
    // border
    std::uniform_real_distribution<double> dens(0.05, 0.7);
    std::uniform_int_distribution<int> small(3, 40), small3(3, 12);
    for (int r = 0; r < rounds * 40; ++r) {
        const int w = small(g), h = small(g);
        auto m = random_mask(g, (size_t)w * h, dens(g));
        if (r & 1) or2d_thin(m.data(), w, h, 0);
        check_graph(m, w, h, 1, false, "2D noise");
        const int w3 = small3(g), h3 = small3(g), d3 = small3(g);
        auto m3 = random_mask(g, (size_t)w3 * h3 * d3, dens(g) * 0.6);
        if (r & 1) or3d_thin(m3.data(), w3, h3, d3, 0);
        check_graph(m3, w3, h3, d3, true, "3D noise");
    }
    // empty and one-pixel frames
    std::vector<uint8_t> one(1, 1), none(1, 0);
    check_segments(one, 1, 1, 1, false, "2D single");
    check_segments(none, 1, 1, 1, false, "2D empty");
    check_segments(one, 1, 1, 1, true, "3D single");
    check_centerline(one, 1, 1, 1, false, 10, 3);
    check_centerline(one, 1, 1, 1, true, 10, 3);
    std::printf("%s: %d rounds, %d failures\n", failures ? "FAILED" : "ok", rounds, failures);
    return failures ? 1 : 0;
}
