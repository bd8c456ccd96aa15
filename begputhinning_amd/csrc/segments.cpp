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
#include <cstdint>
#include <cstdlib>
#include <cstring>
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
void trace_graph(const SegGraph &g, int pw, int ph, bool three_d, Flat &out)
{
    const int nd = g.nd;
    const long sl = (long)pw * ph;
    std::vector<int8_t> code(g.code), orig(g.code);
    auto emit = [&](int r) {
        const long q = (long)g.idx[r];
        const int zz = (int)(q / sl);
        const long rr = q - (long)zz * sl;
        out.x.push_back((int16_t)(rr % pw));
        out.y.push_back((int16_t)(rr / pw));
        if (three_d) out.z.push_back((int16_t)zz);
    };
    auto close = [&](bool end) {
        out.is_end.push_back(end ? 1 : 0);
        out.start.push_back((uint32_t)out.x.size());
    };
    auto live = [&](int r) { return r >= 0 && code[r] != 0; };
    for (int idx = 0; idx < g.npts; ++idx) {
        const int8_t nodes = code[idx];
        if (nodes >= 0) continue;
        code[idx] = 0;
        const int *nb = &g.nbr[(size_t)idx * nd];
        int ni = -1;
        for (int8_t si = 0; si > nodes; --si) {
            bool end = std::abs((int)orig[idx]) < 2;
            emit(idx);
            ++ni;
            while (ni < nd && !live(nb[ni])) ++ni;
            if (ni >= nd) { close(end); continue; }
            int t = nb[ni];
            while (code[t] == 2) {
                code[t] = 0;
                emit(t);
                const int *nt = &g.nbr[(size_t)t * nd];
                for (int k = 0; k < nd; ++k)
                    if (live(nt[k])) { t = nt[k]; break; }
            }
            emit(t);
            if (std::abs((int)orig[t]) < 2) end = true;
            code[t] = (int8_t)(code[t] + 1);
            close(end);
        }
    }
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
    Flat f;
    trace_graph(g, pw, ph, three_d, f);
    return export_segments(f, three_d, out);
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

int centerline_from_segments(const bgt_segments &s, bool three_d, int max_segment_length,
                             int smoothing_span, float **ox, float **oy, float **oz,
                             uint32_t **oidx, uint32_t *nsegs)
{
    const size_t lim = (size_t)(long)max_segment_length;  // negative -> huge, as in the reference
    const int span_cap = smoothing_span < 0 ? 0 : smoothing_span;
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
    size_t o = 0;
    for (size_t k = 0; k < keep.size(); ++k) {
        const int sidx = keep[k];
        const int st = (int)s.seg_start[sidx];
        const int n = (int)(s.seg_start[sidx + 1] - s.seg_start[sidx]);
        const int16_t *px = s.x + st, *py = s.y + st, *pz = three_d ? s.z + st : nullptr;
        I[k] = (uint32_t)o;
        // moving average with a window that grows from the first point, slides,
        // then shrinks into the last point; float32 running sums in the
        // reference's order of additions and subtractions
        const int span = std::min(span_cap, (n - 1) / 2);
        float ax = px[0], ay = py[0], az = pz ? pz[0] : 0.0f;
        int cnt = 1;
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
    I[keep.size()] = (uint32_t)npts;
    *ox = X; *oy = Y;
    if (oz) *oz = Z;
    *oidx = I;
    *nsegs = (uint32_t)keep.size();
    return BGT_OK;
}

}  // namespace bgt
