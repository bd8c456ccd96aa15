// harness3d.cpp -- test-only driver around the reference's own 3D tracer
// (3D/ExtractSegments.cpp, compiled unmodified from the reference tree).
// Never linked into the product.
#include <cstdint>
#include <cstring>
#include <vector>
#include "ExtractSegments.h"

extern "C" int ref3d_extract(const uint8_t *vol, int w, int h, int d, int cap_pts, int cap_segs,
                             int16_t *xs, int16_t *ys, int16_t *zs, uint32_t *starts,
                             uint8_t *is_end, int *npts_out)
{
    std::vector<unsigned char> buf(vol, vol + (size_t)w * h * d);
    std::vector<SSegment *> segs = ExtractSegmentsV<unsigned char>(buf.data(), w, h, d);
    int np = 0, ns = 0, ok = 1;
    if (starts && cap_segs >= 0) starts[0] = 0;
    for (SSegment *s : segs) {
        for (const point3D &p : s->points) {
            if (np < cap_pts) { xs[np] = p.x; ys[np] = p.y; zs[np] = p.z; } else ok = 0;
            np++;
        }
        if (ns < cap_segs) { is_end[ns] = s->is_end_segment; starts[ns + 1] = (uint32_t)np; } else ok = 0;
        ns++;
        delete s;
    }
    *npts_out = np;
    return ok ? ns : -ns - 1;
}

// The reference's ExtractSegments (3D/ExtractSegments.cpp:15-128): point lists only.
extern "C" int ref3d_extract_plain(const uint8_t *vol, int w, int h, int d, int cap_pts,
                                   int cap_segs, int16_t *xs, int16_t *ys, int16_t *zs,
                                   uint32_t *starts, int *npts_out)
{
    std::vector<unsigned char> buf(vol, vol + (size_t)w * h * d);
    std::vector<std::vector<point3D>> segs = ExtractSegments<unsigned char>(buf.data(), w, h, d);
    int np = 0, ns = 0, ok = 1;
    if (starts && cap_segs >= 0) starts[0] = 0;
    for (const auto &s : segs) {
        for (const point3D &p : s) {
            if (np < cap_pts) { xs[np] = p.x; ys[np] = p.y; zs[np] = p.z; } else ok = 0;
            np++;
        }
        if (ns < cap_segs) starts[ns + 1] = (uint32_t)np; else ok = 0;
        ns++;
    }
    *npts_out = np;
    return ok ? ns : -ns - 1;
}
