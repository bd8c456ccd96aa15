// harness2d.cpp -- test-only driver around the reference's own 2D tracer
// (2D/ExtractSegments.cpp, compiled unmodified from the reference tree).
// Exposes ExtractSegmentsV<unsigned char> through a C entry point so the
// oracle's restatement can be compared with it.  Never linked into the product.
#include <cstdint>
#include <cstring>
#include <vector>
#include "ExtractSegments.h"

extern "C" int ref2d_extract(const uint8_t *img, int w, int h, int cap_pts, int cap_segs,
                             int16_t *xs, int16_t *ys, uint32_t *starts, uint8_t *is_end,
                             int *npts_out)
{
    std::vector<unsigned char> buf(img, img + (size_t)w * h);
    std::vector<SSegment *> segs = ExtractSegmentsV<unsigned char>(buf.data(), w, h);
    int np = 0, ns = 0, ok = 1;
    if (starts && cap_segs >= 0) starts[0] = 0;
    for (SSegment *s : segs) {
        for (const point2D &p : s->points) {
            if (np < cap_pts) { xs[np] = p.x; ys[np] = p.y; } else ok = 0;
            np++;
        }
        if (ns < cap_segs) { is_end[ns] = s->is_end_segment; starts[ns + 1] = (uint32_t)np; } else ok = 0;
        ns++;
        delete s;
    }
    *npts_out = np;
    return ok ? ns : -ns - 1;
}
