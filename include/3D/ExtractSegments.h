// ExtractSegments.h (3D) -- drop-in for the reference's 3D/ExtractSegments.h:21-25.
//
// ExtractSegmentsV: traces a thinned volume (nonzero = skeleton) into segments
// in the reference's raster order with its junction / end-point rules over
// 26 neighbours (3D/ExtractSegments.cpp:135-251); the caller deletes every
// SSegment.  ExtractSegments: the same segments as point lists
// (3D/ExtractSegments.cpp:15-128 follows the same steps without the end
// flags).  Voxels on the volume border are not traced (as in the reference).
// On failure: an empty result, message in bgt_last_error().
#pragma once

#include <vector>

#include "../bgt_types.hpp"
#include "Structs.h"

namespace bgt_dropin {
template <typename PixelType, typename F>
inline bool trace3d(PixelType *data, int width, int height, int depth, F &&emit)
{
    bgt_segments s;
    if (bgt_extract_segments3d(dtype_of<PixelType>::value, data, width, height, depth, &s) != BGT_OK)
        return false;
    for (int i = 0; i < s.nseg; ++i) {
        std::vector<point3D> pts;
        pts.reserve(s.seg_start[i + 1] - s.seg_start[i]);
        for (uint32_t k = s.seg_start[i]; k < s.seg_start[i + 1]; ++k)
            pts.push_back(point3D{s.x[k], s.y[k], s.z[k]});
        emit(std::move(pts), s.is_end[i] != 0);
    }
    bgt_segments_free(&s);
    return true;
}
}  // namespace bgt_dropin

template <typename PixelType>
std::vector<std::vector<point3D>> ExtractSegments(PixelType *data, int width, int height, int depth)
{
    std::vector<std::vector<point3D>> out;
    bgt_dropin::trace3d(data, width, height, depth,
                        [&](std::vector<point3D> &&pts, bool) { out.push_back(std::move(pts)); });
    return out;
}

template <typename PixelType>
std::vector<SSegment *> ExtractSegmentsV(PixelType *data, int width, int height, int depth)
{
    std::vector<SSegment *> out;
    bgt_dropin::trace3d(data, width, height, depth, [&](std::vector<point3D> &&pts, bool end) {
        SSegment *seg = new SSegment();
        seg->points = std::move(pts);
        seg->is_end_segment = end;
        out.push_back(seg);
    });
    return out;
}
