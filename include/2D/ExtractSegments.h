// ExtractSegments.h (2D) -- drop-in for the reference's 2D/ExtractSegments.h:16-17.
//
// template<PixelType> std::vector<SSegment*> ExtractSegmentsV(PixelType *data,
//                                                             int width, int height)
// Traces a thinned image (nonzero = skeleton) into segments in the reference's
// raster order with its junction / end-point rules (2D/ExtractSegments.cpp:
// 16-127).  The caller deletes every SSegment.  On failure: an empty vector,
// message in bgt_last_error().
#pragma once

#include <vector>

#include "../bgt_types.hpp"
#include "Structs.h"

template <typename PixelType>
std::vector<SSegment *> ExtractSegmentsV(PixelType *data, int width, int height)
{
    std::vector<SSegment *> out;
    bgt_segments s;
    if (bgt_extract_segments2d(bgt_dropin::dtype_of<PixelType>::value, data, width, height, &s) !=
        BGT_OK)
        return out;
    out.reserve(s.nseg);
    for (int i = 0; i < s.nseg; ++i) {
        SSegment *seg = new SSegment();
        seg->is_end_segment = s.is_end[i] != 0;
        for (uint32_t k = s.seg_start[i]; k < s.seg_start[i + 1]; ++k)
            seg->points.push_back(point2D{s.x[k], s.y[k]});
        out.push_back(seg);
    }
    bgt_segments_free(&s);
    return out;
}
