// CenterlineExtraction.h (2D) -- drop-in for the reference's
// 2D/CenterlineExtraction.h:25-28.
//
// Extracts centerlines from a binary image (data: width x height, host
// memory): pads, thins on the MI355X, traces segments, prunes end segments of
// at most max_segment_length points, smooths with a moving average of
// smoothing_span points (float32, the reference's order of operations) and
// flattens.  Returns the number of segments; out_x/out_y hold all points and
// out_idx the segment starts with out_idx[nsegs] = number of points.  The
// arrays are allocated with new[]; the caller delete[]s them.  On failure:
// returns 0 with out_idx = {0} and empty out_x/out_y, message in
// bgt_last_error() (the reference has no error channel).
#pragma once

#include "../bgt_types.hpp"

template <typename PixelType>
unsigned int CenterlineExtraction(PixelType *data, int width, int height, int max_segment_length,
                                  int smoothing_span, float **out_x, float **out_y,
                                  unsigned int **out_idx)
{
    float *x = nullptr, *y = nullptr;
    uint32_t *idx = nullptr, nsegs = 0;
    if (bgt_centerline2d(bgt_dropin::dtype_of<PixelType>::value, data, width, height,
                         max_segment_length, smoothing_span, &x, &y, &idx, &nsegs) != BGT_OK) {
        *out_x = new float[0];
        *out_y = new float[0];
        *out_idx = new unsigned int[1]{0};
        return 0;
    }
    const size_t npts = idx[nsegs];
    *out_x = bgt_dropin::adopt<float>(x, npts);
    *out_y = bgt_dropin::adopt<float>(y, npts);
    *out_idx = bgt_dropin::adopt<unsigned int>(idx, (size_t)nsegs + 1);
    return nsegs;
}
