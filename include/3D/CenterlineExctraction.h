// CenterlineExctraction.h (3D; the reference's file name, misspelling kept so
// existing #includes resolve) -- drop-in for 3D/CenterlineExctraction.h:26-29.
// CenterlineExtraction.h in this directory is the correctly spelt alias.
//
// Extracts centerlines from a binary volume (data: width x height x depth,
// host memory; foreground iff int(v) > 0): pads, thins on the MI355X, traces
// segments, prunes end segments of at most max_segment_length points, smooths
// with a moving average of smoothing_span points (float32, the reference's
// order of operations) and flattens.  Returns the number of segments;
// out_x/out_y/out_z hold all points and out_idx the segment starts with
// out_idx[nsegs] = number of points.  The arrays are allocated with new[];
// the caller delete[]s them.  On failure: returns 0 with out_idx = {0} and
// empty coordinate arrays, message in bgt_last_error().
#pragma once

#include "../bgt_types.hpp"

template <typename PixelType>
unsigned int CenterlineExtraction(PixelType *data, int width, int height, int depth,
                                  int max_segment_length, int smoothing_span, float **out_x,
                                  float **out_y, float **out_z, unsigned int **out_idx)
{
    float *x = nullptr, *y = nullptr, *z = nullptr;
    uint32_t *idx = nullptr, nsegs = 0;
    if (bgt_centerline3d(bgt_dropin::dtype_of<PixelType>::value, data, width, height, depth,
                         max_segment_length, smoothing_span, &x, &y, &z, &idx, &nsegs) != BGT_OK) {
        *out_x = new float[0];
        *out_y = new float[0];
        *out_z = new float[0];
        *out_idx = new unsigned int[1]{0};
        return 0;
    }
    const size_t npts = idx[nsegs];
    *out_x = bgt_dropin::adopt<float>(x, npts);
    *out_y = bgt_dropin::adopt<float>(y, npts);
    *out_z = bgt_dropin::adopt<float>(z, npts);
    *out_idx = bgt_dropin::adopt<unsigned int>(idx, (size_t)nsegs + 1);
    return nsegs;
}
