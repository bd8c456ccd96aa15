// segments.h -- host tracer and centerline post-processing (internal).
#pragma once

#include <cstdint>

#include "../../include/bgt.h"

namespace bgt {

int extract_segments_mask(const uint8_t *mask, int w, int h, int d, bool three_d,
                          bgt_segments *out);
int extract_segments(int dtype, const void *data, int w, int h, int d, bool three_d,
                     bgt_segments *out);
struct SegGraph;
// segments from the device pre-pass's graph of a pw x ph (x pd) frame
int extract_segments_graph(const SegGraph &g, int pw, int ph, bool three_d, bgt_segments *out);
int centerline_from_segments(const bgt_segments &s, bool three_d, int max_segment_length,
                             int smoothing_span, float **ox, float **oy, float **oz,
                             uint32_t **oidx, uint32_t *nsegs);
// centerline arrays -> the MEX gateway's per-segment n x 2 [y x] (n x 3
// [y x z], z non-null) double matrices, column-major, one after another
int mex_columns(const float *x, const float *y, const float *z, const uint32_t *idx,
                uint32_t nsegs, double *out);

}  // namespace bgt
