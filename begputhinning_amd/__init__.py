"""MI355X-native (gfx950, HIP) implementation of BEGPUThinning's 2D/3D
thinning path behind the reference's API.  See include/bgt.h for the C-ABI."""
from ._lib import BGTError, lib  # noqa: F401
from .thinning import (CenterlineExtraction, ExtractSegmentsV, SSegment, Thinning2D,  # noqa: F401
                       debug_fail_device_allocs,
                       Thinning3D, extract_segments_flat, mex_columns, extract_segments_device, set_config2d,
                       set_config3d, set_engine3d, set_fused_io2d, set_lane_words2d, set_persistent2d,
                       thin2d_device, thin3d_device)
