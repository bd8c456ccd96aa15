// LibThinning2D.h -- drop-in for the reference's 2D/LibThinning2D.h:13-15.
//
// template<PixelType> void Thinning2D(int width, int height, PixelType *data,
//                                     PixelType *out)
// Host pointers, width x height row-major.  Foreground iff (unsigned)v > 0
// (floats saturate: v >= 1); out receives 0/1.  Any width/height >= 1 (the
// reference needs padded widths <= 2048).  The thinning runs on the MI355X
// (bgt_thinning2d).  On failure out is zero-filled and bgt_last_error() holds
// the message (the reference has no error channel).
#pragma once

#include <cstring>

#include "../bgt_types.hpp"

template <typename PixelType>
void Thinning2D(int width, int height, PixelType *data, PixelType *out)
{
    if (bgt_thinning2d(bgt_dropin::dtype_of<PixelType>::value, width, height, data, out) != BGT_OK &&
        out && width > 0 && height > 0)
        std::memset(static_cast<void *>(out), 0, (size_t)width * height * sizeof(PixelType));
}
