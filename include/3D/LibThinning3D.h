// LibThinning3D.h -- drop-in for the reference's 3D/LibThinning3D.h:13-15.
//
// template<PixelType> void Thinning3D(int width, int height, int depth,
//                                     PixelType *data, PixelType *out)
// Host pointers, x fastest, then y, then z.  Foreground iff v > 0 in
// PixelType; out receives 0/1.  Any dimensions >= 1 (the reference needs
// multiples of 32, at least 64).  The thinning runs on the MI355X
// (bgt_thinning3d).  On failure out is zero-filled and bgt_last_error() holds
// the message (the reference waits for a key and exits the process instead).
#pragma once

#include <cstring>

#include "../bgt_types.hpp"

template <typename PixelType>
void Thinning3D(int width, int height, int depth, PixelType *data, PixelType *out)
{
    if (bgt_thinning3d(bgt_dropin::dtype_of<PixelType>::value, width, height, depth, data, out) !=
            BGT_OK &&
        out && width > 0 && height > 0 && depth > 0)
        std::memset(static_cast<void *>(out), 0,
                    (size_t)width * height * depth * sizeof(PixelType));
}
