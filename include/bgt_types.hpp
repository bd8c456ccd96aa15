// bgt_types.hpp -- C++ helpers shared by the drop-in headers in include/2D and
// include/3D: the pixel type -> BGT_* code map and the new[]-owned copies the
// reference's callers delete[].
#ifndef BGT_TYPES_HPP
#define BGT_TYPES_HPP

#include <cstring>
#include <limits>
#include <type_traits>

#include "bgt.h"

namespace bgt_dropin {

// Pixel types of the reference's instantiations (2D/LibThinning2D.cu:134-143,
// 2D/CenterlineExtraction.cpp:146-155); anything else does not compile.
template <typename T>
struct dtype_of {
    static_assert(sizeof(T) == 0, "pixel type not supported by the thinning library");
};
template <> struct dtype_of<unsigned char> { static constexpr int value = BGT_U8; };
template <> struct dtype_of<signed char> { static constexpr int value = BGT_I8; };
template <> struct dtype_of<char> {
    static constexpr int value = std::numeric_limits<char>::is_signed ? BGT_I8 : BGT_U8;
};
template <> struct dtype_of<unsigned short> { static constexpr int value = BGT_U16; };
template <> struct dtype_of<short> { static constexpr int value = BGT_I16; };
template <> struct dtype_of<unsigned int> { static constexpr int value = BGT_U32; };
template <> struct dtype_of<int> { static constexpr int value = BGT_I32; };
template <> struct dtype_of<float> { static constexpr int value = BGT_F32; };
template <> struct dtype_of<double> { static constexpr int value = BGT_F64; };
template <> struct dtype_of<bool> { static constexpr int value = BGT_BOOL; };

// new[] copy (as D) of n library-owned elements of S (released with
// bgt_free); the caller of the drop-in API owns the copy and delete[]s it.
template <typename D, typename S>
inline D *adopt(S *lib_owned, size_t n)
{
    static_assert(sizeof(D) == sizeof(S), "element sizes differ");
    D *r = new D[n];
    if (n && lib_owned) std::memcpy(r, lib_owned, n * sizeof(D));
    bgt_free(lib_owned);
    return r;
}

}  // namespace bgt_dropin

#endif
