// dropin_main.cpp -- test driver written against the drop-in C++ API only
// (include/2D or include/3D, selected by DIM), the way a user of the
// reference's headers calls it.  Built twice by the package build (DIM=2,
// DIM=3) and run by the tests:
//   thin <type> w h [d] in.raw out.raw       Thinning2D/3D<type>
//   trace w h [d] in.raw out.txt             ExtractSegmentsV<unsigned char>
//   trace_plain w h d in.raw out.txt         ExtractSegments<unsigned char> (3D)
//   centerline <type> w h [d] maxlen span in.raw out.txt
// Raw files hold the pixels in the named type, x fastest.  Text output: one
// line per segment ("E" or "-" for is_end_segment, then the points), or for
// centerline the segment count, the index array and the coordinates printed
// with %.9g.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#if DIM == 2
#include "2D/CenterlineExtraction.h"
#include "2D/ExtractSegments.h"
#include "2D/LibThinning2D.h"
#else
#include "3D/CenterlineExctraction.h"
#include "3D/ExtractSegments.h"
#include "3D/LibThinning3D.h"
#endif

static std::vector<char> read_file(const char *path)
{
    std::vector<char> b;
    FILE *f = std::fopen(path, "rb");
    if (!f) return b;
    char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) b.insert(b.end(), tmp, tmp + n);
    std::fclose(f);
    return b;
}

static bool write_file(const char *path, const void *p, size_t n)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, n, f) == n;
    return std::fclose(f) == 0 && ok;
}

struct Args {
    int w = 0, h = 0, d = 1;
    size_t n() const { return (size_t)w * h * d; }
};

template <typename T>
static int do_thin(const Args &a, const char *in, const char *out)
{
    std::vector<char> b = read_file(in);
    if (b.size() != a.n() * sizeof(T)) return 2;
    std::unique_ptr<T[]> src(new T[a.n()]), dst(new T[a.n()]);
    std::memcpy(static_cast<void *>(src.get()), b.data(), b.size());
#if DIM == 2
    Thinning2D<T>(a.w, a.h, src.get(), dst.get());
#else
    Thinning3D<T>(a.w, a.h, a.d, src.get(), dst.get());
#endif
    if (bgt_last_error()[0]) std::fprintf(stderr, "bgt: %s\n", bgt_last_error());
    return write_file(out, dst.get(), a.n() * sizeof(T)) ? 0 : 3;
}

template <typename T>
static int do_centerline(const Args &a, int maxlen, int span, const char *in, const char *out)
{
    std::vector<char> b = read_file(in);
    if (b.size() != a.n() * sizeof(T)) return 2;
    std::unique_ptr<T[]> src(new T[a.n()]);
    std::memcpy(static_cast<void *>(src.get()), b.data(), b.size());
    float *x = nullptr, *y = nullptr, *z = nullptr;
    unsigned int *idx = nullptr;
#if DIM == 2
    const unsigned ns = CenterlineExtraction<T>(src.get(), a.w, a.h, maxlen, span, &x, &y, &idx);
#else
    const unsigned ns =
        CenterlineExtraction<T>(src.get(), a.w, a.h, a.d, maxlen, span, &x, &y, &z, &idx);
#endif
    FILE *f = std::fopen(out, "w");
    if (!f) return 3;
    std::fprintf(f, "%u\n", ns);
    for (unsigned i = 0; i <= ns; ++i) std::fprintf(f, "%u ", idx[i]);
    std::fprintf(f, "\n");
    for (unsigned k = 0; k < idx[ns]; ++k) {
        std::fprintf(f, "%.9g %.9g", x[k], y[k]);
        if (z) std::fprintf(f, " %.9g", z[k]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    delete[] x;
    delete[] y;
    delete[] z;
    delete[] idx;
    return 0;
}

template <typename Seg>
static void print_seg(FILE *f, const Seg &pts, bool end)
{
    std::fprintf(f, "%s", end ? "E" : "-");
    for (const auto &p : pts) {
#if DIM == 2
        std::fprintf(f, " %d,%d", p.x, p.y);
#else
        std::fprintf(f, " %d,%d,%d", p.x, p.y, p.z);
#endif
    }
    std::fprintf(f, "\n");
}

static int do_trace(const Args &a, bool plain, const char *in, const char *out)
{
    std::vector<char> b = read_file(in);
    if (b.size() != a.n()) return 2;
    std::vector<unsigned char> src(b.begin(), b.end());
    FILE *f = std::fopen(out, "w");
    if (!f) return 3;
#if DIM == 2
    (void)plain;
    std::vector<SSegment *> segs = ExtractSegmentsV<unsigned char>(src.data(), a.w, a.h);
#else
    if (plain) {
        std::vector<std::vector<point3D>> ps = ExtractSegments<unsigned char>(src.data(), a.w, a.h, a.d);
        for (const auto &p : ps) print_seg(f, p, false);
        std::fclose(f);
        return 0;
    }
    std::vector<SSegment *> segs = ExtractSegmentsV<unsigned char>(src.data(), a.w, a.h, a.d);
#endif
    for (SSegment *s : segs) {
        print_seg(f, s->points, s->is_end_segment);
        delete s;
    }
    std::fclose(f);
    return 0;
}

#define DISPATCH(type_name, CALL)                                                   \
    do {                                                                            \
        const std::string t = type_name;                                            \
        if (t == "uint8") { using T = unsigned char; return CALL; }                 \
        if (t == "int8") { using T = signed char; return CALL; }                    \
        if (t == "char") { using T = char; return CALL; }                           \
        if (t == "uint16") { using T = unsigned short; return CALL; }               \
        if (t == "int16") { using T = short; return CALL; }                         \
        if (t == "uint32") { using T = unsigned int; return CALL; }                 \
        if (t == "int32") { using T = int; return CALL; }                           \
        if (t == "float32") { using T = float; return CALL; }                       \
        if (t == "float64") { using T = double; return CALL; }                      \
        if (t == "bool") { using T = bool; return CALL; }                           \
        return 4;                                                                   \
    } while (0)

int main(int argc, char **argv)
{
    if (argc < 2) return 1;
    const std::string mode = argv[1];
    int k = 2;
    Args a;
    std::string type;
    if (mode == "thin" || mode == "centerline") type = argv[k++];
    if (argc < k + 2 + (DIM == 3)) return 1;
    a.w = std::atoi(argv[k++]);
    a.h = std::atoi(argv[k++]);
    if (DIM == 3) a.d = std::atoi(argv[k++]);
    if (mode == "thin" && argc == k + 2) DISPATCH(type, (do_thin<T>(a, argv[k], argv[k + 1])));
    if (mode == "centerline" && argc == k + 4)
        DISPATCH(type, (do_centerline<T>(a, std::atoi(argv[k]), std::atoi(argv[k + 1]), argv[k + 2],
                                         argv[k + 3])));
    if ((mode == "trace" || mode == "trace_plain") && argc == k + 2)
        return do_trace(a, mode == "trace_plain", argv[k], argv[k + 1]);
    return 1;
}
