// pack.h -- byte <-> bit conversions between caller buffers (one uint8 per
// pixel, rows at any alignment) and the packed state words.
//
// A run is up to 64 consecutive pixels of one row.  Loads and stores go
// through the naturally aligned dwords that overlap the run, so rows of any
// width and pitch take the same vector path; bytes outside the run are never
// written, and only dwords holding at least one byte of the run are read.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace bgt {

// 16 packed bytes -> 16 bits (bit k set iff byte k is nonzero).  Per dword,
// bit 7 of each byte becomes "byte nonzero" ((b & 0x7f) + 0x7f carries into
// bit 7 unless the low seven bits are zero; or-ing b adds its own bit 7), and
// v_dot4_u32_u8 sums the four flag bytes (128 each) weighted 1, 2, 4, 8 (and
// 16 .. 128 for the next dword): 128 x the flags of eight bytes per chain.
__device__ __forceinline__ uint32_t nz16(uint4 v)
{
    auto f = [](uint32_t x) { return (((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) & 0x80808080u; };
    const uint32_t lo = __builtin_amdgcn_udot4(f(v.y), 0x80402010u,
                                               __builtin_amdgcn_udot4(f(v.x), 0x08040201u, 0u, false), false);
    const uint32_t hi = __builtin_amdgcn_udot4(f(v.w), 0x80402010u,
                                               __builtin_amdgcn_udot4(f(v.z), 0x08040201u, 0u, false), false);
    return (lo >> 7) | (hi << 1);
}

// 4 bits -> 4 bytes of 0/1
__device__ __forceinline__ uint32_t spread4(uint32_t b)
{
    return ((b & 0xfu) * 0x00204081u) & 0x01010101u;
}

// Bits of the n (1..64) bytes at p: bit k set iff p[k] != 0.  Reads the
// 16-byte aligned blocks that overlap the run (at most five uint4 loads; an
// aligned block holding one byte of the run cannot cross a page), turns the
// bytes into an 80-bit mask and shifts it by the run's offset in the first
// block.
__device__ __forceinline__ uint64_t pack_run_u8(const uint8_t *p, int n)
{
    const uintptr_t a = (uintptr_t)p;
    const int r = (int)(a & 15);
    const uint4 *q = reinterpret_cast<const uint4 *>(a - (uintptr_t)r);
    const int nq = (r + n + 15) >> 4;  // blocks overlapping the run, 1..5
    uint32_t m[3] = {0u, 0u, 0u};      // bit i = byte i of the blocks nonzero
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (k < nq) {
            const uint4 v = q[k];
            const uint32_t bits = nz16(v);
            m[k >> 1] |= bits << ((k & 1) * 16);
        }
    }
    const uint32_t lo = __builtin_amdgcn_alignbit(m[1], m[0], (uint32_t)r);
    const uint32_t hi = __builtin_amdgcn_alignbit(m[2], m[1], (uint32_t)r);
    uint64_t v = ((uint64_t)hi << 32) | lo;
    if (n < 64) v &= (1ull << n) - 1ull;
    return v;
}

// The same for the runs of a whole wave, with coalesced loads.  Every lane of
// the wave calls it (`valid` lanes form a prefix), and so does every other
// wave of the workgroup (it holds a barrier).  The valid lanes' runs must
// follow each other in memory (lane l+1's run starts where lane l's ends, as
// the runs of a contiguous image or volume do), so they span at most 4096
// bytes.  The wave loads the 16-byte aligned blocks overlapping that span --
// at most 257, five uint4 instructions of 64 consecutive blocks, each
// covering 1 KiB contiguously -- turns each into a 16-bit mask in LDS (`lds`:
// 320 words per wave), and each lane assembles its run's 80-bit window from
// the masks, as pack_run_u8 does from its own loads.
__device__ __forceinline__ uint64_t pack_runs_wave_u8(const uint8_t *p, int n, bool valid,
                                                      uint32_t *lds)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t vm = __ballot(valid);
    uintptr_t a0 = 0;
    uint32_t nb = 0;
    if (vm) {
        const int last = 63 - __builtin_clzll(vm);
        const uintptr_t s = (uintptr_t)p, e = (uintptr_t)p + (uintptr_t)(valid ? n : 0);
        const uintptr_t s0 = ((uintptr_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)s, 0)) |
                             ((uintptr_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(s >> 32), 0) << 32);
        const uintptr_t e1 = ((uintptr_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)e, last)) |
                             ((uintptr_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(e >> 32), last) << 32);
        a0 = s0 & ~(uintptr_t)15;
        nb = (uint32_t)((((e1 + 15) & ~(uintptr_t)15) - a0) >> 4);
    }
    const uint4 *q = reinterpret_cast<const uint4 *>(a0);
#pragma unroll
    for (uint32_t j = 0; j < 5; ++j) {
        const uint32_t k = 64u * j + lane;
        uint32_t m = 0;
        if (k < nb) {
            const uint4 v = q[k];
            m = nz16(v);
        }
        lds[k] = m;
    }
    __syncthreads();
    uint64_t v = 0;
    if (valid) {
        const uint32_t o = (uint32_t)((uintptr_t)p - a0), kb = o >> 4, r = o & 15u;
        const uint32_t m0 = lds[kb] | lds[kb + 1] << 16, m1 = lds[kb + 2] | lds[kb + 3] << 16;
        const uint32_t m2 = lds[kb + 4];
        const uint32_t lo = __builtin_amdgcn_alignbit(m1, m0, r);
        const uint32_t hi = __builtin_amdgcn_alignbit(m2, m1, r);
        v = ((uint64_t)hi << 32) | lo;
        if (n < 64) v &= (1ull << n) - 1ull;
    }
    return v;
}

// Writes the n (1..64) bytes at p as 0/1 from bits v (bit k -> p[k]).
__device__ __forceinline__ void unpack_run_u8(uint64_t v, uint8_t *p, int n)
{
    const uintptr_t a = (uintptr_t)p;
    const int r = (int)(a & 3);
    uint32_t *d = reinterpret_cast<uint32_t *>(a - (uintptr_t)r);
#pragma unroll
    for (int k = 0; k < 17; ++k) {
        const int p0 = 4 * k - r;  // run index of the dword's first byte
        if (p0 >= n) break;
        if (p0 >= 0 && p0 + 4 <= n) {
            d[k] = spread4((uint32_t)(v >> p0));
        } else {
            const int q0 = p0 < 0 ? 0 : p0;
            const int q1 = p0 + 4 < n ? p0 + 4 : n;
            for (int q = q0; q < q1; ++q) p[q] = (uint8_t)((v >> q) & 1u);
        }
    }
}

// ---- lane order of the 2D state words -------------------------------------
//
// The 2D band kernel keeps pixel k (0..63) of a state word at bit
// lane_bit(k) = (k & 1) * 32 + (k >> 1): the even pixels in the low 32-bit
// half, the odd pixels in the high half.  Then the West neighbour of every odd
// pixel and the East neighbour of every even pixel sit at the same bit of the
// other half, and only the other neighbour of each half needs a shifted copy.
// lane_order turns a word in pixel order (bit k = pixel k, as the packers
// produce it) into lane order; pixel_order turns it back.

__host__ __device__ constexpr int lane_bit(int k) { return (k & 1) * 32 + (k >> 1); }

// even bits of x -> bits 0..15, odd bits -> bits 16..31 (four delta swaps)
__host__ __device__ inline uint32_t unshuffle32(uint32_t x)
{
    uint32_t t;
    t = (x ^ (x >> 1)) & 0x22222222u; x ^= t ^ (t << 1);
    t = (x ^ (x >> 2)) & 0x0c0c0c0cu; x ^= t ^ (t << 2);
    t = (x ^ (x >> 4)) & 0x00f000f0u; x ^= t ^ (t << 4);
    t = (x ^ (x >> 8)) & 0x0000ff00u; x ^= t ^ (t << 8);
    return x;
}

// the inverse: bits 0..15 -> even bits, bits 16..31 -> odd bits
__host__ __device__ inline uint32_t shuffle32(uint32_t x)
{
    uint32_t t;
    t = (x ^ (x >> 8)) & 0x0000ff00u; x ^= t ^ (t << 8);
    t = (x ^ (x >> 4)) & 0x00f000f0u; x ^= t ^ (t << 4);
    t = (x ^ (x >> 2)) & 0x0c0c0c0cu; x ^= t ^ (t << 2);
    t = (x ^ (x >> 1)) & 0x22222222u; x ^= t ^ (t << 1);
    return x;
}

__host__ __device__ inline uint64_t lane_order(uint64_t v)
{
    const uint32_t lo = unshuffle32((uint32_t)v), hi = unshuffle32((uint32_t)(v >> 32));
    const uint32_t even = (lo & 0xffffu) | (hi << 16), odd = (lo >> 16) | (hi & 0xffff0000u);
    return (uint64_t)even | ((uint64_t)odd << 32);
}

__host__ __device__ inline uint64_t pixel_order(uint64_t v)
{
    const uint32_t even = (uint32_t)v, odd = (uint32_t)(v >> 32);
    const uint32_t lo = shuffle32((even & 0xffffu) | (odd << 16));
    const uint32_t hi = shuffle32((even >> 16) | (odd & 0xffff0000u));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

}  // namespace bgt

namespace bgt {

// Coalesced unpacking: the caller's bytes of an item are cut into 16-byte
// aligned chunks, one per thread, so the lanes of a wave store consecutive
// uint4s.  A segment is `n` consecutive bytes at `base` holding pixels
// f = 0..n-1 of lines of width w (pixel f is line f / w, x = f % w): a whole
// contiguous item, or one row of a pitched image.  Chunks that lie inside the
// segment and inside one line take one uint4 store; the rest (line ends,
// segment ends, lines narrower than 16) store their in-segment bytes one by
// one.  Bytes outside the segment are never written.
//
// Src provides   uint32_t bits16(line, x)  (16 pixels from x, x + 16 <= w)
//        and     uint32_t bit(line, x).
template <class Src>
__device__ __forceinline__ void unpack_chunk_u8(uint8_t *base, uint32_t n, uint32_t w, uint32_t c,
                                                const Src &src)
{
    uint8_t *C = reinterpret_cast<uint8_t *>(reinterpret_cast<uintptr_t>(base) & ~(uintptr_t)15) +
                 (size_t)c * 16;
    const long f0 = (long)(C - base);
    if (f0 >= (long)n || f0 + 16 <= 0) return;
    if (f0 >= 0 && f0 + 16 <= (long)n) {
        const uint32_t line = (uint32_t)f0 / w, x0 = (uint32_t)f0 - line * w;
        if (x0 + 16 <= w) {
            const uint32_t b = src.bits16(line, x0);
            *reinterpret_cast<uint4 *>(C) =
                make_uint4(spread4(b), spread4(b >> 4), spread4(b >> 8), spread4(b >> 12));
            return;
        }
    }
    const uint32_t fa = (uint32_t)(f0 < 0 ? 0 : f0);
    const uint32_t fb = (uint32_t)(f0 + 16 < (long)n ? f0 + 16 : (long)n);
    uint32_t line = fa / w, x = fa - line * w;
    for (uint32_t f = fa; f < fb; ++f) {
        C[f - f0] = (uint8_t)src.bit(line, x);
        if (++x == w) {
            x = 0;
            ++line;
        }
    }
}

// Chunks a segment of n bytes at base can touch.
__host__ __device__ inline uint32_t unpack_chunks(uint32_t n) { return n / 16 + 2; }

// The same with one thread per aligned 64-byte block: one division per block,
// the block's 64 pixels gathered from at most two lines, four uint4 stores.
// Blocks at a segment end, or crossing more than one line end (lines
// narrower than 64), store their in-segment bytes one by one.  Src also
// provides uint64_t bits64(line, x): 64 pixels from x of the line (pixels
// past the line's width may hold anything; they are masked here).
template <class Src>
__device__ __forceinline__ void unpack_block64_u8(uint8_t *base, uint32_t n, uint32_t w, uint32_t c,
                                                  const Src &src)
{
    uint8_t *B = reinterpret_cast<uint8_t *>(reinterpret_cast<uintptr_t>(base) & ~(uintptr_t)63) +
                 (size_t)c * 64;
    const long f0 = (long)(B - base);
    if (f0 >= (long)n || f0 + 64 <= 0) return;
    if (f0 >= 0 && f0 + 64 <= (long)n) {
        const uint32_t line = (uint32_t)f0 / w, x = (uint32_t)f0 - line * w;
        bool fast = true;
        uint64_t v;
        if (x + 64 <= w) {
            v = src.bits64(line, x);
        } else if (w >= 64) {
            const uint32_t k1 = w - x;  // 1..63 pixels left on this line
            v = (src.bits64(line, x) & ((1ull << k1) - 1ull)) | (src.bits64(line + 1, 0) << k1);
        } else {
            fast = false;
        }
        if (fast) {
            const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
            uint4 *q = reinterpret_cast<uint4 *>(B);
            q[0] = make_uint4(spread4(lo), spread4(lo >> 4), spread4(lo >> 8), spread4(lo >> 12));
            q[1] = make_uint4(spread4(lo >> 16), spread4(lo >> 20), spread4(lo >> 24), spread4(lo >> 28));
            q[2] = make_uint4(spread4(hi), spread4(hi >> 4), spread4(hi >> 8), spread4(hi >> 12));
            q[3] = make_uint4(spread4(hi >> 16), spread4(hi >> 20), spread4(hi >> 24), spread4(hi >> 28));
            return;
        }
    }
    const uint32_t fa = (uint32_t)(f0 < 0 ? 0 : f0);
    const uint32_t fb = (uint32_t)(f0 + 64 < (long)n ? f0 + 64 : (long)n);
    uint32_t line = fa / w, x = fa - line * w;
    for (uint32_t f = fa; f < fb; ++f) {
        B[f - f0] = (uint8_t)src.bit(line, x);
        if (++x == w) {
            x = 0;
            ++line;
        }
    }
}

// Blocks a segment of n bytes at base can touch.
__host__ __device__ inline uint32_t unpack_blocks(uint32_t n) { return n / 64 + 2; }

// The same blocks with coalesced stores: the 64 lanes of a wave must hold
// blocks c0 .. c0 + 63 of one segment (c0 a multiple of 64) and all call it
// (lanes past the segment included).  Each lane gathers its block's 64 pixels
// as above (one division, two loads); then store instruction j of the wave
// writes blocks 16j .. 16j + 15, lane i the 16-byte piece i % 4 of block
// 16j + i / 4, whose 16 bits it fetches with ds_bpermute.  Every store
// instruction thus covers 1 KiB contiguously instead of 64 pieces 64 bytes
// apart (about twice the store bandwidth on this GPU).  Blocks that are not
// whole inside one segment and two lines store their bytes one by one.
template <class Src>
__device__ __forceinline__ void unpack_block64_wave_u8(uint8_t *base, uint32_t n, uint32_t w,
                                                       uint32_t c, const Src &src)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint8_t *B = reinterpret_cast<uint8_t *>(reinterpret_cast<uintptr_t>(base) & ~(uintptr_t)63) +
                 (size_t)c * 64;
    const long f0 = (long)(B - base);
    bool fast = false;
    uint64_t v = 0;
    if (f0 >= 0 && f0 + 64 <= (long)n) {
        const uint32_t line = (uint32_t)f0 / w, x = (uint32_t)f0 - line * w;
        if (x + 64 <= w) {
            v = src.bits64(line, x);
            fast = true;
        } else if (w >= 64) {
            const uint32_t k1 = w - x;  // 1..63 pixels left on this line
            v = (src.bits64(line, x) & ((1ull << k1) - 1ull)) | (src.bits64(line + 1, 0) << k1);
            fast = true;
        }
    }
    if (!fast && f0 < (long)n && f0 + 64 > 0) {
        const uint32_t fa = (uint32_t)(f0 < 0 ? 0 : f0);
        const uint32_t fb = (uint32_t)(f0 + 64 < (long)n ? f0 + 64 : (long)n);
        uint32_t line = fa / w, x = fa - line * w;
        for (uint32_t f = fa; f < fb; ++f) {
            B[f - f0] = (uint8_t)src.bit(line, x);
            if (++x == w) {
                x = 0;
                ++line;
            }
        }
    }
    const uint64_t fm = __ballot(fast);
    const int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
    uint8_t *B0 = B - (size_t)lane * 64;  // block of lane 0
    const uint32_t piece = lane & 3u;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t from = 16u * j + (lane >> 2);
        const uint32_t a = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(from << 2), lo);
        const uint32_t b = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(from << 2), hi);
        const uint32_t word = piece < 2 ? a : b;
        const uint32_t bits = (piece & 1u) ? (word >> 16) : word;
        if ((fm >> from) & 1ull)
            *reinterpret_cast<uint4 *>(B0 + (size_t)from * 64 + piece * 16) =
                make_uint4(spread4(bits), spread4(bits >> 4), spread4(bits >> 8), spread4(bits >> 12));
    }
}

// Threads of the coalesced block unpacking of a segment of n bytes (whole
// waves of 64 blocks).
__host__ __device__ inline uint32_t unpack_wave_blocks(uint32_t n) { return (n / 64 + 2 + 63) & ~63u; }

// Pixels of an image's rows in its packed state (64-bit words in lane order,
// L per row).
struct RowBits {
    const uint64_t *row0;
    uint32_t L;
    __device__ uint32_t bits16(uint32_t y, uint32_t x) const
    {
        const uint64_t *row = row0 + (size_t)y * L;
        const uint32_t i = x >> 6, s = x & 63;
        uint64_t v = pixel_order(row[i]) >> s;
        if (s > 48) v |= pixel_order(row[i + 1]) << (64 - s);  // x + 16 <= width: word i+1 is in the row
        return (uint32_t)v & 0xffffu;
    }
    __device__ uint32_t bit(uint32_t y, uint32_t x) const
    {
        return (uint32_t)(row0[(size_t)y * L + (x >> 6)] >> lane_bit((int)(x & 63))) & 1u;
    }
    // 64 pixels from x of row y; pixels past the row's last word read as 0
    __device__ uint64_t bits64(uint32_t y, uint32_t x) const
    {
        const uint64_t *row = row0 + (size_t)y * L;
        const uint32_t i = x >> 6, s = x & 63;
        uint64_t v = pixel_order(row[i]) >> s;
        if (s && i + 1 < L) v |= pixel_order(row[i + 1]) << (64 - s);
        return v;
    }
};

// Block t (64 bytes) of a w x h uint8 image at `out` (row pitch `pitch`) from
// its packed rows: the whole image is one segment when `flat` (pitch == w),
// unpacked by whole waves with coalesced stores, else each row is a segment
// of `uchunks` blocks.
__device__ __forceinline__ void unpack_image_u8(const uint64_t *buf, uint32_t L, int w, int h,
                                                uint8_t *out, size_t pitch, uint32_t uchunks,
                                                int flat, uint32_t t)
{
    if (flat) {
        // every lane of a wave takes part (uchunks is whole waves)
        unpack_block64_wave_u8(out, (uint32_t)w * (uint32_t)h, (uint32_t)w, t, RowBits{buf, L});
    } else {
        const uint32_t y = t / uchunks, c = t - y * uchunks;
        if (y >= (uint32_t)h) return;
        unpack_block64_u8(out + (size_t)y * pitch, (uint32_t)w, (uint32_t)w, c,
                          RowBits{buf + (size_t)y * L, L});
    }
}

// Unpack grid of an image: segment mode and chunks per segment.
inline void unpack_image_grid(int w, int h, size_t pitch, int &flat, uint32_t &uchunks,
                              long &threads)
{
    const long px = (long)w * h;
    flat = pitch == (size_t)w && px < 0x7fffffffL;
    uchunks = flat ? unpack_wave_blocks((uint32_t)px) : unpack_blocks((uint32_t)w);
    threads = flat ? (long)uchunks : (long)uchunks * h;
}

}  // namespace bgt
