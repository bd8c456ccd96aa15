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

// 4 packed bytes -> 4 bits (bit k set iff byte k is nonzero)
__device__ __forceinline__ uint32_t nz4(uint32_t x)
{
    uint32_t t = x | (x >> 4);
    t |= t >> 2;
    t |= t >> 1;
    t &= 0x01010101u;
    return (t | (t >> 7) | (t >> 14) | (t >> 21)) & 0xfu;
}

// 4 bits -> 4 bytes of 0/1
__device__ __forceinline__ uint32_t spread4(uint32_t b)
{
    return ((b & 0xfu) * 0x00204081u) & 0x01010101u;
}

// Bits of the n (1..64) bytes at p: bit k set iff p[k] != 0.
__device__ __forceinline__ uint64_t pack_run_u8(const uint8_t *p, int n)
{
    const uintptr_t a = (uintptr_t)p;
    const int r = (int)(a & 3);
    const uint32_t *d = reinterpret_cast<const uint32_t *>(a - (uintptr_t)r);
    const int nd = (r + n + 3) >> 2;  // dwords overlapping the run, 1..17
    uint32_t D[17];
#pragma unroll
    for (int k = 0; k < 17; ++k) D[k] = k < nd ? d[k] : 0u;
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        lo |= nz4(__builtin_amdgcn_alignbyte(D[j + 1], D[j], (uint32_t)r)) << (4 * j);
        hi |= nz4(__builtin_amdgcn_alignbyte(D[j + 9], D[j + 8], (uint32_t)r)) << (4 * j);
    }
    uint64_t v = ((uint64_t)hi << 32) | lo;
    if (n < 64) v &= (1ull << n) - 1ull;
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

}  // namespace bgt
