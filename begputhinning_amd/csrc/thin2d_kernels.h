// thin2d_kernels.h -- gfx950 device code of the 2D thinning path.
//
// State layout in HBM: bit-packed rows, bit k of 64-bit word c of row y is
// pixel x = 64*c + k (the reference packs 32-pixel words the same way,
// 2D/CUDA_DeviceHelperFunctions.cuh:101).  A row has L words, L = the next
// power of two >= ceil(width/64); words past the width are zero and stay zero
// because a subiteration only clears bits.
//
// Work unit: a "band" of consecutive rows of one image, processed by one
// workgroup of NW wave64s.  Lane (g, c) = (lane / L, lane % L) of wave v owns
// word column c of segment q = v*G + g (G = 64 / L segments per wave); a
// segment is R consecutive rows held in registers.  Per subiteration each
// segment exchanges its top and bottom rows with its neighbours through
// double-buffered LDS (one barrier per subiteration), W/E pixel neighbours come
// from the adjacent lane with DPP wave_shr:1 / wave_shl:1 plus v_alignbit, and
// up/down neighbours are the adjacent registers.  K iterations (4K
// subiterations) run per HBM pass; the band carries H = 4K halo rows on each
// side, and only its interior rows are written back, so the result equals the
// global synchronous algorithm exactly (Chebyshev radius 1 per subiteration).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace bgt {

enum { ST_PASS = 0, ST_DONE = 1, ST_ITERS = 2, ST_MASK = 3, ST_PARITY = 4, ST_INTS = 8 };

struct Item2D {
    uint64_t *buf[2];   // ping-pong packed state (L words per row)
    uint8_t *chg;       // per band change flags, [2][nbands] (pass parity)
    int *st;            // item state (ST_*)
    int w, h, L, lgL;   // L lanes per row, lgL = log2(L)
    int nbands;         // bands covering the image
    int band_rows;      // interior rows per band
};

struct Tile2D {
    int item;
    int band;
};

// ---- lane exchange -------------------------------------------------------

// lane i receives lane i-1's value (lane 0 receives 0)
__device__ __forceinline__ uint32_t from_lower_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);
}
// lane i receives lane i+1's value (lane 63 receives 0)
__device__ __forceinline__ uint32_t from_upper_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, false);
}

struct W64 {
    uint32_t lo, hi;
};

// bit k <- pixel x-1 (the reference's ShiftLeft)
__device__ __forceinline__ W64 shift_w(W64 v, uint32_t carry_mask)
{
    uint32_t c = from_lower_lane(v.hi) & carry_mask;
    return {__builtin_amdgcn_alignbit(v.lo, c, 31), __builtin_amdgcn_alignbit(v.hi, v.lo, 31)};
}
// bit k <- pixel x+1 (the reference's ShiftRight)
__device__ __forceinline__ W64 shift_e(W64 v, uint32_t carry_mask)
{
    uint32_t c = from_upper_lane(v.lo) & carry_mask;
    return {__builtin_amdgcn_alignbit(v.hi, v.lo, 1), __builtin_amdgcn_alignbit(c, v.hi, 1)};
}

// ---- the deletion rule ---------------------------------------------------
// North template (2D/CUDA_SubiterationKernels2D.cuh:15-40) over the clockwise
// neighbour labels x0..x7 = N, NE, E, SE, S, SW, W, NW.  The South, East and
// West rules of the reference are this rule with the labels advanced by 4, 2
// and 6 (checked exhaustively against the transcribed reference rules).
__device__ __forceinline__ uint32_t north_rule(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3,
                                               uint32_t x4, uint32_t x5, uint32_t x6, uint32_t x7)
{
    uint32_t a = ~x1 & (x2 ^ x3 ^ ~x4) & (x3 ^ x5) & (x3 ^ x6) & ~x7;
    uint32_t b = x2 & x4 & (x6 | ~x7);
    uint32_t c = x6 & ~x1 & x4;
    uint32_t d = (x1 ^ x5) & (x2 ^ x5) & (x4 ^ ~x5) & ~x6 & ~x7;
    uint32_t e = ~x1 & ~x2 & ~x3 & x6 & x7;
    return ~x0 & (a | b | c | d | e);
}

// direction D: 0 = N, 1 = S, 2 = E, 3 = W (the reference's order,
// 2D/CUDA_ThinningKernelMultiBlock2D.cuh:104-135)
template <int D>
__device__ __forceinline__ uint32_t rule(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3,
                                         uint32_t x4, uint32_t x5, uint32_t x6, uint32_t x7)
{
    if (D == 0) return north_rule(x0, x1, x2, x3, x4, x5, x6, x7);
    if (D == 1) return north_rule(x4, x5, x6, x7, x0, x1, x2, x3);
    if (D == 2) return north_rule(x2, x3, x4, x5, x6, x7, x0, x1);
    return north_rule(x6, x7, x0, x1, x2, x3, x4, x5);
}

// One subiteration of direction D over the R register rows of a segment.
// up_edge / dn_edge are the rows just above / below the segment (new values
// of the previous subiteration, from the neighbouring segments).  Bit r of
// `rows_changed` is set when row r lost a pixel.
template <int D, int R>
__device__ __forceinline__ void subiteration(uint32_t (&lo)[R], uint32_t (&hi)[R], W64 up_edge,
                                             W64 dn_edge, uint32_t mw, uint32_t me,
                                             uint32_t &rows_changed)
{
    W64 up = up_edge, upw = shift_w(up, mw), upe = shift_e(up, me);
    W64 cp = {lo[0], hi[0]}, cpw = shift_w(cp, mw), cpe = shift_e(cp, me);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        W64 dn = (r + 1 < R) ? W64{lo[r + 1 < R ? r + 1 : 0], hi[r + 1 < R ? r + 1 : 0]} : dn_edge;
        W64 dnw = shift_w(dn, mw), dne = shift_e(dn, me);
        // x0=N=up, x1=NE, x2=E, x3=SE, x4=S=dn, x5=SW, x6=W, x7=NW
        const uint32_t mlo = rule<D>(up.lo, upe.lo, cpe.lo, dne.lo, dn.lo, dnw.lo, cpw.lo, upw.lo);
        const uint32_t mhi = rule<D>(up.hi, upe.hi, cpe.hi, dne.hi, dn.hi, dnw.hi, cpw.hi, upw.hi);
        const uint32_t dlo = mlo & cp.lo, dhi = mhi & cp.hi;
        lo[r] = cp.lo ^ dlo;
        hi[r] = cp.hi ^ dhi;
        rows_changed |= ((dlo | dhi) != 0u ? 1u : 0u) << r;
        up = cp; upw = cpw; upe = cpe;
        cp = dn; cpw = dnw; cpe = dne;
        // rows are independent; keep the scheduler from interleaving all of
        // them (that multiplies live registers without adding useful ILP)
        __builtin_amdgcn_sched_barrier(0);
    }
}

}  // namespace bgt
