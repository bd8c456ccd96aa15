// thin2d_kernels.h -- gfx950 device code of the 2D thinning path.
//
// State layout in HBM: bit-packed rows, 64-bit word c of row y holding pixels
// x = 64*c + k, k = 0..63 (the reference packs 32-pixel words of consecutive
// pixels, 2D/CUDA_DeviceHelperFunctions.cuh:101), in lane order: pixel k at
// bit lane_bit(k) = (k & 1) * 32 + (k >> 1), even pixels in the low half, odd
// in the high half (pack.h), so that each half has one of its West/East
// neighbours at the same bit of the other half.  A row has L =
// ceil(width/64) words; bits of pixels past the width are zero and stay zero
// because a subiteration only clears bits.
//
// Work unit: a "band" of consecutive rows of one image, processed by one
// workgroup of NW wave64s.  Lane (g, c) = (lane / L, lane % L) of wave v owns
// word column c of segment q = v*G + g (G = floor(64 / L) segments per wave,
// the lanes past G*L idle); a segment is R consecutive rows held in registers.  Per subiteration each
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

#include "pack.h"

namespace bgt {

// Per-item state words.  ST_MASK collects, over the tiles of a pass, bit j =
// "iteration j of this pass deleted something"; ST_RUN is set by a tile that
// iterated its item to convergence on its own (single-tile items) and then
// holds the number of iterations that deleted something; ST_BASE is the pass
// of a work queue in which the item's first iteration runs (0 in a batch; the
// pass an item joined at in a continuously running queue).
enum {
    ST_PASS = 0, ST_DONE = 1, ST_ITERS = 2, ST_MASK = 3, ST_PARITY = 4, ST_RUN = 5, ST_BASE = 6,
    ST_INTS = 8
};

struct Item2D {
    uint64_t *buf[2];   // ping-pong packed state (L words per row)
    uint8_t *chg;       // per band change flags, [2][nbands] (pass parity)
    int *st;            // item state (ST_*)
    int w, h, L;        // L = 64-bit words per row
    int nbands;         // bands covering the image
    int band_rows;      // interior rows per band
    int tile0;          // id of band 0 in the batch's tile arrays
};

// Lanes a row of L words takes with V words per lane.
__host__ __device__ constexpr int lanes_for(int L, int V) { return (L + V - 1) / V; }

struct Tile2D {
    int item;
    int band;
};

// Caller buffers of an item (w uint8 pixels per row, rows at any alignment
// and pitch) whose first pass packs its band windows from `in` itself.
struct ItemIO2D {
    const uint8_t *in;
    size_t in_pitch;
    uint8_t *out;
    size_t out_pitch;
};

// ---- lane exchange -------------------------------------------------------

// lane i receives lane i-1's value (lane 0 receives 0: DPP wave_shr:1 with
// bound_ctrl, which writes 0 where the source lane is out of range)
__device__ __forceinline__ uint32_t from_lower_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, true);
}
// lane i receives lane i+1's value (lane 63 receives 0: DPP wave_shl:1)
__device__ __forceinline__ uint32_t from_upper_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, true);
}

// The V consecutive 64-bit words a lane holds of one row, in lane order
// (pack.h): h[2j] holds the even pixels of word j (bit b = pixel 2b of the
// word), h[2j + 1] its odd pixels (bit b = pixel 2b + 1).
template <int V>
struct Row {
    uint32_t h[2 * V];
};

// bit <- pixel x-1 (the reference's ShiftLeft).  An odd pixel's West
// neighbour is the even pixel at the same bit; an even pixel 2b's is odd
// pixel 2b - 1, one bit down in the odd half, bit 0 taking bit 31 of the
// previous word's odd half (of the lane below for the first word, masked at
// segment starts).
template <int V>
__device__ __forceinline__ Row<V> shift_w(const Row<V> &v, uint32_t carry_mask)
{
    Row<V> o;
    const uint32_t c = from_lower_lane(v.h[2 * V - 1]) & carry_mask;
#pragma unroll
    for (int j = 0; j < V; ++j) {
        o.h[2 * j] = __builtin_amdgcn_alignbit(v.h[2 * j + 1], j ? v.h[2 * j - 1] : c, 31);
        o.h[2 * j + 1] = v.h[2 * j];
    }
    return o;
}
// bit <- pixel x+1 (the reference's ShiftRight).  An even pixel's East
// neighbour is the odd pixel at the same bit; an odd pixel 2b + 1's is even
// pixel 2b + 2, one bit up in the even half, bit 31 taking bit 0 of the next
// word's even half (of the lane above for the last word, masked at segment
// ends).
template <int V>
__device__ __forceinline__ Row<V> shift_e(const Row<V> &v, uint32_t carry_mask)
{
    Row<V> o;
    const uint32_t c = from_upper_lane(v.h[0]) & carry_mask;
#pragma unroll
    for (int j = 0; j < V; ++j) {
        o.h[2 * j] = v.h[2 * j + 1];
        o.h[2 * j + 1] = __builtin_amdgcn_alignbit(j + 1 < V ? v.h[2 * j + 2] : c, v.h[2 * j], 1);
    }
    return o;
}

// ---- the deletion rule ---------------------------------------------------

// v_bitop3_b32: any bitwise function of three operands in one instruction.
// Bit (a<<2 | b<<1 | c) of the table T is the result for input bits a, b, c.
template <unsigned T>
__device__ __forceinline__ uint32_t lut3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, T & 0xffu);
}

// Pixels of `ctr` that survive the North subiteration, over the clockwise
// neighbour labels x0..x7 = N, NE, E, SE, S, SW, W, NW.  The North template
// of 2D/CUDA_SubiterationKernels2D.cuh:15-40 deletes a pixel when ~x0 & F,
//   F = a | b | c | d | e,  a = ~x1 ~x7 (x2^x3^~x4) (x3^x5) (x3^x6),
//   b = x2 x4 (x6|~x7),  c = x4 x6 ~x1,  d = ~x6 ~x7 (x1^x5) (x2^x5) (x4^~x5),
//   e = ~x1 ~x2 ~x3 x6 x7.
// F, a function of x1..x7, is built here from 8 three-input gates, then
// ctr & (x0 | ~F) from one more.  The network is a decomposition of F's
// 128-entry truth table: F = G(x1, x2, W) with W = H(x4, x6, M), M = x5 ? B : A,
// A and B of (x2, x3, x4, x6, x7) in two and three gates; each gate's table
// takes its value where the outer gates read it and is free elsewhere.  The
// tables are the hex constants (bit a<<2 | b<<1 | c of a table is the result
// for inputs a, b, c); the network equals the template on all 512 inputs.
__device__ __forceinline__ uint32_t north_keep(uint32_t ctr, uint32_t x0, uint32_t x1, uint32_t x2,
                                               uint32_t x3, uint32_t x4, uint32_t x5, uint32_t x6,
                                               uint32_t x7)
{
    const uint32_t t1 = lut3<0x76>(x1, x4, x6);
    const uint32_t t2 = lut3<0xe8>(x2, x3, t1);
    const uint32_t ga = lut3<0x42>(x6, x7, t2);  // F where x5 = 0 (and x1 x2 leave it open)
    const uint32_t t3 = lut3<0x02>(x2, x3, x6);
    const uint32_t gb = lut3<0x1a>(x4, x7, t3);  // F where x5 = 1 (likewise)
    const uint32_t m = lut3<0xca>(x5, gb, ga);   // x5 ? gb : ga
    const uint32_t w = lut3<0xea>(x4, x6, m);
    const uint32_t f = lut3<0x8a>(x1, x2, w);    // F
    return lut3<0xd0>(ctr, x0, f);               // ctr & (x0 | ~F)
}

// direction D: 0 = N, 1 = S, 2 = E, 3 = W (the reference's order,
// 2D/CUDA_ThinningKernelMultiBlock2D.cuh:104-135).  The South, East and West
// rules of the reference are the North rule with the labels advanced by 4, 2
// and 6 (checked exhaustively against the transcribed reference rules).
template <int D>
__device__ __forceinline__ uint32_t keep(uint32_t ctr, uint32_t x0, uint32_t x1, uint32_t x2,
                                         uint32_t x3, uint32_t x4, uint32_t x5, uint32_t x6,
                                         uint32_t x7)
{
    if (D == 0) return north_keep(ctr, x0, x1, x2, x3, x4, x5, x6, x7);
    if (D == 1) return north_keep(ctr, x4, x5, x6, x7, x0, x1, x2, x3);
    if (D == 2) return north_keep(ctr, x2, x3, x4, x5, x6, x7, x0, x1);
    return north_keep(ctr, x6, x7, x0, x1, x2, x3, x4, x5);
}

// One subiteration of direction D over the R register rows of a segment.
// up_edge / dn_edge are the rows just above / below the segment (new values
// of the previous subiteration, from the neighbouring segments).
template <int D, int R, int V>
__device__ __forceinline__ void subiteration(Row<V> (&w)[R], const Row<V> &up_edge,
                                             const Row<V> &dn_edge, uint32_t mw, uint32_t me)
{
    Row<V> up = up_edge, upw = shift_w(up, mw), upe = shift_e(up, me);
    Row<V> cp = w[0], cpw = shift_w(cp, mw), cpe = shift_e(cp, me);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const Row<V> dn = (r + 1 < R) ? w[r + 1 < R ? r + 1 : 0] : dn_edge;
        const Row<V> dnw = shift_w(dn, mw), dne = shift_e(dn, me);
        // x0=N=up, x1=NE, x2=E, x3=SE, x4=S=dn, x5=SW, x6=W, x7=NW
#pragma unroll
        for (int i = 0; i < 2 * V; ++i)
            w[r].h[i] = keep<D>(cp.h[i], up.h[i], upe.h[i], cpe.h[i], dne.h[i], dn.h[i], dnw.h[i],
                                cpw.h[i], upw.h[i]);
        up = cp; upw = cpw; upe = cpe;
        cp = dn; cpw = dnw; cpe = dne;
        // rows are independent; keep the scheduler from interleaving all of
        // them (that multiplies live registers without adding useful ILP)
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Pixels in the rows of `rows` (bit r = register row r).  Subiterations only
// clear pixels, so an iteration deleted one of them iff this count dropped.
template <int R, int V>
__device__ __forceinline__ uint32_t count_rows(const Row<V> (&w)[R], uint32_t rows)
{
    uint32_t n = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < 2 * V; ++i) c += (uint32_t)__popc(w[r].h[i]);
        n += ((rows >> r) & 1u) * c;
    }
    return n;
}

// Advance one item after a pass: converged when some iteration of the pass
// deleted nothing (or a tile iterated the item to convergence on its own).
__device__ __forceinline__ void finalize_state(int *st, int K, int pass, int *live)
{
    if (st[ST_DONE]) return;
    const unsigned mask =
        __hip_atomic_load((unsigned *)&st[ST_MASK], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int run = st[ST_RUN];
    const int m = run ? run - 1 : __popc(mask);  // iterations of this pass that deleted something
    if (run || m < K) {
        st[ST_ITERS] = pass * K + m + 1;
        st[ST_PARITY] = (pass + 1) & 1;
        st[ST_DONE] = 1;
        if (live) atomicSub(live, 1);
    }
    st[ST_PASS] = pass + 1;
    st[ST_MASK] = 0;
}

// LDS words shared by a workgroup processing one band / brick
struct TileInfo {
    unsigned mask;   // iterations (bit j) in which an interior pixel was deleted
    unsigned occ;    // interior non-empty after the pass
    unsigned flag;   // per-iteration change flag (tile-local convergence)
    int t;           // work-queue slot
};

// Run one pass over the interior rows [row0, row0 + nrows) of item `it` (a
// band, or a part of one) with V words per lane: K iterations (4K
// subiterations) from buffer `par` into buffer par^1, or -- when `whole`: the
// rows are the whole image, so that the halo rows are all outside it --
// iterations until one deletes nothing, at most whole_max of them.  The window
// of the launch (waves x G segments x R rows, G = floor(64 / lanes per row))
// must hold nrows + 8K rows, and `edges` 4 * waves * 64 * V words.  Results
// land in `info` (bit j of the mask: iteration j deleted an interior pixel);
// returns the number of iterations that deleted something in the whole case
// (else 0).
//
// IO = 2 (with `io`, the item's caller buffers; the item's first pass): the
// window comes from io->in, the interior rows go to both buffers, as a pack
// would leave them, and an interior found empty skips the iterations
// (nothing can change in it).  Otherwise the packed state alone.
template <int R, int V, int IO = 0>
__device__ __forceinline__ int band_work(const Item2D &it, int row0, int nrows, bool whole, int K,
                                         int par, uint64_t *edges, TileInfo *info,
                                         int whole_max = 0x7fffffff, const ItemIO2D *io = nullptr)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int NW = blockDim.x >> 6;
    // a row takes LV = ceil(L / V) lanes, lane c holding words c V .. c V +
    // V - 1 (those past the row read as zero and are never stored); G =
    // floor(64 / LV) segments per wave; the lanes past G*LV are idle: they
    // hold zeros, load and store nothing and own no interior rows, and their
    // words reach no real lane (segment-end carries are masked)
    const int L = it.L, LV = lanes_for(L, V), G = 64 / LV, S = NW * G, SL = NW * 64;
    const int g = lane / LV, c = lane - g * LV, q = wv * G + g;
    const bool active = g < G;
    const int H = 4 * K;
    const int y_reg0 = row0 - H + q * R;  // image row of register row 0
    uint32_t mw = c == 0 ? 0u : 0xffffffffu;
    uint32_t me = c == LV - 1 ? 0u : 0xffffffffu;
    // opaque to the compiler, so `dpp(v) & mask` stays a VGPR AND that folds
    // into the DPP instruction instead of becoming a lane-mask select
    asm volatile("" : "+v"(mw), "+v"(me));
    // rows of this segment inside the band interior, as a bit mask
    const int ra = max(0, H - q * R), rb = min(R, H + nrows - q * R);
    const uint32_t interior =
        (active && rb > ra) ? (uint32_t)(((1ull << rb) - 1ull) & ~((1ull << ra) - 1ull)) : 0u;
    uint64_t *src = par ? it.buf[1] : it.buf[0];
    uint64_t *dst = par ? it.buf[0] : it.buf[1];
    const int w0 = c * V;  // first word of this lane
    constexpr bool bytes_in = IO == 2;

    Row<V> w[R];
    if (bytes_in) {
        const uint8_t *in = io->in;
        const size_t in_pitch = io->in_pitch;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int y = y_reg0 + r;
            const bool row_ok = active && y >= 0 && y < it.h;
#pragma unroll
            for (int j = 0; j < V; ++j) {
                uint64_t v = 0;
                if (row_ok && w0 + j < L) {
                    const int x0 = 64 * (w0 + j);
                    v = lane_order(pack_run_u8(in + (size_t)y * in_pitch + x0, min(64, it.w - x0)));
                }
                w[r].h[2 * j] = (uint32_t)v;
                w[r].h[2 * j + 1] = (uint32_t)(v >> 32);
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int y = y_reg0 + r;
            const bool row_ok = active && y >= 0 && y < it.h;
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const uint64_t v = row_ok && w0 + j < L ? src[(size_t)y * L + w0 + j] : 0ull;
                w[r].h[2 * j] = (uint32_t)v;
                w[r].h[2 * j + 1] = (uint32_t)(v >> 32);
            }
        }
    }
    // edges written after subiteration s live in buffer (s + 1) & 1; buffer 0
    // holds the initial rows.  Layout [buf][top/bot][NW*64][V]; the entry of
    // lane (g, c) of wave v is v*64 + g*LV + c = tid, so segment q's column c
    // is at (q / G)*64 + (q % G)*LV + c.  Idle lanes only read their own
    // entries.
    const int me_idx = tid;
    const bool has_up = active && q > 0, has_dn = active && q + 1 < S;
    const int up_idx = has_up ? ((q - 1) / G) * 64 + ((q - 1) % G) * LV + c : me_idx;
    const int dn_idx = has_dn ? ((q + 1) / G) * 64 + ((q + 1) % G) * LV + c : me_idx;
    const uint32_t up_ok = has_up ? 0xffffffffu : 0u;
    const uint32_t dn_ok = has_dn ? 0xffffffffu : 0u;
    auto put = [&](int b, int tb, const Row<V> &x) {
        uint64_t *e = edges + ((size_t)(b * 2 + tb) * SL + me_idx) * V;
#pragma unroll
        for (int j = 0; j < V; ++j) e[j] = ((uint64_t)x.h[2 * j + 1] << 32) | x.h[2 * j];
    };
    auto get = [&](int b, int tb, int idx, uint32_t ok) {
        const uint64_t *e = edges + ((size_t)(b * 2 + tb) * SL + idx) * V;
        Row<V> x;
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const uint64_t v = e[j];
            x.h[2 * j] = (uint32_t)v & ok;
            x.h[2 * j + 1] = (uint32_t)(v >> 32) & ok;
        }
        return x;
    };
    put(0, 0, w[0]);
    put(0, 1, w[R - 1]);
    if (tid == 0) {
        info->mask = 0;
        info->occ = 0;
        info->flag = 0;
    }
    __syncthreads();

    unsigned bits = 0;
    int s = 0, changed_iters = 0;
    uint32_t count = count_rows(w, interior);
    // an empty interior stays empty (subiterations only delete)
    const bool empty = IO == 2 && !__syncthreads_or(count != 0u);
    const int max_iters = empty ? 0 : whole ? whole_max : K;
#pragma unroll 1
    for (int itn = 0; itn < max_iters; ++itn) {
#define BGT_SUB(D)                                                                  \
    {                                                                               \
        const int rb_ = s & 1, wb_ = (s + 1) & 1;                                   \
        const Row<V> ue = get(rb_, 1, up_idx, up_ok), de = get(rb_, 0, dn_idx, dn_ok); \
        subiteration<D, R, V>(w, ue, de, mw, me);                                   \
        put(wb_, 0, w[0]);                                                          \
        put(wb_, 1, w[R - 1]);                                                      \
        __syncthreads();                                                            \
        ++s;                                                                        \
    }
        BGT_SUB(0) BGT_SUB(1) BGT_SUB(2) BGT_SUB(3)
#undef BGT_SUB
        const uint32_t now = count_rows(w, interior);
        const bool ch = now != count;
        count = now;
        if (ch && itn < 32) bits |= 1u << itn;
        if (!whole) continue;
        // tile-local convergence: stop after the first iteration that deletes nothing
        if (ch) atomicOr(&info->flag, 1u);
        __syncthreads();
        const unsigned any = info->flag;
        __syncthreads();
        if (!any) break;
        ++changed_iters;
        if (tid == 0) info->flag = 0;
        // the reset above is ordered before the next iteration's first write by
        // the barrier that ends its first subiteration
    }

    uint32_t occ = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = y_reg0 + r;
        if (((interior >> r) & 1u) && y >= 0 && y < it.h) {
#pragma unroll
            for (int j = 0; j < V; ++j)
                if (w0 + j < L) {
                    const uint64_t v = ((uint64_t)w[r].h[2 * j + 1] << 32) | w[r].h[2 * j];
                    dst[(size_t)y * L + w0 + j] = v;
                    if (bytes_in) src[(size_t)y * L + w0 + j] = v;
                    occ |= w[r].h[2 * j] | w[r].h[2 * j + 1];
                }
        }
    }
    if (bits) atomicOr(&info->mask, bits);
    if (occ) atomicOr(&info->occ, 1u);
    __syncthreads();
    return changed_iters;
}

}  // namespace bgt
