// thin3d_kernels.h -- gfx950 device code of the 3D thinning path.
//
// State layout in HBM: bit-packed x-lines.  Line (y, z) holds NH 32-bit words
// (NH = 2 * ceil(width / 64)); bit k of word j is voxel x = 32*j + k.  Words
// past the width are zero and stay zero (subiterations only clear bits).
//
// Work unit: a "brick" = full x-width, 64 y-lines, NW*R z-slices, processed by
// one workgroup of NW wave64s.  Lane l holds y-line l of the brick; wave v
// holds z-slices [v*R, v*R + R) in registers (R x NH words per lane).
// Neighbours: x via v_alignbit with the adjacent word of the same line, y via
// DPP wave_shr:1 / wave_shl:1 (lane -+ 1), z via the adjacent register row or,
// at a wave's first/last row, via double-buffered LDS filled by the waves above
// and below (one barrier per subiteration).  K iterations run per HBM pass;
// the brick carries H = 6K halo lines/slices in y and z and only its interior
// is written back, which is exact (Chebyshev radius 1 per subiteration; the
// protection pre-test reads the same state as the Up subiteration).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "thin2d_kernels.h"  // lane exchange helpers and item state indices

namespace bgt {

struct Item3D {
    uint32_t *buf[2];  // ping-pong packed state
    uint8_t *chg;      // per brick change flags, [2][nbricks]
    int *st;           // item state (ST_*)
    int w, h, d, nh;   // nh = 32-bit words per x-line
    int iy, iz;        // interior lines / slices per brick
    int nby, nbz, nbricks;
    int nhc;           // interior words per x chunk (nh when unchunked)
    int tile0;         // id of brick 0 in its work queue's tile arrays
    int occ0;          // offset of that queue's tiles in the batch-wide tile flags
};

struct Tile3D {
    int item;
    int brick;
};

// Neighbour index k = (dz+1)*9 + (dy+1)*3 + (dx+1) of the data-frame 3x3x3
// block that feeds argument j of the Up rule, for each direction in the
// reference's order U, N, E, D, S, W (3D/CUDA_ThinningKernel3D_Linear.cuh:
// 755-782).  Row d is a proper rotation of the cube taking direction d's face
// neighbour (the one the reference's rule requires to be background: U z-1,
// N y-1, E x-1, D z+1, S y+1, W x+1) to the Up face z-1; the reference's six
// rules (3D/CUDA_SubiterationKernels3D.cuh:14-234) equal the Up rule composed
// with these rotations on all 2^26 configurations.

template <int D, int J>
struct Rot3 {
    static constexpr int tab[6][26] = {
        {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26},
        {18, 19, 20, 9, 10, 11, 0, 1, 2, 21, 22, 23, 12, 14, 3, 4, 5, 24, 25, 26, 15, 16, 17, 6, 7, 8},
        {0, 3, 6, 9, 12, 15, 18, 21, 24, 1, 4, 7, 10, 16, 19, 22, 25, 2, 5, 8, 11, 14, 17, 20, 23, 26},
        {24, 25, 26, 21, 22, 23, 18, 19, 20, 15, 16, 17, 12, 14, 9, 10, 11, 6, 7, 8, 3, 4, 5, 0, 1, 2},
        {6, 7, 8, 15, 16, 17, 24, 25, 26, 3, 4, 5, 12, 14, 21, 22, 23, 0, 1, 2, 9, 10, 11, 18, 19, 20},
        {20, 23, 26, 11, 14, 17, 2, 5, 8, 19, 22, 25, 10, 16, 1, 4, 7, 18, 21, 24, 9, 12, 15, 0, 3, 6},
    };
    static constexpr int value = tab[D][J];
};

// The Up rule (3D/CUDA_SubiterationKernels3D.cuh:14-49): the mask of voxels
// whose 26-neighbourhood m0..m25 (raster (dz, dy, dx) order without the
// centre) matches one of the deletion templates M1..M6 with the z-1 face
// neighbour m4 in the background.
__device__ __forceinline__ uint32_t up_rule(const uint32_t *m)
{
#define M(i) (m[i])
#define Z(i) (~m[i])
    const uint32_t s1 = M(9) | M(10) | M(11) | M(12) | M(13) | M(14) | M(15) | M(16) | M(17) | M(19) | M(23) | M(25);
    const uint32_t s2 = M(20) | M(22);
    const uint32_t s3 = M(18) | M(24);
    const uint32_t q1 = Z(15) & Z(16) & M(18) & Z(24) & Z(25);
    const uint32_t q2 = Z(11) & Z(10) & Z(18) & Z(19) & M(24);
    const uint32_t q3 = Z(13) & Z(22) & M(20);
    uint32_t q4 = Z(12) & Z(20) & M(22);
    const uint32_t q6 = Z(14) & Z(23);
    const uint32_t q5 = q4 & q6;
    q4 &= Z(9) & Z(17);
    const uint32_t m1 = s1 | s2 | s3;
    const uint32_t m5 = ((s1 | s2) & ((Z(14) & Z(23) & q1) | (Z(9) & Z(17) & q2))) |
                        ((s1 | s3) & ((q4 & q6) | (Z(11) & Z(16) & Z(19) & Z(25) & q3)));
    const uint32_t m6 = (Z(24) & Z(15) & M(18) & q5) | (Z(18) & Z(10) & M(24) & q4) | (q3 & q2) | (q3 & q1);
    const uint32_t m4 = (Z(0) & Z(6) & ((M(2) & M(11) & Z(8)) | (M(8) & M(16) & Z(2)))) |
                        (Z(2) & Z(8) & ((M(0) & M(9) & Z(6)) | (M(6) & M(14) & Z(0))));
    const uint32_t m2 = (Z(1) & Z(2) & Z(5) & ((Z(0) & Z(3) & M(15)) | (Z(7) & Z(8) & M(12)))) |
                        (Z(7) & Z(6) & Z(3) & ((Z(5) & Z(8) & M(10)) | (Z(0) & Z(1) & M(13))));
    const uint32_t m3 = (Z(3) & M(13) & ((Z(7) & Z(6) & M(10)) | (Z(0) & Z(1) & M(15)))) |
                        (Z(5) & M(12) & ((Z(1) & Z(2) & M(15)) | (Z(7) & Z(8) & M(10))));
    return Z(4) & (((Z(1) & Z(3) & Z(5) & Z(7)) &
                    ((M(21) & m4) | ((Z(0) & Z(2) & Z(6) & Z(8)) & ((M(21) & m1) | (Z(21) & (m5 | m6)))))) |
                   (M(21) & (m2 | m3)));
#undef M
#undef Z
}

template <int D>
__device__ __forceinline__ uint32_t rule3(const uint32_t (&nb)[27])
{
    const uint32_t m[26] = {
        nb[Rot3<D, 0>::value],  nb[Rot3<D, 1>::value],  nb[Rot3<D, 2>::value],  nb[Rot3<D, 3>::value],
        nb[Rot3<D, 4>::value],  nb[Rot3<D, 5>::value],  nb[Rot3<D, 6>::value],  nb[Rot3<D, 7>::value],
        nb[Rot3<D, 8>::value],  nb[Rot3<D, 9>::value],  nb[Rot3<D, 10>::value], nb[Rot3<D, 11>::value],
        nb[Rot3<D, 12>::value], nb[Rot3<D, 13>::value], nb[Rot3<D, 14>::value], nb[Rot3<D, 15>::value],
        nb[Rot3<D, 16>::value], nb[Rot3<D, 17>::value], nb[Rot3<D, 18>::value], nb[Rot3<D, 19>::value],
        nb[Rot3<D, 20>::value], nb[Rot3<D, 21>::value], nb[Rot3<D, 22>::value], nb[Rot3<D, 23>::value],
        nb[Rot3<D, 24>::value], nb[Rot3<D, 25>::value]};
    return up_rule(m);
}

// 27-neighbourhood of 32-bit word j of the current z-row (zc), with the rows
// below (zm, z-1) and above (zp, z+1) in z; y via the adjacent lanes.
template <int NH>
__device__ __forceinline__ void gather27(const uint32_t (&zm)[NH], const uint32_t (&zc)[NH],
                                         const uint32_t (&zp)[NH], int j, uint32_t (&nb)[27])
{
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
        uint32_t a, b, c;
        if (dz == 0) { a = j > 0 ? zm[j > 0 ? j - 1 : 0] : 0u; b = zm[j]; c = j + 1 < NH ? zm[j + 1 < NH ? j + 1 : 0] : 0u; }
        else if (dz == 1) { a = j > 0 ? zc[j > 0 ? j - 1 : 0] : 0u; b = zc[j]; c = j + 1 < NH ? zc[j + 1 < NH ? j + 1 : 0] : 0u; }
        else { a = j > 0 ? zp[j > 0 ? j - 1 : 0] : 0u; b = zp[j]; c = j + 1 < NH ? zp[j + 1 < NH ? j + 1 : 0] : 0u; }
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
            uint32_t av = a, bv = b, cv = c;
            if (dy == 0) { av = from_lower_lane(a); bv = from_lower_lane(b); cv = from_lower_lane(c); }
            if (dy == 2) { av = from_upper_lane(a); bv = from_upper_lane(b); cv = from_upper_lane(c); }
            const int k = dz * 9 + dy * 3;
            nb[k + 0] = __builtin_amdgcn_alignbit(bv, av, 31);  // x-1
            nb[k + 1] = bv;
            nb[k + 2] = __builtin_amdgcn_alignbit(cv, bv, 1);   // x+1
        }
    }
}

// One subiteration of direction D (0..5 = U, N, E, D, S, W) over the R x NH
// register words of a lane.  The Up step also evaluates the protection mask P
// = voxels deletable in no direction (3D/CUDA_ThinningKernel3D_Linear.cuh:
// 690-702); later steps delete only voxels outside P.  Bit r of rows_changed is
// set when z-row r lost a voxel in one of the interior words [xa, xa + nx).
template <int D, int NH, int R>
__device__ __forceinline__ void step3d(uint32_t (&s)[R][NH], uint32_t *P,
                                       const uint32_t *edges_up, const uint32_t *edges_dn,
                                       uint32_t up_ok, uint32_t dn_ok, int lane, int xa, int nx,
                                       uint32_t &rows_changed)
{
    uint32_t pv[NH];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint32_t zm[NH], zp[NH];
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            if (r == 0) zm[j] = edges_up[j * 64 + lane] & up_ok;
            else zm[j] = pv[j];
            if (r + 1 < R) zp[j] = s[r + 1 < R ? r + 1 : 0][j];
            else zp[j] = edges_dn[j * 64 + lane] & dn_ok;
        }
        uint32_t nw[NH];
        uint32_t racc = 0;
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            uint32_t nb[27];
            gather27<NH>(zm, s[r], zp, j, nb);
            const uint32_t v = s[r][j];
            uint32_t del;
            if (D == 0) {
                const uint32_t cu = rule3<0>(nb);
                const uint32_t any = cu | rule3<1>(nb) | rule3<2>(nb) | rule3<3>(nb) |
                                     rule3<4>(nb) | rule3<5>(nb);
                P[r * NH + j] = v & ~any;
                del = cu & v;
            } else {
                del = rule3<D>(nb) & v & ~P[r * NH + j];
            }
            nw[j] = v ^ del;
            if ((unsigned)(j - xa) < (unsigned)nx) racc |= del;
            __builtin_amdgcn_sched_barrier(0);
        }
        rows_changed |= (racc != 0u ? 1u : 0u) << r;
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            pv[j] = s[r][j];
            s[r][j] = nw[j];
        }
    }
}

// Run one pass over brick `brick` of item `it`: K iterations (6K subiterations
// after the protection pre-test of each) from buffer `par` into buffer par^1.
// `lds` holds the z-edge exchange buffers and the protection masks.
template <int NH, int R>
__device__ void brick_work(const Item3D &it, int brick, int K, int par, uint32_t *edges3,
                           TileInfo *info)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int NW = blockDim.x >> 6;
    // brick index = (bz * nby + by) * nbx + bx
    const int nbx = it.nbricks / (it.nby * it.nbz);
    const int bx = brick % nbx, by = (brick / nbx) % it.nby, bz = brick / (nbx * it.nby);
    if (tid == 0) {
        info->mask = 0;
        info->occ = 0;
    }
    const int H = 6 * K;
    const int xhalo = nbx > 1 ? 1 : 0;
    const int nhc = NH - 2 * xhalo;            // interior words per brick
    const int j_mem0 = bx * nhc - xhalo;       // memory word of register word 0
    const int y = by * it.iy - H + lane;
    const int z_reg0 = bz * it.iz - H + wv * R;
    const bool y_ok = y >= 0 && y < it.h;
    const uint32_t ymask = (lane >= H && lane < H + it.iy) ? 0xffffffffu : 0u;
    const int ra = max(0, H - wv * R), rb = min(R, H + it.iz - wv * R);
    const uint32_t zint = rb > ra ? (uint32_t)(((1ull << rb) - 1ull) & ~((1ull << ra) - 1ull)) : 0u;
    const int xa = xhalo, nx = min(nhc, it.nh - bx * nhc);
    const uint32_t *src = par ? it.buf[1] : it.buf[0];
    uint32_t *dst = par ? it.buf[0] : it.buf[1];

    uint32_t s[R][NH];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int z = z_reg0 + r;
        const bool ok = y_ok && z >= 0 && z < it.d;
        const size_t base = ((size_t)z * it.h + y) * it.nh;
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            const int jm = j_mem0 + j;
            s[r][j] = (ok && jm >= 0 && jm < it.nh) ? src[base + jm] : 0u;
        }
    }
    // protection masks of this lane after the edge buffers: R*NH words per
    // lane at a stride of R*NH+1 words (odd, so the lanes hit distinct banks)
    uint32_t *P = edges3 + 2 * 2 * NW * NH * 64 + tid * (R * NH + 1);
    const int plane = NH * 64;  // words per (buf, tb, wave) block
    auto eidx = [&](int b, int tb, int wave) { return ((b * 2 + tb) * NW + wave) * plane; };
    const int wup = wv > 0 ? wv - 1 : wv, wdn = wv + 1 < NW ? wv + 1 : wv;
    const uint32_t up_ok = wv > 0 ? 0xffffffffu : 0u, dn_ok = wv + 1 < NW ? 0xffffffffu : 0u;
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        edges3[eidx(0, 0, wv) + j * 64 + lane] = s[0][j];
        edges3[eidx(0, 1, wv) + j * 64 + lane] = s[R - 1][j];
    }
    __syncthreads();

    unsigned bits = 0;
    int sidx = 0;
#pragma unroll 1
    for (int itn = 0; itn < K; ++itn) {
        uint32_t rows_changed = 0;
#define BGT_STEP3(D)                                                                          \
{                                                                                         \
    const int rb_ = sidx & 1, wb_ = (sidx + 1) & 1;                                       \
    step3d<D, NH, R>(s, P, edges3 + eidx(rb_, 1, wup), edges3 + eidx(rb_, 0, wdn), up_ok, \
                     dn_ok, lane, xa, nx, rows_changed);                                  \
    _Pragma("unroll") for (int j = 0; j < NH; ++j)                                        \
    {                                                                                     \
        edges3[eidx(wb_, 0, wv) + j * 64 + lane] = s[0][j];                               \
        edges3[eidx(wb_, 1, wv) + j * 64 + lane] = s[R - 1][j];                           \
    }                                                                                     \
    __syncthreads();                                                                      \
    ++sidx;                                                                               \
}
        BGT_STEP3(0) BGT_STEP3(1) BGT_STEP3(2) BGT_STEP3(3) BGT_STEP3(4) BGT_STEP3(5)
#undef BGT_STEP3
        if (ymask && (rows_changed & zint)) bits |= 1u << itn;
    }

    if (ymask && y_ok) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int z = z_reg0 + r;
            if (((zint >> r) & 1u) && z >= 0 && z < it.d) {
                const size_t base = ((size_t)z * it.h + y) * it.nh;
#pragma unroll
                for (int j = 0; j < NH; ++j) {
                    const int jm = j_mem0 + j;
                    if ((unsigned)(j - xa) < (unsigned)nx) dst[base + jm] = s[r][j];
                }
            }
        }
    }
    uint32_t occ = 0;
    if (ymask && y_ok) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int z = z_reg0 + r;
            if (((zint >> r) & 1u) && z >= 0 && z < it.d) {
#pragma unroll
                for (int j = 0; j < NH; ++j)
                    if ((unsigned)(j - xa) < (unsigned)nx) occ |= s[r][j];
            }
        }
    }
    if (bits) atomicOr(&info->mask, bits);
    if (occ) atomicOr(&info->occ, 1u);
    __syncthreads();
}

}  // namespace bgt
