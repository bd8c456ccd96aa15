// thin3d_kernels.h -- the 3D deletion rules as word-parallel device code.
//
// State layout in HBM: bit-packed x-lines.  Line (y, z) holds nh 32-bit words
// (nh = 2 * ceil(width / 64)); bit k of word j is voxel x = 32*j + k.  Words
// past the width are zero and stay zero (subiterations only clear bits).  The
// rules take the 27 words of a 3x3x3 neighbourhood, each already shifted so
// that bit k holds the neighbour of voxel k, and return the mask of voxels the
// direction's templates delete.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "thin2d_kernels.h"  // item state indices (ST_*)

namespace bgt {

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

}  // namespace bgt
