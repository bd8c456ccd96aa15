// harness_pred3d.cpp -- test-only driver around the reference's 3D
// subiteration predicates (3D/CUDA_SubiterationKernels3D.cuh, included
// unmodified; the header's only non-C++ token, the CUDA execution-space
// qualifier __device__, is defined empty on the compiler command line).
// Produces the exhaustive 2^26-entry table of the combined deletion mask of
// one direction.  Never linked into the product.
#include <cstdint>
#include "CUDA_SubiterationKernels3D.cuh"

typedef unsigned int (*pred_fn)(unsigned int, unsigned int *, unsigned int);

// direction order: 0=U 1=N 2=E 3=D 4=S 5=W (the reference's iteration order)
extern "C" void ref3d_table(int dir, uint32_t *out /* 1 << 21 words */)
{
    static const pred_fn fns[6] = {Up3D, North3D, East3D, Down3D, South3D, West3D};
    static const uint32_t lowpat[5] = {0xAAAAAAAAu, 0xCCCCCCCCu, 0xF0F0F0F0u, 0xFF00FF00u, 0xFFFF0000u};
    unsigned int n[26];
    for (int i = 0; i < 5; i++) n[i] = lowpat[i];
    for (uint32_t hi = 0; hi < (1u << 21); hi++) {
        for (int i = 5; i < 26; i++) n[i] = ((hi >> (i - 5)) & 1) ? 0xFFFFFFFFu : 0u;
        // X3D(val, n, prev) = val & ~(combined & ~prev); with val = ~0, prev = 0
        out[hi] = ~fns[dir](0xFFFFFFFFu, n, 0u);
    }
}
