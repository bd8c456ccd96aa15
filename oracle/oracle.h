/*
 * oracle.h -- CPU restatement of BEGPUThinning's thinning path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP product path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links it.
 *
 * Semantics restated (reference file:line in each function's comment in
 * oracle.c):
 *   - 2D directional subiteration predicates N, S, E, W
 *       2D/CUDA_SubiterationKernels2D.cuh:15-119
 *   - 2D iteration: N, S, E, W, each a synchronous parallel update, zero outside
 *       2D/CUDA_ThinningKernelMultiBlock2D.cuh:99-154
 *   - 3D directional predicates Up, Down, South, North, East, West
 *       3D/CUDA_SubiterationKernels3D.cuh:14-234
 *   - 3D iteration: protection mask P, then U, N, E, D, S, W
 *       3D/CUDA_ThinningKernel3D_Linear.cuh:690-782
 *   - ExtractSegmentsV 2D / 3D
 *       2D/ExtractSegments.cpp:16-127, 3D/ExtractSegments.cpp:135-251
 *   - CenterlineExtraction pad / prune / smooth / flatten
 *       2D/CenterlineExtraction.cpp:22-142, 3D/CenterlineExtraction.cpp:22-155
 *
 * Parity status of each function is recorded in the header comment of
 * oracle.c and in the tests that pin it.
 */
#ifndef BGT_ORACLE_H
#define BGT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* direction codes.  2D: 0=N 1=S 2=E 3=W (reference iteration order).
 * 3D: 0=U 1=N 2=E 3=D 4=S 5=W (reference iteration order). */

/* 2D template match on 64-bit words.  up/cp/dn are rows y-1, y, y+1 of one
 * 64-pixel word column; the L/R words are the same rows shifted so that bit k
 * holds pixel x-1 (L, the reference's ShiftLeft) or x+1 (R, ShiftRight).
 * Returns the deletion template mask before it is restricted to cp. */
uint64_t or2d_match(int dir, uint64_t up, uint64_t cp, uint64_t dn,
                    uint64_t Lup, uint64_t Rup, uint64_t Lcp, uint64_t Rcp,
                    uint64_t Ldn, uint64_t Rdn);

/* 512-entry table of one 2D direction: entry c has bit 0..7 = x0..x7
 * (x0=N, x1=NE, x2=E, x3=SE, x4=S, x5=SW, x6=W, x7=NW) and bit 8 = centre;
 * out[c] = 1 iff the centre pixel is deleted. */
void or2d_table(int dir, uint8_t out[512]);

/* Thin a w x h uint8 image (nonzero = foreground) in place to the global
 * fixed point; returns the number of iterations executed including the final
 * no-change one (the reference's iteration counter semantics).  max_iter <= 0
 * means unbounded. */
int or2d_thin(uint8_t *img, int w, int h, int max_iter);

/* Run exactly n_iter iterations (or fewer when converged); returns the number
 * of iterations in which something was deleted. */
int or2d_iterate(uint8_t *img, int w, int h, int n_iter);

/* 3D combined mask C_d for one direction over 64 configurations at once:
 * n[i] is the word of neighbour i (raster (dz,dy,dx) order skipping centre). */
uint64_t or3d_comb(int dir, const uint64_t n[26]);

/* Full 2^26-entry table (bit-packed, 8 MiB) of C_d for one direction:
 * bit c of the table = C_d of configuration c (bit i of c = neighbour i). */
void or3d_table(int dir, uint64_t *out /* 1<<20 words */);

/* Thin a w x h x d uint8 volume (x fastest) in place; returns iterations
 * including the final no-change one. */
int or3d_thin(uint8_t *vol, int w, int h, int d, int max_iter);
int or3d_iterate(uint8_t *vol, int w, int h, int d, int n_iter);

/* Segments as produced by ExtractSegmentsV: points (x,y[,z]) as int16,
 * start offsets seg_start[nseg+1], is_end flag per segment. */
typedef struct {
    int nseg;
    int npts;
    int16_t *x, *y, *z;     /* z is NULL for 2D */
    uint32_t *seg_start;    /* nseg + 1 entries */
    uint8_t *is_end;        /* nseg entries */
} or_segments;

int or2d_extract(const uint8_t *img, int w, int h, or_segments *out);
int or3d_extract(const uint8_t *vol, int w, int h, int d, or_segments *out);
void or_segments_free(or_segments *s);

/* CenterlineExtraction: returns nsegs; out arrays malloc'd (free with
 * or_free).  For 2D, *out_z is left untouched. */
typedef struct {
    int nseg;
    int npts;
    float *x, *y, *z;
    uint32_t *idx;          /* nseg + 1 entries */
    int iterations;         /* oracle iteration count of the thinning */
} or_centerline;

int or2d_centerline(const uint8_t *img, int w, int h, int max_segment_length,
                    int smoothing_span, or_centerline *out);
int or3d_centerline(const uint8_t *vol, int w, int h, int d, int max_segment_length,
                    int smoothing_span, or_centerline *out);
void or_centerline_free(or_centerline *c);

/* Smooth one point list (the moving average of CenterlineExtraction).  n
 * points in px/py/pz (pz may be NULL), output n floats each. */
void or_smooth(const int16_t *px, const int16_t *py, const int16_t *pz, int n,
               int smoothing_span, float *ox, float *oy, float *oz);

void or_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
