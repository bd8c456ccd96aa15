/*
 * oracle.c -- CPU restatement of BEGPUThinning's thinning path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Nothing in the product library
 * includes, links or calls this file.
 *
 * The arithmetic is bit-sliced on uint64_t words so that configuration-size
 * inputs finish in seconds, but every rule is the reference's scalar boolean
 * rule, applied with the reference's global semantics:
 *   - every subiteration is a synchronous parallel update of the whole image;
 *   - everything outside the image is background;
 *   - directions run in the reference's order.
 *
 * How each part is pinned (tests/ hold the checks):
 *   or2d_match / or2d_table : the four directions are transcribed separately
 *       from the reference text; the tests check exhaustively that E, S, W are
 *       N under the 90/180/270 degree rotations of the neighbourhood, that
 *       every deletable configuration is a simple border point of its
 *       direction, and the committed 512-entry tables.
 *   or3d_comb / or3d_table  : exhaustive 2^26 tables against the reference
 *       header compiled on the host (oracle/_ref), plus the rotation
 *       relations between the six directions.
 *   or2d_thin / or3d_thin   : brute force against a scalar per-pixel loop on
 *       small frames; topology (Euler characteristic and component counts)
 *       preserved; fixed point; order sensitivity.
 *   or2d_extract / or3d_extract : the reference ExtractSegments.cpp compiled
 *       on the host (oracle/_ref) on hand-made shapes and seeded skeletons.
 *   or_smooth               : closed form (exact integer window sums divided
 *       in float32), valid while window sums stay below 2^24.
 *   or2d_centerline / or3d_centerline : composition of the above; padding
 *       rule pinned by hand-computed cases.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* 2D predicates: 2D/CUDA_SubiterationKernels2D.cuh                           */
/* ------------------------------------------------------------------------- */

/* North: 2D/CUDA_SubiterationKernels2D.cuh:15-40 */
static uint64_t match_north(uint64_t up, uint64_t cp, uint64_t dn, uint64_t Lup,
                            uint64_t Rup, uint64_t Lcp, uint64_t Rcp, uint64_t Ldn,
                            uint64_t Rdn)
{
    (void)cp;
    uint64_t x6 = Lcp, x1 = Rup, x2 = Rcp, x7 = Lup, x3 = Rdn, x5 = Ldn, x4 = dn;
    uint64_t a = ~x1 & (x2 ^ x3 ^ ~x4) & (x3 ^ x5) & (x3 ^ x6) & ~x7;
    uint64_t b = (x1 ^ x5) & (x2 ^ x5) & (x4 ^ ~x5) & ~x6 & ~x7;
    uint64_t c = ~x1 & ~x2 & ~x3 & x6 & x7;
    uint64_t e = x2 & (x4 & (x6 | ~x7));
    uint64_t f = x6 & (~x1 & x4);
    return (a | e | f | b | c) & ~up;
}

/* East: 2D/CUDA_SubiterationKernels2D.cuh:42-66 */
static uint64_t match_east(uint64_t up, uint64_t cp, uint64_t dn, uint64_t Lup,
                           uint64_t Rup, uint64_t Lcp, uint64_t Rcp, uint64_t Ldn,
                           uint64_t Rdn)
{
    (void)cp;
    uint64_t x6 = Lcp, x3 = Rdn, x1 = Rup, x5 = Ldn, x7 = Lup, x4 = dn, x0 = up;
    uint64_t a = ~x3 & (x4 ^ x5 ^ ~x6) & (x5 ^ x7) & (x5 ^ x0) & ~x1;
    uint64_t b = (x3 ^ x7) & (x4 ^ x7) & (x6 ^ ~x7) & ~x0 & ~x1;
    uint64_t c = ~x3 & ~x4 & ~x5 & x0 & x1;
    uint64_t e = x4 & x6 & (x0 | ~x1);
    uint64_t f = x6 & (~x3 & x0);
    return (a | e | f | b | c) & ~Rcp;
}

/* South: 2D/CUDA_SubiterationKernels2D.cuh:68-93 */
static uint64_t match_south(uint64_t up, uint64_t cp, uint64_t dn, uint64_t Lup,
                            uint64_t Rup, uint64_t Lcp, uint64_t Rcp, uint64_t Ldn,
                            uint64_t Rdn)
{
    (void)cp;
    uint64_t x2 = Rcp, x5 = Ldn, x6 = Lcp, x3 = Rdn, x7 = Lup, x1 = Rup, x0 = up;
    uint64_t a = ~x5 & (x6 ^ x7 ^ ~x0) & (x7 ^ x1) & (x7 ^ x2) & ~x3;
    uint64_t b = (x5 ^ x1) & (x6 ^ x1) & (x0 ^ ~x1) & ~x2 & ~x3;
    uint64_t c = ~x5 & ~x6 & ~x7 & x2 & x3;
    uint64_t e = x6 & (x0 & (x2 | ~x3));
    uint64_t f = x2 & (~x5 & x0);
    return (a | e | f | b | c) & ~dn;
}

/* West: 2D/CUDA_SubiterationKernels2D.cuh:95-119 */
static uint64_t match_west(uint64_t up, uint64_t cp, uint64_t dn, uint64_t Lup,
                           uint64_t Rup, uint64_t Lcp, uint64_t Rcp, uint64_t Ldn,
                           uint64_t Rdn)
{
    (void)cp;
    uint64_t x2 = Rcp, x7 = Lup, x5 = Ldn, x1 = Rup, x3 = Rdn, x0 = up, x4 = dn;
    uint64_t a = ~x7 & (x0 ^ x1 ^ ~x2) & (x1 ^ x3) & (x1 ^ x4) & ~x5;
    uint64_t b = (x7 ^ x3) & (x0 ^ x3) & (x2 ^ ~x3) & ~x4 & ~x5;
    uint64_t c = ~x7 & ~x0 & ~x1 & x4 & x5;
    uint64_t e = x0 & x2 & (x4 | ~x5);
    uint64_t f = x2 & (~x7 & x4);
    return (a | e | f | b | c) & ~Lcp;
}

uint64_t or2d_match(int dir, uint64_t up, uint64_t cp, uint64_t dn, uint64_t Lup,
                    uint64_t Rup, uint64_t Lcp, uint64_t Rcp, uint64_t Ldn, uint64_t Rdn)
{
    switch (dir) {
    case 0: return match_north(up, cp, dn, Lup, Rup, Lcp, Rcp, Ldn, Rdn);
    case 1: return match_south(up, cp, dn, Lup, Rup, Lcp, Rcp, Ldn, Rdn);
    case 2: return match_east(up, cp, dn, Lup, Rup, Lcp, Rcp, Ldn, Rdn);
    default: return match_west(up, cp, dn, Lup, Rup, Lcp, Rcp, Ldn, Rdn);
    }
}

void or2d_table(int dir, uint8_t out[512])
{
    /* one configuration per call, centre at bit 1 of a 3-bit row:
     * bit 0 = x-1, bit 1 = x, bit 2 = x+1 */
    for (int c = 0; c < 512; c++) {
        int xb[8];
        for (int i = 0; i < 8; i++) xb[i] = (c >> i) & 1;
        int p = (c >> 8) & 1;
        uint64_t up = (uint64_t)xb[7] | ((uint64_t)xb[0] << 1) | ((uint64_t)xb[1] << 2);
        uint64_t cp = (uint64_t)xb[6] | ((uint64_t)p << 1) | ((uint64_t)xb[2] << 2);
        uint64_t dn = (uint64_t)xb[5] | ((uint64_t)xb[4] << 1) | ((uint64_t)xb[3] << 2);
        uint64_t m = or2d_match(dir, up, cp, dn, up << 1, up >> 1, cp << 1, cp >> 1,
                                dn << 1, dn >> 1);
        out[c] = (uint8_t)((m & cp) >> 1 & 1);
    }
}

/* ------------------------------------------------------------------------- */
/* 2D iteration: 2D/CUDA_ThinningKernelMultiBlock2D.cuh:99-154 (order N,S,E,W; */
/* each subiteration reads only the state left by the previous one)          */
/* ------------------------------------------------------------------------- */

typedef struct {
    int w, h, nw;     /* nw = 64-bit words per row */
    uint64_t *a, *b;  /* current state, scratch */
} grid2d;

static void pack2d(grid2d *g, const uint8_t *img)
{
    memset(g->a, 0, sizeof(uint64_t) * (size_t)g->nw * g->h);
    for (int y = 0; y < g->h; y++)
        for (int x = 0; x < g->w; x++)
            if (img[(size_t)y * g->w + x]) g->a[(size_t)y * g->nw + (x >> 6)] |= 1ull << (x & 63);
}

static void unpack2d(const grid2d *g, uint8_t *img)
{
    for (int y = 0; y < g->h; y++)
        for (int x = 0; x < g->w; x++)
            img[(size_t)y * g->w + x] = (g->a[(size_t)y * g->nw + (x >> 6)] >> (x & 63)) & 1;
}

static inline uint64_t shl_w(const uint64_t *row, int w, int nw)
{   /* bit k <- pixel x-1 */
    (void)nw;
    return (row[w] << 1) | (w > 0 ? row[w - 1] >> 63 : 0);
}
static inline uint64_t shr_w(const uint64_t *row, int w, int nw)
{   /* bit k <- pixel x+1 */
    return (row[w] >> 1) | (w + 1 < nw ? row[w + 1] << 63 : 0);
}

static int subiter2d(grid2d *g, int dir)
{
    int changed = 0;
    const int nw = g->nw;
    static const uint64_t zero_row_storage[1] = {0};
    (void)zero_row_storage;
    uint64_t *zrow = (uint64_t *)calloc((size_t)nw, sizeof(uint64_t));
    for (int y = 0; y < g->h; y++) {
        const uint64_t *rc = g->a + (size_t)y * nw;
        const uint64_t *ru = y > 0 ? rc - nw : zrow;
        const uint64_t *rd = y + 1 < g->h ? rc + nw : zrow;
        uint64_t *ro = g->b + (size_t)y * nw;
        for (int w = 0; w < nw; w++) {
            uint64_t cp = rc[w];
            if (!cp) { ro[w] = 0; continue; }
            uint64_t m = or2d_match(dir, ru[w], cp, rd[w], shl_w(ru, w, nw), shr_w(ru, w, nw),
                                    shl_w(rc, w, nw), shr_w(rc, w, nw), shl_w(rd, w, nw),
                                    shr_w(rd, w, nw));
            uint64_t del = m & cp;
            ro[w] = cp & ~del;
            if (del) changed = 1;
        }
    }
    free(zrow);
    uint64_t *t = g->a; g->a = g->b; g->b = t;
    return changed;
}

static int run2d(uint8_t *img, int w, int h, int max_iter, int count_only_changed)
{
    if (w <= 0 || h <= 0) return 0;
    grid2d g;
    g.w = w; g.h = h; g.nw = (w + 63) >> 6;
    g.a = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)g.nw * h);
    g.b = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)g.nw * h);
    pack2d(&g, img);
    int it = 0, nchanged = 0;
    for (;;) {
        if (max_iter > 0 && it >= max_iter) break;
        it++;
        int ch = 0;
        ch |= subiter2d(&g, 0); /* N */
        ch |= subiter2d(&g, 1); /* S */
        ch |= subiter2d(&g, 2); /* E */
        ch |= subiter2d(&g, 3); /* W */
        if (!ch) break;
        nchanged++;
    }
    unpack2d(&g, img);
    free(g.a); free(g.b);
    return count_only_changed ? nchanged : it;
}

int or2d_thin(uint8_t *img, int w, int h, int max_iter) { return run2d(img, w, h, max_iter, 0); }
int or2d_iterate(uint8_t *img, int w, int h, int n_iter) { return run2d(img, w, h, n_iter, 1); }

/* ------------------------------------------------------------------------- */
/* 3D predicates: 3D/CUDA_SubiterationKernels3D.cuh (combined masks C_d).     */
/* Neighbour i <-> offset j = i + (i >= 13): dz = j/9-1, dy = (j/3)%3-1,      */
/* dx = j%3-1 (3D/CUDA_ThinningKernel3D_Linear.cuh:57-174).                   */
/* ------------------------------------------------------------------------- */

#define N_(i) (n[i])
#define Z_(i) (~n[i])

/* Up: 3D/CUDA_SubiterationKernels3D.cuh:14-49 */
static uint64_t comb_up(const uint64_t *n)
{
    uint64_t s1 = N_(9) | N_(10) | N_(11) | N_(12) | N_(13) | N_(14) | N_(15) | N_(16) | N_(17) | N_(19) | N_(23) | N_(25);
    uint64_t s2 = N_(20) | N_(22);
    uint64_t s3 = N_(18) | N_(24);
    uint64_t q1 = Z_(15) & Z_(16) & N_(18) & Z_(24) & Z_(25);
    uint64_t q2 = Z_(11) & Z_(10) & Z_(18) & Z_(19) & N_(24);
    uint64_t q3 = Z_(13) & Z_(22) & N_(20);
    uint64_t q4 = Z_(12) & Z_(20) & N_(22);
    uint64_t q6 = Z_(14) & Z_(23);
    uint64_t q5 = q4 & q6;
    q4 &= Z_(9) & Z_(17);
    uint64_t m1 = s1 | s2 | s3;
    uint64_t m5 = ((s1 | s2) & ((Z_(14) & Z_(23) & q1) | (Z_(9) & Z_(17) & q2))) |
                  ((s1 | s3) & ((q4 & q6) | (Z_(11) & Z_(16) & Z_(19) & Z_(25) & q3)));
    uint64_t m6 = (Z_(24) & Z_(15) & N_(18) & q5) | (Z_(18) & Z_(10) & N_(24) & q4) | (q3 & q2) | (q3 & q1);
    uint64_t m4 = (Z_(0) & Z_(6) & ((N_(2) & N_(11) & Z_(8)) | (N_(8) & N_(16) & Z_(2)))) |
                  (Z_(2) & Z_(8) & ((N_(0) & N_(9) & Z_(6)) | (N_(6) & N_(14) & Z_(0))));
    uint64_t m2 = (Z_(1) & Z_(2) & Z_(5) & ((Z_(0) & Z_(3) & N_(15)) | (Z_(7) & Z_(8) & N_(12)))) |
                  (Z_(7) & Z_(6) & Z_(3) & ((Z_(5) & Z_(8) & N_(10)) | (Z_(0) & Z_(1) & N_(13))));
    uint64_t m3 = (Z_(3) & N_(13) & ((Z_(7) & Z_(6) & N_(10)) | (Z_(0) & Z_(1) & N_(15)))) |
                  (Z_(5) & N_(12) & ((Z_(1) & Z_(2) & N_(15)) | (Z_(7) & Z_(8) & N_(10))));
    return Z_(4) & (((Z_(1) & Z_(3) & Z_(5) & Z_(7)) &
                     ((N_(21) & m4) | ((Z_(0) & Z_(2) & Z_(6) & Z_(8)) &
                                       ((N_(21) & m1) | (Z_(21) & (m5 | m6)))))) |
                    (N_(21) & (m2 | m3)));
}

/* Down: 3D/CUDA_SubiterationKernels3D.cuh:51-86 */
static uint64_t comb_down(const uint64_t *n)
{
    uint64_t s1 = N_(11) | N_(10) | N_(9) | N_(13) | N_(12) | N_(16) | N_(15) | N_(14) | N_(2) | N_(0) | N_(8) | N_(6);
    uint64_t s2 = N_(5) | N_(3);
    uint64_t s3 = N_(1) | N_(7);
    uint64_t q1 = Z_(15) & Z_(14) & N_(1) & Z_(7) & Z_(6);
    uint64_t q2 = Z_(9) & Z_(10) & Z_(1) & Z_(0) & N_(7);
    uint64_t q3 = Z_(12) & Z_(3) & N_(5);
    uint64_t q4 = Z_(13) & Z_(5) & N_(3);
    uint64_t q6 = Z_(16) & Z_(8);
    uint64_t q5 = q4 & q6;
    q4 &= Z_(11) & Z_(2);
    uint64_t m1 = s1 | s2 | s3;
    uint64_t m5 = ((s1 | s2) & ((Z_(16) & Z_(8) & q1) | (Z_(11) & Z_(2) & q2))) |
                  ((s1 | s3) & ((q4 & q6) | (Z_(9) & Z_(14) & Z_(0) & Z_(6) & q3)));
    uint64_t m6 = (Z_(7) & Z_(15) & N_(1) & q5) | (Z_(1) & Z_(10) & N_(7) & q4) | (q3 & q2) | (q3 & q1);
    uint64_t m4 = (Z_(19) & Z_(25) & ((N_(17) & N_(9) & Z_(23)) | (N_(23) & N_(14) & Z_(17)))) |
                  (Z_(17) & Z_(23) & ((N_(19) & N_(11) & Z_(25)) | (N_(25) & N_(16) & Z_(19))));
    uint64_t m2 = (Z_(18) & Z_(17) & Z_(20) & ((Z_(19) & Z_(22) & N_(15)) | (Z_(24) & Z_(23) & N_(13)))) |
                  (Z_(24) & Z_(25) & Z_(22) & ((Z_(20) & Z_(23) & N_(10)) | (Z_(19) & Z_(18) & N_(12))));
    uint64_t m3 = (Z_(22) & N_(12) & ((Z_(24) & Z_(25) & N_(10)) | (Z_(19) & Z_(18) & N_(15)))) |
                  (Z_(20) & N_(13) & ((Z_(18) & Z_(17) & N_(15)) | (Z_(24) & Z_(23) & N_(10))));
    return Z_(21) & (((Z_(18) & Z_(22) & Z_(20) & Z_(24)) &
                      ((N_(4) & m4) | ((Z_(19) & Z_(17) & Z_(25) & Z_(23)) &
                                       ((N_(4) & m1) | (Z_(4) & (m5 | m6)))))) |
                     (N_(4) & (m2 | m3)));
}

/* South: 3D/CUDA_SubiterationKernels3D.cuh:88-123 */
static uint64_t comb_south(const uint64_t *n)
{
    uint64_t s1 = N_(3) | N_(4) | N_(5) | N_(12) | N_(13) | N_(20) | N_(21) | N_(22) | N_(0) | N_(2) | N_(17) | N_(19);
    uint64_t s2 = N_(9) | N_(11);
    uint64_t s3 = N_(1) | N_(18);
    uint64_t q1 = Z_(21) & Z_(22) & N_(1) & Z_(18) & Z_(19);
    uint64_t q2 = Z_(5) & Z_(4) & Z_(1) & Z_(2) & N_(18);
    uint64_t q3 = Z_(13) & Z_(11) & N_(9);
    uint64_t q4 = Z_(12) & Z_(9) & N_(11);
    uint64_t q6 = Z_(20) & Z_(17);
    uint64_t q5 = q4 & q6;
    q4 &= Z_(3) & Z_(0);
    uint64_t m1 = s1 | s2 | s3;
    uint64_t m5 = ((s1 | s2) & ((Z_(20) & Z_(17) & q1) | (Z_(3) & Z_(0) & q2))) |
                  ((s1 | s3) & ((q4 & q6) | (Z_(5) & Z_(22) & Z_(2) & Z_(19) & q3)));
    uint64_t m6 = (Z_(18) & Z_(21) & N_(1) & q5) | (Z_(1) & Z_(4) & N_(18) & q4) | (q3 & q2) | (q3 & q1);
    uint64_t m4 = (Z_(6) & Z_(23) & ((N_(8) & N_(5) & Z_(25)) | (N_(25) & N_(22) & Z_(8)))) |
                  (Z_(8) & Z_(25) & ((N_(6) & N_(3) & Z_(23)) | (N_(23) & N_(20) & Z_(6))));
    uint64_t m2 = (Z_(7) & Z_(8) & Z_(16) & ((Z_(6) & Z_(14) & N_(21)) | (Z_(24) & Z_(25) & N_(12)))) |
                  (Z_(24) & Z_(23) & Z_(14) & ((Z_(16) & Z_(25) & N_(4)) | (Z_(6) & Z_(7) & N_(13))));
    uint64_t m3 = (Z_(14) & N_(13) & ((Z_(24) & Z_(23) & N_(4)) | (Z_(6) & Z_(7) & N_(21)))) |
                  (Z_(16) & N_(12) & ((Z_(7) & Z_(8) & N_(21)) | (Z_(24) & Z_(25) & N_(4))));
    return Z_(15) & (((Z_(7) & Z_(14) & Z_(16) & Z_(24)) &
                      ((N_(10) & m4) | ((Z_(6) & Z_(8) & Z_(23) & Z_(25)) &
                                        ((N_(10) & m1) | (Z_(10) & (m5 | m6)))))) |
                     (N_(10) & (m2 | m3)));
}

/* North: 3D/CUDA_SubiterationKernels3D.cuh:125-160 */
static uint64_t comb_north(const uint64_t *n)
{
    uint64_t s1 = N_(20) | N_(21) | N_(22) | N_(12) | N_(13) | N_(3) | N_(4) | N_(5) | N_(23) | N_(25) | N_(6) | N_(8);
    uint64_t s2 = N_(14) | N_(16);
    uint64_t s3 = N_(24) | N_(7);
    uint64_t q1 = Z_(4) & Z_(5) & N_(24) & Z_(7) & Z_(8);
    uint64_t q2 = Z_(22) & Z_(21) & Z_(24) & Z_(25) & N_(7);
    uint64_t q3 = Z_(13) & Z_(16) & N_(14);
    uint64_t q4 = Z_(12) & Z_(14) & N_(16);
    uint64_t q6 = Z_(3) & Z_(6);
    uint64_t q5 = q4 & q6;
    q4 &= Z_(20) & Z_(23);
    uint64_t m1 = s1 | s2 | s3;
    uint64_t m5 = ((s1 | s2) & ((Z_(3) & Z_(6) & q1) | (Z_(20) & Z_(23) & q2))) |
                  ((s1 | s3) & ((q4 & q6) | (Z_(22) & Z_(5) & Z_(25) & Z_(8) & q3)));
    uint64_t m6 = (Z_(7) & Z_(4) & N_(24) & q5) | (Z_(24) & Z_(21) & N_(7) & q4) | (q3 & q2) | (q3 & q1);
    uint64_t m4 = (Z_(17) & Z_(0) & ((N_(19) & N_(22) & Z_(2)) | (N_(2) & N_(5) & Z_(19)))) |
                  (Z_(19) & Z_(2) & ((N_(17) & N_(20) & Z_(0)) | (N_(0) & N_(3) & Z_(17))));
    uint64_t m2 = (Z_(18) & Z_(19) & Z_(11) & ((Z_(17) & Z_(9) & N_(4)) | (Z_(1) & Z_(2) & N_(12)))) |
                  (Z_(1) & Z_(0) & Z_(9) & ((Z_(11) & Z_(2) & N_(21)) | (Z_(17) & Z_(18) & N_(13))));
    uint64_t m3 = (Z_(9) & N_(13) & ((Z_(1) & Z_(0) & N_(21)) | (Z_(17) & Z_(18) & N_(4)))) |
                  (Z_(11) & N_(12) & ((Z_(18) & Z_(19) & N_(4)) | (Z_(1) & Z_(2) & N_(21))));
    return Z_(10) & (((Z_(18) & Z_(9) & Z_(11) & Z_(1)) &
                      ((N_(15) & m4) | ((Z_(17) & Z_(19) & Z_(0) & Z_(2)) &
                                        ((N_(15) & m1) | (Z_(15) & (m5 | m6)))))) |
                     (N_(15) & (m2 | m3)));
}

/* East: 3D/CUDA_SubiterationKernels3D.cuh:162-197 */
static uint64_t comb_east(const uint64_t *n)
{
    uint64_t s1 = N_(18) | N_(10) | N_(1) | N_(21) | N_(4) | N_(24) | N_(15) | N_(7) | N_(19) | N_(2) | N_(25) | N_(8);
    uint64_t s2 = N_(22) | N_(5);
    uint64_t s3 = N_(11) | N_(16);
    uint64_t q1 = Z_(15) & Z_(7) & N_(11) & Z_(16) & Z_(8);
    uint64_t q2 = Z_(1) & Z_(10) & Z_(11) & Z_(2) & N_(16);
    uint64_t q3 = Z_(4) & Z_(5) & N_(22);
    uint64_t q4 = Z_(21) & Z_(22) & N_(5);
    uint64_t q6 = Z_(24) & Z_(25);
    uint64_t q5 = q4 & q6;
    q4 &= Z_(18) & Z_(19);
    uint64_t m1 = s1 | s2 | s3;
    uint64_t m5 = ((s1 | s2) & ((Z_(24) & Z_(25) & q1) | (Z_(18) & Z_(19) & q2))) |
                  ((s1 | s3) & ((q4 & q6) | (Z_(1) & Z_(7) & Z_(2) & Z_(8) & q3)));
    uint64_t m6 = (Z_(16) & Z_(15) & N_(11) & q5) | (Z_(11) & Z_(10) & N_(16) & q4) | (q3 & q2) | (q3 & q1);
    uint64_t m4 = (Z_(17) & Z_(23) & ((N_(0) & N_(1) & Z_(6)) | (N_(6) & N_(7) & Z_(0)))) |
                  (Z_(0) & Z_(6) & ((N_(17) & N_(18) & Z_(23)) | (N_(23) & N_(24) & Z_(17))));
    uint64_t m2 = (Z_(9) & Z_(0) & Z_(3) & ((Z_(17) & Z_(20) & N_(15)) | (Z_(14) & Z_(6) & N_(21)))) |
                  (Z_(14) & Z_(23) & Z_(20) & ((Z_(3) & Z_(6) & N_(10)) | (Z_(17) & Z_(9) & N_(4))));
    uint64_t m3 = (Z_(20) & N_(4) & ((Z_(14) & Z_(23) & N_(10)) | (Z_(17) & Z_(9) & N_(15)))) |
                  (Z_(3) & N_(21) & ((Z_(9) & Z_(0) & N_(15)) | (Z_(14) & Z_(6) & N_(10))));
    return Z_(12) & (((Z_(9) & Z_(20) & Z_(3) & Z_(14)) &
                      ((N_(13) & m4) | ((Z_(17) & Z_(0) & Z_(23) & Z_(6)) &
                                        ((N_(13) & m1) | (Z_(13) & (m5 | m6)))))) |
                     (N_(13) & (m2 | m3)));
}

/* West: 3D/CUDA_SubiterationKernels3D.cuh:199-234 */
static uint64_t comb_west(const uint64_t *n)
{
    uint64_t s1 = N_(1) | N_(10) | N_(18) | N_(4) | N_(21) | N_(7) | N_(15) | N_(24) | N_(0) | N_(17) | N_(6) | N_(23);
    uint64_t s2 = N_(3) | N_(20);
    uint64_t s3 = N_(9) | N_(14);
    uint64_t q1 = Z_(15) & Z_(24) & N_(9) & Z_(14) & Z_(23);
    uint64_t q2 = Z_(18) & Z_(10) & Z_(9) & Z_(17) & N_(14);
    uint64_t q3 = Z_(21) & Z_(20) & N_(3);
    uint64_t q4 = Z_(4) & Z_(3) & N_(20);
    uint64_t q6 = Z_(7) & Z_(6);
    uint64_t q5 = q4 & q6;
    q4 &= Z_(1) & Z_(0);
    uint64_t m1 = s1 | s2 | s3;
    uint64_t m5 = ((s1 | s2) & ((Z_(7) & Z_(6) & q1) | (Z_(1) & Z_(0) & q2))) |
                  ((s1 | s3) & ((q4 & q6) | (Z_(18) & Z_(24) & Z_(17) & Z_(23) & q3)));
    uint64_t m6 = (Z_(14) & Z_(15) & N_(9) & q5) | (Z_(9) & Z_(10) & N_(14) & q4) | (q3 & q2) | (q3 & q1);
    uint64_t m4 = (Z_(2) & Z_(8) & ((N_(19) & N_(18) & Z_(25)) | (N_(25) & N_(24) & Z_(19)))) |
                  (Z_(19) & Z_(25) & ((N_(2) & N_(1) & Z_(8)) | (N_(8) & N_(7) & Z_(2))));
    uint64_t m2 = (Z_(11) & Z_(19) & Z_(22) & ((Z_(2) & Z_(5) & N_(15)) | (Z_(16) & Z_(25) & N_(4)))) |
                  (Z_(16) & Z_(8) & Z_(5) & ((Z_(22) & Z_(25) & N_(10)) | (Z_(2) & Z_(11) & N_(21))));
    uint64_t m3 = (Z_(5) & N_(21) & ((Z_(16) & Z_(8) & N_(10)) | (Z_(2) & Z_(11) & N_(15)))) |
                  (Z_(22) & N_(4) & ((Z_(11) & Z_(19) & N_(15)) | (Z_(16) & Z_(25) & N_(10))));
    return Z_(13) & (((Z_(11) & Z_(5) & Z_(22) & Z_(16)) &
                      ((N_(12) & m4) | ((Z_(2) & Z_(19) & Z_(8) & Z_(25)) &
                                        ((N_(12) & m1) | (Z_(12) & (m5 | m6)))))) |
                     (N_(12) & (m2 | m3)));
}

#undef N_
#undef Z_

uint64_t or3d_comb(int dir, const uint64_t n[26])
{
    switch (dir) {
    case 0: return comb_up(n);
    case 1: return comb_north(n);
    case 2: return comb_east(n);
    case 3: return comb_down(n);
    case 4: return comb_south(n);
    default: return comb_west(n);
    }
}

void or3d_table(int dir, uint64_t *out)
{
    /* 64 configurations per word: bits 0..5 of the configuration index
     * vary inside the word, bits 6..25 are constant per word. */
    static const uint64_t lowpat[6] = {
        0xAAAAAAAAAAAAAAAAull, 0xCCCCCCCCCCCCCCCCull, 0xF0F0F0F0F0F0F0F0ull,
        0xFF00FF00FF00FF00ull, 0xFFFF0000FFFF0000ull, 0xFFFFFFFF00000000ull};
    uint64_t n[26];
    for (int i = 0; i < 6; i++) n[i] = lowpat[i];
    for (uint32_t hi = 0; hi < (1u << 20); hi++) {
        for (int i = 6; i < 26; i++) n[i] = ((hi >> (i - 6)) & 1) ? ~0ull : 0ull;
        out[hi] = or3d_comb(dir, n);
    }
}

/* ------------------------------------------------------------------------- */
/* 3D iteration: 3D/CUDA_ThinningKernel3D_Linear.cuh:690-702 (pre-test),      */
/* :739-782 (U, N, E, D, S, W, each deleting only unprotected voxels).        */
/* State packed along x in 64-bit words.                                      */
/* ------------------------------------------------------------------------- */

typedef struct {
    int w, h, d, nw;
    uint64_t *a, *b, *prot;
} grid3d;

static inline const uint64_t *row3(const grid3d *g, const uint64_t *base, int y, int z,
                                   const uint64_t *zrow)
{
    if (y < 0 || y >= g->h || z < 0 || z >= g->d) return zrow;
    return base + ((size_t)z * g->h + y) * g->nw;
}

static void gather3d(const grid3d *g, const uint64_t *base, int y, int z, int w,
                     const uint64_t *zrow, uint64_t n[26])
{
    int i = 0;
    for (int j = 0; j < 27; j++) {
        if (j == 13) continue;
        int dz = j / 9 - 1, dy = (j / 3) % 3 - 1, dx = j % 3 - 1;
        const uint64_t *r = row3(g, base, y + dy, z + dz, zrow);
        uint64_t v;
        if (dx < 0) v = shl_w(r, w, g->nw);
        else if (dx > 0) v = shr_w(r, w, g->nw);
        else v = r[w];
        n[i++] = v;
    }
}

static int word_has_work(const grid3d *g, const uint64_t *base, int y, int z, int w)
{
    (void)g;
    return base[((size_t)z * g->h + y) * g->nw + w] != 0;
}

static void protect3d(grid3d *g, const uint64_t *zrow)
{
    for (int z = 0; z < g->d; z++)
        for (int y = 0; y < g->h; y++)
            for (int w = 0; w < g->nw; w++) {
                size_t o = ((size_t)z * g->h + y) * g->nw + w;
                uint64_t v = g->a[o];
                if (!v) { g->prot[o] = 0; continue; }
                uint64_t n[26];
                gather3d(g, g->a, y, z, w, zrow, n);
                uint64_t any = 0;
                for (int dir = 0; dir < 6; dir++) any |= or3d_comb(dir, n);
                g->prot[o] = v & ~any;
            }
}

static int subiter3d(grid3d *g, int dir, const uint64_t *zrow)
{
    int changed = 0;
    for (int z = 0; z < g->d; z++)
        for (int y = 0; y < g->h; y++)
            for (int w = 0; w < g->nw; w++) {
                size_t o = ((size_t)z * g->h + y) * g->nw + w;
                uint64_t v = g->a[o];
                if (!word_has_work(g, g->a, y, z, w)) { g->b[o] = 0; continue; }
                uint64_t n[26];
                gather3d(g, g->a, y, z, w, zrow, n);
                uint64_t del = or3d_comb(dir, n) & v & ~g->prot[o];
                g->b[o] = v & ~del;
                if (del) changed = 1;
            }
    uint64_t *t = g->a; g->a = g->b; g->b = t;
    return changed;
}

static int run3d(uint8_t *vol, int w, int h, int d, int max_iter, int count_only_changed)
{
    if (w <= 0 || h <= 0 || d <= 0) return 0;
    grid3d g;
    g.w = w; g.h = h; g.d = d; g.nw = (w + 63) >> 6;
    size_t nwords = (size_t)g.nw * h * d;
    g.a = (uint64_t *)calloc(nwords, sizeof(uint64_t));
    g.b = (uint64_t *)calloc(nwords, sizeof(uint64_t));
    g.prot = (uint64_t *)calloc(nwords, sizeof(uint64_t));
    uint64_t *zrow = (uint64_t *)calloc((size_t)g.nw, sizeof(uint64_t));
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++)
                if (vol[((size_t)z * h + y) * w + x])
                    g.a[((size_t)z * h + y) * g.nw + (x >> 6)] |= 1ull << (x & 63);
    int it = 0, nchanged = 0;
    for (;;) {
        if (max_iter > 0 && it >= max_iter) break;
        it++;
        protect3d(&g, zrow);
        int ch = 0;
        for (int dir = 0; dir < 6; dir++) ch |= subiter3d(&g, dir, zrow);
        if (!ch) break;
        nchanged++;
    }
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++)
                vol[((size_t)z * h + y) * w + x] =
                    (g.a[((size_t)z * h + y) * g.nw + (x >> 6)] >> (x & 63)) & 1;
    free(g.a); free(g.b); free(g.prot); free(zrow);
    return count_only_changed ? nchanged : it;
}

int or3d_thin(uint8_t *vol, int w, int h, int d, int max_iter) { return run3d(vol, w, h, d, max_iter, 0); }
int or3d_iterate(uint8_t *vol, int w, int h, int d, int n_iter) { return run3d(vol, w, h, d, n_iter, 1); }

/* ------------------------------------------------------------------------- */
/* ExtractSegmentsV: 2D/ExtractSegments.cpp:16-127, 3D/ExtractSegments.cpp:   */
/* 135-251.  Same algorithm in 2D (8 neighbours) and 3D (26 neighbours).      */
/* ------------------------------------------------------------------------- */

typedef struct {
    int nseg, npts, cap_pts, cap_seg;
    int16_t *x, *y, *z;
    uint32_t *start;
    uint8_t *is_end;
} segbuf;

static void sb_init(segbuf *b, int three_d)
{
    memset(b, 0, sizeof(*b));
    b->cap_pts = 1024; b->cap_seg = 256;
    b->x = (int16_t *)malloc(sizeof(int16_t) * b->cap_pts);
    b->y = (int16_t *)malloc(sizeof(int16_t) * b->cap_pts);
    b->z = three_d ? (int16_t *)malloc(sizeof(int16_t) * b->cap_pts) : NULL;
    b->start = (uint32_t *)malloc(sizeof(uint32_t) * (b->cap_seg + 1));
    b->is_end = (uint8_t *)malloc(b->cap_seg);
    b->start[0] = 0;
}

static void sb_point(segbuf *b, int x, int y, int z)
{
    if (b->npts == b->cap_pts) {
        b->cap_pts *= 2;
        b->x = (int16_t *)realloc(b->x, sizeof(int16_t) * b->cap_pts);
        b->y = (int16_t *)realloc(b->y, sizeof(int16_t) * b->cap_pts);
        if (b->z) b->z = (int16_t *)realloc(b->z, sizeof(int16_t) * b->cap_pts);
    }
    b->x[b->npts] = (int16_t)x;
    b->y[b->npts] = (int16_t)y;
    if (b->z) b->z[b->npts] = (int16_t)z;
    b->npts++;
}

static void sb_close(segbuf *b, int is_end)
{
    if (b->nseg == b->cap_seg) {
        b->cap_seg *= 2;
        b->start = (uint32_t *)realloc(b->start, sizeof(uint32_t) * (b->cap_seg + 1));
        b->is_end = (uint8_t *)realloc(b->is_end, b->cap_seg);
    }
    b->is_end[b->nseg] = (uint8_t)(is_end != 0);
    b->nseg++;
    b->start[b->nseg] = (uint32_t)b->npts;
}

static void sb_export(segbuf *b, or_segments *out)
{
    out->nseg = b->nseg; out->npts = b->npts;
    out->x = b->x; out->y = b->y; out->z = b->z;
    out->seg_start = b->start; out->is_end = b->is_end;
}

/* Shared tracer.  dims: 2 or 3; delta: neighbour offsets in the reference's
 * raster order; coords of linear index via w, h. */
static int trace_generic(const uint8_t *data, int w, int h, int d, int dims,
                         or_segments *out)
{
    const int nnb = dims == 2 ? 8 : 26;
    const long sl = (long)w * h;
    long delta[26];
    if (dims == 2) {
        const long t[8] = {-w - 1, -w, -w + 1, -1, 1, w - 1, w, w + 1};
        memcpy(delta, t, sizeof(t));
    } else {
        int i = 0;
        for (int j = 0; j < 27; j++) {
            if (j == 13) continue;
            int dz = j / 9 - 1, dy = (j / 3) % 3 - 1, dx = j % 3 - 1;
            delta[i++] = dz * sl + dy * (long)w + dx;
        }
    }
    size_t nel = (size_t)w * h * (dims == 2 ? 1 : d);
    int8_t *nb = (int8_t *)calloc(nel ? nel : 1, 1);
    int8_t *onb = (int8_t *)malloc(nel ? nel : 1);
    long *pts = NULL; size_t npts = 0, cap = 0;

    /* pass 1: 2D/ExtractSegments.cpp:32-50 */
    int zmax = dims == 2 ? 1 : d;
    for (int z = 0; z < zmax; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                long p = (long)z * sl + (long)y * w + x;
                if (!data[p]) continue;
                if (x == 0 || y == 0 || x == w - 1 || y == h - 1) continue;
                if (dims == 3 && (z == 0 || z == d - 1)) continue;
                int nn = 0;
                for (int k = 0; k < nnb; k++) nn += data[p + delta[k]] != 0;
                if (nn == 0) nn = 1;
                if (nn != 2) nn = -nn;
                nb[p] = (int8_t)nn;
                if (npts == cap) {
                    cap = cap ? cap * 2 : 4096;
                    pts = (long *)realloc(pts, sizeof(long) * cap);
                }
                pts[npts++] = p;
            }
    memcpy(onb, nb, nel);

    segbuf sb;
    sb_init(&sb, dims == 3);
#define EMIT_PT(idx)                                                         \
    do {                                                                     \
        long q_ = (idx);                                                     \
        int zz_ = dims == 3 ? (int)(q_ / sl) : 0;                            \
        long r_ = q_ - (long)zz_ * sl;                                       \
        sb_point(&sb, (int)(r_ % w), (int)(r_ / w), zz_);                    \
    } while (0)

    /* pass 2: 2D/ExtractSegments.cpp:56-121 */
    for (size_t t = 0; t < npts; t++) {
        long idx = pts[t];
        int8_t n = nb[idx];
        if (n >= 0) continue;
        nb[idx] = 0;
        int ni = -1;
        for (int8_t si = 0; si > n; si--) {
            int is_end = abs((int)onb[idx]) < 2;
            EMIT_PT(idx);
            ni++;
            for (; ni < nnb; ni++)
                if (nb[idx + delta[ni]] != 0) break;
            if (ni >= nnb) { sb_close(&sb, is_end); continue; }
            long tidx = idx + delta[ni];
            while (nb[tidx] == 2) {
                nb[tidx] = 0;
                EMIT_PT(tidx);
                for (int k = 0; k < nnb; k++)
                    if (nb[tidx + delta[k]] != 0) { tidx += delta[k]; break; }
            }
            EMIT_PT(tidx);
            if (abs((int)onb[tidx]) < 2) is_end = 1;
            nb[tidx] = (int8_t)(nb[tidx] + 1);
            sb_close(&sb, is_end);
        }
    }
#undef EMIT_PT
    free(nb); free(onb); free(pts);
    sb_export(&sb, out);
    return out->nseg;
}

int or2d_extract(const uint8_t *img, int w, int h, or_segments *out)
{
    return trace_generic(img, w, h, 1, 2, out);
}

int or3d_extract(const uint8_t *vol, int w, int h, int d, or_segments *out)
{
    return trace_generic(vol, w, h, d, 3, out);
}

void or_segments_free(or_segments *s)
{
    free(s->x); free(s->y); free(s->z); free(s->seg_start); free(s->is_end);
    memset(s, 0, sizeof(*s));
}

/* ------------------------------------------------------------------------- */
/* Smoothing: 2D/CenterlineExtraction.cpp:60-104 (float32 accumulators in the */
/* reference's exact order; a negative span is treated as 0 -- the reference  */
/* reads out of bounds there).                                                */
/* ------------------------------------------------------------------------- */

void or_smooth(const int16_t *px, const int16_t *py, const int16_t *pz, int n,
               int smoothing_span, float *ox, float *oy, float *oz)
{
    if (n <= 0) return;
    int sspan = smoothing_span < 0 ? 0 : smoothing_span;
    int half = (n - 1) / 2;
    int span = sspan < half ? sspan : half;
    float x = px[0], y = py[0], z = pz ? pz[0] : 0.0f;
    int o = 0, cnt = 1;
    ox[o] = (float)px[0]; oy[o] = (float)py[0]; if (pz) oz[o] = (float)pz[0];
    o++;
    for (int k = 1; k <= span; k++) {
        x += px[2 * k - 1]; y += py[2 * k - 1]; if (pz) z += pz[2 * k - 1];
        x += px[2 * k];     y += py[2 * k];     if (pz) z += pz[2 * k];
        cnt += 2;
        ox[o] = x / (float)cnt; oy[o] = y / (float)cnt; if (pz) oz[o] = z / (float)cnt;
        o++;
    }
    for (int k = span + 1; k < n - span; k++) {
        x += px[k + span]; y += py[k + span]; if (pz) z += pz[k + span];
        x -= px[k - span - 1]; y -= py[k - span - 1]; if (pz) z -= pz[k - span - 1];
        ox[o] = x / (float)cnt; oy[o] = y / (float)cnt; if (pz) oz[o] = z / (float)cnt;
        o++;
    }
    for (int k = n - span, dd = span; k < n; k++, dd--) {
        x -= px[k - dd - 1]; y -= py[k - dd - 1]; if (pz) z -= pz[k - dd - 1];
        x -= px[k - dd];     y -= py[k - dd];     if (pz) z -= pz[k - dd];
        cnt -= 2;
        ox[o] = x / (float)cnt; oy[o] = y / (float)cnt; if (pz) oz[o] = z / (float)cnt;
        o++;
    }
}

/* ------------------------------------------------------------------------- */
/* CenterlineExtraction: 2D/CenterlineExtraction.cpp:22-142,                  */
/* 3D/CenterlineExtraction.cpp:22-155                                         */
/* ------------------------------------------------------------------------- */

static int finish_centerline(or_segments *segs, int max_segment_length, int smoothing_span,
                             int three_d, or_centerline *out)
{
    /* prune: size <= (size_t)max_segment_length && is_end */
    size_t lim = (size_t)(long)max_segment_length;
    uint8_t *keep = (uint8_t *)malloc(segs->nseg ? segs->nseg : 1);
    int nkeep = 0, nkp = 0;
    for (int s = 0; s < segs->nseg; s++) {
        size_t sz = segs->seg_start[s + 1] - segs->seg_start[s];
        int pruned = (sz <= lim) && segs->is_end[s];
        keep[s] = !pruned;
        if (keep[s]) { nkeep++; nkp += (int)sz; }
    }
    out->nseg = nkeep; out->npts = nkp;
    out->x = (float *)malloc(sizeof(float) * (nkp ? nkp : 1));
    out->y = (float *)malloc(sizeof(float) * (nkp ? nkp : 1));
    out->z = three_d ? (float *)malloc(sizeof(float) * (nkp ? nkp : 1)) : NULL;
    out->idx = (uint32_t *)malloc(sizeof(uint32_t) * (nkeep + 1));
    int si = 0, pi = 0;
    for (int s = 0; s < segs->nseg; s++) {
        if (!keep[s]) continue;
        int st = (int)segs->seg_start[s];
        int sz = (int)(segs->seg_start[s + 1] - segs->seg_start[s]);
        out->idx[si++] = (uint32_t)pi;
        or_smooth(segs->x + st, segs->y + st, three_d ? segs->z + st : NULL, sz,
                  smoothing_span, out->x + pi, out->y + pi, three_d ? out->z + pi : NULL);
        pi += sz;
    }
    out->idx[nkeep] = (uint32_t)nkp;
    free(keep);
    return nkeep;
}

int or2d_centerline(const uint8_t *img, int w, int h, int max_segment_length,
                    int smoothing_span, or_centerline *out)
{
    int pw = ((w + 31) / 32) * 32;
    int q = (h + 63) / 128;
    int ph = (q > 1 ? q : 1) * 128 + 64;
    uint8_t *buf = (uint8_t *)calloc((size_t)pw * ph, 1);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) buf[(size_t)y * pw + x] = img[(size_t)y * w + x] != 0;
    out->iterations = or2d_thin(buf, pw, ph, 0);
    or_segments segs;
    or2d_extract(buf, pw, ph, &segs);
    free(buf);
    int r = finish_centerline(&segs, max_segment_length, smoothing_span, 0, out);
    or_segments_free(&segs);
    return r;
}

int or3d_centerline(const uint8_t *vol, int w, int h, int d, int max_segment_length,
                    int smoothing_span, or_centerline *out)
{
    int pw = ((w + 31) / 32) * 32, ph = ((h + 31) / 32) * 32, pd = ((d + 31) / 32) * 32;
    uint8_t *buf = (uint8_t *)calloc((size_t)pw * ph * pd, 1);
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++)
                buf[((size_t)z * ph + y) * pw + x] = vol[((size_t)z * h + y) * w + x] != 0;
    out->iterations = or3d_thin(buf, pw, ph, pd, 0);
    or_segments segs;
    or3d_extract(buf, pw, ph, pd, &segs);
    free(buf);
    int r = finish_centerline(&segs, max_segment_length, smoothing_span, 1, out);
    or_segments_free(&segs);
    return r;
}

void or_centerline_free(or_centerline *c)
{
    free(c->x); free(c->y); free(c->z); free(c->idx);
    memset(c, 0, sizeof(*c));
}

void or_free(void *p) { free(p); }
