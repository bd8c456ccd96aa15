// tile3d.h -- the fused-iteration tile engine of the 3D path (internal).
//
// State layout: the zero-margin word layout of sparse3d.h (bit k of word j of
// x-line (y, z) is voxel x = 32j + k; every line has a zero word before it,
// every plane a zero line above and below, every volume a zero plane before
// and after), in two state arrays shared by all volumes of a run.
//
// Active blocks.  Inside the tiles, activity is tracked per block of 1 word
// by 8 lines by 8 planes: a word whose block is not within one block of a
// block that changed in the previous iteration has the same neighbourhood
// within distance 6 as at that iteration's start, so it repeats that
// iteration's outcome -- it keeps its value through every step -- and is
// left as staged.  A queued tile evaluates only the non-empty words of active
// blocks (all words in iteration 0); the tiles queued are those holding a
// block activated by a change.
//
// Tiles.  A volume is cut into tiles of 2 words (64 voxels) along x by 16
// lines by 16 planes.  One iteration of the reference (the protection
// pre-test, then U, N, E, D, S, W, 3D/CUDA_ThinningKernel3D_Linear.cuh:
// 739-782) is one launch: a workgroup stages a tile with a halo of 6 voxels
// in y and z and one word in x (the region, 4 x 28 x 28 words) in LDS and
// runs all seven steps there.  A voxel's value after an iteration depends on
// the voxels within Chebyshev distance 6 at its start, so the tile's interior
// is exact; step j (1..6) only evaluates the region's non-empty words within
// distance 6 - j of the interior, which are kept sorted by that distance.
//
// Iteration t reads state array t & 1 and writes array (t + 1) & 1.  A tile is
// queued for iteration t + 1 when it holds a block activated in iteration t;
// the other tiles cannot change.  A queued tile writes every interior
// word that is non-zero in either array, so that after iteration t the
// written array holds the whole state: words of tiles that were not queued
// did not change in iterations t - 1 and t, and so hold the same value in
// both arrays.  A tile whose interior is empty in both arrays is never queued
// again (empty words cannot change).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "sparse3d.h"

namespace bgt {

constexpr int kTileWordsX = 2;   // words (32 voxels each) of a tile along x
constexpr int kTileLines = 16;   // lines (y) of a tile
constexpr int kTilePlanes = 16;  // planes (z) of a tile

struct TileItem3D {
    long base;          // first word of the item's padded region
    int w, h, d, nh;    // voxels; data words per x-line (even)
    int P, plane;       // padded line and plane strides in words
    int ntx, nty, ntz;  // tiles along x, y, z
    int tbase;          // first tile of the item
    int nby, nbz;       // activity blocks (1 word x 8 lines x 8 planes) along y, z
    int bbase;          // first block of the item
};

// Control words: [0..2] tile-list counts (by iteration % 3), [3] live items.
enum { TC_COUNT0 = 0, TC_LIVE = 3, TC_INTS = 8 };

// Device pointers of one run (a single volume or all volumes of a batch).
struct Tile3D {
    int nitems = 0;
    const TileItem3D *items = nullptr;
    // pack/unpack grids and caller buffers (the layout of sparse3d_layout)
    const SItem3D *sitems = nullptr;
    const SIO *io = nullptr;
    const int *blk = nullptr;
    int nblk = 0;
    const int *ublk = nullptr;
    int nublk = 0;
    uint32_t *state[2] = {nullptr, nullptr};
    long nwords = 0;
    int ntiles = 0;
    int nblocks = 0;
    // [2][nblocks] iteration + 1 the block is active in, by iteration parity:
    // iteration t reads array t & 1 and activates blocks in array (t + 1) & 1
    int *bstamp = nullptr;
    uint8_t *occ = nullptr;   // [ntiles] interior non-empty in some array
    int *stamp = nullptr;     // [ntiles] iteration + 1 the tile was last queued for
    int *list[3] = {nullptr, nullptr, nullptr};  // [ntiles] each, by iteration % 3
    int *changed = nullptr;   // [3][nitems], by iteration % 3
    int *states = nullptr;    // [nitems][ST_INTS]
    int *ctl = nullptr;       // [TC_INTS]
};

// Per-item tile counts of a layout; returns the total.
long tile3d_items(const Layout3D &lay, std::vector<TileItem3D> &out);

// Carves the engine's arrays from `base` (null: only sizes them) and returns
// the bytes taken; the two state arrays too unless `shared_state` gives them
// (those of a word-list engine the tile engine takes over from).
size_t tile3d_carve(Tile3D &e, char *base, int nitems, long nwords, long ntiles, long nblocks,
                    uint32_t *const *shared_state = nullptr);

// Workgroups of an iteration launch (blocks_per_cu per CU; 0 = as many as fit).
int tile3d_grid(int blocks_per_cu);

// Resets the control words, states and tile flags, zeroes the state arrays
// when `zero` (once per allocation), packs every item into both arrays and
// queues every non-empty tile for iteration 0.
int tile3d_begin(const Tile3D &e, bool zero, hipStream_t s);

// Iteration t over the queued tiles; also closes the items whose iteration
// t - 1 changed nothing (ST_DONE, ST_ITERS = t) and clears the control words
// iteration t + 1 uses.
int tile3d_iteration(const Tile3D &e, int t, int grid, hipStream_t s);

// Writes every item's result to its caller buffer (both arrays hold it once
// the item has converged).
int tile3d_end(const Tile3D &e, hipStream_t s);

// Takes over a run of the word-list engine `w` (sparse3d.h) after its
// iteration k - 1 (k even: its list 0 is current and array 0 holds the
// state), so that tile3d_iteration(e, k, ...) continues it.  `e` shares w's
// state arrays.  w's array 1 differs from array 0 only in words its last
// step changed, which lie in blocks w activated for iteration k: those
// become e's active blocks, their tiles its queue for iteration k; tile
// occupancy comes from w's list of non-empty words (`bound` >= its count),
// item states and the live count from w's.
int tile3d_from_words(const Tile3D &e, const Sparse3D &w, int k, int bound, hipStream_t s);

}  // namespace bgt
