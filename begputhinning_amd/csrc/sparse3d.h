// sparse3d.h -- the non-empty-word list engine of the 3D path (internal).
//
// State layout: every volume is stored bit-packed with a margin of zero words
// around it, in two state arrays shared by all volumes of a run.  An x-line
// holds nh = 2*ceil(w/64) data words (bit k of word j is voxel x = 32j + k)
// after one zero word: P = nh + 1 words per padded line (the next line's zero
// word is the right-hand neighbour of a line's last word).  A padded plane is
// h + 2 lines (a zero line above and below), the volume d + 2 planes and one
// trailing zero word (the right-hand neighbour of the last line's last word).
// Interior word (j, y, z) of an item is at
//     base + ((z + 1) * (h + 2) + (y + 1)) * P + (j + 1),
// so its 26 neighbour words are at constant offsets +-1, +-P, +-plane and
// their sums: no bounds checks, and the margins stay zero for good because
// only listed (interior, non-empty) words are ever written.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace bgt {

struct SItem3D {
    long base;     // first word of the item's padded region
    int w, h, d;   // voxels
    int nh;        // data words per x-line (even)
    int P, plane;  // padded line and plane strides in words
};

// Caller buffers of one item: uint8 voxels, x fastest, then y, then z.
struct SIO {
    const uint8_t *in;
    uint8_t *out;
    int blk0;   // first block of the item in the pack grid
    int ublk0;  // first block of the item in the unpack grid
};

// List of (item, word) entries; *count entries are valid.
struct SList {
    int *item;
    uint32_t *word;  // global word index into the state arrays
    int *count;
};

// Control words of a run: [0] iteration, [1] live items, [2] finished
// workgroups of the current compaction, [3] count of list 0, [4] of list 1.
enum { SC_ITER = 0, SC_LIVE = 1, SC_BLOCKS = 2, SC_COUNT0 = 3, SC_INTS = 8 };

// Device pointers of one run of the engine (a single volume or a batch).
struct Sparse3D {
    int nitems = 0;
    const SItem3D *items = nullptr;  // [nitems]
    const int2 *geo = nullptr;       // [nitems] (P, plane)
    const SIO *io = nullptr;         // [nitems]
    const int *blk = nullptr;        // pack block -> item (null: all item 0)
    int nblk = 0;
    const int *ublk = nullptr;       // unpack block -> item (null: all item 0)
    int nublk = 0;
    uint32_t *state[2] = {nullptr, nullptr};
    long nwords = 0;  // padded words of all items
    SList list[2];
    uint32_t *prot = nullptr;  // [capacity]
    int capacity = 0;          // data words of all items (bounds every list)
    int *states = nullptr;     // [nitems * ST_INTS]
    int *changed = nullptr;    // [nitems]
    int *ctl = nullptr;        // [SC_INTS]
};

// Host-side layout of a set of volumes: items (base, strides), the pack and
// unpack block tables (block -> item, first block of each item) and the
// totals.  Returns BGT_OK or an error when the state would exceed 2^31 words.
struct Layout3D {
    std::vector<SItem3D> items;
    std::vector<int2> geo;
    std::vector<int> blk, blk0, ublk, ublk0;
    long nwords = 0, data_words = 0;
};
int sparse3d_layout(const int *w, const int *h, const int *d, int n, Layout3D &lay);

// Workgroups of the grid-stride step kernel: blocks_per_cu per CU, or the
// occupancy limit when blocks_per_cu is 0.
int sparse3d_grid(int blocks_per_cu);

// Resets the control words and item states, zeroes the state arrays when
// `zero` (needed once per allocation: margins are never written later),
// packs every item into both state arrays and builds list 0.
int sparse3d_begin(const Sparse3D &e, bool zero, hipStream_t s);

// One iteration over list `cur` into list cur^1: six subiteration launches
// and the compaction (which also records per-item convergence).
// `bound` >= the current list's count (lists only shrink); `grid` caps the
// step kernel's workgroups.
int sparse3d_iteration(const Sparse3D &e, int cur, int bound, int grid, hipStream_t s);

// Writes every item's result (state array 0) to its caller buffer.
int sparse3d_end(const Sparse3D &e, hipStream_t s);

}  // namespace bgt
