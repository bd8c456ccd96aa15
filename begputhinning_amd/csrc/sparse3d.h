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
//
// Active words.  A voxel's value after an iteration (six subiterations, each
// reading the 3x3x3 neighbourhood, plus the protection mask taken at its
// start) depends on the voxels within Chebyshev distance 6 at the start of the
// iteration.  So a word none of whose voxels within distance 6 changed in the
// previous iteration, and which did not change itself, repeats that
// iteration's outcome: it stays as it is.  Words are grouped in blocks of one
// word (32 voxels) along x by 8 lines by 8 planes; a word that changes marks
// its block dirty, and the next iteration runs only over the non-empty words
// of the 27 blocks around every dirty block (distance 6 stays inside them).
// The first iteration runs over every non-empty word.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace bgt {

// 64-voxel runs each thread of a pack grid handles (the pack block tables of
// sparse3d_layout give every item whole 256-thread blocks of such threads)
constexpr int kPackRuns3D = 8;

struct SItem3D {
    long base;     // first word of the item's padded region
    int w, h, d;   // voxels
    int nh;        // data words per x-line (even)
    int P, plane;  // padded line and plane strides in words
    int nby, nbz;  // blocks along y and z (nh along x)
    int bbase;     // first block of the item
    int dbase;     // sum of the data words of the items before it
};

// Caller buffers of one item: uint8 voxels, x fastest, then y, then z.
struct SIO {
    const uint8_t *in;
    uint8_t *out;
    int blk0;   // first block of the item in the pack grid
    int ublk0;  // first block of the item in the unpack grid
};

// List of (item, word, block) entries; *count entries are valid.
struct SList {
    int *item;
    uint32_t *word;  // global word index into the state arrays
    int *blk;        // block of the word (global index)
    int *count;
};

// Control words of a run: [0] iteration, [1] live items, [2] finished
// workgroups of the current compaction, [3] count of list 0, [4] of list 1,
// [5 + (t & 1)] count of the active list of iteration t.
enum { SC_ITER = 0, SC_LIVE = 1, SC_BLOCKS = 2, SC_COUNT0 = 3, SC_ACOUNT0 = 5, SC_INTS = 8 };

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
    SList list[2];             // non-empty words of live items, by iteration parity
    SList act;                 // active words of the running iteration (count in ctl)
    uint32_t *prot = nullptr;  // [capacity], by active-list entry
    uint32_t *first = nullptr; // [capacity], the word at the start of the iteration
    int capacity = 0;          // data words of all items (bounds every list)
    int *bstamp = nullptr;     // [nblocks] iteration + 1 of the last activation
    int *dstamp = nullptr;     // [nblocks] iteration + 1 it was last marked dirty
    int nblocks = 0;
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
    long nwords = 0, data_words = 0, nblocks = 0;
};
int sparse3d_layout(const int *w, const int *h, const int *d, int n, Layout3D &lay);

// Workgroups of the grid-stride step kernel: blocks_per_cu per CU, or the
// occupancy limit when blocks_per_cu is 0.
int sparse3d_grid(int blocks_per_cu);

// Resets the control words, item states and block stamps, zeroes the state
// arrays when `zero` (needed once per allocation: margins are never written
// later), packs every item into both state arrays and builds list 0 and the
// first active list (all of list 0).
int sparse3d_begin(const Sparse3D &e, bool zero, hipStream_t s);

// The remaining iterations of every open item in one launch: one workgroup of
// 1024 threads per item runs the six subiterations over the item's active
// words with workgroup barriers between them, marks the blocks its changes
// touch, and gathers the next active words from the non-empty words of the
// blocks around them, until an iteration changes nothing.  For the tail of a
// run, when few words are active: it replaces the per-iteration launches (and
// their dispatch gaps) by one.  Uses the list, prot and first arrays as
// per-item scratch (item i's part starts at its dbase).
int sparse3d_tail(const Sparse3D &e, hipStream_t s);

// One iteration: six subiteration launches over the active list, then the
// compaction of list `cur` into list cur^1 (which also builds the next active
// list and records per-item convergence).  `bound` >= list cur's count and
// `abound` >= the active list's count (both only shrink); `grid` caps the
// step kernel's workgroups.
int sparse3d_iteration(const Sparse3D &e, int cur, int bound, int abound, int grid, hipStream_t s);

// Writes every item's result (state array 0) to its caller buffer.
int sparse3d_end(const Sparse3D &e, hipStream_t s);

}  // namespace bgt
