// sparse3d.h -- the non-empty-word list engine of the 3D path (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bgt {

// One volume: packed state (x-lines of nh 32-bit words, line (y, z) at
// (z*h + y)*nh) in two buffers; the result of every iteration is in buf[0].
struct SItem3D {
    uint32_t *buf[2];
    int *st;  // item state (ST_*)
    int w, h, d, nh;
};

// List of (item, word) entries; *count entries are valid.
struct SList {
    int *item;
    uint32_t *word;
    int *count;
};

// Workgroups of the grid-stride step kernel: blocks_per_cu per CU, or the
// occupancy limit when blocks_per_cu is 0.
int sparse3d_grid(int blocks_per_cu);
int sparse3d_build_list(const SItem3D *items, const long *off, int nitems, long total_words,
                        SList list, hipStream_t s);
// One iteration over `cur` (six subiteration launches, finalize, compaction
// into `next`, whose count is reset first).  max_entries bounds cur's count.
int sparse3d_iteration(const SItem3D *items, int nitems, SList cur, SList next, int max_entries,
                       uint32_t *prot, int *changed, int *iteration, int *live, int grid,
                       hipStream_t s);

}  // namespace bgt
