// queue2d.h -- the 2D band work queue shared by the batch engine and the
// continuously running serving queue (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "thin2d.h"
#include "thin2d_kernels.h"

namespace bgt {

// Work queue of the 2D bands.  Tile ids index `tiles`, `stamp` and `occ`.
// Pass P reads list P % 3 and appends to list (P + 1) % 3; list (P + 2) % 3
// is idle during pass P (read by pass P - 1, appended to by pass P + 1), so
// pass P clears its counters without racing any reader or writer.
struct Queue {
    int *list[3];   // active tile ids, by pass % 3
    int *cnt;       // [3] entries in each list
    int *wctr;      // [3] next list slot to hand out, by pass % 3
    int *stamp;     // per tile: last pass it was queued for
    uint8_t *occ;   // per tile: interior non-empty
    int ntiles;
};

// Per-item convergence of the 2D batch.  The tiles of pass P OR their
// iteration masks into hist[P % 4][item]; pass P + 1 skips the tiles of an
// item whose pass-P mask misses an iteration (that iteration deleted nothing
// in the whole image, so it is a fixed point), and pass P + 2 closes the item
// from that mask.  Slot (P + 1) % 4 is cleared during pass P: its last reader
// is pass P - 1 and its next writer pass P + 1.
struct Close2D {
    unsigned *hist;  // [4][nitems]
    int *states;     // ST_INTS words per item
    int nitems;
    // host-mapped [16] (or null): at P % 16, (P + 1) << 32 | bands queued for
    // pass P, stored when pass P starts
    unsigned long long *report;
    int ncu;         // compute units of the device
    // host-mapped [nitems] (or null): set to 1 when the item is closed
    unsigned *closed;
};

// One pass P = (pbase ? *pbase : 0) + p of the band queue (thin2d_queue_pass
// in batch.hip, every item's bands planned with V words per lane) on `grid`
// band workgroups plus the bookkeeping one.
// `first`: pass 0 of a batch whose items all start there, packing the items
// from the caller's input `io` (thin2d_queue_pass, one word per lane).
int launch_queue_pass2d(const Config2D &c2, int V, int grid, const Item2D *items, const Tile2D *tiles,
                        const Queue &q, const Close2D &c, const int *pbase, int p, hipStream_t s,
                        const ItemIO2D *io = nullptr, bool first = false);
// Band workgroups of a pass: as many as fit (at most per_cu_cap per CU when
// > 0) and no more than ntiles.
int queue_pass_grid2d(const Config2D &c2, int V, int ntiles, int per_cu_cap);

// kernels of thin2d.hip shared with the serving queue
void launch_unpack_item2d(const Item2D *d_item, uint8_t *out, int w, int h, size_t pitch,
                          hipStream_t s);

}  // namespace bgt
