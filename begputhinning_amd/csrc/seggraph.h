// seggraph.h -- device pre-pass of the segment tracer (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace bgt {

// A chain: a maximal run of path points (code 2) linked through each other,
// from end e0 to end e1 (e0 == e1 for a single point), e0 the end of lower
// rank.  x0 / x1 are the points next to e0 / e1 off the chain (-1: none),
// with their packed coordinates; for a single point, its two neighbours that
// are points (in the reference's order).  Runs that close on themselves form
// no chain.
struct SegChain {
    int off, len;  // the chain's points: chain_xyz[off .. off + len), from e0 to e1
    int e0, e1, x0, x1;
    int64_t x0xyz, x1xyz;
};

// The skeleton points of a padded frame as the host's second pass needs them,
// points ranked in raster order and coordinates packed as x | y << 16 |
// z << 32 (sides are at most 32767):
//   - per point, the tracer's code (-1 isolated, 2 path, -n end or junction
//     with n neighbours);
//   - the nodes (points of negative code) in rank order: rank, coordinates,
//     and their neighbours that are points, in the reference's neighbour
//     order, each as rank | slot << 32 | (chain + 1) << 37 (slot: its place
//     among the nd neighbours; chain: the chain of a path point, else -1),
//     node k's at node_ent[node_off[k] .. node_off[k + 1]);
//   - the chains, and their points' coordinates in order.
// The arrays are views of a per-thread host buffer, valid until the thread's
// next segment_graph_device call.
struct SegGraph {
    int npts = 0;
    int nd = 8;
    const int8_t *code = nullptr;
    int nnodes = 0;
    const int *node_rank = nullptr;
    const int64_t *node_xyz = nullptr;
    const int *node_off = nullptr;      // nnodes + 1
    const int64_t *node_ent = nullptr;  // node_off[nnodes]
    int nchains = 0;
    const SegChain *chains = nullptr;
    const int64_t *chain_xyz = nullptr;
};

// d_skel: w x h x d uint8 skeleton in device memory (nonzero = set), the top
// left corner of a zero-padded pw x ph x pd frame (d = pd = 1 in 2D).
int segment_graph_device(const uint8_t *d_skel, int w, int h, int d, int pw, int ph, int pd,
                         bool three_d, hipStream_t s, SegGraph &g);
// The same for a host array of any pixel type (nonzero = set) as a frame of
// its own size: the array is copied to the device first.
int segment_graph_host(int dtype, const void *data, int w, int h, int d, bool three_d, hipStream_t s,
                       SegGraph &g);

// Reserves this thread's pre-pass workspaces and pinned staging for a frame
// of w x h x d with up to npts skeleton points, so that calls up to that size
// allocate nothing (serving tracers call it when they start).
int segment_reserve(int w, int h, int d, bool three_d, long npts);

}  // namespace bgt
