// seggraph.hip -- device pre-pass of the segment tracer: classification and
// stream compaction of skeleton points, so that the host traces only them.
//
// Reference: the tracer's first pass, 2D/ExtractSegments.cpp:32-50 and
// 3D/ExtractSegments.cpp:154-171, scans the whole padded frame for skeleton
// pixels off its border, in raster order, and codes each by its number of
// skeleton neighbours (8 or 26, border pixels included).  Its second pass
// follows segments through those codes.  Here the first pass runs on the
// device over the skeleton as it leaves the thinning kernels, and yields a
// compact graph in raster order: each point's raster index in the padded
// frame, its code, and the ranks of its neighbours that are points themselves
// (in the reference's neighbour order), so that the host's second pass needs
// no frame-sized array.
//
// The skeleton is the unpadded w x h x d uint8 buffer; the padded frame
// (pw x ph x pd, zero outside the skeleton) is virtual: it only decides which
// pixels lie off the border and the raster index.
//
// Kernels, one thread per 64-pixel chunk of a row (chunks in raster order):
//   seg_count  -- the chunk's point bits, their count, and each workgroup's sum;
//   seg_scan   -- one workgroup: exclusive sums of the workgroup sums, total;
//   seg_base   -- each chunk's first rank;
//   seg_place  -- the packed coordinates of the chunk's points;
// then one thread per point: seg_emit -- its code and neighbour ranks;
// then one thread per point (or half-edge) for the chains of path points:
//   seg_link, seg_jump  -- list ranking of the chains by pointer jumping;
//   seg_chain_flags, seg_scan, seg_chain_base -- chain ids and records in
//                          the order of each chain's first end;
//   seg_chain_scatter   -- every chain's points in order;
//   seg_gather          -- the graph's parts into the thread's pinned staging.
// The host's second pass then follows a chain as one run instead of point by
// point.
//
// Nothing goes through the runtime's copy engine: the totals the host waits
// for and the graph itself are written by kernels straight into mapped pinned
// memory.  Under load from other threads (serving), a device-to-host
// hipMemcpyAsync can block in the runtime for milliseconds before it is even
// queued; a kernel launch does not.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "pack.h"
#include "seggraph.h"

namespace bgt {

namespace {

struct Frame {
    const uint8_t *skel;  // w x h x d, rows of w bytes
    int w, h, d;          // skeleton
    int pw, ph, pd;       // padded frame (>= w, h, d)
    int cpr;              // chunks per row: ceil(w / 64)
    int three_d;
};

__device__ __forceinline__ bool off_border(const Frame &f, int x, int y, int z)
{
    return x >= 1 && x <= f.pw - 2 && y >= 1 && y <= f.ph - 2 &&
           (!f.three_d || (z >= 1 && z <= f.pd - 2));
}

// skeleton byte at (x, y, z), zero outside the skeleton's extent
__device__ __forceinline__ bool set_at(const Frame &f, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= f.w || y >= f.h || z >= f.d) return false;
    return f.skel[((size_t)z * f.h + y) * f.w + x] != 0;
}

__global__ __launch_bounds__(256) void seg_count(Frame f, long nchunks, uint64_t *bits,
                                                 int *cnt, int *wsum)
{
    __shared__ int s_sum[4];
    __shared__ uint32_t s_masks[4][320];
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    const bool valid = c < nchunks;
    const long row = valid ? c / f.cpr : 0;
    const int x0 = valid ? (int)(c - row * f.cpr) * 64 : 0;
    const int nx = min(64, f.w - x0);
    // the chunks of a contiguous skeleton follow each other in memory: the
    // wave loads them together
    uint64_t m = pack_runs_wave_u8(f.skel + (size_t)row * f.w + x0, nx, valid, s_masks[threadIdx.x >> 6]);
    int n = 0;
    if (valid) {
        const int y = (int)(row % f.h), z = (int)(row / f.h);
        const bool row_ok = y >= 1 && y <= f.ph - 2 && (!f.three_d || (z >= 1 && z <= f.pd - 2));
        if (!row_ok) m = 0;
        if (x0 == 0) m &= ~1ull;  // x = 0 is on the border
        const int xe = f.pw - 2 - x0;  // the last x off the border, relative to x0
        if (xe < 63) m &= xe < 0 ? 0ull : (2ull << xe) - 1ull;
        bits[c] = m;
        n = __popcll(m);
        cnt[c] = n;
    }
    // workgroup sum
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) wsum[blockIdx.x] = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
}

// one workgroup of 1024 threads: exclusive sums of wsum[0..nw) in place, and
// the total in *total
__global__ __launch_bounds__(1024) void seg_scan(int *wsum, int nw, long *total)
{
    __shared__ long s_part[1024];
    __shared__ long s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int b = 0; b < nw; b += 1024) {
        const int i = b + (int)threadIdx.x;
        const long v = i < nw ? wsum[i] : 0;
        s_part[threadIdx.x] = v;
        __syncthreads();
        // inclusive Hillis-Steele scan over the 1024 values
        for (int o = 1; o < 1024; o <<= 1) {
            const long a = threadIdx.x >= (unsigned)o ? s_part[threadIdx.x - o] : 0;
            __syncthreads();
            s_part[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nw) wsum[i] = (int)(s_carry + s_part[threadIdx.x] - v);
        __syncthreads();
        if (threadIdx.x == 1023) s_carry += s_part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s_carry;
}

// first rank of every chunk: its workgroup's base plus the counts before it
__global__ __launch_bounds__(256) void seg_base(long nchunks, const int *cnt, const int *wbase,
                                                int *base)
{
    __shared__ int s_w[4];
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    const int v = c < nchunks ? cnt[c] : 0;
    // inclusive scan within the wave, then across the four waves
    int s = v;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(s, o);
        if (lane >= o) s += t;
    }
    if (lane == 63) s_w[wv] = s;
    __syncthreads();
    int before = 0;
    for (int k = 0; k < wv; ++k) before += s_w[k];
    if (c < nchunks) base[c] = wbase[blockIdx.x] + before + s - v;
}

__device__ __forceinline__ int rank_of(const Frame &f, const uint64_t *bits, const int *base,
                                       int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= f.w || y >= f.h || z >= f.d) return -1;
    const long c = ((long)z * f.h + y) * f.cpr + (x >> 6);
    const uint64_t m = bits[c];
    const int k = x & 63;
    if (!((m >> k) & 1ull)) return -1;
    return base[c] + __popcll(m & ((1ull << k) - 1ull));
}

// one thread per chunk: the packed coordinates of every point of the chunk
__global__ __launch_bounds__(256) void seg_place(Frame f, long nchunks, const uint64_t *bits,
                                                 const int *base, int64_t *idx)
{
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= nchunks) return;
    uint64_t m = bits[c];
    if (!m) return;
    const long row = c / f.cpr;
    const int64_t x0 = (c - row * f.cpr) * 64;
    const int64_t yz = (int64_t)(row % f.h) << 16 | (int64_t)(row / f.h) << 32;
    int r = base[c];
    while (m) {
        const int k = __ffsll((long long)m) - 1;
        m &= m - 1;
        idx[r++] = (x0 + k) | yz;  // x | y << 16 | z << 32
    }
}

// one thread per point: its code and the ranks of its neighbours
__global__ __launch_bounds__(256) void seg_emit(Frame f, int npts, const uint64_t *bits,
                                                const int *base, const int64_t *idx, int8_t *code,
                                                int *nbr)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= npts) return;
    const uint64_t q = (uint64_t)idx[r];
    const int x = (int)(q & 0xffffu), y = (int)((q >> 16) & 0xffffu), z = (int)(q >> 32);
    const int nd = f.three_d ? 26 : 8;
    int nn = 0, j = 0;
    // the reference's neighbour order: raster order of the offsets
    for (int dz = f.three_d ? -1 : 0; dz <= (f.three_d ? 1 : 0); ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dz && !dy && !dx) continue;
                const int qx = x + dx, qy = y + dy, qz = z + dz;
                nn += set_at(f, qx, qy, qz);
                nbr[(size_t)r * nd + j] =
                    off_border(f, qx, qy, qz) ? rank_of(f, bits, base, qx, qy, qz) : -1;
                ++j;
            }
    if (nn == 0) nn = 1;
    code[r] = (int8_t)(nn == 2 ? 2 : -nn);
}

// ---- chains of path points -------------------------------------------------
//
// A path point (code 2) has at most two neighbours that are points; those that
// are path points too link it into a chain.  Half-edge 2p + s of path point p
// means "leave p through its slot s" (slot s = its s-th point neighbour in the
// reference's order).  It is terminal when that slot holds no path point: the
// walk ends at p, with the slot's point (or -1) as the neighbour off the
// chain.  Otherwise it continues as the half-edge of the next point that does
// not lead back to p.  Pointer jumping (Wyllie's list ranking) gives every
// half-edge its end point, its distance to it and the neighbour off the chain
// there; runs that close on themselves never end.

struct Jump {
    int *nxt;    // next half-edge, -1 terminal, -2 none (not a path point)
    int *dist;   // steps to the end
    int *end;    // end point (valid once terminal)
    int *xo;     // neighbour off the chain at the end
};

// the first two neighbours of p that are points
__device__ __forceinline__ void point_slots(const int *nbr, int nd, int p, int &a, int &b)
{
    a = b = -1;
    const int *row = nbr + (size_t)p * nd;
    for (int j = 0; j < nd; ++j) {
        const int r = row[j];
        if (r < 0) continue;
        if (a < 0) a = r;
        else { b = r; break; }
    }
}

__global__ __launch_bounds__(256) void seg_link(int npts, int nd, const int8_t *code, const int *nbr,
                                                Jump j)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npts) return;
    if (code[p] != 2) {
        j.nxt[2 * p] = j.nxt[2 * p + 1] = -2;
        return;
    }
    int sl[2];
    point_slots(nbr, nd, p, sl[0], sl[1]);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int q = sl[s], h = 2 * p + s;
        if (q >= 0 && code[q] == 2) {
            int qa, qb;
            point_slots(nbr, nd, q, qa, qb);
            j.nxt[h] = 2 * q + (qa == p ? 1 : 0);  // q's slot that does not lead back to p
            j.dist[h] = 1;
            j.end[h] = -1;
            j.xo[h] = -1;
        } else {
            j.nxt[h] = -1;
            j.dist[h] = 0;
            j.end[h] = p;
            j.xo[h] = q;
        }
    }
}

// One launch follows kHops + 1 pointers of the previous launch's state, so
// every half-edge's reach grows (kHops + 1)-fold (twice in the textbook
// round): ceil(log5 n) launches for chains of up to n points.
constexpr int kHops = 4;

__global__ __launch_bounds__(256) void seg_jump(int nh, Jump a, Jump b)
{
    const int h = blockIdx.x * 256 + threadIdx.x;
    if (h >= nh) return;
    int n = a.nxt[h], d = a.dist[h], last = h;
    for (int k = 0; k < kHops && n >= 0; ++k) {
        d += a.dist[n];
        last = n;
        n = a.nxt[n];
    }
    b.nxt[h] = n;
    b.dist[h] = d;
    b.end[h] = a.end[last];
    b.xo[h] = a.xo[last];
}

// exclusive prefix of v over the 256 threads of a workgroup (s_w: 4 ints)
__device__ __forceinline__ int wg_exclusive(int v, int *s_w)
{
    int s = v;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(s, o);
        if (lane >= o) s += t;
    }
    if (lane == 63) s_w[wv] = s;
    __syncthreads();
    int before = 0;
    for (int k = 0; k < wv; ++k) before += s_w[k];
    __syncthreads();
    return before + s - v;
}

struct ChainTmp {
    int *start;  // [npts] the chain's first end (e0) of a chain point, else -1
    int *pos;    // [npts] position from e0
    int *len;    // [npts] chain length, at e0 only (else 0)
    int *nent;   // [npts] a node's neighbours that are points (else 0)
    int *wch, *wlen, *wnode, *went;  // [nwg] workgroup sums of chain starts, lengths, nodes, node entries
    int *cid, *coff;  // [npts] chain id and first position, at e0 only
    int *eoff;   // [npts] a node's first entry
};

// What the host's second pass reads (SegGraph), each array sized for npts
// (node_ent for npts x nd).
struct GraphOut {
    int *node_rank;
    int64_t *node_xyz;
    int *node_off;
    int64_t *node_ent;
    SegChain *rec;
    int64_t *chain_xyz;
};

// per path point: its chain's e0 and its position; e0 also its length; per
// node: its neighbours that are points; workgroup sums of chain starts, chain
// lengths, nodes and node entries
__global__ __launch_bounds__(256) void seg_chain_flags(int npts, int nd, const int8_t *code, const int *nbr,
                                                       Jump j, ChainTmp c)
{
    __shared__ int s_a[4], s_b[4], s_c[4], s_d[4];
    const int p = blockIdx.x * 256 + threadIdx.x;
    int st = 0, ln = 0, nodes = 0, ne = 0;
    if (p < npts) {
        int s0 = -1, ps = 0;
        const int8_t cd = code[p];
        if (cd == 2 && j.nxt[2 * p] == -1 && j.nxt[2 * p + 1] == -1) {
            const int a = j.end[2 * p], b = j.end[2 * p + 1];
            s0 = a < b ? a : b;
            ps = s0 == a ? j.dist[2 * p] : j.dist[2 * p + 1];
            if (p == s0) {
                st = 1;
                ln = j.dist[2 * p] + j.dist[2 * p + 1] + 1;
            }
        }
        nodes = cd < 0;
        if (nodes)
            for (int k = 0; k < nd; ++k) ne += nbr[(size_t)p * nd + k] >= 0;
        c.start[p] = s0;
        c.pos[p] = ps;
        c.len[p] = ln;
        c.nent[p] = ne;
    }
    int x = st, y = ln, z = nodes, w = ne;
    for (int o = 32; o > 0; o >>= 1) {
        x += __shfl_xor(x, o);
        y += __shfl_xor(y, o);
        z += __shfl_xor(z, o);
        w += __shfl_xor(w, o);
    }
    if ((threadIdx.x & 63) == 0) {
        s_a[threadIdx.x >> 6] = x;
        s_b[threadIdx.x >> 6] = y;
        s_c[threadIdx.x >> 6] = z;
        s_d[threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        c.wch[blockIdx.x] = s_a[0] + s_a[1] + s_a[2] + s_a[3];
        c.wlen[blockIdx.x] = s_b[0] + s_b[1] + s_b[2] + s_b[3];
        c.wnode[blockIdx.x] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
        c.went[blockIdx.x] = s_d[0] + s_d[1] + s_d[2] + s_d[3];
    }
}

// chain ids and first positions in the order of e0 and the chain records;
// the nodes in rank order with their coordinates and first entries
__global__ __launch_bounds__(256) void seg_chain_base(int npts, const int64_t *idx, const int8_t *code,
                                                      Jump j, ChainTmp c, GraphOut o)
{
    __shared__ int s_w[4];
    const int p = blockIdx.x * 256 + threadIdx.x;
    const bool st = p < npts && c.start[p] == p;
    const bool node = p < npts && code[p] < 0;
    const int ln = st ? c.len[p] : 0;
    const int id = c.wch[blockIdx.x] + wg_exclusive(st ? 1 : 0, s_w);
    const int off = c.wlen[blockIdx.x] + wg_exclusive(ln, s_w);
    const int nid = c.wnode[blockIdx.x] + wg_exclusive(node ? 1 : 0, s_w);
    const int eo = c.went[blockIdx.x] + wg_exclusive(node ? c.nent[p] : 0, s_w);
    if (node) {
        o.node_rank[nid] = p;
        o.node_xyz[nid] = idx[p];
        o.node_off[nid] = eo;
        c.eoff[p] = eo;
    }
    if (!st) return;
    c.cid[p] = id;
    c.coff[p] = off;
    // e0 = p: one half-edge ends at p itself (its neighbour off the chain at
    // e0), the other at e1 (or at p too, for a single point)
    const int h0 = 2 * p, h1 = 2 * p + 1;
    const bool first_here = j.end[h0] == p && j.dist[h0] == 0;
    SegChain r;
    r.off = off;
    r.len = ln;
    r.e0 = p;
    if (ln == 1) {
        r.e1 = p;
        r.x0 = j.xo[h0];
        r.x1 = j.xo[h1];
    } else if (first_here) {
        r.e1 = j.end[h1];
        r.x0 = j.xo[h0];
        r.x1 = j.xo[h1];
    } else {
        r.e1 = j.end[h0];
        r.x0 = j.xo[h1];
        r.x1 = j.xo[h0];
    }
    r.x0xyz = r.x0 >= 0 ? idx[r.x0] : 0;
    r.x1xyz = r.x1 >= 0 ? idx[r.x1] : 0;
    o.rec[id] = r;
}

// every chain point: its coordinates at its place in the chain's run; every
// node: its entries (neighbours that are points, with their chains)
__global__ __launch_bounds__(256) void seg_chain_scatter(int npts, int nd, const int64_t *idx,
                                                         const int8_t *code, const int *nbr, ChainTmp c,
                                                         GraphOut o)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npts) return;
    const int s0 = c.start[p];
    if (s0 >= 0) o.chain_xyz[c.coff[s0] + c.pos[p]] = idx[p];
    if (code[p] < 0) {
        int e = c.eoff[p];
        for (int k = 0; k < nd; ++k) {
            const int r = nbr[(size_t)p * nd + k];
            if (r < 0) continue;
            const int rs = c.start[r];
            const int64_t ch = rs >= 0 ? (int64_t)c.cid[rs] : -1;
            o.node_ent[e++] = (int64_t)(uint32_t)r | (int64_t)k << 32 | (ch + 1) << 37;
        }
    }
}

// the tracer's foreground of a pixel array: nonzero -> 1
template <typename T>
__global__ __launch_bounds__(256) void nonzero_u8(const T *__restrict__ in, uint8_t *__restrict__ m, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        m[i] = in[i] != (T)0;
}

// Grow-only pinned host staging of one thread's pre-pass results (device to
// host copies into pageable memory go through the runtime's own staging),
// released at thread exit.
// the graph's parts, each from a 256-byte aligned device region to a 16-byte
// aligned place in mapped pinned memory; blockIdx.y picks the part
struct GatherParts {
    const char *src[8];
    char *dst[8];
    size_t bytes[8];
};

__global__ __launch_bounds__(256) void seg_gather(GatherParts g)
{
    const int k = blockIdx.y;
    const size_t nb = g.bytes[k], n16 = nb / 16;
    const uint4 *s = reinterpret_cast<const uint4 *>(g.src[k]);
    uint4 *d = reinterpret_cast<uint4 *>(g.dst[k]);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256) d[i] = s[i];
    if (blockIdx.x == 0 && threadIdx.x < nb - n16 * 16) g.dst[k][n16 * 16 + threadIdx.x] = g.src[k][n16 * 16 + threadIdx.x];
}

struct PinnedStage {
    char *ptr = nullptr;   // host address
    char *dptr = nullptr;  // the same memory as kernels address it
    size_t bytes = 0;
    // outgrown buffers, kept until the thread exits: hipHostFree waits for
    // the device (see Workspace)
    std::vector<char *> retired;
    ~PinnedStage()
    {
        if (on_main_thread()) return;
        if (ptr) (void)hipHostFree(ptr);
        for (char *p : retired) (void)hipHostFree(p);
    }
    int reserve(size_t n)
    {
        if (ptr && bytes >= n) return BGT_OK;
        const size_t want = grow_bytes(n, bytes);
        if (ptr) retired.push_back(ptr);
        ptr = dptr = nullptr;
        bytes = 0;
        // coherent: what kernels write here is in host memory when the
        // stream's work is done, whatever scope the runtime's fences take
        if (hipHostMalloc((void **)&ptr, want, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
            (void)hipGetLastError();
            ptr = nullptr;
            return fail(BGT_ERR_NOMEM, "pinned staging of %zu bytes failed", want);
        }
        if (hipHostGetDevicePointer((void **)&dptr, ptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            retired.push_back(ptr);
            ptr = dptr = nullptr;
            return fail(BGT_ERR_HIP, "pinned staging has no device address");
        }
        bytes = want;
        return BGT_OK;
    }
};

PinnedStage &thread_stage()
{
    thread_local PinnedStage st;
    return st;
}

// Device bytes of the pre-pass's three workspaces: the point count over a
// frame of nchunks 64-pixel chunks (bits, counts, bases, workgroup sums,
// total); the per-point arrays and what the host reads, for n points; the
// chain ranking, for n points.
struct CountLayout {
    size_t bits, ints, wsum;
    explicit CountLayout(long nchunks)
        : bits(round_up(sizeof(uint64_t) * nchunks, 256)), ints(round_up(sizeof(int) * nchunks, 256)),
          wsum(round_up(sizeof(int) * ((nchunks + 255) / 256), 256))
    {
    }
    size_t total() const { return bits + 2 * ints + wsum + 256; }
};

struct GraphLayout {
    size_t idx, code, nbr, ints, xyz, ent, rec;
    GraphLayout(long n, int nd)
        : idx(round_up(sizeof(int64_t) * n, 256)), code(round_up((size_t)n, 256)),
          nbr(round_up(sizeof(int) * (size_t)n * nd, 256)), ints(round_up(sizeof(int) * ((size_t)n + 1), 256)),
          xyz(round_up(sizeof(int64_t) * n, 256)), ent(round_up(sizeof(int64_t) * (size_t)n * nd, 256)),
          rec(round_up(sizeof(SegChain) * (size_t)n, 256))
    {
    }
    size_t total() const { return idx + nbr + code + 2 * ints + 2 * xyz + ent + rec; }
};

struct ChainLayout {
    size_t half, one, wsz;
    explicit ChainLayout(long n)
        : half(round_up(sizeof(int) * 2 * (size_t)n, 256)), one(round_up(sizeof(int) * (size_t)n, 256)),
          wsz(round_up(sizeof(int) * (size_t)((n + 255) / 256), 256))
    {
    }
    size_t total() const { return 8 * half + 7 * one + 4 * wsz + 256; }
};

}  // namespace

int segment_graph_host(int dtype, const void *data, int w, int h, int d, bool three_d, hipStream_t s,
                       SegGraph &g)
{
    g = SegGraph{};
    g.nd = three_d ? 26 : 8;
    const size_t es = dtype_size(dtype);
    if (!data || !es) return fail(BGT_ERR_INVALID, "bad pixel array or type %d", dtype);
    if (w <= 0 || h <= 0 || d <= 0) return fail(BGT_ERR_INVALID, "bad size");
    const size_t n = (size_t)w * h * d;
    const size_t raw = round_up(n * es, 256);
    Workspace &wi = thread_workspace(4);
    BGT_CHECK(workspace_reserve(wi, raw + (es > 1 ? round_up(n, 256) : 0)));
    char *d_raw = (char *)wi.ptr;
    BGT_HIP(hipMemcpyAsync(d_raw, data, n * es, hipMemcpyHostToDevice, s));
    const uint8_t *mask = (const uint8_t *)d_raw;  // one-byte pixels: nonzero bytes are the foreground
    if (es > 1) {
        uint8_t *m = (uint8_t *)(d_raw + raw);
        const unsigned grid = (unsigned)std::min<long>(((long)n + 255) / 256, 65536L);
        switch (dtype) {
        case BGT_U16: hipLaunchKernelGGL(nonzero_u8<uint16_t>, dim3(grid), dim3(256), 0, s, (const uint16_t *)d_raw, m, (long)n); break;
        case BGT_I16: hipLaunchKernelGGL(nonzero_u8<int16_t>, dim3(grid), dim3(256), 0, s, (const int16_t *)d_raw, m, (long)n); break;
        case BGT_U32: hipLaunchKernelGGL(nonzero_u8<uint32_t>, dim3(grid), dim3(256), 0, s, (const uint32_t *)d_raw, m, (long)n); break;
        case BGT_I32: hipLaunchKernelGGL(nonzero_u8<int32_t>, dim3(grid), dim3(256), 0, s, (const int32_t *)d_raw, m, (long)n); break;
        case BGT_F32: hipLaunchKernelGGL(nonzero_u8<float>, dim3(grid), dim3(256), 0, s, (const float *)d_raw, m, (long)n); break;
        case BGT_F64: hipLaunchKernelGGL(nonzero_u8<double>, dim3(grid), dim3(256), 0, s, (const double *)d_raw, m, (long)n); break;
        default: return fail(BGT_ERR_INVALID, "unknown pixel type %d", dtype);
        }
        BGT_HIP(hipGetLastError());
        mask = m;
    }
    return segment_graph_device(mask, w, h, d, w, h, d, three_d, s, g);
}

int segment_reserve(int w, int h, int d, bool three_d, long npts)
{
    if (w <= 0 || h <= 0 || d <= 0 || npts < 0) return fail(BGT_ERR_INVALID, "bad size");
    BGT_CHECK(workspace_reserve(thread_workspace(5), CountLayout((long)d * h * ((w + 63) / 64)).total()));
    if (!npts) return BGT_OK;
    BGT_CHECK(workspace_reserve(thread_workspace(6), GraphLayout(npts, three_d ? 26 : 8).total()));
    BGT_CHECK(workspace_reserve(thread_workspace(7), ChainLayout(npts).total()));
    // a graph takes about 16 bytes per point to the host
    return thread_stage().reserve((size_t)npts * 32 + 4096);
}

int segment_graph_device(const uint8_t *d_skel, int w, int h, int d, int pw, int ph, int pd,
                         bool three_d, hipStream_t s, SegGraph &g)
{
    g = SegGraph{};
    g.nd = three_d ? 26 : 8;
    if (w <= 0 || h <= 0 || d <= 0) return BGT_OK;
    if (pw < w || ph < h || pd < d) return fail(BGT_ERR_INVALID, "padded frame smaller than skeleton");
    if (pw > 32767 || ph > 32767 || pd > 32767)
        return fail(BGT_ERR_UNSUPPORTED, "coordinates are int16 (sides <= 32767)");
    Frame f{d_skel, w, h, d, pw, ph, pd, (w + 63) / 64, three_d ? 1 : 0};
    const long nchunks = (long)d * h * f.cpr;
    const long nwg = (nchunks + 255) / 256;
    if (nwg > 0x7fffffffL) return fail(BGT_ERR_UNSUPPORTED, "skeleton too large for the pre-pass");
    // scratch: bits, counts, bases, workgroup sums, total
    const CountLayout cl(nchunks);
    Workspace &ws = thread_workspace(5);
    BGT_CHECK(workspace_reserve(ws, cl.total()));
    char *p = (char *)ws.ptr;
    uint64_t *bits = (uint64_t *)p;
    p += cl.bits;
    int *cnt = (int *)p;
    p += cl.ints;
    int *base = (int *)p;
    p += cl.ints;
    int *wsum = (int *)p;
    // the point count goes straight to the pinned staging
    PinnedStage &st = thread_stage();
    BGT_CHECK(st.reserve(4096));
    hipLaunchKernelGGL(seg_count, dim3((unsigned)nwg), dim3(256), 0, s, f, nchunks, bits, cnt, wsum);
    hipLaunchKernelGGL(seg_scan, dim3(1), dim3(1024), 0, s, wsum, (int)nwg, (long *)st.dptr);
    hipLaunchKernelGGL(seg_base, dim3((unsigned)nwg), dim3(256), 0, s, nchunks, cnt, wsum, base);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(s));
    long total = 0;
    std::memcpy(&total, st.ptr, sizeof(long));
    if (total > 0x7fffffffL / g.nd) return fail(BGT_ERR_UNSUPPORTED, "too many skeleton points");
    g.npts = (int)total;
    if (!g.npts) return BGT_OK;
    const int n = g.npts;
    // the pre-pass's per-point arrays (device only), then what the host
    // reads, each region sized for the most there can be (only the used part
    // of each is copied)
    const GraphLayout gl(n, g.nd);
    Workspace &wo = thread_workspace(6);
    BGT_CHECK(workspace_reserve(wo, gl.total()));
    char *q = (char *)wo.ptr;
    int64_t *d_idx = (int64_t *)q;
    q += gl.idx;
    int *d_nbr = (int *)q;
    q += gl.nbr;
    int8_t *d_code = (int8_t *)q;
    q += gl.code;
    GraphOut go;
    go.node_rank = (int *)q;
    q += gl.ints;
    go.node_off = (int *)q;
    q += gl.ints;
    go.node_xyz = (int64_t *)q;
    q += gl.xyz;
    go.chain_xyz = (int64_t *)q;
    q += gl.xyz;
    go.node_ent = (int64_t *)q;
    q += gl.ent;
    go.rec = (SegChain *)q;
    hipLaunchKernelGGL(seg_place, dim3((unsigned)nwg), dim3(256), 0, s, f, nchunks, bits, base, d_idx);
    hipLaunchKernelGGL(seg_emit, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, n, bits, base,
                       d_idx, d_code, d_nbr);
    BGT_HIP(hipGetLastError());

    // chains: link, ceil(log5 n) jump launches; chain starts and lengths,
    // nodes and their entries, their prefix sums; records, nodes, runs
    const int pwg = (n + 255) / 256, hwg = (2 * n + 255) / 256;
    const ChainLayout chl(n);
    const size_t half = chl.half, one = chl.one, wsz = chl.wsz;
    Workspace &wt = thread_workspace(7);
    BGT_CHECK(workspace_reserve(wt, chl.total()));
    char *t = (char *)wt.ptr;
    Jump jb[2];
    for (Jump &j : jb) {
        j.nxt = (int *)t; t += half;
        j.dist = (int *)t; t += half;
        j.end = (int *)t; t += half;
        j.xo = (int *)t; t += half;
    }
    ChainTmp ct;
    ct.start = (int *)t; t += one;
    ct.pos = (int *)t; t += one;
    ct.len = (int *)t; t += one;
    ct.nent = (int *)t; t += one;
    ct.cid = (int *)t; t += one;
    ct.coff = (int *)t; t += one;
    ct.eoff = (int *)t; t += one;
    ct.wch = (int *)t; t += wsz;
    ct.wlen = (int *)t; t += wsz;
    ct.wnode = (int *)t; t += wsz;
    ct.went = (int *)t; t += wsz;
    // to the pinned staging: [0] chains, [1] chain points, [2] nodes, [3] node entries
    long *d_tot = (long *)st.dptr;
    hipLaunchKernelGGL(seg_link, dim3((unsigned)pwg), dim3(256), 0, s, n, g.nd, d_code, d_nbr, jb[0]);
    int cur = 0;
    for (long reach = 1; reach < n; reach *= kHops + 1, cur ^= 1)
        hipLaunchKernelGGL(seg_jump, dim3((unsigned)hwg), dim3(256), 0, s, 2 * n, jb[cur], jb[cur ^ 1]);
    hipLaunchKernelGGL(seg_chain_flags, dim3((unsigned)pwg), dim3(256), 0, s, n, g.nd, d_code, d_nbr, jb[cur], ct);
    hipLaunchKernelGGL(seg_scan, dim3(1), dim3(1024), 0, s, ct.wch, pwg, d_tot);
    hipLaunchKernelGGL(seg_scan, dim3(1), dim3(1024), 0, s, ct.wlen, pwg, d_tot + 1);
    hipLaunchKernelGGL(seg_scan, dim3(1), dim3(1024), 0, s, ct.wnode, pwg, d_tot + 2);
    hipLaunchKernelGGL(seg_scan, dim3(1), dim3(1024), 0, s, ct.went, pwg, d_tot + 3);
    hipLaunchKernelGGL(seg_chain_base, dim3((unsigned)pwg), dim3(256), 0, s, n, d_idx, d_code, jb[cur], ct, go);
    hipLaunchKernelGGL(seg_chain_scatter, dim3((unsigned)pwg), dim3(256), 0, s, n, g.nd, d_idx, d_code, d_nbr,
                       ct, go);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(s));
    long tot[4];
    std::memcpy(tot, st.ptr, sizeof(tot));
    if (tot[0] < 0 || tot[0] > n || tot[1] < tot[0] || tot[1] > n || tot[2] < 0 || tot[2] > n || tot[3] < 0 ||
        tot[3] > (long)n * g.nd)
        return fail(BGT_ERR_HIP, "segment pre-pass: %ld chains of %ld points, %ld nodes of %ld entries among %d points",
                    tot[0], tot[1], tot[2], tot[3], n);
    // one pinned buffer: the codes, the nodes' ranks, first entries (and the
    // total after the last), coordinates and entries, the chain records, the
    // chains' coordinates
    const size_t nn = (size_t)tot[2];
    struct Part {
        const void *src;
        size_t bytes;
    } parts[] = {{d_code, (size_t)n},
                 {go.node_rank, sizeof(int) * nn},
                 {go.node_off, sizeof(int) * nn},
                 {go.node_xyz, sizeof(int64_t) * nn},
                 {go.node_ent, sizeof(int64_t) * (size_t)tot[3]},
                 {go.rec, sizeof(SegChain) * (size_t)tot[0]},
                 {go.chain_xyz, sizeof(int64_t) * (size_t)tot[1]}};
    constexpr int kParts = 7;
    size_t at[kParts], bytes = 0;
    for (int k = 0; k < kParts; ++k) {
        at[k] = bytes;
        bytes += round_up(parts[k].bytes + (k == 2 ? sizeof(int) : 0), 16);  // room for node_off[nn]
    }
    BGT_CHECK(st.reserve(bytes));
    GatherParts gp{};
    size_t most = 0;
    for (int k = 0; k < kParts; ++k) {
        gp.src[k] = (const char *)parts[k].src;
        gp.dst[k] = st.dptr + at[k];
        gp.bytes[k] = parts[k].bytes;
        most = std::max(most, parts[k].bytes);
    }
    const unsigned gx = (unsigned)std::min<size_t>(std::max<size_t>((most + 4095) / 4096, 1), 1024);
    hipLaunchKernelGGL(seg_gather, dim3(gx, kParts), dim3(256), 0, s, gp);
    BGT_HIP(hipGetLastError());
    BGT_HIP(hipStreamSynchronize(s));
    // the host trace reads these at random: from ordinary pageable memory
    // (this thread's copy, kept between calls) it runs faster than from the
    // pinned buffer
    thread_local std::vector<char> hv;
    if (hv.size() < bytes) hv.resize(bytes);
    std::memcpy(hv.data(), st.ptr, bytes);
    char *hq = hv.data();
    const int total_ent = (int)tot[3];
    std::memcpy(hq + at[2] + sizeof(int) * nn, &total_ent, sizeof(int));  // node_off[nn]
    g.code = (const int8_t *)(hq + at[0]);
    g.nnodes = (int)nn;
    g.node_rank = (const int *)(hq + at[1]);
    g.node_off = (const int *)(hq + at[2]);
    g.node_xyz = (const int64_t *)(hq + at[3]);
    g.node_ent = (const int64_t *)(hq + at[4]);
    g.nchains = (int)tot[0];
    g.chains = (const SegChain *)(hq + at[5]);
    g.chain_xyz = (const int64_t *)(hq + at[6]);
    return BGT_OK;
}

}  // namespace bgt
