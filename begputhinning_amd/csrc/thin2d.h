// thin2d.h -- host-side interface of the 2D thinning engine (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace bgt {

struct Item2D;
struct Tile2D;

// Launch shape of the fused 2D pass kernel.
struct Config2D {
    int waves = 4;           // wave64s per workgroup
    int rows = 16;           // register rows per lane (8 or 16); 16 fits every width
    int iters_per_pass = 3;  // K: iterations fused per HBM pass (halo 4K rows)
    // batch queue workgroups per CU: 0 = as many as fit, -1 = automatic (two in
    // a batch that also holds volumes, whose chain needs the remaining slots,
    // else as many as fit)
    int blocks_per_cu = -1;
    // 64-bit words per lane (1, 2 or 4), 0 = automatic: the fewest that hold
    // a row in one wave64.  Rows wider than the choice take more.
    int lane_words = 0;
    // batches (one word per lane): 1 = pass 0 packs the images from the
    // caller's bytes itself, 0 = a separate pack launch (faster: the pack
    // runs at full occupancy, pass 0's band loads at the pass's)
    int fused_io = 0;
    // batches: 1 = the whole 2D chain in one persistent launch (passes behind
    // a device-wide barrier), 0 = a launch per pass, driven by the host
    int persistent = 0;
};

// Widest 2D image: a row in one wave64 of lanes of four 64-bit words.
constexpr int kMaxWidth2D = 64 * 64 * 4;

struct Plan2D {
    int w = 0, h = 0, L = 1;  // L = 64-bit words per row
    int V = 1;                // words per lane
    int band_rows = 0, nbands = 0;
    size_t state_words = 0;
};

// Words per lane a w-wide image is planned with under cfg (0: too wide).
int lane_words2d(int w, const Config2D &cfg);
// V = 0: lane_words2d's choice; else V words per lane (at least the image needs)
int plan2d(int w, int h, const Config2D &cfg, Plan2D &pl, int V = 0);
int check_config2d(const Config2D &cfg);
int band_rows_for(int L, int V, const Config2D &cfg);
// LDS bytes of a pass workgroup with V words per lane
size_t pass_lds_bytes2d(const Config2D &cfg, int V);
size_t workspace_bytes2d(const Plan2D &pl);
int setup_item2d(const Plan2D &pl, void *ws, hipStream_t s, Item2D **d_item, unsigned **d_ctr,
                 Item2D *host_item);
// host_poll (may be null: a per-thread pinned buffer) receives ST_INTS words
int run_passes2d(const Config2D &cfg, const Plan2D &pl, Item2D *d_item, unsigned *d_ctr,
                 const int *d_state, hipStream_t s, int *iterations, int *host_poll = nullptr);

int thin2d_device_u8(const uint8_t *d_in, size_t in_pitch, int w, int h, uint8_t *d_out,
                     size_t out_pitch, hipStream_t s, int *iterations, const Config2D &cfg);
int thin2d_host(int dtype, int w, int h, const void *in, void *out, const Config2D &cfg);
// thin2d_device_u8 on the library's fixed per-GPU engine array: no allocation
int thin2d_device_u8_noalloc(const uint8_t *d_in, size_t in_pitch, int w, int h, uint8_t *d_out,
                             size_t out_pitch, hipStream_t s, int *iterations, const Config2D &cfg);
size_t thin2d_noalloc_max_bytes();

Config2D default_config2d();  // a snapshot of the current shape

}  // namespace bgt

namespace bgt {
int thin2d_host_mask(int dtype, int w, int h, const void *in, uint8_t *out, const Config2D &cfg);
// as thin2d_host_mask, but the w x h uint8 skeleton stays on the device, in
// this thread's I/O workspace (valid until the thread's next thinning call)
int thin2d_host_keep(int dtype, int w, int h, const void *in, const uint8_t **d_skel,
                     const Config2D &cfg);
}
