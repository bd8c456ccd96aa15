// thin3d.h -- host-side interface of the 3D thinning engine (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace bgt {

// Launch shape of the 3D word-list engine.
struct Config3D {
    int blocks_per_cu = 2;  // step-kernel workgroups per CU; 0 = occupancy limit
    int max_chunk = 16;     // most iterations queued between two host polls
    // once at most this many words per volume are active on average, the
    // remaining iterations switch to the tile engine (engine 1) or run in one
    // launch, one workgroup per volume (engine 0); 0 = never
    int tail_words = 16384;
    // 0: the non-empty-word list engine (sparse3d.h, six step launches per
    // iteration) with its one-launch tail; 1: that engine, then the tile
    // engine (tile3d.h, one launch per iteration) from tail_words on; 2: the
    // tile engine from the first iteration
    int engine = 1;
};

int check_config3d(const Config3D &c);
int thin3d_device_u8(const uint8_t *d_in, int w, int h, int d, uint8_t *d_out, hipStream_t s,
                     int *iterations, const Config3D &cfg);
int thin3d_host(int dtype, int w, int h, int d, const void *in, void *out, const Config3D &cfg);
// thins the foreground int(v) > 0 of a host volume of any pixel type (the
// 3D centerline's threshold); the w x h x d uint8 skeleton stays on the
// device, in this thread's I/O workspace (valid until its next thinning call)
int thin3d_keep_int_positive(int dtype, int w, int h, int d, const void *in,
                             const uint8_t **d_skel, const Config3D &cfg);

Config3D default_config3d();  // a snapshot of the current shape

}  // namespace bgt
