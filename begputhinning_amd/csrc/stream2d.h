// stream2d.h -- a continuously running 2D band queue for serving (internal).
//
// Images join a running queue at pass boundaries instead of waiting for a
// batch to finish: the engine holds `slots` item slots, each with its own
// state buffers (sized for the largest image it admits) and band range, and a
// driver thread keeps the band-queue passes of batch.hip flowing while any
// slot is live.  Before queueing pass P it admits waiting images into free
// slots -- three small launches on the pass stream, ordered after pass P - 1:
// the slot's item and state (first pass P), the pack into both buffers, and
// the image's non-empty bands appended to pass P's queue.  Each item then
// converges on its own: the pass kernel counts its iterations from its first
// pass and closes it two passes after the pass whose mask misses an
// iteration, flagging it in host-mapped memory; the driver unpacks it on a
// side stream, after the passes queued so far (the closing pass's state
// writes are then visible), and reports the request done.  Results equal the
// single-image path's.
//
// No counterpart in the reference (it thins one image per blocking call,
// 2D/LibThinning2D.cu:39-130).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>

#include "thin2d.h"

namespace bgt {

struct Stream2D;

struct Stream2DConfig {
    int slots = 64;        // images in the queue at once
    int max_w = 4096;      // largest image a slot holds
    int max_h = 4096;
    int blocks_per_cu = 2; // band workgroups per CU of a pass
};

// Called on the driver thread when a request completes (status BGT_OK, or the
// error that ended it), with the tag given at submission.
using Stream2DDone = std::function<void(void *tag, int status, int iterations)>;

// Allocates the slots (stream2d_device_bytes) and starts the driver thread.
int stream2d_create(const Stream2DConfig &cfg, const Config2D &c2, Stream2DDone done,
                    Stream2D **out);
// Finishes the running images, fails those not yet admitted, stops the driver
// and frees everything.
void stream2d_destroy(Stream2D *e);
// Whether a w x h image fits a slot under the engine's launch shape.
bool stream2d_fits(const Stream2D *e, int w, int h);
// Queues an image (device buffers, inputs complete on the device).
int stream2d_submit(Stream2D *e, const uint8_t *in, size_t in_pitch, int w, int h, uint8_t *out,
                    size_t out_pitch, void *tag);
size_t stream2d_device_bytes(const Stream2D *e);
// Passes run and images admitted so far.
void stream2d_stats(const Stream2D *e, uint64_t *passes, uint64_t *admitted);

}  // namespace bgt
