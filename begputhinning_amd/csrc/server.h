// server.h -- continuous-batching request loop (internal).
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/bgt.h"
#include "thin2d.h"
#include "thin3d.h"

namespace bgt {

struct ServerConfig {
    size_t budget = 0;    // device bytes the engines of admitted requests may use
    int workers2d = 2;    // worker threads (each a stream + batch engine) for 2D
    int workers3d = 1;    // and for 3D requests
    int max_items = 64;   // most requests per batch
    // frames held by the continuously running 2D queue at once (0: none; all
    // frames then go to the 2D workers)
    int stream_slots = 64;
};

struct Server;
int server_create(const ServerConfig &cfg, const Config2D &c2, const Config3D &c3, Server **out);
void server_destroy(Server *sv);
// cls 0 = 2D (d ignored), 1 = 3D (pitches ignored: volumes are contiguous)
int server_submit(Server *sv, int cls, const uint8_t *in, size_t in_pitch, uint8_t *out,
                  size_t out_pitch, int w, int h, int d, uint64_t *ticket);
int server_wait(Server *sv, uint64_t ticket, int *iterations, double *latency_ms);
int server_poll(Server *sv, uint64_t ticket, int *done);
void server_stats(Server *sv, uint64_t *batches, uint64_t *completed, size_t *peak_bytes);
// `threads` tracer threads trace every served skeleton into its segment list
// (before any submit)
int server_enable_segments(Server *sv, int threads);
int server_wait_segments(Server *sv, uint64_t ticket, int *iterations, double *latency_ms,
                         double *segments_ms, bgt_segments *out);

}  // namespace bgt
