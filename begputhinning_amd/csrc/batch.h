// batch.h -- host interface of the batched engine (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "thin2d.h"
#include "thin3d.h"

namespace bgt {

struct Batch;
Batch *batch_new();
void batch_delete(Batch *b);
int batch_add2d(Batch &b, const uint8_t *in, size_t in_pitch, int w, int h, uint8_t *out,
                size_t out_pitch);
int batch_add3d(Batch &b, const uint8_t *in, int w, int h, int d, uint8_t *out);
int batch_set_config(Batch &b, const Config2D &c2, const Config3D &c3);
int batch_run(Batch &b, hipStream_t s, bool use_graph);
int batch_iterations(Batch &b, int *out);
size_t batch_device_bytes(Batch &b);
int batch_size(const Batch &b);
void batch_set_timing(Batch &b, bool on);
void batch_kernel_time(const Batch &b, double *ms2d, int *n2d, double *ms3d, int *n3d);
// Device bytes a run of the current items needs (no allocation).
int batch_plan_bytes(Batch &b, size_t *bytes);
// Lay later runs out in caller-owned device memory instead of an allocation
// of the batch's own (mem == nullptr restores that).
void batch_set_arena(Batch &b, void *mem, size_t bytes);
// Remove all items (keeps streams, events and the arena).
void batch_clear(Batch &b);
// Time the batch under candidate launch shapes and keep the fastest.
int batch_autotune(Batch &b, hipStream_t s, int reps, Config2D *best2, Config3D *best3,
                   double *best_ms, int *tried);

}  // namespace bgt
