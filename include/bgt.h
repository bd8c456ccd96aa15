/*
 * bgt.h -- C-ABI of the MI355X-native BEGPUThinning thinning path.
 *
 * Plain C: pointers, sizes and status codes; no HIP or torch types (streams are
 * passed as `void *`, a hipStream_t; NULL means the default stream).  Every
 * function returns BGT_OK (0) or a BGT_ERR_* code; bgt_last_error() gives the
 * message of the calling thread's last failure.
 *
 * The drop-in C++ API of the reference (CenterlineExtraction.h,
 * LibThinning2D.h, LibThinning3D.h, ExtractSegments.h) is provided header-only
 * in include/2D and include/3D on top of these entry points.
 */
#ifndef BGT_H
#define BGT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BGT_OK = 0,
    BGT_ERR_INVALID = 1,       /* bad argument (size, pointer, type) */
    BGT_ERR_HIP = 2,           /* HIP runtime error */
    BGT_ERR_NOMEM = 3,         /* device or host allocation failed */
    BGT_ERR_UNSUPPORTED = 4,   /* shape outside what this build handles */
    BGT_ERR_NOT_CONVERGED = 5, /* iteration cap reached */
    BGT_ERR_ADMISSION = 6      /* serving: request does not fit the memory budget */
};

/* Pixel types of the reference's template instantiations
 * (2D/LibThinning2D.cu:134-143, 2D/CenterlineExtraction.cpp:146-155).  `char`
 * is BGT_I8 on targets where char is signed. */
enum {
    BGT_U8 = 0, BGT_I8 = 1, BGT_U16 = 2, BGT_I16 = 3, BGT_U32 = 4,
    BGT_I32 = 5, BGT_F32 = 6, BGT_F64 = 7, BGT_BOOL = 8
};

const char *bgt_last_error(void);
int bgt_version(void);

/* Launch shapes (no reference counterpart: the reference hard-codes them,
 * 2D/LibThinning2D.cu:59-60, 3D/LibThinning3D.cu:50).  2D: waves per
 * workgroup (1-8), register rows per lane (8 or 16), iterations fused per
 * HBM pass (1-16), workgroups per CU of the batch pass kernel (0-32, 0 = the
 * occupancy limit, -1 = automatic, the default: 2 in a batch that also holds
 * volumes, whose kernels take the other slots, else the occupancy limit).  3D: workgroups per CU of the subiteration kernel (0-32,
 * 0 = the occupancy limit), the most iterations queued between two host
 * convergence polls (1-64), and the average active words per volume at or
 * below which the word-list engine hands the remaining iterations over: to
 * the tile engine (engine 1, the default; 16384) or to one launch, one
 * workgroup per volume (engine 0); 0 = never.  Applies to later calls of the
 * calling process and to batches created afterwards. */
int bgt_set_config2d(int waves, int rows, int iters_per_pass, int blocks_per_cu);
int bgt_set_config3d(int blocks_per_cu, int max_chunk, int tail_words);
/* 3D engine (no reference counterpart; results do not depend on it): 0 = the
 * non-empty-word lists, six step launches per iteration, finishing in one
 * launch per batch; 1 (default) = the word lists, then from tail_words on the
 * tiles, one launch per iteration with the whole iteration staged in LDS;
 * 2 = the tiles from the first iteration. */
int bgt_set_engine3d(int engine);
/* 64-bit words per lane of the 2D pass kernel (no reference counterpart;
 * results do not depend on it): 1, 2 or 4, 0 = automatic (the default: the
 * fewest that hold a row in one wave64, i.e. 1 up to 4096 pixels, 2 up to
 * 8192, 4 up to 16384, the widest 2D image).  Images wider than the choice
 * take more; a batch's images all take those of its widest. */
int bgt_set_lane_words2d(int words);
/* Batches of 2D images up to 4096 wide (no reference counterpart; results do
 * not depend on it): 1 = the first band pass reads the caller's bytes
 * itself, the conversion of EncodeBitsKernel
 * (2D/CUDA_DeviceHelperFunctions.cuh:85-130) fused into the first thinning
 * pass; 0 (default, 3% faster on the mixed batch) = a separate pack launch.
 * Batches take it when created or when bgt_batch_use_current_config is
 * called. */
int bgt_set_fused_io2d(int on);
/* Batches of 2D images (no reference counterpart; results do not depend on
 * it): 1 = the whole band-pass chain in one persistent launch, its passes
 * separated by a device-wide barrier (bounded: a wait past about a second
 * fails the run); 0 (default) = a launch per pass, followed by the host
 * through per-pass reports in host-mapped memory.  Replaces the reference's
 * launch + blocking copy + host sum per pass (2D/LibThinning2D.cu:102-112).
 * Batches take it when created or when bgt_batch_use_current_config is
 * called. */
int bgt_set_persistent2d(int on);
/* Test hook (no reference counterpart; the reference aborts on CUDA errors,
 * 3D/LibThinning3D.cu:21-29): the next n device allocations the library makes
 * fail as if the device were out of memory, and the calls that needed them
 * return BGT_ERR_NOMEM.  0 clears it. */
int bgt_debug_fail_device_allocs(int n);

/* ------------------------------------------------------------------------- */
/* Drop-in thinning on host buffers.                                          */
/* ------------------------------------------------------------------------- */

/* Replaces template<class T> void Thinning2D(int width, int height, T *data,
 * T *out) (2D/LibThinning2D.h:13-15, 2D/LibThinning2D.cu:39-130).
 * Foreground iff (unsigned)v > 0 with saturating conversion of floats
 * (2D/CUDA_DeviceHelperFunctions.cuh:101); out holds T(0) / T(1).  Widths
 * 1..16384, heights >= 1 (the reference needs padded widths <= 2048); wider
 * images fail with BGT_ERR_UNSUPPORTED. */
int bgt_thinning2d(int dtype, int width, int height, const void *in, void *out);

/* Replaces template<class T> void Thinning3D(int width, int height, int depth,
 * T *data, T *out) (3D/LibThinning3D.h:13-15, 3D/LibThinning3D.cu:34-90).
 * Foreground iff v > 0 in T (3D/CUDA_DeviceHelperFunctions.cuh:162).  Any
 * dimensions >= 1 (the reference needs multiples of 32, >= 64). */
int bgt_thinning3d(int dtype, int width, int height, int depth, const void *in, void *out);

/* ------------------------------------------------------------------------- */
/* Device-resident thinning (inputs already in HBM).  uint8 masks, nonzero =  */
/* foreground; out receives 0/1.  *iterations (may be NULL) receives the      */
/* number of iterations executed including the final no-change one.           */
/* ------------------------------------------------------------------------- */
int bgt_thin2d_device(const uint8_t *d_in, size_t in_pitch, int width, int height,
                      uint8_t *d_out, size_t out_pitch, void *stream, int *iterations);
int bgt_thin3d_device(const uint8_t *d_in, int width, int height, int depth, uint8_t *d_out,
                      void *stream, int *iterations);

/* bgt_thin2d_device without any allocation (no counterpart in the reference,
 * which allocates on every call, 2D/LibThinning2D.cu:73-75).  For callers
 * that cannot allocate device memory: code under stream capture, tight
 * serving loops, harnesses that forbid hipMalloc.  The engine state lives in
 * one fixed array per GPU of bgt_thin2d_noalloc_max_bytes() bytes (16 MiB,
 * held until the process ends); an image whose state needs more
 * (2 * ceil(width / 64) * height * 8 bytes plus a few KiB) is rejected with
 * BGT_ERR_UNSUPPORTED.  Same results as bgt_thin2d_device.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (from several
 * threads or streams) race and must be serialised by the caller. */
int bgt_thin2d_device_noalloc(const uint8_t *d_in, size_t in_pitch, int width, int height,
                              uint8_t *d_out, size_t out_pitch, void *stream, int *iterations);
size_t bgt_thin2d_noalloc_max_bytes(void);

/* ------------------------------------------------------------------------- */
/* Segment tracing (host), ExtractSegmentsV.                                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nseg;
    int npts;
    int16_t *x, *y, *z;   /* z is NULL for 2D */
    uint32_t *seg_start;  /* nseg + 1 entries, seg_start[nseg] = npts */
    uint8_t *is_end;      /* nseg entries */
} bgt_segments;

/* Replaces template<class T> std::vector<SSegment*> ExtractSegmentsV(T *data,
 * int width, int height) (2D/ExtractSegments.h:16-17, 2D/ExtractSegments.cpp:16-127). */
int bgt_extract_segments2d(int dtype, const void *img, int width, int height, bgt_segments *out);
/* Replaces ExtractSegmentsV(T *data, int width, int height, int depth)
 * (3D/ExtractSegments.h:21-22, 3D/ExtractSegments.cpp:135-251). */
int bgt_extract_segments3d(int dtype, const void *vol, int width, int height, int depth,
                           bgt_segments *out);
/* ExtractSegmentsV over a skeleton in device memory (no reference
 * counterpart; the reference traces host arrays only, 2D/ExtractSegments.cpp:16-127,
 * 3D/ExtractSegments.cpp:135-251).  d_skel holds width x height (x depth)
 * uint8 values, nonzero = set, rows contiguous; it is the top-left corner of
 * a zero-padded padded_width x padded_height (x padded_depth) frame, and the
 * result equals the host tracer's on that frame.  The skeleton points are
 * classified and compacted on the GPU; the host follows only them.  `stream`
 * is a hipStream_t (NULL = default stream), synchronised before returning. */
int bgt_extract_segments2d_device(const void *d_skel, int width, int height, int padded_width,
                                  int padded_height, void *stream, bgt_segments *out);
int bgt_extract_segments3d_device(const void *d_skel, int width, int height, int depth,
                                  int padded_width, int padded_height, int padded_depth,
                                  void *stream, bgt_segments *out);
void bgt_segments_free(bgt_segments *s);

/* ------------------------------------------------------------------------- */
/* Centerline extraction (pad, thin on the GPU, trace, prune, smooth,         */
/* flatten).  Outputs are malloc'd by the library: release with bgt_free.     */
/* ------------------------------------------------------------------------- */

/* Replaces template<class T> unsigned CenterlineExtraction(T *data, int width,
 * int height, int max_segment_length, int smoothing_span, float **out_x,
 * float **out_y, unsigned **out_idx) (2D/CenterlineExtraction.h:25-28,
 * 2D/CenterlineExtraction.cpp:22-142).  A negative smoothing_span is treated
 * as 0 (the reference indexes out of bounds there). */
int bgt_centerline2d(int dtype, const void *img, int width, int height, int max_segment_length,
                     int smoothing_span, float **out_x, float **out_y, uint32_t **out_idx,
                     uint32_t *nsegs);
/* Replaces CenterlineExtraction(T *data, int width, int height, int depth, ...)
 * (3D/CenterlineExctraction.h:26-29, 3D/CenterlineExtraction.cpp:22-155);
 * foreground iff int(v) > 0 (3D/CenterlineExtraction.cpp:44). */
int bgt_centerline3d(int dtype, const void *vol, int width, int height, int depth,
                     int max_segment_length, int smoothing_span, float **out_x, float **out_y,
                     float **out_z, uint32_t **out_idx, uint32_t *nsegs);
void bgt_free(void *p);

/* The MEX gateways' output layout (2D/Matlab2DThinning.cpp:163-175,
 * 3D/Matlab3DThinning.cpp:157-170): segment i of a centerline (out_idx as
 * returned above) becomes an n_i x 2 double matrix with columns [y x], or
 * n_i x 3 with columns [y x z] when z is non-NULL, stored column-major.  The
 * matrices are written one after another into `out`, which holds
 * out_idx[nsegs] * (z ? 3 : 2) doubles; matrix i starts at out_idx[i] * (z ?
 * 3 : 2).  A gateway copies each into its mxCreateDoubleMatrix cell. */
int bgt_mex_columns(const float *x, const float *y, const float *z, const uint32_t *out_idx,
                    uint32_t nsegs, double *out);

/* ------------------------------------------------------------------------- */
/* Batched thinning (no reference counterpart: the reference thins one image  */
/* per call, 2D/LibThinning2D.cu:39-130).  Many device-resident images and    */
/* volumes of uneven size are thinned together; each item converges on its    */
/* own and its result equals the single-item result.                          */
/* ------------------------------------------------------------------------- */
typedef struct bgt_batch_s *bgt_batch;

int bgt_batch_create(bgt_batch *out);
/* Take the launch shapes current at this call (bgt_set_config2d/3d). */
int bgt_batch_use_current_config(bgt_batch b);
int bgt_batch_add2d(bgt_batch b, const uint8_t *d_in, size_t in_pitch, int width, int height,
                    uint8_t *d_out, size_t out_pitch);
int bgt_batch_add3d(bgt_batch b, const uint8_t *d_in, int width, int height, int depth,
                    uint8_t *d_out);
/* Thin every item (pack, passes to convergence, unpack); returns when done.
 * use_graph != 0 replays the pass chunk as a captured hipGraph.  A batch can
 * be run again (inputs are re-read). */
int bgt_batch_run(bgt_batch b, void *stream, int use_graph);
/* Iterations of every item (in the order added), after a run. */
int bgt_batch_iterations(bgt_batch b, int *iterations);
/* Device bytes the batch holds (0 before its first run). */
int bgt_batch_device_bytes(bgt_batch b, size_t *bytes);
/* Per-launch kernel timing with HIP events on the run's stream (direct
 * launches only; a run with timing on does not use the graph).  After a run:
 * total milliseconds and launch counts of the 2D and 3D pass kernels. */
int bgt_batch_set_timing(bgt_batch b, int on);
int bgt_batch_kernel_time(bgt_batch b, double *ms2d, int *launches2d, double *ms3d,
                          int *launches3d);
/* Launch-shape autotuner (no reference counterpart; the reference hard-codes
 * its shapes, 2D/LibThinning2D.cu:59-60, 3D/LibThinning3D.cu:50): runs the
 * batch under candidate 2D shapes (waves, rows, iterations per pass,
 * workgroups per CU) and 3D step grids, `reps` timed runs each after one
 * layout run (their median counts), keeps the fastest for the batch and
 * reports it (any output
 * pointer may be NULL).  Results do not depend on the shape.  The batch's
 * outputs hold a complete result afterwards. */
int bgt_batch_autotune(bgt_batch b, void *stream, int reps, int *waves, int *rows,
                       int *iters_per_pass, int *blocks_per_cu_2d, int *blocks_per_cu_3d,
                       double *best_ms);
int bgt_batch_destroy(bgt_batch b);

/* ------------------------------------------------------------------------- */
/* Serving (no reference counterpart: the reference thins one image per       */
/* blocking call, 2D/LibThinning2D.cu:39-130).  Device-resident requests are  */
/* admitted by the device memory their thinning needs (at most hbm_budget     */
/* bytes in use at once), batched as they arrive and thinned by worker        */
/* threads (workers2d for images, workers3d for volumes; each its own stream  */
/* and batch engine), so requests arriving while others run overlap them on  */
/* the device.  A result equals the single-item result whatever it ran with.  */
/* Inputs must be ready on the device when submitted; outputs are complete    */
/* when bgt_server_wait returns.  The budget bounds the engine memory of the  */
/* requests in flight; each worker also keeps a grow-only arena (the largest  */
/* batch it ran, plus a quarter) and per-thread workspaces until the server   */
/* is destroyed, which are not counted against it.                            */
/* Frames up to 2048 x 2048 join a continuously running 2D band queue at its  */
/* next pass boundary (up to max_items at once, in slots allocated at create  */
/* within half the budget and counted against it); larger frames and volumes */
/* go to the workers.  BGT_SERVE_QUEUE=0 in the environment sends every frame */
/* to the workers.                                                            */
/* ------------------------------------------------------------------------- */
typedef struct bgt_server_s *bgt_server;

/* Launch shapes are those current at this call (bgt_set_config2d/3d). */
int bgt_server_create(bgt_server *out, size_t hbm_budget, int workers2d, int workers3d,
                      int max_items);
/* BGT_ERR_ADMISSION when the request alone needs more than the budget. */
int bgt_server_submit2d(bgt_server s, const uint8_t *d_in, size_t in_pitch, int width, int height,
                        uint8_t *d_out, size_t out_pitch, uint64_t *ticket);
int bgt_server_submit3d(bgt_server s, const uint8_t *d_in, int width, int height, int depth,
                        uint8_t *d_out, uint64_t *ticket);
/* Blocks until the request is done and releases its ticket; latency_ms is
 * submit-to-done on the host's steady clock.  Either pointer may be NULL. */
int bgt_server_wait(bgt_server s, uint64_t ticket, int *iterations, double *latency_ms);
int bgt_server_poll(bgt_server s, uint64_t ticket, int *done);
/* Segment lists of served skeletons (no reference counterpart; the
 * reference traces host arrays after a blocking call, 2D/CenterlineExtraction.cpp:39-42):
 * `threads` tracer threads of the server take each request as its skeleton
 * is done and trace it -- the device pre-pass on the thread's own stream,
 * then the host trace -- the result that of bgt_extract_segments2d/3d_device
 * on the output (a frame of the output's size; 2D outputs need contiguous
 * rows).  Call before the first submit.  bgt_server_wait then also waits for
 * the trace (and drops the list).  Each tracer thread reserves its
 * workspaces (about 80 MB of device memory and 5 MB of pinned host memory,
 * enough for 2^17 points) and traces two tiny skeletons before this returns,
 * so that the first requests do not pay for either. */
int bgt_server_enable_segments(bgt_server s, int threads);
/* Blocks until the request's segment list is ready; hands it over in *out
 * (release with bgt_segments_free).  latency_ms: submit-to-skeleton,
 * segments_ms: submit-to-segments, on the host's steady clock.  Any pointer
 * but `out` may be NULL. */
int bgt_server_wait_segments(bgt_server s, uint64_t ticket, int *iterations, double *latency_ms,
                             double *segments_ms, bgt_segments *out);
int bgt_server_stats(bgt_server s, uint64_t *batches, uint64_t *completed, size_t *peak_bytes);
/* Requests not yet started fail; batches in flight finish first.  Call after
 * every wait has returned. */
int bgt_server_destroy(bgt_server s);

/* ------------------------------------------------------------------------- */
/* Synthetic inputs (seeded, deterministic) for tests and benchmarks.         */
/* ------------------------------------------------------------------------- */
uint64_t bgt_synth_item_seed(int config, int item);
int bgt_synth_polylines(uint8_t *img, int w, int h, uint64_t seed, int n_lines, double p_salt);
int bgt_synth_tree2d(uint8_t *img, int w, int h, uint64_t seed, int depth, double p_salt,
                     double target_fraction);
int bgt_synth_tree3d(uint8_t *vol, int w, int h, int d, uint64_t seed, int depth,
                     int n_ellipsoids, double target_fraction);
int bgt_synth_batch_dims(uint64_t seed, int n2, int lo2, int hi2, int *dims2, int n3, int lo3,
                         int hi3, int *dims3);
int bgt_synth_arrivals(uint64_t seed, int n, double qps, double *gaps, double *mix);

#ifdef __cplusplus
}
#endif
#endif
