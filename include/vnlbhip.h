/*
 * vnlbhip: MI355X (gfx950) VNLB hot path behind a plain C ABI.
 *
 * Replaces, for one GPU, the stage pipeline the reference drives from its
 * VapourSynth plugins: the CPU stage process_basic_anchor_no_flow /
 * process_final_anchor_no_flow (src/core/core.hpp:107-116) followed by
 * vnlb.Aggregate (src/vs/plugin.cpp:966-1045), as the CUDA plugin returns it
 * in one call (render_frame, src/vs/cuda_plugin.cpp:1635-1864, with
 * StagePipeline::enqueue src/cuda/pipeline.hpp:82-137).
 *
 * Video layout: contiguous float32 [frames][channels][height][width],
 * channels 1 (GrayS) or 3 (YUV444PS).  All functions return 0 on success and
 * a negative code on failure; vnlbhip_last_error() then holds the message.
 * No exception crosses this boundary.
 */
#ifndef VNLBHIP_H
#define VNLBHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VNLBHIP_ABI_VERSION 7

enum {
    VNLBHIP_STAGE_BASIC = 0,
    VNLBHIP_STAGE_FINAL = 1,
};

enum {
    VNLBHIP_OK = 0,
    VNLBHIP_ERROR_INVALID = -1,  /* parameters, geometry or limits */
    VNLBHIP_ERROR_DEVICE = -2,   /* HIP runtime failure */
    VNLBHIP_ERROR_MODEL = -3,    /* a group could not cover the rank */
};

/* Stage parameters with the meaning and units of StageParameters
 * (src/core/core.hpp:17-41): sigma and sigma_basic in [0,1] sample units,
 * tau as a squared distance in those units (the plugin front end divides
 * 8-bit values by 255 and 255^2, src/vs/plugin.cpp:291-312). */
typedef struct vnlbhip_params {
    float sigma;
    int patch_size;
    int patch_time;
    int search_window;
    int search_bwd;
    int search_fwd;
    int similar;
    int rank;
    float similar_cap_factor;
    float model_cap_factor;
    float beta;
    float tau;
    float variance_threshold;
    float sigma_basic;
    float gamma;
    float weight_alpha;
    float weight_beta;
    float weight_gamma;
    float weight_epsilon;
    float membership_noise_floor;
    int proc_step;
    int flat_areas;
    int couple_channels;
} vnlbhip_params;

typedef struct vnlbhip_stats {
    int64_t anchors;              /* grid anchors matched */
    int64_t groups;               /* anchors that survived the paste mask */
    int64_t unconverged_groups;   /* eigen problems that hit the sweep limit */
    int32_t batches;              /* anchor-frame batches of the last call */
    int32_t parts;                /* pipeline parts per batch of the last call */
    int64_t retained;             /* retained samples summed over the groups */
    int64_t refined;              /* FAST form: eigen problems re-solved in fp64 */
    int64_t graph_replays;        /* stage graphs replayed by the context so far */
    int64_t reallocations;        /* device buffers the last call (re)allocated */
    int64_t proposed;             /* anchors matched by the learned proposer */
    int64_t exhaustive_fallbacks; /* of those, anchors whose score margin was too low */
    int32_t tile_rows;            /* FAST form: anchor tile of the MFMA matcher, grid rows x */
    int32_t tile_cols;            /* anchors per row (0 x 0 when it did not run) */
} vnlbhip_stats;

typedef struct vnlbhip_context vnlbhip_context;

/* Defaults of vnlb.Basic (stage 0) or vnlb.Final (stage 1): StageParameters
 * defaults (src/core/core.hpp:17-41) with the Final overrides of
 * parse_stage_parameters (src/vs/plugin.cpp:331-339); sigma is left 0. */
int vnlbhip_default_params(int stage, vnlbhip_params* out);

/* Validation of validate_stage_configuration (src/core/core.cpp:1997-2024)
 * plus this build's kernel limits.  Needs no GPU. */
int vnlbhip_validate(int stage, const vnlbhip_params* params, int width,
                     int height, int frames, int channels, char* message,
                     size_t message_size);

/* One context per stream/thread (the reference's reserve/enqueue objects are
 * likewise not thread-safe, src/cuda/pipeline.hpp:82-126).  Contexts share
 * nothing: each owns its device buffers, plans, graphs and streams. */
int vnlbhip_create(int device, vnlbhip_context** out);
void vnlbhip_destroy(vnlbhip_context* context);
const char* vnlbhip_last_error(const vnlbhip_context* context);

/* Sizes every device buffer a vnlbhip_stage call of this stage, parameter
 * set, geometry and the context's current form needs, launching nothing
 * (StagePipeline::reserve, src/cuda/pipeline.hpp:82-107).  Later stage,
 * stack and debug calls of that shape or smaller then neither allocate nor
 * synchronise the device for buffers (vnlbhip_stats.reallocations stays 0);
 * without it the first call of a larger shape allocates and waits for the
 * device.  Blocks until done. */
int vnlbhip_reserve(vnlbhip_context* context, int stage, const vnlbhip_params* params, int width,
                    int height, int frames, int channels);

/* Whole-clip stage on device memory: every anchor frame's matching, paste
 * mask, group model and weighted aggregation, then normalisation with the
 * noisy clip as fallback, written to out.  ref is the Final stage's guide
 * clip (the Basic output), NULL for Basic.  stream is a hipStream_t or NULL.
 * Blocks until the stage has finished (group counts are read back). */
int vnlbhip_stage(vnlbhip_context* context, int stage,
                  const vnlbhip_params* params, const float* noisy,
                  const float* ref, float* out, int width, int height,
                  int frames, int channels, void* stream);

/* Size of the per-device scratch array of vnlbhip_stage_scratch. */
#define VNLBHIP_SCRATCH_BYTES (128u * 1024u * 1024u)

/* vnlbhip_stage without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as vnlbhip_stage; every internal buffer
 * (match lists, paste tables, models, accumulators) is carved from one
 * fixed scratch array of VNLBHIP_SCRATCH_BYTES per GPU, which each GPU the
 * library is set up on holds until the process ends (its device address is
 * taken once per device when a context is created).  The caller passes
 * inputs and output.  A stage that needs more than the array fails with
 * VNLBHIP_ERROR_INVALID before any kernel runs.  The call allocates no
 * device or pinned memory, creates no stream or event, captures no graph,
 * keeps no profiling timings and runs no autotuning; the context's own
 * buffers are left as they were.  Every call on one GPU shares its array,
 * which keeps the last call's contents (each call clears what it needs):
 * calls that overlap on one GPU -- from two threads or two contexts --
 * race, so serialise them.  Blocks until the stage has finished. */
int vnlbhip_stage_scratch(vnlbhip_context* context, int stage,
                          const vnlbhip_params* params, const float* noisy,
                          const float* ref, float* out, int width, int height,
                          int frames, int channels, void* stream);

/* One anchor frame's contribution stack on device memory, as a vnlb.Basic /
 * vnlb.Final output frame holds it (src/vs/plugin.cpp:264-271,625-641; the
 * CUDA plugin's pack_kernel, src/cuda/aggregation.cu:508-528, for
 * vnlb.Aggregate, src/vs/plugin.cpp:966-1045): slots = search_bwd +
 * search_fwd + patch_time; for channel c and slot s (output frame
 * anchor_frame - search_bwd + s) the numerator plane, then the weight plane,
 * at stack + ((c * slots + s) * 2 + {0, 1}) * height * width.  Slots outside
 * the clip, and every slot of an anchor frame past frames - patch_time, are
 * zero.  Summing the stacks of the anchor frames that reach an output frame
 * and dividing gives vnlbhip_stage's output.  Blocks until done. */
int vnlbhip_anchor_stack(vnlbhip_context* context, int stage,
                         const vnlbhip_params* params, const float* noisy,
                         const float* ref, int width, int height, int frames,
                         int channels, int anchor_frame, float* stack,
                         void* stream);

/* The same over a window of a longer clip, as the reference's frame
 * requests run it (render_stage_frame, src/vs/plugin.cpp:616-758, with
 * make_geometry(vi, range.first, local_frames) and VideoGeometry's
 * first_frame / total_frames, src/core/video.hpp:19-47): noisy / ref hold
 * `frames` frames starting at absolute frame first_frame of a clip of
 * total_frames; anchor_frame is absolute.  Grid phases and temporal ranges
 * follow the absolute indices and the clip length.  The window must hold
 * the anchor frame's input frames (input_frame_range_for_anchor,
 * src/core/core.cpp:2026-2034), else VNLBHIP_ERROR_INVALID; an anchor past
 * total_frames - patch_time gives a zero stack.  At most 256 frames per
 * window; the clip may be of any length. */
int vnlbhip_anchor_stack_window(vnlbhip_context* context, int stage,
                                const vnlbhip_params* params, const float* noisy,
                                const float* ref, int width, int height, int frames,
                                int channels, int first_frame, int total_frames,
                                int anchor_frame, float* stack, void* stream);

/* Same with host buffers: copies in, runs, copies out. */
int vnlbhip_stage_host(vnlbhip_context* context, int stage,
                       const vnlbhip_params* params, const float* noisy,
                       const float* ref, float* out, int width, int height,
                       int frames, int channels);

/* Diagnostics for one anchor frame (device input pointers, host outputs):
 * the grid anchors in raster order and, per anchor, the match list
 * (positions as linear clip indices ((frame*height)+y)*width+x, distances),
 * the paste-mask decision, aggregation weights (active anchors), model
 * eigenvalues (first model, active modelled anchors), group flags
 * (bit0 equal, bit1 flat, bit2 modelled) and the dual Gram XX^T / B the
 * eigen solver received (first model of active modelled dual groups with
 * B <= gram_dim, at grams + (a * gram_dim + r) * gram_dim + c).  Per-anchor
 * arrays have capacity_anchors entries, match arrays capacity_anchors * cap
 * entries with cap from vnlbhip_retained_cap; optional outputs may be
 * NULL. */
int vnlbhip_debug_frame(vnlbhip_context* context, int stage,
                        const vnlbhip_params* params, const float* noisy,
                        const float* ref, int width, int height, int frames,
                        int channels, int anchor_frame, int capacity_anchors,
                        int* num_anchors, int* anchor_x, int* anchor_y,
                        int* counts, int* positions, float* distances,
                        unsigned char* active, float* weights,
                        float* eigenvalues, int* group_flags, int gram_dim,
                        float* grams);

int vnlbhip_retained_cap(const vnlbhip_params* params);

/* The staggered anchor grid of one anchor frame (absolute index) of a clip
 * of total_frames, in raster order (reference_processing_mask_selected /
 * initialize_frame_processing_mask, src/core/core.cpp:934-967): writes up
 * to capacity positions to xs / ys (either may be NULL) and returns the
 * anchor count.  Needs no GPU. */
int vnlbhip_anchor_grid(const vnlbhip_params* params, int width, int height, int total_frames,
                        int anchor_frame, int capacity, int* xs, int* ys);

/* The aggregation window a(u) of prepare_aggregation_window
 * (src/core/core.cpp:296-324), patch_size^2 floats, row-major.  Needs no
 * GPU. */
int vnlbhip_aggregation_window(const vnlbhip_params* params, float* window);
int vnlbhip_get_stats(const vnlbhip_context* context, vnlbhip_stats* out);

/* Per-kernel device time, measured with hipEvents recorded on the stage's
 * stream around every launch while profiling is enabled (no counterpart in
 * the reference, whose benchmarks time with cudaEvents,
 * tools/benchmark_cuda_block_match.cpp:512-526).  Enabling resets the
 * totals. */
enum {
    VNLBHIP_KERNEL_MATCH = 0,
    VNLBHIP_KERNEL_PASTE = 1,
    VNLBHIP_KERNEL_PREPARE = 2,
    VNLBHIP_KERNEL_JACOBI = 3,
    VNLBHIP_KERNEL_FILTER = 4,
    VNLBHIP_KERNEL_NORMALIZE = 5,
    VNLBHIP_KERNEL_SELECT = 6,  /* FAST form: k-nearest selection (PLAIN selects inside MATCH) */
    VNLBHIP_KERNEL_ENERGY = 7,  /* FAST form: patch energies, once per stage */
    VNLBHIP_KERNEL_COUNT = 8,
};

typedef struct vnlbhip_kernel_times {
    double milliseconds[VNLBHIP_KERNEL_COUNT];
    int64_t launches[VNLBHIP_KERNEL_COUNT];
} vnlbhip_kernel_times;

int vnlbhip_set_profiling(vnlbhip_context* context, int enable);
int vnlbhip_get_kernel_times(const vnlbhip_context* context, vnlbhip_kernel_times* out);

/* Arithmetic form of the stages run on a context (no counterpart in the
 * reference).  PLAIN, the default, sums every distance and model product in
 * fp32 in the CPU path's order.  FAST takes the candidate distances from
 * bf16x3 products on the matrix cores (8x8 patches, windows up to 25 wide;
 * other shapes keep the PLAIN matcher), solves group
 * models of at most 16 rows with a parallel-order fp32 Jacobi, and filters
 * dual-PCA groups (one model, at most 16 model rows, at most 64 retained
 * samples, 8x8 patches) through their bf16x3 cross-Gram; the selection
 * rule, group flags, means and Grams are the PLAIN ones.  Returns
 * VNLBHIP_ERROR_INVALID for an unknown form. */
enum {
    VNLBHIP_FORM_PLAIN = 0,
    VNLBHIP_FORM_FAST = 1,
};

int vnlbhip_set_form(vnlbhip_context* context, int form);

/* FAST form: group models whose eigenvalue gap at the rank cutoff (the
 * rank-th minus the next eigenvalue, both of the fp32 solve) is below
 * relative_gap times the largest eigenvalue are re-solved in fp64 (no
 * counterpart in the reference, whose solver runs at one precision).
 * Default 1e-3; 0 disables; negative or non-finite values are rejected with
 * VNLBHIP_ERROR_INVALID.  Only models with rank < rows have a cutoff. */
int vnlbhip_set_eigen_refine(vnlbhip_context* context, float relative_gap);

/* 1: vnlbhip_stage captures a stage's launch sequence (every kernel, both
 * streams) into a hipGraph after an eager run and replays it on later calls
 * with the same stage, form, pipeline parts, geometry, parameters and
 * noisy / ref / out addresses; a replay is skipped (eager run and a new
 * capture) after any device buffer of the context was reallocated.  Only
 * stages whose anchor frames fit one batch are captured.  Off while kernel
 * profiling is on.  0 (default): eager launches.  No counterpart in the
 * reference.  Values other than 0 and 1 are rejected. */
int vnlbhip_set_graph(vnlbhip_context* context, int enable);

/* 1: vnlbhip_stage times the stage once per pipeline part count (1..4
 * parts of consecutive anchor frames per batch; a part's serial paste scan
 * runs beside the next part's matching) on the first call for a stage,
 * form, geometry and parameter set, and keeps the fastest for later calls;
 * the first call then runs the stage six times.  0 (default): two parts.
 * vnlbhip_stats.parts reports the count used.  No counterpart in the
 * reference.  Values other than 0 and 1 are rejected. */
int vnlbhip_set_autotune(vnlbhip_context* context, int enable);

/* FAST form: anchor tiles of the MFMA matcher, `rows` grid rows x `cols`
 * anchors per row (1 to 4 each; default 4 x 4), for every call the
 * autotuner does not set (with autotuning on, the first vnlbhip_stage call
 * of a stage, form, geometry and parameter set also times every shape of
 * 4x4, 3x4, 2x4, 4x3, 4x2 and 1x4 whose window union fits the kernel and
 * keeps the fastest).  A shape whose union does not fit falls back to
 * 2 x 4.  vnlbhip_stats.tile_rows / tile_cols report the shape used.  No
 * counterpart in the reference. */
int vnlbhip_set_tile_shape(vnlbhip_context* context, int rows, int cols);

/* Per-anchor search centres from a motion field (MVToolsFlowProvider's
 * centres, src/flow/flow.hpp:321-390, as a dense input; the reference
 * centres each scheduled frame's window on them, src/core/core.cpp:1328-1340).
 * offsets is a device array of int16 pairs (dx, dy), laid out
 * [frames][temporal][height][width][2] for the clip of the next
 * vnlbhip_stage / vnlbhip_anchor_stack / vnlbhip_debug_frame calls, with
 * temporal = search_bwd + search_fwd + 1; index ti is the candidate frame's
 * place in the anchor's schedule (the anchor frame, then later frames
 * ascending, then earlier frames descending).  For ti > 0 the window is
 * centred on (x + dx, y + dy) clamped to the frame and then shifted inside
 * it like an unguided window; ti = 0 keeps the anchor.  Guided stages run
 * the plain matcher and filter in either form.  NULL (default) clears the
 * field.  The array must stay valid while it is set; frames, height and
 * width are the shape of the clip it was made for.  Non-positive sizes with
 * a non-NULL array are rejected with VNLBHIP_ERROR_INVALID, and a stage
 * whose search window or clip does not match the field fails with
 * VNLBHIP_ERROR_INVALID before any kernel reads it. */
int vnlbhip_set_search_offsets(vnlbhip_context* context, const int16_t* offsets, int temporal,
                               int frames, int height, int width);

/* Per-output-frame rendering, the frame-request model of the VapourSynth
 * plugins (vnlb.Basic / vnlb.Final produce one anchor frame's contribution
 * stack per request, src/vs/plugin.cpp:616-758; vnlb.Aggregate sums the
 * stacks of the anchors around an output frame, :966-1045; the CUDA plugin
 * keeps device frames in a FrameCache, src/cuda/frame_cache.hpp:73-177).
 * A cache holds `capacity` anchor stacks in device memory (each
 * channels * (search_bwd + search_fwd + patch_time) * 2 * width * height
 * floats); rendering output frame n computes the missing stacks of anchors
 * [n - search_fwd - patch_time + 1, n + search_bwd] (replacing the least
 * recently used stacks of other anchors) and aggregates them into out
 * ([channels][height][width], device).  `frames` at creation is the
 * clip's length.  vnlbhip_frame_cache_render takes the whole clip on the
 * device, as vnlbhip_stage does (at most 256 frames); calls with other clip
 * pointers start from an empty cache.  vnlbhip_frame_cache_render_window
 * takes a window of the clip instead: noisy / ref hold window_frames frames
 * from absolute frame first_frame, which must include frame n and the
 * input frames of every anchor stack the call computes (a clip of any
 * length; cached stacks are keyed by frame number, so a frame's content
 * must not change while its stacks are cached -- VapourSynth frames do not;
 * vnlbhip_frame_cache_invalidate drops every stack).  A change of the
 * context's form, eigen re-solve gap, selection placement or search offsets
 * also starts from an empty cache.  capacity must be at least search_bwd +
 * search_fwd + patch_time (at most 32 stacks contribute to a frame).
 * Rendering every frame in order computes each anchor stack once.  Renders
 * of one cache are serialised by a mutex inside it, but its context must
 * not be used by another thread meanwhile.  A render returns before its
 * kernel ends; a later render on another stream that replaces a stack waits
 * for the renders that read it.  The cache belongs to its context (destroy
 * it first). */
typedef struct vnlbhip_frame_cache vnlbhip_frame_cache;

typedef struct vnlbhip_frame_cache_stats {
    int64_t hits;       /* stacks found in the cache */
    int64_t misses;     /* stacks computed */
    int64_t evictions;  /* cached stacks replaced */
} vnlbhip_frame_cache_stats;

int vnlbhip_frame_cache_create(vnlbhip_context* context, int stage, const vnlbhip_params* params,
                               int width, int height, int frames, int channels, int capacity,
                               vnlbhip_frame_cache** out);
int vnlbhip_frame_cache_render(vnlbhip_frame_cache* cache, int frame, const float* noisy,
                               const float* ref, float* out, void* stream);
int vnlbhip_frame_cache_render_window(vnlbhip_frame_cache* cache, int frame, const float* noisy,
                                      const float* ref, int first_frame, int window_frames,
                                      float* out, void* stream);
int vnlbhip_frame_cache_invalidate(vnlbhip_frame_cache* cache);
int vnlbhip_frame_cache_get_stats(const vnlbhip_frame_cache* cache, vnlbhip_frame_cache_stats* out);
void vnlbhip_frame_cache_destroy(vnlbhip_frame_cache* cache);

/* The grid of one MVTools vector clip: the MVTools_MVAnalysisData fields
 * that MVToolsVectorGrid::nearest_vector reads (src/flow/flow.hpp:38-56,
 * 176-222; parsed by parse_mvtools_analysis, :196-258). */
typedef struct vnlbhip_mv_shape {
    int32_t block_size_x;
    int32_t block_size_y;
    int32_t overlap_x;
    int32_t overlap_y;
    int32_t block_count_x;
    int32_t block_count_y;
    int32_t pel;
} vnlbhip_mv_shape;

/* Fills the offset field of vnlbhip_set_search_offsets from MVTools vector
 * clips, as MVToolsFlowProvider::center_for chains them
 * (src/flow/flow.hpp:340-390): towards an earlier candidate frame the grids
 * of the forward clip (mvfw) of the frames from the anchor frame down to one
 * after the candidate, towards a later one the grids of the backward clip
 * (mvbw) of the frames from the anchor frame up to one before the
 * candidate (the provider as src/vs/plugin.cpp:700-726 builds it); each
 * usable grid moves the centre by the vector of the block nearest to
 * centre + patch_size / 2, rounded from 1/pel pixel units, and clamps it to
 * the frame.  Device arrays: *_vectors [frames][block_count_y]
 * [block_count_x][2] int32 (x, y) of the finest level, *_valid [frames]
 * uint8 (0: the frame's grid is not used -- an invalid vector frame or a
 * scene change, src/vs/plugin.cpp:476-485,516-519); NULL vectors: no clip
 * for that direction.  offsets: device, 4-byte aligned,
 * [frames][search_bwd + search_fwd + 1][height][width][2] int16, written on
 * stream; the schedule follows params (patch_size, patch_time, search_bwd,
 * search_fwd).  Invalid arguments: VNLBHIP_ERROR_INVALID. */
int vnlbhip_mvtools_offsets(vnlbhip_context* context, const vnlbhip_params* params,
                            const vnlbhip_mv_shape* fw_shape, const int32_t* fw_vectors,
                            const uint8_t* fw_valid, const vnlbhip_mv_shape* bw_shape,
                            const int32_t* bw_vectors, const uint8_t* bw_valid, int16_t* offsets,
                            int width, int height, int frames, void* stream);

/* FAST form: 1 (default) runs the k-nearest selection inside the MFMA
 * matcher where the retained cap allows (at most 64); 0 stages every
 * candidate distance in device memory and selects in a second kernel with
 * the PLAIN form's selection (no counterpart in the reference).  Both give
 * the same match lists; 0 exists to check that.  Values other than 0 and 1
 * are rejected with VNLBHIP_ERROR_INVALID. */
int vnlbhip_set_fused_selection(vnlbhip_context* context, int enable);

/* Eigen solver limits: max_sweeps (JacobiEigenOptions::max_sweeps,
 * src/linalg/linalg.hpp:154-158, default 64; 1 .. 10000) for every Jacobi
 * solver of the context, and what an unconverged problem does: 0 (default)
 * its result is used and vnlbhip_stats.unconverged_groups counts it, as the
 * CPU path does (topk_symmetric_eigen_jacobi's status is not checked by
 * src/core/core.cpp); 1 the stage fails with VNLBHIP_ERROR_MODEL, as the
 * CUDA plugin's solver status check does (src/vs/cuda_plugin.cpp:1831-1846). */
int vnlbhip_set_jacobi(vnlbhip_context* context, int max_sweeps, int fail_unconverged);

/* ---- Learned candidate proposal (no counterpart in the reference) ----
 *
 * The reference scans every candidate of the search window
 * (find_similar_patches, src/core/core.cpp:1245-1362; the CUDA matcher
 * src/cuda/block_match.cu:592-912).  A proposer is a small MLP (input 128
 * features of a patch pair: 2x2 box means of the two 8x8 patches, their
 * absolute differences, the offset, the box-mean SSD; two hidden layers of
 * 256; one score) trained on the GPU on the plain matcher's exact k nearest.
 * Set on a context, it scores every candidate of each window and only the
 * top_k scored candidates get exact fp32 distances (each equal to the
 * exhaustive scan's bit for bit); when the score margin between the K-th
 * and the (top_k + 1)-th best is below `margin`, that anchor's window is
 * scanned exhaustively.  The selection rules then run unchanged over the
 * proposals.  Parameters are one flat fp32 vector of
 * VNLBHIP_PROPOSER_PARAMS values: W1 [256][128], b1 [256], W2 [256][256],
 * b2 [256], w3 [256], b3. */
#define VNLBHIP_PROPOSER_INPUT 128
#define VNLBHIP_PROPOSER_HIDDEN 256
#define VNLBHIP_PROPOSER_PARAMS 99073
#define VNLBHIP_PRECISION_BF16 0
#define VNLBHIP_PRECISION_FP8 1

typedef struct vnlbhip_proposer vnlbhip_proposer;

/* A proposer on `device` with He-uniform weights from `seed` (bf16 copies
 * made); errors are reported through the return code only. */
int vnlbhip_proposer_create(int device, uint64_t seed, vnlbhip_proposer** out);
void vnlbhip_proposer_destroy(vnlbhip_proposer* proposer);
const char* vnlbhip_proposer_last_error(const vnlbhip_proposer* proposer);
/* Host copies of the fp32 parameters (VNLBHIP_PROPOSER_PARAMS floats);
 * set_params also refreshes the bf16 copies and drops fp8 calibration.
 * Both synchronise the device. */
int vnlbhip_proposer_get_params(const vnlbhip_proposer* proposer, float* params);
int vnlbhip_proposer_set_params(vnlbhip_proposer* proposer, const float* params);
/* Input rows of (anchor, candidate) pairs of a device clip: anchors and
 * candidates are packed positions (x | y << 12 | frame << 24, frames of the
 * clip, 8x8 patches of patch_time frames inside it), out: bf16
 * [rows][VNLBHIP_PROPOSER_INPUT] (device, 16-byte aligned). */
int vnlbhip_proposer_features(vnlbhip_proposer* proposer, const float* clip, int width, int height,
                              int frames, int channels, int patch_time, const uint32_t* anchors,
                              const uint32_t* candidates, int rows, void* out, void* stream);
/* Forward, listwise loss and gradients of rows of inputs x (bf16, device)
 * in groups of `group` consecutive rows with targets (device, summing to 1
 * per group): grads (device, VNLBHIP_PROPOSER_PARAMS) and loss_groups
 * (device, rows / group).  rows must be a multiple of group. */
int vnlbhip_proposer_gradients(vnlbhip_proposer* proposer, const void* x, const float* targets, int rows,
                               int group, float* grads, float* loss_groups, void* stream);
/* The same, then one Adam step (beta 0.9 / 0.999, eps 1e-8) at
 * learning_rate on the proposer's parameters and its bf16 copies. */
int vnlbhip_proposer_train_step(vnlbhip_proposer* proposer, const void* x, const float* targets, int rows,
                                int group, float learning_rate, float* loss_groups, void* stream);
/* Scores of rows of inputs with the matcher's inference code (precision
 * VNLBHIP_PRECISION_BF16, or _FP8 after calibration). */
int vnlbhip_proposer_scores(vnlbhip_proposer* proposer, const void* x, int rows, int precision, float* scores,
                            void* stream);
/* FP8 (e4m3) inference copies and their scales from sample input rows
 * (per-input and per-output-row scales, a static hidden scale). */
int vnlbhip_proposer_calibrate(vnlbhip_proposer* proposer, const void* x, int rows, void* stream);
/* Matching of the context's stages through the proposer (NULL clears):
 * 8x8 patches, patch_time 1 or 2, unguided windows; top_k in [K, 255].
 * The proposer must outlive its use by the context. */
int vnlbhip_set_proposer(vnlbhip_context* context, const vnlbhip_proposer* proposer, int top_k, float margin,
                         int precision);
/* The match lists of one anchor frame (matching only, no paste or model),
 * with the context's form and proposer: anchors (packed positions), counts
 * and the retained matches [anchor][cap] (cap = vnlbhip_retained_cap;
 * positions packed, distances) into device buffers of `capacity` anchors;
 * *anchor_count (host) receives the frame's anchor count.  Blocks. */
int vnlbhip_match_lists(vnlbhip_context* context, int stage, const vnlbhip_params* params, const float* noisy,
                        const float* ref, int width, int height, int frames, int channels, int anchor_frame,
                        int capacity, int* anchor_count, uint32_t* anchors, int* counts, uint32_t* positions,
                        float* distances, void* stream);
/* Search windows of anchors (host arrays; compute_temporal_range and
 * compute_spatial_range, src/core/core.cpp:873-932): per anchor x_low,
 * y_low, nx, ny, t_low, t_high (absolute frames).  Needs no GPU. */
int vnlbhip_search_windows(const vnlbhip_params* params, int width, int height, int total_frames, int count,
                           const uint32_t* anchors, int32_t* windows);

#ifdef __cplusplus
}
#endif

#endif /* VNLBHIP_H */
