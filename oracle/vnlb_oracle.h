/*
 * VNLB CPU oracle: C interface.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the CPU hot path of
 * yuygfgg/Vapoursynth-VNLB (src/core/core.cpp, src/linalg/linalg.cpp,
 * src/aggregate/aggregate.cpp, src/vs/plugin.cpp) so that the HIP path can be
 * checked against it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Video layout everywhere: contiguous float32 [frames][channels][height][width].
 * Patch positions are linear clip indices: ((frame * height) + y) * width + x.
 */
#ifndef VNLB_ORACLE_H
#define VNLB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field meaning and defaults follow StageParameters (src/core/core.hpp:17-41).
 * sigma, sigma_basic are in [0,1] sample units; tau is a squared distance in
 * the same units (the plugin divides the 8-bit values by 255 and 255^2). */
typedef struct vo_params {
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
} vo_params;

enum {
    VO_STAGE_BASIC = 0,
    VO_STAGE_FINAL = 1,
};

enum {
    /* Accumulate contribution stacks in double and round once per pixel
     * (the order-independent form the HIP path uses).  Without it the stacks
     * are float, accumulated in raster order as the reference does. */
    VO_ACCUM_DOUBLE = 1,
    /* Select matches with the reference's sequential heap / bounded buffer /
     * std::partial_sort procedure instead of the equivalent set rule. */
    VO_SELECT_FAITHFUL = 2,
    /* Run the group model (mean .. projection, weights) in double. */
    VO_MODEL_DOUBLE = 4,
};

/* Whole-clip stage: every anchor frame, then the vnlb.Aggregate
 * normalisation with the noisy clip as fallback.  ref is the Final stage's
 * matching/model clip (ignored for Basic).  Returns 0 or -1 (see
 * vo_last_error). */
int vo_stage(int stage, const vo_params* params, const float* noisy,
             const float* ref, int width, int height, int frames,
             int channels, float* out, int flags, int threads,
             int64_t* groups_out);

/* Per anchor frame diagnostics.  Anchors are the staggered grid in raster
 * order; arrays are sized [capacity_anchors] and [capacity_anchors * cap]
 * where cap = retained group cap.  Optional outputs may be NULL.
 *   counts[a], positions[a*cap + i], distances[a*cap + i]: match lists of
 *     every grid anchor (matching does not depend on the paste mask);
 *   active[a]: 1 if the anchor survives the ordered paste mask;
 *   weights[a*cap + i]: final aggregation weight of each retained patch
 *     (active anchors only);
 *   eigenvalues[a*rank + j]: model eigenvalues (active, modelled groups);
 *   group_flags[a]: bit0 equal group, bit1 flat group, bit2 modelled;
 *   grams[(a*gram_dim + r)*gram_dim + c]: the dual Gram XX^T / B of the
 *     first model (active, modelled, dual groups with B <= gram_dim), in
 *     the model's precision (double with VO_MODEL_DOUBLE). */
int vo_frame_debug(int stage, const vo_params* params, const float* noisy,
                   const float* ref, int width, int height, int frames,
                   int channels, int anchor_frame, int capacity_anchors,
                   int* num_anchors, int* anchor_x, int* anchor_y,
                   int* counts, int* positions, float* distances,
                   unsigned char* active, float* weights, float* eigenvalues,
                   int* group_flags, int flags, int gram_dim, double* grams);

/* Search-centre guidance for the following calls (MVToolsFlowProvider's
 * centres, src/flow/flow.hpp:321-390, as a dense input): offsets
 * [frames][temporal][height][width][2] int16 (dx, dy) per anchor position and
 * scheduled candidate frame (anchor, future ascending, past descending);
 * temporal = search_bwd + search_fwd + 1.  The window of candidate frame ti
 * is centred on the anchor plus the offset clamped to the frame; ti = 0
 * keeps the anchor.  NULL clears it.  The pointer must stay valid while set. */
void vo_set_search_offsets(const int16_t* offsets, int temporal);

/* JacobiEigenOptions::max_sweeps of every eigen solve (src/linalg/linalg.hpp:
 * 153-158); 0 or less restores the default 64. */
void vo_set_max_sweeps(int max_sweeps);

/* The fields of one MVTools vector clip that MVToolsVectorGrid::nearest_vector
 * reads (src/flow/flow.hpp:38-56,176-222). */
typedef struct vo_mv_shape {
    int block_size_x;
    int block_size_y;
    int overlap_x;
    int overlap_y;
    int block_count_x;
    int block_count_y;
    int pel;
} vo_mv_shape;

/* MVToolsFlowProvider::center_for (src/flow/flow.hpp:340-390) over dense
 * vector clips: fw_vectors / bw_vectors [grid_frames][block_count_y]
 * [block_count_x][2] int32 (x, y) in 1/pel pixel units, fw_valid / bw_valid
 * [grid_frames] (0: the frame's grid is unusable, as for an invalid vector
 * frame or a scene change).  The forward clip (mvfw) supplies the grids
 * walked towards earlier frames, the backward clip (mvbw) those walked
 * towards later frames (src/vs/plugin.cpp:700-726).  NULL vectors: no clip
 * for that direction.  Writes the centre to out_xy[0..1]. */
void vo_center_for(const vo_mv_shape* fw, const int32_t* fw_vectors, const unsigned char* fw_valid,
                   const vo_mv_shape* bw, const int32_t* bw_vectors, const unsigned char* bw_valid,
                   int grid_frames, int anchor_x, int anchor_y, int anchor_frame,
                   int candidate_frame, int patch_size, int width, int height, int* out_xy);

/* The dense offset field of vo_set_search_offsets for a clip of `frames`
 * frames: offsets[((f * temporal + ti) * height + y) * width + x] = centre -
 * anchor for anchor (x, y, f) and the ti-th frame of its schedule (0 where
 * ti is past the schedule).  temporal = search_bwd + search_fwd + 1.
 * Returns 0, or -1 on invalid arguments. */
int vo_mvtools_offsets(const vo_params* params, const vo_mv_shape* fw, const int32_t* fw_vectors,
                       const unsigned char* fw_valid, const vo_mv_shape* bw,
                       const int32_t* bw_vectors, const unsigned char* bw_valid, int width,
                       int height, int frames, int16_t* offsets);

/* Contribution stack of one anchor frame in the vnlb.Basic / vnlb.Final
 * output-frame layout (src/vs/plugin.cpp:264-271,625-641; the CUDA
 * pack_kernel, src/cuda/aggregation.cu:508-528): per channel c and slot s =
 * output frame - anchor_frame + search_bwd, a numerator plane then the
 * weight plane, at ((c * slots + s) * 2 + {0, 1}) * height * width, slots =
 * search_bwd + search_fwd + patch_time.  An anchor frame past
 * frames - patch_time gives zeros.  Returns 0 or -1. */
int vo_anchor_stack(int stage, const vo_params* params, const float* noisy,
                    const float* ref, int width, int height, int frames,
                    int channels, int anchor_frame, float* packed, int flags);

/* The same over a window: noisy / ref hold `frames` frames from absolute
 * frame first_frame of a clip of total_frames; anchor_frame is absolute
 * (VideoGeometry::first_frame / total_frames, src/core/video.hpp:19-47). */
int vo_anchor_stack_window(int stage, const vo_params* params, const float* noisy,
                           const float* ref, int width, int height, int frames,
                           int channels, int first_frame, int total_frames,
                           int anchor_frame, float* packed, int flags);

/* One group's aggregation weights from its centred membership rows
 * (similar x dim, row-major) and `rank` eigen pairs (basis rows of dim):
 * reset, `models` times accumulate_aggregation_model, then
 * finalize_aggregation_weights (src/core/core.cpp:235-294,326-412);
 * models = 0 takes the default model of equal groups.  precision 0 runs the
 * model arithmetic in float (the stage's), 1 in double. */
int vo_aggregation_weights(int stage, const vo_params* params, const double* rows, int similar,
                           int dim, const double* eigenvalues, const double* basis, int rank,
                           int basis_similar, int models, int precision, float* weights);

/* prepare_aggregation_window (src/core/core.cpp:296-324): patch_size^2
 * values, row-major. */
int vo_aggregation_window(const vo_params* params, float* window);

/* The staggered anchor grid of one frame (reference_processing_mask_selected,
 * src/core/core.cpp:934-950) as a width x height byte mask; returns the
 * anchor count. */
int vo_grid_mask(const vo_params* params, int width, int height, int total_frames, int anchor_frame,
                 unsigned char* mask);

/* Retained group cap (max(K, ceil(K * cap_factor)) limited by the window). */
int vo_retained_cap(const vo_params* params);

/* Staggered grid anchor count of one anchor frame. */
int vo_grid_count(const vo_params* params, int width, int height, int frames,
                  int anchor_frame);

/* Dense helpers exposed for the reference's linalg known-answer tests.
 * All matrices are row-major double. */
int vo_linalg_mean(const double* samples, int rows, int cols, double* mean);
int vo_linalg_gram_centered(const double* centered, int rows, int cols,
                            double scale, double* gram);
int vo_linalg_covariance_centered(const double* centered, int rows, int cols,
                                  double scale, double* covariance);
/* Top-rank eigenpairs of a symmetric dim x dim matrix by cyclic Jacobi.
 * Returns the sweep count (>= 0) and sets *converged, or -1 on error. */
int vo_linalg_topk_eigen(const double* matrix, int dim, int rank,
                         double* values, double* vectors, int* converged);
int vo_linalg_topk_eigen_f32(const float* matrix, int dim, int rank,
                             float* values, float* vectors, int* converged);
int vo_linalg_dual_basis(const double* centered, int rows, int cols,
                         const double* gram_values,
                         const double* gram_vectors, int rank,
                         double* basis);

const char* vo_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VNLB_ORACLE_H */
