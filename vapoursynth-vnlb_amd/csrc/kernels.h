// Launch entry points of the stage kernels (defined in the .hip files).
#pragma once

#include "common.h"
#include "proposer_types.h"

#include <stddef.h>

namespace vnlbhip {

// guided: per-frame search windows from offsets ([frames][offsets_temporal]
// [height][width][2] int16 (dx, dy) per anchor and scheduled frame), else
// NULL and every window centred on the anchor.
size_t match_shared_bytes(const StageDesc& d, bool guided = false);

// Anchor stacks that contribute to one output frame: stack k's slot[k]
// holds that frame (render_frame_kernel).
constexpr int kMaxStackSet = 32;
struct StackSet {
    const float* stack[kMaxStackSet];
    int slot[kMaxStackSet];
    int count;
};
// One output frame [C][H][W] from the stacks; source: that frame of the
// clip (fallback where no weight landed).
hipError_t launch_render_frame(const StageDesc& d, const StackSet& set, const float* source, float* out,
                               hipStream_t stream);

// The grid fields of an MVTools vector clip (flow.hip).
struct MvShape {
    int block_size_x;
    int block_size_y;
    int overlap_x;
    int overlap_y;
    int block_count_x;
    int block_count_y;
    int pel;
};
// Dense offsets [frames][temporal][H][W][2] from vector clips (either
// vectors pointer may be NULL: no clip for that direction).
hipError_t launch_mvtools_offsets(const StageDesc& d, const MvShape& fw, const int32_t* fw_vectors,
                                  const uint8_t* fw_valid, const MvShape& bw, const int32_t* bw_vectors,
                                  const uint8_t* bw_valid, int16_t* offsets, hipStream_t stream);
hipError_t launch_match(const StageDesc& d, const float* source, const uint32_t* anchors,
                        int anchor_count, int* counts, uint32_t* positions, float* distances,
                        const int16_t* offsets, int offsets_temporal, hipStream_t stream);

// MFMA form of the matcher (match_fast.hip).  Tiles hold 16 anchor indices
// (-1 for empty slots) of one frame, slot 4 * grid row + column (at most 4 x
// 4); max_cols / max_rows bound the union of the tiles' search windows.
bool fast_match_supported(const StageDesc& d, int max_cols, int max_rows);
size_t fast_match_shared_bytes(const StageDesc& d, int max_cols, int max_rows);
size_t fast_match_stride(const StageDesc& d);
// offsets [4] channel offsets, flags [1] (set where an energy is not finite
// or too large for the bounded epilogue), energy [frames][H][W], split
// [frames][match channels][H][W] packed bf16 hi << 16 | lo of the offset
// samples (its buffer holds the row sums on the way), partials
// [kFastPartialDoubles] (8-byte aligned) for the offsets' partial sums.
constexpr int kFastPartialDoubles = 2 * kMaxChannels * 16;
hipError_t launch_fast_energy(const StageDesc& d, const float* source, float* offsets, unsigned* flags,
                              float* energy, uint32_t* split, double* partials, hipStream_t stream);
// Distances of anchor index i land at staged[(i - anchor_base) * stride].
// fused: k-nearest selection inside the matcher (fast_match_fused), results
// straight into counts / positions / distances; otherwise distances are
// staged for launch_fast_select.
bool fast_match_fused(const StageDesc& d, int max_cols, int max_rows);
hipError_t launch_fast_match(const StageDesc& d, int max_cols, int max_rows, const uint32_t* split,
                             const unsigned* flags, const float* energy, const uint32_t* anchors,
                             const int* tiles, int tile_count, int anchor_base, float* staged,
                             bool fused, int* counts, uint32_t* positions, float* distances,
                             hipStream_t stream);
hipError_t launch_fast_select(const StageDesc& d, const float* staged, const uint32_t* anchors,
                              int anchor_count, int* counts, uint32_t* positions, float* distances,
                              hipStream_t stream);

size_t paste_shared_bytes(const StageDesc& d, int max_anchors);
int paste_target_stride(const StageDesc& d);
// Entries of the targets buffer for anchor_count anchors (rows plus the
// padding the walk reads past the last row).
size_t paste_targets_size(const StageDesc& d, int anchor_count);
// frame_offsets holds frame_count + 1 clip-wide anchor indices (the range
// may start past anchor 0); anchor_count is their span.  dep (64-bit, indexed
// by clip-wide anchor) must be zero on entry over that span.  active holds
// at least active_pad + 64 bytes: bytes from active_pad on are scratch for
// the walk's stores past a frame's last anchor (active_pad >= every anchor
// index of the batch).
hipError_t launch_paste(const StageDesc& d, int frame_count, const int* frame_offsets,
                        const int* frame_ids, const int* row_tables, const int* counts,
                        const uint32_t* positions, unsigned char* active, int max_anchors,
                        int anchor_count, unsigned long long* dep, int* target_counts,
                        int* targets, int active_pad, hipStream_t stream);
// Appends the active anchors of [anchor_begin, anchor_begin + anchor_count)
// to group_list at atomically reserved places of *group_count.
hipError_t launch_compact(const unsigned char* active, int anchor_begin, int anchor_count,
                          int* group_list, int* group_count, hipStream_t stream);

size_t prepare_shared_bytes(const StageDesc& d);
size_t filter_shared_bytes(const StageDesc& d);
// prepare, filter and fast filter: persistent grids that take groups from
// a work queue at work_counter (zero on entry), so workgroups that become
// resident late (beside other kernels) take fewer groups instead of
// stretching the launch.  prepare and the fast filter use per-XCD queues
// (kXcdCounterWords ints: each XCD's contiguous eighth of the list first);
// the filter uses work_counter[0].  The model kernels read the group count from
// *group_count_dev (written by the compaction on the device); grids and
// buffers are sized by the capacity, an upper bound (the part's anchors).
hipError_t launch_prepare(const StageDesc& d, const float* noisy, const float* basic,
                          const uint32_t* anchors, const int* counts, const uint32_t* positions,
                          const int* group_list, const int* group_count_dev, int group_capacity,
                          int problem_stride, int* group_flags, int* problem_dims, float* means,
                          float* problems, bool mfma_gram, int* work_counter, hipStream_t stream);
hipError_t launch_jacobi(const StageDesc& d, const int* group_count_dev, int problem_capacity,
                         int problem_stride, const int* group_flags, const int* problem_dims,
                         const float* problems, float* eigenvalues, float* vectors, int* status,
                         hipStream_t stream);
// FAST form: parallel-order Jacobi for problems of at most 16 rows.
bool jacobi_parallel_supported(int problem_stride);
// refine_list / refine_count (zeroed by the caller; may be null): problems
// whose relative eigenvalue gap at the rank cutoff is below refine_gap are
// listed and re-solved in fp64 by a second kernel on the same stream.
hipError_t launch_jacobi_parallel(const StageDesc& d, const int* group_count_dev, int problem_capacity,
                                  int problem_stride, const int* group_flags, const int* problem_dims,
                                  const float* problems, float* eigenvalues, float* vectors,
                                  int* status, float refine_gap, int* refine_list, int* refine_count,
                                  int* refined_total, hipStream_t stream);
hipError_t launch_filter(const StageDesc& d, const float* noisy, const float* basic,
                         const uint32_t* anchors, const int* counts, const uint32_t* positions,
                         const int* group_list, const int* group_count_dev, int group_capacity,
                         int problem_stride, const int* group_flags, const int* problem_dims,
                         const float* means, const float* eigenvalues, const float* vectors,
                         double* accumulators, float* weights_out, int* work_counter,
                         hipStream_t stream);

// MFMA form of the filter (filter_fast.hip): dual-PCA groups of one model,
// 8x8 patches, at most 16 model rows and components, at most 64 samples.
bool fast_filter_supported(const StageDesc& d);
size_t fast_filter_shared_bytes(const StageDesc& d);
hipError_t launch_fast_filter(const StageDesc& d, const float* noisy, const float* basic,
                              const uint32_t* anchors, const int* counts, const uint32_t* positions,
                              const int* group_list, const int* group_count_dev, int group_capacity,
                              int problem_stride, const int* group_flags, const float* means,
                              const float* eigenvalues, const float* vectors, const float* grams,
                              float* accumulators, float* weights_out, int* work_counter,
                              hipStream_t stream);

// One anchor frame's contribution stack (accumulator slots first_frame ..
// first_frame + slots - 1) in the packed vnlb output-frame layout.
hipError_t launch_pack_stack(const StageDesc& d, const double* acc, int first_frame, int slots,
                             float* stack, hipStream_t stream);
hipError_t launch_pack_stack(const StageDesc& d, const float* acc, int first_frame, int slots,
                             float* stack, hipStream_t stream);
hipError_t launch_fill_zero(double* data, size_t count, hipStream_t stream);
hipError_t launch_fill_zero_int(int* data, size_t count, hipStream_t stream);
hipError_t launch_normalize(const StageDesc& d, const double* acc, const float* source, float* out,
                            hipStream_t stream);
// FAST form accumulators (fp32, same layout).
hipError_t launch_normalize(const StageDesc& d, const float* acc, const float* source, float* out,
                            hipStream_t stream);

// Learned candidate proposer (proposer.hip; network in proposer.h).
size_t proposer_workspace_floats(int rows);
hipError_t launch_proposer_init(float* params, uint64_t seed, hipStream_t stream);
hipError_t launch_proposer_refresh(const float* params, __bf16* w1, __bf16* w2, hipStream_t stream);
hipError_t launch_proposer_features(const float* clip, int width, int height, int channels, int patch_time,
                                    const uint32_t* anchors, const uint32_t* candidates, int rows, __bf16* out,
                                    hipStream_t stream);
hipError_t launch_proposer_scores(const __bf16* x, int rows, const ProposerWeights& wt, int precision,
                                  float* scores, hipStream_t stream);
// FP8 inference copies from sample rows: per-input scales (x_scale), W1 with
// them folded in and per-row scales, the hidden scale, W2 per-row scales.
hipError_t launch_proposer_calibrate(const float* params, const __bf16* w1, const __bf16* x, int rows,
                                     float* workspace, uint8_t* w1_fp8, uint8_t* w2_fp8, float* scales,
                                     hipStream_t stream);
hipError_t launch_proposer_gradients(const float* params, const __bf16* w1, const __bf16* w2, const __bf16* x,
                                     const float* targets, int rows, int group, float* workspace, float* grads,
                                     float* loss_groups, hipStream_t stream);
hipError_t launch_proposer_adam(float* params, const float* grads, float* m, float* v, float lr, int step,
                                hipStream_t stream);
// Proposing matcher (match.hip): 8x8 patches, unguided windows; the
// network scores every candidate, exact distances are taken for the top_k
// scored ones (or for the whole window when the score margin between the
// K-th and the (top_k + 1)-th best is below `margin`), then the plain
// selection.  fallbacks[0] counts the anchors searched exhaustively.
bool learned_match_supported(const StageDesc& d);
hipError_t launch_learned_match(const StageDesc& d, const float* source, const uint32_t* anchors, int anchor_count,
                                int* counts, uint32_t* positions, float* distances, const ProposerWeights& wt,
                                int top_k, float margin, int precision, int* fallbacks, hipStream_t stream);

// The fixed per-device scratch array of vnlbhip_stage_scratch (scratch.hip):
// its size, and its device address on `device`, which must be the current
// device (looked up once per device, then cached).
size_t scratch_bytes();
hipError_t scratch_address(int device, void** out);

}  // namespace vnlbhip
