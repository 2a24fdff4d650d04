// Shared device/host definitions for the vnlbhip stage kernels.
//
// Video layout: contiguous float32 [frames][channels][height][width].
// Patch positions are packed as x | y << 12 | frame << 24 (width, height up to
// 4096, at most 256 frames in the device window of one stage call; the clip
// itself may be longer, see StageDesc::first_frame).
//
// Floating-point order: every decision-relevant quantity follows the 16-lane
// order the CPU restatement fixes for the reference's Highway kernels
// (src/linalg/kernels_highway.hpp, src/core/distance_highway.hpp): lane-wise
// fused multiply-add chains, a halving-tree ReduceSum, separately rounded
// scalar tails.  The library is compiled with -ffp-contract=off so that only
// the explicit fmaf calls fuse.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace vnlbhip {

constexpr int kLanes = 16;
constexpr int kMaxPatch = 16;
constexpr int kMaxDim = 1024;          // C * Pt * P^2
constexpr int kMaxModelRows = 128;     // B = min(S, ceil(K * model_cap))
constexpr int kMaxCap = 1024;          // retained patches per group
constexpr int kMaxRank = 64;
constexpr int kMaxCandidates = 32768;  // window^2 * temporal slots (the matchers' LDS budget binds first)
constexpr int kMaxChannels = 4;
constexpr int kMaxTemporal = 64;
constexpr int kMaxFrames = 256;
constexpr int kMaxSide = 4096;
// XCDs of the GPU and the layout of a per-XCD work queue: one counter per XCD,
// 16 ints (a 64-byte line) apart.
constexpr int kXcds = 8;
constexpr int kXcdCounterStride = 16;
constexpr int kXcdCounterWords = kXcds * kXcdCounterStride;

enum Stage : int {
    kBasic = 0,
    kFinal = 1,
};

// Everything a stage kernel needs to know, passed by value.
struct StageDesc {
    int stage;
    int width;
    int height;
    int frames;         // frames held in device memory (the local window)
    int channels;
    int first_frame;    // absolute clip index of local frame 0
    int total_frames;   // frames of the whole clip (>= first_frame + frames)
    int patch;          // P
    int patch_time;     // Pt
    int patch_area;     // P^2
    int patch_dim;      // Pt * P^2
    int dim;            // C * patch_dim (sample row length)
    int coupled;        // couple_channels && channels > 1
    int models;         // coupled ? 1 : channels
    int model_dim;      // coupled ? dim : patch_dim
    int use_dual;       // model_dim >= 64 && model rows < model_dim
    int search_window;
    int search_bwd;     // clamped to total_frames - Pt (same ranges, no overflow)
    int search_fwd;
    int similar;        // K
    int cap;            // retained cap (output stride)
    int soft_cap;       // max(K, ceil(K * cap_factor)), INT_MAX if disabled
    int model_cap;      // max(K, ceil(K * model_cap_factor))
    int rank;
    int slots;          // bwd + fwd + Pt
    int step;           // anchor grid step
    uint32_t step_magic;  // floor(2^18 / step) + 1: x / step == (x * magic) >> 18 for x < 4096
    int flat_areas;
    int weights_enabled;
    int windowed;       // weight_gamma != 0
    int fused_width;    // coupled && !windowed ? 16 * floor(P / 16) : 0
    int max_rows;       // max model rows over groups (<= model_cap, <= cap)
    int jacobi_dim;     // n of the eigen problems (dual: model rows, else model_dim)
    int max_sweeps;     // Jacobi sweep limit (64 in the reference)
    int row_chunk;      // retained rows staged in LDS at a time (>= max_rows)
    float tau;
    float sigma;
    float sigma_basic;
    float beta;
    float variance_threshold;
    float gamma;
    float weight_alpha;
    float weight_beta;
    float weight_gamma;
    float weight_epsilon;
    float membership_noise_floor;
    float window[kMaxPatch * kMaxPatch];
};

__host__ __device__ inline uint32_t pack_position(int x, int y, int frame) {
    return static_cast<uint32_t>(x) | (static_cast<uint32_t>(y) << 12) |
           (static_cast<uint32_t>(frame) << 24);
}
__host__ __device__ inline int position_x(uint32_t p) {
    return static_cast<int>(p & 0xFFFu);
}
__host__ __device__ inline int position_y(uint32_t p) {
    return static_cast<int>((p >> 12) & 0xFFFu);
}
__host__ __device__ inline int position_frame(uint32_t p) {
    return static_cast<int>(p >> 24);
}

// Per-group flag bits written by the prepare kernel.
enum GroupFlag : int {
    kFlagEqual = 1,
    kFlagFlat = 2,
    kFlagModelled = 4,
};

}  // namespace vnlbhip
