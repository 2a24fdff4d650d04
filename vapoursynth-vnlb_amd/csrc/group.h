// Helpers shared by the per-group kernels (prepare, jacobi, filter).
#pragma once

#include "common.h"

namespace vnlbhip {

constexpr int kModelThreads = 256;
constexpr int kTeams = kModelThreads / kLanes;  // 16-lane teams per block

struct GroupRefs {
    const uint32_t* anchors;
    const int* counts;
    const uint32_t* positions;
    const int* group_list;
};

// Offset of the patch origin in a [F][C][H][W] video.
__device__ inline size_t patch_base(const StageDesc& d, int x, int y, int frame) {
    return (static_cast<size_t>(frame) * d.channels * d.height + y) * d.width + x;
}

__device__ inline size_t patch_base(const StageDesc& d, uint32_t position) {
    return patch_base(d, position_x(position), position_y(position), position_frame(position));
}

// Transcendentals are evaluated in double and rounded once to float, on the
// CPU restatement and here, so both sides agree to the last bit.
__device__ inline float log_via_double(float x) {
    return static_cast<float>(log(static_cast<double>(x)));
}
__device__ inline float log1p_via_double(float x) {
    return static_cast<float>(log1p(static_cast<double>(x)));
}
__device__ inline float exp_via_double(float x) {
    return static_cast<float>(exp(static_cast<double>(x)));
}

// Rows that enter the model of a group of `similar` samples
// (basis_similar, core.cpp:642-643 with model_similar_cap :159-172).
__device__ inline int model_rows_of(const StageDesc& d, int similar) {
    return similar < d.model_cap ? similar : d.model_cap;
}

// Dual PCA when the model dimension is at least 64 and exceeds the rows
// (use_dual_pca, core.cpp:656-658).
__device__ inline bool group_uses_dual(const StageDesc& d, int rows) {
    return d.model_dim >= 64 && rows < d.model_dim;
}

// Slot layout of a model row: lane l of a 16-lane team owns the elements
// e = 16 k + l (k < chunks: vector part; k == chunks, l < tail: scalar
// tail).  Rows are stored per lane with a pitch of 4-float units, odd in
// count so that the 16 lanes' 128-bit reads hit distinct banks.
__host__ __device__ inline int slot_count(int model_dim) {
    return model_dim / kLanes + (model_dim % kLanes != 0 ? 1 : 0);
}

// Register slice length the kernels are instantiated for.
__host__ __device__ inline int template_slots(int model_dim) {
    const int slots = slot_count(model_dim);
    return slots <= 4 ? 4 : slots <= 8 ? 8 : slots <= 16 ? 16 : slots <= 24 ? 24 : slots <= 32 ? 32 : 64;
}

__host__ __device__ inline int slot_pitch(int model_dim) {
    const int slots = template_slots(model_dim);
    return slots + (((slots / 4) % 2 == 0) ? 4 : 0);
}

// Scatter metadata of a model-row element: t | c << 8 | fused | valid.
constexpr int kMetaValid = 1 << 17;
constexpr int kMetaFused = 1 << 16;

struct Accumulators {
    double* data;  // [frames][channels + 1][height][width]
};

// One weighted contribution (write_sample_contributions_coupled*,
// core.cpp:1635-1834): value and weight into the accumulator cell, the
// weight only once per pixel (channel 0).  Samples whose frame falls outside
// the anchor's slot range are dropped.
__device__ inline void scatter(const StageDesc& d, double* __restrict__ cell_base, int slot_base,
                               int meta, int acc_off, float win, float value, float w) {
    if ((meta & kMetaValid) == 0) {
        return;
    }
    const int t = meta & 0xFF;
    const int c = (meta >> 8) & 0xFF;
    const int slot = slot_base + t;
    if (slot < 0 || slot >= d.slots) {
        return;
    }
    const float cw = w * win;
    double addend;
    if ((meta & kMetaFused) != 0) {
        addend = static_cast<double>(w) * static_cast<double>(value);
    } else {
        const float product = cw * value;
        addend = static_cast<double>(product);
    }
    double* cell = cell_base + acc_off;
    unsafeAtomicAdd(cell, addend);
    if (c == 0) {
        const size_t plane = static_cast<size_t>(d.width) * d.height;
        unsafeAtomicAdd(cell + static_cast<size_t>(d.channels) * plane, static_cast<double>(cw));
    }
}

}  // namespace vnlbhip
