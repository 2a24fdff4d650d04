// Search geometry of an anchor, shared by the kernels and the host.
#pragma once

#include "common.h"

namespace vnlbhip {

// compute_temporal_range (reference src/core/core.cpp:873-896) on a window
// of the clip: `anchor`, `low` and `high` are local frame indices (absolute
// index - first_frame); the range is shifted into the clip's patch origins
// [0, total_frames - Pt] in absolute terms, as VideoGeometry::first_frame /
// source_frames make the reference do (src/core/video.hpp:19-47).
__host__ __device__ inline void temporal_range(const StageDesc& d, int anchor,
                                               int& low, int& high) {
    const int min_origin = -d.first_frame;
    const int max_origin = d.total_frames - d.patch_time - d.first_frame;
    const int start = anchor - d.search_bwd;
    const int end = anchor + d.search_fwd;
    const int shift = (start < min_origin ? start - min_origin : 0) +
                      (end - max_origin > 0 ? end - max_origin : 0);
    low = start - shift > min_origin ? start - shift : min_origin;
    high = end - shift < max_origin ? end - shift : max_origin;
}

// compute_spatial_range (reference src/core/core.cpp:910-932).
__host__ __device__ inline void spatial_range(const StageDesc& d, int center,
                                              int axis, int& low, int& high) {
    const int max_origin = axis - d.patch;
    const int half = (d.search_window - 1) / 2;
    const int start = center - half;
    const int end = center + half;
    const int shift = (start < 0 ? start : 0) + (end - max_origin > 0 ? end - max_origin : 0);
    low = start - shift > 0 ? start - shift : 0;
    high = end - shift < max_origin ? end - shift : max_origin;
}

// schedule_temporal_range (reference src/core/core.cpp:898-908): the anchor
// frame, the future frames ascending, then the past frames descending.
__host__ __device__ inline int scheduled_frame(int anchor, int low, int high,
                                               int index) {
    if (index == 0) {
        return anchor;
    }
    const int future = high - anchor;
    if (index <= future) {
        return anchor + index;
    }
    return anchor - (index - future);
}

}  // namespace vnlbhip
