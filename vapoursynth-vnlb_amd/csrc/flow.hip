// Search centres from MVTools vector clips, as the dense offset field the
// guided matcher reads.
//
// MVToolsFlowProvider::center_for (src/flow/flow.hpp:340-390) walks from the
// anchor frame to the candidate frame one frame at a time; each usable grid
// on the way moves the centre by the vector of the block nearest to the
// centre's patch middle (rounded from 1/pel to whole pixels, halves away
// from zero) and clamps it to the frame.  Walking to frame a + k repeats the
// walk to a + k - 1 and adds one step, so one thread per anchor position
// extends the walk frame by frame and writes every scheduled frame's offset:
// O(temporal) steps per anchor instead of O(temporal^2).
//
// Layout: vectors [frames][block_count_y][block_count_x][2] int32 (x, y);
// valid [frames] uint8; offsets [frames][temporal][height][width][2] int16
// with temporal = search_bwd + search_fwd + 1 and index ti the candidate's
// place in the anchor's schedule (anchor, later frames ascending, earlier
// frames descending); entries past the schedule are 0.
#include "geometry.h"
#include "kernels.h"

namespace vnlbhip {
namespace {

// MVToolsVectorGrid::nearest_block_index (src/flow/flow.hpp:208-222).
__device__ inline int nearest_block(int coordinate, int first_center, int step, int count) {
    if (count <= 1) {
        return 0;
    }
    const int shifted = coordinate - first_center;
    const int index = shifted >= 0 ? (shifted + step / 2) / step : -((-shifted + step / 2) / step);
    return min(max(index, 0), count - 1);
}

// detail::round_subpixel_to_pixel (src/flow/flow.hpp:185-193).
__device__ inline int pel_to_pixel(int value, int pel) {
    if (pel <= 1) {
        return value;
    }
    return value >= 0 ? (value + pel / 2) / pel : -((-value + pel / 2) / pel);
}

__device__ inline void follow_vector(const MvShape& s, const int32_t* __restrict__ vectors, int frame,
                                     int patch_center, int width, int height, int& x, int& y) {
    const int step_x = max(1, s.block_size_x - max(0, s.overlap_x));
    const int step_y = max(1, s.block_size_y - max(0, s.overlap_y));
    const int bx = nearest_block(x + patch_center, s.block_size_x / 2, step_x, s.block_count_x);
    const int by = nearest_block(y + patch_center, s.block_size_y / 2, step_y, s.block_count_y);
    const size_t at = ((static_cast<size_t>(frame) * s.block_count_y + by) * s.block_count_x + bx) * 2;
    x = min(max(x + pel_to_pixel(vectors[at], s.pel), 0), width - 1);
    y = min(max(y + pel_to_pixel(vectors[at + 1], s.pel), 0), height - 1);
}

__global__ __launch_bounds__(256) void mvtools_offsets_kernel(
    StageDesc d, MvShape fw, const int32_t* __restrict__ fw_vectors,
    const uint8_t* __restrict__ fw_valid, MvShape bw, const int32_t* __restrict__ bw_vectors,
    const uint8_t* __restrict__ bw_valid, short2* __restrict__ offsets) {
    const size_t plane = static_cast<size_t>(d.width) * d.height;
    const size_t count = plane * d.frames;
    const int temporal = d.search_bwd + d.search_fwd + 1;
    const int patch_center = d.patch / 2;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        const int f = static_cast<int>(i / plane);
        const size_t pixel = i - static_cast<size_t>(f) * plane;
        const int y = static_cast<int>(pixel / d.width);
        const int x = static_cast<int>(pixel - static_cast<size_t>(y) * d.width);
        short2* out = offsets + static_cast<size_t>(f) * temporal * plane + pixel;
        int low = 0;
        int high = 0;
        temporal_range(d, f, low, high);
        // Frames past frames - Pt are never anchors, and their range ends
        // below them: their schedule is the frame itself, then the earlier
        // frames, cut at temporal entries.
        const int future = max(high - f, 0);
        const int past = min(f - low, temporal - 1 - future);
        out[0] = make_short2(0, 0);
        // Later frames: the backward clip's grids of frames f .. f + k - 1.
        int cx = x;
        int cy = y;
        for (int k = 1; k <= future; ++k) {
            const int frame = f + k - 1;
            if (bw_vectors != nullptr && bw_valid[frame] != 0) {
                follow_vector(bw, bw_vectors, frame, patch_center, d.width, d.height, cx, cy);
            }
            out[static_cast<size_t>(k) * plane] = make_short2(static_cast<short>(cx - x), static_cast<short>(cy - y));
        }
        // Earlier frames: the forward clip's grids of frames f .. f - k + 1.
        cx = x;
        cy = y;
        for (int k = 1; k <= past; ++k) {
            const int frame = f - k + 1;
            if (fw_vectors != nullptr && fw_valid[frame] != 0) {
                follow_vector(fw, fw_vectors, frame, patch_center, d.width, d.height, cx, cy);
            }
            out[static_cast<size_t>(future + k) * plane] =
                make_short2(static_cast<short>(cx - x), static_cast<short>(cy - y));
        }
        for (int ti = future + past + 1; ti < temporal; ++ti) {
            out[static_cast<size_t>(ti) * plane] = make_short2(0, 0);
        }
    }
}

}  // namespace

hipError_t launch_mvtools_offsets(const StageDesc& d, const MvShape& fw, const int32_t* fw_vectors,
                                  const uint8_t* fw_valid, const MvShape& bw, const int32_t* bw_vectors,
                                  const uint8_t* bw_valid, int16_t* offsets, hipStream_t stream) {
    const size_t count = static_cast<size_t>(d.width) * d.height * d.frames;
    const size_t blocks = (count + 255) / 256;
    const unsigned grid = static_cast<unsigned>(blocks < 8192 ? blocks : 8192);
    hipLaunchKernelGGL(mvtools_offsets_kernel, dim3(grid), dim3(256), 0, stream, d, fw, fw_vectors,
                       fw_valid, bw, bw_vectors, bw_valid, reinterpret_cast<short2*>(offsets));
    return hipGetLastError();
}

}  // namespace vnlbhip
