// Ordered paste mask (reference apply_frame_paste_mask, src/core/core.cpp:
// 969-1008, applied in the raster loop of process_anchor_impl :1851-1903).
//
// Matching does not read the mask, so all grid anchors are matched first and
// this kernel replays the serial decision: walking the grid in raster order,
// an anchor whose bit is still set is processed and clears the bits of every
// match in the anchor frame (plus its 4-neighbours away from the border).
// One wavefront owns one anchor frame; the mask holds one bit per grid
// anchor in LDS, and match lists stream through a double-buffered LDS stage
// so global latency overlaps the serial walk.
#include "common.h"
#include "device_util.h"

namespace vnlbhip {
namespace {

constexpr int kPasteThreads = 64;

struct GridRows {
    const int* first;   // [height] grid index of the row's first anchor, -1 off grid
    const int* phase;   // [height] column phase (mod step)
    const int* count;   // [height] anchors in the row
};

// Grid index of (x, y) in the anchor frame's staggered grid, or -1.
__device__ inline int grid_index(const StageDesc& d, const int* s_first,
                                 const int* s_phase, const int* s_count, int x,
                                 int y) {
    if (x < 0 || x >= d.width || y < 0 || y >= d.height) {
        return -1;
    }
    const int first = s_first[y];
    if (first < 0) {
        return -1;
    }
    const int max_x = d.width - d.patch;
    if (x > max_x) {
        return -1;
    }
    if (x == 0) {
        return first;
    }
    if (x == max_x) {
        return first + s_count[y] - 1;
    }
    const int phase = s_phase[y];
    if (x % d.step != phase) {
        return -1;
    }
    const int first_column = phase > 0 ? phase : d.step;
    return first + 1 + (x - first_column) / d.step;
}

__device__ inline void clear_bit(unsigned* mask, int index) {
    if (index >= 0) {
        atomicAnd(&mask[index >> 5], ~(1u << (index & 31)));
    }
}

template <int kRounds>  // ceil(cap / 64)
__global__ __launch_bounds__(kPasteThreads) void paste_kernel(
    StageDesc d, const int* __restrict__ frame_offsets,
    const int* __restrict__ frame_ids, const int* __restrict__ row_tables,
    const int* __restrict__ counts, const uint32_t* __restrict__ positions,
    unsigned char* __restrict__ active) {
    extern __shared__ unsigned s_mem[];
    constexpr int kBatch = 32 / kRounds;

    const int slot = blockIdx.x;
    const int first_anchor = frame_offsets[slot];
    const int anchor_count = frame_offsets[slot + 1] - first_anchor;
    const int anchor_frame = frame_ids[slot];
    const int words = (anchor_count + 31) >> 5;
    const int lane = threadIdx.x;

    unsigned* s_mask = s_mem;
    int* s_first = reinterpret_cast<int*>(s_mask + words);
    int* s_phase = s_first + d.height;
    int* s_count = s_phase + d.height;
    uint32_t* s_stage = reinterpret_cast<uint32_t*>(s_count + d.height);
    int* s_stage_count = reinterpret_cast<int*>(s_stage + 2 * kBatch * 64 * kRounds);

    const int* rows = row_tables + static_cast<size_t>(slot) * 3 * d.height;
    for (int y = lane; y < d.height; y += kPasteThreads) {
        s_first[y] = rows[y];
        s_phase[y] = rows[d.height + y];
        s_count[y] = rows[2 * d.height + y];
    }
    for (int w = lane; w < words; w += kPasteThreads) {
        const int remaining = anchor_count - w * 32;
        s_mask[w] = remaining >= 32 ? 0xFFFFFFFFu : ((1u << remaining) - 1u);
    }

    const int cap = d.cap;
    const int P = d.patch;
    uint32_t staged[kBatch * kRounds];
    int staged_count = 0;
    const auto load_batch = [&](int batch_begin) {
        const int batch_n = anchor_count - batch_begin < kBatch ? anchor_count - batch_begin : kBatch;
        staged_count = lane < batch_n ? counts[first_anchor + batch_begin + lane] : 0;
#pragma unroll
        for (int i = 0; i < kBatch; ++i) {
#pragma unroll
            for (int r = 0; r < kRounds; ++r) {
                const int entry = r * 64 + lane;
                uint32_t value = 0;
                if (i < batch_n && entry < cap) {
                    value = positions[static_cast<size_t>(first_anchor + batch_begin + i) * cap +
                                      entry];
                }
                staged[i * kRounds + r] = value;
            }
        }
    };

    load_batch(0);
    int buffer = 0;
    for (int batch_begin = 0; batch_begin < anchor_count; batch_begin += kBatch) {
        uint32_t* stage = s_stage + buffer * kBatch * 64 * kRounds;
        int* stage_count = s_stage_count + buffer * kBatch;
#pragma unroll
        for (int i = 0; i < kBatch; ++i) {
#pragma unroll
            for (int r = 0; r < kRounds; ++r) {
                stage[(i * kRounds + r) * 64 + lane] = staged[i * kRounds + r];
            }
        }
        if (lane < kBatch) {
            stage_count[lane] = staged_count;
        }
        __syncthreads();
        if (batch_begin + kBatch < anchor_count) {
            load_batch(batch_begin + kBatch);
        }
        const int batch_n = anchor_count - batch_begin < kBatch ? anchor_count - batch_begin : kBatch;
        for (int i = 0; i < batch_n; ++i) {
            const int a = batch_begin + i;
            const bool set = ((s_mask[a >> 5] >> (a & 31)) & 1u) != 0;
            if (lane == 0) {
                active[first_anchor + a] = set ? 1 : 0;
            }
            if (set) {
                const int similar = stage_count[i];
#pragma unroll
                for (int r = 0; r < kRounds; ++r) {
                    const int entry = r * 64 + lane;
                    if (entry < similar) {
                        const uint32_t p = stage[(i * kRounds + r) * 64 + lane];
                        if (position_frame(p) == anchor_frame) {
                            const int mx = position_x(p);
                            const int my = position_y(p);
                            clear_bit(s_mask, grid_index(d, s_first, s_phase, s_count, mx, my));
                            if (my > 2 * P) {
                                clear_bit(s_mask, grid_index(d, s_first, s_phase, s_count, mx, my - 1));
                            }
                            if (my < d.height - 2 * P) {
                                clear_bit(s_mask, grid_index(d, s_first, s_phase, s_count, mx, my + 1));
                            }
                            if (mx > 2 * P) {
                                clear_bit(s_mask, grid_index(d, s_first, s_phase, s_count, mx - 1, my));
                            }
                            if (mx < d.width - 2 * P) {
                                clear_bit(s_mask, grid_index(d, s_first, s_phase, s_count, mx + 1, my));
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        buffer ^= 1;
        __syncthreads();
    }
}

// Appends the active anchors of a batch to a group list (order is not
// significant: every later stage is per group and accumulation is
// order-independent).
__global__ __launch_bounds__(256) void compact_kernel(
    const unsigned char* __restrict__ active, int anchor_count,
    int* __restrict__ group_list, int* __restrict__ group_count) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const bool take = a < anchor_count && active[a] != 0;
    const unsigned long long ballot = __ballot(take);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == 0 && ballot != 0) {
        base = atomicAdd(group_count, static_cast<int>(__popcll(ballot)));
    }
    base = __shfl(base, 0, 64);
    if (take) {
        const unsigned long long below = lane == 0 ? 0ull : (ballot & ((1ull << lane) - 1ull));
        group_list[base + static_cast<int>(__popcll(below))] = a;
    }
}

}  // namespace

size_t paste_shared_bytes(const StageDesc& d, int max_anchors) {
    const int rounds = (d.cap + 63) / 64;
    const int batch = 32 / (rounds <= 1 ? 1 : (rounds <= 2 ? 2 : 4));
    const size_t words = static_cast<size_t>((max_anchors + 31) / 32);
    return words * 4 + static_cast<size_t>(3) * d.height * 4 +
           static_cast<size_t>(2) * batch * 64 * (rounds <= 1 ? 1 : (rounds <= 2 ? 2 : 4)) * 4 +
           static_cast<size_t>(2) * batch * 4;
}

hipError_t launch_paste(const StageDesc& d, int frame_count,
                        const int* frame_offsets, const int* frame_ids,
                        const int* row_tables, const int* counts,
                        const uint32_t* positions, unsigned char* active,
                        int max_anchors, hipStream_t stream) {
    if (frame_count <= 0) {
        return hipSuccess;
    }
    const size_t shared = paste_shared_bytes(d, max_anchors);
    const dim3 grid(static_cast<unsigned>(frame_count));
    const dim3 block(kPasteThreads);
    if (d.cap <= 64) {
        hipLaunchKernelGGL(paste_kernel<1>, grid, block, shared, stream, d, frame_offsets,
                           frame_ids, row_tables, counts, positions, active);
    } else if (d.cap <= 128) {
        hipLaunchKernelGGL(paste_kernel<2>, grid, block, shared, stream, d, frame_offsets,
                           frame_ids, row_tables, counts, positions, active);
    } else {
        hipLaunchKernelGGL(paste_kernel<4>, grid, block, shared, stream, d, frame_offsets,
                           frame_ids, row_tables, counts, positions, active);
    }
    return hipGetLastError();
}

hipError_t launch_compact(const unsigned char* active, int anchor_count,
                          int* group_list, int* group_count, hipStream_t stream) {
    if (anchor_count <= 0) {
        return hipSuccess;
    }
    const dim3 grid(static_cast<unsigned>((anchor_count + 255) / 256));
    hipLaunchKernelGGL(compact_kernel, grid, dim3(256), 0, stream, active, anchor_count,
                       group_list, group_count);
    return hipGetLastError();
}

}  // namespace vnlbhip
