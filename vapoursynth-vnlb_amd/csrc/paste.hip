// Ordered paste mask (reference apply_frame_paste_mask, src/core/core.cpp:
// 969-1008, applied in the raster loop of process_anchor_impl :1851-1903).
//
// Matching does not read the mask, so all grid anchors are matched first and
// this kernel replays the serial decision: walking the grid in raster order,
// an anchor whose bit is still set is processed and clears the bits of every
// match in the anchor frame (plus its 4-neighbours away from the border).
// One wavefront owns one anchor frame; the mask holds one bit per grid
// anchor in LDS.
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

__device__ inline int div_step(const StageDesc& d, int x) {
    return static_cast<int>((static_cast<uint32_t>(x) * d.step_magic) >> 18);
}

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
    const int q = div_step(d, x);
    if (x - q * d.step != phase) {
        return -1;
    }
    // Columns phase (or step when phase is 0), +step, ... between 0 and
    // max_x follow the x = 0 anchor.
    return first + q + (phase > 0 ? 1 : 0);
}

// Visits the grid indices an anchor's matches clear (the match position and
// its 4-neighbours away from the border, apply_frame_paste_mask
// core.cpp:980-1008).  Only matches in the anchor frame clear anything.
template <class Fn>
__device__ inline void for_each_clear(const StageDesc& d, const int* s_first, const int* s_phase,
                                      const int* s_count, uint32_t p, int anchor_frame, Fn&& fn) {
    if (position_frame(p) != anchor_frame) {
        return;
    }
    const int P = d.patch;
    const int mx = position_x(p);
    const int my = position_y(p);
    fn(grid_index(d, s_first, s_phase, s_count, mx, my));
    if (my > 2 * P) {
        fn(grid_index(d, s_first, s_phase, s_count, mx, my - 1));
    }
    if (my < d.height - 2 * P) {
        fn(grid_index(d, s_first, s_phase, s_count, mx, my + 1));
    }
    if (mx > 2 * P) {
        fn(grid_index(d, s_first, s_phase, s_count, mx - 1, my));
    }
    if (mx < d.width - 2 * P) {
        fn(grid_index(d, s_first, s_phase, s_count, mx + 1, my));
    }
}

// The serial raster walk, 64 anchors at a time.  Clearing an anchor that is
// already decided has no effect, so within a chunk only clears of later
// anchors of the same chunk matter: each lane records, for the anchors its
// matches clear, an in-chunk predecessor bit.  A scalar scan over the 64
// lanes then makes the sequential decisions exactly, and the survivors'
// clears of anchors beyond the chunk are applied to the mask before the next
// chunk starts.
__global__ __launch_bounds__(kPasteThreads) void paste_kernel(
    StageDesc d, const int* __restrict__ frame_offsets, const int* __restrict__ frame_ids,
    const int* __restrict__ row_tables, const int* __restrict__ counts,
    const uint32_t* __restrict__ positions, unsigned char* __restrict__ active) {
    extern __shared__ unsigned long long s_mem64[];
    unsigned* s_mem = reinterpret_cast<unsigned*>(s_mem64);
    const int slot = blockIdx.x;
    const int first_anchor = frame_offsets[slot];
    const int anchor_count = frame_offsets[slot + 1] - first_anchor;
    const int anchor_frame = frame_ids[slot];
    const int words = (anchor_count + 31) >> 5;
    const int lane = threadIdx.x;
    const int cap = d.cap;
    const int stride = cap + 1;  // padded: lanes read their own list conflict-free

    // 8-byte aligned predecessor masks first, then the mask bits, row tables
    // and the chunk's match lists.
    unsigned long long* s_dep = reinterpret_cast<unsigned long long*>(s_mem);
    unsigned* s_mask = s_mem + 2 * 64;
    int* s_first = reinterpret_cast<int*>(s_mask + words);
    int* s_phase = s_first + d.height;
    int* s_count = s_phase + d.height;
    uint32_t* s_lists = reinterpret_cast<uint32_t*>(s_count + d.height);

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
    __syncthreads();

    for (int base = 0; base < anchor_count; base += 64) {
        const int n = anchor_count - base < 64 ? anchor_count - base : 64;
        const size_t list_base = static_cast<size_t>(first_anchor + base) * cap;
        for (int i = lane; i < n * cap; i += kPasteThreads) {
            const int j = i / cap;
            s_lists[j * stride + (i - j * cap)] = positions[list_base + i];
        }
        const int a = base + lane;
        const int similar = lane < n ? counts[first_anchor + a] : 0;
        s_dep[lane] = 0ull;
        __syncthreads();

        // In-chunk predecessors.
        const uint32_t* list = s_lists + lane * stride;
        const int chunk_end = base + n;
        for (int k = 0; k < similar; ++k) {
            for_each_clear(d, s_first, s_phase, s_count, list[k], anchor_frame, [&](int t) {
                if (t > a && t < chunk_end) {
                    atomicOr(&s_dep[t - base], 1ull << lane);
                }
            });
        }
        __syncthreads();

        // Sequential decisions of the chunk (uniform scalar loop).
        const bool start = lane < n && ((s_mask[a >> 5] >> (a & 31)) & 1u) != 0;
        const unsigned long long start_bits = __ballot(start);
        const unsigned long long my_dep = s_dep[lane];
        const unsigned dep_lo = static_cast<unsigned>(my_dep);
        const unsigned dep_hi = static_cast<unsigned>(my_dep >> 32);
        unsigned long long active_bits = 0ull;
        for (int i = 0; i < n; ++i) {
            const unsigned long long dep_i =
                (static_cast<unsigned long long>(
                     static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(dep_hi), i)))
                 << 32) |
                static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(dep_lo), i));
            if (((start_bits >> i) & 1ull) != 0 && (dep_i & active_bits) == 0) {
                active_bits |= 1ull << i;
            }
        }
        const bool is_active = ((active_bits >> lane) & 1ull) != 0;
        if (lane < n) {
            active[first_anchor + a] = is_active ? 1 : 0;
        }

        // Survivors clear anchors beyond the chunk.
        if (is_active) {
            for (int k = 0; k < similar; ++k) {
                for_each_clear(d, s_first, s_phase, s_count, list[k], anchor_frame, [&](int t) {
                    if (t >= chunk_end) {
                        atomicAnd(&s_mask[t >> 5], ~(1u << (t & 31)));
                    }
                });
            }
        }
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
    const size_t words = static_cast<size_t>((max_anchors + 31) / 32);
    const size_t list_words = static_cast<size_t>(64) * (d.cap + 1);
    return 64 * 8 + words * 4 + static_cast<size_t>(3) * d.height * 4 + list_words * 4;
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
    hipLaunchKernelGGL(paste_kernel, dim3(static_cast<unsigned>(frame_count)), dim3(kPasteThreads),
                       shared, stream, d, frame_offsets, frame_ids, row_tables, counts, positions,
                       active);
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
