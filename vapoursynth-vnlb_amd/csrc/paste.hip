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

// Clear targets of every anchor, computed in parallel (they do not depend
// on the mask): a clear of a later anchor in the same 64-anchor chunk sets a
// predecessor bit in that anchor's mask word; clears beyond the chunk are
// listed.  Clears of earlier anchors cannot change any decision and are
// dropped.
__global__ __launch_bounds__(256) void paste_targets_kernel(
    StageDesc d, int frame_count, const int* __restrict__ frame_offsets,
    const int* __restrict__ frame_ids, const int* __restrict__ row_tables,
    const int* __restrict__ counts, const uint32_t* __restrict__ positions,
    unsigned long long* __restrict__ dep, int* __restrict__ target_counts,
    int* __restrict__ targets, int target_stride) {
    // frame_offsets are clip-wide anchor indices; the range may start past 0.
    const int a = frame_offsets[0] + blockIdx.x * blockDim.x + threadIdx.x;
    const int total = frame_offsets[frame_count];
    if (a >= total) {
        return;
    }
    int slot = 0;
    while (slot + 1 < frame_count && frame_offsets[slot + 1] <= a) {
        ++slot;
    }
    const int first = frame_offsets[slot];
    const int count = frame_offsets[slot + 1] - first;
    const int local = a - first;
    const int chunk_end = min((local & ~63) + 64, count);
    const int anchor_frame = frame_ids[slot];
    const int* rows = row_tables + static_cast<size_t>(slot) * 3 * d.height;
    const int* r_first = rows;
    const int* r_phase = rows + d.height;
    const int* r_count = rows + 2 * d.height;
    const int similar = counts[a];
    const uint32_t* list = positions + static_cast<size_t>(a) * d.cap;
    int* out = targets + static_cast<size_t>(a) * target_stride;
    int n = 0;
    for (int k = 0; k < similar; ++k) {
        for_each_clear(d, r_first, r_phase, r_count, list[k], anchor_frame, [&](int t) {
            if (t <= local) {
                return;
            }
            if (t < chunk_end) {
                atomicOr(&dep[first + t], 1ull << (local & 63));
            } else if (n < target_stride) {
                out[n++] = t;
            }
        });
    }
    // An anchor without matches (every candidate distance NaN) is not a
    // group and clears nothing (process_anchor_impl skips it,
    // core.cpp:1851-1903): marked with a negative target count.
    target_counts[a] = similar > 0 ? n : -1;
}

// The serial raster walk, 64 anchors at a time: wave-wide rounds over the
// chunk's predecessor masks make the sequential decisions exactly, then the
// survivors clear their targets beyond the chunk.  The inputs of the next
// kPasteDepth - 1 chunks are in flight while the current one is decided.
constexpr int kPrefetchTargets = 24;
constexpr int kPasteDepth = 4;
static_assert(kPasteThreads == 64, "the paste walk is one wavefront per frame");

__global__ __launch_bounds__(kPasteThreads) void paste_kernel(
    StageDesc d, const int* __restrict__ frame_offsets,
    const unsigned long long* __restrict__ dep, const int* __restrict__ target_counts,
    const int* __restrict__ targets, int target_stride, unsigned char* __restrict__ active,
    int active_pad) {
    extern __shared__ unsigned s_mask[];
    const int slot = blockIdx.x;
    const int first = frame_offsets[slot];
    const int count = frame_offsets[slot + 1] - first;
    const int words = (count + 31) >> 5;
    const int lane = threadIdx.x;
    for (int w = lane; w < words; w += kPasteThreads) {
        const int remaining = count - w * 32;
        s_mask[w] = remaining >= 32 ? 0xFFFFFFFFu : ((1u << remaining) - 1u);
    }
    __syncthreads();

    struct Chunk {
        unsigned long long dep;
        int n;
        int t[kPrefetchTargets];
    };
    // Every load and store of the walk issues unconditionally (indices
    // clamped into the frame, rows read past their targets into the next row
    // or the buffer's padding, stores past the frame into active's padding),
    // so the wait for a chunk counts exactly the operations issued after it.
    const auto load = [&](int base, Chunk& c) {
        const int a = base + lane;
        const bool valid = a < count;
        const int at = first + (valid ? a : count - 1);
        // Raw values: lanes past the frame are never read (decide uses lanes
        // below n only), and a select here would wait for the load.
        c.dep = dep[at];
        c.n = target_counts[at];
        const int4* list = reinterpret_cast<const int4*>(targets + static_cast<size_t>(at) * target_stride);
#pragma unroll
        for (int h = 0; h < kPrefetchTargets / 4; ++h) {
            const int4 v = list[h];
            c.t[4 * h + 0] = v.x;
            c.t[4 * h + 1] = v.y;
            c.t[4 * h + 2] = v.z;
            c.t[4 * h + 3] = v.w;
        }
    };
    // A ring of kPasteDepth chunks in flight: chunk c is decided from
    // ring[c % kPasteDepth], which is then refilled with chunk
    // c + kPasteDepth.  The ring is indexed by unrolled constants (registers,
    // no copies), and the block is one wavefront, so the end of a chunk
    // needs only a wave barrier: LDS operations of one wave complete in
    // order, and no fence drains the loads in flight.
    if (count <= 0) {
        return;
    }
    Chunk ring[kPasteDepth];
#pragma unroll
    for (int k = 0; k < kPasteDepth; ++k) {
        load(64 * k, ring[k]);
    }
    const auto decide = [&](const Chunk& c, int base) {
        const int n = min(max(count - base, 0), 64);
        const int a = base + lane;
        const bool start = lane < n && c.n >= 0 && ((s_mask[a >> 5] >> (a & 31)) & 1u) != 0;
        const unsigned long long start_bits = __ballot(start);
        // The serial rule -- a candidate (bit still set) is processed unless
        // an earlier processed anchor of the chunk clears it -- resolved in
        // rounds across the wave: a candidate is dropped once an earlier
        // processed one clears it, and processed once every earlier
        // candidate clearing it is decided and none was processed.  Each
        // round decides at least the lowest undecided candidate, so this
        // equals the raster-order walk; rounds follow the longest chain of
        // clears inside the chunk.
        const unsigned long long preds = c.dep & start_bits;
        unsigned long long active_bits = 0ull;
        unsigned long long undecided = start_bits;
        while (undecided != 0ull) {
            const bool mine = ((undecided >> lane) & 1ull) != 0;
            const bool drop = mine && (preds & active_bits) != 0;
            const bool take = mine && !drop && (preds & undecided) == 0;
            const unsigned long long taken = __ballot(take);
            const unsigned long long dropped = __ballot(drop);
            active_bits |= taken;
            undecided &= ~(taken | dropped);
        }
        const bool is_active = ((active_bits >> lane) & 1ull) != 0;
        active[lane < n ? first + a : active_pad + lane] = is_active ? 1 : 0;
        if (is_active) {
#pragma unroll
            for (int i = 0; i < kPrefetchTargets; ++i) {
                if (i < c.n) {
                    const int t = c.t[i];
                    atomicAnd(&s_mask[t >> 5], ~(1u << (t & 31)));
                }
            }
            const int* list = targets + static_cast<size_t>(first + a) * target_stride;
            for (int i = kPrefetchTargets; i < c.n; ++i) {
                const int t = list[i];
                atomicAnd(&s_mask[t >> 5], ~(1u << (t & 31)));
            }
        }
        __builtin_amdgcn_wave_barrier();
    };
    for (int base = 0; base < count; base += 64 * kPasteDepth) {
#pragma unroll
        for (int k = 0; k < kPasteDepth; ++k) {
            // Chunks past the frame decide nothing (n = 0).
            const int chunk = base + 64 * k;
            decide(ring[k], chunk);
            load(chunk + 64 * kPasteDepth, ring[k]);
        }
    }
}

// Appends the active anchors of [anchor_begin, anchor_begin + anchor_count)
// to a group list (order is not significant: every later stage is per group
// and accumulation is order-independent).
__global__ __launch_bounds__(256) void compact_kernel(
    const unsigned char* __restrict__ active, int anchor_begin, int anchor_count,
    int* __restrict__ group_list, int* __restrict__ group_count) {
    const int local = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = anchor_begin + local;
    const bool take = local < anchor_count && active[a] != 0;
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
    (void)d;
    return static_cast<size_t>((max_anchors + 31) / 32) * 4;
}

int paste_target_stride(const StageDesc& d) {
    return (5 * d.cap + 3) / 4 * 4;  // rows of 16-byte multiples
}

size_t paste_targets_size(const StageDesc& d, int anchor_count) {
    // The walk reads kPrefetchTargets entries of every row, the last row's
    // past the end of the rows.
    return static_cast<size_t>(anchor_count) * paste_target_stride(d) + kPrefetchTargets;
}

hipError_t launch_paste(const StageDesc& d, int frame_count, const int* frame_offsets,
                        const int* frame_ids, const int* row_tables, const int* counts,
                        const uint32_t* positions, unsigned char* active, int max_anchors,
                        int anchor_count, unsigned long long* dep, int* target_counts,
                        int* targets, int active_pad, hipStream_t stream) {
    if (frame_count <= 0 || anchor_count <= 0) {
        return hipSuccess;
    }
    const int stride = paste_target_stride(d);
    hipLaunchKernelGGL(paste_targets_kernel, dim3(static_cast<unsigned>((anchor_count + 255) / 256)),
                       dim3(256), 0, stream, d, frame_count, frame_offsets, frame_ids, row_tables,
                       counts, positions, dep, target_counts, targets, stride);
    hipError_t status = hipGetLastError();
    if (status != hipSuccess) {
        return status;
    }
    hipLaunchKernelGGL(paste_kernel, dim3(static_cast<unsigned>(frame_count)), dim3(kPasteThreads),
                       paste_shared_bytes(d, max_anchors), stream, d, frame_offsets, dep,
                       target_counts, targets, stride, active, active_pad);
    return hipGetLastError();
}

hipError_t launch_compact(const unsigned char* active, int anchor_begin, int anchor_count,
                          int* group_list, int* group_count, hipStream_t stream) {
    if (anchor_count <= 0) {
        return hipSuccess;
    }
    const dim3 grid(static_cast<unsigned>((anchor_count + 255) / 256));
    hipLaunchKernelGGL(compact_kernel, grid, dim3(256), 0, stream, active, anchor_begin,
                       anchor_count, group_list, group_count);
    return hipGetLastError();
}

}  // namespace vnlbhip
