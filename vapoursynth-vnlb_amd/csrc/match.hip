// Block matching: one workgroup per anchor.
//
// Follows find_similar_patches / select_similar_patches of the reference CPU
// path (src/core/core.cpp:1200-1362): squared distances over the scheduled
// frames and the boundary-shifted search window, the K nearest, then every
// candidate within T = max(tau, d_K), capped at min(soft cap, window size).
// The reference's sequential heap and bounded buffer keep exactly that set
// (the CPU restatement checks the equivalence); the output order is the K
// nearest ascending by (distance, scan position), then the rest in scan
// order.  Distances are summed in the reference's row order
// (scan_*_spatial_distances core.cpp:1033-1150) with the 16-lane row kernel.
#include "common.h"
#include "device_util.h"

namespace vnlbhip {
namespace {

constexpr int kMatchThreads = 256;

__device__ inline uint32_t distance_key(float d) {
    // Distances are sums of squares (>= +0) or +inf; NaN sorts last.
    return d != d ? 0xFFFFFFFFu : __float_as_uint(d);
}

// Squared distance of one row of P samples added to `distance`
// (add_squared_row_distance_highway, distance_highway.hpp:11-58, 16 lanes).
template <int kP>
__device__ inline float add_row_distance(float distance, const float* ref_row,
                                         const float* __restrict__ cand_row,
                                         int runtime_p) {
    const int p = kP > 0 ? kP : runtime_p;
    if (p == kLanes) {
        float sq[kLanes];
#pragma unroll
        for (int i = 0; i < kLanes; ++i) {
            const float diff = ref_row[i] - cand_row[i];
            sq[i] = diff * diff;
        }
#pragma unroll
        for (int width = kLanes / 2; width >= 1; width >>= 1) {
#pragma unroll
            for (int i = 0; i < width; ++i) {
                sq[i] = sq[i] + sq[i + width];
            }
        }
        return distance + sq[0];
    }
    if (kP > 0) {
#pragma unroll
        for (int i = 0; i < kP; ++i) {
            const float diff = ref_row[i] - cand_row[i];
            const float square = diff * diff;
            distance = distance + square;
        }
    } else {
        for (int i = 0; i < p; ++i) {
            const float diff = ref_row[i] - cand_row[i];
            const float square = diff * diff;
            distance = distance + square;
        }
    }
    return distance;
}

// k-th smallest key (1-based rank) among n distances by MSB-first radix
// select over 8-bit digits.
__device__ uint32_t radix_select(const float* dist, int n, int rank,
                                 unsigned* hist, int* shared) {
    uint32_t prefix = 0;
    uint32_t mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = threadIdx.x; i < 256; i += blockDim.x) {
            hist[i] = 0;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const uint32_t key = distance_key(dist[i]);
            if ((key & mask) == prefix) {
                atomicAdd(&hist[(key >> shift) & 0xFFu], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            const unsigned h0 = hist[lane * 4 + 0];
            const unsigned h1 = hist[lane * 4 + 1];
            const unsigned h2 = hist[lane * 4 + 2];
            const unsigned h3 = hist[lane * 4 + 3];
            const int local = static_cast<int>(h0 + h1 + h2 + h3);
            const int inclusive = wave_inclusive_scan(local);
            const int exclusive = inclusive - local;
            if (exclusive < rank && rank <= inclusive) {
                int before = exclusive;
                int bin = lane * 4;
                const unsigned hs[4] = {h0, h1, h2, h3};
                for (int j = 0; j < 4; ++j) {
                    if (before + static_cast<int>(hs[j]) >= rank) {
                        bin = lane * 4 + j;
                        break;
                    }
                    before += static_cast<int>(hs[j]);
                }
                shared[0] = bin;
                shared[1] = before;
            }
        }
        __syncthreads();
        const int bin = shared[0];
        rank -= shared[1];
        prefix |= static_cast<uint32_t>(bin) << shift;
        mask |= 0xFFu << shift;
        __syncthreads();
    }
    return prefix;
}

template <int kP>
__global__ __launch_bounds__(kMatchThreads) void match_kernel(
    StageDesc d, const float* __restrict__ source,
    const uint32_t* __restrict__ anchors, int anchor_count,
    int* __restrict__ counts, uint32_t* __restrict__ positions,
    float* __restrict__ distances) {
    __shared__ float s_dist[kMaxCandidates];
    __shared__ float s_ref[kMaxDim];
    __shared__ unsigned s_hist[256];
    __shared__ int s_int[16];
    __shared__ int s_scan[8];
    __shared__ float s_sel_dist[kMaxCap];
    __shared__ int s_sel_index[kMaxCap];
    __shared__ int s_sel_rank[kMaxCap];

    const int a = blockIdx.x;
    if (a >= anchor_count) {
        return;
    }
    const uint32_t anchor = anchors[a];
    const int ax = position_x(anchor);
    const int ay = position_y(anchor);
    const int af = position_frame(anchor);
    const size_t out_base = static_cast<size_t>(a) * d.cap;

    if (d.similar <= 1) {
        if (threadIdx.x == 0) {
            counts[a] = 1;
            positions[out_base] = anchor;
            distances[out_base] = 0.0F;
        }
        return;
    }

    const bool final_stage = d.stage == kFinal;
    const int match_channels = final_stage ? d.channels : 1;
    const int P = kP > 0 ? kP : d.patch;
    const int area = P * P;

    int t_low = 0;
    int t_high = 0;
    temporal_range(d, af, t_low, t_high);
    const int nt = t_high - t_low + 1;
    int x_low = 0;
    int x_high = 0;
    int y_low = 0;
    int y_high = 0;
    spatial_range(d, ax, d.width, x_low, x_high);
    spatial_range(d, ay, d.height, y_low, y_high);
    const int nx = x_high - x_low + 1;
    const int ny = y_high - y_low + 1;
    const int per_frame = nx * ny;
    const int n = nt * per_frame;

    // Reference patch [c][t][y][x] (load_reference_patch_* core.cpp:1152-1198).
    const int ref_len = match_channels * d.patch_time * area;
    for (int i = threadIdx.x; i < ref_len; i += blockDim.x) {
        const int c = i / (d.patch_time * area);
        const int rem = i - c * d.patch_time * area;
        const int t = rem / area;
        const int yx = rem - t * area;
        const int y = yx / P;
        const int x = yx - y * P;
        s_ref[i] = video_at(source, d, ax + x, ay + y, af + t, c);
    }
    __syncthreads();

    // Distances in scan order.
    int valid_local = 0;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const int ti = c / per_frame;
        const int rem = c - ti * per_frame;
        const int yy = rem / nx;
        const int cx = x_low + (rem - yy * nx);
        const int cy = y_low + yy;
        const int frame = scheduled_frame(af, t_low, t_high, ti);
        float distance = 0.0F;
        for (int ch = 0; ch < match_channels; ++ch) {
            for (int t = 0; t < d.patch_time; ++t) {
                const float* plane =
                    source + (static_cast<size_t>(frame + t) * d.channels + ch) *
                                 d.height * d.width;
                const float* ref_plane = s_ref + (ch * d.patch_time + t) * area;
                for (int y = 0; y < P; ++y) {
                    distance = add_row_distance<kP>(
                        distance, ref_plane + y * P,
                        plane + static_cast<size_t>(cy + y) * d.width + cx, P);
                }
            }
        }
        s_dist[c] = distance;
        valid_local += distance == distance ? 1 : 0;
    }
    if (threadIdx.x == 0) {
        s_int[2] = 0;
        s_int[3] = 0;
        s_int[4] = 0;
    }
    __syncthreads();
    atomicAdd(&s_int[2], valid_local);
    __syncthreads();
    const int valid = s_int[2];
    if (valid == 0) {
        if (threadIdx.x == 0) {
            counts[a] = 0;
        }
        return;
    }
    const int k = d.similar < valid ? d.similar : valid;

    // T = max(tau, d_K) (select_similar_patches core.cpp:1217-1218).
    const uint32_t kth_key = radix_select(s_dist, n, k, s_hist, s_int);
    const float kth = __uint_as_float(kth_key);
    const float threshold = d.tau < kth ? kth : d.tau;

    int within_local = 0;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        within_local += s_dist[c] <= threshold ? 1 : 0;
    }
    atomicAdd(&s_int[3], within_local);
    __syncthreads();
    const int within = s_int[3];

    // Buffer capacity of find_similar_patches (core.cpp:1279-1287).
    const long long window_cap =
        static_cast<long long>(d.search_window) * d.search_window * nt;
    const int buffer_cap =
        static_cast<int>(window_cap < d.soft_cap ? window_cap : d.soft_cap);

    // Selection predicate: d < limit, or d == limit within the first
    // `need_equal` ties in scan order.
    float limit = threshold;
    int need_equal = 0x7FFFFFFF;
    if (within > buffer_cap) {
        const uint32_t cap_key = radix_select(s_dist, n, buffer_cap, s_hist, s_int);
        limit = __uint_as_float(cap_key);
        int less_local = 0;
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            less_local += s_dist[c] < limit ? 1 : 0;
        }
        atomicAdd(&s_int[4], less_local);
        __syncthreads();
        need_equal = buffer_cap - s_int[4];
    }

    // Ordered compaction over contiguous chunks of the scan order.
    const int chunk = (n + blockDim.x - 1) / blockDim.x;
    const int begin = threadIdx.x * chunk < n ? threadIdx.x * chunk : n;
    const int end = begin + chunk < n ? begin + chunk : n;
    int equal_local = 0;
    for (int c = begin; c < end; ++c) {
        equal_local += s_dist[c] == limit ? 1 : 0;
    }
    int equal_total = 0;
    int equal_rank = block_exclusive_scan(equal_local, equal_total, s_scan);
    int selected_local = 0;
    for (int c = begin; c < end; ++c) {
        const float dc = s_dist[c];
        if (dc < limit) {
            ++selected_local;
        } else if (dc == limit) {
            if (equal_rank < need_equal) {
                ++selected_local;
            }
            ++equal_rank;
        }
    }
    int selected = 0;
    int out_index = block_exclusive_scan(selected_local, selected, s_scan);
    equal_rank -= equal_local;
    for (int c = begin; c < end; ++c) {
        const float dc = s_dist[c];
        bool take = false;
        if (dc < limit) {
            take = true;
        } else if (dc == limit) {
            take = equal_rank < need_equal;
            ++equal_rank;
        }
        if (take) {
            s_sel_dist[out_index] = dc;
            s_sel_index[out_index] = c;
            ++out_index;
        }
    }
    __syncthreads();

    // Rank by (distance, scan position); the K nearest lead, ascending.
    const int k_out = k < selected ? k : selected;
    for (int e = threadIdx.x; e < selected; e += blockDim.x) {
        const float de = s_sel_dist[e];
        int rank = 0;
        for (int f = 0; f < selected; ++f) {
            const float df = s_sel_dist[f];
            rank += (df < de || (df == de && f < e)) ? 1 : 0;
        }
        s_sel_rank[e] = rank;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < selected; e += blockDim.x) {
        const int rank = s_sel_rank[e];
        int slot = rank;
        if (rank >= k_out) {
            int leading_before = 0;
            for (int f = 0; f < e; ++f) {
                leading_before += s_sel_rank[f] < k_out ? 1 : 0;
            }
            slot = k_out + (e - leading_before);
        }
        const int c = s_sel_index[e];
        const int ti = c / per_frame;
        const int rem = c - ti * per_frame;
        const int yy = rem / nx;
        const int cx = x_low + (rem - yy * nx);
        const int cy = y_low + yy;
        const int frame = scheduled_frame(af, t_low, t_high, ti);
        positions[out_base + slot] = pack_position(cx, cy, frame);
        distances[out_base + slot] = s_sel_dist[e];
    }
    if (threadIdx.x == 0) {
        counts[a] = selected;
    }
}

}  // namespace

hipError_t launch_match(const StageDesc& d, const float* source,
                        const uint32_t* anchors, int anchor_count, int* counts,
                        uint32_t* positions, float* distances,
                        hipStream_t stream) {
    if (anchor_count <= 0) {
        return hipSuccess;
    }
    const dim3 grid(static_cast<unsigned>(anchor_count));
    const dim3 block(kMatchThreads);
    if (d.patch == 8) {
        hipLaunchKernelGGL(match_kernel<8>, grid, block, 0, stream, d, source,
                           anchors, anchor_count, counts, positions, distances);
    } else if (d.patch == 4) {
        hipLaunchKernelGGL(match_kernel<4>, grid, block, 0, stream, d, source,
                           anchors, anchor_count, counts, positions, distances);
    } else {
        hipLaunchKernelGGL(match_kernel<0>, grid, block, 0, stream, d, source,
                           anchors, anchor_count, counts, positions, distances);
    }
    return hipGetLastError();
}

}  // namespace vnlbhip
