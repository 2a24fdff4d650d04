// k-nearest selection of one anchor's candidates (shared by the exact and
// the MFMA matchers).
//
// select_similar_patches (src/core/core.cpp:1200-1230) over the distances of
// the scanned window: the K nearest, then every candidate within
// T = max(tau, d_K), capped at the buffer capacity of find_similar_patches
// (core.cpp:1279-1287) with ties kept in scan order.  Output order: the K
// nearest ascending by (distance, scan position), then the rest in scan
// order.
#pragma once

#include "common.h"
#include "device_util.h"

namespace vnlbhip {

__device__ inline uint32_t distance_key(float d) {
    // Distances are sums of squares (>= +0) or +inf; NaN sorts last.
    return d != d ? 0xFFFFFFFFu : __float_as_uint(d);
}

// k-th smallest key (1-based rank) among n distances by MSB-first radix
// select over 8-bit digits.
__device__ inline uint32_t radix_select(const float* dist, int n, int rank,
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

// Selection over s_dist (scan order) into the anchor's output list.
__device__ inline void select_and_write(const StageDesc& d, const float* s_dist, int n, int af,
                                 int t_low, int t_high, int x_low, int y_low, int nx,
                                 int per_frame, int nt, unsigned* s_hist, int* s_int, int* s_scan,
                                 float* s_sel_dist, int* s_sel_index, int* s_sel_rank, int a,
                                 size_t out_base, int* __restrict__ counts,
                                 uint32_t* __restrict__ positions,
                                 float* __restrict__ distances) {
    int valid_local = 0;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const float v = s_dist[c];
        valid_local += v == v ? 1 : 0;
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
    const long long window_cap = static_cast<long long>(d.search_window) * d.search_window * nt;
    const int buffer_cap = static_cast<int>(window_cap < d.soft_cap ? window_cap : d.soft_cap);

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

}  // namespace vnlbhip
