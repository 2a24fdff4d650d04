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

// Short path of the selection.  Split the candidates into G >= k disjoint
// subsets (the threads' strided elements, G groups of threads): the largest
// subset minimum T0 bounds the k-th smallest distance, so the candidates
// with key <= T0 (typically a few dozen) contain the k nearest.  When they
// fit one per thread and T = max(tau, d_k) <= T0 (so T does not reach past
// them), the whole selection, cap included, is decided within that list,
// which stays in scan order.  Returns false, with nothing written, when a
// condition fails; every decision is block-uniform.
__device__ inline bool select_short(const StageDesc& d, const float* s_dist, int n, int k,
                                    int buffer_cap, unsigned* s_u, int* s_int, int* s_scan,
                                    float* s_sel_dist, int* s_sel_index, int* s_list_index,
                                    int& selected) {
    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int groups = k <= 16 ? 16 : (k <= 64 ? 64 : 0);
    if (groups == 0 || groups > nthreads) {
        return false;
    }
    uint32_t m = 0xFFFFFFFFu;
    for (int c = tid; c < n; c += nthreads) {
        const uint32_t key = distance_key(s_dist[c]);
        m = key < m ? key : m;
    }
    s_u[tid] = m;
    if (tid == 0) {
        s_int[10] = 0;
        s_int[12] = 0;
    }
    __syncthreads();
    if (tid < 64) {
        uint32_t gm = 0;
        if (tid < groups) {
            const int per = nthreads / groups;
            gm = 0xFFFFFFFFu;
            for (int i = 0; i < per; ++i) {
                const uint32_t v = s_u[tid * per + i];
                gm = v < gm ? v : gm;
            }
        }
#pragma unroll
        for (int offset = 32; offset >= 1; offset >>= 1) {
            const uint32_t other = static_cast<uint32_t>(__shfl_xor(static_cast<int>(gm), offset, 64));
            gm = other > gm ? other : gm;
        }
        if (tid == 0) {
            s_int[8] = static_cast<int>(gm);
        }
    }
    __syncthreads();
    const uint32_t t0 = static_cast<uint32_t>(s_int[8]);

    // Candidates with key <= T0, compacted in scan order.
    const int chunk = (n + nthreads - 1) / nthreads;
    const int begin = tid * chunk < n ? tid * chunk : n;
    const int end = begin + chunk < n ? begin + chunk : n;
    int local = 0;
    for (int c = begin; c < end; ++c) {
        local += distance_key(s_dist[c]) <= t0 ? 1 : 0;
    }
    int list = 0;
    int at = block_exclusive_scan(local, list, s_scan);
    if (list > nthreads) {
        return false;
    }
    for (int c = begin; c < end; ++c) {
        const uint32_t key = distance_key(s_dist[c]);
        if (key <= t0) {
            s_u[at] = key;
            s_list_index[at] = c;
            ++at;
        }
    }
    __syncthreads();

    // Rank of each entry by (key, scan position); d_k has rank k - 1.
    uint32_t key_e = 0xFFFFFFFFu;
    int rank = 0;
    if (tid < list) {
        key_e = s_u[tid];
        for (int f = 0; f < list; ++f) {
            const uint32_t kf = s_u[f];
            rank += (kf < key_e || (kf == key_e && f < tid)) ? 1 : 0;
        }
        if (rank == k - 1) {
            s_int[9] = static_cast<int>(key_e);
        }
    }
    __syncthreads();
    const float kth = __uint_as_float(static_cast<uint32_t>(s_int[9]));
    const float threshold = d.tau < kth ? kth : d.tau;
    const uint32_t threshold_key = distance_key(threshold);
    if (threshold_key > t0) {
        return false;
    }
    if (tid < list && key_e <= threshold_key) {
        atomicAdd(&s_int[10], 1);
    }
    __syncthreads();
    const int within = s_int[10];
    uint32_t limit_key = threshold_key;
    int need_equal = 0x7FFFFFFF;
    if (within > buffer_cap) {
        if (tid < list && rank == buffer_cap - 1) {
            s_int[11] = static_cast<int>(key_e);
        }
        __syncthreads();
        limit_key = static_cast<uint32_t>(s_int[11]);
        if (tid < list && key_e < limit_key) {
            atomicAdd(&s_int[12], 1);
        }
        __syncthreads();
        need_equal = buffer_cap - s_int[12];
    }
    const bool equal = tid < list && key_e == limit_key;
    int equal_total = 0;
    const int before_equal = block_exclusive_scan(equal ? 1 : 0, equal_total, s_scan);
    const bool take = tid < list && (key_e < limit_key || (equal && before_equal < need_equal));
    const int out = block_exclusive_scan(take ? 1 : 0, selected, s_scan);
    if (take) {
        s_sel_dist[out] = __uint_as_float(key_e);
        s_sel_index[out] = s_list_index[tid];
    }
    __syncthreads();
    return true;
}

// The general path: radix selections over all candidates.  Fills
// s_sel_dist / s_sel_index in scan order and returns the selected count.
__device__ inline int select_general(const StageDesc& d, const float* s_dist, int n, int k,
                                     int buffer_cap, unsigned* s_hist, int* s_int, int* s_scan,
                                     float* s_sel_dist, int* s_sel_index) {
    if (threadIdx.x == 0) {
        s_int[3] = 0;
        s_int[4] = 0;
    }
    __syncthreads();
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
    return selected;
}

// Ranks the selected candidates by (distance, scan position), the K nearest
// first and ascending, the rest in scan order, and writes the anchor's list.
__device__ inline void write_selected(const float* s_sel_dist, const int* s_sel_index,
                                      int* s_sel_rank, int selected, int k, int af, int t_low,
                                      int t_high, int x_low, int y_low, int nx, int per_frame, int a,
                                      size_t out_base, int* __restrict__ counts,
                                      uint32_t* __restrict__ positions,
                                      float* __restrict__ distances,
                                      const int* frame_x_low, const int* frame_y_low) {
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
        // Guided search: each scheduled frame has its own window origin.
        const int cx = (frame_x_low != nullptr ? frame_x_low[ti] : x_low) + (rem - yy * nx);
        const int cy = (frame_y_low != nullptr ? frame_y_low[ti] : y_low) + yy;
        const int frame = scheduled_frame(af, t_low, t_high, ti);
        positions[out_base + slot] = pack_position(cx, cy, frame);
        distances[out_base + slot] = s_sel_dist[e];
    }
    if (threadIdx.x == 0) {
        counts[a] = selected;
    }
}

// Selection over s_dist (scan order) into the anchor's output list.
__device__ inline void select_and_write(const StageDesc& d, const float* s_dist, int n, int af,
                                 int t_low, int t_high, int x_low, int y_low, int nx,
                                 int per_frame, int nt, unsigned* s_hist, int* s_int, int* s_scan,
                                 float* s_sel_dist, int* s_sel_index, int* s_sel_rank, int a,
                                 size_t out_base, int* __restrict__ counts,
                                 uint32_t* __restrict__ positions,
                                 float* __restrict__ distances,
                                 const int* frame_x_low = nullptr,
                                 const int* frame_y_low = nullptr) {
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

    // Buffer capacity of find_similar_patches (core.cpp:1279-1287).
    const long long window_cap = static_cast<long long>(d.search_window) * d.search_window * nt;
    const int buffer_cap = static_cast<int>(window_cap < d.soft_cap ? window_cap : d.soft_cap);

    int selected = 0;
    if (!select_short(d, s_dist, n, k, buffer_cap, s_hist, s_int, s_scan, s_sel_dist, s_sel_index,
                      s_sel_rank, selected)) {
        selected = select_general(d, s_dist, n, k, buffer_cap, s_hist, s_int, s_scan, s_sel_dist,
                                  s_sel_index);
    }
    write_selected(s_sel_dist, s_sel_index, s_sel_rank, selected, k, af, t_low, t_high, x_low,
                   y_low, nx, per_frame, a, out_base, counts, positions, distances, frame_x_low,
                   frame_y_low);
}

}  // namespace vnlbhip
