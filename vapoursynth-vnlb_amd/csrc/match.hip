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
#include "kernels.h"
#include "proposer.h"
#include "select.h"

namespace vnlbhip {
namespace {

constexpr int kMatchThreads = 256;

// Row distance of 4 horizontally adjacent candidates against one reference
// row, each added to its running distance in the reference's element order.
template <int kP>
__device__ inline void add_row_distance4(float (&distance)[4], const float* ref_row,
                                         const float* cand_row) {
    float values[kP + 3];
#pragma unroll
    for (int i = 0; i < kP + 3; ++i) {
        values[i] = cand_row[i];
    }
    float ref[kP];
#pragma unroll
    for (int i = 0; i < kP; ++i) {
        ref[i] = ref_row[i];
    }
    if (kP == kLanes) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float sq[kLanes];
#pragma unroll
            for (int i = 0; i < kLanes; ++i) {
                const float diff = ref[i] - values[j + i];
                sq[i] = diff * diff;
            }
#pragma unroll
            for (int width = kLanes / 2; width >= 1; width >>= 1) {
#pragma unroll
                for (int i = 0; i < width; ++i) {
                    sq[i] = sq[i] + sq[i + width];
                }
            }
            distance[j] = distance[j] + sq[0];
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int i = 0; i < kP; ++i) {
                const float diff = ref[i] - values[j + i];
                const float square = diff * diff;
                distance[j] = distance[j] + square;
            }
        }
    }
}

// Geometry of the LDS search region of one anchor.
struct Region {
    int rows;      // ny + P - 1
    int cols;      // nx + P - 1
    int frames;    // nt + Pt - 1
    int floats;    // frames * rows * cols (+ padding)
};

// Unguided: the scheduled frames share one window, frames t_low ..
// t_high + Pt - 1.  Guided: every scheduled frame has its own window, so
// each (frame, patch frame) pair has its own plane.
__host__ __device__ inline int region_floats(int window, int patch, int temporal, int patch_time,
                                             bool guided) {
    const int side = window + patch - 1;
    return (guided ? temporal * patch_time : temporal + patch_time - 1) * side * side + 8;
}

// Candidate distances of a full window: window^2 per scheduled frame.
__host__ __device__ inline int window_candidates(const StageDesc& d) {
    return d.search_window * d.search_window * (d.search_bwd + d.search_fwd + 1);
}

// Entries of each of the three selection arrays after the distances: the
// selected list (at most the cap) and the short path's candidate list (at
// most one per thread).
__host__ __device__ inline int selection_entries(const StageDesc& d) {
    return d.cap > kMatchThreads ? d.cap : kMatchThreads;
}

// Distances (scan order) with the search region staged in LDS one channel
// at a time; each thread owns 4 horizontally adjacent candidates of one
// row so every loaded sample feeds 4 distances.  The per-candidate order of
// additions is exactly that of the CPU scan: channel, frame offset, row,
// column.
template <int kP, bool kGuided>
__global__ __launch_bounds__(kMatchThreads) void match_kernel(
    StageDesc d, const float* __restrict__ source, const uint32_t* __restrict__ anchors,
    int anchor_count, int* __restrict__ counts, uint32_t* __restrict__ positions,
    float* __restrict__ distances, const int16_t* __restrict__ offsets, int offsets_temporal) {
    extern __shared__ float s_dyn[];
    // Guided search: the window origin of each scheduled frame.
    __shared__ int s_frame_x_low[kMaxTemporal];
    __shared__ int s_frame_y_low[kMaxTemporal];
    __shared__ float s_ref[kMaxDim];
    __shared__ unsigned s_hist[256];
    __shared__ int s_int[16];
    __shared__ int s_scan[8];

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
    constexpr int P = kP;
    constexpr int area = kP * kP;

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
    const int rcols = nx + P - 1;
    const int rrows = ny + P - 1;
    // Guided search (MVToolsFlowProvider::center_for, src/flow/flow.hpp:340-390, from
    // a dense offset field): the window of scheduled frame ti is centred on
    // the anchor plus its offset clamped to the frame (the anchor frame keeps
    // the anchor); the window size does not depend on the centre, since
    // compute_spatial_range shifts it inside the frame.
    constexpr bool guided = kGuided;
    if (guided) {
        for (int ti = threadIdx.x; ti < nt; ti += blockDim.x) {
            int cx = ax;
            int cy = ay;
            if (ti > 0) {
                const size_t at =
                    ((static_cast<size_t>(af) * offsets_temporal + ti) * d.height + ay) * d.width + ax;
                cx = min(max(ax + static_cast<int>(offsets[2 * at]), 0), d.width - 1);
                cy = min(max(ay + static_cast<int>(offsets[2 * at + 1]), 0), d.height - 1);
            }
            int low = 0;
            int high = 0;
            spatial_range(d, cx, d.width, low, high);
            s_frame_x_low[ti] = low;
            spatial_range(d, cy, d.height, low, high);
            s_frame_y_low[ti] = low;
        }
    }
    const int rframes = guided ? nt * d.patch_time : nt + d.patch_time - 1;
    const int plane = rrows * rcols;
    const int region_size = rframes * plane;

    float* s_region = s_dyn;
    float* s_dist = s_dyn + region_floats(d.search_window, P, d.search_bwd + d.search_fwd + 1,
                                          d.patch_time, guided);
    // Selection scratch after the candidate distances.
    float* s_sel_dist = s_dist + window_candidates(d);
    int* s_sel_index = reinterpret_cast<int*>(s_sel_dist + selection_entries(d));
    int* s_sel_rank = s_sel_index + selection_entries(d);

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
    for (int i = region_size + threadIdx.x; i < region_size + 8; i += blockDim.x) {
        s_region[i] = 0.0F;
    }

    const int groups_x = (nx + 3) >> 2;
    const int items = nt * ny * groups_x;
    for (int ch = 0; ch < match_channels; ++ch) {
        __syncthreads();
        for (int i = threadIdx.x; i < region_size; i += blockDim.x) {
            const int fr = i / plane;
            const int rem = i - fr * plane;
            const int ry = rem / rcols;
            const int rx = rem - ry * rcols;
            if (guided) {
                const int ti = fr / d.patch_time;
                const int t = fr - ti * d.patch_time;
                s_region[i] = video_at(source, d, s_frame_x_low[ti] + rx, s_frame_y_low[ti] + ry,
                                       scheduled_frame(af, t_low, t_high, ti) + t, ch);
            } else {
                s_region[i] = video_at(source, d, x_low + rx, y_low + ry, t_low + fr, ch);
            }
        }
        __syncthreads();
        for (int item = threadIdx.x; item < items; item += blockDim.x) {
            const int gx = item % groups_x;
            const int rest = item / groups_x;
            const int yy = rest % ny;
            const int ti = rest / ny;
            const int xx0 = gx * 4;
            const int frame = scheduled_frame(af, t_low, t_high, ti);
            const int base_c = ti * per_frame + yy * nx + xx0;
            const int valid = nx - xx0 < 4 ? nx - xx0 : 4;
            float dist[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dist[j] = (ch == 0 || j >= valid) ? 0.0F : s_dist[base_c + j];
            }
            for (int t = 0; t < d.patch_time; ++t) {
                const float* region_plane =
                    s_region + (guided ? ti * d.patch_time + t : frame - t_low + t) * plane;
                const float* ref_plane = s_ref + (ch * d.patch_time + t) * area;
#pragma unroll
                for (int y = 0; y < P; ++y) {
                    add_row_distance4<kP>(dist, ref_plane + y * P,
                                          region_plane + (yy + y) * rcols + xx0);
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (j < valid) {
                    s_dist[base_c + j] = dist[j];
                }
            }
        }
    }
    __syncthreads();
    select_and_write(d, s_dist, n, af, t_low, t_high, x_low, y_low, nx, per_frame, nt, s_hist,
                     s_int, s_scan, s_sel_dist, s_sel_index, s_sel_rank, a, out_base, counts,
                     positions, distances, guided ? s_frame_x_low : nullptr,
                     guided ? s_frame_y_low : nullptr);
}

// Rank-th smallest (1-based) of n unsigned keys key(i), MSB-first radix
// select over 8-bit digits (all threads of the block call it).
template <class KeyFn>
__device__ inline uint32_t radix_select_keys(KeyFn key, int n, int rank, unsigned* hist, int* shared) {
    uint32_t prefix = 0;
    uint32_t mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = threadIdx.x; i < 256; i += blockDim.x) {
            hist[i] = 0;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const uint32_t k = key(i);
            if ((k & mask) == prefix) {
                atomicAdd(&hist[(k >> shift) & 0xFFu], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            const unsigned h[4] = {hist[lane * 4], hist[lane * 4 + 1], hist[lane * 4 + 2], hist[lane * 4 + 3]};
            const int local = static_cast<int>(h[0] + h[1] + h[2] + h[3]);
            const int inclusive = wave_inclusive_scan(local);
            const int exclusive = inclusive - local;
            if (exclusive < rank && rank <= inclusive) {
                int before = exclusive;
                int bin = lane * 4 + 3;
                for (int j = 0; j < 4; ++j) {
                    if (before + static_cast<int>(h[j]) >= rank) {
                        bin = lane * 4 + j;
                        break;
                    }
                    before += static_cast<int>(h[j]);
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

// Proposing matcher (8x8 patches, unguided windows).  Channel 0 of the
// window is staged as in match_kernel; the network scores every candidate
// (tiles of 64, features from the staged region); the top_k scored
// candidates get exact distances in match_kernel's order (so each equals
// the exhaustive scan's value bit for bit) and the rest +inf, unless the
// score margin between the K-th and the (top_k + 1)-th best is below
// `margin`: then the whole window is scanned exactly.  The selection is
// match_kernel's.
template <bool kFp8>
__global__ __launch_bounds__(kMatchThreads) void learned_match_kernel(
    StageDesc d, const float* __restrict__ source, const uint32_t* __restrict__ anchors, int anchor_count,
    int* __restrict__ counts, uint32_t* __restrict__ positions, float* __restrict__ distances, ProposerWeights wt,
    int top_k, float margin, int* __restrict__ fallbacks) {
    constexpr int P = 8;
    constexpr int area = P * P;
    extern __shared__ float s_dyn[];
    __shared__ float s_ref[kMaxDim];
    __shared__ unsigned s_hist[256];
    __shared__ int s_int[16];
    __shared__ int s_scan[8];
    __shared__ __attribute__((aligned(16))) __bf16 s_x[kPropTile * kPropXPitch];
    __shared__ __attribute__((aligned(16))) __bf16 s_h[kPropTile * kPropHPitch];
    __shared__ float s_cbox[kPropTile * kPropBoxes];
    __shared__ float s_abox[kPropBoxes];
    __shared__ float s_partial[4 * 64];
    __shared__ float s_tile[kPropTile];
    __shared__ int s_prop[kPropMaxProposed];
    __shared__ float s_pdist[kPropMaxProposed];
    __shared__ int s_mode[2];  // fallback flag, proposal count

    const int a = blockIdx.x;
    if (a >= anchor_count) {
        return;
    }
    const int tid = threadIdx.x;
    const uint32_t anchor = anchors[a];
    const int ax = position_x(anchor);
    const int ay = position_y(anchor);
    const int af = position_frame(anchor);
    const size_t out_base = static_cast<size_t>(a) * d.cap;
    const bool final_stage = d.stage == kFinal;
    const int match_channels = final_stage ? d.channels : 1;

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
    const int rcols = nx + P - 1;
    const int rrows = ny + P - 1;
    const int rframes = nt + d.patch_time - 1;
    const int plane = rrows * rcols;
    const int region_size = rframes * plane;
    float* s_region = s_dyn;
    float* s_dist = s_dyn + region_floats(d.search_window, P, d.search_bwd + d.search_fwd + 1, d.patch_time, false);
    float* s_sel_dist = s_dist + window_candidates(d);
    int* s_sel_index = reinterpret_cast<int*>(s_sel_dist + selection_entries(d));
    int* s_sel_rank = s_sel_index + selection_entries(d);

    const int ref_len = match_channels * d.patch_time * area;
    for (int i = tid; i < ref_len; i += blockDim.x) {
        const int c = i / (d.patch_time * area);
        const int rem = i - c * d.patch_time * area;
        const int t = rem / area;
        const int yx = rem - t * area;
        s_ref[i] = video_at(source, d, ax + yx % P, ay + yx / P, af + t, c);
    }
    for (int i = region_size + tid; i < region_size + 8; i += blockDim.x) {
        s_region[i] = 0.0F;
    }
    if (tid == 0) {
        s_mode[0] = n <= top_k ? 1 : 0;  // a window no larger than top_k is scanned whole
        s_mode[1] = 0;
    }
    const int boxes = 16 * d.patch_time;
    const int groups_x = (nx + 3) >> 2;
    const int items = nt * ny * groups_x;
    for (int ch = 0; ch < match_channels; ++ch) {
        __syncthreads();
        for (int i = tid; i < region_size; i += blockDim.x) {
            const int fr = i / plane;
            const int rem = i - fr * plane;
            const int ry = rem / rcols;
            const int rx = rem - ry * rcols;
            s_region[i] = video_at(source, d, x_low + rx, y_low + ry, t_low + fr, ch);
        }
        __syncthreads();
        if (ch == 0 && s_mode[0] == 0) {
            // Scores of every candidate (s_dist holds them until the
            // proposals are chosen).
            if (tid < d.patch_time) {
                patch_box_means(s_ref + tid * area, P, s_abox + 16 * tid);
            }
            for (int base = 0; base < n; base += kPropTile) {
                {
                    const int r = tid >> 2;
                    const int q = tid & 3;
                    const int i = base + r;
                    if (i < n) {
                        const int ti = i / per_frame;
                        const int rem = i - ti * per_frame;
                        const int yy = rem / nx;
                        const int xx = rem - yy * nx;
                        const int frame = scheduled_frame(af, t_low, t_high, ti);
                        for (int t = 0; t < d.patch_time; ++t) {
                            const float* at = s_region + (frame - t_low + t) * plane + (yy + 2 * q) * rcols + xx;
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const float* p = at + 2 * j;
                                s_cbox[r * kPropBoxes + t * 16 + q * 4 + j] =
                                    ((p[0] + p[1]) + (p[rcols] + p[rcols + 1])) * 0.25F;
                            }
                        }
                    }
                }
                __syncthreads();
                {
                    const int r = tid >> 2;
                    const int q = tid & 3;
                    const int i = base + r;
                    float v[32];
                    if (i >= n) {
                        for (int k = 0; k < 32; ++k) {
                            v[k] = 0.0F;
                        }
                    } else {
                        const float* cb = s_cbox + r * kPropBoxes;
                        if (q < 3) {
                            for (int k = 0; k < 32; ++k) {
                                float value = 0.0F;
                                if (k < boxes) {
                                    value = q == 0 ? s_abox[k] - 0.5F
                                                   : (q == 1 ? cb[k] - 0.5F : fabsf(s_abox[k] - cb[k]) * 4.0F);
                                }
                                v[k] = value;
                            }
                        } else {
                            const int ti = i / per_frame;
                            const int rem = i - ti * per_frame;
                            const int yy = rem / nx;
                            const int xx = rem - yy * nx;
                            const int frame = scheduled_frame(af, t_low, t_high, ti);
                            float ssd = 0.0F;
                            for (int k = 0; k < boxes; ++k) {
                                const float diff = s_abox[k] - cb[k];
                                ssd = ssd + diff * diff;
                            }
                            v[0] = static_cast<float>(x_low + xx - ax) * (1.0F / 16.0F);
                            v[1] = static_cast<float>(y_low + yy - ay) * (1.0F / 16.0F);
                            v[2] = static_cast<float>(frame - af) * 0.25F;
                            v[3] = ssd * 4.0F;
                            for (int k = 4; k < 32; ++k) {
                                v[k] = 0.0F;
                            }
                        }
                    }
                    store_features32<kFp8>(s_x, r, 32 * q, v, kFp8 ? wt.fp8.scales + kPropScaleX : nullptr);
                }
                __syncthreads();
                if constexpr (kFp8) {
                    proposer_tile_scores_fp8(reinterpret_cast<const uint8_t*>(s_x), reinterpret_cast<uint8_t*>(s_h),
                                             s_partial, s_tile, wt);
                } else {
                    proposer_tile_scores(s_x, s_h, s_partial, s_tile, wt);
                }
                if (tid < kPropTile && base + tid < n) {
                    s_dist[base + tid] = s_tile[tid];
                }
            }
            __syncthreads();
            // Margin between the K-th and the (top_k + 1)-th best score.
            const auto key = [&](int i) { return score_rank_key(s_dist[i]); };
            const int k_rank = d.similar < n ? d.similar : n;
            const uint32_t key_k = radix_select_keys(key, n, k_rank, s_hist, s_int);
            const uint32_t key_top = radix_select_keys(key, n, top_k, s_hist, s_int);
            const uint32_t key_next = radix_select_keys(key, n, top_k + 1, s_hist, s_int);
            const float gap = score_from_rank_key(key_k) - score_from_rank_key(key_next);
            bool fallback = !(gap >= margin);
            if (!fallback) {
                // Proposals: every candidate scored at least the top_k-th
                // best, in scan order (more than top_k only on score ties).
                int count = 0;
                for (int c0 = 0; c0 < n; c0 += blockDim.x) {
                    const int c = c0 + tid;
                    const int take = c < n && key(c) <= key_top ? 1 : 0;
                    int total = 0;
                    const int at = block_exclusive_scan(take, total, s_scan);
                    if (take != 0 && count + at < kPropMaxProposed) {
                        s_prop[count + at] = c;
                    }
                    count += total;
                }
                fallback = count > kPropMaxProposed;
                if (tid == 0) {
                    s_mode[1] = count;
                }
            }
            __syncthreads();
            if (tid == 0) {
                s_mode[0] = fallback ? 1 : 2;
                if (fallback) {
                    atomicAdd(fallbacks, 1);
                }
            }
            __syncthreads();
        }
        const int mode = s_mode[0];
        if (mode == 2) {
            // Exact distances of the proposals, channel by channel in
            // match_kernel's order (frame offset, row, column).
            const int count = s_mode[1];
            for (int j = tid; j < count; j += blockDim.x) {
                const int i = s_prop[j];
                const int ti = i / per_frame;
                const int rem = i - ti * per_frame;
                const int yy = rem / nx;
                const int xx = rem - yy * nx;
                const int frame = scheduled_frame(af, t_low, t_high, ti);
                float dist = ch == 0 ? 0.0F : s_pdist[j];
                for (int t = 0; t < d.patch_time; ++t) {
                    const float* region_plane = s_region + (frame - t_low + t) * plane;
                    const float* ref_plane = s_ref + (ch * d.patch_time + t) * area;
                    for (int y = 0; y < P; ++y) {
                        const float* cand = region_plane + (yy + y) * rcols + xx;
                        const float* ref = ref_plane + y * P;
#pragma unroll
                        for (int x = 0; x < P; ++x) {
                            const float diff = ref[x] - cand[x];
                            const float square = diff * diff;
                            dist = dist + square;
                        }
                    }
                }
                s_pdist[j] = dist;
            }
        } else {
            for (int item = tid; item < items; item += blockDim.x) {
                const int gx = item % groups_x;
                const int rest = item / groups_x;
                const int yy = rest % ny;
                const int ti = rest / ny;
                const int xx0 = gx * 4;
                const int frame = scheduled_frame(af, t_low, t_high, ti);
                const int base_c = ti * per_frame + yy * nx + xx0;
                const int valid = nx - xx0 < 4 ? nx - xx0 : 4;
                float dist[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    dist[j] = (ch == 0 || j >= valid) ? 0.0F : s_dist[base_c + j];
                }
                for (int t = 0; t < d.patch_time; ++t) {
                    const float* region_plane = s_region + (frame - t_low + t) * plane;
                    const float* ref_plane = s_ref + (ch * d.patch_time + t) * area;
#pragma unroll
                    for (int y = 0; y < P; ++y) {
                        add_row_distance4<P>(dist, ref_plane + y * P, region_plane + (yy + y) * rcols + xx0);
                    }
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (j < valid) {
                        s_dist[base_c + j] = dist[j];
                    }
                }
            }
        }
    }
    __syncthreads();
    if (s_mode[0] == 2) {
        for (int i = tid; i < n; i += blockDim.x) {
            s_dist[i] = __builtin_huge_valf();
        }
        __syncthreads();
        const int count = s_mode[1];
        for (int j = tid; j < count; j += blockDim.x) {
            s_dist[s_prop[j]] = s_pdist[j];
        }
        __syncthreads();
    }
    select_and_write(d, s_dist, n, af, t_low, t_high, x_low, y_low, nx, per_frame, nt, s_hist, s_int, s_scan,
                     s_sel_dist, s_sel_index, s_sel_rank, a, out_base, counts, positions, distances);
}

}  // namespace

bool learned_match_supported(const StageDesc& d) {
    return d.patch == 8 && d.patch_time >= 1 && d.patch_time <= 2 && d.similar > 1;
}

hipError_t launch_learned_match(const StageDesc& d, const float* source, const uint32_t* anchors, int anchor_count,
                                int* counts, uint32_t* positions, float* distances, const ProposerWeights& wt,
                                int top_k, float margin, int precision, int* fallbacks, hipStream_t stream) {
    if (anchor_count <= 0) {
        return hipSuccess;
    }
    if (!learned_match_supported(d) || top_k < d.similar || top_k >= kPropMaxProposed || precision < 0 ||
        precision > 1 || (precision == 1 && (wt.fp8.w1 == nullptr || wt.fp8.w2 == nullptr || wt.fp8.scales == nullptr))) {
        return hipErrorInvalidValue;
    }
    const size_t shared = match_shared_bytes(d, false);
    if (precision == 1) {
        hipLaunchKernelGGL(learned_match_kernel<true>, dim3(static_cast<unsigned>(anchor_count)), dim3(kMatchThreads),
                           shared, stream, d, source, anchors, anchor_count, counts, positions, distances, wt, top_k,
                           margin, fallbacks);
    } else {
        hipLaunchKernelGGL(learned_match_kernel<false>, dim3(static_cast<unsigned>(anchor_count)), dim3(kMatchThreads),
                           shared, stream, d, source, anchors, anchor_count, counts, positions, distances, wt, top_k,
                           margin, fallbacks);
    }
    return hipGetLastError();
}

size_t match_shared_bytes(const StageDesc& d, bool guided) {
    const int temporal = d.search_bwd + d.search_fwd + 1;
    return (static_cast<size_t>(region_floats(d.search_window, d.patch, temporal, d.patch_time, guided)) +
            static_cast<size_t>(window_candidates(d)) + 3 * static_cast<size_t>(selection_entries(d))) *
           sizeof(float);
}

hipError_t launch_match(const StageDesc& d, const float* source,
                        const uint32_t* anchors, int anchor_count, int* counts,
                        uint32_t* positions, float* distances, const int16_t* offsets,
                        int offsets_temporal, hipStream_t stream) {
    if (anchor_count <= 0) {
        return hipSuccess;
    }
    const dim3 grid(static_cast<unsigned>(anchor_count));
    const dim3 block(kMatchThreads);
    const size_t shared = match_shared_bytes(d, offsets != nullptr);
    switch (d.patch) {
#define VNLBHIP_MATCH_CASE(PV)                                                                  \
    case PV:                                                                                    \
        if (offsets != nullptr) {                                                               \
            hipLaunchKernelGGL((match_kernel<PV, true>), grid, block, shared, stream, d, source, \
                               anchors, anchor_count, counts, positions, distances, offsets,   \
                               offsets_temporal);                                               \
        } else {                                                                                \
            hipLaunchKernelGGL((match_kernel<PV, false>), grid, block, shared, stream, d, source,\
                               anchors, anchor_count, counts, positions, distances, offsets,   \
                               offsets_temporal);                                               \
        }                                                                                       \
        break;
        VNLBHIP_MATCH_CASE(1)
        VNLBHIP_MATCH_CASE(2)
        VNLBHIP_MATCH_CASE(3)
        VNLBHIP_MATCH_CASE(4)
        VNLBHIP_MATCH_CASE(5)
        VNLBHIP_MATCH_CASE(6)
        VNLBHIP_MATCH_CASE(7)
        VNLBHIP_MATCH_CASE(8)
        VNLBHIP_MATCH_CASE(9)
        VNLBHIP_MATCH_CASE(10)
        VNLBHIP_MATCH_CASE(11)
        VNLBHIP_MATCH_CASE(12)
        VNLBHIP_MATCH_CASE(13)
        VNLBHIP_MATCH_CASE(14)
        VNLBHIP_MATCH_CASE(15)
        VNLBHIP_MATCH_CASE(16)
#undef VNLBHIP_MATCH_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace vnlbhip
