// Filter stage of a group: dual-basis map, Wiener coefficients, aggregation
// weights, reconstruction and the weighted scatter of every sample into the
// per-output-frame double accumulators.
//
// Semantics follow the reference's filter path (src/core/core.cpp:184-871,
// src/linalg/linalg.cpp:153-595) with the 16-lane summation order of the CPU
// restatement: a dot product is 16 lane-wise fma chains over the vector part
// of the row, a halving tree over the lanes, then the scalar tail in order.
//
// Slot layout.  Lane l of a 16-lane team owns the elements e = 16 k + l of a
// model row: slots k < chunks are the vector part, slot k == chunks (lanes
// l < tail) the scalar tail.  Per-element tables, the group means and the
// basis live in LDS per lane with a padded pitch, so a lane reads its slots
// with 128-bit loads and the 16 lanes of a team hit distinct banks; the four
// teams of a wavefront read the same basis words (broadcast).  Sample rows
// never go through LDS: a team loads its row straight into registers.
//
// Workgroups are persistent: each builds the stage-constant tables once and
// then walks the groups with a stride coprime with their count, so that
// concurrently running groups scatter into distant parts of the clip.
#include "common.h"
#include "device_util.h"
#include "group.h"
#include "kernels.h"

#include <float.h>

namespace vnlbhip {
namespace {

// LDS layout of the filter kernel in 4-byte words (host computed).
struct FilterLayout {
    int pitch;       // words per lane slot row (>= template slots, 16-byte multiple)
    int row;         // 16 * pitch: words per model row
    int chunks;      // vector slots per lane
    int tail;        // lanes holding a tail element
    int u_pitch;     // words per dual-eigenvector sample row (multiple of 4)
    int tab_in;      // [models][row] input offsets from the patch origin
    int tab_acc;     // [models][row] accumulator offsets from the patch's cell
    int tab_meta;    // [models][row] t | c << 8 | fused << 16 | valid << 17
    int tab_win;     // [models][row] aggregation window weight
    int mean_n;      // [models][row] noisy mean
    int mean_b;      // [models][row] basic mean (Final)
    int basis;       // [models][rank][row]
    int u;           // [models][max_rows][u_pitch] dual eigenvectors, sample-major
    int proj;        // [cap][models][rank] projections kept between the passes
    int pos;         // [cap] packed sample positions
    int logw;        // [cap]
    int w;           // [cap]
    int score;       // [models][cap]
    int proj_cache;  // Basic: scores and reconstruction share the projections
    int words;
};

// Template slot count chosen for a stage (register array length).
int filter_template_slots(const StageDesc& d) {
    return template_slots(d.model_dim);
}

FilterLayout filter_layout(const StageDesc& d) {
    FilterLayout L{};
    L.pitch = slot_pitch(d.model_dim);
    L.row = kLanes * L.pitch;
    L.chunks = d.model_dim / kLanes;
    L.tail = d.model_dim % kLanes;
    L.u_pitch = (d.rank + 3) / 4 * 4;
    const int models = d.models;
    int at = 0;
    const auto take = [&at](int words) {
        const int start = at;
        at += (words + 3) / 4 * 4;
        return start;
    };
    L.tab_in = take(models * L.row);
    L.tab_acc = take(models * L.row);
    L.tab_meta = take(models * L.row);
    L.tab_win = take(models * L.row);
    L.mean_n = take(models * L.row);
    L.mean_b = take(models * L.row);
    L.basis = take(models * d.rank * L.row);
    // Any group with fewer rows than a model dimension of at least 64 is dual.
    L.u = take(d.model_dim >= 64 ? models * d.max_rows * L.u_pitch : 0);
    const int proj_words = d.cap * models * d.rank;
    L.proj_cache = (d.stage == kBasic && proj_words <= 4096) ? 1 : 0;
    L.proj = take(L.proj_cache ? proj_words : 0);
    L.pos = take(d.cap);
    L.logw = take(d.cap);
    L.w = take(d.cap);
    L.score = take(models * d.cap);
    L.words = at;
    return L;
}

// Loads a lane's slots of an LDS row (slot layout) into registers with
// 128-bit reads; bt receives the tail slot.
template <int kS>
__device__ inline void load_slots(float (&b)[kS], float& bt, const float* row, int chunks,
                                  bool tail_lane) {
    const float4* b4 = reinterpret_cast<const float4*>(row);
#pragma unroll
    for (int q = 0; q < kS / 4; ++q) {
        if (4 * q < chunks) {
            const float4 v = b4[q];
            b[4 * q + 0] = v.x;
            b[4 * q + 1] = v.y;
            b[4 * q + 2] = v.z;
            b[4 * q + 3] = v.w;
        } else {
            b[4 * q + 0] = 0.0F;
            b[4 * q + 1] = 0.0F;
            b[4 * q + 2] = 0.0F;
            b[4 * q + 3] = 0.0F;
        }
    }
    bt = tail_lane ? row[chunks] : 0.0F;
}

// Dot product of two register row slices in the 16-lane order
// (dot_contiguous_highway, kernels_highway.hpp:12-44): lane-wise fma chains
// over the vector slots, a halving tree over the team, then the tail
// elements one by one.
template <int kS>
__device__ inline float team_dot(const float (&x)[kS], float xt, const float (&b)[kS], float bt,
                                 int chunks, int tail, int lane) {
    float part = 0.0F;
#pragma unroll
    for (int k = 0; k < kS; ++k) {
        if (k < chunks) {
            part = fmaf(x[k], b[k], part);
        }
    }
    float total = segment16_sum(part);
    if (tail > 0) {
        const float product = lane < tail ? xt * bt : 0.0F;
        for (int i = 0; i < tail; ++i) {
            total = total + __shfl(product, i, kLanes);
        }
    }
    return total;
}

// Squared norm of a register row slice in the same order.
template <int kS>
__device__ inline float team_norm2(const float (&x)[kS], float xt, int chunks, int tail,
                                   int lane) {
    float part = 0.0F;
#pragma unroll
    for (int k = 0; k < kS; ++k) {
        if (k < chunks) {
            part = fmaf(x[k], x[k], part);
        }
    }
    float total = segment16_sum(part);
    if (tail > 0) {
        const float product = lane < tail ? xt * xt : 0.0F;
        for (int i = 0; i < tail; ++i) {
            total = total + __shfl(product, i, kLanes);
        }
    }
    return total;
}

// Loads a lane's slots of one sample row, centered by an LDS mean in slot
// layout.  Slots outside the row stay zero.
template <int kS>
__device__ inline void load_centered(float (&x)[kS], float& xt, const float* __restrict__ video,
                                     size_t base, const int* tab_in, const float* mean,
                                     int chunks, bool tail_lane) {
#pragma unroll
    for (int k = 0; k < kS; ++k) {
        x[k] = 0.0F;
        if (k < chunks) {
            x[k] = video[base + tab_in[k]] - mean[k];
        }
    }
    xt = 0.0F;
    if (tail_lane) {
        xt = video[base + tab_in[chunks]] - mean[chunks];
    }
}

// kExact: the model row is exactly 16 * kS elements (no tail, no slot
// guards at run time).
template <int kS, bool kExact>
__global__ __launch_bounds__(kModelThreads, kS <= 24 ? 3 : 1) void filter_kernel(
    StageDesc d, FilterLayout L, const float* __restrict__ noisy,
    const float* __restrict__ basic, GroupRefs refs, const int* __restrict__ group_count_dev,
    int problem_stride, const int* __restrict__ group_flags, const float* __restrict__ means,
    const float* __restrict__ eigenvalues, const float* __restrict__ vectors, Accumulators acc,
    float* __restrict__ weights_out, int* __restrict__ work_counter) {
    extern __shared__ float4 s_dyn4[];
    float* s_dyn = reinterpret_cast<float*>(s_dyn4);
    __shared__ float s_lambda[kMaxChannels * kMaxRank];
    __shared__ float s_coef[kMaxChannels * kMaxRank];
    __shared__ float s_lhat[kMaxChannels * kMaxRank];
    __shared__ float s_trace_term[kMaxChannels * kMaxRank];
    __shared__ float s_volume_term[kMaxChannels * kMaxRank];
    __shared__ float s_scalar[4 * kMaxChannels];
    __shared__ float s_red[kMaxChannels][kModelThreads / 64];

    int* tab_in = reinterpret_cast<int*>(s_dyn + L.tab_in);
    int* tab_acc = reinterpret_cast<int*>(s_dyn + L.tab_acc);
    int* tab_meta = reinterpret_cast<int*>(s_dyn + L.tab_meta);
    float* tab_win = s_dyn + L.tab_win;
    float* s_mean_n = s_dyn + L.mean_n;
    float* s_mean_b = s_dyn + L.mean_b;
    float* s_basis = s_dyn + L.basis;
    float* s_u = s_dyn + L.u;
    float* s_proj = s_dyn + L.proj;
    uint32_t* s_pos = reinterpret_cast<uint32_t*>(s_dyn + L.pos);
    float* s_logw = s_dyn + L.logw;
    float* s_w = s_dyn + L.w;
    float* s_score = s_dyn + L.score;

    const int tid = threadIdx.x;
    const int team = tid / kLanes;
    const int lane = tid % kLanes;
    const int models = d.models;
    const int md = d.model_dim;
    const int R = d.rank;
    const int row = L.row;
    const int pitch = L.pitch;
    const int chunks = kExact ? kS : L.chunks;
    const int tail = kExact ? 0 : L.tail;
    const bool tail_lane = lane < tail;
    const bool final_stage = d.stage == kFinal;

    // Stage-constant element tables in slot layout.
    for (int i = tid; i < models * row; i += kModelThreads) {
        const int m = i / row;
        const int rem = i - m * row;
        const int l = rem / pitch;
        const int k = rem - l * pitch;
        const bool valid = k < chunks || (k == chunks && l < tail);
        int in_off = 0;
        int acc_off = 0;
        int meta = 0;
        float win = 0.0F;
        if (valid) {
            const int e = (d.coupled ? 0 : m * d.patch_dim) + kLanes * k + l;
            const int c = e / d.patch_dim;
            const int wi = e - c * d.patch_dim;
            const int t = wi / d.patch_area;
            const int yx = wi - t * d.patch_area;
            const int y = yx / d.patch;
            const int x = yx - y * d.patch;
            in_off = ((t * d.channels + c) * d.height + y) * d.width + x;
            acc_off = ((t * (d.channels + 1) + c) * d.height + y) * d.width + x;
            meta = t | (c << 8) | (x < d.fused_width ? kMetaFused : 0) | kMetaValid;
            win = d.window[yx];
        }
        tab_in[i] = in_off;
        tab_acc[i] = acc_off;
        tab_meta[i] = meta;
        tab_win[i] = win;
    }

    const float sigma = d.sigma;
    const float beta = d.beta;
    const float sigma2 = (beta * sigma) * sigma;
    const float model_noise2 = final_stage ? (beta * d.sigma_basic) * d.sigma_basic : sigma2;
    const float estimate_noise2 = sigma2;
    const float match_sigma = final_stage ? d.sigma_basic : sigma;
    const float match_noise2 = match_sigma * match_sigma;
    const float floor_noise2 = (d.membership_noise_floor * sigma) * sigma;
    const float membership_noise2 = match_noise2 > floor_noise2 ? match_noise2 : floor_noise2;
    const float dual_floor = 128.0F * FLT_EPSILON;
    const float* member_video = final_stage ? basic : noisy;
    const float* member_mean_all = final_stage ? s_mean_b : s_mean_n;

    // Group indices are taken in order from the work counter (the next one
    // requested while the current group is processed) and spread over the
    // list by the largest odd stride near 0.618 * count coprime with the
    // count (uniform: every thread computes it from the device count).
    const int group_count = *group_count_dev;
    unsigned group_stride = static_cast<unsigned>(static_cast<double>(group_count) * 0.6180339887) | 1u;
    while (group_stride > 1) {
        unsigned x = group_stride;
        unsigned y = static_cast<unsigned>(group_count);
        while (y != 0) {
            const unsigned r = x % y;
            x = y;
            y = r;
        }
        if (x == 1) {
            break;
        }
        group_stride -= 2;
    }
    if (group_stride == 0) {
        group_stride = 1;
    }
    __shared__ int s_next;
    if (threadIdx.x == 0) {
        s_next = atomicAdd(work_counter, 1);
    }
    int following = 0;
    const auto advance = [&]() {
        __syncthreads();  // the group is done with the LDS and s_next
        if (threadIdx.x == 0) {
            s_next = following;
        }
        __syncthreads();
        return s_next;
    };
    __syncthreads();
    for (int gi = s_next; gi < group_count; gi = advance()) {
        if (threadIdx.x == 0) {
            following = atomicAdd(work_counter, 1);
        }
        const int g = static_cast<int>((static_cast<unsigned long long>(gi) * group_stride) %
                                       static_cast<unsigned>(group_count));
        const int a = refs.group_list[g];
        const int af = position_frame(refs.anchors[a]);
        const int similar = refs.counts[a];
        const size_t list = static_cast<size_t>(a) * d.cap;
        const int flags = group_flags[g];
        const bool equal = (flags & kFlagEqual) != 0;
        const bool flat = (flags & kFlagFlat) != 0;
        const bool modelled = (flags & kFlagModelled) != 0;
        const int rows = model_rows_of(d, similar);
        const bool dual = group_uses_dual(d, rows);
        const bool fit = modelled && !equal;

        for (int r = tid; r < similar; r += kModelThreads) {
            s_pos[r] = refs.positions[list + r];
            s_logw[r] = 0.0F;
            s_w[r] = 1.0F;
        }
        const float* g_mean = means + static_cast<size_t>(g) * 2 * d.dim;
        for (int i = tid; i < models * row; i += kModelThreads) {
            const int m = i / row;
            const int rem = i - m * row;
            const int l = rem / pitch;
            const int k = rem - l * pitch;
            const bool valid = k < chunks || (k == chunks && l < tail);
            const int e = (d.coupled ? 0 : m * d.patch_dim) + kLanes * k + l;
            s_mean_n[i] = valid ? g_mean[e] : 0.0F;
            s_mean_b[i] = (valid && final_stage) ? g_mean[d.dim + e] : 0.0F;
        }
        if (fit) {
            for (int i = tid; i < models * R; i += kModelThreads) {
                s_lambda[(i / R) * kMaxRank + i % R] = eigenvalues[static_cast<size_t>(g) * models * R + i];
            }
            if (dual) {
                // Eigenvectors of the Gram matrix, sample-major: u[m][s][j].
                for (int i = tid; i < models * R * rows; i += kModelThreads) {
                    const int m = i / (R * rows);
                    const int rem = i - m * R * rows;
                    const int j = rem / rows;
                    const int s = rem - j * rows;
                    const size_t problem = static_cast<size_t>(g) * models + m;
                    s_u[(m * d.max_rows + s) * L.u_pitch + j] =
                        vectors[(problem * R + j) * problem_stride + s];
                }
            } else {
                // Covariance eigenvectors are the basis itself.
                for (int i = tid; i < models * R * row; i += kModelThreads) {
                    const int q = i / row;  // m * R + j
                    const int rem = i - q * row;
                    const int l = rem / pitch;
                    const int k = rem - l * pitch;
                    const bool valid = k < chunks || (k == chunks && l < tail);
                    s_basis[i] = valid ? vectors[static_cast<size_t>(g * models * R + q) * problem_stride +
                                                 kLanes * k + l]
                                       : 0.0F;
                }
            }
        }
        __syncthreads();

        if (fit && dual) {
            // map_dual_eigenvectors_to_basis (linalg.cpp:552-595): basis row j
            // is sum_s u[s][j] (x_s - mean) / sqrt(lambda_j), then normalized.
            for (int i = tid; i < models * md; i += kModelThreads) {
                const int m = i / md;
                const int e = i - m * md;
                const int k = e / kLanes;
                const int slot = m * row + (e % kLanes) * pitch + k;
                const bool fused = k < chunks;
                const int in_off = tab_in[slot];
                const float mean_e = member_mean_all[slot];
                for (int j0 = 0; j0 < R; j0 += 16) {
                    float accs[16];
#pragma unroll
                    for (int jj = 0; jj < 16; ++jj) {
                        accs[jj] = 0.0F;
                    }
                    for (int s = 0; s < rows; ++s) {
                        const float x = member_video[patch_base(d, s_pos[s]) + in_off] - mean_e;
                        const float4* u4 =
                            reinterpret_cast<const float4*>(s_u + (m * d.max_rows + s) * L.u_pitch + j0);
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            if (j0 + 4 * q < R) {
                                const float4 u = u4[q];
                                const float uu[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                                for (int h = 0; h < 4; ++h) {
                                    const int jj = 4 * q + h;
                                    if (j0 + jj < R) {
                                        if (fused) {
                                            accs[jj] = fmaf(uu[h], x, accs[jj]);
                                        } else {
                                            const float product = uu[h] * x;
                                            accs[jj] = accs[jj] + product;
                                        }
                                    }
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int jj = 0; jj < 16; ++jj) {
                        const int j = j0 + jj;
                        if (j < R) {
                            const float lambda = s_lambda[m * kMaxRank + j];
                            float value = 0.0F;
                            if (!(lambda <= dual_floor)) {
                                const float inv_sqrt = 1.0F / sqrtf(lambda);
                                value = accs[jj] * inv_sqrt;
                            }
                            s_basis[(m * R + j) * row + slot - m * row] = value;
                        }
                    }
                }
            }
            __syncthreads();
            // normalize_or_zero (linalg.cpp:235-246).
            for (int q = team; q < models * R; q += kTeams) {
                const int m = q / R;
                const int j = q - m * R;
                if (s_lambda[m * kMaxRank + j] <= dual_floor) {
                    continue;
                }
                float* brow = s_basis + q * row + lane * pitch;
                float v[kS];
#pragma unroll
                for (int k = 0; k < kS; ++k) {
                    v[k] = k < chunks ? brow[k] : 0.0F;
                }
                const float vt = tail_lane ? brow[chunks] : 0.0F;
                const float norm2 = team_norm2(v, vt, chunks, tail, lane);
                const bool zero = norm2 <= dual_floor;
                const float scale = zero ? 0.0F : 1.0F / sqrtf(norm2);
#pragma unroll
                for (int k = 0; k < kS; ++k) {
                    if (k < chunks) {
                        brow[k] = zero ? 0.0F : v[k] * scale;
                    }
                }
                if (tail_lane) {
                    brow[chunks] = zero ? 0.0F : vt * scale;
                }
            }
        }

        // compute_filter_coefficients (core.cpp:616-636) and the per-model
        // terms of accumulate_aggregation_model (core.cpp:326-412).
        if (fit) {
            const float sigma_basic2 = final_stage ? (beta * d.sigma_basic) * d.sigma_basic : sigma2;
            const float threshold = d.variance_threshold * sigma2;
            for (int i = tid; i < models * R; i += kModelThreads) {
                const int m = i / R;
                const int j = i - m * R;
                const float lambda = s_lambda[m * kMaxRank + j];
                float variance = lambda;
                if (sigma_basic2 > 0.0F) {
                    variance = variance - fminf(variance, sigma_basic2);
                }
                s_coef[m * kMaxRank + j] =
                    variance > threshold ? 1.0F / (1.0F + (sigma2 / variance)) : 0.0F;
                const float lhat_raw = lambda - model_noise2;
                const float lhat = lhat_raw > 0.0F ? lhat_raw : 0.0F;
                s_lhat[m * kMaxRank + j] = lhat;
                s_trace_term[m * kMaxRank + j] = lhat > 0.0F ? lhat / (lhat + estimate_noise2) : 0.0F;
                s_volume_term[m * kMaxRank + j] =
                    lhat > 0.0F ? log1p_via_double(lhat / membership_noise2) : 0.0F;
            }
        }
        __syncthreads();
        if (!equal && tid < models) {
            const int m = tid;
            const int rank_used = modelled ? R : 0;
            float trace = 0.0F;
            float volume = 0.0F;
            for (int j = 0; j < rank_used; ++j) {
                if (s_lhat[m * kMaxRank + j] > 0.0F) {
                    trace = trace + s_trace_term[m * kMaxRank + j];
                    volume = volume + s_volume_term[m * kMaxRank + j];
                }
            }
            const float mean_term = static_cast<float>(md) / static_cast<float>(rows);
            const float tau = (d.weight_epsilon + mean_term) + trace;
            s_scalar[4 * m + 0] = d.weight_alpha == 0.0F ? 0.0F : -d.weight_alpha * log_via_double(tau);
            s_scalar[4 * m + 1] = volume / static_cast<float>(md);
        }
        __syncthreads();

        // Aggregation weights (accumulate_aggregation_model core.cpp:326-412,
        // finalize_aggregation_weights :276-294).
        const bool weights_need_scores = !equal && d.weight_beta != 0.0F;
        const bool cached = L.proj_cache != 0 && weights_need_scores && modelled;
        int model_count = 0;
        if (!equal) {
            if (d.weight_alpha == 0.0F && d.weight_beta == 0.0F) {
                model_count = models;
            } else if (!weights_need_scores) {
                for (int r = tid; r < similar; r += kModelThreads) {
                    float lw = s_logw[r];
                    for (int m = 0; m < models; ++m) {
                        lw = lw + s_scalar[4 * m + 0];
                    }
                    s_logw[r] = lw;
                }
                model_count = models;
            } else {
                const int rank_used = modelled ? R : 0;
                const float score_scale = sqrtf(2.0F * static_cast<float>(md));
                float local_min[kMaxChannels];
#pragma unroll
                for (int m = 0; m < kMaxChannels; ++m) {
                    local_min[m] = INFINITY;
                }
                for (int r = team; r < similar; r += kTeams) {
                    const size_t base = patch_base(d, s_pos[r]);
                    for (int m = 0; m < models; ++m) {
                        const int lane_slot = m * row + lane * pitch;
                        float x[kS];
                        float xt;
                        load_centered(x, xt, member_video, base, tab_in + lane_slot,
                                      member_mean_all + lane_slot, chunks, tail_lane);
                        const float norm2 = team_norm2(x, xt, chunks, tail, lane);
                        float distance = 0.0F;
                        float projected = 0.0F;
                        for (int j = 0; j < rank_used; ++j) {
                            float b[kS];
                            float bt;
                            load_slots(b, bt, s_basis + (m * R + j) * row + lane * pitch, chunks, tail_lane);
                            const float projection = team_dot(x, xt, b, bt, chunks, tail, lane);
                            if (cached && lane == 0) {
                                s_proj[(r * models + m) * R + j] = projection;
                            }
                            const float projection2 = projection * projection;
                            distance = distance +
                                       projection2 / (s_lhat[m * kMaxRank + j] + membership_noise2);
                            projected = projected + projection2;
                        }
                        const float residual = norm2 - projected > 0.0F ? norm2 - projected : 0.0F;
                        distance = distance + residual / membership_noise2;
                        const float score = (distance - static_cast<float>(md)) / score_scale;
                        if (lane == 0) {
                            s_score[m * d.cap + r] = score;
                        }
#pragma unroll
                        for (int mm = 0; mm < kMaxChannels; ++mm) {
                            if (mm == m) {
                                local_min[mm] = fminf(local_min[mm], score);
                            }
                        }
                    }
                }
#pragma unroll
                for (int m = 0; m < kMaxChannels; ++m) {
                    if (m < models) {
                        const float v = wave_min(local_min[m]);
                        if ((tid & 63) == 0) {
                            s_red[m][tid >> 6] = v;
                        }
                    }
                }
                __syncthreads();
                for (int r = tid; r < similar; r += kModelThreads) {
                    float lw = s_logw[r];
                    for (int m = 0; m < models; ++m) {
                        float min_score = s_red[m][0];
                        for (int w = 1; w < kModelThreads / 64; ++w) {
                            min_score = fminf(min_score, s_red[m][w]);
                        }
                        const float excess = s_score[m * d.cap + r] - min_score;
                        const float membership = (excess > 0.0F ? excess : 0.0F) + s_scalar[4 * m + 1];
                        const float delta = s_scalar[4 * m + 0] - (d.weight_beta * membership);
                        lw = lw + delta;
                    }
                    s_logw[r] = lw;
                }
                model_count = models;
            }
        }
        if (d.weights_enabled) {
            if (model_count == 0) {
                const float tau = d.weight_epsilon + (static_cast<float>(md) / static_cast<float>(rows));
                const float log_group = d.weight_alpha == 0.0F ? 0.0F : -d.weight_alpha * log_via_double(tau);
                for (int r = tid; r < similar; r += kModelThreads) {
                    s_logw[r] = s_logw[r] + log_group;
                }
                model_count = 1;
            }
            const float inv_models = 1.0F / static_cast<float>(model_count);
            for (int r = tid; r < similar; r += kModelThreads) {
                float value = s_logw[r] * inv_models;
                value = value < -80.0F ? -80.0F : (value > 80.0F ? 80.0F : value);
                s_w[r] = exp_via_double(value);
            }
        }
        __syncthreads();
        if (weights_out != nullptr) {
            for (int r = tid; r < similar; r += kModelThreads) {
                weights_out[list + r] = s_w[r];
            }
        }

        // Reconstruction (project_bayes_estimate core.cpp:578-614) and the
        // weighted scatter of every sample.
        const float* center_all = (final_stage && flat) ? s_mean_b : s_mean_n;
        for (int r = team; r < similar; r += kTeams) {
            const float w = s_w[r];
            const uint32_t p = s_pos[r];
            const int pf = position_frame(p);
            const size_t base = patch_base(d, p);
            double* cell_base =
                acc.data + (static_cast<size_t>(pf) * (d.channels + 1) * d.height + position_y(p)) * d.width +
                position_x(p);
            const int slot_base = pf - af + d.search_bwd;
            for (int m = 0; m < models; ++m) {
                const int lane_slot = m * row + lane * pitch;
                const int* t_in = tab_in + lane_slot;
                const int* t_acc = tab_acc + lane_slot;
                const int* t_meta = tab_meta + lane_slot;
                const float* t_win = tab_win + lane_slot;
                const float* center = center_all + lane_slot;
                float out[kS];
                float out_t = 0.0F;
                if (equal) {
#pragma unroll
                    for (int k = 0; k < kS; ++k) {
                        out[k] = k < chunks ? noisy[base + t_in[k]] : 0.0F;
                    }
                    if (tail_lane) {
                        out_t = noisy[base + t_in[chunks]];
                    }
                } else if (!modelled) {
#pragma unroll
                    for (int k = 0; k < kS; ++k) {
                        out[k] = k < chunks ? center[k] : 0.0F;
                    }
                    if (tail_lane) {
                        out_t = center[chunks];
                    }
                } else {
                    float y[kS] = {};
                    float yt = 0.0F;
                    if (!cached) {
                        load_centered(y, yt, noisy, base, t_in, center, chunks, tail_lane);
                    }
#pragma unroll
                    for (int k = 0; k < kS; ++k) {
                        out[k] = 0.0F;
                    }
                    // Per element the terms arrive in basis order, fused over
                    // the vector slots and rounded product-then-sum in the tail.
                    for (int j = 0; j < R; ++j) {
                        const float coefficient = s_coef[m * kMaxRank + j];
                        if (coefficient == 0.0F) {
                            continue;
                        }
                        float b[kS];
                        float bt;
                        load_slots(b, bt, s_basis + (m * R + j) * row + lane * pitch, chunks, tail_lane);
                        const float projection = cached ? s_proj[(r * models + m) * R + j]
                                                        : team_dot(y, yt, b, bt, chunks, tail, lane);
                        const float sp = coefficient * projection;
#pragma unroll
                        for (int k = 0; k < kS; ++k) {
                            if (k < chunks) {
                                out[k] = fmaf(sp, b[k], out[k]);
                            }
                        }
                        if (tail_lane) {
                            const float product = sp * bt;
                            out_t = out_t + product;
                        }
                    }
#pragma unroll
                    for (int k = 0; k < kS; ++k) {
                        if (k < chunks) {
                            out[k] = out[k] + center[k];
                        }
                    }
                    if (tail_lane) {
                        out_t = out_t + center[chunks];
                    }
                }
#pragma unroll
                for (int k = 0; k < kS; ++k) {
                    if (k < chunks) {
                        scatter(d, cell_base, slot_base, t_meta[k], t_acc[k], t_win[k], out[k], w);
                    }
                }
                if (tail_lane) {
                    scatter(d, cell_base, slot_base, t_meta[chunks], t_acc[chunks], t_win[chunks], out_t, w);
                }
            }
        }
    }
}

}  // namespace

size_t filter_shared_bytes(const StageDesc& d) {
    return static_cast<size_t>(filter_layout(d).words) * sizeof(float);
}

hipError_t launch_filter(const StageDesc& d, const float* noisy, const float* basic,
                         const uint32_t* anchors, const int* counts,
                         const uint32_t* positions, const int* group_list,
                         const int* group_count_dev, int group_capacity,
                         int problem_stride, const int* group_flags, const int* problem_dims,
                         const float* means, const float* eigenvalues, const float* vectors,
                         double* accumulators, float* weights_out, int* work_counter,
                         hipStream_t stream) {
    (void)problem_dims;
    if (group_capacity <= 0) {
        return hipSuccess;
    }
    const GroupRefs refs{anchors, counts, positions, group_list};
    const Accumulators acc{accumulators};
    const FilterLayout L = filter_layout(d);
    const size_t shared = static_cast<size_t>(L.words) * sizeof(float);
    int device = 0;
    int units = 0;
    hipError_t status = hipGetDevice(&device);
    if (status == hipSuccess) {
        status = hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device);
    }
    if (status != hipSuccess) {
        return status;
    }
    const int slots = filter_template_slots(d);
    const bool exact = d.model_dim == kLanes * slots;
#define VNLBHIP_FILTER(KS, EXACT)                                                                 \
    do {                                                                                          \
        int per_unit = 0;                                                                         \
        status = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_unit, filter_kernel<KS, EXACT>, \
                                                              kModelThreads, shared);             \
        if (status != hipSuccess) {                                                               \
            return status;                                                                        \
        }                                                                                         \
        const long long resident = static_cast<long long>(per_unit > 0 ? per_unit : 1) * units;   \
        const unsigned blocks =                                                                   \
            static_cast<unsigned>(resident < group_capacity ? resident : group_capacity);         \
        hipLaunchKernelGGL((filter_kernel<KS, EXACT>), dim3(blocks), dim3(kModelThreads), shared, \
                           stream, d, L, noisy, basic, refs, group_count_dev,                     \
                           problem_stride, group_flags, means, eigenvalues, vectors, acc,         \
                           weights_out, work_counter);                                            \
    } while (0)
#define VNLBHIP_FILTER_SLOTS(KS)          \
    if (exact) {                          \
        VNLBHIP_FILTER(KS, true);         \
    } else {                              \
        VNLBHIP_FILTER(KS, false);        \
    }
    switch (slots) {
        case 4: VNLBHIP_FILTER_SLOTS(4); break;
        case 8: VNLBHIP_FILTER_SLOTS(8); break;
        case 16: VNLBHIP_FILTER_SLOTS(16); break;
        case 24: VNLBHIP_FILTER_SLOTS(24); break;
        case 32: VNLBHIP_FILTER_SLOTS(32); break;
        default: VNLBHIP_FILTER_SLOTS(64); break;
    }
#undef VNLBHIP_FILTER_SLOTS
#undef VNLBHIP_FILTER
    return hipGetLastError();
}

}  // namespace vnlbhip
