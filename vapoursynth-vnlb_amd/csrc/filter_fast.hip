// Filter stage, MFMA form, for dual-PCA groups: model rows B <= 16 below a
// model dimension that is a multiple of 32, 8x8 patches, one model (coupled
// channels or a single plane), at most 64 retained samples.
//
// With the centered model rows Xc (B x D), the Gram G = Xc Xc^T / B, its
// eigenpairs (lambda_j, u_j) and the dual basis
//     b_j = Xc^T u_j / sqrt(B lambda_j)
// (map_dual_eigenvectors_to_basis, linalg.cpp:552-595, normalises
// Xc^T u_j / sqrt(lambda_j), whose norm is sqrt(B)), every quantity of the
// filter follows from the cross-Gram C = Y Xc^T of the rows Y that are
// projected:
//     projections  p_rj = sum_s C[r][s] u_j[s] / sqrt(B lambda_j)
//     estimate     y^_r = center + sum_s W[r][s] Xc[s],
//                  W[r][s] = sum_j p_rj coef_j u_j[s] / sqrt(B lambda_j)
// so no basis vector of the model dimension is formed.  C is a
// 16 x 16 x D product on the matrix cores in bf16x3 (hi*hi + hi*lo + lo*hi,
// fp32 accumulation), its operands read straight from the clip, one 8-sample
// patch row per lane.  The weights, coefficients and scatter are those of
// the exact filter (filter.hip).  One wavefront owns one group; its
// contributions are combined in the wavefront's LDS tile (one channel and
// one output frame of the group's window at a time) and reach the
// accumulators once per touched cell: the retained patches overlap, so this
// takes about a third of the per-sample atomics.
#include "common.h"
#include "device_util.h"
#include "group.h"
#include "kernels.h"

#include <float.h>
#include <hip/hip_bf16.h>

#include <type_traits>

namespace vnlbhip {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int kWaves = kModelThreads / 64;
constexpr int kFastRows = 64;   // retained samples per group
constexpr int kFastModel = 16;  // model rows and rank

// Per-wavefront LDS region (4-byte words).
struct WaveLayout {
    int gram;   // [16][kFastRows] C[r][s] at s * kFastRows + r; then W, row-major in sorted order
    int norm;   // [kFastRows] squared norms of the member rows
    int u;      // [16][16] eigenvectors u_j[s]
    int pos;    // [kFastRows]
    int sorted; // per sample in source-slot order: [kFastRows] position, tile cell, weight
    int begin;  // [kMaxTemporal + 8] first sorted sample of each source slot
    int tile;   // [side][side][2] value plane and weight plane of one output frame
    int words;
};

// Side of the aggregation tile: the search window plus a patch.
__host__ __device__ inline int tile_side(const StageDesc& d) {
    return d.search_window + d.patch - 1;
}

__host__ __device__ inline WaveLayout wave_layout(const StageDesc& d) {
    WaveLayout L{};
    int at = 0;
    L.gram = at;
    at += 16 * kFastRows;
    L.norm = at;
    at += kFastRows;
    L.u = at;
    at += 16 * 16;
    L.pos = at;
    at += kFastRows;
    L.sorted = at;
    at += 3 * kFastRows;
    L.begin = at;
    at += (kMaxTemporal + 8 + 3) / 4 * 4;  // slots + patch frames <= kMaxTemporal + 4
    L.tile = at;
    at += 2 * tile_side(d) * tile_side(d);
    L.words = (at + 3) / 4 * 4;
    return L;
}

// Stage-constant element table: input offset of every model element.
__host__ __device__ inline int table_words(int dim) {
    return (dim + 3) / 4 * 4;
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Eight consecutive floats from a 4-byte aligned address (dwordx4 loads).
__device__ inline void load8(const float* p, float (&v)[8]) {
    float4 lo;
    float4 hi;
    __builtin_memcpy(&lo, p, sizeof(lo));
    __builtin_memcpy(&hi, p + 4, sizeof(hi));
    v[0] = lo.x;
    v[1] = lo.y;
    v[2] = lo.z;
    v[3] = lo.w;
    v[4] = hi.x;
    v[5] = hi.y;
    v[6] = hi.z;
    v[7] = hi.w;
}

__device__ inline void split8(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const __bf16 h = static_cast<__bf16>(x[v]);
        hi[v] = h;
        lo[v] = static_cast<__bf16>(x[v] - static_cast<float>(h));
    }
}

// C[r][s] = <rows_r - rows_mean, model_s - xmean> for r < nrows, s < B, and
// norm[r] = |rows_r - rows_mean|^2, into LDS (C column-major, s * kFastRows
// + r).  rows and model share the sample positions s_pos (model rows are the
// first B samples).  One k step of 32 model elements reads every operand of
// the four 16-row tiles at once: 8 consecutive samples of one patch row per
// lane, the model operand and the means shared by the tiles.
__device__ inline void cross_gram(const StageDesc& d, const float* __restrict__ rows_video,
                                  const float* rows_mean, const float* __restrict__ model_video,
                                  const float* xmean, const uint32_t* s_pos, int nrows, int B,
                                  const int* tab_in, float* s_gram, float* s_norm, int lane) {
    constexpr int kRowTiles = kFastRows / 16;
    const int md = d.model_dim;
    const int kb = lane >> 4;
    const int col = lane & 15;
    const int row_tiles = (nrows + 15) / 16;
    const size_t model_base = col < B ? patch_base(d, s_pos[col]) : 0;
    size_t row_base[kRowTiles];
#pragma unroll
    for (int mt = 0; mt < kRowTiles; ++mt) {
        const int r = 16 * mt + col;
        row_base[mt] = r < nrows ? patch_base(d, s_pos[r]) : 0;
    }
    floatx4 acc[kRowTiles];
    float norm[kRowTiles];
#pragma unroll
    for (int mt = 0; mt < kRowTiles; ++mt) {
        acc[mt] = floatx4{0.0F, 0.0F, 0.0F, 0.0F};
        norm[mt] = 0.0F;
    }
    for (int q = 0; q * 32 < md; ++q) {
        const int e0 = 32 * q + 8 * kb;
        const int off = tab_in[e0];  // 8 samples of one patch row from here
        float b[8];
        float bm[8];
        float am[8];
        float a[kRowTiles][8];
        load8(xmean + e0, bm);
        load8(rows_mean + e0, am);
        if (col < B) {
            load8(model_video + model_base + off, b);
        }
#pragma unroll
        for (int mt = 0; mt < kRowTiles; ++mt) {
            if (mt < row_tiles && 16 * mt + col < nrows) {
                load8(rows_video + row_base[mt] + off, a[mt]);
            }
        }
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            b[v] = col < B ? b[v] - bm[v] : 0.0F;
        }
        bf16x8 b_hi;
        bf16x8 b_lo;
        split8(b, b_hi, b_lo);
#pragma unroll
        for (int mt = 0; mt < kRowTiles; ++mt) {
            if (mt < row_tiles) {
                const bool row_ok = 16 * mt + col < nrows;
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    a[mt][v] = row_ok ? a[mt][v] - am[v] : 0.0F;
                    norm[mt] = fmaf(a[mt][v], a[mt][v], norm[mt]);
                }
                bf16x8 a_hi;
                bf16x8 a_lo;
                split8(a[mt], a_hi, a_lo);
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, b_hi, acc[mt], 0, 0, 0);
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, b_lo, acc[mt], 0, 0, 0);
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo, b_hi, acc[mt], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int mt = 0; mt < kRowTiles; ++mt) {
        if (mt < row_tiles) {
            // acc[v] = C[16 mt + 4 kb + v][col]
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = 16 * mt + 4 * kb + v;
                if (rr < nrows) {
                    s_gram[col * kFastRows + rr] = acc[mt][v];
                }
            }
            float n = norm[mt];
            n = n + __shfl_xor(n, 16, 64);
            n = n + __shfl_xor(n, 32, 64);
            const int r = 16 * mt + col;
            if (kb == 0 && r < nrows) {
                s_norm[r] = n;
            }
        }
    }
}

// Adds one contribution to the wavefront's tile cell (value plane; weight
// plane too when with_weight).  The 64 lanes of a call address 64 distinct
// cells (the pixels of one patch) and the wavefront's LDS operations
// complete in issue order, so a plain read-modify-write is exact, in the
// order of the calls, without LDS atomics.
__device__ inline void tile_add(float* s_tile, int cell, float value, float weight, bool with_weight) {
    if (with_weight) {
        float2* t2 = reinterpret_cast<float2*>(s_tile) + cell;
        float2 t = *t2;
        t.x = t.x + value;
        t.y = t.y + weight;
        *t2 = t;
    } else {
        s_tile[2 * cell] = s_tile[2 * cell] + value;
    }
}

// sum_s c[s] u[s] in ascending s, u a 16-float LDS row read as four float4.
__device__ inline float dot16(const float (&c)[16], const float* u) {
    const float4* u4 = reinterpret_cast<const float4*>(u);
    float p = 0.0F;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const float4 uv = u4[h];
        p = fmaf(c[4 * h + 0], uv.x, p);
        p = fmaf(c[4 * h + 1], uv.y, p);
        p = fmaf(c[4 * h + 2], uv.z, p);
        p = fmaf(c[4 * h + 3], uv.w, p);
    }
    return p;
}

// kPt: frames of a patch (64-element chunks of one channel's model row).
template <int kPt>
__global__ __launch_bounds__(kModelThreads, 3) void filter_fast_kernel(
    StageDesc d, const float* __restrict__ noisy, const float* __restrict__ basic, GroupRefs refs,
    const int* __restrict__ group_count_dev, int problem_stride, const int* __restrict__ group_flags,
    const float* __restrict__ means, const float* __restrict__ eigenvalues,
    const float* __restrict__ vectors, const float* __restrict__ grams, float* __restrict__ acc,
    float* __restrict__ weights_out, int* __restrict__ work_counter) {
    extern __shared__ float4 s_dyn4[];
    float* s_dyn = reinterpret_cast<float*>(s_dyn4);
    const int dim = d.dim;  // == model dimension (one model)
    const int twords = table_words(dim);
    int* tab_in = reinterpret_cast<int*>(s_dyn);
    const WaveLayout W = wave_layout(d);
    const int side = tile_side(d);
    const int cells = side * side;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    float* s_wave = s_dyn + twords + wave * W.words;
    float* s_gram = s_wave + W.gram;
    float* s_norm = s_wave + W.norm;
    float* s_u = s_wave + W.u;
    uint32_t* s_pos = reinterpret_cast<uint32_t*>(s_wave + W.pos);
    uint32_t* s_kpos = reinterpret_cast<uint32_t*>(s_wave + W.sorted);
    int* s_kcell = reinterpret_cast<int*>(s_wave + W.sorted + kFastRows);
    float* s_kw = s_wave + W.sorted + 2 * kFastRows;
    int* s_begin = reinterpret_cast<int*>(s_wave + W.begin);
    float* s_tile = s_wave + W.tile;
    // Element e = c * patch_dim + 64 t + lane of an 8x8 patch sits at
    // (lane & 7, lane >> 3) for every channel and patch frame.
    const int lane_cell = (lane >> 3) * side + (lane & 7);
    const float lane_window = d.window[lane];

    for (int e = tid; e < dim; e += kModelThreads) {
        const int c = e / d.patch_dim;
        const int wi = e - c * d.patch_dim;
        const int t = wi / d.patch_area;
        const int yx = wi - t * d.patch_area;
        const int y = yx / d.patch;
        const int x = yx - y * d.patch;
        tab_in[e] = ((t * d.channels + c) * d.height + y) * d.width + x;
    }
    for (int i = lane; i < 2 * cells; i += 64) {
        s_tile[i] = 0.0F;  // every flush leaves the tile cleared
    }
    __syncthreads();

    const bool final_stage = d.stage == kFinal;
    const int md = d.model_dim;
    const int R = d.rank;
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
    const float* model_video = final_stage ? basic : noisy;

    const int group_count = *group_count_dev;
    // Groups are taken one request per wavefront from this XCD's contiguous
    // eighth of the list (then from the others'), so the groups in flight on
    // an XCD are neighbours that share its L2; the next index is requested
    // while the current group runs.
    const int home = blockIdx.x % kXcds;
    int visited = 0;
    int request = 0;
    if (lane == 0) {
        request = xcd_next(work_counter, group_count, home, visited);
    }
    int following = 0;
    // A group's anchor, flags, retained count and positions (sample `lane`)
    // are loaded while the wave's previous group is processed, as vector
    // loads that nothing waits for until they are used.
    int next_a = 0;
    int next_flags = 0;
    int next_similar = 0;
    uint32_t next_anchor = 0;
    uint32_t next_pos = 0;
    const auto fetch_group = [&](int index) {
        int v = 0;
        asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(index));
        next_a = refs.group_list[v];
        next_flags = group_flags[v];
    };
    const auto fetch_members = [&]() {
        next_anchor = refs.anchors[next_a];
        next_similar = refs.counts[next_a];
        if (lane < d.cap) {
            next_pos = refs.positions[static_cast<size_t>(next_a) * d.cap + lane];
        }
    };
    int g = __builtin_amdgcn_readlane(request, 0);
    if (g < group_count) {
        fetch_group(g);
        fetch_members();
    }
    while (g < group_count) {
        if (lane == 0) {
            following = xcd_next(work_counter, group_count, home, visited);
        }
        wave_sync();  // the previous group is done with this wave's LDS
        const int a = __builtin_amdgcn_readfirstlane(next_a);
        const uint32_t anchor_p = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(next_anchor)));
        const int af = position_frame(anchor_p);
        const int similar = __builtin_amdgcn_readfirstlane(next_similar);
        const size_t list = static_cast<size_t>(a) * d.cap;
        const int flags = __builtin_amdgcn_readfirstlane(next_flags);
        const uint32_t group_pos = next_pos;
        const bool equal = (flags & kFlagEqual) != 0;
        const bool flat = (flags & kFlagFlat) != 0;
        const bool modelled = (flags & kFlagModelled) != 0;
        const int B = model_rows_of(d, similar);
        const bool fit = modelled && !equal;

        uint32_t my_pos = 0;
        if (lane < similar) {
            my_pos = group_pos;
            s_pos[lane] = my_pos;
        }
        // Samples sorted by source slot q = frame - af + search_bwd over
        // [1 - kPt, slots): the sample feeds output slot q + t through patch
        // frame t, and output slots outside [0, slots) are dropped (scatter).
        // Lane r keeps its sorted index; the sorted tables hold the position
        // and the tile cell of the patch origin.
        int x_low = 0;
        int x_high = 0;
        int y_low = 0;
        int y_high = 0;
        spatial_range(d, position_x(anchor_p), d.width, x_low, x_high);
        spatial_range(d, position_y(anchor_p), d.height, y_low, y_high);
        const int sorted_slots = d.slots + kPt - 1;
        int my_k = -1;
        {
            const int my_slot = lane < similar ? position_frame(my_pos) - af + d.search_bwd + kPt - 1 : -1;
            int at = 0;
            for (int q = 0; q < sorted_slots; ++q) {
                const uint64_t members = __ballot(my_slot == q);
                if (lane == 0) {
                    s_begin[q] = at;
                }
                if (my_slot == q) {
                    my_k = at + __popcll(members & ((1ULL << lane) - 1));
                }
                at += __popcll(members);
            }
            if (lane == 0) {
                s_begin[sorted_slots] = at;
            }
            if (my_k >= 0) {
                s_kpos[my_k] = my_pos;
                s_kcell[my_k] = (position_y(my_pos) - y_low) * side + position_x(my_pos) - x_low;
            }
        }
        // Centre of the estimate rows and mean of the model / member rows.
        const float* g_mean = means + static_cast<size_t>(g) * 2 * dim;
        const float* center_of = (final_stage && flat) ? g_mean + dim : g_mean;
        const float* xmean_of = final_stage ? g_mean + dim : g_mean;
        // Per component j (lane j): eigenpair terms.
        float lambda = 0.0F;
        float inv = 0.0F;
        float coef = 0.0F;
        float lhat = 0.0F;
        float trace_term = 0.0F;
        float volume_term = 0.0F;
        if (fit && lane < R) {
            lambda = eigenvalues[static_cast<size_t>(g) * R + lane];
            if (!(lambda <= dual_floor)) {
                inv = 1.0F / sqrtf(static_cast<float>(B) * lambda);
            }
            const float sigma_basic2 = final_stage ? (beta * d.sigma_basic) * d.sigma_basic : sigma2;
            float variance = lambda;
            if (sigma_basic2 > 0.0F) {
                variance = variance - fminf(variance, sigma_basic2);
            }
            coef = variance > d.variance_threshold * sigma2 ? 1.0F / (1.0F + (sigma2 / variance)) : 0.0F;
            const float lhat_raw = lambda - model_noise2;
            lhat = lhat_raw > 0.0F ? lhat_raw : 0.0F;
            trace_term = lhat > 0.0F ? lhat / (lhat + estimate_noise2) : 0.0F;
            volume_term = lhat > 0.0F ? log1p_via_double(lhat / membership_noise2) : 0.0F;
        }
        if (fit) {
            // U zero-padded to 16 x 16 so rows are read as four float4.
            for (int i = lane; i < 16 * 16; i += 64) {
                const int j = i >> 4;
                const int s = i & 15;
                s_u[i] = (j < R && s < B) ? vectors[(static_cast<size_t>(g) * R + j) * problem_stride + s] : 0.0F;
            }
        }
        // accumulate_aggregation_model (core.cpp:326-412): group terms, in
        // component order.
        float log_group = 0.0F;
        float volume_md = 0.0F;
        if (!equal) {
            const int rank_used = modelled ? R : 0;
            float trace = 0.0F;
            float volume = 0.0F;
            for (int j = 0; j < rank_used; ++j) {
                const float lh = __shfl(lhat, j, 64);
                const float tt = __shfl(trace_term, j, 64);
                const float vt = __shfl(volume_term, j, 64);
                if (lh > 0.0F) {
                    trace = trace + tt;
                    volume = volume + vt;
                }
            }
            const float mean_term = static_cast<float>(md) / static_cast<float>(B);
            const float tau = (d.weight_epsilon + mean_term) + trace;
            log_group = d.weight_alpha == 0.0F ? 0.0F : -d.weight_alpha * log_via_double(tau);
            volume_md = volume / static_cast<float>(md);
        }
        wave_sync();
        // The wave's next group: its index, anchor and flags are requested now.
        const int g_next = __builtin_amdgcn_readlane(following, 0);
        if (g_next < group_count) {
            fetch_group(g_next);
        }

        // Aggregation weights (finalize_aggregation_weights core.cpp:276-294).
        const bool weights_need_scores = !equal && d.weight_beta != 0.0F;
        const bool estimate_is_member = !final_stage;  // Basic: same rows, same centre
        float p_est[kFastModel];
#pragma unroll
        for (int j = 0; j < kFastModel; ++j) {
            p_est[j] = 0.0F;
        }
        bool have_estimate = false;
        float logw = 0.0F;
        int model_count = 0;
        if (!equal) {
            if (d.weight_alpha == 0.0F && d.weight_beta == 0.0F) {
                model_count = 1;
            } else if (!weights_need_scores) {
                logw = log_group;
                model_count = 1;
            } else {
                const int rank_used = modelled ? R : 0;
                if (modelled && similar == B) {
                    // Every member is a model row: C = B G with the Gram the
                    // model kernel built from the same centered rows.
                    const float* gram = grams + static_cast<size_t>(g) * problem_stride * problem_stride;
                    const float scale_b = static_cast<float>(B);
                    for (int i = lane; i < B * B; i += 64) {
                        const int r = i / B;
                        const int s = i - r * B;
                        s_gram[s * kFastRows + r] = scale_b * gram[r * problem_stride + s];
                    }
                    if (lane < B) {
                        s_norm[lane] = scale_b * gram[lane * problem_stride + lane];
                    }
                } else {
                    cross_gram(d, model_video, xmean_of, model_video, xmean_of, s_pos, similar, B, tab_in,
                               s_gram, s_norm, lane);
                }
                wave_sync();
                float score = INFINITY;
                if (lane < similar) {
                    float distance = 0.0F;
                    float projected = 0.0F;
                    float crow[kFastModel];
#pragma unroll
                    for (int s = 0; s < kFastModel; ++s) {
                        crow[s] = s < B ? s_gram[s * kFastRows + lane] : 0.0F;
                    }
                    // Unrolled and branch-free (U is zero-padded), so the
                    // rows' LDS reads are issued ahead of the products;
                    // components past rank_used add exact zeros.
#pragma unroll
                    for (int j = 0; j < kFastModel; ++j) {
                        const float inv_j = __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                                          __builtin_bit_cast(int, inv), j));
                        const float lh = __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                                       __builtin_bit_cast(int, lhat), j));
                        const bool used = j < rank_used;
                        const float dot = dot16(crow, s_u + j * 16);
                        const float p = used && inv_j != 0.0F ? dot * inv_j : 0.0F;
                        if (estimate_is_member) {
                            p_est[j] = p;
                        }
                        const float p2 = p * p;
                        distance = distance + (used ? p2 / (lh + membership_noise2) : 0.0F);
                        projected = projected + (used ? p2 : 0.0F);
                    }
                    const float norm2 = s_norm[lane];
                    const float residual = norm2 - projected > 0.0F ? norm2 - projected : 0.0F;
                    distance = distance + residual / membership_noise2;
                    score = (distance - static_cast<float>(md)) / sqrtf(2.0F * static_cast<float>(md));
                }
                have_estimate = estimate_is_member && fit;
                const float min_score = wave_min(score);
                const float excess = score - min_score;
                const float membership = (excess > 0.0F ? excess : 0.0F) + volume_md;
                logw = log_group - (d.weight_beta * membership);
                model_count = 1;
            }
        }
        float w = 1.0F;
        if (d.weights_enabled) {
            if (model_count == 0) {
                const float tau = d.weight_epsilon + (static_cast<float>(md) / static_cast<float>(B));
                logw = d.weight_alpha == 0.0F ? 0.0F : -d.weight_alpha * log_via_double(tau);
                model_count = 1;
            }
            float value = logw / static_cast<float>(model_count);
            value = value < -80.0F ? -80.0F : (value > 80.0F ? 80.0F : value);
            w = exp_via_double(value);
        }
        if (lane < similar) {
            if (my_k >= 0) {
                s_kw[my_k] = w;
            }
            if (weights_out != nullptr) {
                weights_out[list + lane] = w;
            }
        }

        // Projections of the estimate rows and W = P diag(coef / sqrt(B lambda)) U^T:
        // lane r keeps row r of W.
        float wrow[kFastModel];
#pragma unroll
        for (int s = 0; s < kFastModel; ++s) {
            wrow[s] = 0.0F;
        }
        if (fit) {
            if (!have_estimate) {
                wave_sync();
                cross_gram(d, noisy, center_of, model_video, xmean_of, s_pos, similar, B, tab_in, s_gram,
                           s_norm, lane);
                wave_sync();
                if (lane < similar) {
                    float crow[kFastModel];
#pragma unroll
                    for (int s = 0; s < kFastModel; ++s) {
                        crow[s] = s < B ? s_gram[s * kFastRows + lane] : 0.0F;
                    }
#pragma unroll
                    for (int j = 0; j < kFastModel; ++j) {
                        const float inv_j = __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                                          __builtin_bit_cast(int, inv), j));
                        const float dot = dot16(crow, s_u + j * 16);  // branch-free, as above
                        p_est[j] = j < R && inv_j != 0.0F ? dot * inv_j : 0.0F;
                    }
                }
            }
            wave_sync();
            if (lane < similar) {
#pragma unroll
                for (int j = 0; j < kFastModel; ++j) {
                    const float scale_j = __builtin_bit_cast(
                        float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, coef * inv), j));
                    if (j < R && scale_j != 0.0F) {
                        const float pj = p_est[j] * scale_j;
                        const float4* u4 = reinterpret_cast<const float4*>(s_u + j * 16);
#pragma unroll
                        for (int h = 0; h < kFastModel / 4; ++h) {
                            const float4 uv = u4[h];
                            wrow[4 * h + 0] = fmaf(pj, uv.x, wrow[4 * h + 0]);
                            wrow[4 * h + 1] = fmaf(pj, uv.y, wrow[4 * h + 1]);
                            wrow[4 * h + 2] = fmaf(pj, uv.z, wrow[4 * h + 2]);
                            wrow[4 * h + 3] = fmaf(pj, uv.w, wrow[4 * h + 3]);
                        }
                    }
                }
#pragma unroll
                for (int s = 0; s < kFastModel; ++s) {
                    wrow[s] = s < B ? wrow[s] : 0.0F;
                }
                if (my_k >= 0) {
                    // Row-major in sorted order for the broadcast reads below.
                    float4* w4 = reinterpret_cast<float4*>(s_gram + my_k * kFastModel);
#pragma unroll
                    for (int h = 0; h < kFastModel / 4; ++h) {
                        w4[h] = float4{wrow[4 * h + 0], wrow[4 * h + 1], wrow[4 * h + 2], wrow[4 * h + 3]};
                    }
                }
            }
        }
        wave_sync();
        if (g_next < group_count) {
            fetch_members();
        }
        // The sorted tables in registers, sample k in lane k, so the loops
        // below take them with v_readlane instead of LDS broadcasts.
        const int sorted_count = __builtin_amdgcn_readfirstlane(s_begin[sorted_slots]);
        int kcell_l = 0;
        float kw_l = 0.0F;
        int ky_l = 0;  // tile row and column of the sample's patch origin
        int kx_l = 0;
        if (lane < sorted_count) {
            kcell_l = s_kcell[lane];
            kw_l = s_kw[lane];
            ky_l = kcell_l / side;
            kx_l = kcell_l - ky_l * side;
        }

        // Estimate and weighted aggregation: per channel and output frame of
        // the group's window, contributions are combined in the tile (value
        // plane; weight plane for channel 0) and flushed once per cell.
        const size_t plane = static_cast<size_t>(d.width) * d.height;
        // The lane's model-row samples of a channel, every load issued at
        // once (rows beyond B read row 0 and are zeroed when centered); the
        // next channel's are issued after the last estimates of this one,
        // so they arrive during its last flush.
        float x[kPt][kFastModel];
        float center[kPt];
        float xmean[kPt];
        const auto issue_loads = [&](int c) {
#pragma unroll
            for (int tt = 0; tt < kPt; ++tt) {
                const int e = c * d.patch_dim + 64 * tt + lane;
                center[tt] = center_of[e];
                xmean[tt] = 0.0F;
#pragma unroll
                for (int s2 = 0; s2 < kFastModel; ++s2) {
                    x[tt][s2] = 0.0F;
                }
            }
            if (fit) {
#pragma unroll
                for (int tt = 0; tt < kPt; ++tt) {
                    const int e = c * d.patch_dim + 64 * tt + lane;
                    const int in_off = tab_in[e];
                    xmean[tt] = xmean_of[e];
#pragma unroll
                    for (int s2 = 0; s2 < kFastModel; ++s2) {
                        const uint32_t ps = s_pos[s2 < B ? s2 : 0];
                        x[tt][s2] = model_video[patch_base(d, ps) + in_off];
                    }
                }
            }
        };
        // The last output frame that receives samples (its flush is the
        // one the next channel's loads overlap).
        int last_active = -1;
        for (int slot = d.slots - 1; slot >= 0 && last_active < 0; --slot) {
            if (__builtin_amdgcn_readfirstlane(s_begin[slot]) != __builtin_amdgcn_readfirstlane(s_begin[slot + kPt])) {
                last_active = slot;
            }
        }
        issue_loads(0);
        for (int c = 0; c < d.channels; ++c) {
            if (fit) {
#pragma unroll
                for (int tt = 0; tt < kPt; ++tt) {
#pragma unroll
                    for (int s2 = 0; s2 < kFastModel; ++s2) {
                        x[tt][s2] = s2 < B ? x[tt][s2] - xmean[tt] : 0.0F;
                    }
                }
            }
            for (int slot = 0; slot < d.slots; ++slot) {
                // Samples of source slot slot - t feed this output frame
                // through patch frame t; s_begin is indexed by q + kPt - 1.
                const int k_first = __builtin_amdgcn_readfirstlane(s_begin[slot]);
                const int k_last = __builtin_amdgcn_readfirstlane(s_begin[slot + kPt]);
                if (k_first == k_last) {
                    continue;
                }
#pragma unroll
                for (int tt = 0; tt < kPt; ++tt) {
                    const int source_slot = slot - tt + kPt - 1;
                    const int k_begin = s_begin[source_slot];
                    const int k_end = s_begin[source_slot + 1];
                    if (equal) {
                        // Bypass: the patch itself (the only global loads of
                        // this loop, kept out of the modelled one).
                        const int in_off = tab_in[c * d.patch_dim + 64 * tt + lane];
                        for (int k = k_begin; k < k_end; ++k) {
                            const int cell = s_kcell[k] + lane_cell;
                            const float cw = s_kw[k] * lane_window;
                            const float value = noisy[patch_base(d, s_kpos[k]) + in_off];
                            tile_add(s_tile, cell, cw * value, cw, c == 0);
                        }
                        continue;
                    }
                    const int kb0 = __builtin_amdgcn_readfirstlane(k_begin);
                    const int ke0 = __builtin_amdgcn_readfirstlane(k_end);
                    // Modelled groups and the weight plane as compile-time
                    // cases, so the two estimates of a pair have no branch
                    // between them.
                    const auto run = [&](auto model_case, auto weight_case) {
                        constexpr bool kModelled = decltype(model_case)::value;
                        constexpr bool kWeight = decltype(weight_case)::value;
                        const auto estimate = [&](int k) -> float {
                            float value = center[tt];
                            if constexpr (kModelled) {
                                const float4* w4 = reinterpret_cast<const float4*>(s_gram + k * kFastModel);
                                const float4 wa = w4[0];
                                const float4 wb = w4[1];
                                const float4 wc = w4[2];
                                const float4 wd = w4[3];
                                float s0 = wa.x * x[tt][0];
                                float s1 = wc.x * x[tt][8];
                                s0 = fmaf(wa.y, x[tt][1], s0);
                                s1 = fmaf(wc.y, x[tt][9], s1);
                                s0 = fmaf(wa.z, x[tt][2], s0);
                                s1 = fmaf(wc.z, x[tt][10], s1);
                                s0 = fmaf(wa.w, x[tt][3], s0);
                                s1 = fmaf(wc.w, x[tt][11], s1);
                                s0 = fmaf(wb.x, x[tt][4], s0);
                                s1 = fmaf(wd.x, x[tt][12], s1);
                                s0 = fmaf(wb.y, x[tt][5], s0);
                                s1 = fmaf(wd.y, x[tt][13], s1);
                                s0 = fmaf(wb.z, x[tt][6], s0);
                                s1 = fmaf(wd.z, x[tt][14], s1);
                                s0 = fmaf(wb.w, x[tt][7], s0);
                                s1 = fmaf(wd.w, x[tt][15], s1);
                                value = value + (s0 + s1);
                            }
                            return value;
                        };
                        int k = kb0;
                        for (; k + 1 < ke0; k += 2) {
                            const int cell0 = __builtin_amdgcn_readlane(kcell_l, k) + lane_cell;
                            const int cell1 = __builtin_amdgcn_readlane(kcell_l, k + 1) + lane_cell;
                            const float cw0 = __builtin_bit_cast(
                                                  float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, kw_l), k)) *
                                              lane_window;
                            const float cw1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                                            __builtin_bit_cast(int, kw_l), k + 1)) *
                                              lane_window;
                            const float v0 = estimate(k);
                            const float v1 = estimate(k + 1);
                            tile_add(s_tile, cell0, cw0 * v0, cw0, kWeight);
                            tile_add(s_tile, cell1, cw1 * v1, cw1, kWeight);
                        }
                        if (k < ke0) {
                            const int cell = __builtin_amdgcn_readlane(kcell_l, k) + lane_cell;
                            const float cw = __builtin_bit_cast(
                                                 float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, kw_l), k)) *
                                             lane_window;
                            const float value = estimate(k);
                            tile_add(s_tile, cell, cw * value, cw, kWeight);
                        }
                    };
                    if (modelled) {
                        if (c == 0) {
                            run(std::true_type{}, std::true_type{});
                        } else {
                            run(std::true_type{}, std::false_type{});
                        }
                    } else {
                        if (c == 0) {
                            run(std::false_type{}, std::true_type{});
                        } else {
                            run(std::false_type{}, std::false_type{});
                        }
                    }
                }
                if (slot == last_active && c + 1 < d.channels) {
                    issue_loads(c + 1);
                }
                wave_sync();
                // Flush.  A global atomic costs one memory-side operation per
                // 64-byte line an instruction touches, whatever the element
                // size, so the accumulators are fp32 (16 per line) and the
                // tile is flushed over the bounding box of the patches that
                // fed this pass in blocks of 4 rows x 16 line-aligned pixels
                // (lane = row * 16 + pixel; aligned when the width is a
                // multiple of 16): every touched cell is read, cleared and
                // added once.
                const int frame = af - d.search_bwd + slot;
                float* value_plane = acc + (static_cast<size_t>(frame) * (d.channels + 1) + c) * plane;
                float* weight_plane = value_plane + static_cast<size_t>(d.channels - c) * plane;
                const bool feeds = lane >= k_first && lane < k_last;
                const int ty0 = __builtin_amdgcn_readfirstlane(wave_min_int(feeds ? ky_l : side));
                const int ty1 = __builtin_amdgcn_readfirstlane(wave_max_int(feeds ? ky_l : -1)) + 7;  // the 8x8 footprint
                const int tx0 = __builtin_amdgcn_readfirstlane(wave_min_int(feeds ? kx_l : side));
                const int tx1 = __builtin_amdgcn_readfirstlane(wave_max_int(feeds ? kx_l : -1)) + 7;
                const int gx_first = (x_low + tx0) & ~15;  // line-aligned image column
                const int gx_last = x_low + tx1;
                unsigned long long* tile64 = reinterpret_cast<unsigned long long*>(s_tile);
                for (int ty_b = ty0; ty_b <= ty1; ty_b += 4) {
                    const int ty = ty_b + (lane >> 4);
                    for (int gx_b = gx_first; gx_b <= gx_last; gx_b += 16) {
                        const int tx = gx_b + (lane & 15) - x_low;
                        unsigned long long taken = 0ULL;
                        if (ty <= ty1 && tx >= tx0 && tx <= tx1) {
                            unsigned long long* slot64 = tile64 + ty * side + tx;
                            taken = *slot64;
                            *slot64 = 0ULL;
                        }
                        if (taken != 0ULL) {
                            const float2 v = __builtin_bit_cast(float2, taken);
                            const size_t offset = static_cast<size_t>(y_low + ty) * d.width + x_low + tx;
                            unsafeAtomicAdd(value_plane + offset, v.x);
                            if (c == 0) {
                                unsafeAtomicAdd(weight_plane + offset, v.y);
                            }
                        }
                    }
                }
                wave_sync();
            }
        }
        g = g_next;
    }
}

}  // namespace

bool fast_filter_supported(const StageDesc& d) {
    return d.use_dual != 0 && d.models == 1 && d.max_rows <= kFastModel && d.rank <= kFastModel &&
           d.patch == 8 && d.patch_time >= 1 && d.patch_time <= 4 && d.cap <= kFastRows &&
           d.slots <= kMaxTemporal && d.fused_width == 0 && fast_filter_shared_bytes(d) <= 64 * 1024;
}

size_t fast_filter_shared_bytes(const StageDesc& d) {
    return (static_cast<size_t>(table_words(d.dim)) + static_cast<size_t>(kWaves) * wave_layout(d).words) *
           sizeof(float);
}

hipError_t launch_fast_filter(const StageDesc& d, const float* noisy, const float* basic,
                              const uint32_t* anchors, const int* counts, const uint32_t* positions,
                              const int* group_list, const int* group_count_dev, int group_capacity,
                              int problem_stride, const int* group_flags, const float* means,
                              const float* eigenvalues, const float* vectors, const float* grams,
                              float* accumulators, float* weights_out, int* work_counter,
                              hipStream_t stream) {
    if (group_capacity <= 0) {
        return hipSuccess;
    }
    const GroupRefs refs{anchors, counts, positions, group_list};
    const size_t shared = fast_filter_shared_bytes(d);
    int device = 0;
    int units = 0;
    hipError_t status = hipGetDevice(&device);
    if (status == hipSuccess) {
        status = hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device);
    }
    if (status != hipSuccess) {
        return status;
    }
    const long long wanted = (group_capacity + kWaves - 1) / kWaves;
#define VNLBHIP_FAST_FILTER(KPT)                                                                     \
    do {                                                                                             \
        int per_unit = 0;                                                                            \
        status = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_unit, filter_fast_kernel<KPT>,    \
                                                              kModelThreads, shared);                \
        if (status != hipSuccess) {                                                                  \
            return status;                                                                           \
        }                                                                                            \
        const long long resident = static_cast<long long>(per_unit > 0 ? per_unit : 1) * units;      \
        const unsigned blocks = static_cast<unsigned>(resident < wanted ? resident : wanted);        \
        hipLaunchKernelGGL(filter_fast_kernel<KPT>, dim3(blocks), dim3(kModelThreads), shared,       \
                           stream, d, noisy, basic, refs, group_count_dev, problem_stride, group_flags, \
                           means, eigenvalues, vectors, grams, accumulators, weights_out, work_counter);                    \
    } while (0)
    switch (d.patch_time) {
        case 1: VNLBHIP_FAST_FILTER(1); break;
        case 2: VNLBHIP_FAST_FILTER(2); break;
        case 3: VNLBHIP_FAST_FILTER(3); break;
        default: VNLBHIP_FAST_FILTER(4); break;
    }
#undef VNLBHIP_FAST_FILTER
    return hipGetLastError();
}

}  // namespace vnlbhip
