// Per-group Bayesian model and aggregation scatter.
//
//   prepare_kernel  flags (equal / flat), model means, centered model rows,
//                   Gram (dual PCA) or covariance matrices
//   jacobi_kernel   batched cyclic Jacobi eigen solve, several small
//                   problems per wavefront
//   filter_kernel   dual-basis map, Wiener coefficients, aggregation
//                   weights, reconstruction and the weighted scatter into
//                   per-output-frame double accumulators
//
// Semantics follow filter_vnlb_samples and its callees in the reference
// (src/core/core.cpp:184-871, src/linalg/linalg.cpp:153-595) with the
// 16-lane summation order of the CPU restatement.
#include "common.h"
#include "device_util.h"

#include <float.h>

namespace vnlbhip {
namespace {

constexpr int kModelThreads = 256;
constexpr int kTeams = kModelThreads / kLanes;  // 16-lane teams per block

struct GroupRefs {
    const uint32_t* anchors;
    const int* counts;
    const uint32_t* positions;
    const int* group_list;
};

// Per-element offset of a sample row element (layout [c][t][y][x]) from the
// patch origin in a [F][C][H][W] video, built once per workgroup so that an
// element read is base + offset.
__device__ inline void build_input_offsets(const StageDesc& d, int* s_off) {
    for (int e = threadIdx.x; e < d.dim; e += blockDim.x) {
        const int c = e / d.patch_dim;
        const int w = e - c * d.patch_dim;
        const int t = w / d.patch_area;
        const int yx = w - t * d.patch_area;
        const int y = yx / d.patch;
        const int x = yx - y * d.patch;
        s_off[e] = ((t * d.channels + c) * d.height + y) * d.width + x;
    }
}

__device__ inline size_t patch_base(const StageDesc& d, int x, int y, int frame) {
    return (static_cast<size_t>(frame) * d.channels * d.height + y) * d.width + x;
}

__device__ inline float log_via_double(float x) {
    return static_cast<float>(log(static_cast<double>(x)));
}
__device__ inline float log1p_via_double(float x) {
    return static_cast<float>(log1p(static_cast<double>(x)));
}
__device__ inline float exp_via_double(float x) {
    return static_cast<float>(exp(static_cast<double>(x)));
}

__device__ inline int model_rows_of(const StageDesc& d, int similar) {
    return similar < d.model_cap ? similar : d.model_cap;
}

__device__ inline bool group_uses_dual(const StageDesc& d, int rows) {
    return d.model_dim >= 64 && rows < d.model_dim;
}

// ---------------------------------------------------------------------------
// prepare
// ---------------------------------------------------------------------------

// Loads rows [first, first + n) of a group's samples into s_rows
// ([n][dim]) with consecutive threads on consecutive elements.
__device__ inline void load_rows(const StageDesc& d, const float* __restrict__ video,
                                 const size_t* s_base, const int* s_off, int first, int n,
                                 float* s_rows) {
    const int total = n * d.dim;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int r = i / d.dim;
        const int e = i - r * d.dim;
        s_rows[i] = video[s_base[first + r] + s_off[e]];
    }
}

// Flags (equal / flat), model means and the Gram (dual PCA) or covariance
// matrix of one group.  The retained rows stream through LDS in chunks of
// d.row_chunk rows; each element's sums run over the rows in order.
__global__ __launch_bounds__(kModelThreads) void prepare_kernel(
    StageDesc d, const float* __restrict__ noisy, const float* __restrict__ basic,
    GroupRefs refs, int group_count, int problem_stride,
    int* __restrict__ group_flags, int* __restrict__ problem_dims,
    float* __restrict__ means, float* __restrict__ problems) {
    extern __shared__ float s_dyn[];  // [row_chunk][dim] | row0 [dim] | offsets [dim]
    __shared__ size_t s_base[kMaxCap];
    __shared__ int s_flag[4];
    __shared__ double s_sum[kModelThreads / 64][kMaxChannels][2];

    const int g = blockIdx.x;
    if (g >= group_count) {
        return;
    }
    const int a = refs.group_list[g];
    const int similar = refs.counts[a];
    const size_t list = static_cast<size_t>(a) * d.cap;
    const int dim = d.dim;
    const int chunk = d.row_chunk;
    float* s_rows = s_dyn;
    float* s_row0 = s_rows + static_cast<size_t>(chunk) * dim;
    int* s_off = reinterpret_cast<int*>(s_row0 + dim);
    for (int i = threadIdx.x; i < similar; i += blockDim.x) {
        const uint32_t p = refs.positions[list + i];
        s_base[i] = patch_base(d, position_x(p), position_y(p), position_frame(p));
    }
    build_input_offsets(d, s_off);
    if (threadIdx.x == 0) {
        s_flag[0] = 1;
    }
    __syncthreads();

    const bool final_stage = d.stage == kFinal;
    const int rows = model_rows_of(d, similar);
    const bool want_flat = final_stage && d.flat_areas != 0;
    constexpr int kPerThread = kMaxDim / kModelThreads;  // elements per thread

    // Pass over the noisy rows: equality against row 0 (core.cpp:516-531),
    // per-channel double sums for the flat test (core.cpp:533-559) and the
    // noisy mean over the first `rows` rows (compute_mean linalg.cpp:290-311).
    float mean_acc[kPerThread];
    double sums[kMaxChannels][2];
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
        mean_acc[k] = 0.0F;
    }
#pragma unroll
    for (int c = 0; c < kMaxChannels; ++c) {
        sums[c][0] = 0.0;
        sums[c][1] = 0.0;
    }
    bool equal_local = true;
    for (int first = 0; first < similar; first += chunk) {
        const int n = similar - first < chunk ? similar - first : chunk;
        load_rows(d, noisy, s_base, s_off, first, n, s_rows);
        __syncthreads();
        if (first == 0) {
            for (int e = threadIdx.x; e < dim; e += blockDim.x) {
                s_row0[e] = s_rows[e];
            }
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const int e = threadIdx.x + k * kModelThreads;
            if (e < dim) {
                const float v0 = s_row0[e];
                const int c = e / d.patch_dim;
                for (int r = 0; r < n; ++r) {
                    const float v = s_rows[r * dim + e];
                    equal_local = equal_local && (v == v0);
                    if (first + r < rows) {
                        mean_acc[k] = mean_acc[k] + v;
                    }
                    if (want_flat) {
#pragma unroll
                        for (int cc = 0; cc < kMaxChannels; ++cc) {
                            if (cc == c) {
                                sums[cc][0] += static_cast<double>(v);
                                sums[cc][1] += static_cast<double>(v) * static_cast<double>(v);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    if (!equal_local) {
        s_flag[0] = 0;
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (want_flat) {
#pragma unroll
        for (int c = 0; c < kMaxChannels; ++c) {
            const double s0 = wave_sum_double(sums[c][0]);
            const double s1 = wave_sum_double(sums[c][1]);
            if (lane == 0) {
                s_sum[wave][c][0] = s0;
                s_sum[wave][c][1] = s1;
            }
        }
    }
    __syncthreads();
    const bool equal = similar > 1 && s_flag[0] != 0;
    int flags = equal ? kFlagEqual : 0;
    if (!equal && want_flat) {
        const int per_channel = similar * d.patch_dim;
        if (per_channel > 1) {
            float variance = 0.0F;
            const double count = static_cast<double>(per_channel);
            for (int c = 0; c < d.channels; ++c) {
                double total = 0.0;
                double total2 = 0.0;
                for (int w = 0; w < kModelThreads / 64; ++w) {
                    total += s_sum[w][c][0];
                    total2 += s_sum[w][c][1];
                }
                variance = variance +
                           static_cast<float>((total2 - (total * total / count)) / (count - 1.0));
            }
            variance = variance / static_cast<float>(d.channels);
            if (variance < (d.sigma * d.sigma) * d.gamma) {
                flags |= kFlagFlat;
            }
        }
    }

    float* mean_noisy = means + static_cast<size_t>(g) * 2 * dim;
    float* mean_basic = mean_noisy + dim;
    const float inv_rows = 1.0F / static_cast<float>(rows);
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
        const int e = threadIdx.x + k * kModelThreads;
        if (e < dim) {
            mean_acc[k] = mean_acc[k] * inv_rows;
            mean_noisy[e] = mean_acc[k];
        }
    }

    const bool modelled = !equal && d.rank > 0;
    const bool dual = group_uses_dual(d, rows);
    const int n = dual ? rows : d.model_dim;
    if (threadIdx.x == 0) {
        if (rows < (d.model_dim < d.rank ? d.model_dim : d.rank)) {
            flags |= 8;  // rows cannot cover the rank: reported by the host
        }
        group_flags[g] = flags | (modelled ? kFlagModelled : 0);
        problem_dims[g] = n;
    }
    // The basic mean is needed by the filter even when the group is not
    // modelled (flat centre, rank 0); the model rows are the first `rows`
    // basic rows (Final) or noisy rows (Basic), centered.
    if (final_stage) {
        load_rows(d, basic, s_base, s_off, 0, rows, s_rows);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const int e = threadIdx.x + k * kModelThreads;
            if (e < dim) {
                float acc = 0.0F;
                for (int r = 0; r < rows; ++r) {
                    acc = acc + s_rows[r * dim + e];
                }
                acc = acc * inv_rows;
                mean_basic[e] = acc;
                for (int r = 0; r < rows; ++r) {
                    s_rows[r * dim + e] = s_rows[r * dim + e] - acc;
                }
            }
        }
    } else {
        if (!modelled) {
            return;
        }
        if (similar > chunk) {
            load_rows(d, noisy, s_base, s_off, 0, rows, s_rows);
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const int e = threadIdx.x + k * kModelThreads;
            if (e < dim) {
                for (int r = 0; r < rows; ++r) {
                    s_rows[r * dim + e] = s_rows[r * dim + e] - mean_acc[k];
                }
            }
        }
    }
    if (!modelled) {
        return;
    }
    __syncthreads();

    const int md = d.model_dim;
    const float scale = 1.0F / static_cast<float>(rows);
    for (int m = 0; m < d.models; ++m) {
        const int offset = d.coupled ? 0 : m * d.patch_dim;
        const float* model = s_rows + offset;  // row stride dim
        float* out = problems + (static_cast<size_t>(g) * d.models + m) * problem_stride *
                                    problem_stride;
        if (dual) {
            // compute_gram_centered (linalg.cpp:326-354): scale * dot.
            const int team = threadIdx.x / kLanes;
            const int tl = threadIdx.x % kLanes;
            const int pairs = rows * (rows + 1) / 2;
            const int chunks = md / kLanes;
            for (int q = team; q < pairs; q += kTeams) {
                int r = 0;
                while ((r + 1) * (r + 2) / 2 <= q) {
                    ++r;
                }
                const int c = q - r * (r + 1) / 2;
                const float* left = model + r * dim;
                const float* right = model + c * dim;
                float acc = 0.0F;
                for (int k = 0; k < chunks; ++k) {
                    acc = fmaf(left[k * kLanes + tl], right[k * kLanes + tl], acc);
                }
                float total = segment16_sum(acc);
                for (int i = chunks * kLanes; i < md; ++i) {
                    const float product = left[i] * right[i];
                    total = total + product;
                }
                if (tl == 0) {
                    const float value = scale * total;
                    out[r * problem_stride + c] = value;
                    out[c * problem_stride + r] = value;
                }
            }
        } else {
            // compute_covariance_centered (linalg.cpp:377-414): per entry a
            // chain over samples, fused where add_scaled's vector part
            // covers the column.
            const int pairs = md * (md + 1) / 2;
            for (int q = threadIdx.x; q < pairs; q += blockDim.x) {
                int c0 = 0;
                int remaining = q;
                while (remaining >= md - c0) {
                    remaining -= md - c0;
                    ++c0;
                }
                const int c1 = c0 + remaining;
                const int vector_end = ((md - c0) / kLanes) * kLanes;
                const bool fused = (c1 - c0) < vector_end;
                float acc = 0.0F;
                for (int s2 = 0; s2 < rows; ++s2) {
                    const float left = model[s2 * dim + c0];
                    const float right = model[s2 * dim + c1];
                    if (fused) {
                        acc = fmaf(left, right, acc);
                    } else {
                        const float product = left * right;
                        acc = acc + product;
                    }
                }
                const float value = acc * scale;
                out[c0 * problem_stride + c1] = value;
                out[c1 * problem_stride + c0] = value;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// jacobi
// ---------------------------------------------------------------------------

// topk_symmetric_eigen_jacobi (linalg.cpp:447-519): cyclic row-by-row
// rotations above threshold = max(0, 64 eps max(1, max|A|)), at most
// max_sweeps sweeps, "converged" also when the off-diagonal is below
// 8 * threshold; eigenpairs sorted descending (stable) and each vector made
// unit length with its largest-magnitude entry positive
// (canonicalize_vector_row :204-233).  Each lane owns one matrix row.
template <int kLP>  // lanes per problem (power of two >= n)
__global__ __launch_bounds__(64) void jacobi_kernel(
    StageDesc d, int problem_count, int problem_stride,
    const int* __restrict__ group_flags, const int* __restrict__ problem_dims,
    const float* __restrict__ problems, float* __restrict__ eigenvalues,
    float* __restrict__ vectors, int* __restrict__ status) {
    constexpr int kPerWave = 64 / kLP;
    constexpr int kStride = kLP + 1;
    __shared__ float s_m[kPerWave][kLP * kStride];
    __shared__ float s_v[kPerWave][kLP * kStride];
    __shared__ int s_order[kPerWave][kLP];

    const int sub = threadIdx.x / kLP;
    const int r = threadIdx.x % kLP;
    const int problem = blockIdx.x * kPerWave + sub;
    const int group = problem / d.models;
    bool valid = problem < problem_count;
    int n = 0;
    if (valid) {
        valid = (group_flags[group] & kFlagModelled) != 0;
        n = problem_dims[group];
    }
    float* m = s_m[sub];
    float* v = s_v[sub];
    const float* in = problems + static_cast<size_t>(problem) * problem_stride * problem_stride;

    float row_max = 0.0F;
    for (int c = 0; c < kLP; ++c) {
        float value = 0.0F;
        if (valid && r < n && c < n) {
            value = in[r * problem_stride + c];
            row_max = fmaxf(row_max, fabsf(value));
        }
        m[r * kStride + c] = value;
        v[r * kStride + c] = (r == c) ? 1.0F : 0.0F;
    }
    // Segment max over the problem's lanes.
    for (int offset = kLP / 2; offset >= 1; offset >>= 1) {
        row_max = fmaxf(row_max, __shfl_xor(row_max, offset, kLP));
    }
    const float matrix_scale = row_max > 1.0F ? row_max : 1.0F;
    const float threshold = (64.0F * FLT_EPSILON) * matrix_scale;
    __syncthreads();

    const auto offdiag_max = [&]() {
        float value = 0.0F;
        if (r < n) {
            for (int c = r + 1; c < n; ++c) {
                value = fmaxf(value, fabsf(m[r * kStride + c]));
            }
        }
        for (int offset = kLP / 2; offset >= 1; offset >>= 1) {
            value = fmaxf(value, __shfl_xor(value, offset, kLP));
        }
        return value;
    };

    bool done = !valid || n <= 1;
    bool converged = true;
    int sweeps = 0;
    for (int sweep = 0; sweep < d.max_sweeps; ++sweep) {
        if (!done) {
            sweeps = sweep + 1;
        }
        const float off = offdiag_max();
        if (!done && off <= threshold) {
            done = true;
        }
        if (__all(done)) {
            break;
        }
        for (int p = 0; p + 1 < kLP; ++p) {
            for (int q = p + 1; q < kLP; ++q) {
                bool rotate = !done && q < n;
                float apq = 0.0F;
                if (rotate) {
                    apq = m[p * kStride + q];
                    rotate = fabsf(apq) > threshold;
                }
                float c = 1.0F;
                float s = 0.0F;
                float new_pp = 0.0F;
                float new_qq = 0.0F;
                float new_rp = 0.0F;
                float new_rq = 0.0F;
                float new_vp = 0.0F;
                float new_vq = 0.0F;
                if (rotate) {
                    // jacobi_rotate (linalg.cpp:153-202).
                    const float app = m[p * kStride + p];
                    const float aqq = m[q * kStride + q];
                    const float tau = (aqq - app) / (2.0F * apq);
                    const float tau_abs = fabsf(tau);
                    const float t_abs = 1.0F / (tau_abs + sqrtf(1.0F + (tau * tau)));
                    const float t = tau < 0.0F ? -t_abs : t_abs;
                    c = 1.0F / sqrtf(1.0F + (t * t));
                    s = t * c;
                    const float cc = c * c;
                    const float ss = s * s;
                    const float cs = c * s;
                    const float two_cs = 2.0F * cs;
                    new_pp = ((cc * app) - (two_cs * apq)) + (ss * aqq);
                    new_qq = ((ss * app) + (two_cs * apq)) + (cc * aqq);
                    if (r < n && r != p && r != q) {
                        const float arp = m[r * kStride + p];
                        const float arq = m[r * kStride + q];
                        new_rp = (c * arp) - (s * arq);
                        new_rq = (s * arp) + (c * arq);
                    }
                    if (r < n) {
                        const float vrp = v[r * kStride + p];
                        const float vrq = v[r * kStride + q];
                        new_vp = (c * vrp) - (s * vrq);
                        new_vq = (s * vrp) + (c * vrq);
                    }
                }
                __syncthreads();
                if (rotate) {
                    if (r < n && r != p && r != q) {
                        m[r * kStride + p] = new_rp;
                        m[p * kStride + r] = new_rp;
                        m[r * kStride + q] = new_rq;
                        m[q * kStride + r] = new_rq;
                    }
                    if (r == 0) {
                        m[p * kStride + p] = new_pp;
                        m[q * kStride + q] = new_qq;
                        m[p * kStride + q] = 0.0F;
                        m[q * kStride + p] = 0.0F;
                    }
                    if (r < n) {
                        v[r * kStride + p] = new_vp;
                        v[r * kStride + q] = new_vq;
                    }
                }
                __syncthreads();
            }
        }
    }
    const float final_off = offdiag_max();
    if (!done) {
        sweeps = d.max_sweeps;
        converged = final_off <= 8.0F * threshold;
    }

    // Stable descending order of the diagonal.
    if (valid && r < n) {
        const float dr = m[r * kStride + r];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const float dj = m[j * kStride + j];
            rank += (dj > dr || (j < r && !(dr > dj))) ? 1 : 0;
        }
        s_order[sub][rank] = r;
    }
    __syncthreads();
    if (valid && r < d.rank && r < n) {
        const int col = s_order[sub][r];
        float* vec_out = vectors + (static_cast<size_t>(problem) * d.rank + r) * problem_stride;
        float norm_squared = 0.0F;
        float max_abs = 0.0F;
        int pivot = 0;
        for (int i = 0; i < n; ++i) {
            const float value = v[i * kStride + col];
            const float square = value * value;
            norm_squared = norm_squared + square;
            const float value_abs = fabsf(value);
            if (value_abs > max_abs) {
                max_abs = value_abs;
                pivot = i;
            }
        }
        float scale = 1.0F;
        const bool rescale = norm_squared > FLT_MIN;
        if (rescale) {
            scale = 1.0F / sqrtf(norm_squared);
            if (v[pivot * kStride + col] < 0.0F) {
                scale = -scale;
            }
        }
        for (int i = 0; i < n; ++i) {
            const float value = v[i * kStride + col];
            vec_out[i] = rescale ? value * scale : value;
        }
        eigenvalues[static_cast<size_t>(problem) * d.rank + r] = m[col * kStride + col];
    }
    if (valid && r == 0) {
        status[problem] = sweeps | (converged ? 0 : (1 << 16));
    }
}

// ---------------------------------------------------------------------------
// filter + scatter
// ---------------------------------------------------------------------------

struct Accumulators {
    double* data;  // [frames][channels + 1][height][width]
};

__device__ inline void scatter_value(const StageDesc& d, Accumulators acc, size_t acc_base,
                                     int pf, int anchor_frame, int acc_off, int meta, float win,
                                     float value, float w) {
    const int t = meta & 0xFF;
    const int c = (meta >> 8) & 0xFF;
    const int slot = pf + t - anchor_frame + d.search_bwd;
    if (slot < 0 || slot >= d.slots) {
        return;
    }
    const float cw = w * win;
    double addend;
    if ((meta >> 16) != 0) {
        addend = static_cast<double>(w) * static_cast<double>(value);
    } else {
        const float product = cw * value;
        addend = static_cast<double>(product);
    }
    double* cell = acc.data + acc_base + acc_off;
    unsafeAtomicAdd(cell, addend);
    if (c == 0) {
        const size_t plane = static_cast<size_t>(d.width) * d.height;
        unsafeAtomicAdd(cell + static_cast<size_t>(d.channels) * plane, static_cast<double>(cw));
    }
}

// Per-lane register slice of a row: element k * 16 + lane for k < kR.
template <int kR>
struct RowSlice {
    float v[kR];
};

// Dot product of a centered LDS row slice with an LDS vector in the 16-lane
// order (dot_contiguous_highway, kernels_highway.hpp:12-44): lane-wise fma
// chains, halving tree, then the sequential tail.
template <int kR>
__device__ inline float team_dot(const RowSlice<kR>& row, const float* vec, int chunks, int lane,
                                 const float* row_tail, const float* tail_mean, int md) {
    float part = 0.0F;
#pragma unroll
    for (int k = 0; k < kR; ++k) {
        if (k < chunks) {
            part = fmaf(row.v[k], vec[k * kLanes + lane], part);
        }
    }
    float total = segment16_sum(part);
    for (int i = chunks * kLanes; i < md; ++i) {
        const float x = row_tail[i] - tail_mean[i];
        const float product = x * vec[i];
        total = total + product;
    }
    return total;
}

template <int kR>
__device__ inline void load_slice(RowSlice<kR>& row, const float* src, const float* mean,
                                  int chunks, int lane) {
#pragma unroll
    for (int k = 0; k < kR; ++k) {
        row.v[k] = 0.0F;
        if (k < chunks) {
            const int e = k * kLanes + lane;
            row.v[k] = src[e] - mean[e];
        }
    }
}

// Dual-basis map, Wiener coefficients, aggregation weights, reconstruction
// and the weighted scatter of one group.  Rows stream through LDS in chunks
// of d.row_chunk.
template <int kR>
__global__ __launch_bounds__(kModelThreads) void filter_kernel(
    StageDesc d, const float* __restrict__ noisy, const float* __restrict__ basic,
    GroupRefs refs, int group_count, int problem_stride,
    const int* __restrict__ group_flags, const int* __restrict__ problem_dims,
    const float* __restrict__ means, const float* __restrict__ eigenvalues,
    const float* __restrict__ vectors, Accumulators acc, float* __restrict__ weights_out,
    unsigned group_stride) {
    extern __shared__ float s_dyn[];
    __shared__ size_t s_base[kMaxCap];
    __shared__ size_t s_acc_base[kMaxCap];
    __shared__ int s_f[kMaxCap];
    __shared__ float s_logw[kMaxCap];
    __shared__ float s_w[kMaxCap];
    __shared__ float s_score[kMaxCap];
    __shared__ float s_coef[kMaxChannels * kMaxRank];
    __shared__ float s_lhat[kMaxChannels * kMaxRank];
    __shared__ float s_sp[kTeams][kMaxRank];
    __shared__ float s_red[8];
    __shared__ float s_scalar[4 * kMaxChannels];

    // Workgroups visit the groups with a stride coprime with their count so
    // that concurrently running groups scatter into distant parts of the
    // clip instead of one band of rows.
    if (static_cast<int>(blockIdx.x) >= group_count) {
        return;
    }
    const int g = static_cast<int>((static_cast<unsigned long long>(blockIdx.x) * group_stride) %
                                   static_cast<unsigned>(group_count));
    const int a = refs.group_list[g];
    const int af = position_frame(refs.anchors[a]);
    const int similar = refs.counts[a];
    const size_t list = static_cast<size_t>(a) * d.cap;
    const int dim = d.dim;
    const int md = d.model_dim;
    const int R = d.rank;
    const int chunk = d.row_chunk;
    float* s_rows = s_dyn;                                      // [chunk][dim]
    float* s_basis = s_rows + static_cast<size_t>(chunk) * dim;  // [models][R][md]
    float* s_mean_n = s_basis + static_cast<size_t>(R) * dim;
    float* s_mean_b = s_mean_n + dim;
    int* s_off = reinterpret_cast<int*>(s_mean_b + dim);
    int* s_acc_off = s_off + dim;
    int* s_meta = s_acc_off + dim;
    float* s_win = reinterpret_cast<float*>(s_meta + dim);

    for (int i = threadIdx.x; i < similar; i += blockDim.x) {
        const uint32_t p = refs.positions[list + i];
        const int px = position_x(p);
        const int py = position_y(p);
        const int pf = position_frame(p);
        s_base[i] = patch_base(d, px, py, pf);
        s_acc_base[i] =
            (static_cast<size_t>(pf) * (d.channels + 1) * d.height + py) * d.width + px;
        s_f[i] = pf;
        s_logw[i] = 0.0F;
        s_w[i] = 1.0F;
    }
    build_input_offsets(d, s_off);
    // Scatter tables: accumulator offset from the patch's accumulator base,
    // frame offset t, channel, fused flag and window weight per element.
    const bool final_stage = d.stage == kFinal;
    const float* g_mean = means + static_cast<size_t>(g) * 2 * dim;
    for (int e = threadIdx.x; e < dim; e += blockDim.x) {
        const int c = e / d.patch_dim;
        const int w = e - c * d.patch_dim;
        const int t = w / d.patch_area;
        const int yx = w - t * d.patch_area;
        const int y = yx / d.patch;
        const int x = yx - y * d.patch;
        s_acc_off[e] = ((t * (d.channels + 1) + c) * d.height + y) * d.width + x;
        s_meta[e] = t | (c << 8) | ((x < d.fused_width ? 1 : 0) << 16);
        s_win[e] = d.window[yx];
        s_mean_n[e] = g_mean[e];
        s_mean_b[e] = final_stage ? g_mean[dim + e] : 0.0F;
    }
    const int flags = group_flags[g];
    const bool equal = (flags & kFlagEqual) != 0;
    const bool flat = (flags & kFlagFlat) != 0;
    const bool modelled = (flags & kFlagModelled) != 0;
    const int rows = model_rows_of(d, similar);
    const bool dual = group_uses_dual(d, rows);
    __syncthreads();

    const int team = threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    const int chunks = md / kLanes;
    const float* center = (final_stage && flat) ? s_mean_b : s_mean_n;
    const float* member_video = final_stage ? basic : noisy;
    const float* member_mean = final_stage ? s_mean_b : s_mean_n;

    const float sigma = d.sigma;
    const float beta = d.beta;
    const float model_noise2 =
        final_stage ? (beta * d.sigma_basic) * d.sigma_basic : (beta * sigma) * sigma;
    const float estimate_noise2 = (beta * sigma) * sigma;
    const bool weights_need_scores = !equal && d.weight_beta != 0.0F;
    int model_count = 0;
    // Which source rows s_rows holds: 0 none, 1 member rows [0, resident),
    // 2 noisy rows [0, resident).
    int resident_kind = 0;
    int resident = 0;

    if (!equal) {
        if (modelled && dual) {
            load_rows(d, member_video, s_base, s_off, 0, rows, s_rows);
            resident_kind = 1;
            resident = rows;
            __syncthreads();
        }
        for (int m = 0; m < d.models; ++m) {
            const int offset = d.coupled ? 0 : m * d.patch_dim;
            float* basis = s_basis + static_cast<size_t>(m) * R * md;
            const size_t problem = static_cast<size_t>(g) * d.models + m;
            const float* lambdas = eigenvalues + problem * R;
            const float* vecs = vectors + problem * R * problem_stride;
            if (modelled) {
                if (dual) {
                    // map_dual_eigenvectors_to_basis (linalg.cpp:552-595).
                    const float floor = 128.0F * FLT_EPSILON;
                    const int fused_end = chunks * kLanes;
                    for (int e = threadIdx.x; e < md; e += blockDim.x) {
                        const bool fused = e < fused_end;
                        const float mean_e = member_mean[offset + e];
                        for (int j0 = 0; j0 < R; j0 += 16) {
                            float accs[16];
#pragma unroll
                            for (int jj = 0; jj < 16; ++jj) {
                                accs[jj] = 0.0F;
                            }
                            for (int s = 0; s < rows; ++s) {
                                const float x = s_rows[s * dim + offset + e] - mean_e;
#pragma unroll
                                for (int jj = 0; jj < 16; ++jj) {
                                    const int j = j0 + jj;
                                    if (j < R) {
                                        const float u = vecs[static_cast<size_t>(j) * problem_stride + s];
                                        if (fused) {
                                            accs[jj] = fmaf(u, x, accs[jj]);
                                        } else {
                                            const float product = u * x;
                                            accs[jj] = accs[jj] + product;
                                        }
                                    }
                                }
                            }
#pragma unroll
                            for (int jj = 0; jj < 16; ++jj) {
                                const int j = j0 + jj;
                                if (j < R) {
                                    const float lambda = lambdas[j];
                                    float value = 0.0F;
                                    if (!(lambda <= floor)) {
                                        const float inv_sqrt = 1.0F / sqrtf(lambda);
                                        value = accs[jj] * inv_sqrt;
                                    }
                                    basis[static_cast<size_t>(j) * md + e] = value;
                                }
                            }
                        }
                    }
                    __syncthreads();
                    // normalize_or_zero (linalg.cpp:235-246).
                    for (int j = team; j < R; j += kTeams) {
                        float* row = basis + static_cast<size_t>(j) * md;
                        const float lambda = lambdas[j];
                        if (lambda <= floor) {
                            continue;
                        }
                        float part = 0.0F;
                        for (int k = 0; k < chunks; ++k) {
                            const float value = row[k * kLanes + lane];
                            part = fmaf(value, value, part);
                        }
                        float norm2 = segment16_sum(part);
                        for (int i = chunks * kLanes; i < md; ++i) {
                            const float product = row[i] * row[i];
                            norm2 = norm2 + product;
                        }
                        const bool zero = norm2 <= floor;
                        const float scale = zero ? 0.0F : 1.0F / sqrtf(norm2);
                        for (int i = lane; i < md; i += kLanes) {
                            row[i] = zero ? 0.0F : row[i] * scale;
                        }
                    }
                } else {
                    for (int i = threadIdx.x; i < R * md; i += blockDim.x) {
                        const int j = i / md;
                        const int e = i - j * md;
                        basis[i] = vecs[static_cast<size_t>(j) * problem_stride + e];
                    }
                }
                // compute_filter_coefficients (core.cpp:616-636).
                if (threadIdx.x < R) {
                    const int j = threadIdx.x;
                    const float sigma2 = (beta * sigma) * sigma;
                    const float sigma_basic2 =
                        final_stage ? (beta * d.sigma_basic) * d.sigma_basic : sigma2;
                    const float threshold = d.variance_threshold * sigma2;
                    float variance = lambdas[j];
                    if (sigma_basic2 > 0.0F) {
                        variance = variance - fminf(variance, sigma_basic2);
                    }
                    s_coef[m * kMaxRank + j] =
                        variance > threshold ? 1.0F / (1.0F + (sigma2 / variance)) : 0.0F;
                    const float lhat = lambdas[j] - model_noise2;
                    s_lhat[m * kMaxRank + j] = lhat > 0.0F ? lhat : 0.0F;
                }
            }
            // accumulate_aggregation_model (core.cpp:326-412): group terms.
            if (threadIdx.x == 0) {
                const int rank_used = modelled ? R : 0;
                float trace = 0.0F;
                float volume = 0.0F;
                const float match_sigma = final_stage ? d.sigma_basic : sigma;
                const float match_noise2 = match_sigma * match_sigma;
                const float floor_noise2 = (d.membership_noise_floor * sigma) * sigma;
                const float membership_noise2 =
                    match_noise2 > floor_noise2 ? match_noise2 : floor_noise2;
                for (int j = 0; j < rank_used; ++j) {
                    const float lambda = s_lhat[m * kMaxRank + j];
                    if (lambda > 0.0F) {
                        trace = trace + lambda / (lambda + estimate_noise2);
                        volume = volume + log1p_via_double(lambda / membership_noise2);
                    }
                }
                const float mean_term = static_cast<float>(md) / static_cast<float>(rows);
                const float tau = (d.weight_epsilon + mean_term) + trace;
                s_scalar[4 * m + 0] =
                    d.weight_alpha == 0.0F ? 0.0F : -d.weight_alpha * log_via_double(tau);
                s_scalar[4 * m + 1] = volume / static_cast<float>(md);
                s_scalar[4 * m + 2] = membership_noise2;
            }
        }
        __syncthreads();

        // Membership scores over all rows, all models (the min over rows is
        // needed before any weight is final).
        if (d.weight_alpha == 0.0F && d.weight_beta == 0.0F) {
            model_count = d.models;
        } else if (!weights_need_scores) {
            for (int r = threadIdx.x; r < similar; r += blockDim.x) {
                float lw = s_logw[r];
                for (int m = 0; m < d.models; ++m) {
                    lw = lw + s_scalar[4 * m + 0];
                }
                s_logw[r] = lw;
            }
            model_count = d.models;
        } else {
            for (int m = 0; m < d.models; ++m) {
                const int offset = d.coupled ? 0 : m * d.patch_dim;
                const float* basis = s_basis + static_cast<size_t>(m) * R * md;
                const int rank_used = modelled ? R : 0;
                const float volume = s_scalar[4 * m + 1];
                const float membership_noise2 = s_scalar[4 * m + 2];
                const float score_scale = sqrtf(2.0F * static_cast<float>(md));
                float local_min = INFINITY;
                for (int first = 0; first < similar; first += chunk) {
                    const int n = similar - first < chunk ? similar - first : chunk;
                    if (!(resident_kind == 1 && first == 0 && n <= resident)) {
                        __syncthreads();
                        load_rows(d, member_video, s_base, s_off, first, n, s_rows);
                        resident_kind = first == 0 ? 1 : 0;
                        resident = first == 0 ? n : 0;
                        __syncthreads();
                    }
                    for (int rr = team; rr < n; rr += kTeams) {
                        const int r = first + rr;
                        const float* src = s_rows + rr * dim + offset;
                        const float* mean = member_mean + offset;
                        RowSlice<kR> row;
                        load_slice(row, src, mean, chunks, lane);
                        float part = 0.0F;
#pragma unroll
                        for (int k = 0; k < kR; ++k) {
                            if (k < chunks) {
                                part = fmaf(row.v[k], row.v[k], part);
                            }
                        }
                        float norm2 = segment16_sum(part);
                        for (int i = chunks * kLanes; i < md; ++i) {
                            const float x = src[i] - mean[i];
                            const float product = x * x;
                            norm2 = norm2 + product;
                        }
                        float distance = 0.0F;
                        float projected = 0.0F;
                        for (int j = 0; j < rank_used; ++j) {
                            const float projection = team_dot(row, basis + static_cast<size_t>(j) * md,
                                                              chunks, lane, src, mean, md);
                            const float projection2 = projection * projection;
                            distance = distance +
                                       projection2 / (s_lhat[m * kMaxRank + j] + membership_noise2);
                            projected = projected + projection2;
                        }
                        const float residual = norm2 - projected > 0.0F ? norm2 - projected : 0.0F;
                        distance = distance + residual / membership_noise2;
                        const float score = (distance - static_cast<float>(md)) / score_scale;
                        if (lane == 0) {
                            s_score[r] = score;
                        }
                        local_min = fminf(local_min, score);
                    }
                }
                local_min = wave_min(local_min);
                if ((threadIdx.x & 63) == 0) {
                    s_red[threadIdx.x >> 6] = local_min;
                }
                __syncthreads();
                float min_score = s_red[0];
                for (int w = 1; w < (kModelThreads >> 6); ++w) {
                    min_score = fminf(min_score, s_red[w]);
                }
                const float log_group = s_scalar[4 * m + 0];
                for (int r = threadIdx.x; r < similar; r += blockDim.x) {
                    const float excess = s_score[r] - min_score;
                    const float membership = (excess > 0.0F ? excess : 0.0F) + volume;
                    const float delta = log_group - (d.weight_beta * membership);
                    s_logw[r] = s_logw[r] + delta;
                }
                __syncthreads();
            }
            model_count = d.models;
        }
    }

    // finalize_aggregation_weights (core.cpp:276-294).
    if (d.weights_enabled) {
        if (model_count == 0) {
            const float tau = d.weight_epsilon +
                              (static_cast<float>(md) / static_cast<float>(rows));
            const float log_group =
                d.weight_alpha == 0.0F ? 0.0F : -d.weight_alpha * log_via_double(tau);
            for (int r = threadIdx.x; r < similar; r += blockDim.x) {
                s_logw[r] = s_logw[r] + log_group;
            }
            model_count = 1;
            __syncthreads();
        }
        const float inv_models = 1.0F / static_cast<float>(model_count);
        for (int r = threadIdx.x; r < similar; r += blockDim.x) {
            float value = s_logw[r] * inv_models;
            value = value < -80.0F ? -80.0F : (value > 80.0F ? 80.0F : value);
            s_w[r] = exp_via_double(value);
        }
    }
    __syncthreads();
    if (weights_out != nullptr) {
        for (int r = threadIdx.x; r < similar; r += blockDim.x) {
            weights_out[list + r] = s_w[r];
        }
    }

    // Reconstruction (project_bayes_estimate core.cpp:578-614) and the
    // weighted scatter of every sample (write_*_contributions).
    const bool noisy_resident_ok = !final_stage;  // Basic's member rows are the noisy rows
    for (int first = 0; first < similar; first += chunk) {
        const int n = similar - first < chunk ? similar - first : chunk;
        const bool reuse = modelled && !equal && noisy_resident_ok && resident_kind == 1 &&
                           first == 0 && n <= resident;
        if (!reuse && (equal || modelled)) {
            __syncthreads();
            load_rows(d, noisy, s_base, s_off, first, n, s_rows);
            __syncthreads();
        }
        for (int rr = team; rr < n; rr += kTeams) {
            const int r = first + rr;
            const float w = s_w[r];
            const float* row_n = s_rows + rr * dim;
            for (int m = 0; m < d.models; ++m) {
                const int offset = d.coupled ? 0 : m * d.patch_dim;
                if (equal) {
                    for (int e = lane; e < md; e += kLanes) {
                        scatter_value(d, acc, s_acc_base[r], s_f[r], af, s_acc_off[offset + e],
                                      s_meta[offset + e], s_win[offset + e], row_n[offset + e], w);
                    }
                    continue;
                }
                if (!modelled) {
                    for (int e = lane; e < md; e += kLanes) {
                        scatter_value(d, acc, s_acc_base[r], s_f[r], af, s_acc_off[offset + e],
                                      s_meta[offset + e], s_win[offset + e], center[offset + e], w);
                    }
                    continue;
                }
                const float* basis = s_basis + static_cast<size_t>(m) * R * md;
                const float* src = row_n + offset;
                const float* cmean = center + offset;
                RowSlice<kR> row;
                load_slice(row, src, cmean, chunks, lane);
                for (int j = 0; j < R; ++j) {
                    const float coefficient = s_coef[m * kMaxRank + j];
                    float scaled = 0.0F;
                    if (coefficient != 0.0F) {
                        const float projection = team_dot(row, basis + static_cast<size_t>(j) * md,
                                                          chunks, lane, src, cmean, md);
                        scaled = coefficient * projection;
                    }
                    if (lane == 0) {
                        s_sp[team][j] = scaled;
                    }
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                for (int e = lane; e < md; e += kLanes) {
                    const bool fused = e < chunks * kLanes;
                    float out = 0.0F;
                    for (int j = 0; j < R; ++j) {
                        if (s_coef[m * kMaxRank + j] == 0.0F) {
                            continue;
                        }
                        const float sp = s_sp[team][j];
                        const float b = basis[static_cast<size_t>(j) * md + e];
                        if (fused) {
                            out = fmaf(sp, b, out);
                        } else {
                            const float product = sp * b;
                            out = out + product;
                        }
                    }
                    out = out + cmean[e];
                    scatter_value(d, acc, s_acc_base[r], s_f[r], af, s_acc_off[offset + e],
                                  s_meta[offset + e], s_win[offset + e], out, w);
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
    }
}

}  // namespace

size_t prepare_shared_bytes(const StageDesc& d) {
    return (static_cast<size_t>(d.row_chunk) * d.dim + 2 * static_cast<size_t>(d.dim)) *
           sizeof(float);
}

size_t filter_shared_bytes(const StageDesc& d) {
    return (static_cast<size_t>(d.row_chunk) * d.dim + static_cast<size_t>(d.rank) * d.dim +
            6 * static_cast<size_t>(d.dim)) *
           sizeof(float);
}

hipError_t launch_prepare(const StageDesc& d, const float* noisy, const float* basic,
                          const uint32_t* anchors, const int* counts,
                          const uint32_t* positions, const int* group_list,
                          int group_count, int problem_stride, int* group_flags,
                          int* problem_dims, float* means, float* problems,
                          hipStream_t stream) {
    if (group_count <= 0) {
        return hipSuccess;
    }
    const GroupRefs refs{anchors, counts, positions, group_list};
    hipLaunchKernelGGL(prepare_kernel, dim3(static_cast<unsigned>(group_count)),
                       dim3(kModelThreads), prepare_shared_bytes(d), stream, d, noisy, basic,
                       refs, group_count, problem_stride, group_flags, problem_dims, means,
                       problems);
    return hipGetLastError();
}

hipError_t launch_jacobi(const StageDesc& d, int problem_count, int problem_stride,
                         const int* group_flags, const int* problem_dims,
                         const float* problems, float* eigenvalues, float* vectors,
                         int* status, hipStream_t stream) {
    if (problem_count <= 0) {
        return hipSuccess;
    }
    const int n = problem_stride;
#define VNLBHIP_JACOBI(LP)                                                                      \
    hipLaunchKernelGGL(jacobi_kernel<LP>,                                                       \
                       dim3(static_cast<unsigned>((problem_count + (64 / LP) - 1) / (64 / LP))), \
                       dim3(64), 0, stream, d, problem_count, problem_stride, group_flags,       \
                       problem_dims, problems, eigenvalues, vectors, status)
    if (n <= 4) {
        VNLBHIP_JACOBI(4);
    } else if (n <= 8) {
        VNLBHIP_JACOBI(8);
    } else if (n <= 16) {
        VNLBHIP_JACOBI(16);
    } else if (n <= 32) {
        VNLBHIP_JACOBI(32);
    } else {
        VNLBHIP_JACOBI(64);
    }
#undef VNLBHIP_JACOBI
    return hipGetLastError();
}

hipError_t launch_filter(const StageDesc& d, const float* noisy, const float* basic,
                         const uint32_t* anchors, const int* counts,
                         const uint32_t* positions, const int* group_list, int group_count,
                         int problem_stride, const int* group_flags, const int* problem_dims,
                         const float* means, const float* eigenvalues, const float* vectors,
                         double* accumulators, float* weights_out, hipStream_t stream) {
    if (group_count <= 0) {
        return hipSuccess;
    }
    const GroupRefs refs{anchors, counts, positions, group_list};
    const Accumulators acc{accumulators};
    const int chunks = (d.model_dim + kLanes - 1) / kLanes;
    const size_t shared = filter_shared_bytes(d);
    const dim3 grid(static_cast<unsigned>(group_count));
    const dim3 block(kModelThreads);
    // Largest stride near 0.618 * count that is coprime with the count.
    unsigned stride = static_cast<unsigned>(static_cast<double>(group_count) * 0.6180339887) | 1u;
    const auto gcd = [](unsigned x, unsigned y) {
        while (y != 0) {
            const unsigned r = x % y;
            x = y;
            y = r;
        }
        return x;
    };
    while (stride > 1 && gcd(stride, static_cast<unsigned>(group_count)) != 1) {
        stride -= 2;
    }
    if (stride == 0) {
        stride = 1;
    }
#define VNLBHIP_FILTER(KR)                                                                  \
    hipLaunchKernelGGL(filter_kernel<KR>, grid, block, shared, stream, d, noisy, basic,     \
                       refs, group_count, problem_stride, group_flags, problem_dims, means, \
                       eigenvalues, vectors, acc, weights_out, stride)
    if (chunks <= 4) {
        VNLBHIP_FILTER(4);
    } else if (chunks <= 8) {
        VNLBHIP_FILTER(8);
    } else if (chunks <= 16) {
        VNLBHIP_FILTER(16);
    } else if (chunks <= 24) {
        VNLBHIP_FILTER(24);
    } else if (chunks <= 32) {
        VNLBHIP_FILTER(32);
    } else {
        VNLBHIP_FILTER(64);
    }
#undef VNLBHIP_FILTER
    return hipGetLastError();
}

}  // namespace vnlbhip
