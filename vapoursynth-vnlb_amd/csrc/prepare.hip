// Per-group model input: group flags (equal / flat), the noisy and basic
// means, the centered model rows and their Gram (dual PCA) or covariance
// matrix, for the eigen solver.
//
// Semantics follow the reference (equality bypass core.cpp:516-531, flat
// test :533-559, compute_mean linalg.cpp:290-311, center :313-324,
// compute_gram_centered :326-354, compute_covariance_centered :377-414) in
// the 16-lane order of the CPU restatement.
//
// Layout as in the filter kernel: model rows live in LDS in slot layout
// (lane l of a 16-lane team owns elements 16 k + l), one team loads one row
// with consecutive lanes on consecutive elements, and a Gram entry is a
// team dot product read with 128-bit LDS loads.  Workgroups are persistent
// and build the element tables once.
#include "common.h"
#include "device_util.h"
#include "group.h"
#include "kernels.h"

#include <hip/hip_bf16.h>

namespace vnlbhip {
namespace {

struct PrepareLayout {
    int pitch;     // words per lane slot row
    int row;       // 16 * pitch
    int width;     // models * row: words per sample row
    int chunks;
    int tail;
    int tab_in;    // [models][row] input offsets
    int tab_c;     // [models][row] channel of the element (-1 outside the row)
    int rows;      // [max_rows][width] model rows
    int pos;       // [cap] packed positions
    int gram;      // [waves][4][64] partial Gram tiles of the matrix-core form
    int words;
};

PrepareLayout prepare_layout(const StageDesc& d) {
    PrepareLayout L{};
    L.pitch = slot_pitch(d.model_dim);
    L.row = kLanes * L.pitch;
    L.width = d.models * L.row;
    L.chunks = d.model_dim / kLanes;
    L.tail = d.model_dim % kLanes;
    int at = 0;
    const auto take = [&at](int words) {
        const int start = at;
        at += (words + 3) / 4 * 4;
        return start;
    };
    // The channel table is read only by the flat-area test (Final) and by
    // rows that do not fill whole 16-element slots; without either it is
    // left out (-1), which lets one more workgroup fit on a compute unit.
    const bool exact = d.model_dim == kLanes * template_slots(d.model_dim);
    const bool flat = d.stage == kFinal && d.flat_areas != 0;
    L.tab_in = take(L.width);
    L.tab_c = exact && !flat ? -1 : take(L.width);
    L.rows = take(d.max_rows * L.width);
    L.pos = take(d.cap);
    // Per-wave partial Grams of the matrix-core form: over the model rows
    // (after a barrier: the products have read them) when there is one model.
    constexpr int kGramWords = kModelThreads / 64 * 4 * 64;
    L.gram = d.models == 1 && d.max_rows * L.width >= kGramWords ? L.rows : take(kGramWords);
    L.words = at;
    return L;
}

// Loads sample row r (all models) into LDS slot layout: team-wide, lane l
// reads the elements 16 k + l of each model.  Every load of a lane is issued
// before the first LDS store (kExact: kS slots, all full), so a row costs
// one memory round trip.
template <int kS, bool kExact>
__device__ inline void load_row(const StageDesc& d, const PrepareLayout& L,
                                const float* __restrict__ video, size_t base, const int* tab_in,
                                float* dst, int lane) {
    for (int m = 0; m < d.models; ++m) {
        const int at = m * L.row + lane * L.pitch;
        if (kExact) {
            const int4* t4 = reinterpret_cast<const int4*>(tab_in + at);
            float v[kS];
#pragma unroll
            for (int h = 0; h < kS / 4; ++h) {
                const int4 o = t4[h];
                v[4 * h + 0] = video[base + o.x];
                v[4 * h + 1] = video[base + o.y];
                v[4 * h + 2] = video[base + o.z];
                v[4 * h + 3] = video[base + o.w];
            }
            float4* d4 = reinterpret_cast<float4*>(dst + at);
#pragma unroll
            for (int h = 0; h < kS / 4; ++h) {
                d4[h] = float4{v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]};
            }
        } else {
            const int n = L.chunks + (lane < L.tail ? 1 : 0);
            for (int k0 = 0; k0 < n; k0 += 8) {
                float v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    v[k] = k0 + k < n ? video[base + tab_in[at + k0 + k]] : 0.0F;
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (k0 + k < n) {
                        dst[at + k0 + k] = v[k];
                    }
                }
            }
        }
    }
}

// Means over the model rows of four consecutive slots per thread (kExact:
// every slot k < kS holds an element), in row order as the scalar loop does,
// written to mean_out; optionally the rows are centered and the equality of
// every row with row 0 is tracked.
template <int kS>
__device__ inline void slot_means4(const StageDesc& d, const PrepareLayout& L, float* s_rows, int rows,
                                   float inv_rows, float* mean_out, bool center, bool& equal_local,
                                   int tid) {
    constexpr int kGroups = kS / 4;  // float4 groups per lane slot row
    const int width = L.width;
    const int total = d.models * kLanes * kGroups;
    for (int t = tid; t < total; t += kModelThreads) {
        const int m = t / (kLanes * kGroups);
        const int rem = t - m * kLanes * kGroups;
        const int l = rem / kGroups;
        const int k = 4 * (rem - l * kGroups);
        const int i = m * L.row + l * L.pitch + k;
        const float4 v0 = *reinterpret_cast<const float4*>(s_rows + i);
        float4 acc = {0.0F, 0.0F, 0.0F, 0.0F};
        for (int r = 0; r < rows; ++r) {
            const float4 v = *reinterpret_cast<const float4*>(s_rows + r * width + i);
            equal_local = equal_local && v.x == v0.x && v.y == v0.y && v.z == v0.z && v.w == v0.w;
            acc.x = acc.x + v.x;
            acc.y = acc.y + v.y;
            acc.z = acc.z + v.z;
            acc.w = acc.w + v.w;
        }
        const float4 mean = {acc.x * inv_rows, acc.y * inv_rows, acc.z * inv_rows, acc.w * inv_rows};
        float* out = mean_out + (d.coupled ? 0 : m * d.patch_dim) + kLanes * k + l;
        out[0] = mean.x;
        out[kLanes] = mean.y;
        out[2 * kLanes] = mean.z;
        out[3 * kLanes] = mean.w;
        if (center) {
            for (int r = 0; r < rows; ++r) {
                float4* v = reinterpret_cast<float4*>(s_rows + r * width + i);
                float4 x = *v;
                x.x = x.x - mean.x;
                x.y = x.y - mean.y;
                x.z = x.z - mean.z;
                x.w = x.w - mean.w;
                *v = x;
            }
        }
    }
}

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

// kMfma (FAST form): the dual Gram on the matrix cores in bf16x3 instead of
// the 16-lane fp32 team dots.  At least four workgroups per compute unit
// (at most 128 VGPRs): a group is a chain of dependent memory round trips,
// so groups in flight set the rate.
template <int kS, bool kExact, bool kMfma, bool kFlat>
__global__ __launch_bounds__(kModelThreads, kFlat ? 4 : 5) void prepare_kernel(
    StageDesc d, PrepareLayout L, const float* __restrict__ noisy,
    const float* __restrict__ basic, GroupRefs refs, const int* __restrict__ group_count_dev,
    int problem_stride,
    int* __restrict__ group_flags, int* __restrict__ problem_dims, float* __restrict__ means,
    float* __restrict__ problems, int* __restrict__ work_counter) {
    extern __shared__ float4 s_dyn4[];
    float* s_dyn = reinterpret_cast<float*>(s_dyn4);
    __shared__ int s_equal;
    __shared__ int s_next;
    __shared__ double s_sum[kModelThreads / 64][kMaxChannels][2];

    int* tab_in = reinterpret_cast<int*>(s_dyn + L.tab_in);
    int* tab_c = L.tab_c >= 0 ? reinterpret_cast<int*>(s_dyn + L.tab_c) : nullptr;
    float* s_rows = s_dyn + L.rows;
    uint32_t* s_pos = reinterpret_cast<uint32_t*>(s_dyn + L.pos);

    const int tid = threadIdx.x;
    const int team = tid / kLanes;
    const int lane = tid % kLanes;
    const int width = L.width;
    const int row = L.row;
    const int pitch = L.pitch;
    const int chunks = kExact ? kS : L.chunks;
    const int tail = kExact ? 0 : L.tail;
    const int md = d.model_dim;
    const bool final_stage = d.stage == kFinal;
    // kFlat: the flat-area test runs (Final with flat_areas): the sums it
    // takes over every retained row are compiled only there.
    constexpr bool want_flat = kFlat;
    // Slots per channel when one model covers whole 16-element slots of each
    // channel (coupled, patch_dim a multiple of 16), else 0.
    const int channel_runs = (d.models == 1 && d.patch_dim % kLanes == 0) ? d.patch_dim / kLanes : 0;

    for (int i = tid; i < width; i += kModelThreads) {
        const int m = i / row;
        const int rem = i - m * row;
        const int l = rem / pitch;
        const int k = rem - l * pitch;
        const bool valid = k < chunks || (k == chunks && l < tail);
        int in_off = 0;
        int channel = -1;
        if (valid) {
            const int e = (d.coupled ? 0 : m * d.patch_dim) + kLanes * k + l;
            const int c = e / d.patch_dim;
            const int wi = e - c * d.patch_dim;
            const int t = wi / d.patch_area;
            const int yx = wi - t * d.patch_area;
            const int y = yx / d.patch;
            const int x = yx - y * d.patch;
            in_off = ((t * d.channels + c) * d.height + y) * d.width + x;
            channel = c;
        }
        tab_in[i] = in_off;
        if (tab_c != nullptr) {
            tab_c[i] = channel;
        }
    }

    const int group_count = *group_count_dev;
    // Groups come from this XCD's contiguous eighth of the list (then from
    // the others'), so the groups in flight on an XCD are neighbours that
    // share its L2; the next index is requested while the current group is
    // processed.
    const int home = blockIdx.x % kXcds;
    int visited = 0;
    if (tid == 0) {
        s_next = xcd_next(work_counter, group_count, home, visited);
    }
    int following = 0;
    const auto advance = [&]() {
        __syncthreads();  // the group is done with the LDS and s_next
        if (tid == 0) {
            s_next = following;
        }
        __syncthreads();
        return s_next;
    };
    __syncthreads();
    for (int g = s_next; g < group_count; g = advance()) {
        if (tid == 0) {
            following = xcd_next(work_counter, group_count, home, visited);
        }
        const int a = refs.group_list[g];
        const int similar = refs.counts[a];
        const size_t list = static_cast<size_t>(a) * d.cap;
        const int rows = model_rows_of(d, similar);
        for (int r = tid; r < similar; r += kModelThreads) {
            s_pos[r] = refs.positions[list + r];
        }
        if (tid == 0) {
            s_equal = 1;
        }
        __syncthreads();

        // Noisy model rows into LDS; flat sums (double, per channel) over
        // every retained noisy row.
        double sums[kMaxChannels][2];
#pragma unroll
        for (int c = 0; c < kMaxChannels; ++c) {
            sums[c][0] = 0.0;
            sums[c][1] = 0.0;
        }
        for (int r = team; r < rows; r += kTeams) {
            float* dst = s_rows + r * width;
            load_row<kS, kExact>(d, L, noisy, patch_base(d, s_pos[r]), tab_in, dst, lane);
            if (want_flat && channel_runs > 0) {
                // One model over whole 16-element slots per channel: the
                // lane's slots k of channel c are c * runs .. (c + 1) * runs - 1
                // (same per-channel order as below).
                const int at = lane * pitch;
#pragma unroll
                for (int cc = 0; cc < kMaxChannels; ++cc) {
                    if (cc < d.channels) {
                        double s0 = sums[cc][0];
                        double s1 = sums[cc][1];
                        for (int k = cc * channel_runs; k < (cc + 1) * channel_runs; ++k) {
                            const double v = static_cast<double>(dst[at + k]);
                            s0 += v;
                            s1 += v * v;
                        }
                        sums[cc][0] = s0;
                        sums[cc][1] = s1;
                    }
                }
            } else if (want_flat) {
                for (int m = 0; m < d.models; ++m) {
                    const int at = m * row + lane * pitch;
                    for (int k = 0; k <= chunks; ++k) {
                        if (k < chunks || lane < tail) {
                            const int c = tab_c[at + k];
                            const double v = static_cast<double>(dst[at + k]);
#pragma unroll
                            for (int cc = 0; cc < kMaxChannels; ++cc) {
                                if (cc == c) {
                                    sums[cc][0] += v;
                                    sums[cc][1] += v * v;
                                }
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();

        // Per element (one thread per slot): equality of the model rows
        // with row 0 and the noisy mean over the model rows in row order.
        const float inv_rows = 1.0F / static_cast<float>(rows);
        float* mean_noisy = means + static_cast<size_t>(g) * 2 * d.dim;
        float* mean_basic = mean_noisy + d.dim;
        bool equal_local = true;
        if (kExact) {
            slot_means4<kS>(d, L, s_rows, rows, inv_rows, mean_noisy, !final_stage, equal_local, tid);
        }
        for (int i = tid; i < (kExact ? 0 : width); i += kModelThreads) {
            if (tab_c[i] < 0) {
                continue;
            }
            const float v0 = s_rows[i];
            float acc = 0.0F;
            for (int r = 0; r < rows; ++r) {
                const float v = s_rows[r * width + i];
                equal_local = equal_local && (v == v0);
                acc = acc + v;
            }
            const float mean = acc * inv_rows;
            const int m = i / row;
            const int rem = i - m * row;
            const int l = rem / pitch;
            const int k = rem - l * pitch;
            mean_noisy[(d.coupled ? 0 : m * d.patch_dim) + kLanes * k + l] = mean;
            if (!final_stage) {
                // Basic: the model rows are the noisy rows, centered.
                for (int r = 0; r < rows; ++r) {
                    s_rows[r * width + i] = s_rows[r * width + i] - mean;
                }
            }
        }
        if (!equal_local) {
            s_equal = 0;
        }
        __syncthreads();
        // Retained rows beyond the model rows: equality (only while all
        // rows so far equal row 0) and the flat sums.
        if (similar > rows && (want_flat || s_equal != 0)) {
            const bool check_equal = s_equal != 0;
            const size_t base0 = patch_base(d, s_pos[0]);
            for (int r = rows + team; r < similar; r += kTeams) {
                const size_t base = patch_base(d, s_pos[r]);
                for (int m = 0; m < d.models; ++m) {
                    const int at = m * row + lane * pitch;
                    bool same = true;
                    // The row's samples (kExact: every load issued first).
                    float staged[kExact ? kS : 1];
                    if (kExact) {
                        const int4* t4 = reinterpret_cast<const int4*>(tab_in + at);
#pragma unroll
                        for (int h = 0; h < kS / 4; ++h) {
                            const int4 o = t4[h];
                            staged[4 * h + 0] = noisy[base + o.x];
                            staged[4 * h + 1] = noisy[base + o.y];
                            staged[4 * h + 2] = noisy[base + o.z];
                            staged[4 * h + 3] = noisy[base + o.w];
                        }
                    }
#pragma unroll
                    for (int k = 0; k <= chunks; ++k) {
                        if (k < chunks || lane < tail) {
                            const float v = kExact ? staged[k < kS ? k : 0] : noisy[base + tab_in[at + k]];
                            if (want_flat) {
                                const int c = tab_c[at + k];
                                const double dv = static_cast<double>(v);
#pragma unroll
                                for (int cc = 0; cc < kMaxChannels; ++cc) {
                                    if (cc == c) {
                                        sums[cc][0] += dv;
                                        sums[cc][1] += dv * dv;
                                    }
                                }
                            }
                            if (check_equal) {
                                // Row 0 from the source: Basic rows in LDS
                                // are centered by now.
                                same = same && v == noisy[base0 + tab_in[at + k]];
                            }
                        }
                    }
                    if (check_equal && !same) {
                        s_equal = 0;
                    }
                }
            }
        }
        const int wave = tid >> 6;
        if (want_flat) {
#pragma unroll
            for (int c = 0; c < kMaxChannels; ++c) {
                const double s0 = wave_sum_double(sums[c][0]);
                const double s1 = wave_sum_double(sums[c][1]);
                if ((tid & 63) == 0) {
                    s_sum[wave][c][0] = s0;
                    s_sum[wave][c][1] = s1;
                }
            }
        }
        __syncthreads();

        const bool equal = similar > 1 && s_equal != 0;
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
        const bool modelled = !equal && d.rank > 0;
        const bool dual = group_uses_dual(d, rows);
        if (tid == 0) {
            if (rows < (md < d.rank ? md : d.rank)) {
                flags |= 8;  // rows cannot cover the rank: reported by the host
            }
            group_flags[g] = flags | (modelled ? kFlagModelled : 0);
            problem_dims[g] = dual ? rows : md;
        }

        if (final_stage) {
            // The basic mean is needed by the filter even when the group is
            // not modelled (flat centre); the model rows are the basic rows.
            for (int r = team; r < rows; r += kTeams) {
                load_row<kS, kExact>(d, L, basic, patch_base(d, s_pos[r]), tab_in, s_rows + r * width, lane);
            }
            __syncthreads();
            if (kExact) {
                bool unused = true;
                slot_means4<kS>(d, L, s_rows, rows, inv_rows, mean_basic, true, unused, tid);
            }
            for (int i = tid; i < (kExact ? 0 : width); i += kModelThreads) {
                if (tab_c[i] < 0) {
                    continue;
                }
                float acc = 0.0F;
                for (int r = 0; r < rows; ++r) {
                    acc = acc + s_rows[r * width + i];
                }
                acc = acc * inv_rows;
                const int m = i / row;
                const int rem = i - m * row;
                const int l = rem / pitch;
                const int k = rem - l * pitch;
                mean_basic[(d.coupled ? 0 : m * d.patch_dim) + kLanes * k + l] = acc;
                for (int r = 0; r < rows; ++r) {
                    s_rows[r * width + i] = s_rows[r * width + i] - acc;
                }
            }
            __syncthreads();
        }
        if (!modelled) {
            continue;
        }

        const float scale = 1.0F / static_cast<float>(rows);
        for (int m = 0; m < d.models; ++m) {
            float* out = problems + (static_cast<size_t>(g) * d.models + m) * problem_stride * problem_stride;
            if (kMfma && dual) {
                // G = X X^T / rows with X the centered model rows (at most 16):
                // v_mfma_f32_16x16x32_bf16 over the elements in blocks of 8
                // slots of one lane column (any element order gives the same
                // sum; A and B share the lane's registers since B = A^T),
                // three products hi.hi + hi.lo + lo.hi.  Wavefront w takes the
                // k steps q = w mod 4; the four partial tiles are summed in
                // wavefront order through LDS.
                {
                    const int wave = tid >> 6;
                    const int col = tid & 15;
                    const int kb = (tid & 63) >> 4;
                    const int blocks_per_lane = kS / 8;
                    floatx4 acc = {0.0F, 0.0F, 0.0F, 0.0F};
                    for (int q = wave; q < kS / 2; q += kModelThreads / 64) {
                        const int b = 4 * q + kb;
                        const int src_lane = b / blocks_per_lane;
                        const int k0 = 8 * (b - src_lane * blocks_per_lane);
                        float4 lo4 = {0.0F, 0.0F, 0.0F, 0.0F};
                        float4 hi4 = {0.0F, 0.0F, 0.0F, 0.0F};
                        if (col < rows) {  // rows past the model rows are not staged
                            const float4* src = reinterpret_cast<const float4*>(
                                s_rows + col * width + m * row + src_lane * pitch + k0);
                            lo4 = src[0];
                            hi4 = src[1];
                        }
                        const float v[8] = {lo4.x, lo4.y, lo4.z, lo4.w, hi4.x, hi4.y, hi4.z, hi4.w};
                        bf16x8 a_hi;
                        bf16x8 a_lo;
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const float x = v[e];
                            const __bf16 h = static_cast<__bf16>(x);
                            a_hi[e] = h;
                            a_lo[e] = static_cast<__bf16>(x - static_cast<float>(h));
                        }
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, a_hi, acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, a_lo, acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo, a_hi, acc, 0, 0, 0);
                    }
                    // acc[v] = partial G[4 kb + v][col]
                    float* s_gram = s_dyn + L.gram;
                    if (L.gram == L.rows) {
                        __syncthreads();  // every wave is done reading the rows
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        s_gram[(wave * 4 + v) * 64 + (tid & 63)] = acc[v];
                    }
                    __syncthreads();
                    if (tid < 64) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            float total = s_gram[v * 64 + tid];
                            for (int w = 1; w < kModelThreads / 64; ++w) {
                                total = total + s_gram[(w * 4 + v) * 64 + tid];
                            }
                            const int r = 4 * kb + v;
                            if (r < rows && col < rows) {
                                out[r * problem_stride + col] = scale * total;
                            }
                        }
                    }
                    __syncthreads();  // the partials are rewritten by the next model
                }
            } else if (dual) {
                // compute_gram_centered (linalg.cpp:326-354): scale * dot in
                // the 16-lane order, one team per entry.
                const int pairs = rows * (rows + 1) / 2;
                for (int q = team; q < pairs; q += kTeams) {
                    // Row of the triangular index q: the float estimate,
                    // corrected to the exact integer.
                    int r = static_cast<int>((sqrtf(8.0F * static_cast<float>(q) + 1.0F) - 1.0F) * 0.5F);
                    while ((r + 1) * (r + 2) / 2 <= q) {
                        ++r;
                    }
                    while (r > 0 && r * (r + 1) / 2 > q) {
                        --r;
                    }
                    const int c = q - r * (r + 1) / 2;
                    const float* left = s_rows + r * width + m * row + lane * pitch;
                    const float* right = s_rows + c * width + m * row + lane * pitch;
                    const float4* l4 = reinterpret_cast<const float4*>(left);
                    const float4* r4 = reinterpret_cast<const float4*>(right);
                    float part = 0.0F;
#pragma unroll
                    for (int h = 0; h < kS / 4; ++h) {
                        if (4 * h < chunks) {
                            const float4 x = l4[h];
                            const float4 y = r4[h];
                            if (4 * h + 0 < chunks) part = fmaf(x.x, y.x, part);
                            if (4 * h + 1 < chunks) part = fmaf(x.y, y.y, part);
                            if (4 * h + 2 < chunks) part = fmaf(x.z, y.z, part);
                            if (4 * h + 3 < chunks) part = fmaf(x.w, y.w, part);
                        }
                    }
                    float total = segment16_sum(part);
                    if (tail > 0) {
                        const float product = lane < tail ? left[chunks] * right[chunks] : 0.0F;
                        for (int i = 0; i < tail; ++i) {
                            total = total + __shfl(product, i, kLanes);
                        }
                    }
                    if (lane == 0) {
                        const float value = scale * total;
                        out[r * problem_stride + c] = value;
                        out[c * problem_stride + r] = value;
                    }
                }
            } else {
                // compute_covariance_centered (linalg.cpp:377-414): per entry
                // a chain over samples, fused where add_scaled's vector part
                // covers the column.
                const int pairs = md * (md + 1) / 2;
                for (int q = tid; q < pairs; q += kModelThreads) {
                    int c0 = 0;
                    int remaining = q;
                    while (remaining >= md - c0) {
                        remaining -= md - c0;
                        ++c0;
                    }
                    const int c1 = c0 + remaining;
                    const int vector_end = ((md - c0) / kLanes) * kLanes;
                    const bool fused = (c1 - c0) < vector_end;
                    const int s0 = m * row + (c0 % kLanes) * pitch + c0 / kLanes;
                    const int s1 = m * row + (c1 % kLanes) * pitch + c1 / kLanes;
                    float acc = 0.0F;
                    for (int s2 = 0; s2 < rows; ++s2) {
                        const float left = s_rows[s2 * width + s0];
                        const float right = s_rows[s2 * width + s1];
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
}

}  // namespace

size_t prepare_shared_bytes(const StageDesc& d) {
    return static_cast<size_t>(prepare_layout(d).words) * sizeof(float);
}

hipError_t launch_prepare(const StageDesc& d, const float* noisy, const float* basic,
                          const uint32_t* anchors, const int* counts, const uint32_t* positions,
                          const int* group_list, const int* group_count_dev, int group_capacity,
                          int problem_stride, int* group_flags, int* problem_dims, float* means,
                          float* problems, bool mfma_gram, int* work_counter, hipStream_t stream) {
    if (group_capacity <= 0) {
        return hipSuccess;
    }
    const GroupRefs refs{anchors, counts, positions, group_list};
    const PrepareLayout L = prepare_layout(d);
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
    const int slots = template_slots(d.model_dim);
    const bool exact = d.model_dim == kLanes * slots;
    // The matrix-core Gram needs whole 8-slot blocks and at most 16 rows.
    const bool mfma = mfma_gram && exact && slots % 8 == 0 && d.max_rows <= 16;
    const bool flat = d.stage == kFinal && d.flat_areas != 0;
#define VNLBHIP_PREPARE4(KS, EXACT, MFMA, FLAT)                                                    \
    do {                                                                                           \
        int per_unit = 0;                                                                          \
        status = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_unit,                           \
                                                              prepare_kernel<KS, EXACT, MFMA, FLAT>, \
                                                              kModelThreads, shared);              \
        if (status != hipSuccess) {                                                                \
            return status;                                                                         \
        }                                                                                          \
        const long long resident = static_cast<long long>(per_unit > 0 ? per_unit : 1) * units;    \
        const unsigned blocks =                                                                    \
            static_cast<unsigned>(resident < group_capacity ? resident : group_capacity);          \
        hipLaunchKernelGGL((prepare_kernel<KS, EXACT, MFMA, FLAT>), dim3(blocks), dim3(kModelThreads), shared, \
                           stream, d, L, noisy, basic, refs, group_count_dev, problem_stride,      \
                           group_flags, problem_dims, means, problems, work_counter);              \
    } while (0)
#define VNLBHIP_PREPARE3(KS, EXACT, MFMA)                                                          \
    if (flat) {                                                                                    \
        VNLBHIP_PREPARE4(KS, EXACT, MFMA, true);                                                   \
    } else {                                                                                       \
        VNLBHIP_PREPARE4(KS, EXACT, MFMA, false);                                                  \
    }
#define VNLBHIP_PREPARE(KS, EXACT)                                                                 \
    if (EXACT && mfma) {                                                                           \
        VNLBHIP_PREPARE3(KS, EXACT, true);                                                         \
    } else {                                                                                       \
        VNLBHIP_PREPARE3(KS, EXACT, false);                                                        \
    }
#define VNLBHIP_PREPARE_SLOTS(KS)      \
    if (exact) {                       \
        VNLBHIP_PREPARE(KS, true);     \
    } else {                           \
        VNLBHIP_PREPARE(KS, false);    \
    }
    switch (slots) {
        case 4: VNLBHIP_PREPARE_SLOTS(4); break;
        case 8: VNLBHIP_PREPARE_SLOTS(8); break;
        case 16: VNLBHIP_PREPARE_SLOTS(16); break;
        case 24: VNLBHIP_PREPARE_SLOTS(24); break;
        case 32: VNLBHIP_PREPARE_SLOTS(32); break;
        default: VNLBHIP_PREPARE_SLOTS(64); break;
    }
#undef VNLBHIP_PREPARE_SLOTS
#undef VNLBHIP_PREPARE
#undef VNLBHIP_PREPARE3
#undef VNLBHIP_PREPARE4
    return hipGetLastError();
}

}  // namespace vnlbhip
