// Block matching, MFMA form.
//
// The squared distance of a candidate patch c to an anchor patch r is
//
//     ||r - c||^2 = E(r) + E(c) - 2 <r, c>
//
// with E the patch energy (fp32, one value per patch origin) and the cross
// term accumulated in fp32 on the matrix cores from three bf16 products
// (x = hi + lo, <r, c> ~ hi.hi + hi.lo + lo.hi).  Samples are offset by a
// per-channel constant first, which leaves distances unchanged and keeps the
// energies close to the distances they are subtracted from.  The candidates
// kept are then chosen by the same k-nearest selection as the exact form
// (select.h), so only the distance values differ from it.
//
// Anchors are matched in tiles of up to 16 anchors of one frame (4 grid rows
// x 4 anchors; 2 x 4 where the union of wider windows would not fit): the
// tile's candidates are the union of the anchors' windows.
// One v_mfma_f32_16x16x32_bf16 multiplies 16 candidates of a column
// (consecutive rows) by the 16 anchors over 4 patch rows.  The search region
// of one (frame, channel) plane is staged in LDS as bf16 hi / lo, four
// copies shifted by 0..3 samples so that every lane reads its 8 samples with
// 8-byte aligned loads.  Supported for 8x8 patches.
#include "common.h"
#include "device_util.h"
#include "kernels.h"
#include "select.h"

#include <float.h>
#include <hip/hip_bf16.h>

namespace vnlbhip {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef __attribute__((__vector_size__(4 * sizeof(unsigned)))) unsigned u32x4;

constexpr int kFastThreads = 256;
constexpr int kFastWaves = kFastThreads / 64;
constexpr int kFastPatch = 8;
constexpr int kTileAnchors = 16;
constexpr int kCopies = 4;
constexpr int kFillPerThread = 3;          // fill items (4 output columns) per thread
constexpr int kFillSpan = 4 + kCopies - 1;  // samples behind one item
constexpr int kEnergyPerThread = 5;         // energy box items per thread

// Per-channel offset: the mean of the finite values of a strided sample of
// the channel (any finite offset leaves the distances unchanged; non-finite
// samples only poison the patches that hold them).
__global__ __launch_bounds__(256) void channel_offset_kernel(const float* __restrict__ source,
                                                             int channels, size_t plane,
                                                             int frames, float* __restrict__ offsets) {
    __shared__ double s_part[4];
    const int c = blockIdx.x;
    const size_t per_channel = plane * frames;
    const size_t stride = per_channel / 65536 + 1;
    double sum = 0.0;
    size_t count = 0;
    for (size_t i = threadIdx.x * stride; i < per_channel; i += blockDim.x * stride) {
        const size_t f = i / plane;
        const size_t rem = i - f * plane;
        const float v = source[(f * channels + c) * plane + rem];
        if (v == v && fabsf(v) <= FLT_MAX) {
            sum += static_cast<double>(v);
            ++count;
        }
    }
    sum = wave_sum_double(sum);
    double n = wave_sum_double(static_cast<double>(count));
    if ((threadIdx.x & 63) == 0) {
        s_part[threadIdx.x >> 6] = sum;
    }
    __syncthreads();
    __shared__ double s_count[4];
    if ((threadIdx.x & 63) == 0) {
        s_count[threadIdx.x >> 6] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double total = 0.0;
        double total_n = 0.0;
        for (int w = 0; w < 4; ++w) {
            total += s_part[w];
            total_n += s_count[w];
        }
        offsets[c] = total_n > 0.0 ? static_cast<float>(total / total_n) : 0.0F;
    }
}

// Horizontal sums of squares: rows[f][c][y][x] = sum_{px<P} (v - offset)^2.
__global__ __launch_bounds__(256) void row_energy_kernel(StageDesc d, const float* __restrict__ source,
                                                         const float* __restrict__ offsets,
                                                         int match_channels, float* __restrict__ rows) {
    const size_t plane = static_cast<size_t>(d.width) * d.height;
    const size_t total = plane * match_channels * d.frames;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) {
        return;
    }
    const size_t f = i / (plane * match_channels);
    const size_t rem = i - f * plane * match_channels;
    const int c = static_cast<int>(rem / plane);
    const size_t p = rem - static_cast<size_t>(c) * plane;
    const int y = static_cast<int>(p / d.width);
    const int x = static_cast<int>(p - static_cast<size_t>(y) * d.width);
    float sum = 0.0F;
    if (x <= d.width - d.patch) {
        const float mu = offsets[c];
        const float* src = source + ((f * d.channels + c) * d.height + y) * d.width + x;
        for (int px = 0; px < d.patch; ++px) {
            const float v = src[px] - mu;
            sum = fmaf(v, v, sum);
        }
    }
    rows[i] = sum;
}

// Patch energies: energy[f][y][x] = sum over channels, frames and rows of
// the row sums, for every patch origin.
__global__ __launch_bounds__(256) void patch_energy_kernel(StageDesc d, const float* __restrict__ rows,
                                                           int match_channels, float* __restrict__ energy) {
    const size_t plane = static_cast<size_t>(d.width) * d.height;
    const size_t total = plane * d.frames;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) {
        return;
    }
    const int f = static_cast<int>(i / plane);
    const size_t p = i - static_cast<size_t>(f) * plane;
    const int y = static_cast<int>(p / d.width);
    const int x = static_cast<int>(p - static_cast<size_t>(y) * d.width);
    float sum = 0.0F;
    if (x <= d.width - d.patch && y <= d.height - d.patch && f <= d.frames - d.patch_time) {
        for (int t = 0; t < d.patch_time; ++t) {
            for (int c = 0; c < match_channels; ++c) {
                const float* r = rows + ((static_cast<size_t>(f + t) * match_channels + c) * d.height + y) * d.width + x;
                for (int py = 0; py < d.patch; ++py) {
                    sum += r[static_cast<size_t>(py) * d.width];
                }
            }
        }
    }
    energy[i] = sum;
}

struct FastShape {
    int row_tiles;     // 16-row candidate tiles per column
    int region_rows;   // 16 * row_tiles + P - 1
    int pitch;         // bf16 samples per LDS copy row (multiple of 4)
    int copy_elems;    // region_rows * pitch
    int stride;        // distance slots per anchor (window^2 * temporal)
    int per_frame;     // window^2
    int ebox_pitch;    // energy box column stride: 16 * row_tiles + 4 (16-byte rows, banks spread)
    int ebox_words;    // energy box, column-major: max cols x ebox_pitch
    int scratch_keys;  // compaction list per anchor (window^2 16-bit entries, rounded to 2)
};

__device__ inline uint2 pack4(__bf16 a, __bf16 b, __bf16 c, __bf16 d) {
    const unsigned lo = static_cast<unsigned>(__builtin_bit_cast(unsigned short, a)) |
                        (static_cast<unsigned>(__builtin_bit_cast(unsigned short, b)) << 16);
    const unsigned hi = static_cast<unsigned>(__builtin_bit_cast(unsigned short, c)) |
                        (static_cast<unsigned>(__builtin_bit_cast(unsigned short, d)) << 16);
    return uint2{lo, hi};
}

__device__ inline bf16x8 load_bf16x8(const __bf16* p) {
    // 8-byte aligned: two 64-bit LDS reads.
    const uint2* q = reinterpret_cast<const uint2*>(p);
    const uint2 a = q[0];
    const uint2 b = q[1];
    const u32x4 v = {a.x, a.y, b.x, b.y};
    return __builtin_bit_cast(bf16x8, v);
}

// One compare-exchange step of a bitonic network over the wavefront:
// partner lane ^ kJ, blocks of kK lanes alternately ascending (kK = 64: the
// whole wavefront ascending).
template <int kK, int kJ>
__device__ inline uint64_t bitonic_step(uint64_t key, int lane) {
    const uint64_t other = xor64<kJ>(key);
    // Keep the smaller key where the lane is the lower end of an ascending
    // pair or the upper end of a descending one (keys are distinct).
    const bool want_low = ((lane & kJ) == 0) == ((lane & kK) == 0);
    return (other < key) == want_low ? other : key;
}

// Wave-wide bitonic sort of 64 keys, one per lane, ascending by lane.
__device__ inline uint64_t wave_sort64(uint64_t key, int lane) {
    key = bitonic_step<2, 1>(key, lane);
    key = bitonic_step<4, 2>(key, lane);
    key = bitonic_step<4, 1>(key, lane);
    key = bitonic_step<8, 4>(key, lane);
    key = bitonic_step<8, 2>(key, lane);
    key = bitonic_step<8, 1>(key, lane);
    key = bitonic_step<16, 8>(key, lane);
    key = bitonic_step<16, 4>(key, lane);
    key = bitonic_step<16, 2>(key, lane);
    key = bitonic_step<16, 1>(key, lane);
    key = bitonic_step<32, 16>(key, lane);
    key = bitonic_step<32, 8>(key, lane);
    key = bitonic_step<32, 4>(key, lane);
    key = bitonic_step<32, 2>(key, lane);
    key = bitonic_step<32, 1>(key, lane);
    key = bitonic_step<64, 32>(key, lane);
    key = bitonic_step<64, 16>(key, lane);
    key = bitonic_step<64, 8>(key, lane);
    key = bitonic_step<64, 4>(key, lane);
    key = bitonic_step<64, 2>(key, lane);
    key = bitonic_step<64, 1>(key, lane);
    return key;
}

// The 64 smallest of two ascending wave-wide lists, ascending: the first
// half-cleaner of the bitonic merge of r and reversed c, then the merge.
__device__ inline uint64_t wave_merge_lower64(uint64_t r, uint64_t c, int lane) {
    const uint64_t reversed = __shfl(c, 63 - lane, 64);
    uint64_t key = reversed < r ? reversed : r;
    key = bitonic_step<64, 32>(key, lane);
    key = bitonic_step<64, 16>(key, lane);
    key = bitonic_step<64, 8>(key, lane);
    key = bitonic_step<64, 4>(key, lane);
    key = bitonic_step<64, 2>(key, lane);
    key = bitonic_step<64, 1>(key, lane);
    return key;
}

// Four independent lists at once (one per anchor of a wavefront): the
// exchanges of a step are issued together so their latencies overlap.
template <int kK, int kJ>
__device__ inline void bitonic_step4(uint64_t (&key)[4], int lane) {
    uint64_t other[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        other[q] = xor64<kJ>(key[q]);
    }
    const bool want_low = ((lane & kJ) == 0) == ((lane & kK) == 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        key[q] = (other[q] < key[q]) == want_low ? other[q] : key[q];
    }
}

__device__ inline void wave_sort64x4(uint64_t (&key)[4], int lane) {
    bitonic_step4<2, 1>(key, lane);
    bitonic_step4<4, 2>(key, lane);
    bitonic_step4<4, 1>(key, lane);
    bitonic_step4<8, 4>(key, lane);
    bitonic_step4<8, 2>(key, lane);
    bitonic_step4<8, 1>(key, lane);
    bitonic_step4<16, 8>(key, lane);
    bitonic_step4<16, 4>(key, lane);
    bitonic_step4<16, 2>(key, lane);
    bitonic_step4<16, 1>(key, lane);
    bitonic_step4<32, 16>(key, lane);
    bitonic_step4<32, 8>(key, lane);
    bitonic_step4<32, 4>(key, lane);
    bitonic_step4<32, 2>(key, lane);
    bitonic_step4<32, 1>(key, lane);
    bitonic_step4<64, 32>(key, lane);
    bitonic_step4<64, 16>(key, lane);
    bitonic_step4<64, 8>(key, lane);
    bitonic_step4<64, 4>(key, lane);
    bitonic_step4<64, 2>(key, lane);
    bitonic_step4<64, 1>(key, lane);
}

// best[q] <- the 64 smallest of best[q] and sorted[q], ascending.
__device__ inline void wave_merge_lower64x4(uint64_t (&best)[4], const uint64_t (&sorted)[4], int lane) {
    uint64_t key[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint64_t reversed = __shfl(sorted[q], 63 - lane, 64);
        key[q] = reversed < best[q] ? reversed : best[q];
    }
    bitonic_step4<64, 32>(key, lane);
    bitonic_step4<64, 16>(key, lane);
    bitonic_step4<64, 8>(key, lane);
    bitonic_step4<64, 4>(key, lane);
    bitonic_step4<64, 2>(key, lane);
    bitonic_step4<64, 1>(key, lane);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        best[q] = key[q];
    }
}

constexpr uint64_t kNoKey = ~static_cast<uint64_t>(0);

// Keys per anchor up to which a frame's candidates are inserted one by one
// rather than sorted and merged.
constexpr int kInsertKeys = 24;

// `key` into a list of distinct keys ascending over the lanes (the 64
// smallest so far; kNoKey for empty places): the keys below it stay, the
// rest move up one lane and the largest drops out.
__device__ inline uint64_t wave_insert(uint64_t list, uint64_t key, int lane) {
    const int place = __popcll(__ballot(list < key));
    const uint64_t before = __shfl_up(list, 1, 64);
    return lane < place ? list : (lane == place ? key : before);
}

__device__ inline uint64_t readlane64(uint64_t v, int lane) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v & 0xFFFFFFFFu), lane));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), lane));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Exclusive bound on the keys that can still be selected, for a running
// list `best` (ascending over the lanes): a key at or above the 64th can not
// enter it, and a key above both the K-th so far and tau can not be
// selected (the final threshold max(tau, d_K) only falls as candidates
// arrive, and k = min(K, valid) <= K).
__device__ inline uint64_t admit_bound(uint64_t best, int similar, uint64_t tau_key) {
    const uint64_t last = readlane64(best, 63);
    const uint64_t kth = readlane64(best, similar - 1);
    const uint64_t reach = kth > tau_key ? kth : tau_key;
    // keys <= reach may be selected: as an exclusive bound, reach + 1
    // (reach < kNoKey unless nothing is held yet).
    const uint64_t open = reach == kNoKey ? kNoKey : reach + 1;
    return open < last ? open : last;
}

// Cross terms of one anchor tile against its candidates, all candidate
// frames.  kFused = false: distances land in `out` ([anchor][stride], scan
// order) for select_fast_kernel.  kFused = true (retained cap <= 64): the
// k-nearest selection runs here.  Each wavefront keeps, for four anchors of
// the tile, the 64 smallest keys (distance key << 32 | scan position) seen
// so far in lane order; after every candidate frame the candidates below the
// 64th key are compacted, sorted and merged in.  The selected set of
// select.h is a prefix of the (distance, scan position) order no longer
// than the retained cap, so it lies within those 64 keys, and the lists
// written are those of select_and_write.
template <int kTilesPerWave, bool kFused>
__global__ __launch_bounds__(kFastThreads) void match_fast_kernel(
    StageDesc d, FastShape S, const float* __restrict__ source, const float* __restrict__ offsets,
    const float* __restrict__ energy, const uint32_t* __restrict__ anchors,
    const int* __restrict__ tiles, int anchor_base, float* __restrict__ out, int* __restrict__ counts,
    uint32_t* __restrict__ positions, float* __restrict__ distances) {
    extern __shared__ uint4 s_dyn16[];
    __bf16* s_copy = reinterpret_cast<__bf16*>(s_dyn16);  // [buffer][hi/lo][copy][rows][pitch]
    // Epilogue staging after the two plane buffers.
    float* s_stage = reinterpret_cast<float*>(s_copy + static_cast<size_t>(4) * kCopies * S.copy_elems);
    // Patch energies of the candidate frame over the tile's window union
    // go into the plane buffer that the epilogue's plane has finished with.
    __shared__ int s_index[kTileAnchors];
    __shared__ int s_ax[kTileAnchors];
    __shared__ int s_ay[kTileAnchors];
    __shared__ int s_xl[kTileAnchors];
    __shared__ int s_yl[kTileAnchors];
    __shared__ int s_nx[kTileAnchors];
    __shared__ int s_ny[kTileAnchors];
    __shared__ float s_energy[kTileAnchors];
    __shared__ int s_box[5];

    const int tid = threadIdx.x;
    // Uniform per wavefront: tile indices derived from it stay scalar.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int* tile = tiles + static_cast<size_t>(blockIdx.x) * kTileAnchors;
    const bool final_stage = d.stage == kFinal;
    const int match_channels = final_stage ? d.channels : 1;

    if (tid < kTileAnchors) {
        const int index = tile[tid];
        s_index[tid] = index;
        if (index >= 0) {
            const uint32_t p = anchors[index];
            const int ax = position_x(p);
            const int ay = position_y(p);
            int xl = 0;
            int xh = 0;
            int yl = 0;
            int yh = 0;
            spatial_range(d, ax, d.width, xl, xh);
            spatial_range(d, ay, d.height, yl, yh);
            s_ax[tid] = ax;
            s_ay[tid] = ay;
            s_xl[tid] = xl;
            s_yl[tid] = yl;
            s_nx[tid] = xh - xl + 1;
            s_ny[tid] = yh - yl + 1;
            s_energy[tid] = energy[(static_cast<size_t>(position_frame(p)) * d.height + ay) * d.width + ax];
        }
    }
    __syncthreads();
    if (tid == 0) {
        int x0 = 1 << 30;
        int x1 = -1;
        int y0 = 1 << 30;
        int y1 = -1;
        int first = -1;
        for (int n = 0; n < kTileAnchors; ++n) {
            if (s_index[n] >= 0) {
                first = first < 0 ? s_index[n] : first;
                x0 = min(x0, s_xl[n]);
                x1 = max(x1, s_xl[n] + s_nx[n] - 1);
                y0 = min(y0, s_yl[n]);
                y1 = max(y1, s_yl[n] + s_ny[n] - 1);
            }
        }
        s_box[0] = x0;
        s_box[1] = x1;
        s_box[2] = y0;
        s_box[3] = y1;
        s_box[4] = first;
    }
    __syncthreads();
    if (s_box[4] < 0) {
        return;  // empty tile (the host never emits one)
    }
    const int af = position_frame(anchors[s_box[4]]);
    const int ux0 = s_box[0];
    const int uy0 = s_box[2];
    const int ucols = s_box[1] - ux0 + 1;
    const int urows = s_box[3] - uy0 + 1;
    const int row_tiles = (urows + 15) / 16;  // <= S.row_tiles (host checked)
    const int mtiles = ucols * row_tiles;     // <= kTilesPerWave * waves (host checked)
    // col = mt / row_tiles as (mt * magic) >> 16, exact for mt < 32768.
    const int tile_magic = (65536 + row_tiles - 1) / row_tiles;
    int t_low = 0;
    int t_high = 0;
    temporal_range(d, af, t_low, t_high);
    const int nt = t_high - t_low + 1;
    const size_t plane = static_cast<size_t>(d.width) * d.height;

    const int n_lane = lane & 15;  // anchor (B column) / candidate row (A row)
    const int kb = lane >> 4;      // 8-sample block of the 32-wide k step
    const int my_anchor = s_index[n_lane];
    const int my_ax = my_anchor >= 0 ? s_ax[n_lane] : 0;
    const int my_ay = my_anchor >= 0 ? s_ay[n_lane] : 0;

    // Planes (candidate frame ti, channel c, patch frame t) in order, the
    // region of plane p + 1 loaded into registers while plane p is on the
    // matrix cores, then stored into the other LDS buffer.
    const int planes_per_frame = match_channels * d.patch_time;
    const int total_planes = nt * planes_per_frame;
    const int groups4 = S.pitch / 4;
    const int fill_items = S.region_rows * groups4;  // <= kFillPerThread * threads (host checked)
    float pre[kFillPerThread][kFillSpan];

    const auto plane_of = [&](int p, int& frame, int& c, int& t) {
        const int ti = p / planes_per_frame;
        const int rem = p - ti * planes_per_frame;
        c = rem / d.patch_time;
        t = rem - c * d.patch_time;
        frame = scheduled_frame(af, t_low, t_high, ti);
    };
    const auto prefetch = [&](int p) {
        int frame = 0;
        int c = 0;
        int t = 0;
        plane_of(p, frame, c, t);
        const float mu = offsets[c];
        const float* src = source + (static_cast<size_t>(frame + t) * d.channels + c) * plane;
#pragma unroll
        for (int k = 0; k < kFillPerThread; ++k) {
            const int item = tid + k * kFastThreads;
            const int ry = item / groups4;
            const int x0 = ux0 + 4 * (item - ry * groups4);
            const int y = uy0 + ry;
            const bool row_ok = item < fill_items && y < d.height;
            const float* row_src = src + static_cast<size_t>(y) * d.width + x0;
            if (row_ok && x0 + 8 <= d.width) {
                // Interior: two 16-byte loads (4-byte aligned) cover the span.
                float4 a;
                float4 b;
                __builtin_memcpy(&a, row_src, sizeof(a));
                __builtin_memcpy(&b, row_src + 4, sizeof(b));
                const float span[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
                for (int v = 0; v < kFillSpan; ++v) {
                    pre[k][v] = span[v] - mu;
                }
            } else {
#pragma unroll
                for (int v = 0; v < kFillSpan; ++v) {
                    float value = 0.0F;
                    if (row_ok && x0 + v < d.width) {
                        value = row_src[v] - mu;
                    }
                    pre[k][v] = value;
                }
            }
        }
    };
    const auto store = [&](int buffer) {
        __bf16* base = s_copy + static_cast<size_t>(buffer) * 2 * kCopies * S.copy_elems;
#pragma unroll
        for (int k = 0; k < kFillPerThread; ++k) {
            const int item = tid + k * kFastThreads;
            if (item < fill_items) {
                const int ry = item / groups4;
                const int j = 4 * (item - ry * groups4);
                __bf16 hi[kFillSpan];
                __bf16 lo[kFillSpan];
#pragma unroll
                for (int v = 0; v < kFillSpan; ++v) {
                    hi[v] = static_cast<__bf16>(pre[k][v]);
                    lo[v] = static_cast<__bf16>(pre[k][v] - static_cast<float>(hi[v]));
                }
#pragma unroll
                for (int q = 0; q < kCopies; ++q) {
                    // copy q holds plane[ry][j + q .. j + q + 3] at columns j .. j + 3
                    const int at = q * S.copy_elems + ry * S.pitch + j;
                    *reinterpret_cast<uint2*>(base + at) = pack4(hi[q], hi[q + 1], hi[q + 2], hi[q + 3]);
                    *reinterpret_cast<uint2*>(base + kCopies * S.copy_elems + at) =
                        pack4(lo[q], lo[q + 1], lo[q + 2], lo[q + 3]);
                }
            }
        }
    };

    // Anchor operands (anchor n_lane, patch rows 4 ks + kb of plane
    // (af + t, c), offset by the channel mean) for the next plane.
    float pre_anchor[2][8];
    const auto prefetch_anchor = [&](int p) {
        int frame = 0;
        int c = 0;
        int t = 0;
        plane_of(p, frame, c, t);
        const float mu = offsets[c];
        const float* anchor_plane = source + (static_cast<size_t>(af + t) * d.channels + c) * plane;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int py = 4 * ks + kb;
            // The anchor patch lies inside the frame: 8 samples, two 16-byte loads.
            float4 a = {0.0F, 0.0F, 0.0F, 0.0F};
            float4 b = {0.0F, 0.0F, 0.0F, 0.0F};
            if (my_anchor >= 0) {
                const float* row = anchor_plane + static_cast<size_t>(my_ay + py) * d.width + my_ax;
                __builtin_memcpy(&a, row, sizeof(a));
                __builtin_memcpy(&b, row + 4, sizeof(b));
            }
            const float span[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                pre_anchor[ks][v] = my_anchor >= 0 ? span[v] - mu : 0.0F;
            }
        }
    };
    // Energies of a candidate frame over the union box, kEnergyPerThread
    // items per thread (host checked), for the epilogue.
    const int box_items = urows * ucols;
    float pre_energy[kEnergyPerThread];
    const auto prefetch_energy = [&](int frame) {
        const float* frame_energy = energy + static_cast<size_t>(frame) * plane;
#pragma unroll
        for (int k = 0; k < kEnergyPerThread; ++k) {
            int item = tid + k * kFastThreads;
            // Opaque: the box offsets are recomputed per frame instead of
            // being held in registers across the plane loop.
            asm volatile("" : "+v"(item));
            item = item < box_items ? item : box_items - 1;  // a cell of the box
            const int ry = item / ucols;
            const int cx = item - ry * ucols;
            pre_energy[k] = frame_energy[(uy0 + ry) * d.width + ux0 + cx];
        }
    };

    // Operands of tile i, patch rows 4 ks .. 4 ks + 3: mt clamped to a valid
    // tile so every LDS read stays in the staged region (accumulators of
    // tiles past mtiles are never used).
    const auto load_tile = [&](const __bf16* copies, int i, int ks, bf16x8& a_hi, bf16x8& a_lo) {
        int mt = wave + kFastWaves * i;
        // Opaque to loop-invariant motion: the few scalar address operations
        // are redone per plane instead of holding every tile's addresses.
        asm volatile("" : "+s"(mt));
        mt = mt < mtiles ? mt : mtiles - 1;
        const int col = (mt * tile_magic) >> 16;
        const int rt = mt - col * row_tiles;
        const int shift = col & (kCopies - 1);
        const int j = col - shift;
        const int ry = 16 * rt + n_lane + 4 * ks + kb;
        const int at = shift * S.copy_elems + ry * S.pitch + j;
        a_hi = load_bf16x8(copies + at);
        a_lo = load_bf16x8(copies + kCopies * S.copy_elems + at);
    };

    floatx4 acc[kTilesPerWave];
    uint64_t best[4] = {kNoKey, kNoKey, kNoKey, kNoKey};  // kFused: anchors 4 wave + q
    // Keys at or below tau (any scan position) may be selected whatever d_K is.
    const uint64_t tau_key = (static_cast<uint64_t>(distance_key(d.tau)) << 32) | 0xFFFFFFFFull;
    int valid_count[4] = {0, 0, 0, 0};
    prefetch(0);
    prefetch_anchor(0);
    store(0);
    float cur_anchor[2][8];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            cur_anchor[ks][v] = pre_anchor[ks][v];
        }
    }
    __syncthreads();
    for (int p = 0; p < total_planes; ++p) {
        int frame = 0;
        int c = 0;
        int t = 0;
        plane_of(p, frame, c, t);
        const bool frame_done = c == match_channels - 1 && t == d.patch_time - 1;
        if (c == 0 && t == 0) {
#pragma unroll
            for (int i = 0; i < kTilesPerWave; ++i) {
                acc[i] = floatx4{0.0F, 0.0F, 0.0F, 0.0F};
            }
        }
        if (p + 1 < total_planes) {
            prefetch(p + 1);
            prefetch_anchor(p + 1);
        }
        if (frame_done) {
            prefetch_energy(frame);
        }
        bf16x8 b_hi[2];
        bf16x8 b_lo[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const float x = cur_anchor[ks][v];
                const __bf16 hi = static_cast<__bf16>(x);
                b_hi[ks][v] = hi;
                b_lo[ks][v] = static_cast<__bf16>(x - static_cast<float>(hi));
            }
        }
        const __bf16* copies = s_copy + static_cast<size_t>(p & 1) * 2 * kCopies * S.copy_elems;
        // Pairs of tiles, their products alternating so consecutive MFMAs
        // are independent; the LDS reads of the next pair's half (ks) go out
        // as soon as the current pair's products of that half are issued.
        bf16x8 cur_hi[2][2];
        bf16x8 cur_lo[2][2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            load_tile(copies, 0, ks, cur_hi[0][ks], cur_lo[0][ks]);
            load_tile(copies, 1, ks, cur_hi[1][ks], cur_lo[1][ks]);
        }
#pragma unroll
        for (int i = 0; i < kTilesPerWave; i += 2) {
            int mt = wave + kFastWaves * i;
            asm volatile("" : "+s"(mt));
            if (mt >= mtiles) {
                break;  // uniform: later pairs are past the tile count too
            }
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur_hi[0][ks], b_hi[ks], acc[i], 0, 0, 0);
                acc[i + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur_hi[1][ks], b_hi[ks], acc[i + 1], 0, 0, 0);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur_hi[0][ks], b_lo[ks], acc[i], 0, 0, 0);
                acc[i + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur_hi[1][ks], b_lo[ks], acc[i + 1], 0, 0, 0);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur_lo[0][ks], b_hi[ks], acc[i], 0, 0, 0);
                acc[i + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur_lo[1][ks], b_hi[ks], acc[i + 1], 0, 0, 0);
                if (i + 2 < kTilesPerWave) {
                    load_tile(copies, i + 2, ks, cur_hi[0][ks], cur_lo[0][ks]);
                    load_tile(copies, i + 3, ks, cur_hi[1][ks], cur_lo[1][ks]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (p + 1 < total_planes) {
            store((p + 1) & 1);
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    cur_anchor[ks][v] = pre_anchor[ks][v];
                }
            }
        }
        if (frame_done) {
            // Epilogue: this candidate frame's distances into LDS per anchor
            // (window scan order), then out to global memory.  The energy
            // box takes the buffer of plane p once every wave is past it.
            float* s_ebox = reinterpret_cast<float*>(s_copy + static_cast<size_t>(p & 1) * 2 * kCopies * S.copy_elems);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kEnergyPerThread; ++k) {
                int item = tid + k * kFastThreads;
                asm volatile("" : "+v"(item));
                if (item < box_items) {
                    const int ry = item / ucols;
                    s_ebox[(item - ry * ucols) * S.ebox_pitch + ry] = pre_energy[k];
                }
            }
            __syncthreads();
            const int ti = p / planes_per_frame;
            const bool anchor_frame = frame == af;
            // The lane's anchor window in registers for the loop (an empty
            // window for an empty slot).
            const float my_energy = my_anchor >= 0 ? s_energy[n_lane] : 0.0F;
            const int my_xl = my_anchor >= 0 ? s_xl[n_lane] : 0;
            const int my_yl = my_anchor >= 0 ? s_yl[n_lane] : 0;
            const int my_nx = my_anchor >= 0 ? s_nx[n_lane] : 0;
            const int my_ny = my_anchor >= 0 ? s_ny[n_lane] : 0;
            // The anchor against itself: exactly 0 (NaN for a patch holding
            // NaN, as the exact sum gives).
            const float my_self = my_energy == my_energy ? 0.0F : my_energy;
            // Lane (n_lane, kb) holds candidates 16 rt + 4 kb + v of column
            // col against anchor n_lane; their energies are consecutive in
            // the column-major box (one broadcast 16-byte read).
#pragma unroll
            for (int i = 0; i < kTilesPerWave; ++i) {
                int mt = wave + kFastWaves * i;
                asm volatile("" : "+s"(mt));
                if (mt >= mtiles) {
                    break;  // uniform
                }
                const int col = (mt * tile_magic) >> 16;
                const int rt = mt - col * row_tiles;
                const int ry0 = 16 * rt + 4 * kb;
                const float4 ec = *reinterpret_cast<const float4*>(s_ebox + col * S.ebox_pitch + ry0);
                const float ecv[4] = {ec.x, ec.y, ec.z, ec.w};
                const int rel_x = ux0 + col - my_xl;
                const int rel_y0 = uy0 + ry0 - my_yl;
                const bool x_ok = static_cast<unsigned>(rel_x) < static_cast<unsigned>(my_nx);
                const bool self_x = anchor_frame && ux0 + col == my_ax;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rel_y = rel_y0 + v;
                    float dist = (my_energy + ecv[v]) - 2.0F * acc[i][v];
                    dist = dist < 0.0F ? 0.0F : dist;  // rounding below 0; NaN stays NaN
                    dist = self_x && uy0 + ry0 + v == my_ay ? my_self : dist;
                    if (x_ok && static_cast<unsigned>(rel_y) < static_cast<unsigned>(my_ny)) {
                        s_stage[n_lane * S.per_frame + __mul24(rel_y, my_nx) + rel_x] = dist;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            __syncthreads();
            if constexpr (kFused) {
                // Compacted candidates (window positions of the keys below the
                // anchor's 64th), one list per anchor, in the idle plane
                // buffer after the energy box.
                uint16_t* scratch = reinterpret_cast<uint16_t*>(s_ebox + S.ebox_words) + wave * 4 * S.scratch_keys;
                // First candidate frame: the lists are still empty, so every
                // key would be admitted.  Each lane's smallest key (over the
                // candidates i = lane mod 64) goes in first, which bounds the
                // K-th key by the K-th of those minima; the compaction below
                // then skips the lane's own minimum (keys are unique).
                uint64_t own_min[4] = {kNoKey, kNoKey, kNoKey, kNoKey};
                if (ti == 0) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int n = 4 * wave + q;
                        if (s_index[n] < 0) {
                            continue;
                        }
                        const int count = s_nx[n] * s_ny[n];
                        for (int i = lane; i < count; i += 64) {
                            const uint64_t key =
                                (static_cast<uint64_t>(distance_key(s_stage[n * S.per_frame + i])) << 32) |
                                static_cast<uint32_t>(i);
                            own_min[q] = key < own_min[q] ? key : own_min[q];
                        }
                    }
                    uint64_t first[4] = {own_min[0], own_min[1], own_min[2], own_min[3]};
                    wave_sort64x4(first, lane);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        best[q] = first[q];  // the lists were empty
                    }
                }
                int m[4] = {0, 0, 0, 0};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int n = 4 * wave + q;
                    if (s_index[n] < 0) {
                        continue;
                    }
                    const int count = s_nx[n] * s_ny[n];
                    const uint32_t base_index = static_cast<uint32_t>(ti * count);
                    const uint64_t threshold = admit_bound(best[q], d.similar, tau_key);
                    uint16_t* list = scratch + q * S.scratch_keys;
                    for (int c0 = 0; c0 < count; c0 += 64) {
                        const int i = c0 + lane;
                        bool pass = false;
                        bool valid = false;
                        uint64_t key = kNoKey;
                        if (i < count) {
                            const float dist = s_stage[n * S.per_frame + i];
                            valid = dist == dist;
                            key = (static_cast<uint64_t>(distance_key(dist)) << 32) | (base_index + i);
                            pass = key < threshold && key != own_min[q];
                        }
                        valid_count[q] += __popcll(__ballot(valid));
                        const uint64_t mask = __ballot(pass);
                        if (pass) {
                            list[m[q] + __popcll(mask & ((1ULL << lane) - 1))] = static_cast<uint16_t>(i);
                        }
                        m[q] += __popcll(mask);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const int most = max(max(m[0], m[1]), max(m[2], m[3]));
                if (most <= kInsertKeys) {
                    // Few keys (the usual case once the lists hold a frame's
                    // worth): each goes into its sorted list directly, a ballot
                    // finding its place and one lane shift making room; the
                    // lists end as the sort-and-merge rounds below leave them.
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int n = 4 * wave + q;
                        const uint32_t base_index = static_cast<uint32_t>(ti * s_nx[n] * s_ny[n]);
                        for (int j = 0; j < m[q]; ++j) {
                            const int i = scratch[q * S.scratch_keys + j];
                            const uint64_t key =
                                (static_cast<uint64_t>(distance_key(s_stage[n * S.per_frame + i])) << 32) |
                                (base_index + i);
                            best[q] = wave_insert(best[q], key, lane);
                        }
                    }
                }
                // Otherwise rounds of 64 compacted keys, the four anchors side
                // by side; keys no longer below an anchor's 64th are dropped.
                const int rounds = most <= kInsertKeys ? 0 : (most + 63) / 64;
                for (int r = 0; r < rounds; ++r) {
                    uint64_t key[4];
                    bool any = false;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int at = 64 * r + lane;
                        const uint64_t current = admit_bound(best[q], d.similar, tau_key);
                        uint64_t v = kNoKey;
                        if (at < m[q]) {
                            const int n = 4 * wave + q;
                            const int i = scratch[q * S.scratch_keys + at];
                            const uint32_t base_index = static_cast<uint32_t>(ti * s_nx[n] * s_ny[n]);
                            v = (static_cast<uint64_t>(distance_key(s_stage[n * S.per_frame + i])) << 32) |
                                (base_index + i);
                        }
                        v = v < current ? v : kNoKey;
                        key[q] = v;
                        any = any || __ballot(v != kNoKey) != 0;
                    }
                    if (!any) {
                        continue;
                    }
                    wave_sort64x4(key, lane);
                    wave_merge_lower64x4(best, key, lane);
                }
            } else {
                for (int n = 0; n < kTileAnchors; ++n) {
                    const int index = s_index[n];
                    if (index < 0) {
                        continue;
                    }
                    const int count = s_nx[n] * s_ny[n];
                    float* dst = out + static_cast<size_t>(index - anchor_base) * S.stride +
                                 static_cast<size_t>(ti) * count;
                    for (int i = tid; i < count; i += kFastThreads) {
                        dst[i] = s_stage[n * S.per_frame + i];
                    }
                }
            }
        }
        __syncthreads();
    }
    if constexpr (kFused) {
        // select_and_write over the retained keys: k = min(K, valid), the
        // threshold max(tau, d_k), the prefix within it (at most the buffer
        // capacity of find_similar_patches, core.cpp:1279-1287), the K
        // nearest first, the rest in scan order.
        const long long window_cap = static_cast<long long>(d.search_window) * d.search_window * nt;
        const int buffer_cap = static_cast<int>(window_cap < d.soft_cap ? window_cap : d.soft_cap);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int n = 4 * wave + q;
            const int a = s_index[n];
            if (a < 0) {
                continue;
            }
            if (valid_count[q] == 0) {
                if (lane == 0) {
                    counts[a] = 0;
                }
                continue;
            }
            const int k = d.similar < valid_count[q] ? d.similar : valid_count[q];
            const uint64_t key = best[q];
            const uint32_t dist_key = static_cast<uint32_t>(key >> 32);
            const uint32_t index = static_cast<uint32_t>(key & 0xFFFFFFFFu);
            const float kth = __uint_as_float(static_cast<uint32_t>(readlane64(key, k - 1) >> 32));
            const float threshold = d.tau < kth ? kth : d.tau;
            const uint32_t threshold_key = distance_key(threshold);
            const bool selected = lane < buffer_cap && key != kNoKey && dist_key <= threshold_key;
            const int total = __popcll(__ballot(selected));
            int slot = lane;
            if (lane >= k) {
                int before = 0;
                for (int f = k; f < total; ++f) {
                    const uint32_t other = static_cast<uint32_t>(__shfl(static_cast<int>(index), f, 64));
                    before += other < index ? 1 : 0;
                }
                slot = k + before;
            }
            if (selected) {
                const int per_frame = s_nx[n] * s_ny[n];
                const int ti = static_cast<int>(index) / per_frame;
                const int rem = static_cast<int>(index) - ti * per_frame;
                const int yy = rem / s_nx[n];
                const int cx = s_xl[n] + (rem - yy * s_nx[n]);
                const int cy = s_yl[n] + yy;
                const int frame = scheduled_frame(af, t_low, t_high, ti);
                const size_t out_base = static_cast<size_t>(a) * d.cap;
                positions[out_base + slot] = pack_position(cx, cy, frame);
                distances[out_base + slot] = __uint_as_float(dist_key);
            }
            if (lane == 0) {
                counts[a] = total;
            }
        }
    }
}

// Selection of every anchor from its staged distances.
__global__ __launch_bounds__(256) void select_fast_kernel(
    StageDesc d, int stride, const float* __restrict__ staged, const uint32_t* __restrict__ anchors,
    int anchor_count, int* __restrict__ counts, uint32_t* __restrict__ positions,
    float* __restrict__ distances) {
    extern __shared__ float s_dist[];
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
    const float* src = staged + static_cast<size_t>(a) * stride;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        s_dist[i] = src[i];
    }
    __syncthreads();
    select_and_write(d, s_dist, n, af, t_low, t_high, x_low, y_low, nx, per_frame, nt, s_hist, s_int,
                     s_scan, s_sel_dist, s_sel_index, s_sel_rank, a, out_base, counts, positions,
                     distances);
}

FastShape fast_shape(const StageDesc& d, int max_cols, int max_rows) {
    FastShape S{};
    S.row_tiles = (max_rows + 15) / 16;
    S.region_rows = 16 * S.row_tiles + kFastPatch - 1;
    // Copy rows cover samples j .. j + 7 for j < max_cols rounded to 4.
    S.pitch = (max_cols + kFastPatch - 1 + 3) / 4 * 4;
    if ((S.pitch / 4) % 2 == 0) {
        S.pitch += 4;  // odd number of 8-byte units: rows spread over banks
    }
    S.copy_elems = S.region_rows * S.pitch;
    S.per_frame = d.search_window * d.search_window;
    S.stride = S.per_frame * (d.search_bwd + d.search_fwd + 1);
    S.ebox_pitch = 16 * S.row_tiles + 4;
    S.ebox_words = max_cols * S.ebox_pitch;
    S.scratch_keys = (S.per_frame + 1) / 2 * 2;
    return S;
}

constexpr int kMaxTilesPerWave = 24;

}  // namespace

bool fast_match_supported(const StageDesc& d, int max_cols, int max_rows) {
    if (d.patch != kFastPatch || d.similar <= 1) {
        return false;
    }
    const int row_tiles = (max_rows + 15) / 16;
    const FastShape S = fast_shape(d, max_cols, max_rows);
    return static_cast<long long>(max_cols) * row_tiles <= static_cast<long long>(kMaxTilesPerWave) * kFastWaves &&
           S.region_rows * (S.pitch / 4) <= kFillPerThread * kFastThreads &&
           static_cast<long long>(max_cols) * max_rows <= static_cast<long long>(kEnergyPerThread) * kFastThreads &&
           static_cast<long long>(S.ebox_words) * 4 <= static_cast<long long>(2) * kCopies * S.copy_elems * 2 &&
           fast_match_shared_bytes(d, max_cols, max_rows) <= 150 * 1024;  // of 160 KB per CU
}

bool fast_match_fused(const StageDesc& d, int max_cols, int max_rows) {
    // The running lists hold 64 keys per anchor; the compaction scratch of
    // four wavefronts shares the idle plane buffer with the energy box.
    const FastShape S = fast_shape(d, max_cols, max_rows);
    const long long scratch = static_cast<long long>(S.ebox_words) * 4 + 2LL * kTileAnchors * S.scratch_keys;
    return d.cap <= 64 && d.similar <= 64 && scratch <= static_cast<long long>(2) * kCopies * S.copy_elems * 2;
}

size_t fast_match_shared_bytes(const StageDesc& d, int max_cols, int max_rows) {
    const FastShape S = fast_shape(d, max_cols, max_rows);
    const size_t buffers = static_cast<size_t>(4) * kCopies * S.copy_elems * sizeof(__bf16);
    const size_t staging = static_cast<size_t>(kTileAnchors) * S.per_frame * sizeof(float);
    return buffers + staging;
}

size_t fast_match_stride(const StageDesc& d) {
    return static_cast<size_t>(d.search_window) * d.search_window * (d.search_bwd + d.search_fwd + 1);
}

hipError_t launch_fast_energy(const StageDesc& d, const float* source, float* offsets, float* rows,
                              float* energy, hipStream_t stream) {
    const int match_channels = d.stage == kFinal ? d.channels : 1;
    const size_t plane = static_cast<size_t>(d.width) * d.height;
    hipLaunchKernelGGL(channel_offset_kernel, dim3(static_cast<unsigned>(match_channels)), dim3(256), 0,
                       stream, source, d.channels, plane, d.frames, offsets);
    const size_t row_total = plane * match_channels * d.frames;
    hipLaunchKernelGGL(row_energy_kernel, dim3(static_cast<unsigned>((row_total + 255) / 256)), dim3(256), 0,
                       stream, d, source, offsets, match_channels, rows);
    const size_t total = plane * d.frames;
    hipLaunchKernelGGL(patch_energy_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                       stream, d, rows, match_channels, energy);
    return hipGetLastError();
}

hipError_t launch_fast_match(const StageDesc& d, int max_cols, int max_rows, const float* source,
                             const float* offsets, const float* energy, const uint32_t* anchors,
                             const int* tiles, int tile_count, int anchor_base, float* staged,
                             bool fused, int* counts, uint32_t* positions, float* distances,
                             hipStream_t stream) {
    if (tile_count <= 0) {
        return hipSuccess;
    }
    const FastShape S = fast_shape(d, max_cols, max_rows);
    const size_t shared = fast_match_shared_bytes(d, max_cols, max_rows);
    const int mtiles = max_cols * S.row_tiles;
    const int per_wave = (mtiles + kFastWaves - 1) / kFastWaves;
#define VNLBHIP_FAST(T)                                                                            \
    do {                                                                                           \
        if (fused) {                                                                               \
            hipLaunchKernelGGL((match_fast_kernel<T, true>), dim3(static_cast<unsigned>(tile_count)), \
                               dim3(kFastThreads), shared, stream, d, S, source, offsets, energy,  \
                               anchors, tiles, anchor_base, staged, counts, positions, distances); \
        } else {                                                                                   \
            hipLaunchKernelGGL((match_fast_kernel<T, false>), dim3(static_cast<unsigned>(tile_count)), \
                               dim3(kFastThreads), shared, stream, d, S, source, offsets, energy,  \
                               anchors, tiles, anchor_base, staged, counts, positions, distances); \
        }                                                                                          \
    } while (0)
    if (per_wave <= 8) {
        VNLBHIP_FAST(8);
    } else if (per_wave <= 12) {
        VNLBHIP_FAST(12);
    } else if (per_wave <= 16) {
        VNLBHIP_FAST(16);
    } else if (per_wave <= 18) {
        VNLBHIP_FAST(18);
    } else if (per_wave <= 20) {
        VNLBHIP_FAST(20);
    } else {
        VNLBHIP_FAST(24);
    }
#undef VNLBHIP_FAST
    return hipGetLastError();
}

hipError_t launch_fast_select(const StageDesc& d, const float* staged, const uint32_t* anchors,
                              int anchor_count, int* counts, uint32_t* positions, float* distances,
                              hipStream_t stream) {
    if (anchor_count <= 0) {
        return hipSuccess;
    }
    const size_t stride = fast_match_stride(d);
    hipLaunchKernelGGL(select_fast_kernel, dim3(static_cast<unsigned>(anchor_count)), dim3(256),
                       stride * sizeof(float), stream, d, static_cast<int>(stride), staged, anchors,
                       anchor_count, counts, positions, distances);
    return hipGetLastError();
}

}  // namespace vnlbhip
