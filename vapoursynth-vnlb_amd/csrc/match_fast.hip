// Block matching, MFMA form.
//
// The squared distance of a candidate patch c to an anchor patch r is
//
//     ||r - c||^2 = E(r) + E(c) - 2 <r, c>
//
// with E the patch energy (fp32, one value per patch origin) and the cross
// term accumulated in fp32 on the matrix cores from three bf16 products
// (x = hi + lo, <r, c> ~ hi.hi + lo.hi + hi.lo).  Samples are offset by a
// per-channel constant first, which leaves distances unchanged and keeps the
// energies close to the distances they are subtracted from.  The candidates
// kept are then chosen by the same k-nearest selection as the exact form
// (select.h), so only the distance values differ from it.
//
// Once per stage, every matched plane is split into packed bf16 pairs
// (hi << 16 | lo of the offset sample), the format the matcher stages.
//
// Anchors are matched in tiles of up to 4 grid rows x 4 anchors of one frame
// (slot 4 * row + column); the tile's candidates are the union of the
// anchors' windows.  One v_mfma_f32_16x16x32_bf16 multiplies the 16 anchor
// slots (A) by 16 candidates of a union column (B, consecutive rows) over 4
// patch rows, so lane (n, kb) of the result holds candidate n against the
// four anchors of grid row kb.  The search region of one (frame, channel)
// plane is staged in LDS in groups of 4 columns: per row and group, four
// copies shifted by 0..3 columns, each 4 bf16 hi then their 4 lo.
// Wavefront w takes the union columns w mod 4, so it reads only copy w, at
// group offsets fixed at compile time.
//
// After each candidate frame the wavefronts test their products against a
// per-anchor bound (the largest key the running k-nearest lists can still
// take) and write out only the candidates within it; the wavefront owning an
// anchor merges them into its list.  The anchor frame, which comes first,
// and clips whose energies are not all finite and small enough that no sum
// overflows write every window distance and select from all of them.
// Supported for 8x8 patches.
#include "common.h"
#include "device_util.h"
#include "kernels.h"
#include "select.h"

#include <float.h>

#include <algorithm>
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
constexpr int kItemsPerThread = 3;    // staged (row, group) items per thread
constexpr int kEnergyPerThread = 6;   // energy box items per thread
// Energies above this bound (or not finite) switch a stage to the exhaustive
// epilogue: below it no energy sum or doubled cross term overflows.
constexpr float kEnergyBound = 8.0e37F;

// Per-channel offset: the mean of the finite values of a strided sample of
// the channel (any finite offset leaves the distances unchanged; non-finite
// samples only poison the patches that hold them).  The sample is split
// over kOffsetBlocks workgroups per channel, each holding sixteen loads per
// thread in flight; their (sum, count) partials are added in block order by
// channel_offset_kernel, which also clears the stage's energy flag.
constexpr int kOffsetBlocks = 16;
constexpr int kOffsetBatch = 16;
__global__ __launch_bounds__(256) void channel_partial_kernel(const float* __restrict__ source, int channels,
                                                              size_t plane, int frames,
                                                              double* __restrict__ partials) {
    __shared__ double s_part[4];
    __shared__ double s_count[4];
    const int c = blockIdx.y;
    const size_t per_channel = plane * frames;
    const size_t stride = per_channel / 65536 + 1;
    const size_t step = static_cast<size_t>(kOffsetBlocks) * blockDim.x * stride;
    double sum = 0.0;
    size_t count = 0;
    for (size_t i0 = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * stride; i0 < per_channel;
         i0 += kOffsetBatch * step) {
        float v[kOffsetBatch];
#pragma unroll
        for (int b = 0; b < kOffsetBatch; ++b) {
            const size_t i = i0 + b * step;
            const size_t f = i / plane;
            const size_t rem = i - f * plane;
            v[b] = i < per_channel ? source[(f * channels + c) * plane + rem] : 0.0F;
        }
#pragma unroll
        for (int b = 0; b < kOffsetBatch; ++b) {
            if (i0 + b * step < per_channel && v[b] == v[b] && fabsf(v[b]) <= FLT_MAX) {
                sum += static_cast<double>(v[b]);
                ++count;
            }
        }
    }
    sum = wave_sum_double(sum);
    double n = wave_sum_double(static_cast<double>(count));
    if ((threadIdx.x & 63) == 0) {
        s_part[threadIdx.x >> 6] = sum;
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
        double* out = partials + 2 * (static_cast<size_t>(c) * kOffsetBlocks + blockIdx.x);
        out[0] = total;
        out[1] = total_n;
    }
}

__global__ __launch_bounds__(64) void channel_offset_kernel(const double* __restrict__ partials, int channels,
                                                            float* __restrict__ offsets,
                                                            unsigned* __restrict__ flags) {
    const int c = threadIdx.x;
    if (c == 0) {
        flags[0] = 0u;
    }
    if (c < channels) {
        double total = 0.0;
        double total_n = 0.0;
        for (int b = 0; b < kOffsetBlocks; ++b) {
            total += partials[2 * (c * kOffsetBlocks + b)];
            total_n += partials[2 * (c * kOffsetBlocks + b) + 1];
        }
        offsets[c] = total_n > 0.0 ? static_cast<float>(total / total_n) : 0.0F;
    }
}

// Grids of the kernels below: x over the items of a plane (pixels or strips
// of patch origins; 32-bit, a plane holds at most 4096 x 4096), y over
// frame (x channel).

// Horizontal sums of squares: rows[f][c][y][x] = sum_{px<P} (v - offset)^2,
// one origin per thread (consecutive lanes on consecutive columns: each
// load instruction reads a few cache lines).
__global__ __launch_bounds__(256) void row_energy_kernel(StageDesc d, const float* __restrict__ source,
                                                         const float* __restrict__ offsets,
                                                         int match_channels, float* __restrict__ rows) {
    const int plane = d.width * d.height;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= plane) {
        return;
    }
    const int fc = blockIdx.y;  // f * match_channels + c
    const int f = fc / match_channels;
    const int c = fc - f * match_channels;
    const int y = p / d.width;
    const int x = p - y * d.width;
    float sum = 0.0F;
    if (x <= d.width - d.patch) {
        const float mu = offsets[c];
        const float* src = source + (static_cast<size_t>(f) * d.channels + c) * plane + p;
        for (int px = 0; px < d.patch; ++px) {
            const float v = src[px] - mu;
            sum = fmaf(v, v, sum);
        }
    }
    rows[static_cast<size_t>(fc) * plane + p] = sum;
}

// Patch energies: energy[f][y][x] = sum over patch frames t, channels c and
// rows py (in that order) of the row sums, for every patch origin (0
// elsewhere).  flags[0] becomes 1 where an energy is not finite or above
// kEnergyBound.  One thread per column and strip of kEnergyStrip origins: it
// reads the strip's rows once per (t, c) and keeps one running sum per
// origin, so each origin's sum has the terms and order of a sum over its own
// 8 rows with an eighth of the loads.  8x8 patches (the matcher's).
constexpr int kEnergyStrip = 8;
__global__ __launch_bounds__(256) void patch_energy_kernel(StageDesc d, const float* __restrict__ rows,
                                                           int match_channels, float* __restrict__ energy,
                                                           unsigned* __restrict__ flags) {
    const int strips = (d.height + kEnergyStrip - 1) / kEnergyStrip;
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= d.width * strips) {
        return;
    }
    const int f = blockIdx.y;
    const int strip = item / d.width;
    const int x = item - strip * d.width;
    const int y0 = strip * kEnergyStrip;
    const int plane = d.width * d.height;
    const bool column_ok = x <= d.width - kFastPatch && f <= d.frames - d.patch_time;
    float sum[kEnergyStrip];
#pragma unroll
    for (int j = 0; j < kEnergyStrip; ++j) {
        sum[j] = 0.0F;
    }
    if (column_ok) {
        for (int t = 0; t < d.patch_time; ++t) {
            for (int c = 0; c < match_channels; ++c) {
                const float* r = rows + (static_cast<size_t>(f + t) * match_channels + c) * plane + x;
                float v[kEnergyStrip + kFastPatch - 1];
#pragma unroll
                for (int k = 0; k < kEnergyStrip + kFastPatch - 1; ++k) {
                    const int y = y0 + k;
                    v[k] = y < d.height ? r[static_cast<size_t>(y) * d.width] : 0.0F;
                }
#pragma unroll
                for (int j = 0; j < kEnergyStrip; ++j) {
#pragma unroll
                    for (int py = 0; py < kFastPatch; ++py) {
                        sum[j] += v[j + py];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kEnergyStrip; ++j) {
        const int y = y0 + j;
        if (y < d.height) {
            const bool origin = column_ok && y <= d.height - kFastPatch;
            const float value = origin ? sum[j] : 0.0F;
            if (origin && !(value <= kEnergyBound)) {
                atomicOr(flags, 1u);
            }
            energy[(static_cast<size_t>(f) * d.height + y) * d.width + x] = value;
        }
    }
}

// split[f][c][y][x] = bf16(v) << 16 | bf16(v - bf16(v)) of v = sample - offset.
__global__ __launch_bounds__(256) void split_planes_kernel(StageDesc d, const float* __restrict__ source,
                                                           const float* __restrict__ offsets,
                                                           int match_channels, uint32_t* __restrict__ split) {
    const int plane = d.width * d.height;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= plane) {
        return;
    }
    const int fc = blockIdx.y;  // f * match_channels + c
    const int f = fc / match_channels;
    const int c = fc - f * match_channels;
    const size_t i = static_cast<size_t>(fc) * plane + p;
    const float v = source[(static_cast<size_t>(f) * d.channels + c) * plane + p] - offsets[c];
    const __bf16 hi = static_cast<__bf16>(v);
    const __bf16 lo = static_cast<__bf16>(v - static_cast<float>(hi));
    split[i] = (static_cast<uint32_t>(__builtin_bit_cast(unsigned short, hi)) << 16) |
               static_cast<uint32_t>(__builtin_bit_cast(unsigned short, lo));
}

// The bf16 hi (lo) halves of two packed samples as one dword, a first:
// v_perm_b32 picks bytes 2, 3 (0, 1) of a, then of b.
__device__ inline uint32_t hi_pair(uint32_t a, uint32_t b) {
    return __builtin_amdgcn_perm(b, a, 0x07060302u);
}
__device__ inline uint32_t lo_pair(uint32_t a, uint32_t b) {
    return __builtin_amdgcn_perm(b, a, 0x05040100u);
}

// One compare-exchange step of a bitonic network over the wavefront:
// partner lane ^ kJ, blocks of kK lanes alternately ascending (kK = 64: the
// whole wavefront ascending).  Four independent lists at once (one per
// anchor of a wavefront): the exchanges of a step are issued together so
// their latencies overlap.
template <int kK, int kJ>
__device__ inline void bitonic_step4(uint64_t (&key)[4], int lane) {
    uint64_t other[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        other[q] = xor64<kJ>(key[q]);
    }
    // Keep the smaller key where the lane is the lower end of an ascending
    // pair or the upper end of a descending one (keys are distinct).
    const bool want_low = ((lane & kJ) == 0) == ((lane & kK) == 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        key[q] = (other[q] < key[q]) == want_low ? other[q] : key[q];
    }
}

// Wave-wide bitonic sort of 64 keys per list, one per lane, ascending by lane.
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

// best[q] <- the 64 smallest of best[q] and sorted[q], ascending: the first
// half-cleaner of the bitonic merge of best and reversed sorted, then the
// merge.
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

// The value of lane l - 1 in lane l (DPP wave_shr:1; bound_ctrl gives lane
// 0, which has no source, a 0).
__device__ inline uint64_t from_lane_below(uint64_t v) {
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(v & 0xFFFFFFFFu), 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(v >> 32), 0x138, 0xF, 0xF, true);
    return (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo);
}

// `key` into a list of distinct keys ascending over the lanes (the 64
// smallest so far; kNoKey for empty places): the keys below it stay, the
// rest move up one lane and the largest drops out.
__device__ inline uint64_t wave_insert(uint64_t list, uint64_t key, int lane) {
    const int place = __popcll(__ballot(list < key));
    const uint64_t before = from_lane_below(list);
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

// LDS rows of one staged plane: the bf16 hi of groups of 4 columns, four
// copies of each group (shifted by 0..3 columns, 8 bytes each), then the lo
// in the same order, then 8 bytes: rows 2 x odd dwords apart, so the 8-byte
// operand reads of 16 consecutive rows fall on distinct banks.
__host__ __device__ constexpr int staged_row_bytes(int groups) {
    return 2 * 8 * kCopies * groups + 8;
}
// The anchors' packed samples of one plane after its rows, 16-byte pieces
// [patch row][half][slot] (16 slots read one piece each, on distinct banks).
constexpr int kAnchorBytes = kTileAnchors * kFastPatch * kFastPatch * 4;

// LDS layout of the matcher for one launch: every size follows from the
// instantiation's row tiles (kR) and union columns per wavefront (kA), the
// window and the widest union.  Two plane buffers; after a candidate frame's
// products the epilogue takes the space of both.
struct TileLayout {
    int groups;        // 4-column groups per staged row: kA + 1
    int row_bytes;     // staged_row_bytes(groups)
    int rows_bytes;    // (16 kR + P - 1) rows
    int buffer_bytes;  // rows, then the anchors' samples
    int stride;        // distance slots per anchor (window^2 * temporal), staged mode
    int per_frame;     // window^2
    int ebox_pitch;    // floats per union column of the energy box (16 kR + 1: odd, columns on distinct banks)
    int ebox_bytes;    // energy box, column-major, 16-byte aligned
    int stage_bytes;   // window distances of the 16 anchor slots
    int scratch_keys;  // candidate list entries per anchor (window^2 rounded to 2)
    int sub_keys;      // entries of one wavefront's part of a list (its union columns of a window)
    int shared_bytes;  // max(two buffers, energy box + distances + lists)
};

// Distances of one anchor tile against its candidates, all candidate frames.
// fused = 0: distances land in `out` ([anchor][stride], scan order) for
// select_fast_kernel.  fused = 1 (retained cap and K <= 64): the k-nearest
// selection runs here.  Each wavefront keeps, for anchors 4 w .. 4 w + 3 of
// the tile, the 64 smallest keys (distance key << 32 | scan position) seen
// so far in lane order.  The selected set of select.h is a prefix of the
// (distance, scan position) order no longer than the retained cap, so it
// lies within those 64 keys, and the lists written are those of
// select_and_write.
template <int kR, int kA>
__global__ __launch_bounds__(kFastThreads) __attribute__((amdgpu_waves_per_eu(kR * kA <= 18 ? 2 : 1)))
void match_fast_kernel(
    StageDesc d, TileLayout L, const uint32_t* __restrict__ split, const unsigned* __restrict__ flags,
    const float* __restrict__ energy, const uint32_t* __restrict__ anchors,
    const int* __restrict__ tiles, int anchor_base, int fused, float* __restrict__ out,
    int* __restrict__ counts, uint32_t* __restrict__ positions, float* __restrict__ distances) {
    constexpr int kTiles = kR * kA;  // tile i: union column 4 (i / kR) + wave, row tile i % kR
    static_assert(kTiles % 2 == 0, "tiles go through the matrix cores in pairs");
    constexpr int kGroups = kA + 1;
    constexpr int kRowBytes = staged_row_bytes(kGroups);
    constexpr int kRegionRows = 16 * kR + kFastPatch - 1;
    constexpr int kRowsBytes = kRegionRows * kRowBytes;
    constexpr int kItems = (kRegionRows * kGroups + kFastThreads - 1) / kFastThreads;
    static_assert(kItems <= kItemsPerThread, "staging items per thread");
    static_assert(kAnchorBytes == 16 * kFastThreads, "one 16-byte anchor piece per thread");
    extern __shared__ uint4 s_dyn16[];
    unsigned char* s_bytes = reinterpret_cast<unsigned char*>(s_dyn16);  // [buffer][rows | anchors]
    __shared__ int s_index[kTileAnchors];
    __shared__ int s_ax[kTileAnchors];
    __shared__ int s_ay[kTileAnchors];
    __shared__ int s_xl[kTileAnchors];
    __shared__ int s_yl[kTileAnchors];
    __shared__ int s_nx[kTileAnchors];
    __shared__ int s_ny[kTileAnchors];
    __shared__ float s_energy[kTileAnchors];
    // Largest distance that can still enter each anchor's list (NaN: every
    // window distance is needed) and each wavefront's count of the frame's
    // candidates per anchor.
    __shared__ __align__(16) float s_bound[kTileAnchors];
    __shared__ int s_parts[kTileAnchors * kFastWaves];
    // Between candidate frames (fused): the running lists of the anchors of
    // each wavefront and their counts of valid (non-NaN) distances.
    __shared__ uint64_t s_best[kFastWaves][4][64];
    __shared__ int s_valid[kTileAnchors];
    __shared__ int s_box[5];
    __shared__ int s_vx[8];  // union x range of the anchors of each tile column
    __shared__ int s_full;

    const int tid = threadIdx.x;
    // Uniform per wavefront: tile indices derived from it stay scalar.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    // Each XCD takes a contiguous run of tiles (neighbouring anchors, whose
    // window unions overlap in its L2).
    const int* tile = tiles + static_cast<size_t>(xcd_block(blockIdx.x, gridDim.x)) * kTileAnchors;
    const bool final_stage = d.stage == kFinal;
    const int match_channels = final_stage ? d.channels : 1;
    const bool finite = flags[0] == 0u;

    if (tid < kTileAnchors) {
        const int index = tile[tid];
        s_index[tid] = index;
        int ax = 0;
        int ay = 0;
        int xl = 0;
        int xh = -1;
        int yl = 0;
        int yh = -1;
        float e = 0.0F;
        if (index >= 0) {
            const uint32_t p = anchors[index];
            ax = position_x(p);
            ay = position_y(p);
            spatial_range(d, ax, d.width, xl, xh);
            spatial_range(d, ay, d.height, yl, yh);
            e = energy[(static_cast<size_t>(position_frame(p)) * d.height + ay) * d.width + ax];
        }
        s_ax[tid] = ax;
        s_ay[tid] = ay;
        s_xl[tid] = xl;
        s_yl[tid] = yl;
        s_nx[tid] = xh - xl + 1;  // 0 for an empty slot
        s_ny[tid] = yh - yl + 1;
        s_energy[tid] = e;
        s_bound[tid] = __builtin_nanf("");
    }
    __syncthreads();
    if (tid == 0) {
        int x0 = 1 << 30;
        int x1 = -1;
        int y0 = 1 << 30;
        int y1 = -1;
        int first = -1;
        for (int v = 0; v < 4; ++v) {
            s_vx[v] = 1 << 30;
            s_vx[4 + v] = -1;
        }
        for (int n = 0; n < kTileAnchors; ++n) {
            if (s_index[n] >= 0) {
                first = first < 0 ? s_index[n] : first;
                x0 = min(x0, s_xl[n]);
                x1 = max(x1, s_xl[n] + s_nx[n] - 1);
                y0 = min(y0, s_yl[n]);
                y1 = max(y1, s_yl[n] + s_ny[n] - 1);
                s_vx[n & 3] = min(s_vx[n & 3], s_xl[n]);
                s_vx[4 + (n & 3)] = max(s_vx[4 + (n & 3)], s_xl[n] + s_nx[n] - 1);
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
    const int ucols = s_box[1] - ux0 + 1;  // <= 4 kA (host checked)
    const int urows = s_box[3] - uy0 + 1;
    const int row_tiles = (urows + 15) / 16;  // <= kR (host checked)
    int t_low = 0;
    int t_high = 0;
    temporal_range(d, af, t_low, t_high);
    const int nt = t_high - t_low + 1;
    const size_t plane = static_cast<size_t>(d.width) * d.height;

    const int n_lane = lane & 15;  // anchor slot (A row) / candidate of the tile (B column)
    const int kb = lane >> 4;      // 8-sample block of the 32-wide k step / grid row of the result
    // Union columns of this wavefront (wave, wave + 4, ...) and row tiles.
    const int wave_cols = __builtin_amdgcn_readfirstlane(ucols > wave ? (ucols - wave + 3) / 4 : 0);
    const int wave_rows = __builtin_amdgcn_readfirstlane(row_tiles);
    const auto tile_ok = [&](int i) {
        return i / kR < wave_cols && i % kR < wave_rows;
    };

    // Planes (candidate frame ti, channel c, patch frame t) in order.
    const int planes_per_frame = match_channels * d.patch_time;
    const int total_planes = nt * planes_per_frame;

    // Staged items: (row r, group g) holds columns 4 g .. 4 g + 6 of the
    // region for the four copies.  Rows past the union's patches and groups
    // past its columns are not staged (only tiles past the union read them).
    // Each thread also stages one 16-byte piece of the anchors' samples:
    // slot tid % 16, patch row tid / 32, columns 4 ((tid / 16) % 2) .. + 3.
    const int box_items = urows * ucols;
    const int groups_used = (ucols - 1) / 4 + 2;
    const int rows_used = urows + kFastPatch - 1;
    int item_src[kItems];  // sample offset in a plane of column 4 g, or -1
    int item_lds[kItems];  // byte offset in a buffer
    int item_x[kItems];    // frame column of 4 g
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int item = tid + k * kFastThreads;
        const int r = item / kGroups;
        const int g = item - r * kGroups;
        const bool ok = r < rows_used && g < groups_used;
        item_x[k] = ux0 + 4 * g;
        item_src[k] = ok ? (uy0 + r) * d.width + item_x[k] : -1;
        item_lds[k] = r * kRowBytes + 8 * kCopies * g;
    }
    const int piece_slot = tid & 15;
    const int piece_src = s_index[piece_slot] >= 0
                              ? (s_ay[piece_slot] + (tid >> 5)) * d.width + s_ax[piece_slot] + 4 * ((tid >> 4) & 1)
                              : -1;
    uint32_t pre[kItems][7];
    uint4 pre_piece;
    const auto prefetch = [&](int p) {
        const int ti = p / planes_per_frame;
        const int rem = p - ti * planes_per_frame;
        const int c = rem / d.patch_time;
        const int t = rem - c * d.patch_time;
        const int frame = scheduled_frame(af, t_low, t_high, ti);
        const uint32_t* src = split + (static_cast<size_t>(frame + t) * match_channels + c) * plane;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (item_src[k] >= 0) {
                if (item_x[k] + 7 <= d.width) {
                    uint4 a;
                    uint2 b;
                    __builtin_memcpy(&a, src + item_src[k], sizeof(a));
                    __builtin_memcpy(&b, src + item_src[k] + 4, sizeof(b));
                    pre[k][0] = a.x;
                    pre[k][1] = a.y;
                    pre[k][2] = a.z;
                    pre[k][3] = a.w;
                    pre[k][4] = b.x;
                    pre[k][5] = b.y;
                    pre[k][6] = src[item_src[k] + 6];
                } else {
                    // Right border: columns past the frame stage as zeros.
#pragma unroll
                    for (int e = 0; e < 7; ++e) {
                        pre[k][e] = item_x[k] + e < d.width ? src[item_src[k] + e] : 0u;
                    }
                }
            }
        }
        // The anchor patch lies inside the frame.
        const uint32_t* anchor_src = split + (static_cast<size_t>(af + t) * match_channels + c) * plane;
        pre_piece = uint4{0u, 0u, 0u, 0u};
        if (piece_src >= 0) {
            __builtin_memcpy(&pre_piece, anchor_src + piece_src, sizeof(pre_piece));
        }
    };
    const auto store = [&](int p) {
        unsigned char* base = s_bytes + static_cast<size_t>(p & 1) * L.buffer_bytes;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (item_src[k] >= 0) {
                // Copies 0..3 of group g (copy q: columns 4 g + q .. 4 g + q + 3),
                // hi, then the same for lo in the row's second half.
                const uint32_t* s = pre[k];
                uint2* at = reinterpret_cast<uint2*>(base + item_lds[k]);
#pragma unroll
                for (int q = 0; q < kCopies; ++q) {
                    at[q] = uint2{hi_pair(s[q], s[q + 1]), hi_pair(s[q + 2], s[q + 3])};
                    at[kCopies * kGroups + q] = uint2{lo_pair(s[q], s[q + 1]), lo_pair(s[q + 2], s[q + 3])};
                }
            }
        }
        reinterpret_cast<uint4*>(base + kRowsBytes)[tid] = pre_piece;
    };

    // Anchor operands of plane p: slot n_lane, patch rows 4 ks + kb, split
    // into hi / lo.
    bf16x8 a_hi[2];
    bf16x8 a_lo[2];
    const auto load_anchor = [&](int p) {
        const uint4* piece = reinterpret_cast<const uint4*>(s_bytes + static_cast<size_t>(p & 1) * L.buffer_bytes +
                                                            kRowsBytes) + n_lane;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const uint4 a = piece[32 * (4 * ks + kb)];
            const uint4 b = piece[32 * (4 * ks + kb) + 16];
            const u32x4 hi = {hi_pair(a.x, a.y), hi_pair(a.z, a.w), hi_pair(b.x, b.y), hi_pair(b.z, b.w)};
            const u32x4 lo = {lo_pair(a.x, a.y), lo_pair(a.z, a.w), lo_pair(b.x, b.y), lo_pair(b.z, b.w)};
            a_hi[ks] = __builtin_bit_cast(bf16x8, hi);
            a_lo[ks] = __builtin_bit_cast(bf16x8, lo);
        }
    };

    // Candidate operands of tile i, patch rows 4 ks .. 4 ks + 3: row
    // 16 rt + n_lane + 4 ks + kb, copy `wave` of groups a and a + 1
    // (a = i / kR), hi and lo: the group offsets are immediates of the LDS
    // reads.  Reads of tiles past the union stay inside the rows (groups
    // kA + 1, rows 16 kR + 7).
    const unsigned lane_bytes = static_cast<unsigned>((n_lane + kb) * kRowBytes + 8 * wave);
    const auto load_tile = [&](const unsigned (&base)[kR][2], int i, int ks, bf16x8& b_hi, bf16x8& b_lo) {
        const uint2* at = reinterpret_cast<const uint2*>(s_bytes + base[i % kR][ks] + 8 * kCopies * (i / kR));
        const uint2 h0 = at[0];
        const uint2 h1 = at[kCopies];
        const uint2 l0 = at[kCopies * kGroups];
        const uint2 l1 = at[kCopies * kGroups + kCopies];
        const u32x4 hi = {h0.x, h0.y, h1.x, h1.y};
        const u32x4 lo = {l0.x, l0.y, l1.x, l1.y};
        b_hi = __builtin_bit_cast(bf16x8, hi);
        b_lo = __builtin_bit_cast(bf16x8, lo);
    };

    // Epilogue space (both plane buffers once a frame's products are done):
    // the energy box, the window distances per anchor slot, the candidate
    // lists per anchor slot.
    float* s_ebox = reinterpret_cast<float*>(s_bytes);
    float* s_stage = reinterpret_cast<float*>(s_bytes + L.ebox_bytes);
    uint16_t* s_list = reinterpret_cast<uint16_t*>(s_bytes + L.ebox_bytes + L.stage_bytes);

    // Keys at or below tau (any scan position) may be selected whatever d_K is.
    const uint64_t tau_key = (static_cast<uint64_t>(distance_key(d.tau)) << 32) | 0xFFFFFFFFull;
    if (tid < kTileAnchors) {
        s_valid[tid] = 0;
    }
    prefetch(0);
    store(0);
    __syncthreads();
    for (int ti = 0; ti < nt; ++ti) {
        const int frame = scheduled_frame(af, t_low, t_high, ti);
        floatx4 acc[kTiles];
#pragma unroll
        for (int i = 0; i < kTiles; ++i) {
            acc[i] = floatx4{0.0F, 0.0F, 0.0F, 0.0F};
        }
        const int p_last = (ti + 1) * planes_per_frame - 1;
        // Products of plane p on the matrix cores (accumulated into acc).
        const auto products = [&](int p) {
            load_anchor(p);
            unsigned base[kR][2];
            const unsigned buffer_base = static_cast<unsigned>((p & 1) * L.buffer_bytes) + lane_bytes;
#pragma unroll
            for (int rt = 0; rt < kR; ++rt) {
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    base[rt][ks] = buffer_base + static_cast<unsigned>((16 * rt + 4 * ks) * kRowBytes);
                }
            }
            // Pairs of tiles, their products alternating so consecutive MFMAs
            // are independent; the operands of the pair's second half (ks = 1)
            // and of the next pair's first half are read while the products run.
            bf16x8 cur_hi[2][2];
            bf16x8 cur_lo[2][2];
            load_tile(base, 0, 0, cur_hi[0][0], cur_lo[0][0]);
            load_tile(base, 1, 0, cur_hi[1][0], cur_lo[1][0]);
#pragma unroll
            for (int i = 0; i < kTiles; i += 2) {
                const int j = i + 1;
                const bool ok = tile_ok(i) || tile_ok(j);
                load_tile(base, i, 1, cur_hi[0][1], cur_lo[0][1]);
                load_tile(base, j, 1, cur_hi[1][1], cur_lo[1][1]);
                if (ok) {
                    // (A pair with one tile past the union computes it anyway:
                    // its products are never read.)
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi[0], cur_hi[0][0], acc[i], 0, 0, 0);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi[0], cur_hi[1][0], acc[j], 0, 0, 0);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo[0], cur_hi[0][0], acc[i], 0, 0, 0);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo[0], cur_hi[1][0], acc[j], 0, 0, 0);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi[0], cur_lo[0][0], acc[i], 0, 0, 0);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi[0], cur_lo[1][0], acc[j], 0, 0, 0);
                }
                if (i + 2 < kTiles) {
                    load_tile(base, i + 2, 0, cur_hi[0][0], cur_lo[0][0]);
                    load_tile(base, i + 3, 0, cur_hi[1][0], cur_lo[1][0]);
                }
                if (ok) {
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi[1], cur_hi[0][1], acc[i], 0, 0, 0);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi[1], cur_hi[1][1], acc[j], 0, 0, 0);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo[1], cur_hi[0][1], acc[i], 0, 0, 0);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo[1], cur_hi[1][1], acc[j], 0, 0, 0);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi[1], cur_lo[0][1], acc[i], 0, 0, 0);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi[1], cur_lo[1][1], acc[j], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        // The frame's energy box goes into registers while its last plane
        // is on the matrix cores (its LDS space is free only after them);
        // tile shapes above 18 products per wave have no registers to spare
        // for it and load it after the products.
        constexpr bool kEarlyBox = kTiles <= 18;
        float ebox_pre[kEarlyBox ? kEnergyPerThread : 1];
        const auto load_box = [&]() {
#pragma unroll
            for (int k = 0; k < kEnergyPerThread; ++k) {
                const int item = tid + k * kFastThreads;
                ebox_pre[k % (kEarlyBox ? kEnergyPerThread : 1)] = 0.0F;
                if (item < box_items) {
                    const int ry = item / ucols;
                    const int cx = item - ry * ucols;
                    ebox_pre[k % (kEarlyBox ? kEnergyPerThread : 1)] =
                        energy[(static_cast<size_t>(frame) * d.height + uy0 + ry) * d.width + ux0 + cx];
                }
            }
        };
        if constexpr (kEarlyBox) {
            for (int p = ti * planes_per_frame; p < p_last; ++p) {
                // The next plane of the frame goes into registers while this
                // one is on the matrix cores, then into the other buffer.
                prefetch(p + 1);
                products(p);
                store(p + 1);
                __syncthreads();
            }
            load_box();
            products(p_last);
            __syncthreads();
        } else {
            for (int p = ti * planes_per_frame; p <= p_last; ++p) {
                if (p < p_last) {
                    prefetch(p + 1);
                }
                products(p);
                if (p < p_last) {
                    store(p + 1);
                }
                __syncthreads();
            }
        }

        // Epilogue of the candidate frame in the space of both buffers.
#pragma unroll
        for (int k = 0; k < kEnergyPerThread; ++k) {
            int item = tid + k * kFastThreads;
            if constexpr (kEarlyBox) {
                if (item < box_items) {
                    const int ry = item / ucols;
                    const int cx = item - ry * ucols;
                    s_ebox[cx * L.ebox_pitch + ry] = ebox_pre[k % (kEarlyBox ? kEnergyPerThread : 1)];
                }
            } else {
                // Opaque: the box offsets are recomputed per frame instead of
                // being held in registers across the plane loop.
                asm volatile("" : "+v"(item));
                if (item < box_items) {
                    const int ry = item / ucols;
                    const int cx = item - ry * ucols;
                    s_ebox[cx * L.ebox_pitch + ry] =
                        energy[(static_cast<size_t>(frame) * d.height + uy0 + ry) * d.width + ux0 + cx];
                }
            }
        }
        if (wave == 0) {
            const bool open = lane < kTileAnchors && s_index[lane] >= 0 && s_bound[lane] != s_bound[lane];
            const unsigned long long any = __ballot(open);
            if (lane == 0) {
                s_full = any != 0ull ? 1 : 0;
            }
        }
        __syncthreads();
        const bool anchor_frame = ti == 0;  // the schedule starts with the anchor frame
        const bool full = fused == 0 || anchor_frame || !finite || s_full != 0;
        {
            // Result lanes: candidate n_lane against slots 4 kb + v (grid row
            // kb, one y range; an empty slot has an empty x range).
            float ea[4];
            int exl[4];
            int enx[4];
            int eyl = 0;
            int eny = 0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int slot = 4 * kb + v;
                ea[v] = s_energy[slot];
                exl[v] = s_xl[slot];
                enx[v] = s_nx[slot];
                if (s_index[slot] >= 0) {
                    eyl = s_yl[slot];
                    eny = s_ny[slot];
                }
            }
            if (full) {
                // Every window distance of the frame into s_stage (scan
                // order per anchor slot; the owners set the anchors' own
                // distances).
#pragma unroll
                for (int i = 0; i < kTiles; ++i) {
                    if (!tile_ok(i)) {
                        continue;  // uniform
                    }
                    const int col = wave + 4 * (i / kR);
                    const int rel_x0 = ux0 + col;
                    const int ry = 16 * (i % kR) + n_lane;
                    const float ec = s_ebox[col * L.ebox_pitch + ry];
                    const int rel_y = uy0 + ry - eyl;
                    const bool y_ok = static_cast<unsigned>(rel_y) < static_cast<unsigned>(eny);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int rel_x = rel_x0 - exl[v];
                        float dist = (ea[v] + ec) - 2.0F * acc[i][v];
                        dist = dist < 0.0F ? 0.0F : dist;  // rounding below 0; NaN stays NaN
                        if (y_ok && static_cast<unsigned>(rel_x) < static_cast<unsigned>(enx[v])) {
                            s_stage[(4 * kb + v) * L.per_frame + __mul24(rel_y, enx[v]) + rel_x] = dist;
                        }
                    }
                }
            } else {
                // Only candidates within the anchor's bound: their distance
                // into s_stage and their window position onto the list of
                // the anchor.  Every wavefront fills its own part of each
                // list, counted in registers (the lanes of a grid row keep
                // the same counts).  (Finite data: rows outside the y range
                // get an infinite energy, which no finite bound admits.)
                const float4 bound4 = *reinterpret_cast<const float4*>(s_bound + 4 * kb);
                const float bound[4] = {bound4.x, bound4.y, bound4.z, bound4.w};
                int vx0[4];
                int vx1[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    vx0[v] = __builtin_amdgcn_readfirstlane(s_vx[v]);
                    vx1[v] = __builtin_amdgcn_readfirstlane(s_vx[4 + v]);
                }
                const uint64_t row_lanes = 0xFFFFull << (16 * kb);
                const uint64_t row_below = row_lanes & ((1ull << lane) - 1ull);
                int appended[4] = {0, 0, 0, 0};
#pragma unroll
                for (int i = 0; i < kTiles; ++i) {
                    if (!tile_ok(i)) {
                        continue;  // uniform
                    }
                    const int col = wave + 4 * (i / kR);
                    const int x = ux0 + col;
                    const int ry = 16 * (i % kR) + n_lane;
                    const int rel_y = uy0 + ry - eyl;
                    float ec = s_ebox[col * L.ebox_pitch + ry];
                    ec = static_cast<unsigned>(rel_y) < static_cast<unsigned>(eny) ? ec : __builtin_inff();
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        if (x < vx0[v] || x > vx1[v]) {
                            continue;  // uniform: outside every window of the column
                        }
                        const int rel_x = x - exl[v];
                        // (ea + ec) - 2 acc with one rounding of the sum and
                        // one of the difference, as in the exhaustive path.
                        const float dist = fmaf(-2.0F, acc[i][v], ea[v] + ec);
                        const bool pass =
                            static_cast<unsigned>(rel_x) < static_cast<unsigned>(enx[v]) && dist <= bound[v];
                        const uint64_t mask = __ballot(pass);
                        if (mask == 0ull) {
                            continue;  // uniform
                        }
                        if (pass) {
                            const int slot = 4 * kb + v;
                            const int at = __mul24(rel_y, enx[v]) + rel_x;
                            s_stage[slot * L.per_frame + at] = dist < 0.0F ? 0.0F : dist;
                            s_list[(slot * kFastWaves + wave) * L.sub_keys + appended[v] + __popcll(mask & row_below)] =
                                static_cast<uint16_t>(at);
                        }
                        appended[v] += __popcll(mask & row_lanes);
                    }
                }
                if (n_lane == 0) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        s_parts[(4 * kb + v) * kFastWaves + wave] = appended[v];
                    }
                }
            }
        }
        __syncthreads();
        if (fused != 0) {
            // Keys of the frame per anchor (4 wave + q): m[q] window
            // positions, in one list of the owner (full frames) or in the
            // four wavefronts' parts of it (filtered frames; part_end[q][w]:
            // end of part w in the list order).
            uint16_t* scratch = s_list + wave * 4 * L.scratch_keys;
            uint64_t best[4];
            int valid_count[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                best[q] = anchor_frame ? kNoKey : s_best[wave][q][lane];
                valid_count[q] = s_valid[4 * wave + q];
            }
            int m[4] = {0, 0, 0, 0};
            int part_end[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            if (full) {
                if (anchor_frame && lane < 4) {
                    // The anchor against itself: exactly 0 (NaN for a patch
                    // holding NaN, as the exact sum gives).
                    const int n = 4 * wave + lane;
                    if (s_index[n] >= 0) {
                        const float e = s_energy[n];
                        s_stage[n * L.per_frame + (s_ay[n] - s_yl[n]) * s_nx[n] + s_ax[n] - s_xl[n]] =
                            e == e ? 0.0F : e;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
                                (static_cast<uint64_t>(distance_key(s_stage[n * L.per_frame + i])) << 32) |
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
                // Compaction: window positions of the keys below each anchor's
                // bound, one list per anchor.
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int n = 4 * wave + q;
                    if (s_index[n] < 0) {
                        continue;
                    }
                    const int count = s_nx[n] * s_ny[n];
                    const uint32_t base_index = static_cast<uint32_t>(ti * count);
                    const uint64_t threshold = admit_bound(best[q], d.similar, tau_key);
                    uint16_t* list = scratch + q * L.scratch_keys;
                    for (int c0 = 0; c0 < count; c0 += 64) {
                        const int i = c0 + lane;
                        bool pass = false;
                        bool valid = false;
                        if (i < count) {
                            const float dist = s_stage[n * L.per_frame + i];
                            valid = dist == dist;
                            const uint64_t key = (static_cast<uint64_t>(distance_key(dist)) << 32) | (base_index + i);
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
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int n = 4 * wave + q;
                    if (s_index[n] >= 0) {
                        int end = 0;
#pragma unroll
                        for (int w = 0; w < kFastWaves; ++w) {
                            end += __builtin_amdgcn_readfirstlane(s_parts[n * kFastWaves + w]);
                            if (w < kFastWaves - 1) {
                                part_end[q][w] = end;
                            }
                        }
                        m[q] = end;
                        valid_count[q] += s_nx[n] * s_ny[n];  // finite data: every distance is valid
                    }
                }
            }
            // Window position of key j of anchor q.
            const auto position = [&](int q, int j) -> int {
                if (full) {
                    return scratch[q * L.scratch_keys + j];
                }
                const int n = 4 * wave + q;
                const int w = (j >= part_end[q][0] ? 1 : 0) + (j >= part_end[q][1] ? 1 : 0) +
                              (j >= part_end[q][2] ? 1 : 0);
                const int start = j >= part_end[q][2] ? part_end[q][2]
                                  : j >= part_end[q][1] ? part_end[q][1]
                                  : j >= part_end[q][0] ? part_end[q][0]
                                                        : 0;
                return s_list[(n * kFastWaves + w) * L.sub_keys + j - start];
            };
            const auto key_of = [&](int q, int i) -> uint64_t {
                const int n = 4 * wave + q;
                const uint32_t base_index = static_cast<uint32_t>(ti * s_nx[n] * s_ny[n]);
                return (static_cast<uint64_t>(distance_key(s_stage[n * L.per_frame + i])) << 32) | (base_index + i);
            };
            const int most = max(max(m[0], m[1]), max(m[2], m[3]));
            if (most <= kInsertKeys) {
                // Few keys (the usual case once the lists hold a frame's
                // worth): lane j holds key j of each anchor; the keys go into
                // the four sorted lists one at a time, side by side, a ballot
                // finding each one's place and one lane shift making room.
                // Keys no longer below a list's bound change nothing.
                uint64_t mine[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    mine[q] = lane < m[q] ? key_of(q, position(q, lane)) : kNoKey;
                }
                for (int j = 0; j < most; ++j) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (j < m[q]) {
                            best[q] = wave_insert(best[q], readlane64(mine[q], j), lane);
                        }
                    }
                }
            }
            // Otherwise rounds of 64 listed keys, the four anchors side by
            // side; keys no longer below an anchor's bound are dropped.
            const int rounds = most <= kInsertKeys ? 0 : (most + 63) / 64;
            for (int r = 0; r < rounds; ++r) {
                uint64_t key[4];
                bool any = false;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int at = 64 * r + lane;
                    const uint64_t current = admit_bound(best[q], d.similar, tau_key);
                    uint64_t v = at < m[q] ? key_of(q, position(q, at)) : kNoKey;
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
            // Bounds for the next frame: the largest distance a key below the
            // admit bound can have (NaN where that is not finite).
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int n = 4 * wave + q;
                const uint32_t top = static_cast<uint32_t>(admit_bound(best[q], d.similar, tau_key) >> 32);
                if (lane == 0 && s_index[n] >= 0) {
                    s_bound[n] = top < 0x7F800000u ? __uint_as_float(top) : __builtin_nanf("");
                    s_valid[n] = valid_count[q];
                }
                s_best[wave][q][lane] = best[q];
            }
        } else {
            for (int n = 0; n < kTileAnchors; ++n) {
                const int index = s_index[n];
                if (index < 0) {
                    continue;
                }
                const int count = s_nx[n] * s_ny[n];
                float* dst = out + static_cast<size_t>(index - anchor_base) * L.stride +
                             static_cast<size_t>(ti) * count;
                // The anchor against itself: exactly 0 (NaN for a patch
                // holding NaN, as the exact sum gives).
                const int self = anchor_frame ? (s_ay[n] - s_yl[n]) * s_nx[n] + s_ax[n] - s_xl[n] : -1;
                const float e = s_energy[n];
                for (int i = tid; i < count; i += kFastThreads) {
                    dst[i] = i == self ? (e == e ? 0.0F : e) : s_stage[n * L.per_frame + i];
                }
            }
        }
        // The next frame's first plane once every wave is past the epilogue.
        __syncthreads();
        if (p_last + 1 < total_planes) {
            prefetch(p_last + 1);
            store(p_last + 1);
            __syncthreads();
        }
    }
    if (fused != 0) {
        // select_and_write over the retained keys: k = min(K, valid), the
        // threshold max(tau, d_k), the prefix within it (at most the buffer
        // capacity of find_similar_patches, core.cpp:1279-1287), the K
        // nearest first, the rest in scan order.
        const long long window_cap = static_cast<long long>(d.search_window) * d.search_window * nt;
        const int buffer_cap = static_cast<int>(window_cap < d.soft_cap ? window_cap : d.soft_cap);
        int valid_count[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            valid_count[q] = s_valid[4 * wave + q];
        }
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
            const uint64_t key = s_best[wave][q][lane];
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

// Instantiations: row tiles x union columns per wavefront, by cost.
struct Variant {
    int rows;
    int cols;
};
constexpr Variant kVariants[] = {{1, 6}, {1, 10}, {2, 6}, {2, 8}, {2, 9}, {2, 10}, {3, 10}};
constexpr int kVariantCount = static_cast<int>(sizeof(kVariants) / sizeof(kVariants[0]));

int pick_variant(int max_cols, int max_rows) {
    const int rows = (max_rows + 15) / 16;
    const int cols = (max_cols + 3) / 4;
    for (int v = 0; v < kVariantCount; ++v) {
        if (kVariants[v].rows >= rows && kVariants[v].cols >= cols) {
            return v;
        }
    }
    return -1;
}

TileLayout tile_layout(const StageDesc& d, int max_cols, int variant) {
    const Variant V = kVariants[variant];
    TileLayout L{};
    L.groups = V.cols + 1;
    L.row_bytes = staged_row_bytes(L.groups);
    L.rows_bytes = (16 * V.rows + kFastPatch - 1) * L.row_bytes;
    L.buffer_bytes = L.rows_bytes + kAnchorBytes;
    L.per_frame = d.search_window * d.search_window;
    L.stride = L.per_frame * (d.search_bwd + d.search_fwd + 1);
    L.ebox_pitch = 16 * V.rows + 1;
    L.ebox_bytes = (max_cols * L.ebox_pitch * 4 + 15) / 16 * 16;
    L.stage_bytes = kTileAnchors * L.per_frame * 4;
    L.scratch_keys = (L.per_frame + 1) / 2 * 2;
    L.sub_keys = (d.search_window + 3) / 4 * d.search_window;
    const int lists = 2 * kTileAnchors * std::max(L.scratch_keys, kFastWaves * L.sub_keys);
    const int epilogue = L.ebox_bytes + L.stage_bytes + lists;
    L.shared_bytes = std::max(2 * L.buffer_bytes, epilogue);
    return L;
}

}  // namespace

bool fast_match_supported(const StageDesc& d, int max_cols, int max_rows) {
    if (d.patch != kFastPatch || d.similar <= 1 || max_cols < 1 || max_rows < 1) {
        return false;
    }
    const int variant = pick_variant(max_cols, max_rows);
    if (variant < 0) {
        return false;
    }
    const TileLayout L = tile_layout(d, max_cols, variant);
    return static_cast<long long>(max_cols) * max_rows <= static_cast<long long>(kEnergyPerThread) * kFastThreads &&
           L.shared_bytes <= 150 * 1024;  // of 160 KB per CU
}

bool fast_match_fused(const StageDesc& d, int max_cols, int max_rows) {
    // The running lists hold 64 keys per anchor.
    return fast_match_supported(d, max_cols, max_rows) && d.cap <= 64 && d.similar <= 64;
}

size_t fast_match_shared_bytes(const StageDesc& d, int max_cols, int max_rows) {
    const int variant = pick_variant(max_cols, max_rows);
    return variant < 0 ? 0 : static_cast<size_t>(tile_layout(d, max_cols, variant).shared_bytes);
}

size_t fast_match_stride(const StageDesc& d) {
    return static_cast<size_t>(d.search_window) * d.search_window * (d.search_bwd + d.search_fwd + 1);
}

hipError_t launch_fast_energy(const StageDesc& d, const float* source, float* offsets, unsigned* flags,
                              float* energy, uint32_t* split, double* partials, hipStream_t stream) {
    if (d.patch != kFastPatch) {
        return hipErrorInvalidValue;  // the patch energies are summed over 8 rows
    }
    const int match_channels = d.stage == kFinal ? d.channels : 1;
    const size_t plane = static_cast<size_t>(d.width) * d.height;
    static_assert(2 * kMaxChannels * kOffsetBlocks <= kFastPartialDoubles, "offset partials");
    hipLaunchKernelGGL(channel_partial_kernel, dim3(kOffsetBlocks, static_cast<unsigned>(match_channels)),
                       dim3(256), 0, stream, source, d.channels, plane, d.frames, partials);
    hipLaunchKernelGGL(channel_offset_kernel, dim3(1), dim3(64), 0, stream, partials, match_channels, offsets,
                       flags);
    // The row sums use the split planes' buffer, which is written last.
    float* rows = reinterpret_cast<float*>(split);
    const unsigned plane_blocks = static_cast<unsigned>((plane + 255) / 256);
    hipLaunchKernelGGL(row_energy_kernel, dim3(plane_blocks, static_cast<unsigned>(d.frames * match_channels)),
                       dim3(256), 0, stream, d, source, offsets, match_channels, rows);
    const size_t strip_items = static_cast<size_t>(d.width) * ((d.height + kEnergyStrip - 1) / kEnergyStrip);
    hipLaunchKernelGGL(patch_energy_kernel, dim3(static_cast<unsigned>((strip_items + 255) / 256),
                                                 static_cast<unsigned>(d.frames)),
                       dim3(256), 0, stream, d, rows, match_channels, energy, flags);
    hipLaunchKernelGGL(split_planes_kernel, dim3(plane_blocks, static_cast<unsigned>(d.frames * match_channels)),
                       dim3(256), 0, stream, d, source, offsets, match_channels, split);
    return hipGetLastError();
}

hipError_t launch_fast_match(const StageDesc& d, int max_cols, int max_rows, const uint32_t* split,
                             const unsigned* flags, const float* energy, const uint32_t* anchors,
                             const int* tiles, int tile_count, int anchor_base, float* staged,
                             bool fused, int* counts, uint32_t* positions, float* distances,
                             hipStream_t stream) {
    if (tile_count <= 0) {
        return hipSuccess;
    }
    const int variant = pick_variant(max_cols, max_rows);
    if (variant < 0) {
        return hipErrorInvalidValue;
    }
    const TileLayout L = tile_layout(d, max_cols, variant);
    const size_t shared = static_cast<size_t>(L.shared_bytes);
    const int fused_flag = fused ? 1 : 0;
#define VNLBHIP_FAST(R, A)                                                                              \
    hipLaunchKernelGGL((match_fast_kernel<R, A>), dim3(static_cast<unsigned>(tile_count)), dim3(kFastThreads), \
                       shared, stream, d, L, split, flags, energy, anchors, tiles, anchor_base, fused_flag, \
                       staged, counts, positions, distances)
    switch (variant) {
        case 0: VNLBHIP_FAST(1, 6); break;
        case 1: VNLBHIP_FAST(1, 10); break;
        case 2: VNLBHIP_FAST(2, 6); break;
        case 3: VNLBHIP_FAST(2, 8); break;
        case 4: VNLBHIP_FAST(2, 9); break;
        case 5: VNLBHIP_FAST(2, 10); break;
        default: VNLBHIP_FAST(3, 10); break;
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
