// Device helpers shared by the stage kernels: geometry rules of the
// reference CPU path and wave/block reductions for 64-lane wavefronts.
#pragma once

#include "common.h"
#include "geometry.h"

namespace vnlbhip {

// XCD-aware work distribution (kXcds, counters: common.h).  The GPU
// dispatches the workgroups of a launch round-robin over its 8 XCDs (block b
// on XCD b % 8), and each XCD has its own L2: work that shares data
// (neighbouring anchors read overlapping patches) should run on one XCD.

// Remap of block b of n: the blocks of XCD x take the contiguous range
// [x * (n / 8) + min(x, n % 8), ...) of the work items, in order.
__device__ inline int xcd_block(int b, int n) {
    const int x = b % kXcds;
    const int per = n / kXcds;
    const int rem = n % kXcds;
    return x * per + (x < rem ? x : rem) + b / kXcds;
}

// Group queue over kXcdCounterWords zeroed ints: XCD x hands out the items
// of its contiguous eighth [count x / 8, count (x + 1) / 8) of [0, count) in
// order, then a caller whose eighth is exhausted takes from the following
// XCDs' eighths.  `visited` (0 on the first call) is the caller's own state.
// Returns count when every eighth is exhausted.  Called by one lane.
__device__ inline int xcd_next(int* counters, int count, int home, int& visited) {
    while (visited < kXcds) {
        const int x = (home + visited) % kXcds;
        const int begin = static_cast<int>(static_cast<long long>(count) * x / kXcds);
        const int end = static_cast<int>(static_cast<long long>(count) * (x + 1) / kXcds);
        if (begin < end) {
            const int g = begin + atomicAdd(counters + x * kXcdCounterStride, 1);
            if (g < end) {
                return g;
            }
        }
        ++visited;
    }
    return count;
}

__device__ inline float video_at(const float* __restrict__ video,
                                 const StageDesc& d, int x, int y, int frame,
                                 int channel) {
    return video[((static_cast<size_t>(frame) * d.channels + channel) * d.height +
                  y) *
                     d.width +
                 x];
}

// Value of the lane kShift places further along the same 16-lane row
// (DPP row_ror: a VALU operand modifier, no LDS traffic).
template <int kShift>
__device__ inline float row_rotate(float v) {
    static_assert(kShift >= 1 && kShift <= 15, "row_ror takes 1..15");
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x120 + kShift, 0xF, 0xF,
                                           true));
}

// Halving-tree sum over a 16-lane segment (the ReduceSum order of the CPU
// restatement); every lane of the segment ends with the total.  Level h adds
// lane l and lane l + 16 / 2^h of the previous level: after level h the
// values repeat with period 16 / 2^h, so the rotation by that period pairs
// lane l with lane l XOR (16 / 2^h) of the halving tree, and float addition
// is commutative, so every lane holds exactly the tree's sums.
__device__ inline float segment16_sum(float v) {
    v = v + row_rotate<8>(v);
    v = v + row_rotate<4>(v);
    v = v + row_rotate<2>(v);
    v = v + row_rotate<1>(v);
    return v;
}

__device__ inline int wave_inclusive_scan(int value) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int offset = 1; offset < 64; offset <<= 1) {
        const int other = __shfl_up(value, offset, 64);
        if (lane >= offset) {
            value += other;
        }
    }
    return value;
}

// Exclusive prefix sum over the block in thread order.  scratch holds one
// int per wave.  All threads must call it.
__device__ inline int block_exclusive_scan(int value, int& total, int* scratch) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int waves = (blockDim.x + 63) >> 6;
    const int inclusive = wave_inclusive_scan(value);
    if (lane == 63) {
        scratch[wave] = inclusive;
    }
    __syncthreads();
    int base = 0;
    int sum = 0;
    for (int w = 0; w < waves; ++w) {
        const int v = scratch[w];
        if (w < wave) {
            base += v;
        }
        sum += v;
    }
    __syncthreads();
    total = sum;
    return base + inclusive - value;
}

__device__ inline float wave_max(float v) {
#pragma unroll
    for (int offset = 32; offset >= 1; offset >>= 1) {
        v = fmaxf(v, __shfl_xor(v, offset, 64));
    }
    return v;
}

__device__ inline float wave_min(float v) {
#pragma unroll
    for (int offset = 32; offset >= 1; offset >>= 1) {
        v = fminf(v, __shfl_xor(v, offset, 64));
    }
    return v;
}

__device__ inline int wave_min_int(int v) {
#pragma unroll
    for (int offset = 32; offset >= 1; offset >>= 1) {
        v = min(v, __shfl_xor(v, offset, 64));
    }
    return v;
}

__device__ inline int wave_max_int(int v) {
#pragma unroll
    for (int offset = 32; offset >= 1; offset >>= 1) {
        v = max(v, __shfl_xor(v, offset, 64));
    }
    return v;
}

__device__ inline double wave_max_double(double v) {
#pragma unroll
    for (int offset = 32; offset >= 1; offset >>= 1) {
        v = fmax(v, __shfl_xor(v, offset, 64));
    }
    return v;
}

__device__ inline double wave_sum_double(double v) {
#pragma unroll
    for (int offset = 32; offset >= 1; offset >>= 1) {
        v = v + __shfl_xor(v, offset, 64);
    }
    return v;
}

// v from lane (lane ^ kMask) of the same 16-lane row, with DPP moves only:
// row_mirror is xor 15, row_half_mirror xor 7 and quad_perm any xor below 4.
// Every lane has a source in these patterns, so the moves are issued with
// bound_ctrl set: the "old" operand is then dead and no zero needs to be
// materialised before each move.
template <int kMask>
__device__ inline int xor16_bits(int x) {
    static_assert(kMask >= 0 && kMask < 16, "xor within a row of 16 lanes");
    constexpr int kAfterMirror = (kMask & 8) != 0 ? (kMask ^ 15) : kMask;
    constexpr int kQuad = (kAfterMirror & 4) != 0 ? (kAfterMirror ^ 7) : kAfterMirror;
    constexpr int kQuadCtrl = (0 ^ kQuad) | ((1 ^ kQuad) << 2) | ((2 ^ kQuad) << 4) | ((3 ^ kQuad) << 6);
    if constexpr ((kMask & 8) != 0) {
        x = __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, true);  // row_mirror
    }
    if constexpr ((kAfterMirror & 4) != 0) {
        x = __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, true);  // row_half_mirror
    }
    if constexpr (kQuad != 0) {
        x = __builtin_amdgcn_update_dpp(0, x, kQuadCtrl, 0xF, 0xF, true);
    }
    return x;
}

template <int kMask>
__device__ inline float xor16(float value) {
    return __builtin_bit_cast(float, xor16_bits<kMask>(__builtin_bit_cast(int, value)));
}

// 64-bit value from lane (lane ^ kMask) of the wavefront: DPP within rows of
// 16 lanes, ds_bpermute across rows.
template <int kMask>
__device__ inline uint64_t xor64(uint64_t value) {
    if constexpr (kMask < 16) {
        const uint32_t lo = static_cast<uint32_t>(xor16_bits<kMask>(static_cast<int>(value & 0xFFFFFFFFu)));
        const uint32_t hi = static_cast<uint32_t>(xor16_bits<kMask>(static_cast<int>(value >> 32)));
        return (static_cast<uint64_t>(hi) << 32) | lo;
    } else {
        return __shfl_xor(value, kMask, 64);
    }
}

}  // namespace vnlbhip
