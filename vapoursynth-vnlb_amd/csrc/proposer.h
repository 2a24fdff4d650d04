// Learned candidate proposer: the network and its input, shared by the
// training kernels (proposer.hip) and the proposing matcher (match.hip).
//
// The reference scans every candidate of the search window
// (find_similar_patches, src/core/core.cpp:1245-1362).  The proposing
// matcher first scores the window with a small MLP and computes exact
// distances only for the best scored candidates.  The network:
//
//   x  = features(anchor, candidate)                        128 floats
//   h1 = relu(W1 x + b1)        W1 [256][128]  (bf16 operands, fp32 sums)
//   h2 = relu(W2 h1 + b2)       W2 [256][256]  (bf16 operands, fp32 sums)
//   s  = w3 . h2 + b3           w3 [256]       (fp32)
//
// Features of an 8x8 patch pair (patch time 1 or 2; channel 0 of the
// matched clip): the 2x2 box means of each patch frame (4x4 per frame,
// index t * 16 + i * 4 + j), the anchor's and the candidate's means minus
// 0.5, their absolute differences times 4, the offset (dx / 16, dy / 16,
// dt / 4), the box-mean SSD times 4, zeros up to 128.
#pragma once

#include "common.h"
#include "device_util.h"
#include "proposer_types.h"

namespace vnlbhip {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 prop_bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float prop_floatx4;

constexpr int kPropBoxes = 32;          // box means per patch (patch time <= 2)
constexpr int kPropTile = 64;           // candidates scored per tile
constexpr int kPropXPitch = kPropIn + 8;        // bf16 row pitch of the input tile
constexpr int kPropHPitch = kPropHidden + 8;    // bf16 row pitch of the hidden tile

// 2x2 box means of an 8x8 patch frame whose top-left sample is at(0, 0) and
// whose rows are `pitch` floats apart: out[i * 4 + j].
__device__ inline void patch_box_means(const float* at, int pitch, float* out) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float* p = at + (2 * i) * pitch + 2 * j;
            const float p00 = p[0];
            const float p01 = p[1];
            const float p10 = p[pitch];
            const float p11 = p[pitch + 1];
            out[i * 4 + j] = ((p00 + p01) + (p10 + p11)) * 0.25F;
        }
    }
}

// The 128 input features of one pair from the box means of the anchor (a)
// and the candidate (c), `boxes` = 16 * patch time of them.
__device__ inline void proposer_features(const float* a, const float* c, int boxes, int dx, int dy, int dt,
                                         float* x) {
    float ssd = 0.0F;
    for (int k = 0; k < kPropBoxes; ++k) {
        if (k < boxes) {
            const float diff = a[k] - c[k];
            x[k] = a[k] - 0.5F;
            x[32 + k] = c[k] - 0.5F;
            x[64 + k] = fabsf(diff) * 4.0F;
            ssd = ssd + diff * diff;
        } else {
            x[k] = 0.0F;
            x[32 + k] = 0.0F;
            x[64 + k] = 0.0F;
        }
    }
    x[96] = static_cast<float>(dx) * (1.0F / 16.0F);
    x[97] = static_cast<float>(dy) * (1.0F / 16.0F);
    x[98] = static_cast<float>(dt) * 0.25F;
    x[99] = ssd * 4.0F;
    for (int k = 100; k < kPropIn; ++k) {
        x[k] = 0.0F;
    }
}

// Order-preserving map of a score to an unsigned key that sorts the
// highest score first.
__device__ inline uint32_t score_rank_key(float s) {
    const uint32_t bits = __float_as_uint(s);
    const uint32_t ascending = (bits & 0x80000000u) != 0 ? ~bits : (bits | 0x80000000u);
    return ~ascending;
}

__device__ inline float score_from_rank_key(uint32_t key) {
    const uint32_t ascending = ~key;
    const uint32_t bits = (ascending & 0x80000000u) != 0 ? (ascending & 0x7FFFFFFFu) : ~ascending;
    return __uint_as_float(bits);
}

// Scores of one tile of kPropTile input rows (bf16 in LDS, row pitch
// kPropXPitch), all 256 threads of the block (4 waves): layer 1 into the
// bf16 hidden tile h (pitch kPropHPitch), layer 2 and the output layer in
// registers.  Wave w owns hidden units [64 w, 64 w + 64).  partial: 4 * 64
// floats of LDS.  Ends with the scores in scores[0 .. 63] (LDS) after a
// barrier.
__device__ inline void proposer_tile_scores(const __bf16* __restrict__ x, __bf16* __restrict__ h,
                                            float* __restrict__ partial, float* __restrict__ scores,
                                            const ProposerWeights& wt) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15;
    const int kb = lane >> 4;
    const int n0 = wave * 64;
    // Layer 1: Z1[64 x 64 of this wave] = X[64 x 128] W1^T.
    prop_floatx4 acc[4][4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            acc[rt][ct] = prop_floatx4{0.0F, 0.0F, 0.0F, 0.0F};
        }
    }
#pragma unroll
    for (int ks = 0; ks < kPropIn / 32; ++ks) {
        prop_bf16x8 b[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            b[ct] = *reinterpret_cast<const prop_bf16x8*>(wt.w1 + static_cast<size_t>(n0 + 16 * ct + col) * kPropIn +
                                                          32 * ks + 8 * kb);
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const prop_bf16x8 a =
                *reinterpret_cast<const prop_bf16x8*>(x + (16 * rt + col) * kPropXPitch + 32 * ks + 8 * kb);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[ct], acc[rt][ct], 0, 0, 0);
            }
        }
    }
    // acc[rt][ct][v] = Z1[16 rt + 4 kb + v][n0 + 16 ct + col]
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int n = n0 + 16 * ct + col;
        const float bias = wt.b1[n];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float z = acc[rt][ct][v] + bias;
                h[(16 * rt + 4 * kb + v) * kPropHPitch + n] = static_cast<__bf16>(z > 0.0F ? z : 0.0F);
            }
        }
    }
    __syncthreads();
    // Layer 2: Z2[64 x 64 of this wave] = H1[64 x 256] W2^T.
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            acc[rt][ct] = prop_floatx4{0.0F, 0.0F, 0.0F, 0.0F};
        }
    }
#pragma unroll 2
    for (int ks = 0; ks < kPropHidden / 32; ++ks) {
        prop_bf16x8 b[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            b[ct] = *reinterpret_cast<const prop_bf16x8*>(wt.w2 + static_cast<size_t>(n0 + 16 * ct + col) *
                                                                      kPropHidden + 32 * ks + 8 * kb);
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const prop_bf16x8 a =
                *reinterpret_cast<const prop_bf16x8*>(h + (16 * rt + col) * kPropHPitch + 32 * ks + 8 * kb);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[ct], acc[rt][ct], 0, 0, 0);
            }
        }
    }
    // Output layer: per row, sum over this wave's 64 units of relu(z2) w3,
    // in unit order within a lane's four columns, then the 16 columns of a
    // row segment (halving tree).
    float w3v[4];
    float b2v[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        w3v[ct] = wt.w3[n0 + 16 * ct + col];
        b2v[ct] = wt.b2[n0 + 16 * ct + col];
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float sum = 0.0F;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const float z = acc[rt][ct][v] + b2v[ct];
                sum = sum + (z > 0.0F ? z : 0.0F) * w3v[ct];
            }
            sum = segment16_sum(sum);
            if (col == 0) {
                partial[wave * 64 + 16 * rt + 4 * kb + v] = sum;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < kPropTile) {
        const int r = threadIdx.x;
        scores[r] = ((partial[r] + partial[64 + r]) + (partial[128 + r] + partial[192 + r])) + wt.b3[0];
    }
    __syncthreads();
}

// ---- FP8 (OCP e4m3) inference ----

constexpr int kPropXPitch8 = kPropIn + 16;      // bytes per row of the fp8 input tile
constexpr int kPropHPitch8 = kPropHidden + 16;  // bytes per row of the fp8 hidden tile
constexpr float kFp8Max = 448.0F;

__device__ inline float fp8_clamp(float v) { return fminf(fmaxf(v, -kFp8Max), kFp8Max); }

// Four floats as four e4m3 bytes (byte i = value i), rounded by the
// hardware conversion; values beyond the format's range are clamped first.
__device__ inline uint32_t pack_fp8x4(float a, float b, float c, float d) {
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(fp8_clamp(a), fp8_clamp(b), 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(fp8_clamp(c), fp8_clamp(d), w, true);
    return static_cast<uint32_t>(w);
}

__device__ inline uint8_t to_fp8(float v) { return static_cast<uint8_t>(pack_fp8x4(v, 0.0F, 0.0F, 0.0F) & 0xFFu); }

// proposer_tile_scores with fp8 operands (v_mfma_f32_16x16x32_fp8_fp8, fp32
// accumulation): x holds the scaled input rows (row pitch kPropXPitch8
// bytes), h receives the scaled hidden rows (pitch kPropHPitch8).
__device__ inline void proposer_tile_scores_fp8(const uint8_t* __restrict__ x, uint8_t* __restrict__ h,
                                                float* __restrict__ partial, float* __restrict__ scores,
                                                const ProposerWeights& wt) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15;
    const int kb = lane >> 4;
    const int n0 = wave * 64;
    const float* scales = wt.fp8.scales;
    const float h_scale = scales[kPropScaleH];
    prop_floatx4 acc[4][4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            acc[rt][ct] = prop_floatx4{0.0F, 0.0F, 0.0F, 0.0F};
        }
    }
#pragma unroll
    for (int ks = 0; ks < kPropIn / 32; ++ks) {
        long b[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            b[ct] = *reinterpret_cast<const long*>(wt.fp8.w1 + static_cast<size_t>(n0 + 16 * ct + col) * kPropIn +
                                                   32 * ks + 8 * kb);
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const long a = *reinterpret_cast<const long*>(x + (16 * rt + col) * kPropXPitch8 + 32 * ks + 8 * kb);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b[ct], acc[rt][ct], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int n = n0 + 16 * ct + col;
        const float inv = scales[kPropScaleW1 + n];
        const float bias = wt.b1[n];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float z = acc[rt][ct][v] * inv + bias;
                h[(16 * rt + 4 * kb + v) * kPropHPitch8 + n] = to_fp8((z > 0.0F ? z : 0.0F) * h_scale);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            acc[rt][ct] = prop_floatx4{0.0F, 0.0F, 0.0F, 0.0F};
        }
    }
#pragma unroll 2
    for (int ks = 0; ks < kPropHidden / 32; ++ks) {
        long b[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            b[ct] = *reinterpret_cast<const long*>(wt.fp8.w2 + static_cast<size_t>(n0 + 16 * ct + col) * kPropHidden +
                                                   32 * ks + 8 * kb);
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const long a = *reinterpret_cast<const long*>(h + (16 * rt + col) * kPropHPitch8 + 32 * ks + 8 * kb);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b[ct], acc[rt][ct], 0, 0, 0);
            }
        }
    }
    float w3v[4];
    float b2v[4];
    float inv2[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int n = n0 + 16 * ct + col;
        w3v[ct] = wt.w3[n];
        b2v[ct] = wt.b2[n];
        inv2[ct] = scales[kPropScaleW2 + n];
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float sum = 0.0F;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const float z = acc[rt][ct][v] * inv2[ct] + b2v[ct];
                sum = sum + (z > 0.0F ? z : 0.0F) * w3v[ct];
            }
            sum = segment16_sum(sum);
            if (col == 0) {
                partial[wave * 64 + 16 * rt + 4 * kb + v] = sum;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < kPropTile) {
        const int r = threadIdx.x;
        scores[r] = ((partial[r] + partial[64 + r]) + (partial[128 + r] + partial[192 + r])) + wt.b3[0];
    }
    __syncthreads();
}

// Stores 32 consecutive input features of one row into the input tile:
// bf16, or fp8 scaled by x_scale[k0 + i].
template <bool kFp8>
__device__ inline void store_features32(void* tile, int row, int k0, const float* v, const float* x_scale) {
    if constexpr (kFp8) {
        uint32_t* dst = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(tile) + row * kPropXPitch8 + k0);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            dst[i] = pack_fp8x4(v[4 * i] * x_scale[k0 + 4 * i], v[4 * i + 1] * x_scale[k0 + 4 * i + 1],
                                v[4 * i + 2] * x_scale[k0 + 4 * i + 2], v[4 * i + 3] * x_scale[k0 + 4 * i + 3]);
        }
    } else {
        __bf16* dst = static_cast<__bf16*>(tile) + row * kPropXPitch + k0;
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            dst[i] = static_cast<__bf16>(v[i]);
        }
    }
}

}  // namespace vnlbhip
