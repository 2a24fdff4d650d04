// Plain data of the learned candidate proposer shared by host and device
// code: parameter layout and the weights as kernels read them.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace vnlbhip {

constexpr int kPropIn = 128;
constexpr int kPropHidden = 256;
constexpr int kPropMaxProposed = 256;  // proposals (with score ties) per anchor

// Offsets of the flat fp32 parameter vector.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kPropHidden * kPropIn;
constexpr int kOffW2 = kOffB1 + kPropHidden;
constexpr int kOffB2 = kOffW2 + kPropHidden * kPropHidden;
constexpr int kOffW3 = kOffB2 + kPropHidden;
constexpr int kOffB3 = kOffW3 + kPropHidden;
constexpr int kPropParams = kOffB3 + 1;

// FP8 (OCP e4m3) inference copies: the input row is scaled by x_scale[k]
// and W1 holds W1[n][k] / x_scale[k] times a per-row scale r1[n]; the hidden
// row is scaled by h_scale and W2 rows by r2[n].  The products are taken
// back with w1_inv[n] = 1 / r1[n] and w2_inv[n] = 1 / (h_scale r2[n]).
// scales: x_scale[128], w1_inv[256], w2_inv[256], h_scale.
constexpr int kPropScaleX = 0;
constexpr int kPropScaleW1 = kPropScaleX + kPropIn;
constexpr int kPropScaleW2 = kPropScaleW1 + kPropHidden;
constexpr int kPropScaleH = kPropScaleW2 + kPropHidden;
constexpr int kPropScales = kPropScaleH + 1;

struct ProposerFp8 {
    const uint8_t* w1;      // [256][128]
    const uint8_t* w2;      // [256][256]
    const float* scales;    // kPropScales
};

// The network's weights as the kernels read them.
struct ProposerWeights {
    const __bf16* w1;   // [256][128]
    const float* b1;    // [256]
    const __bf16* w2;   // [256][256]
    const float* b2;    // [256]
    const float* w3;    // [256]
    const float* b3;    // [1]
    ProposerFp8 fp8;
};

}  // namespace vnlbhip
