// Training and scoring kernels of the learned candidate proposer (the network
// and its features are described in proposer.h).
//
// Training minimises, over groups of rows (one anchor's candidates each),
// the listwise loss -sum_r t_r log softmax(s)_r with targets t summing to 1
// per group (the exact k nearest of the plain matcher share the mass).
// Matrix products run on the matrix cores (v_mfma_f32_16x16x32_bf16: bf16
// operands, fp32 accumulation) in one generic tiled kernel; the fp32 master
// weights take Adam steps and are copied to bf16 for the products.
#include <cfloat>

#include "kernels.h"
#include "proposer.h"

namespace vnlbhip {
namespace {

constexpr int kGemmTile = 64;
constexpr int kGemmK = 32;
constexpr int kGemmPitch = kGemmK + 8;  // bf16 per staged row (16-byte pad)

__device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// He-uniform weights (limit sqrt(6 / fan_in)), zero biases; the output
// weights at a tenth of that so the first scores are small.
__global__ void init_kernel(float* __restrict__ params, uint64_t seed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kPropParams) {
        return;
    }
    const double u = static_cast<double>(splitmix64(seed ^ (0xA24BAED4963EE407ull * (i + 1))) >> 11) *
                     (1.0 / 9007199254740992.0);
    float limit = 0.0F;
    if (i < kOffB1) {
        limit = sqrtf(6.0F / kPropIn);
    } else if (i >= kOffW2 && i < kOffB2) {
        limit = sqrtf(6.0F / kPropHidden);
    } else if (i >= kOffW3 && i < kOffB3) {
        limit = 0.1F * sqrtf(6.0F / kPropHidden);
    }
    params[i] = static_cast<float>((2.0 * u - 1.0) * limit);
}

__global__ void to_bf16_kernel(const float* __restrict__ src, __bf16* __restrict__ dst, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        dst[i] = static_cast<__bf16>(src[i]);
    }
}

// One input row per thread from global memory (channel 0 of the clip).
__global__ void features_kernel(const float* __restrict__ clip, int width, int height, int channels,
                                int patch_time, const uint32_t* __restrict__ anchors,
                                const uint32_t* __restrict__ candidates, int rows, __bf16* __restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) {
        return;
    }
    const uint32_t a = anchors[r];
    const uint32_t c = candidates[r];
    const size_t plane = static_cast<size_t>(width) * height;
    float am[kPropBoxes];
    float cm[kPropBoxes];
    for (int t = 0; t < patch_time; ++t) {
        const float* ap = clip + (static_cast<size_t>(position_frame(a) + t) * channels) * plane +
                          static_cast<size_t>(position_y(a)) * width + position_x(a);
        const float* cp = clip + (static_cast<size_t>(position_frame(c) + t) * channels) * plane +
                          static_cast<size_t>(position_y(c)) * width + position_x(c);
        patch_box_means(ap, width, am + 16 * t);
        patch_box_means(cp, width, cm + 16 * t);
    }
    float x[kPropIn];
    proposer_features(am, cm, 16 * patch_time, position_x(c) - position_x(a), position_y(c) - position_y(a),
                      position_frame(c) - position_frame(a), x);
    __bf16* row = out + static_cast<size_t>(r) * kPropIn;
    for (int k = 0; k < kPropIn; ++k) {
        row[k] = static_cast<__bf16>(x[k]);
    }
}

template <class T> __device__ inline float load_as_float(const T* p) { return static_cast<float>(*p); }

// C[m][n] = sum_k A(m, k) B(k, n) with A(m, k) = a[m * a_rs + k * a_cs] and
// B(k, n) = b[k * b_rs + n * b_cs], operands rounded to bf16 when staged.
// 64 x 64 output tiles, four waves of 32 x 32.  Epilogue: + bias[n] when
// given; fp32 result to c32 (ldc) and/or relu(result) as bf16 to c16.
template <class TA, class TB>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const TA* __restrict__ a, long long a_rs,
                                                   long long a_cs, const TB* __restrict__ b, long long b_rs,
                                                   long long b_cs, const float* __restrict__ bias,
                                                   float* __restrict__ c32, __bf16* __restrict__ c16, int ldc) {
    __shared__ __attribute__((aligned(16))) __bf16 s_a[kGemmTile * kGemmPitch];
    __shared__ __attribute__((aligned(16))) __bf16 s_b[kGemmTile * kGemmPitch];
    const int m0 = blockIdx.y * kGemmTile;
    const int n0 = blockIdx.x * kGemmTile;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15;
    const int kb = lane >> 4;
    const int wm = (wave >> 1) * 32;
    const int wn = (wave & 1) * 32;
    prop_floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = prop_floatx4{0.0F, 0.0F, 0.0F, 0.0F};
        }
    }
    // Staging order: k fastest when the operand is k-contiguous, else the
    // row index fastest.
    const bool a_k_fast = a_cs == 1;
    const bool b_k_fast = b_rs == 1;
    for (int k0 = 0; k0 < K; k0 += kGemmK) {
        for (int e = threadIdx.x; e < kGemmTile * kGemmK; e += blockDim.x) {
            const int i = a_k_fast ? e / kGemmK : e % kGemmTile;
            const int k = a_k_fast ? e % kGemmK : e / kGemmTile;
            const int m = m0 + i;
            const int kk = k0 + k;
            float v = 0.0F;
            if (m < M && kk < K) {
                v = load_as_float(a + m * a_rs + kk * a_cs);
            }
            s_a[i * kGemmPitch + k] = static_cast<__bf16>(v);
        }
        for (int e = threadIdx.x; e < kGemmTile * kGemmK; e += blockDim.x) {
            const int j = b_k_fast ? e / kGemmK : e % kGemmTile;
            const int k = b_k_fast ? e % kGemmK : e / kGemmTile;
            const int n = n0 + j;
            const int kk = k0 + k;
            float v = 0.0F;
            if (n < N && kk < K) {
                v = load_as_float(b + kk * b_rs + n * b_cs);
            }
            s_b[j * kGemmPitch + k] = static_cast<__bf16>(v);
        }
        __syncthreads();
        prop_bf16x8 av[2];
        prop_bf16x8 bv[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            av[i] = *reinterpret_cast<const prop_bf16x8*>(s_a + (wm + 16 * i + col) * kGemmPitch + 8 * kb);
            bv[i] = *reinterpret_cast<const prop_bf16x8*>(s_b + (wn + 16 * i + col) * kGemmPitch + 8 * kb);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // acc[i][j][v] = C[wm + 16 i + 4 kb + v][wn + 16 j + col]
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int n = n0 + wn + 16 * j + col;
        if (n >= N) {
            continue;
        }
        const float add = bias != nullptr ? bias[n] : 0.0F;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + 16 * i + 4 * kb + v;
                if (m < M) {
                    const float z = acc[i][j][v] + add;
                    if (c32 != nullptr) {
                        c32[static_cast<size_t>(m) * ldc + n] = z;
                    }
                    if (c16 != nullptr) {
                        c16[static_cast<size_t>(m) * ldc + n] = static_cast<__bf16>(z > 0.0F ? z : 0.0F);
                    }
                }
            }
        }
    }
}

// s[m] = sum_n relu(z2[m][n]) w3[n] + b3 (one wave per row; lane l sums
// units 4 l .. 4 l + 3 in order, then a butterfly over the wave).
__global__ void output_kernel(const float* __restrict__ z2, const float* __restrict__ params, int rows,
                              float* __restrict__ s) {
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) {
        return;
    }
    const float* w3 = params + kOffW3;
    float sum = 0.0F;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int n = 4 * lane + q;
        const float z = z2[static_cast<size_t>(r) * kPropHidden + n];
        sum = sum + (z > 0.0F ? z : 0.0F) * w3[n];
    }
    for (int offset = 32; offset >= 1; offset >>= 1) {
        sum = sum + __shfl_xor(sum, offset, 64);
    }
    if (lane == 0) {
        s[r] = sum + params[kOffB3];
    }
}

// Listwise loss per group (one wave per group): ds = (softmax(s) - t) / G,
// loss_out[g] = -sum t log softmax(s).
__global__ void loss_kernel(const float* __restrict__ s, const float* __restrict__ targets, int groups,
                            int group, float* __restrict__ ds, float* __restrict__ loss_out) {
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= groups) {
        return;
    }
    const float* sg = s + static_cast<size_t>(g) * group;
    const float* tg = targets + static_cast<size_t>(g) * group;
    float mx = -FLT_MAX;
    for (int r = lane; r < group; r += 64) {
        mx = fmaxf(mx, sg[r]);
    }
    mx = wave_max(mx);
    float z = 0.0F;
    for (int r = lane; r < group; r += 64) {
        z = z + expf(sg[r] - mx);
    }
    for (int offset = 32; offset >= 1; offset >>= 1) {
        z = z + __shfl_xor(z, offset, 64);
    }
    const float log_z = logf(z);
    const float inv_groups = 1.0F / static_cast<float>(groups);
    float loss = 0.0F;
    for (int r = lane; r < group; r += 64) {
        const float logp = sg[r] - mx - log_z;
        const float t = tg[r];
        ds[static_cast<size_t>(g) * group + r] = (expf(logp) - t) * inv_groups;
        loss = loss - t * logp;
    }
    for (int offset = 32; offset >= 1; offset >>= 1) {
        loss = loss + __shfl_xor(loss, offset, 64);
    }
    if (lane == 0) {
        loss_out[g] = loss;
    }
}

// dz2[m][n] = ds[m] w3[n] (z2 > 0)
__global__ void output_grad_kernel(const float* __restrict__ z2, const float* __restrict__ ds,
                                   const float* __restrict__ params, int rows, float* __restrict__ dz2) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(rows) * kPropHidden) {
        return;
    }
    const int n = static_cast<int>(i % kPropHidden);
    const size_t m = i / kPropHidden;
    dz2[i] = z2[i] > 0.0F ? ds[m] * params[kOffW3 + n] : 0.0F;
}

// dz[m][n] *= (z[m][n] > 0)
__global__ void relu_mask_kernel(const float* __restrict__ z, float* __restrict__ dz, size_t count) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count && !(z[i] > 0.0F)) {
        dz[i] = 0.0F;
    }
}

constexpr int kColChunks = 64;

// Column sums in two deterministic passes: part[chunk][n] = sum over the
// chunk's rows (in order) of f(a[m][n]) * (w ? w[m] : 1), f = relu when
// relu_a; then out[n] = sum over chunks in order.
__global__ void column_partial_kernel(const float* __restrict__ a, int rows, int cols, const float* __restrict__ w,
                                      int relu_a, float* __restrict__ part) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (n >= cols) {
        return;
    }
    const int per = (rows + kColChunks - 1) / kColChunks;
    const int lo = chunk * per;
    const int hi = min(rows, lo + per);
    float sum = 0.0F;
    for (int m = lo; m < hi; ++m) {
        float v = a[static_cast<size_t>(m) * cols + n];
        if (relu_a != 0) {
            v = v > 0.0F ? v : 0.0F;
        }
        sum = sum + (w != nullptr ? v * w[m] : v);
    }
    part[static_cast<size_t>(chunk) * cols + n] = sum;
}

__global__ void column_total_kernel(const float* __restrict__ part, int cols, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= cols) {
        return;
    }
    float sum = 0.0F;
    for (int c = 0; c < kColChunks; ++c) {
        sum = sum + part[static_cast<size_t>(c) * cols + n];
    }
    out[n] = sum;
}

// db3 = sum ds (one block, fixed order).
__global__ void sum_kernel(const float* __restrict__ v, int count, float* __restrict__ out) {
    __shared__ float s_part[256];
    float sum = 0.0F;
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        sum = sum + v[i];
    }
    s_part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float total = 0.0F;
        for (int i = 0; i < static_cast<int>(blockDim.x); ++i) {
            total = total + s_part[i];
        }
        out[0] = total;
    }
}

__global__ void adam_kernel(float* __restrict__ params, const float* __restrict__ grads, float* __restrict__ m,
                            float* __restrict__ v, int count, float lr, float beta1, float beta2, float eps,
                            float correction1, float correction2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) {
        return;
    }
    const float g = grads[i];
    const float mi = beta1 * m[i] + (1.0F - beta1) * g;
    const float vi = beta2 * v[i] + (1.0F - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] = params[i] - lr * (mi / correction1) / (sqrtf(vi / correction2) + eps);
}

// Scores of rows of inputs with the inference tile (the matcher's code).
template <bool kFp8>
__global__ __launch_bounds__(256) void scores_kernel(const __bf16* __restrict__ x, int rows, ProposerWeights wt,
                                                     float* __restrict__ scores) {
    __shared__ __attribute__((aligned(16))) __bf16 s_x[kPropTile * kPropXPitch];
    __shared__ __attribute__((aligned(16))) __bf16 s_h[kPropTile * kPropHPitch];
    __shared__ float s_partial[4 * 64];
    __shared__ float s_scores[kPropTile];
    const int r0 = blockIdx.x * kPropTile;
    {
        // Thread (r, q) stages inputs 32 q .. 32 q + 31 of row r.
        const int r = threadIdx.x >> 2;
        const int q = threadIdx.x & 3;
        float v[32];
        for (int k = 0; k < 32; ++k) {
            v[k] = r0 + r < rows ? static_cast<float>(x[static_cast<size_t>(r0 + r) * kPropIn + 32 * q + k]) : 0.0F;
        }
        store_features32<kFp8>(s_x, r, 32 * q, v, kFp8 ? wt.fp8.scales + kPropScaleX : nullptr);
    }
    __syncthreads();
    if constexpr (kFp8) {
        proposer_tile_scores_fp8(reinterpret_cast<const uint8_t*>(s_x), reinterpret_cast<uint8_t*>(s_h), s_partial,
                                 s_scores, wt);
    } else {
        proposer_tile_scores(s_x, s_h, s_partial, s_scores, wt);
    }
    if (threadIdx.x < kPropTile && r0 + static_cast<int>(threadIdx.x) < rows) {
        scores[r0 + threadIdx.x] = s_scores[threadIdx.x];
    }
}

// ---- FP8 calibration ----

// x_scale[k] = 224 / max_r |x[r][k]| (1 where the column is zero): half the
// format's range, room for inputs beyond the sample.
__global__ void input_scale_kernel(const __bf16* __restrict__ x, int rows, float* __restrict__ scales) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kPropIn) {
        return;
    }
    float mx = 0.0F;
    for (int r = 0; r < rows; ++r) {
        mx = fmaxf(mx, fabsf(static_cast<float>(x[static_cast<size_t>(r) * kPropIn + k])));
    }
    scales[kPropScaleX + k] = mx > 0.0F ? 0.5F * kFp8Max / mx : 1.0F;
}

// One row of a weight matrix per block: w'[k] = w[n][k] * fold[k] (fold:
// 1 / x_scale, or none), row scale r = 448 / max |w'| (1 for a zero row),
// out[n][k] = fp8(w'[k] r), inv[n] = 1 / (r * extra) with extra = h_scale
// for W2 (read from scales) or 1.
__global__ __launch_bounds__(256) void quantize_rows_kernel(const float* __restrict__ w, int cols,
                                                            const float* __restrict__ x_scale, int use_h_scale,
                                                            float* __restrict__ scales, int inv_offset,
                                                            uint8_t* __restrict__ out) {
    __shared__ float s_max[4];
    const int n = blockIdx.x;
    const float* row = w + static_cast<size_t>(n) * cols;
    float mx = 0.0F;
    for (int k = threadIdx.x; k < cols; k += blockDim.x) {
        const float v = x_scale != nullptr ? row[k] / x_scale[k] : row[k];
        mx = fmaxf(mx, fabsf(v));
    }
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) {
        s_max[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    const float row_max = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
    const float r = row_max > 0.0F ? kFp8Max / row_max : 1.0F;
    for (int k = threadIdx.x; k < cols; k += blockDim.x) {
        const float v = x_scale != nullptr ? row[k] / x_scale[k] : row[k];
        out[static_cast<size_t>(n) * cols + k] = to_fp8(v * r);
    }
    if (threadIdx.x == 0) {
        const float extra = use_h_scale != 0 ? scales[kPropScaleH] : 1.0F;
        scales[inv_offset + n] = 1.0F / (r * extra);
    }
}

// h_scale = 224 / max relu(z1) over the sample (one block).
__global__ __launch_bounds__(256) void hidden_scale_kernel(const float* __restrict__ z1, size_t count,
                                                           float* __restrict__ scales) {
    __shared__ float s_max[4];
    float mx = 0.0F;
    for (size_t i = threadIdx.x; i < count; i += blockDim.x) {
        mx = fmaxf(mx, z1[i]);
    }
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) {
        s_max[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float m = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
        scales[kPropScaleH] = m > 0.0F ? 0.5F * kFp8Max / m : 1.0F;
    }
}

unsigned blocks_for(size_t count, int threads) {
    return static_cast<unsigned>((count + threads - 1) / threads);
}

template <class TA, class TB>
hipError_t gemm(int M, int N, int K, const TA* a, long long a_rs, long long a_cs, const TB* b, long long b_rs,
                long long b_cs, const float* bias, float* c32, __bf16* c16, int ldc, hipStream_t stream) {
    const dim3 grid(static_cast<unsigned>((N + kGemmTile - 1) / kGemmTile),
                    static_cast<unsigned>((M + kGemmTile - 1) / kGemmTile));
    hipLaunchKernelGGL((gemm_kernel<TA, TB>), grid, dim3(256), 0, stream, M, N, K, a, a_rs, a_cs, b, b_rs, b_cs,
                       bias, c32, c16, ldc);
    return hipGetLastError();
}

hipError_t column_sums(const float* a, int rows, int cols, const float* w, bool relu_a, float* part, float* out,
                       hipStream_t stream) {
    hipLaunchKernelGGL(column_partial_kernel, dim3(blocks_for(cols, 64), kColChunks), dim3(64), 0, stream, a, rows,
                       cols, w, relu_a ? 1 : 0, part);
    hipLaunchKernelGGL(column_total_kernel, dim3(blocks_for(cols, 64)), dim3(64), 0, stream, part, cols, out);
    return hipGetLastError();
}

}  // namespace

size_t proposer_workspace_floats(int rows) {
    // z1, z2, dz2, dz1 (fp32 [rows][256]); h1 (bf16 [rows][256], as floats);
    // s, ds; column partials.
    return static_cast<size_t>(rows) * kPropHidden * 4 + static_cast<size_t>(rows) * kPropHidden / 2 +
           static_cast<size_t>(rows) * 2 + static_cast<size_t>(kColChunks) * kPropHidden + 64;
}

hipError_t launch_proposer_init(float* params, uint64_t seed, hipStream_t stream) {
    hipLaunchKernelGGL(init_kernel, dim3(blocks_for(kPropParams, 256)), dim3(256), 0, stream, params, seed);
    return hipGetLastError();
}

hipError_t launch_proposer_refresh(const float* params, __bf16* w1, __bf16* w2, hipStream_t stream) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(kPropHidden * kPropIn, 256)), dim3(256), 0, stream,
                       params + kOffW1, w1, kPropHidden * kPropIn);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(kPropHidden * kPropHidden, 256)), dim3(256), 0, stream,
                       params + kOffW2, w2, kPropHidden * kPropHidden);
    return hipGetLastError();
}

hipError_t launch_proposer_features(const float* clip, int width, int height, int channels, int patch_time,
                                    const uint32_t* anchors, const uint32_t* candidates, int rows, __bf16* out,
                                    hipStream_t stream) {
    if (rows <= 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(features_kernel, dim3(blocks_for(rows, 64)), dim3(64), 0, stream, clip, width, height,
                       channels, patch_time, anchors, candidates, rows, out);
    return hipGetLastError();
}

hipError_t launch_proposer_scores(const __bf16* x, int rows, const ProposerWeights& wt, int precision,
                                  float* scores, hipStream_t stream) {
    if (rows <= 0) {
        return hipSuccess;
    }
    if (precision == 1) {
        if (wt.fp8.w1 == nullptr || wt.fp8.w2 == nullptr || wt.fp8.scales == nullptr) {
            return hipErrorInvalidValue;
        }
        hipLaunchKernelGGL(scores_kernel<true>, dim3(blocks_for(rows, kPropTile)), dim3(256), 0, stream, x, rows, wt,
                           scores);
    } else {
        hipLaunchKernelGGL(scores_kernel<false>, dim3(blocks_for(rows, kPropTile)), dim3(256), 0, stream, x, rows, wt,
                           scores);
    }
    return hipGetLastError();
}

hipError_t launch_proposer_calibrate(const float* params, const __bf16* w1, const __bf16* x, int rows,
                                     float* workspace, uint8_t* w1_fp8, uint8_t* w2_fp8, float* scales,
                                     hipStream_t stream) {
    if (rows <= 0) {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(input_scale_kernel, dim3(blocks_for(kPropIn, 64)), dim3(64), 0, stream, x, rows, scales);
    hipLaunchKernelGGL(quantize_rows_kernel, dim3(kPropHidden), dim3(256), 0, stream, params + kOffW1, kPropIn,
                       scales + kPropScaleX, 0, scales, kPropScaleW1, w1_fp8);
    // Hidden activations of the sample with the bf16 weights.
    hipError_t e = gemm(rows, kPropHidden, kPropIn, x, kPropIn, 1, w1, 1, kPropIn, params + kOffB1, workspace,
                        static_cast<__bf16*>(nullptr), kPropHidden, stream);
    if (e != hipSuccess) {
        return e;
    }
    hipLaunchKernelGGL(hidden_scale_kernel, dim3(1), dim3(256), 0, stream, workspace,
                       static_cast<size_t>(rows) * kPropHidden, scales);
    hipLaunchKernelGGL(quantize_rows_kernel, dim3(kPropHidden), dim3(256), 0, stream, params + kOffW2, kPropHidden,
                       static_cast<const float*>(nullptr), 1, scales, kPropScaleW2, w2_fp8);
    return hipGetLastError();
}

// Forward, loss and gradients of one batch: grads (flat, kPropParams) and
// loss_groups[g] (per group).  w1, w2: the bf16 copies of params.
hipError_t launch_proposer_gradients(const float* params, const __bf16* w1, const __bf16* w2, const __bf16* x,
                                     const float* targets, int rows, int group, float* workspace, float* grads,
                                     float* loss_groups, hipStream_t stream) {
    const size_t wide = static_cast<size_t>(rows) * kPropHidden;
    float* z1 = workspace;
    float* z2 = z1 + wide;
    float* dz2 = z2 + wide;
    float* dz1 = dz2 + wide;
    __bf16* h1 = reinterpret_cast<__bf16*>(dz1 + wide);
    float* s = dz1 + wide + wide / 2;
    float* ds = s + rows;
    float* part = ds + rows;
    const int groups = rows / group;
    hipError_t e;
    // Forward.
    if ((e = gemm(rows, kPropHidden, kPropIn, x, kPropIn, 1, w1, 1, kPropIn, params + kOffB1, z1, h1,
                  kPropHidden, stream)) != hipSuccess) {
        return e;
    }
    if ((e = gemm(rows, kPropHidden, kPropHidden, h1, kPropHidden, 1, w2, 1, kPropHidden, params + kOffB2, z2,
                  static_cast<__bf16*>(nullptr), kPropHidden, stream)) != hipSuccess) {
        return e;
    }
    hipLaunchKernelGGL(output_kernel, dim3(blocks_for(rows, 4)), dim3(256), 0, stream, z2, params, rows, s);
    hipLaunchKernelGGL(loss_kernel, dim3(blocks_for(groups, 4)), dim3(256), 0, stream, s, targets, groups, group,
                       ds, loss_groups);
    // Output layer.
    if ((e = column_sums(z2, rows, kPropHidden, ds, true, part, grads + kOffW3, stream)) != hipSuccess) {
        return e;
    }
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, stream, ds, rows, grads + kOffB3);
    hipLaunchKernelGGL(output_grad_kernel, dim3(blocks_for(wide, 256)), dim3(256), 0, stream, z2, ds, params, rows,
                       dz2);
    // Layer 2: dW2 = dz2^T h1, db2 = column sums of dz2, dh1 = dz2 W2.
    if ((e = column_sums(dz2, rows, kPropHidden, nullptr, false, part, grads + kOffB2, stream)) != hipSuccess) {
        return e;
    }
    if ((e = gemm(kPropHidden, kPropHidden, rows, dz2, 1, kPropHidden, h1, kPropHidden, 1,
                  static_cast<const float*>(nullptr), grads + kOffW2, static_cast<__bf16*>(nullptr), kPropHidden,
                  stream)) != hipSuccess) {
        return e;
    }
    if ((e = gemm(rows, kPropHidden, kPropHidden, dz2, kPropHidden, 1, w2, kPropHidden, 1,
                  static_cast<const float*>(nullptr), dz1, static_cast<__bf16*>(nullptr), kPropHidden,
                  stream)) != hipSuccess) {
        return e;
    }
    hipLaunchKernelGGL(relu_mask_kernel, dim3(blocks_for(wide, 256)), dim3(256), 0, stream, z1, dz1, wide);
    // Layer 1: dW1 = dz1^T x, db1 = column sums of dz1.
    if ((e = column_sums(dz1, rows, kPropHidden, nullptr, false, part, grads + kOffB1, stream)) != hipSuccess) {
        return e;
    }
    return gemm(kPropHidden, kPropIn, rows, dz1, 1, kPropHidden, x, kPropIn, 1, static_cast<const float*>(nullptr),
                grads + kOffW1, static_cast<__bf16*>(nullptr), kPropIn, stream);
}

hipError_t launch_proposer_adam(float* params, const float* grads, float* m, float* v, float lr, int step,
                                hipStream_t stream) {
    const float beta1 = 0.9F;
    const float beta2 = 0.999F;
    const float c1 = 1.0F - powf(beta1, static_cast<float>(step));
    const float c2 = 1.0F - powf(beta2, static_cast<float>(step));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(kPropParams, 256)), dim3(256), 0, stream, params, grads, m, v,
                       kPropParams, lr, beta1, beta2, 1e-8F, c1, c2);
    return hipGetLastError();
}

}  // namespace vnlbhip
