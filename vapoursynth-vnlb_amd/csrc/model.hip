// Batched cyclic Jacobi eigen solve of the group models, several small
// problems per wavefront.
//
// Semantics follow filter_vnlb_samples and its callees in the reference
// (src/core/core.cpp:184-871, src/linalg/linalg.cpp:153-595) with the
// 16-lane summation order of the CPU restatement.
#include "common.h"
#include "device_util.h"
#include "group.h"
#include "kernels.h"

#include <float.h>

namespace vnlbhip {
namespace {

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

}  // namespace

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

}  // namespace vnlbhip
