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

#include <utility>

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
// kLP = 128 (n up to 128, the dual Grams of up to 128 model rows): one
// problem per block of two wavefronts; the segment reductions then go
// through LDS and the rotation loops stop at the problem's n (uniform in the
// block).
template <int kLP>  // lanes per problem (power of two >= n)
__global__ __launch_bounds__(128) void jacobi_kernel(
    StageDesc d, const int* __restrict__ group_count_dev, int problem_stride,
    const int* __restrict__ group_flags, const int* __restrict__ problem_dims,
    const float* __restrict__ problems, float* __restrict__ eigenvalues,
    float* __restrict__ vectors, int* __restrict__ status) {
    const int problem_count = *group_count_dev * d.models;
    constexpr int kPerWave = kLP >= 64 ? 1 : 64 / kLP;
    constexpr int kStride = kLP + 1;
    extern __shared__ float s_jacobi[];
    __shared__ int s_order[kPerWave][kLP];
    __shared__ float s_reduce[2];

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
    float* m = s_jacobi + static_cast<size_t>(sub) * 2 * kLP * kStride;
    float* v = m + kLP * kStride;
    const float* in = problems + static_cast<size_t>(problem) * problem_stride * problem_stride;
    // Maximum over the problem's lanes (all threads of the block call it).
    const auto problem_max = [&](float value) {
        for (int offset = (kLP < 64 ? kLP : 64) / 2; offset >= 1; offset >>= 1) {
            value = fmaxf(value, __shfl_xor(value, offset, kLP < 64 ? kLP : 64));
        }
        if constexpr (kLP > 64) {
            __syncthreads();
            if ((threadIdx.x & 63) == 0) {
                s_reduce[threadIdx.x >> 6] = value;
            }
            __syncthreads();
            value = fmaxf(s_reduce[0], s_reduce[1]);
        }
        return value;
    };

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
    row_max = problem_max(row_max);
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
        return problem_max(value);
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
        // With one problem per block its n bounds the pairs; with several
        // per wavefront every lane walks all kLP (the barriers are shared).
        const int pairs_n = kPerWave == 1 ? n : kLP;
        for (int p = 0; p + 1 < pairs_n; ++p) {
            for (int q = p + 1; q < pairs_n; ++q) {
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
// parallel-order Jacobi (FAST form)
// ---------------------------------------------------------------------------

// Column pair (j, j ^ k) of the relative row: column x = r ^ j pairs with
// x ^ k and takes the rotation of lane x.  Both lanes of a pair hold the
// same c and opposite sigma, so lane x ^ k's values are c and -sigma of lane
// x: one exchange per column pair instead of two.
template <int kK, int kJ>
__device__ inline void rotate_columns(float (&a)[16], float (&v)[16], const float (&b)[16], float c,
                                      float sigma) {
    constexpr int kHigh = kK >= 8 ? 8 : (kK >= 4 ? 4 : (kK >= 2 ? 2 : 1));
    if constexpr ((kJ & kHigh) == 0) {
        constexpr int kJJ = kJ ^ kK;
        const float c0 = xor16<kJ>(c);
        const float s0 = xor16<kJ>(sigma);
        const float c1 = c0;
        const float s1 = -s0;
        const float v0 = v[kJ];
        const float v1 = v[kJJ];
        a[kJ] = fmaf(s0, b[kJJ], c0 * b[kJ]);
        a[kJJ] = fmaf(s1, b[kJ], c1 * b[kJJ]);
        v[kJ] = fmaf(s0, v1, c0 * v0);
        v[kJJ] = fmaf(s1, v0, c1 * v1);
    }
}

template <int kK, int... kJ>
__device__ inline void rotate_rows_and_columns(float (&a)[16], float (&v)[16], float c, float sigma,
                                               std::integer_sequence<int, kJ...>) {
    // Rows: B[r][.] = c A[r][.] + sigma A[r ^ k][.]; the partner's
    // A[r ^ k][r ^ j] is its a[j ^ k].
    float b[16];
    ((b[kJ] = fmaf(sigma, xor16<kK>(a[kJ ^ kK]), c * a[kJ])), ...);
    (rotate_columns<kK, kJ>(a, v, b, c, sigma), ...);
}

// One rotation round of the parallel ordering for problems of at most 16
// rows: in round k every index r is paired with r ^ k (rounds 1..15 cover
// each pair exactly once per sweep), and the eight disjoint rotations of a
// problem are applied together as A <- J^T A J, V <- V J.  Lane r of a
// 16-lane row holds row r of A and V in relative order, a[j] = A[r][r ^ j],
// so the diagonal is a[0], the pivot A[r][r ^ k] is a[k] and every register
// index and lane exchange is a compile-time constant.  The rotation is that
// of the cyclic solver (jacobi_rotate, linalg.cpp:153-202) with hardware
// reciprocal and square-root approximations; both lanes of a pair compute it
// from the lower index's values, so they agree bit for bit.  Rounds in which
// no problem of the wavefront rotates are skipped.
template <int kK>
__device__ inline void parallel_round(float (&a)[16], float (&v)[16], int r, bool active,
                                      float threshold) {
    constexpr int kHigh = kK >= 8 ? 8 : (kK >= 4 ? 4 : (kK >= 2 ? 2 : 1));
    const bool is_p = (r & kHigh) == 0;  // r is the lower index of its pair
    const float own_diag = a[0];
    const float own_off = a[kK];
    const float partner_diag = xor16<kK>(own_diag);
    const float partner_off = xor16<kK>(own_off);
    const float app = is_p ? own_diag : partner_diag;
    const float aqq = is_p ? partner_diag : own_diag;
    const float apq = is_p ? own_off : partner_off;
    const bool rotate = active && fabsf(apq) > threshold;
    if (!__any(rotate)) {
        return;
    }
    // Branch-free: every lane evaluates the rotation (apq is replaced by 1
    // where no rotation happens), then selects.
    const float safe_apq = rotate ? apq : 1.0F;
    const float tau = (aqq - app) * __builtin_amdgcn_rcpf(2.0F * safe_apq);
    const float tau_abs = fabsf(tau);
    const float t_abs = __builtin_amdgcn_rcpf(tau_abs + __builtin_amdgcn_sqrtf(1.0F + (tau * tau)));
    const float t = tau < 0.0F ? -t_abs : t_abs;
    const float c_rot = __builtin_amdgcn_rsqf(1.0F + (t * t));
    const float s_rot = t * c_rot;
    const float cc = c_rot * c_rot;
    const float ss = s_rot * s_rot;
    const float two_cs = 2.0F * (c_rot * s_rot);
    const float new_pp = ((cc * app) - (two_cs * apq)) + (ss * aqq);
    const float new_qq = ((ss * app) + (two_cs * apq)) + (cc * aqq);
    const float c = rotate ? c_rot : 1.0F;
    // -s for the lower index, +s for the upper one.
    const float sigma = rotate ? (is_p ? -s_rot : s_rot) : 0.0F;
    const float new_diag = is_p ? new_pp : new_qq;
    rotate_rows_and_columns<kK>(a, v, c, sigma, std::make_integer_sequence<int, 16>{});
    a[0] = rotate ? new_diag : a[0];
    a[kK] = rotate ? 0.0F : a[kK];
    // Rounds are scheduled one after the other (bounded register use).
    __builtin_amdgcn_sched_barrier(0);
}

template <int... kKs>
__device__ inline void parallel_sweep(float (&a)[16], float (&v)[16], int r, bool active,
                                      float threshold) {
    (parallel_round<kKs>(a, v, r, active, threshold), ...);
}

// Same sweep structure and canonical vectors as jacobi_kernel, with the
// parallel ordering above: 16 lanes per problem, four per wavefront.  The
// sweeps stop once every off-diagonal entry is at most 2^-14 of
// max(1, max |A|) (the exact form stops at 64 float epsilons, about 2^-17):
// the sweeps this saves (7 % of the solver's time on 16-row models) only
// refine eigenpairs already within the fast form's tolerance (eigenvalues
// within 1e-3 of the largest, checked against the fp64 oracle).
__global__ __launch_bounds__(64) void jacobi_parallel_kernel(
    StageDesc d, const int* __restrict__ group_count_dev, int problem_stride,
    const int* __restrict__ group_flags, const int* __restrict__ problem_dims,
    const float* __restrict__ problems, float* __restrict__ eigenvalues, float* __restrict__ vectors,
    int* __restrict__ status, float refine_gap, int* __restrict__ refine_list,
    int* __restrict__ refine_count) {
    const int problem_count = *group_count_dev * d.models;
    constexpr int kPerWave = 4;
    constexpr int kStride = 17;
    __shared__ float s_v[kPerWave][16 * kStride];
    __shared__ float s_diag[kPerWave][16];
    __shared__ int s_order[kPerWave][16];

    const int sub = threadIdx.x / 16;
    const int r = threadIdx.x % 16;
    const int problem = blockIdx.x * kPerWave + sub;
    const int group = problem / d.models;
    bool valid = problem < problem_count;
    int n = 0;
    if (valid) {
        valid = (group_flags[group] & kFlagModelled) != 0;
        n = problem_dims[group];
    }
    const float* in = problems + static_cast<size_t>(problem) * problem_stride * problem_stride;
    float a[16];
    float v[16];
    float row_max = 0.0F;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int col = r ^ j;
        float value = 0.0F;
        if (valid && r < n && col < n) {
            value = in[r * problem_stride + col];
            row_max = fmaxf(row_max, fabsf(value));
        }
        a[j] = value;
        v[j] = j == 0 ? 1.0F : 0.0F;
    }
    row_max = fmaxf(row_max, xor16<8>(row_max));
    row_max = fmaxf(row_max, xor16<4>(row_max));
    row_max = fmaxf(row_max, xor16<2>(row_max));
    row_max = fmaxf(row_max, xor16<1>(row_max));
    const float matrix_scale = row_max > 1.0F ? row_max : 1.0F;
    const float threshold = 0x1p-14F * matrix_scale;

    const auto offdiag_max = [&]() {
        float value = 0.0F;
#pragma unroll
        for (int j = 1; j < 16; ++j) {
            value = fmaxf(value, fabsf(a[j]));
        }
        value = fmaxf(value, xor16<8>(value));
        value = fmaxf(value, xor16<4>(value));
        value = fmaxf(value, xor16<2>(value));
        value = fmaxf(value, xor16<1>(value));
        return value;
    };

    bool done = !valid || n <= 1;
    bool converged = true;
    int sweeps = 0;
#pragma nounroll
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
        parallel_sweep<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15>(a, v, r, !done, threshold);
    }
    const float final_off = offdiag_max();
    if (!done) {
        sweeps = d.max_sweeps;
        converged = final_off <= 8.0F * threshold;
    }

    // V[r][r ^ j] = v[j] into LDS; the diagonal in row order.
    float* vs = s_v[sub];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        vs[r * kStride + (r ^ j)] = v[j];
    }
    s_diag[sub][r] = a[0];
    __syncthreads();
    if (valid && r < n) {
        const float dr = s_diag[sub][r];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const float dj = s_diag[sub][j];
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
            const float value = vs[i * kStride + col];
            norm_squared = norm_squared + value * value;
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
            if (vs[pivot * kStride + col] < 0.0F) {
                scale = -scale;
            }
        }
        for (int i = 0; i < n; ++i) {
            const float value = vs[i * kStride + col];
            vec_out[i] = rescale ? value * scale : value;
        }
        eigenvalues[static_cast<size_t>(problem) * d.rank + r] = s_diag[sub][col];
    }
    if (valid && r == 0) {
        status[problem] = sweeps | (converged ? 0 : (1 << 16));
        // A rank cutoff inside the spectrum whose gap is small against the
        // largest eigenvalue: the kept subspace is re-solved in fp64.
        const int rank = d.rank;
        if (rank < n && rank > 0 && refine_list != nullptr) {
            const float top = fabsf(s_diag[sub][s_order[sub][0]]);
            const float gap = s_diag[sub][s_order[sub][rank - 1]] - s_diag[sub][s_order[sub][rank]];
            if (!(gap >= refine_gap * fmaxf(top, FLT_MIN))) {
                refine_list[atomicAdd(refine_count, 1)] = problem;
            }
        }
    }
}

// fp64 re-solve of the flagged problems (cyclic Jacobi in double, the
// reference's solver at its double instantiation, linalg.cpp:447-550: same
// rotation, threshold 64 eps max(1, max|A|) with the double eps, stopping
// rule, ordering and canonical vectors), results rounded to float.  One
// problem per wavefront at a time, lane r owning row r; the list length is
// read on the device, so no host round trip is needed.
__global__ __launch_bounds__(64) void jacobi_refine_kernel(
    StageDesc d, int problem_stride, const int* __restrict__ problem_dims,
    const float* __restrict__ problems, const int* __restrict__ refine_list,
    const int* __restrict__ refine_count, float* __restrict__ eigenvalues,
    float* __restrict__ vectors, int* __restrict__ refined_total) {
    constexpr int kN = 16;
    constexpr int kStride = kN + 1;
    __shared__ double s_m[kN * kStride];
    __shared__ double s_v[kN * kStride];
    __shared__ int s_order[kN];
    const int r = threadIdx.x;
    const int total = *refine_count;
    if (blockIdx.x == 0 && r == 0 && refined_total != nullptr) {
        atomicAdd(refined_total, total);
    }
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const int problem = refine_list[item];
        const int group = problem / d.models;
        const int n = problem_dims[group];
        const float* in = problems + static_cast<size_t>(problem) * problem_stride * problem_stride;
        __syncthreads();
        double row_max = 0.0;
        if (r < kN) {
            for (int c = 0; c < kN; ++c) {
                double value = 0.0;
                if (r < n && c < n) {
                    value = static_cast<double>(in[r * problem_stride + c]);
                    row_max = fmax(row_max, fabs(value));
                }
                s_m[r * kStride + c] = value;
                s_v[r * kStride + c] = (r == c) ? 1.0 : 0.0;
            }
        }
        row_max = wave_max_double(row_max);
        const double threshold = (64.0 * DBL_EPSILON) * (row_max > 1.0 ? row_max : 1.0);
        __syncthreads();
        for (int sweep = 0; sweep < d.max_sweeps; ++sweep) {
            double off = 0.0;
            if (r < n) {
                for (int c = r + 1; c < n; ++c) {
                    off = fmax(off, fabs(s_m[r * kStride + c]));
                }
            }
            off = wave_max_double(off);
            if (off <= threshold) {
                break;
            }
            for (int p = 0; p + 1 < n; ++p) {
                for (int q = p + 1; q < n; ++q) {
                    const double apq = s_m[p * kStride + q];
                    const bool rotate = fabs(apq) > threshold;  // uniform
                    double c = 1.0;
                    double sn = 0.0;
                    double new_pp = 0.0;
                    double new_qq = 0.0;
                    double new_rp = 0.0;
                    double new_rq = 0.0;
                    double new_vp = 0.0;
                    double new_vq = 0.0;
                    if (rotate) {
                        const double app = s_m[p * kStride + p];
                        const double aqq = s_m[q * kStride + q];
                        const double tau = (aqq - app) / (2.0 * apq);
                        const double t_abs = 1.0 / (fabs(tau) + sqrt(1.0 + tau * tau));
                        const double t = tau < 0.0 ? -t_abs : t_abs;
                        c = 1.0 / sqrt(1.0 + t * t);
                        sn = t * c;
                        const double cc = c * c;
                        const double ss = sn * sn;
                        const double two_cs = 2.0 * (c * sn);
                        new_pp = ((cc * app) - (two_cs * apq)) + (ss * aqq);
                        new_qq = ((ss * app) + (two_cs * apq)) + (cc * aqq);
                        if (r < n && r != p && r != q) {
                            const double arp = s_m[r * kStride + p];
                            const double arq = s_m[r * kStride + q];
                            new_rp = (c * arp) - (sn * arq);
                            new_rq = (sn * arp) + (c * arq);
                        }
                        if (r < n) {
                            const double vrp = s_v[r * kStride + p];
                            const double vrq = s_v[r * kStride + q];
                            new_vp = (c * vrp) - (sn * vrq);
                            new_vq = (sn * vrp) + (c * vrq);
                        }
                    }
                    __syncthreads();
                    if (rotate) {
                        if (r < n && r != p && r != q) {
                            s_m[r * kStride + p] = new_rp;
                            s_m[p * kStride + r] = new_rp;
                            s_m[r * kStride + q] = new_rq;
                            s_m[q * kStride + r] = new_rq;
                        }
                        if (r == 0) {
                            s_m[p * kStride + p] = new_pp;
                            s_m[q * kStride + q] = new_qq;
                            s_m[p * kStride + q] = 0.0;
                            s_m[q * kStride + p] = 0.0;
                        }
                        if (r < n) {
                            s_v[r * kStride + p] = new_vp;
                            s_v[r * kStride + q] = new_vq;
                        }
                    }
                    __syncthreads();
                }
            }
        }
        if (r < n) {
            const double dr = s_m[r * kStride + r];
            int rank = 0;
            for (int j = 0; j < n; ++j) {
                const double dj = s_m[j * kStride + j];
                rank += (dj > dr || (j < r && !(dr > dj))) ? 1 : 0;
            }
            s_order[rank] = r;
        }
        __syncthreads();
        if (r < d.rank && r < n) {
            const int col = s_order[r];
            float* vec_out = vectors + (static_cast<size_t>(problem) * d.rank + r) * problem_stride;
            double norm_squared = 0.0;
            double max_abs = 0.0;
            int pivot = 0;
            for (int i = 0; i < n; ++i) {
                const double value = s_v[i * kStride + col];
                norm_squared += value * value;
                if (fabs(value) > max_abs) {
                    max_abs = fabs(value);
                    pivot = i;
                }
            }
            double scale = 1.0;
            const bool rescale = norm_squared > DBL_MIN;
            if (rescale) {
                scale = 1.0 / sqrt(norm_squared);
                if (s_v[pivot * kStride + col] < 0.0) {
                    scale = -scale;
                }
            }
            for (int i = 0; i < n; ++i) {
                const double value = s_v[i * kStride + col];
                vec_out[i] = static_cast<float>(rescale ? value * scale : value);
            }
            eigenvalues[static_cast<size_t>(problem) * d.rank + r] = static_cast<float>(s_m[col * kStride + col]);
        }
    }
}

}  // namespace

hipError_t launch_jacobi(const StageDesc& d, const int* group_count_dev, int problem_capacity,
                         int problem_stride, const int* group_flags, const int* problem_dims,
                         const float* problems, float* eigenvalues, float* vectors,
                         int* status, hipStream_t stream) {
    if (problem_capacity <= 0) {
        return hipSuccess;
    }
    const int n = problem_stride;
#define VNLBHIP_JACOBI(LP)                                                                      \
    do {                                                                                        \
        constexpr int kPer = LP >= 64 ? 1 : 64 / LP;                                            \
        const size_t shared = static_cast<size_t>(kPer) * 2 * LP * (LP + 1) * sizeof(float);    \
        hipLaunchKernelGGL(jacobi_kernel<LP>,                                                   \
                           dim3(static_cast<unsigned>((problem_capacity + kPer - 1) / kPer)),   \
                           dim3(LP > 64 ? LP : 64), shared, stream, d, group_count_dev,         \
                           problem_stride, group_flags, problem_dims, problems, eigenvalues,    \
                           vectors, status);                                                    \
    } while (0)
    if (n > 128) {
        return hipErrorInvalidValue;
    }
    if (n <= 4) {
        VNLBHIP_JACOBI(4);
    } else if (n <= 8) {
        VNLBHIP_JACOBI(8);
    } else if (n <= 16) {
        VNLBHIP_JACOBI(16);
    } else if (n <= 32) {
        VNLBHIP_JACOBI(32);
    } else if (n <= 64) {
        VNLBHIP_JACOBI(64);
    } else {
        VNLBHIP_JACOBI(128);
    }
#undef VNLBHIP_JACOBI
    return hipGetLastError();
}

}  // namespace vnlbhip

namespace vnlbhip {

bool jacobi_parallel_supported(int problem_stride) {
    return problem_stride <= 16;
}

hipError_t launch_jacobi_parallel(const StageDesc& d, const int* group_count_dev, int problem_capacity,
                                  int problem_stride, const int* group_flags, const int* problem_dims,
                                  const float* problems, float* eigenvalues, float* vectors,
                                  int* status, float refine_gap, int* refine_list, int* refine_count,
                                  int* refined_total, hipStream_t stream) {
    if (problem_capacity <= 0) {
        return hipSuccess;
    }
    if (!jacobi_parallel_supported(problem_stride)) {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(jacobi_parallel_kernel, dim3(static_cast<unsigned>((problem_capacity + 3) / 4)),
                       dim3(64), 0, stream, d, group_count_dev, problem_stride, group_flags,
                       problem_dims, problems, eigenvalues, vectors, status, refine_gap, refine_list,
                       refine_count);
    if (refine_list != nullptr && d.rank > 0) {
        const int blocks = problem_capacity < 2048 ? problem_capacity : 2048;
        hipLaunchKernelGGL(jacobi_refine_kernel, dim3(static_cast<unsigned>(blocks)), dim3(64), 0, stream, d,
                           problem_stride, problem_dims, problems, refine_list, refine_count, eigenvalues,
                           vectors, refined_total);
    }
    return hipGetLastError();
}

}  // namespace vnlbhip
