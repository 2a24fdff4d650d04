// VNLB CPU oracle (test infrastructure only; see vnlb_oracle.h).
//
// A restatement, in one translation unit, of the reference CPU hot path:
//   src/core/core.cpp            stage driver, matching, selection, model,
//                                weights, contributions, paste mask
//   src/core/distance_highway.hpp  row SSD
//   src/linalg/linalg.cpp        mean, centering, Gram, covariance, Jacobi,
//                                dual basis map
//   src/linalg/kernels_highway.hpp dot / add_scaled / scale row kernels
//   src/aggregate/aggregate.cpp  contribution layout
//   src/vs/plugin.cpp:966-1045   vnlb.Aggregate normalisation
//
// Floating-point order.  The reference vectorises with Google Highway, which
// is not available here, so its lane-level summation order is not pinned by
// any reference test.  This oracle fixes one documented order: a 16-lane
// target (what a -march=native build selects on an AVX-512 host) whose
// ReduceSum is a halving tree, MulAdd is a fused multiply-add, and scalar
// tails are separately rounded multiply then add (GCC in ISO mode does not
// contract).  The HIP path follows exactly the same order.
//
// Transcendentals (log, log1p, exp) are evaluated in double and rounded to
// float, on both sides, so that they agree bit for bit; sqrt and division are
// IEEE correctly rounded.
#include "vnlb_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

std::mutex g_error_lock;
std::string g_error;

void set_error(const std::string& message) {
    std::lock_guard<std::mutex> lock(g_error_lock);
    g_error = message;
}

void require(bool condition, const char* message) {
    if (!condition) {
        throw std::invalid_argument(message);
    }
}

constexpr int kLanes = 16;

// ---------------------------------------------------------------------------
// Highway kernel model (src/linalg/kernels_highway.hpp,
// src/core/distance_highway.hpp) at 16 lanes.
// ---------------------------------------------------------------------------

template <class T> T reduce_lanes(T* v) {
    for (int width = kLanes / 2; width >= 1; width >>= 1) {
        for (int lane = 0; lane < width; ++lane) {
            v[lane] = v[lane] + v[lane + width];
        }
    }
    return v[0];
}

// dot_contiguous_highway (kernels_highway.hpp:12-44).  The count == lanes and
// count == 2*lanes branches equal the generic loop because fma(a, b, 0) == a*b.
template <class T> T dot(const T* a, const T* b, std::size_t n) {
    T acc[kLanes];
    for (T& v : acc) {
        v = T{};
    }
    std::size_t i = 0;
    for (; i + kLanes <= n; i += kLanes) {
        for (int lane = 0; lane < kLanes; ++lane) {
            acc[lane] = std::fma(a[i + lane], b[i + lane], acc[lane]);
        }
    }
    T total = reduce_lanes(acc);
    for (; i < n; ++i) {
        const T product = a[i] * b[i];
        total = total + product;
    }
    return total;
}

// add_scaled_contiguous_highway (kernels_highway.hpp:160-189).
template <class T>
void add_scaled(T* out, const T* in, T scale, std::size_t n) {
    std::size_t i = 0;
    for (; i + kLanes <= n; i += kLanes) {
        for (int lane = 0; lane < kLanes; ++lane) {
            out[i + lane] = std::fma(scale, in[i + lane], out[i + lane]);
        }
    }
    for (; i < n; ++i) {
        const T product = scale * in[i];
        out[i] = out[i] + product;
    }
}

// scale_contiguous_highway (kernels_highway.hpp:259-283).
template <class T> void scale_row(T* row, T scale, std::size_t n) {
    for (std::size_t i = 0; i < n; ++i) {
        row[i] = row[i] * scale;
    }
}

// add_squared_row_distance_highway (distance_highway.hpp:11-58).
float add_row_distance(float distance, const float* a, const float* b, int n) {
    float acc[kLanes];
    for (float& v : acc) {
        v = 0.0F;
    }
    int i = 0;
    for (; i + kLanes <= n; i += kLanes) {
        for (int lane = 0; lane < kLanes; ++lane) {
            const float diff = a[i + lane] - b[i + lane];
            acc[lane] = std::fma(diff, diff, acc[lane]);
        }
    }
    distance = distance + reduce_lanes(acc);
    for (; i < n; ++i) {
        const float diff = a[i] - b[i];
        const float square = diff * diff;
        distance = distance + square;
    }
    return distance;
}

template <class T> T t_log(T x) {
    if constexpr (std::is_same_v<T, float>) {
        return static_cast<float>(std::log(static_cast<double>(x)));
    } else {
        return std::log(x);
    }
}
template <class T> T t_log1p(T x) {
    if constexpr (std::is_same_v<T, float>) {
        return static_cast<float>(std::log1p(static_cast<double>(x)));
    } else {
        return std::log1p(x);
    }
}
template <class T> T t_exp(T x) {
    if constexpr (std::is_same_v<T, float>) {
        return static_cast<float>(std::exp(static_cast<double>(x)));
    } else {
        return std::exp(x);
    }
}

// ---------------------------------------------------------------------------
// Dense linear algebra (src/linalg/linalg.cpp).
// ---------------------------------------------------------------------------

// compute_mean (linalg.cpp:290-311): sequential row sum, times 1/rows.
template <class T>
void compute_mean(const T* samples, int rows, int cols, T* mean) {
    for (int c = 0; c < cols; ++c) {
        mean[c] = T{};
    }
    for (int r = 0; r < rows; ++r) {
        const T* row = samples + static_cast<std::size_t>(r) * cols;
        for (int c = 0; c < cols; ++c) {
            mean[c] = mean[c] + row[c];
        }
    }
    const T inv_count = T{1} / static_cast<T>(rows);
    for (int c = 0; c < cols; ++c) {
        mean[c] = mean[c] * inv_count;
    }
}

// compute_gram_centered (linalg.cpp:326-354), non-BLAS branch.
template <class T>
void compute_gram(const T* centered, int rows, int cols, T scale, T* gram) {
    for (int r = 0; r < rows; ++r) {
        const T* left = centered + static_cast<std::size_t>(r) * cols;
        for (int c = 0; c <= r; ++c) {
            const T value =
                scale * dot(left, centered + static_cast<std::size_t>(c) * cols,
                            static_cast<std::size_t>(cols));
            gram[r * rows + c] = value;
            gram[c * rows + r] = value;
        }
    }
}

// compute_covariance_centered (linalg.cpp:377-414), non-BLAS branch.
template <class T>
void compute_covariance(const T* centered, int rows, int cols, T scale,
                        T* cov) {
    std::fill(cov, cov + static_cast<std::size_t>(cols) * cols, T{});
    for (int s = 0; s < rows; ++s) {
        const T* row = centered + static_cast<std::size_t>(s) * cols;
        for (int c0 = 0; c0 < cols; ++c0) {
            const T left = row[c0];
            add_scaled(cov + static_cast<std::size_t>(c0) * cols + c0, row + c0,
                       left, static_cast<std::size_t>(cols - c0));
        }
    }
    for (int r = 0; r < cols; ++r) {
        for (int c = r; c < cols; ++c) {
            const T value = cov[r * cols + c] * scale;
            cov[r * cols + c] = value;
            cov[c * cols + r] = value;
        }
    }
}

// jacobi_rotate (linalg.cpp:153-202).
template <class T>
void jacobi_rotate(T* m, T* v, int n, int p, int q) {
    const T app = m[p * n + p];
    const T aqq = m[q * n + q];
    const T apq = m[p * n + q];
    if (apq == T{}) {
        return;
    }
    const T tau = (aqq - app) / (static_cast<T>(2) * apq);
    const T tau_abs = std::abs(tau);
    const T t_abs = T{1} / (tau_abs + std::sqrt(T{1} + (tau * tau)));
    const T t = tau < T{} ? -t_abs : t_abs;
    const T c = T{1} / std::sqrt(T{1} + (t * t));
    const T s = t * c;
    const T cc = c * c;
    const T ss = s * s;
    const T cs = c * s;
    for (int r = 0; r < n; ++r) {
        if (r == p || r == q) {
            continue;
        }
        const T arp = m[r * n + p];
        const T arq = m[r * n + q];
        const T new_rp = (c * arp) - (s * arq);
        const T new_rq = (s * arp) + (c * arq);
        m[r * n + p] = new_rp;
        m[p * n + r] = new_rp;
        m[r * n + q] = new_rq;
        m[q * n + r] = new_rq;
    }
    const T two_cs = static_cast<T>(2) * cs;
    m[p * n + p] = ((cc * app) - (two_cs * apq)) + (ss * aqq);
    m[q * n + q] = ((ss * app) + (two_cs * apq)) + (cc * aqq);
    m[p * n + q] = T{};
    m[q * n + p] = T{};
    for (int r = 0; r < n; ++r) {
        const T vrp = v[r * n + p];
        const T vrq = v[r * n + q];
        v[r * n + p] = (c * vrp) - (s * vrq);
        v[r * n + q] = (s * vrp) + (c * vrq);
    }
}

// canonicalize_vector_row (linalg.cpp:204-233).
template <class T> void canonicalize(T* row, int n) {
    T norm_squared{};
    T max_abs{};
    int pivot = 0;
    for (int c = 0; c < n; ++c) {
        const T value = row[c];
        const T square = value * value;
        norm_squared = norm_squared + square;
        const T value_abs = std::abs(value);
        if (value_abs > max_abs) {
            max_abs = value_abs;
            pivot = c;
        }
    }
    if (norm_squared <= std::numeric_limits<T>::min()) {
        return;
    }
    T scale = T{1} / std::sqrt(norm_squared);
    if (row[pivot] < T{}) {
        scale = -scale;
    }
    for (int c = 0; c < n; ++c) {
        row[c] = row[c] * scale;
    }
}

struct EigenStatus {
    int sweeps = 0;
    bool converged = true;
};

// JacobiEigenOptions::max_sweeps (linalg.hpp:153-158): 64 unless
// vo_set_max_sweeps changed it.
int g_max_sweeps = 64;

// topk_symmetric_eigen_jacobi (linalg.cpp:447-519) with the default
// JacobiEigenOptions (linalg.hpp:153-158): g_max_sweeps sweeps, relative
// tolerance 64*eps, absolute tolerance 0.  values[rank], vectors[rank][n].
template <class T>
EigenStatus topk_eigen(const T* sym, int n, int rank, T* values, T* vectors,
                       std::vector<T>& m, std::vector<T>& v,
                       std::vector<int>& order) {
    m.assign(sym, sym + static_cast<std::size_t>(n) * n);
    v.assign(static_cast<std::size_t>(n) * n, T{});
    for (int i = 0; i < n; ++i) {
        v[i * n + i] = T{1};
    }
    T max_abs{};
    for (int i = 0; i < n * n; ++i) {
        max_abs = std::max(max_abs, std::abs(m[i]));
    }
    const T matrix_scale = std::max(T{1}, max_abs);
    const T threshold =
        std::max(T{}, static_cast<T>(64) * std::numeric_limits<T>::epsilon() *
                          matrix_scale);
    const auto max_offdiag = [&]() {
        T value{};
        for (int r = 0; r < n; ++r) {
            for (int c = r + 1; c < n; ++c) {
                value = std::max(value, std::abs(m[r * n + c]));
            }
        }
        return value;
    };

    EigenStatus status;
    if (n > 1) {
        status.converged = false;
        for (int sweep = 0; sweep < g_max_sweeps; ++sweep) {
            status.sweeps = sweep + 1;
            if (max_offdiag() <= threshold) {
                status.converged = true;
                break;
            }
            for (int p = 0; p + 1 < n; ++p) {
                for (int q = p + 1; q < n; ++q) {
                    if (std::abs(m[p * n + q]) > threshold) {
                        jacobi_rotate(m.data(), v.data(), n, p, q);
                    }
                }
            }
        }
        if (!status.converged && max_offdiag() <= static_cast<T>(8) * threshold) {
            status.converged = true;
        }
    }

    order.resize(static_cast<std::size_t>(n));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int l, int r) {
        return m[l * n + l] > m[r * n + r];
    });
    for (int out = 0; out < rank; ++out) {
        const int col = order[static_cast<std::size_t>(out)];
        values[out] = m[col * n + col];
        for (int r = 0; r < n; ++r) {
            vectors[out * n + r] = v[r * n + col];
        }
        canonicalize(vectors + static_cast<std::size_t>(out) * n, n);
    }
    return status;
}

// normalize_or_zero (linalg.cpp:235-246).
template <class T> void normalize_or_zero(T* row, int n, T floor) {
    const T norm_squared = dot(row, row, static_cast<std::size_t>(n));
    if (norm_squared <= floor) {
        std::fill(row, row + n, T{});
        return;
    }
    const T scale = T{1} / std::sqrt(norm_squared);
    scale_row(row, scale, static_cast<std::size_t>(n));
}

// map_dual_eigenvectors_to_basis (linalg.cpp:552-595); floor = 128*eps.
template <class T>
void dual_basis(const T* centered, int rows, int cols, const T* values,
                const T* vectors, int rank, T* basis) {
    const T floor = static_cast<T>(128) * std::numeric_limits<T>::epsilon();
    for (int b = 0; b < rank; ++b) {
        T* row = basis + static_cast<std::size_t>(b) * cols;
        std::fill(row, row + cols, T{});
        const T eigenvalue = values[b];
        if (eigenvalue <= floor) {
            continue;
        }
        for (int s = 0; s < rows; ++s) {
            add_scaled(row, centered + static_cast<std::size_t>(s) * cols,
                       vectors[b * rows + s], static_cast<std::size_t>(cols));
        }
        const T inv_sqrt = T{1} / std::sqrt(eigenvalue);
        scale_row(row, inv_sqrt, static_cast<std::size_t>(cols));
        normalize_or_zero(row, cols, floor);
    }
}

// ---------------------------------------------------------------------------
// Stage definitions (src/core/core.cpp).
// ---------------------------------------------------------------------------

// VideoGeometry (src/core/video.hpp:19-47): `frames` frames in memory, the
// window [first_frame, first_frame + frames) of a clip of total_frames (0:
// the window is the clip).  Frame arguments below are absolute indices.
struct Geometry {
    int width = 0;
    int height = 0;
    int frames = 0;
    int channels = 0;
    int first_frame = 0;
    int total_frames = 0;
    int source_frames() const { return total_frames > 0 ? total_frames : frames; }
    std::size_t local(int frame) const { return static_cast<std::size_t>(frame - first_frame); }
};

// soft_similar_cap / model_similar_cap (core.cpp:139-172).
int capped_similar(int similar, float factor) {
    if (factor == 0.0F) {
        return std::numeric_limits<int>::max();
    }
    const float max_factor = static_cast<float>(std::numeric_limits<int>::max()) /
                             static_cast<float>(similar);
    if (factor >= max_factor) {
        return std::numeric_limits<int>::max();
    }
    const float cap = std::ceil(static_cast<float>(similar) * factor);
    return std::max(similar, static_cast<int>(cap));
}

int normalized_proc_step(const vo_params& p) {  // core.cpp:40-48
    if (p.proc_step > 0) {
        return p.proc_step;
    }
    return p.patch_size <= 1 ? 1 : p.patch_size;
}

long long raw_candidate_count(const vo_params& p) {
    return static_cast<long long>(p.search_window) * p.search_window *
           (static_cast<long long>(p.search_bwd) + p.search_fwd + 1);
}

int retained_cap(const vo_params& p) {  // checked_retained_group_count
    const long long raw = raw_candidate_count(p);
    const long long soft = capped_similar(p.similar, p.similar_cap_factor);
    return static_cast<int>(std::min(soft, raw));
}

// validate_stage_parameters / validate_stage_configuration (core.cpp:1948-2024).
void validate(const vo_params& p, const Geometry& g) {
    require(std::isfinite(p.sigma) && p.sigma > 0.0F,
            "sigma must be finite and positive");
    require(p.patch_size > 0, "block_size must be positive");
    require(p.patch_time > 0, "patch_time must be positive");
    require(p.search_window > 0 && p.search_window % 2 == 1,
            "bm_range must describe a positive odd search window");
    require(p.search_bwd >= 0 && p.search_fwd >= 0,
            "search_bwd and search_fwd must be non-negative");
    require(p.similar > 0, "group_size must be positive");
    require(p.rank >= 0, "rank must be non-negative");
    require(std::isfinite(p.similar_cap_factor) &&
                (p.similar_cap_factor == 0.0F || p.similar_cap_factor >= 1.0F),
            "cap_factor must be zero or at least 1.0");
    require(std::isfinite(p.model_cap_factor) &&
                (p.model_cap_factor == 0.0F || p.model_cap_factor >= 1.0F),
            "model_cap_factor must be zero or at least 1.0");
    require(std::isfinite(p.beta) && p.beta > 0.0F,
            "beta must be finite and positive");
    require(std::isfinite(p.tau) && p.tau >= 0.0F,
            "tau must be finite and non-negative");
    require(std::isfinite(p.variance_threshold),
            "variance_threshold must be finite");
    require(std::isfinite(p.sigma_basic) && p.sigma_basic >= 0.0F,
            "sigma_basic must be finite and non-negative");
    require(std::isfinite(p.gamma) && p.gamma > 0.0F,
            "gamma must be finite and positive");
    require(std::isfinite(p.weight_alpha) && p.weight_alpha >= 0.0F,
            "weight_alpha must be finite and non-negative");
    require(std::isfinite(p.weight_beta) && p.weight_beta >= 0.0F,
            "weight_beta must be finite and non-negative");
    require(std::isfinite(p.weight_gamma) && p.weight_gamma >= 0.0F,
            "weight_gamma must be finite and non-negative");
    require(std::isfinite(p.weight_epsilon) && p.weight_epsilon > 0.0F,
            "weight_epsilon must be finite and positive");
    require(std::isfinite(p.membership_noise_floor) &&
                p.membership_noise_floor > 0.0F,
            "membership_noise_floor must be finite and positive");
    require(p.proc_step >= 0, "block_step must be non-negative; use 0 for auto");
    require(g.width > 0 && g.height > 0 && g.frames > 0 && g.channels > 0,
            "video geometry must be non-empty");
    require(g.first_frame >= 0, "first frame must be non-negative");
    require(static_cast<long long>(g.source_frames()) >= static_cast<long long>(g.frames) + g.first_frame,
            "local frame window must fit inside the source frame count");
    require(p.patch_size <= g.width && p.patch_size <= g.height,
            "block_size must fit inside the frame");
    require(p.patch_time <= g.source_frames(), "patch_time must fit inside the clip");
    const long long estimator_dim =
        static_cast<long long>(p.patch_size) * p.patch_size * p.patch_time *
        ((p.couple_channels != 0) ? g.channels : 1);
    require(estimator_dim <= std::numeric_limits<int>::max(),
            "estimator dimension overflows int");
    require(p.rank <= estimator_dim,
            "rank must not exceed the effective patch dimension");
    require(p.similar >= std::min<long long>(estimator_dim, p.rank),
            "group_size must cover the effective VNLB rank");
}

// compute_temporal_range (core.cpp:873-896).
void temporal_range(const vo_params& p, const Geometry& g, int anchor, int& low,
                    int& high) {
    const int max_origin = g.source_frames() - p.patch_time;
    const long long start = static_cast<long long>(anchor) - p.search_bwd;
    const long long end = static_cast<long long>(anchor) + p.search_fwd;
    const long long shift = std::min(0LL, start) + std::max(0LL, end - max_origin);
    low = static_cast<int>(std::max(0LL, start - shift));
    high = static_cast<int>(std::min<long long>(max_origin, end - shift));
}

// schedule_temporal_range (core.cpp:898-908): anchor, future, past.
void schedule_frames(int anchor, int low, int high, std::vector<int>& out) {
    out.clear();
    out.push_back(anchor);
    for (int f = anchor + 1; f <= high; ++f) {
        out.push_back(f);
    }
    for (int f = anchor - 1; f >= low; --f) {
        out.push_back(f);
    }
}

// compute_spatial_range (core.cpp:910-932).
void spatial_range(const vo_params& p, int center, int axis, int& low,
                   int& high) {
    const int max_origin = axis - p.patch_size;
    const int half = (p.search_window - 1) / 2;
    const int start = center - half;
    const int end = center + half;
    const int shift = std::min(0, start) + std::max(0, end - max_origin);
    low = std::max(0, start - shift);
    high = std::min(max_origin, end - shift);
}

// Search-centre guidance (MVToolsFlowProvider::center_for, src/flow/flow.hpp:340-390,
// as a dense input): offsets[((anchor_frame * temporal + ti) * height + y) *
// width + x] holds (dx, dy) for the anchor at (x, y) and the ti-th frame of
// its schedule; the centre is the anchor plus the offset, clamped to the
// frame, and the anchor frame itself (ti = 0) keeps the anchor.  Null: the
// anchor is the centre of every window.
const int16_t* g_offsets = nullptr;
int g_offsets_temporal = 0;

void search_window(const vo_params& p, const Geometry& g, int ax, int ay, int af, int ti,
                   int& x_low, int& x_high, int& y_low, int& y_high) {
    int cx = ax;
    int cy = ay;
    if (g_offsets != nullptr && ti > 0) {
        const std::size_t at =
            ((static_cast<std::size_t>(af) * g_offsets_temporal + ti) * g.height + ay) * g.width + ax;
        cx = std::clamp(ax + g_offsets[2 * at], 0, g.width - 1);
        cy = std::clamp(ay + g_offsets[2 * at + 1], 0, g.height - 1);
    }
    spatial_range(p, cx, g.width, x_low, x_high);
    spatial_range(p, cy, g.height, y_low, y_high);
}

// reference_processing_mask_selected (core.cpp:934-950).
bool grid_selected(const vo_params& p, const Geometry& g, int anchor_frame,
                   int x, int y) {
    const int step = normalized_proc_step(p);
    const int max_x = g.width - p.patch_size;
    const int max_y = g.height - p.patch_size;
    const int max_frame = g.source_frames() - p.patch_time;
    const int phase_y = anchor_frame == max_frame ? 0 : anchor_frame;
    if (y != 0 && y != max_y && y % step != phase_y % step) {
        return false;
    }
    const int phase_x = y == max_y ? 0 : phase_y + (y / step);
    return x == 0 || x == max_x || x % step == phase_x % step;
}

// prepare_aggregation_window (core.cpp:296-324).
std::vector<float> aggregation_window(const vo_params& p) {
    const int n = p.patch_size;
    std::vector<float> window(static_cast<std::size_t>(n) * n, 1.0F);
    const float gamma = p.weight_gamma;
    if (gamma == 0.0F || n <= 1) {
        return window;
    }
    const float center = 0.5F * static_cast<float>(n - 1);
    const float radius = std::max(center, 0.5F);
    float max_weight = 0.0F;
    for (int y = 0; y < n; ++y) {
        const float dy = (static_cast<float>(y) - center) / radius;
        for (int x = 0; x < n; ++x) {
            const float dx = (static_cast<float>(x) - center) / radius;
            const float raw = std::exp(-0.5F * ((dx * dx) + (dy * dy)));
            window[static_cast<std::size_t>(y * n + x)] = raw;
            max_weight = std::max(max_weight, raw);
        }
    }
    const float inv_max = max_weight > 0.0F ? 1.0F / max_weight : 1.0F;
    for (float& value : window) {
        value = std::pow(value * inv_max, gamma);
    }
    return window;
}

struct Match {
    float distance = 0.0F;
    int x = 0;
    int y = 0;
    int frame = 0;
    int scan = 0;
};

struct StageSetup {
    int stage = VO_STAGE_BASIC;
    vo_params p{};
    Geometry g{};
    const float* noisy = nullptr;
    const float* ref = nullptr;
    int patch_area = 0;
    int patch_dim = 0;
    bool coupled = false;     // couple_channels && channels > 1
    int model_dim = 0;        // dimension of one model (coupled: C*patch_dim)
    int models = 0;           // models per group (coupled: 1, else channels)
    int slots = 0;
    int soft_cap = 0;
    int model_cap = 0;
    int cap = 0;              // retained_cap
    std::vector<float> window;
    int flags = 0;

    float sample(const float* video, int x, int y, int f, int c) const {
        return video[((g.local(f) * g.channels + c) * g.height +
                      y) *
                         g.width +
                     x];
    }
};

StageSetup make_setup(int stage, const vo_params& p, const float* noisy,
                      const float* ref, const Geometry& g, int flags) {
    validate(p, g);
    require(noisy != nullptr, "video data pointer must not be null");
    if (stage == VO_STAGE_FINAL) {
        require(ref != nullptr, "Final needs a reference clip");
    }
    StageSetup s;
    s.stage = stage;
    s.p = p;
    s.g = g;
    s.noisy = noisy;
    s.ref = ref;
    s.patch_area = p.patch_size * p.patch_size;
    s.patch_dim = s.patch_area * p.patch_time;
    s.coupled = p.couple_channels != 0 && g.channels > 1;
    s.model_dim = s.coupled ? g.channels * s.patch_dim : s.patch_dim;
    s.models = s.coupled ? 1 : g.channels;
    s.slots = p.search_bwd + p.search_fwd + p.patch_time;
    s.soft_cap = capped_similar(p.similar, p.similar_cap_factor);
    s.model_cap = capped_similar(p.similar, p.model_cap_factor);
    s.cap = retained_cap(p);
    s.window = aggregation_window(p);
    s.flags = flags;
    return s;
}

struct Workspace {
    std::vector<Match> candidates;
    std::vector<Match> matches;
    std::vector<int> scheduled;
    std::vector<float> reference_patch;
    std::vector<float> row_buffer;
    std::vector<int> order;
    std::vector<float> heap;
};

// Candidate enumeration in the CPU scan order (core.cpp:1296-1359): scheduled
// frames, then y, then x, each with its full distance.  The reference's early
// exit (core.cpp:1064-1077) only abandons candidates whose partial sum already
// exceeds the limit; partial sums of squares never decrease, so the recorded
// set and distances are those of the full sums computed here.
void enumerate_candidates(const StageSetup& s, int ax, int ay, int af,
                          Workspace& ws) {
    const vo_params& p = s.p;
    const int P = p.patch_size;
    const bool final_stage = s.stage == VO_STAGE_FINAL;
    const float* source = final_stage ? s.ref : s.noisy;
    const int match_channels = final_stage ? s.g.channels : 1;

    // load_reference_patch_basic / _final (core.cpp:1152-1198).
    ws.reference_patch.resize(static_cast<std::size_t>(match_channels) *
                              s.patch_dim);
    for (int c = 0; c < match_channels; ++c) {
        for (int t = 0; t < p.patch_time; ++t) {
            for (int y = 0; y < P; ++y) {
                for (int x = 0; x < P; ++x) {
                    ws.reference_patch[static_cast<std::size_t>(c) *
                                           s.patch_dim +
                                       (t * P + y) * P + x] =
                        s.sample(source, ax + x, ay + y, af + t, c);
                }
            }
        }
    }

    int low = 0;
    int high = 0;
    temporal_range(p, s.g, af, low, high);
    schedule_frames(af, low, high, ws.scheduled);
    ws.candidates.clear();
    ws.row_buffer.resize(static_cast<std::size_t>(P));
    int scan = 0;
    for (std::size_t ti = 0; ti < ws.scheduled.size(); ++ti) {
        const int frame = ws.scheduled[ti];
        int x_low = 0;
        int x_high = 0;
        int y_low = 0;
        int y_high = 0;
        search_window(p, s.g, ax, ay, af, static_cast<int>(ti), x_low, x_high, y_low, y_high);
        for (int cy = y_low; cy <= y_high; ++cy) {
            for (int cx = x_low; cx <= x_high; ++cx) {
                float distance = 0.0F;
                for (int c = 0; c < match_channels; ++c) {
                    for (int t = 0; t < p.patch_time; ++t) {
                        for (int y = 0; y < P; ++y) {
                            const float* reference_row =
                                ws.reference_patch.data() +
                                static_cast<std::size_t>(c) * s.patch_dim +
                                (t * P + y) * P;
                            const float* candidate_row =
                                source +
                                ((s.g.local(frame + t) *
                                      s.g.channels +
                                  c) *
                                     s.g.height +
                                 cy + y) *
                                    s.g.width +
                                cx;
                            distance = add_row_distance(distance, reference_row,
                                                        candidate_row, P);
                        }
                    }
                }
                ws.candidates.push_back(Match{distance, cx, cy, frame, scan});
                ++scan;
            }
        }
    }
}

// Set rule for find_similar_patches + select_similar_patches
// (core.cpp:1200-1362).  With T = max(tau, d_K), d_K the K-th smallest
// distance, the reference's bounded buffer ends up holding every candidate
// with d <= T when there are at most cap of them, and otherwise the cap
// smallest; select keeps exactly those, the K smallest first in ascending
// distance.  Exact distance ties are ordered by scan position here; the
// reference leaves their order to std::partial_sort.
int select_set_rule(const StageSetup& s, int ax, int ay, int af,
                    Workspace& ws) {
    const vo_params& p = s.p;
    ws.matches.clear();
    if (p.similar <= 1) {
        ws.matches.push_back(Match{0.0F, ax, ay, af, 0});
        return 1;
    }
    std::vector<Match>& all = ws.candidates;
    ws.order.clear();
    for (int i = 0; i < static_cast<int>(all.size()); ++i) {
        if (!std::isnan(all[static_cast<std::size_t>(i)].distance)) {
            ws.order.push_back(i);
        }
    }
    const auto less = [&](int l, int r) {
        const Match& a = all[static_cast<std::size_t>(l)];
        const Match& b = all[static_cast<std::size_t>(r)];
        if (a.distance != b.distance) {
            return a.distance < b.distance;
        }
        return a.scan < b.scan;
    };
    std::sort(ws.order.begin(), ws.order.end(), less);
    const int n = static_cast<int>(ws.order.size());
    if (n == 0) {
        return 0;
    }
    const int k = std::min(p.similar, n);
    const float kth =
        all[static_cast<std::size_t>(ws.order[static_cast<std::size_t>(k - 1)])]
            .distance;
    const float threshold = std::max(p.tau, kth);
    int within = 0;
    while (within < n &&
           all[static_cast<std::size_t>(ws.order[static_cast<std::size_t>(within)])]
                   .distance <= threshold) {
        ++within;
    }
    const long long window_cap = static_cast<long long>(p.search_window) *
                                 p.search_window *
                                 static_cast<long long>(ws.scheduled.size());
    const int buffer_cap =
        static_cast<int>(std::min<long long>(s.soft_cap, window_cap));
    const int retained = std::min(within, buffer_cap);
    std::vector<char> keep(all.size(), 0);
    for (int i = 0; i < retained; ++i) {
        keep[static_cast<std::size_t>(ws.order[static_cast<std::size_t>(i)])] = 1;
    }
    for (int i = 0; i < k; ++i) {
        const int index = ws.order[static_cast<std::size_t>(i)];
        ws.matches.push_back(all[static_cast<std::size_t>(index)]);
        keep[static_cast<std::size_t>(index)] = 0;
    }
    for (std::size_t i = 0; i < all.size(); ++i) {
        if (keep[i] != 0) {
            ws.matches.push_back(all[i]);
        }
    }
    return static_cast<int>(ws.matches.size());
}

// The reference's sequential procedure, stated literally
// (find_similar_patches core.cpp:1245-1362 with the early-exit scan of
// scan_*_spatial_distances :1033-1150, select_similar_patches
// :1200-1243), with libstdc++'s heap and partial_sort.
int select_faithful(const StageSetup& s, int ax, int ay, int af,
                    Workspace& ws) {
    const vo_params& p = s.p;
    ws.matches.clear();
    if (p.similar <= 1) {
        ws.matches.push_back(Match{0.0F, ax, ay, af, 0});
        return 1;
    }
    const int P = p.patch_size;
    const bool final_stage = s.stage == VO_STAGE_FINAL;
    const float* source = final_stage ? s.ref : s.noisy;
    const int match_channels = final_stage ? s.g.channels : 1;
    ws.reference_patch.resize(static_cast<std::size_t>(match_channels) * s.patch_dim);
    for (int c = 0; c < match_channels; ++c) {
        for (int t = 0; t < p.patch_time; ++t) {
            for (int y = 0; y < P; ++y) {
                for (int x = 0; x < P; ++x) {
                    ws.reference_patch[static_cast<std::size_t>(c) * s.patch_dim +
                                       (t * P + y) * P + x] =
                        s.sample(source, ax + x, ay + y, af + t, c);
                }
            }
        }
    }
    int low = 0;
    int high = 0;
    temporal_range(p, s.g, af, low, high);
    schedule_frames(af, low, high, ws.scheduled);
    const long long window_cap = static_cast<long long>(p.search_window) *
                                 p.search_window *
                                 static_cast<long long>(ws.scheduled.size());
    const int output_cap =
        static_cast<int>(std::min<long long>(s.soft_cap, window_cap));
    std::vector<float>& top = ws.heap;
    top.clear();
    float soft_worst = std::numeric_limits<float>::infinity();
    std::size_t soft_worst_index = 0;
    const auto farther = [](float l, float r) { return l < r; };
    const auto by_distance = [](const Match& l, const Match& r) {
        return l.distance < r.distance;
    };
    const auto refresh = [&]() {
        const auto worst =
            std::max_element(ws.matches.begin(), ws.matches.end(), by_distance);
        soft_worst_index = static_cast<std::size_t>(worst - ws.matches.begin());
        soft_worst = worst->distance;
    };
    int scan = 0;
    for (std::size_t ti = 0; ti < ws.scheduled.size(); ++ti) {
        const int frame = ws.scheduled[ti];
        int x_low = 0;
        int x_high = 0;
        int y_low = 0;
        int y_high = 0;
        search_window(p, s.g, ax, ay, af, static_cast<int>(ti), x_low, x_high, y_low, y_high);
        for (int cy = y_low; cy <= y_high; ++cy) {
            for (int cx = x_low; cx <= x_high; ++cx, ++scan) {
                float limit = static_cast<int>(top.size()) >= p.similar
                                  ? std::max(p.tau, top.front())
                                  : std::numeric_limits<float>::infinity();
                if (static_cast<int>(ws.matches.size()) >= output_cap) {
                    limit = std::min(limit, soft_worst);
                }
                float distance = 0.0F;
                for (int c = 0; c < match_channels && !(distance > limit); ++c) {
                    for (int t = 0; t < p.patch_time && !(distance > limit); ++t) {
                        for (int y = 0; y < P; ++y) {
                            const float* reference_row =
                                ws.reference_patch.data() +
                                static_cast<std::size_t>(c) * s.patch_dim + (t * P + y) * P;
                            const float* candidate_row =
                                source +
                                ((s.g.local(frame + t) * s.g.channels + c) *
                                     s.g.height +
                                 cy + y) *
                                    s.g.width +
                                cx;
                            distance = add_row_distance(distance, reference_row,
                                                        candidate_row, P);
                            if (distance > limit) {
                                break;
                            }
                        }
                    }
                }
                if (!(distance <= limit)) {
                    continue;
                }
                const Match candidate{distance, cx, cy, frame, scan};
                if (static_cast<int>(ws.matches.size()) < output_cap) {
                    ws.matches.push_back(candidate);
                    if (static_cast<int>(ws.matches.size()) == output_cap) {
                        refresh();
                    }
                } else if (candidate.distance < soft_worst) {
                    ws.matches[soft_worst_index] = candidate;
                    refresh();
                }
                if (static_cast<int>(top.size()) < p.similar) {
                    top.push_back(candidate.distance);
                    std::push_heap(top.begin(), top.end(), farther);
                } else if (candidate.distance < top.front()) {
                    std::pop_heap(top.begin(), top.end(), farther);
                    top.back() = candidate.distance;
                    std::push_heap(top.begin(), top.end(), farther);
                }
            }
        }
    }
    if (ws.matches.empty()) {
        return 0;
    }
    const int selected =
        std::min(p.similar, static_cast<int>(ws.matches.size()));
    std::partial_sort(ws.matches.begin(), ws.matches.begin() + selected,
                      ws.matches.end(), by_distance);
    const float kth = ws.matches[static_cast<std::size_t>(selected - 1)].distance;
    const float threshold = std::max(p.tau, kth);
    auto output = ws.matches.begin();
    for (auto it = ws.matches.begin(); it != ws.matches.end(); ++it) {
        if (it->distance <= threshold) {
            *output = *it;
            ++output;
        }
    }
    const int cap = std::min(s.soft_cap, static_cast<int>(ws.matches.size()));
    if (output > ws.matches.begin() + cap) {
        std::partial_sort(ws.matches.begin(), ws.matches.begin() + cap, output,
                          by_distance);
        output = ws.matches.begin() + cap;
    }
    ws.matches.resize(static_cast<std::size_t>(output - ws.matches.begin()));
    return static_cast<int>(ws.matches.size());
}

int find_matches(const StageSetup& s, int ax, int ay, int af, Workspace& ws) {
    if (s.p.similar <= 1) {
        ws.matches.assign(1, Match{0.0F, ax, ay, af, 0});
        return 1;
    }
    if ((s.flags & VO_SELECT_FAITHFUL) != 0) {
        return select_faithful(s, ax, ay, af, ws);
    }
    enumerate_candidates(s, ax, ay, af, ws);
    return select_set_rule(s, ax, ay, af, ws);
}

// ---------------------------------------------------------------------------
// Group model (core.cpp:184-294, 326-412, 491-871).
// ---------------------------------------------------------------------------

template <class T> struct ModelScratch {
    std::vector<T> noisy;       // S x d (as T)
    std::vector<T> basic;       // S x d
    std::vector<T> centered_noisy;
    std::vector<T> centered_basic;
    std::vector<T> mean_noisy;
    std::vector<T> mean_basic;
    std::vector<T> gram;
    std::vector<T> eigenvalues;
    std::vector<T> dual_vectors;
    std::vector<T> basis;
    std::vector<T> coefficients;
    std::vector<T> filtered;    // S x d
    std::vector<T> scores;
    std::vector<T> jm;
    std::vector<T> jv;
    std::vector<int> order;
};

struct GroupWeights {
    std::vector<float> log_weights;
    std::vector<float> weights;
    int models = 0;
};

template <class T> T positive_or_zero(T value) {
    return value > T{} ? value : T{};
}

struct ModelInfo {
    bool modelled = false;
    std::vector<float> eigenvalues;
    int gram_rows = 0;         // dual models: the B x B Gram XX^T / B, else 0
    std::vector<double> gram;  // in the model's precision, widened
};

// accumulate_aggregation_model (core.cpp:326-412): adds one model's log
// weights to log_weights[0, similar).  rows: the S centred membership rows
// (the basic rows in Final, the noisy rows in Basic), d values each;
// eigenvalues / basis: the model's `rank` leading pairs (basis rows of d).
// Group precision tau = eps + d / B + sum lambda^ / (lambda^ + beta sigma^2)
// with lambda^ = max(lambda - model noise, 0) (group_precision_tau,
// :222-242); membership scores Z and the volume term as in README.md:122-185.
template <class T>
void accumulate_model_weights(const vo_params& p, bool final_stage, const T* rows, int similar, int d,
                              const T* eigenvalues, const T* basis, int rank, int basis_similar,
                              std::vector<float>& log_weights, std::vector<T>& scores) {
    if (p.weight_alpha == 0.0F && p.weight_beta == 0.0F) {
        return;
    }
    const T sigma = static_cast<T>(p.sigma);
    const T sigma_basic = static_cast<T>(p.sigma_basic);
    const T beta = static_cast<T>(p.beta);
    const T model_noise2 = final_stage ? (beta * sigma_basic) * sigma_basic : (beta * sigma) * sigma;
    const T estimate_noise2 = (beta * sigma) * sigma;
    const T alpha = static_cast<T>(p.weight_alpha);
    T trace{};
    for (int j = 0; j < rank; ++j) {
        const T lambda = positive_or_zero(eigenvalues[j] - model_noise2);
        if (lambda > T{}) {
            trace = trace + lambda / (lambda + estimate_noise2);
        }
    }
    const T mean_term = static_cast<T>(d) / static_cast<T>(basis_similar);
    const T tau = (static_cast<T>(p.weight_epsilon) + mean_term) + trace;
    const T log_group = p.weight_alpha == 0.0F ? T{} : -alpha * t_log(tau);
    if (p.weight_beta == 0.0F) {
        for (int r = 0; r < similar; ++r) {
            log_weights[static_cast<std::size_t>(r)] =
                static_cast<float>(static_cast<T>(log_weights[static_cast<std::size_t>(r)]) + log_group);
        }
        return;
    }
    const T match_sigma = final_stage ? sigma_basic : sigma;
    const T match_noise2 = match_sigma * match_sigma;
    const T floor_noise2 = (static_cast<T>(p.membership_noise_floor) * sigma) * sigma;
    const T membership_noise2 = std::max(match_noise2, floor_noise2);
    T volume{};
    for (int j = 0; j < rank; ++j) {
        const T lambda = positive_or_zero(eigenvalues[j] - model_noise2);
        if (lambda > T{}) {
            volume = volume + t_log1p(lambda / membership_noise2);
        }
    }
    volume = volume / static_cast<T>(d);
    T min_score = std::numeric_limits<T>::infinity();
    const T score_scale = std::sqrt(static_cast<T>(2) * static_cast<T>(d));
    scores.resize(static_cast<std::size_t>(similar));
    for (int r = 0; r < similar; ++r) {
        const T* row = rows + static_cast<std::size_t>(r) * d;
        const T norm2 = dot(row, row, static_cast<std::size_t>(d));
        T distance{};
        T projected_norm2{};
        for (int j = 0; j < rank; ++j) {
            const T projection = dot(row, basis + static_cast<std::size_t>(j) * d, static_cast<std::size_t>(d));
            const T projection2 = projection * projection;
            const T lambda = positive_or_zero(eigenvalues[j] - model_noise2);
            distance = distance + projection2 / (lambda + membership_noise2);
            projected_norm2 = projected_norm2 + projection2;
        }
        const T residual = positive_or_zero(norm2 - projected_norm2);
        distance = distance + residual / membership_noise2;
        const T score = (distance - static_cast<T>(d)) / score_scale;
        scores[static_cast<std::size_t>(r)] = score;
        min_score = std::min(min_score, score);
    }
    const T weight_beta = static_cast<T>(p.weight_beta);
    for (int r = 0; r < similar; ++r) {
        const T membership_value = positive_or_zero(scores[static_cast<std::size_t>(r)] - min_score) + volume;
        const T delta = log_group - (weight_beta * membership_value);
        log_weights[static_cast<std::size_t>(r)] =
            static_cast<float>(static_cast<T>(log_weights[static_cast<std::size_t>(r)]) + delta);
    }
}

// finalize_aggregation_weights (core.cpp:276-294) with
// accumulate_default_aggregation_model (:245-274) when no model added log
// weights (equal groups): w = exp(clamp(mean log weight, +-80)), all 1 when
// alpha = beta_w = gamma_w = 0.
void finalize_weights(const vo_params& p, int model_dim, int basis_similar, int models,
                      std::vector<float>& log_weights, std::vector<float>& weights, int similar) {
    const bool enabled = p.weight_alpha != 0.0F || p.weight_beta != 0.0F || p.weight_gamma != 0.0F;
    weights.assign(static_cast<std::size_t>(similar), 1.0F);
    if (!enabled) {
        return;
    }
    if (models == 0) {
        const float tau = p.weight_epsilon + (static_cast<float>(model_dim) / static_cast<float>(basis_similar));
        const float log_group = p.weight_alpha == 0.0F ? 0.0F : -p.weight_alpha * t_log(tau);
        for (int r = 0; r < similar; ++r) {
            log_weights[static_cast<std::size_t>(r)] = log_weights[static_cast<std::size_t>(r)] + log_group;
        }
        models = 1;
    }
    const float inv_models = 1.0F / static_cast<float>(models);
    for (int r = 0; r < similar; ++r) {
        const float value = log_weights[static_cast<std::size_t>(r)] * inv_models;
        weights[static_cast<std::size_t>(r)] = t_exp(std::clamp(value, -80.0F, 80.0F));
    }
}

// filter_vnlb_samples (core.cpp:638-792) including compute_filter_coefficients
// (:616-636), accumulate_aggregation_model (:326-412) and
// project_bayes_estimate (:578-614).  Inputs are the S rows of one model;
// writes S filtered rows and adds the model's log weights.
template <class T>
void filter_samples(const StageSetup& s, ModelScratch<T>& m, int similar,
                    int d, bool flat, std::vector<float>& log_weights,
                    int& model_count, ModelInfo* info) {
    const vo_params& p = s.p;
    const bool final_stage = s.stage == VO_STAGE_FINAL;
    const int rank = p.rank;
    const int basis_similar = std::min(similar, s.model_cap);
    if (basis_similar < std::min(d, rank)) {
        throw std::runtime_error("group_size must cover the effective VNLB rank");
    }
    const std::size_t sd = static_cast<std::size_t>(similar) * d;
    m.mean_noisy.assign(static_cast<std::size_t>(d), T{});
    m.mean_basic.assign(static_cast<std::size_t>(d), T{});
    m.centered_noisy.resize(sd);
    m.centered_basic.resize(sd);
    const bool use_dual = d >= 64 && basis_similar < d;

    const auto center_with = [&](const std::vector<T>& rows, const std::vector<T>& mean,
                                 std::vector<T>& out) {
        for (int r = 0; r < similar; ++r) {
            for (int c = 0; c < d; ++c) {
                const std::size_t i = static_cast<std::size_t>(r) * d + c;
                out[i] = rows[i] - mean[static_cast<std::size_t>(c)];
            }
        }
    };

    if (final_stage) {
        compute_mean(m.basic.data(), basis_similar, d, m.mean_basic.data());
        center_with(m.basic, m.mean_basic, m.centered_basic);
        if (flat) {
            center_with(m.noisy, m.mean_basic, m.centered_noisy);
        } else {
            compute_mean(m.noisy.data(), basis_similar, d, m.mean_noisy.data());
            center_with(m.noisy, m.mean_noisy, m.centered_noisy);
        }
    } else {
        compute_mean(m.noisy.data(), basis_similar, d, m.mean_noisy.data());
        center_with(m.noisy, m.mean_noisy, m.centered_noisy);
    }
    const std::vector<T>& center =
        (final_stage && flat) ? m.mean_basic : m.mean_noisy;
    const std::vector<T>& membership =
        final_stage ? m.centered_basic : m.centered_noisy;
    const std::vector<T>& model_rows =
        final_stage ? m.centered_basic : m.centered_noisy;

    const T sigma = static_cast<T>(p.sigma);
    const T sigma_basic = static_cast<T>(p.sigma_basic);
    const T beta = static_cast<T>(p.beta);

    m.filtered.resize(sd);
    m.eigenvalues.assign(static_cast<std::size_t>(std::max(rank, 1)), T{});
    m.basis.assign(static_cast<std::size_t>(rank) * d, T{});
    if (rank > 0) {
        const T scale = T{1} / static_cast<T>(basis_similar);
        if (use_dual) {
            m.gram.resize(static_cast<std::size_t>(basis_similar) * basis_similar);
            compute_gram(model_rows.data(), basis_similar, d, scale, m.gram.data());
            m.dual_vectors.resize(static_cast<std::size_t>(rank) * basis_similar);
            if (info != nullptr) {
                info->gram_rows = basis_similar;
                info->gram.assign(m.gram.begin(), m.gram.end());
            }
            topk_eigen(m.gram.data(), basis_similar, rank, m.eigenvalues.data(),
                       m.dual_vectors.data(), m.jm, m.jv, m.order);
            dual_basis(model_rows.data(), basis_similar, d, m.eigenvalues.data(),
                       m.dual_vectors.data(), rank, m.basis.data());
        } else {
            m.gram.resize(static_cast<std::size_t>(d) * d);
            compute_covariance(model_rows.data(), basis_similar, d, scale,
                               m.gram.data());
            topk_eigen(m.gram.data(), d, rank, m.eigenvalues.data(),
                       m.basis.data(), m.jm, m.jv, m.order);
        }
        // compute_filter_coefficients (core.cpp:616-636).
        const T sigma2 = (beta * sigma) * sigma;
        const T sigma_basic2 = final_stage ? (beta * sigma_basic) * sigma_basic : sigma2;
        const T threshold = static_cast<T>(p.variance_threshold) * sigma2;
        m.coefficients.resize(static_cast<std::size_t>(rank));
        for (int j = 0; j < rank; ++j) {
            T variance = m.eigenvalues[static_cast<std::size_t>(j)];
            if (sigma_basic2 > T{}) {
                variance = variance - std::min(variance, sigma_basic2);
            }
            m.coefficients[static_cast<std::size_t>(j)] =
                variance > threshold ? T{1} / (T{1} + (sigma2 / variance)) : T{};
        }
        if (info != nullptr) {
            info->modelled = true;
            info->eigenvalues.assign(m.eigenvalues.begin(),
                                     m.eigenvalues.begin() + rank);
        }
    }

    // accumulate_aggregation_model (core.cpp:326-412).
    accumulate_model_weights<T>(p, final_stage, membership.data(), similar, d, m.eigenvalues.data(),
                                m.basis.data(), rank, basis_similar, log_weights, m.scores);
    ++model_count;

    // project_bayes_estimate (core.cpp:578-614), or the centre for rank 0.
    for (int r = 0; r < similar; ++r) {
        T* out = m.filtered.data() + static_cast<std::size_t>(r) * d;
        if (rank == 0) {
            for (int c = 0; c < d; ++c) {
                out[c] = center[static_cast<std::size_t>(c)];
            }
            continue;
        }
        const T* input = m.centered_noisy.data() + static_cast<std::size_t>(r) * d;
        std::fill(out, out + d, T{});
        for (int j = 0; j < rank; ++j) {
            const T coefficient = m.coefficients[static_cast<std::size_t>(j)];
            if (coefficient == T{}) {
                continue;
            }
            const T* basis = m.basis.data() + static_cast<std::size_t>(j) * d;
            const T projection = dot(input, basis, static_cast<std::size_t>(d));
            const T scaled = coefficient * projection;
            add_scaled(out, basis, scaled, static_cast<std::size_t>(d));
        }
        for (int c = 0; c < d; ++c) {
            out[c] = out[c] + center[static_cast<std::size_t>(c)];
        }
    }
}

// Equality of all retained rows (samples_are_equal core.cpp:516-531,
// group_is_equal core.cpp:110-131).
bool rows_equal(const std::vector<float>& rows, int similar, int dim) {
    if (similar <= 1) {
        return false;
    }
    for (int r = 1; r < similar; ++r) {
        for (int c = 0; c < dim; ++c) {
            if (rows[static_cast<std::size_t>(r) * dim + c] != rows[static_cast<std::size_t>(c)]) {
                return false;
            }
        }
    }
    return true;
}

// group_variance / samples_channel_variance (core.cpp:491-559): per channel
// unbiased variance over all retained values, double accumulators, mean over
// channels.  rows are sample-major [S][C*patch_dim].
float channel_variance(const std::vector<float>& rows, int similar, int channels,
                       int patch_dim, bool group_order) {
    const int values = patch_dim * similar;
    if (values <= 1) {
        return 0.0F;
    }
    const int dim = channels * patch_dim;
    float variance = 0.0F;
    for (int c = 0; c < channels; ++c) {
        double sum = 0.0;
        double sum2 = 0.0;
        if (group_order) {  // [position][patch]
            for (int pos = 0; pos < patch_dim; ++pos) {
                for (int r = 0; r < similar; ++r) {
                    const float v = rows[static_cast<std::size_t>(r) * dim + c * patch_dim + pos];
                    sum += v;
                    sum2 += static_cast<double>(v) * static_cast<double>(v);
                }
            }
        } else {  // [patch][position]
            for (int r = 0; r < similar; ++r) {
                for (int pos = 0; pos < patch_dim; ++pos) {
                    const float v = rows[static_cast<std::size_t>(r) * dim + c * patch_dim + pos];
                    sum += v;
                    sum2 += static_cast<double>(v) * static_cast<double>(v);
                }
            }
        }
        const double count = static_cast<double>(values);
        variance = variance + static_cast<float>((sum2 - (sum * sum / count)) / (count - 1.0));
    }
    return variance / static_cast<float>(channels);
}

// ---------------------------------------------------------------------------
// Per anchor frame driver: process_anchor_impl (core.cpp:1820-1906).
// ---------------------------------------------------------------------------

template <class Acc> struct Stack {
    int slots = 0;
    int planes = 0;  // channels + 1 (weight last)
    int width = 0;
    int height = 0;
    std::vector<Acc> data;
    void reset(int s, int c, int w, int h) {
        slots = s;
        planes = c + 1;
        width = w;
        height = h;
        data.assign(static_cast<std::size_t>(slots) * planes * w * h, Acc{});
    }
    Acc& at(int slot, int plane, int x, int y) {
        return data[((static_cast<std::size_t>(slot) * planes + plane) * height + y) * width + x];
    }
};

struct FrameDebug {
    int capacity = 0;
    int count = 0;
    int* anchor_x = nullptr;
    int* anchor_y = nullptr;
    int* counts = nullptr;
    int* positions = nullptr;
    float* distances = nullptr;
    unsigned char* active = nullptr;
    float* weights = nullptr;
    float* eigenvalues = nullptr;
    int* group_flags = nullptr;
    int gram_dim = 0;          // grams: [capacity][gram_dim][gram_dim]
    double* grams = nullptr;
};

template <class T, class Acc>
long long process_anchor_frame(const StageSetup& s, int anchor_frame, Stack<Acc>& stack,
                               FrameDebug* debug) {
    const vo_params& p = s.p;
    const Geometry& g = s.g;
    const int P = p.patch_size;
    stack.reset(s.slots, g.channels, g.width, g.height);  // clear_contributions
    if (anchor_frame < 0 || anchor_frame > g.source_frames() - p.patch_time) {
        if (debug != nullptr) {
            debug->count = 0;
        }
        return 0;
    }
    {
        // The window must hold the anchor's input frames
        // (input_frame_range_for_anchor, core.cpp:2026-2034).
        int low = 0;
        int high = 0;
        temporal_range(p, g, anchor_frame, low, high);
        require(low >= g.first_frame && high + p.patch_time - 1 < g.first_frame + g.frames,
                "the frame window does not hold the anchor frame's input frames");
    }
    const int max_x = g.width - P;
    const int max_y = g.height - P;
    std::vector<unsigned char> mask(static_cast<std::size_t>(g.width) * g.height, 0);
    std::vector<int> grid_index(static_cast<std::size_t>(g.width) * g.height, -1);
    int grid_count = 0;
    for (int y = 0; y <= max_y; ++y) {
        for (int x = 0; x <= max_x; ++x) {
            if (grid_selected(p, g, anchor_frame, x, y)) {
                mask[static_cast<std::size_t>(y) * g.width + x] = 1;
                grid_index[static_cast<std::size_t>(y) * g.width + x] = grid_count++;
            }
        }
    }
    const int cap = s.cap;
    Workspace ws;
    if (debug != nullptr) {
        require(grid_count <= debug->capacity, "debug capacity too small");
        debug->count = grid_count;
        // Matching does not read the mask, so every grid anchor's list is
        // recorded up front.
        for (int y = 0; y <= max_y; ++y) {
            for (int x = 0; x <= max_x; ++x) {
                const int a = grid_index[static_cast<std::size_t>(y) * g.width + x];
                if (a < 0) {
                    continue;
                }
                const int similar = find_matches(s, x, y, anchor_frame, ws);
                if (debug->anchor_x != nullptr) {
                    debug->anchor_x[a] = x;
                    debug->anchor_y[a] = y;
                }
                if (debug->counts != nullptr) {
                    debug->counts[a] = similar;
                }
                for (int i = 0; i < similar && i < cap; ++i) {
                    const Match& mt = ws.matches[static_cast<std::size_t>(i)];
                    if (debug->positions != nullptr) {
                        debug->positions[static_cast<std::size_t>(a) * cap + i] =
                            (mt.frame * g.height + mt.y) * g.width + mt.x;
                    }
                    if (debug->distances != nullptr) {
                        debug->distances[static_cast<std::size_t>(a) * cap + i] = mt.distance;
                    }
                }
                if (debug->active != nullptr) {
                    debug->active[a] = 0;
                }
                if (debug->group_flags != nullptr) {
                    debug->group_flags[a] = 0;
                }
            }
        }
    }

    const int channels = g.channels;
    const int patch_dim = s.patch_dim;
    const int dim = channels * patch_dim;
    std::vector<float> rows_noisy;
    std::vector<float> rows_basic;
    std::vector<float> filtered;
    GroupWeights gw;
    ModelScratch<T> scratch;
    ModelInfo info;
    long long groups = 0;
    const bool final_stage = s.stage == VO_STAGE_FINAL;

    for (int ay = 0; ay <= max_y; ++ay) {
        for (int ax = 0; ax <= max_x; ++ax) {
            if (mask[static_cast<std::size_t>(ay) * g.width + ax] == 0) {
                continue;
            }
            const int similar = find_matches(s, ax, ay, anchor_frame, ws);
            if (similar <= 0) {
                continue;
            }
            // Gather sample-major rows [patch][c][t][y][x]
            // (gather_*_coupled core.cpp:1364-1585; the per-channel group
            // layout is the same data transposed).
            rows_noisy.resize(static_cast<std::size_t>(similar) * dim);
            if (final_stage) {
                rows_basic.resize(static_cast<std::size_t>(similar) * dim);
            }
            for (int r = 0; r < similar; ++r) {
                const Match& mt = ws.matches[static_cast<std::size_t>(r)];
                for (int c = 0; c < channels; ++c) {
                    for (int t = 0; t < p.patch_time; ++t) {
                        for (int y = 0; y < P; ++y) {
                            for (int x = 0; x < P; ++x) {
                                const std::size_t i = static_cast<std::size_t>(r) * dim +
                                                      c * patch_dim + (t * P + y) * P + x;
                                rows_noisy[i] = s.sample(s.noisy, mt.x + x, mt.y + y,
                                                         mt.frame + t, c);
                                if (final_stage) {
                                    rows_basic[i] = s.sample(s.ref, mt.x + x, mt.y + y,
                                                             mt.frame + t, c);
                                }
                            }
                        }
                    }
                }
            }

            // reset_aggregation_weights (core.cpp:235-243).
            gw.log_weights.assign(static_cast<std::size_t>(similar), 0.0F);
            gw.weights.assign(static_cast<std::size_t>(similar), 1.0F);
            gw.models = 0;
            info.modelled = false;
            info.gram_rows = 0;
            int group_flag = 0;

            filtered = rows_noisy;
            if (rows_equal(rows_noisy, similar, dim)) {
                group_flag |= 1;  // bypass: output = input
            } else {
                bool flat = false;
                if (final_stage && p.flat_areas != 0) {
                    const float threshold = (p.sigma * p.sigma) * p.gamma;
                    flat = channel_variance(rows_noisy, similar, channels, patch_dim,
                                            !s.coupled) < threshold;
                }
                if (flat) {
                    group_flag |= 2;
                }
                for (int model = 0; model < s.models; ++model) {
                    const int d = s.model_dim;
                    const int first = s.coupled ? 0 : model;  // channel offset
                    scratch.noisy.resize(static_cast<std::size_t>(similar) * d);
                    scratch.basic.resize(static_cast<std::size_t>(similar) * d);
                    for (int r = 0; r < similar; ++r) {
                        for (int c = 0; c < d; ++c) {
                            const std::size_t src =
                                static_cast<std::size_t>(r) * dim + first * patch_dim + c;
                            scratch.noisy[static_cast<std::size_t>(r) * d + c] =
                                static_cast<T>(rows_noisy[src]);
                            if (final_stage) {
                                scratch.basic[static_cast<std::size_t>(r) * d + c] =
                                    static_cast<T>(rows_basic[src]);
                            }
                        }
                    }
                    filter_samples<T>(s, scratch, similar, d, flat, gw.log_weights,
                                      gw.models, model == 0 ? &info : nullptr);
                    for (int r = 0; r < similar; ++r) {
                        for (int c = 0; c < d; ++c) {
                            filtered[static_cast<std::size_t>(r) * dim + first * patch_dim + c] =
                                static_cast<float>(
                                    scratch.filtered[static_cast<std::size_t>(r) * d + c]);
                        }
                    }
                }
                if (info.modelled) {
                    group_flag |= 4;
                }
            }

            finalize_weights(p, s.model_dim, std::min(similar, s.model_cap), gw.models, gw.log_weights,
                             gw.weights, similar);

            // Contributions (write_group_contributions core.cpp:1587-1636,
            // write_sample_contributions_coupled_* :1638-1795).
            const bool windowed = p.weight_gamma != 0.0F;
            const int fused_width = (s.coupled && !windowed) ? (P / kLanes) * kLanes : 0;
            const bool accum_double = (s.flags & VO_ACCUM_DOUBLE) != 0;
            for (int r = 0; r < similar; ++r) {
                const Match& mt = ws.matches[static_cast<std::size_t>(r)];
                const float w = gw.weights[static_cast<std::size_t>(r)];
                for (int t = 0; t < p.patch_time; ++t) {
                    const int offset = mt.frame + t - anchor_frame;
                    const int slot = offset + p.search_bwd;
                    if (slot < 0 || slot >= s.slots) {
                        continue;
                    }
                    for (int y = 0; y < P; ++y) {
                        for (int x = 0; x < P; ++x) {
                            const int ox = mt.x + x;
                            const int oy = mt.y + y;
                            const float cw = w * s.window[static_cast<std::size_t>(y * P + x)];
                            const bool fused = x < fused_width;
                            for (int c = 0; c < channels; ++c) {
                                const float value =
                                    filtered[static_cast<std::size_t>(r) * dim + c * patch_dim +
                                             (t * P + y) * P + x];
                                Acc& acc = stack.at(slot, c, ox, oy);
                                if (accum_double) {
                                    if (fused) {
                                        acc = acc + static_cast<Acc>(static_cast<double>(w) *
                                                                     static_cast<double>(value));
                                    } else {
                                        const float product = cw * value;
                                        acc = acc + static_cast<Acc>(product);
                                    }
                                } else {
                                    if (fused) {
                                        acc = static_cast<Acc>(std::fma(w, value, static_cast<float>(acc)));
                                    } else {
                                        const float product = cw * value;
                                        acc = static_cast<Acc>(static_cast<float>(acc) + product);
                                    }
                                }
                            }
                            Acc& wacc = stack.at(slot, channels, ox, oy);
                            if (accum_double) {
                                wacc = wacc + static_cast<Acc>(cw);
                            } else {
                                wacc = static_cast<Acc>(static_cast<float>(wacc) + cw);
                            }
                        }
                    }
                }
            }
            ++groups;

            if (debug != nullptr) {
                const int a = grid_index[static_cast<std::size_t>(ay) * g.width + ax];
                if (debug->active != nullptr) {
                    debug->active[a] = 1;
                }
                if (debug->weights != nullptr) {
                    for (int r = 0; r < similar && r < cap; ++r) {
                        debug->weights[static_cast<std::size_t>(a) * cap + r] =
                            gw.weights[static_cast<std::size_t>(r)];
                    }
                }
                if (debug->eigenvalues != nullptr && info.modelled) {
                    for (int j = 0; j < p.rank; ++j) {
                        debug->eigenvalues[static_cast<std::size_t>(a) * p.rank + j] =
                            info.eigenvalues[static_cast<std::size_t>(j)];
                    }
                }
                if (debug->group_flags != nullptr) {
                    debug->group_flags[a] = group_flag;
                }
                if (debug->grams != nullptr && info.modelled && info.gram_rows > 0 &&
                    info.gram_rows <= debug->gram_dim) {
                    const int n = info.gram_rows;
                    for (int r = 0; r < n; ++r) {
                        for (int c2 = 0; c2 < n; ++c2) {
                            debug->grams[(static_cast<std::size_t>(a) * debug->gram_dim + r) *
                                             debug->gram_dim + c2] =
                                info.gram[static_cast<std::size_t>(r) * n + c2];
                        }
                    }
                }
            }

            // apply_frame_paste_mask (core.cpp:980-1008).
            const auto clear = [&](int frame, int x, int y) {
                if (frame != anchor_frame || x < 0 || x >= g.width || y < 0 ||
                    y >= g.height) {
                    return;
                }
                mask[static_cast<std::size_t>(y) * g.width + x] = 0;
            };
            for (int r = 0; r < similar; ++r) {
                const Match& mt = ws.matches[static_cast<std::size_t>(r)];
                clear(mt.frame, mt.x, mt.y);
                if (mt.y > 2 * P) {
                    clear(mt.frame, mt.x, mt.y - 1);
                }
                if (mt.y < g.height - 2 * P) {
                    clear(mt.frame, mt.x, mt.y + 1);
                }
                if (mt.x > 2 * P) {
                    clear(mt.frame, mt.x - 1, mt.y);
                }
                if (mt.x < g.width - 2 * P) {
                    clear(mt.frame, mt.x + 1, mt.y);
                }
            }
        }
    }
    return groups;
}

// vnlb.Aggregate (src/vs/plugin.cpp:966-1045): output t sums the stacks of
// anchors [max(0, t - fwd - (Pt - 1)), min(n - 1, t + bwd)] at slot
// t - anchor + bwd, then numerator / weight or the source sample.
template <class Acc>
void aggregate(const StageSetup& s, std::vector<Stack<Acc>>& stacks, float* out) {
    const Geometry& g = s.g;
    const vo_params& p = s.p;
    const std::size_t plane = static_cast<std::size_t>(g.width) * g.height;
    for (int t = 0; t < g.frames; ++t) {
        const int first = std::max(0, t - (p.search_fwd + p.patch_time - 1));
        const int last = std::min(g.frames - 1, t + p.search_bwd);
        for (int y = 0; y < g.height; ++y) {
            for (int x = 0; x < g.width; ++x) {
                Acc numerator[8] = {};
                Acc weight{};
                for (int a = first; a <= last; ++a) {
                    const int slot = t - a + p.search_bwd;
                    Stack<Acc>& st = stacks[static_cast<std::size_t>(a)];
                    for (int c = 0; c < g.channels; ++c) {
                        numerator[c] = numerator[c] + st.at(slot, c, x, y);
                    }
                    weight = weight + st.at(slot, g.channels, x, y);
                }
                const float wf = static_cast<float>(weight);
                for (int c = 0; c < g.channels; ++c) {
                    const std::size_t index =
                        (static_cast<std::size_t>(t) * g.channels + c) * plane +
                        static_cast<std::size_t>(y) * g.width + x;
                    out[index] = wf > 0.0F ? static_cast<float>(numerator[c]) / wf
                                           : s.noisy[index];
                }
            }
        }
    }
}

template <class T, class Acc>
long long run_stage(const StageSetup& s, float* out, int threads) {
    require(s.g.channels <= 8, "at most eight channels are supported");
    std::vector<Stack<Acc>> stacks(static_cast<std::size_t>(s.g.frames));
    std::atomic<int> next{0};
    std::atomic<long long> groups{0};
    std::mutex error_lock;
    std::string error;
    const auto worker = [&]() {
        try {
            for (;;) {
                const int frame = next.fetch_add(1);
                if (frame >= s.g.frames) {
                    break;
                }
                groups += process_anchor_frame<T, Acc>(
                    s, frame, stacks[static_cast<std::size_t>(frame)], nullptr);
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lock(error_lock);
            error = e.what();
            next = s.g.frames;
        }
    };
    const int count = std::max(1, std::min(threads, s.g.frames));
    std::vector<std::thread> pool;
    for (int i = 1; i < count; ++i) {
        pool.emplace_back(worker);
    }
    worker();
    for (auto& th : pool) {
        th.join();
    }
    if (!error.empty()) {
        throw std::runtime_error(error);
    }
    aggregate(s, stacks, out);
    return groups.load();
}

}  // namespace

extern "C" {

const char* vo_last_error(void) {
    std::lock_guard<std::mutex> lock(g_error_lock);
    return g_error.c_str();
}

int vo_retained_cap(const vo_params* params) {
    if (params == nullptr || params->similar <= 0) {
        return -1;
    }
    return retained_cap(*params);
}

int vo_aggregation_weights(int stage, const vo_params* params, const double* rows, int similar,
                           int dim, const double* eigenvalues, const double* basis, int rank,
                           int basis_similar, int models, int precision, float* weights) {
    try {
        require(params != nullptr && weights != nullptr && similar > 0 && dim > 0 && rank >= 0 &&
                    basis_similar > 0 && models >= 0,
                "bad arguments");
        require(rank == 0 || (eigenvalues != nullptr && basis != nullptr), "rank needs eigen pairs");
        require(rows != nullptr, "rows must not be null");
        const bool final_stage = stage == VO_STAGE_FINAL;
        std::vector<float> log_weights(static_cast<std::size_t>(similar), 0.0F);
        std::vector<float> out;
        const auto run = [&](auto zero) {
            using T = decltype(zero);
            const std::size_t sd = static_cast<std::size_t>(similar) * dim;
            std::vector<T> r(rows, rows + sd);
            std::vector<T> e(eigenvalues, eigenvalues + rank);
            std::vector<T> b(basis, basis + static_cast<std::size_t>(rank) * dim);
            std::vector<T> scores;
            for (int m = 0; m < models; ++m) {
                accumulate_model_weights<T>(*params, final_stage, r.data(), similar, dim, e.data(), b.data(),
                                            rank, basis_similar, log_weights, scores);
            }
        };
        if (precision == 0) {
            run(0.0F);
        } else {
            run(0.0);
        }
        finalize_weights(*params, dim, basis_similar, models, log_weights, out, similar);
        std::copy(out.begin(), out.end(), weights);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int vo_aggregation_window(const vo_params* params, float* window) {
    if (params == nullptr || window == nullptr || params->patch_size <= 0) {
        return -1;
    }
    const std::vector<float> w = aggregation_window(*params);
    std::copy(w.begin(), w.end(), window);
    return 0;
}

int vo_grid_mask(const vo_params* params, int width, int height, int total_frames, int anchor_frame,
                 unsigned char* mask) {
    if (params == nullptr || mask == nullptr || width < params->patch_size || height < params->patch_size ||
        total_frames < params->patch_time) {
        return -1;
    }
    const Geometry g{width, height, total_frames, 1, 0, total_frames};
    int count = 0;
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            const bool inside = x <= width - params->patch_size && y <= height - params->patch_size;
            const bool on = inside && grid_selected(*params, g, anchor_frame, x, y);
            mask[static_cast<std::size_t>(y) * width + x] = on ? 1 : 0;
            count += on ? 1 : 0;
        }
    }
    return count;
}

int vo_grid_count(const vo_params* params, int width, int height, int frames,
                  int anchor_frame) {
    if (params == nullptr) {
        return -1;
    }
    const Geometry g{width, height, frames, 1};
    if (anchor_frame < 0 || anchor_frame > frames - params->patch_time) {
        return 0;
    }
    int count = 0;
    for (int y = 0; y <= height - params->patch_size; ++y) {
        for (int x = 0; x <= width - params->patch_size; ++x) {
            count += grid_selected(*params, g, anchor_frame, x, y) ? 1 : 0;
        }
    }
    return count;
}

void vo_set_search_offsets(const int16_t* offsets, int temporal) {
    g_offsets = offsets;
    g_offsets_temporal = offsets != nullptr ? temporal : 0;
}

void vo_set_max_sweeps(int max_sweeps) { g_max_sweeps = max_sweeps > 0 ? max_sweeps : 64; }

namespace {

// MVToolsVectorGrid::nearest_block_index (src/flow/flow.hpp:208-222): the
// block whose centre first_center + i * step is nearest, halves rounded
// away from the first block, clamped to the grid.
int nearest_block(int coordinate, int first_center, int step, int count) {
    if (count <= 1) {
        return 0;
    }
    const int shifted = coordinate - first_center;
    const int index = shifted >= 0 ? (shifted + step / 2) / step : -((-shifted + step / 2) / step);
    return std::clamp(index, 0, count - 1);
}

// detail::round_subpixel_to_pixel (src/flow/flow.hpp:185-193): nearest
// whole pixel, halves away from zero.
int pel_to_pixel(int value, int pel) {
    if (pel <= 1) {
        return value;
    }
    return value >= 0 ? (value + pel / 2) / pel : -((-value + pel / 2) / pel);
}

// One step of the chain: the grid of `frame` moves the centre by the vector
// of the block nearest to the centre's patch middle
// (center_for's loop bodies, src/flow/flow.hpp:359-386).
void follow_vector(const vo_mv_shape& s, const int32_t* vectors, int frame, int patch_center,
                   int width, int height, int& x, int& y) {
    const int step_x = std::max(1, s.block_size_x - std::max(0, s.overlap_x));
    const int step_y = std::max(1, s.block_size_y - std::max(0, s.overlap_y));
    const int bx = nearest_block(x + patch_center, s.block_size_x / 2, step_x, s.block_count_x);
    const int by = nearest_block(y + patch_center, s.block_size_y / 2, step_y, s.block_count_y);
    const std::size_t at =
        ((static_cast<std::size_t>(frame) * s.block_count_y + by) * s.block_count_x + bx) * 2;
    x += pel_to_pixel(vectors[at], s.pel);
    y += pel_to_pixel(vectors[at + 1], s.pel);
    if (width > 0) {
        x = std::clamp(x, 0, width - 1);
    }
    if (height > 0) {
        y = std::clamp(y, 0, height - 1);
    }
}

}  // namespace

void vo_center_for(const vo_mv_shape* fw, const int32_t* fw_vectors, const unsigned char* fw_valid,
                   const vo_mv_shape* bw, const int32_t* bw_vectors, const unsigned char* bw_valid,
                   int grid_frames, int anchor_x, int anchor_y, int anchor_frame,
                   int candidate_frame, int patch_size, int width, int height, int* out_xy) {
    int x = anchor_x;
    int y = anchor_y;
    const int patch_center = std::max(0, patch_size / 2);
    if (candidate_frame < anchor_frame) {
        for (int frame = anchor_frame; frame > candidate_frame; --frame) {
            if (fw_vectors != nullptr && frame >= 0 && frame < grid_frames && fw_valid[frame] != 0) {
                follow_vector(*fw, fw_vectors, frame, patch_center, width, height, x, y);
            }
        }
    } else {
        for (int frame = anchor_frame; frame < candidate_frame; ++frame) {
            if (bw_vectors != nullptr && frame >= 0 && frame < grid_frames && bw_valid[frame] != 0) {
                follow_vector(*bw, bw_vectors, frame, patch_center, width, height, x, y);
            }
        }
    }
    out_xy[0] = x;
    out_xy[1] = y;
}

int vo_mvtools_offsets(const vo_params* params, const vo_mv_shape* fw, const int32_t* fw_vectors,
                       const unsigned char* fw_valid, const vo_mv_shape* bw,
                       const int32_t* bw_vectors, const unsigned char* bw_valid, int width,
                       int height, int frames, int16_t* offsets) {
    if (params == nullptr || offsets == nullptr || width <= 0 || height <= 0 || frames <= 0) {
        return -1;
    }
    const vo_params& p = *params;
    const Geometry g{width, height, frames, 1};
    const int temporal = p.search_bwd + p.search_fwd + 1;
    std::vector<int> schedule;
    for (int f = 0; f < frames; ++f) {
        int low = 0;
        int high = 0;
        temporal_range(p, g, f, low, high);
        schedule_frames(f, low, high, schedule);
        for (int ti = 0; ti < temporal; ++ti) {
            for (int y = 0; y < height; ++y) {
                for (int x = 0; x < width; ++x) {
                    int c[2] = {x, y};
                    if (ti < static_cast<int>(schedule.size())) {
                        vo_center_for(fw, fw_vectors, fw_valid, bw, bw_vectors, bw_valid, frames, x,
                                      y, f, schedule[ti], p.patch_size, width, height, c);
                    }
                    const std::size_t at =
                        ((static_cast<std::size_t>(f) * temporal + ti) * height + y) * width + x;
                    offsets[2 * at] = static_cast<int16_t>(c[0] - x);
                    offsets[2 * at + 1] = static_cast<int16_t>(c[1] - y);
                }
            }
        }
    }
    return 0;
}

int vo_anchor_stack(int stage, const vo_params* params, const float* noisy,
                    const float* ref, int width, int height, int frames,
                    int channels, int anchor_frame, float* packed, int flags) {
    return vo_anchor_stack_window(stage, params, noisy, ref, width, height, frames, channels, 0,
                                  frames, anchor_frame, packed, flags);
}

int vo_anchor_stack_window(int stage, const vo_params* params, const float* noisy,
                           const float* ref, int width, int height, int frames,
                           int channels, int first_frame, int total_frames, int anchor_frame,
                           float* packed, int flags) {
    try {
        require(params != nullptr && packed != nullptr, "null argument");
        require(stage == VO_STAGE_BASIC || stage == VO_STAGE_FINAL, "bad stage");
        require(anchor_frame >= 0 && anchor_frame < total_frames, "anchor frame out of range");
        const Geometry g{width, height, frames, channels, first_frame, total_frames};
        const StageSetup s = make_setup(stage, *params, noisy, ref, g, flags);
        const auto pack = [&](auto& stack) {
            const std::size_t pixels = static_cast<std::size_t>(width) * height;
            for (int c = 0; c < channels; ++c) {
                for (int slot = 0; slot < s.slots; ++slot) {
                    float* num = packed + (static_cast<std::size_t>(c) * s.slots + slot) * 2 * pixels;
                    float* wgt = num + pixels;
                    for (int y = 0; y < height; ++y) {
                        for (int x = 0; x < width; ++x) {
                            num[static_cast<std::size_t>(y) * width + x] =
                                static_cast<float>(stack.at(slot, c, x, y));
                            wgt[static_cast<std::size_t>(y) * width + x] =
                                static_cast<float>(stack.at(slot, channels, x, y));
                        }
                    }
                }
            }
        };
        const bool model_double = (flags & VO_MODEL_DOUBLE) != 0;
        if ((flags & VO_ACCUM_DOUBLE) != 0) {
            Stack<double> stack;
            if (model_double) {
                process_anchor_frame<double, double>(s, anchor_frame, stack, nullptr);
            } else {
                process_anchor_frame<float, double>(s, anchor_frame, stack, nullptr);
            }
            pack(stack);
        } else {
            Stack<float> stack;
            if (model_double) {
                process_anchor_frame<double, float>(s, anchor_frame, stack, nullptr);
            } else {
                process_anchor_frame<float, float>(s, anchor_frame, stack, nullptr);
            }
            pack(stack);
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int vo_stage(int stage, const vo_params* params, const float* noisy,
             const float* ref, int width, int height, int frames, int channels,
             float* out, int flags, int threads, int64_t* groups_out) {
    try {
        require(params != nullptr && out != nullptr, "null argument");
        require(stage == VO_STAGE_BASIC || stage == VO_STAGE_FINAL, "bad stage");
        const Geometry g{width, height, frames, channels};
        const StageSetup s = make_setup(stage, *params, noisy, ref, g, flags);
        long long groups = 0;
        const bool model_double = (flags & VO_MODEL_DOUBLE) != 0;
        const bool accum_double = (flags & VO_ACCUM_DOUBLE) != 0;
        if (model_double) {
            groups = accum_double ? run_stage<double, double>(s, out, threads)
                                  : run_stage<double, float>(s, out, threads);
        } else {
            groups = accum_double ? run_stage<float, double>(s, out, threads)
                                  : run_stage<float, float>(s, out, threads);
        }
        if (groups_out != nullptr) {
            *groups_out = groups;
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int vo_frame_debug(int stage, const vo_params* params, const float* noisy,
                   const float* ref, int width, int height, int frames,
                   int channels, int anchor_frame, int capacity_anchors,
                   int* num_anchors, int* anchor_x, int* anchor_y, int* counts,
                   int* positions, float* distances, unsigned char* active,
                   float* weights, float* eigenvalues, int* group_flags,
                   int flags, int gram_dim, double* grams) {
    try {
        require(params != nullptr, "null argument");
        const Geometry g{width, height, frames, channels};
        const StageSetup s = make_setup(stage, *params, noisy, ref, g, flags);
        FrameDebug debug;
        debug.capacity = capacity_anchors;
        debug.anchor_x = anchor_x;
        debug.anchor_y = anchor_y;
        debug.counts = counts;
        debug.positions = positions;
        debug.distances = distances;
        debug.active = active;
        debug.weights = weights;
        debug.eigenvalues = eigenvalues;
        debug.group_flags = group_flags;
        debug.gram_dim = gram_dim;
        debug.grams = grams;
        if ((flags & VO_MODEL_DOUBLE) != 0) {
            Stack<double> stack;
            process_anchor_frame<double, double>(s, anchor_frame, stack, &debug);
        } else {
            Stack<double> stack;
            process_anchor_frame<float, double>(s, anchor_frame, stack, &debug);
        }
        if (num_anchors != nullptr) {
            *num_anchors = debug.count;
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int vo_linalg_mean(const double* samples, int rows, int cols, double* mean) {
    if (samples == nullptr || mean == nullptr || rows <= 0 || cols <= 0) {
        return -1;
    }
    compute_mean(samples, rows, cols, mean);
    return 0;
}

int vo_linalg_gram_centered(const double* centered, int rows, int cols,
                            double scale, double* gram) {
    if (centered == nullptr || gram == nullptr || rows <= 0 || cols <= 0) {
        return -1;
    }
    compute_gram(centered, rows, cols, scale, gram);
    return 0;
}

int vo_linalg_covariance_centered(const double* centered, int rows, int cols,
                                  double scale, double* covariance) {
    if (centered == nullptr || covariance == nullptr || rows <= 0 || cols <= 0) {
        return -1;
    }
    compute_covariance(centered, rows, cols, scale, covariance);
    return 0;
}

int vo_linalg_topk_eigen(const double* matrix, int dim, int rank,
                         double* values, double* vectors, int* converged) {
    if (matrix == nullptr || values == nullptr || vectors == nullptr || dim <= 0 ||
        rank < 0 || rank > dim) {
        return -1;
    }
    std::vector<double> m;
    std::vector<double> v;
    std::vector<int> order;
    const EigenStatus status = topk_eigen(matrix, dim, rank, values, vectors, m, v, order);
    if (converged != nullptr) {
        *converged = status.converged ? 1 : 0;
    }
    return status.sweeps;
}

int vo_linalg_topk_eigen_f32(const float* matrix, int dim, int rank,
                             float* values, float* vectors, int* converged) {
    if (matrix == nullptr || values == nullptr || vectors == nullptr || dim <= 0 ||
        rank < 0 || rank > dim) {
        return -1;
    }
    std::vector<float> m;
    std::vector<float> v;
    std::vector<int> order;
    const EigenStatus status = topk_eigen(matrix, dim, rank, values, vectors, m, v, order);
    if (converged != nullptr) {
        *converged = status.converged ? 1 : 0;
    }
    return status.sweeps;
}

int vo_linalg_dual_basis(const double* centered, int rows, int cols,
                         const double* gram_values, const double* gram_vectors,
                         int rank, double* basis) {
    if (centered == nullptr || gram_values == nullptr || gram_vectors == nullptr ||
        basis == nullptr || rows <= 0 || cols <= 0 || rank < 0) {
        return -1;
    }
    dual_basis(centered, rows, cols, gram_values, gram_vectors, rank, basis);
    return 0;
}

}  // extern "C"
