// Host side of the vnlbhip C ABI: parameter validation, the staggered anchor
// grid, anchor-frame batching and the kernel sequence of one stage.
//
// Per batch of anchor frames (whole clip when it fits the budget):
//   match -> paste mask -> compact active anchors -> prepare -> jacobi ->
//   filter/scatter
// then one normalisation over all output frames.  The reference runs the
// same stage per anchor frame on the CPU (process_anchor_impl,
// src/core/core.cpp:1820-1906) or per chunk of groups on CUDA
// (StagePipeline::enqueue, src/cuda/pipeline.cu:554-667).
#include "vnlbhip.h"

#include "geometry.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace {

using vnlbhip::StageDesc;

struct InvalidArgument : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct DeviceFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct ModelFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void require(bool condition, const char* message) {
    if (!condition) {
        throw InvalidArgument(message);
    }
}

void check(hipError_t status, const char* what) {
    if (status != hipSuccess) {
        throw DeviceFailure(std::string(what) + ": " + hipGetErrorString(status));
    }
}

// A clip, or a window of it held in device memory: `frames` local frames
// starting at absolute frame `first_frame` of a clip of `total_frames`
// (VideoGeometry, src/core/video.hpp:19-47; 0 means the window is the clip).
struct Geometry {
    int width = 0;
    int height = 0;
    int frames = 0;
    int channels = 0;
    int first_frame = 0;
    int total_frames = 0;
    int source_frames() const { return total_frames > 0 ? total_frames : frames; }
};

// soft_similar_cap / model_similar_cap (src/core/core.cpp:139-172).
int capped_similar(int similar, float factor) {
    if (factor == 0.0F) {
        return INT_MAX;
    }
    const float max_factor = static_cast<float>(INT_MAX) / static_cast<float>(similar);
    if (factor >= max_factor) {
        return INT_MAX;
    }
    const float cap = std::ceil(static_cast<float>(similar) * factor);
    return std::max(similar, static_cast<int>(cap));
}

int retained_cap(const vnlbhip_params& p) {
    const long long raw = static_cast<long long>(p.search_window) * p.search_window *
                          (static_cast<long long>(p.search_bwd) + p.search_fwd + 1);
    return static_cast<int>(std::min<long long>(capped_similar(p.similar, p.similar_cap_factor), raw));
}

// validate_stage_parameters / validate_stage_configuration
// (src/core/core.cpp:1948-2024), same messages.
void validate_reference(const vnlbhip_params& p, const Geometry& g) {
    require(std::isfinite(p.sigma) && p.sigma > 0.0F, "sigma must be finite and positive");
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
    require(std::isfinite(p.beta) && p.beta > 0.0F, "beta must be finite and positive");
    require(std::isfinite(p.tau) && p.tau >= 0.0F, "tau must be finite and non-negative");
    require(std::isfinite(p.variance_threshold), "variance_threshold must be finite");
    require(std::isfinite(p.sigma_basic) && p.sigma_basic >= 0.0F,
            "sigma_basic must be finite and non-negative");
    require(std::isfinite(p.gamma) && p.gamma > 0.0F, "gamma must be finite and positive");
    require(std::isfinite(p.weight_alpha) && p.weight_alpha >= 0.0F,
            "weight_alpha must be finite and non-negative");
    require(std::isfinite(p.weight_beta) && p.weight_beta >= 0.0F,
            "weight_beta must be finite and non-negative");
    require(std::isfinite(p.weight_gamma) && p.weight_gamma >= 0.0F,
            "weight_gamma must be finite and non-negative");
    require(std::isfinite(p.weight_epsilon) && p.weight_epsilon > 0.0F,
            "weight_epsilon must be finite and positive");
    require(std::isfinite(p.membership_noise_floor) && p.membership_noise_floor > 0.0F,
            "membership_noise_floor must be finite and positive");
    require(p.proc_step >= 0, "block_step must be non-negative; use 0 for auto");
    require(g.width > 0 && g.height > 0 && g.frames > 0 && g.channels > 0,
            "video geometry must be non-empty");
    require(g.first_frame >= 0, "first frame must be non-negative");
    require(g.total_frames >= 0 && static_cast<long long>(g.source_frames()) >=
                                       static_cast<long long>(g.frames) + g.first_frame,
            "local frame window must fit inside the source frame count");
    require(p.patch_size <= g.width && p.patch_size <= g.height,
            "block_size must fit inside the frame");
    require(p.patch_time <= g.source_frames(), "patch_time must fit inside the clip");
    const long long estimator_dim = static_cast<long long>(p.patch_size) * p.patch_size *
                                    p.patch_time * (p.couple_channels != 0 ? g.channels : 1);
    require(estimator_dim <= INT_MAX, "estimator dimension overflows int");
    require(p.rank <= estimator_dim, "rank must not exceed the effective patch dimension");
    require(p.similar >= std::min<long long>(estimator_dim, p.rank),
            "group_size must cover the effective VNLB rank");
}

// prepare_aggregation_window (src/core/core.cpp:296-324).
void aggregation_window(const vnlbhip_params& p, float* window) {
    const int n = p.patch_size;
    for (int i = 0; i < n * n; ++i) {
        window[i] = 1.0F;
    }
    const float gamma = p.weight_gamma;
    if (gamma == 0.0F || n <= 1) {
        return;
    }
    const float center = 0.5F * static_cast<float>(n - 1);
    const float radius = std::max(center, 0.5F);
    float max_weight = 0.0F;
    for (int y = 0; y < n; ++y) {
        const float dy = (static_cast<float>(y) - center) / radius;
        for (int x = 0; x < n; ++x) {
            const float dx = (static_cast<float>(x) - center) / radius;
            const float raw = std::exp(-0.5F * ((dx * dx) + (dy * dy)));
            window[y * n + x] = raw;
            max_weight = std::max(max_weight, raw);
        }
    }
    const float inv_max = max_weight > 0.0F ? 1.0F / max_weight : 1.0F;
    for (int i = 0; i < n * n; ++i) {
        window[i] = std::pow(window[i] * inv_max, gamma);
    }
}

StageDesc make_desc(int stage, const vnlbhip_params& p, const Geometry& g, bool matching_only = false) {
    require(stage == VNLBHIP_STAGE_BASIC || stage == VNLBHIP_STAGE_FINAL,
            "stage must be 0 (Basic) or 1 (Final)");
    validate_reference(p, g);

    StageDesc d{};
    d.stage = stage;
    d.width = g.width;
    d.height = g.height;
    d.frames = g.frames;
    d.channels = g.channels;
    d.first_frame = g.first_frame;
    d.total_frames = g.source_frames();
    d.patch = p.patch_size;
    d.patch_time = p.patch_time;
    d.patch_area = p.patch_size * p.patch_size;
    d.patch_dim = d.patch_area * p.patch_time;
    d.dim = d.patch_dim * g.channels;
    d.coupled = (p.couple_channels != 0 && g.channels > 1) ? 1 : 0;
    d.models = d.coupled ? 1 : g.channels;
    d.model_dim = d.coupled ? d.dim : d.patch_dim;
    d.search_window = p.search_window;
    // A radius beyond the clip's last patch origin gives the same temporal
    // ranges as that origin (compute_temporal_range shifts the window back
    // inside), so the kernels see radii that cannot overflow.
    const int max_origin = g.source_frames() - p.patch_time;
    d.search_bwd = std::min(p.search_bwd, max_origin);
    d.search_fwd = std::min(p.search_fwd, max_origin);
    d.similar = p.similar;
    d.cap = retained_cap(p);
    d.soft_cap = capped_similar(p.similar, p.similar_cap_factor);
    d.model_cap = capped_similar(p.similar, p.model_cap_factor);
    d.rank = p.rank;
    d.slots = d.search_bwd + d.search_fwd + p.patch_time;
    d.step = p.proc_step > 0 ? p.proc_step : (p.patch_size <= 1 ? 1 : p.patch_size);
    d.step_magic = (1u << 18) / static_cast<uint32_t>(d.step) + 1u;
    require(d.step <= 64, "block_step above 64 is not supported");
    d.flat_areas = p.flat_areas != 0 ? 1 : 0;
    d.weights_enabled =
        (p.weight_alpha != 0.0F || p.weight_beta != 0.0F || p.weight_gamma != 0.0F) ? 1 : 0;
    d.windowed = p.weight_gamma != 0.0F ? 1 : 0;
    d.fused_width = (d.coupled && !d.windowed) ? (p.patch_size / vnlbhip::kLanes) * vnlbhip::kLanes : 0;
    d.max_rows = std::min(d.cap, d.model_cap);
    d.use_dual = (d.model_dim >= 64 && d.max_rows < d.model_dim) ? 1 : 0;
    d.jacobi_dim = d.use_dual ? d.max_rows : d.model_dim;
    d.max_sweeps = 64;
    d.row_chunk = std::max(d.max_rows, std::min(d.cap, 16));
    d.tau = p.tau;
    d.sigma = p.sigma;
    d.sigma_basic = p.sigma_basic;
    d.beta = p.beta;
    d.variance_threshold = p.variance_threshold;
    d.gamma = p.gamma;
    d.weight_alpha = p.weight_alpha;
    d.weight_beta = p.weight_beta;
    d.weight_gamma = p.weight_gamma;
    d.weight_epsilon = p.weight_epsilon;
    d.membership_noise_floor = p.membership_noise_floor;

    // Kernel limits of this build.
    require(g.width <= vnlbhip::kMaxSide && g.height <= vnlbhip::kMaxSide,
            "frame width and height must be at most 4096");
    require(g.frames <= vnlbhip::kMaxFrames,
            "at most 256 frames in device memory per stage call (pass longer clips as windows)");
    require(g.channels == 1 || g.channels == 3, "only GrayS and YUV444PS clips are supported");
    require(p.patch_size <= vnlbhip::kMaxPatch, "block_size above 16 is not supported");
    require(d.dim <= vnlbhip::kMaxDim, "patch dimension above 1024 is not supported");
    require(d.cap <= vnlbhip::kMaxCap, "retained group cap above 1024 is not supported");
    // Model limits do not bind a call that only matches (vnlbhip_match_lists).
    require(matching_only || d.rank <= vnlbhip::kMaxRank, "rank above 64 is not supported");
    require(matching_only || d.jacobi_dim <= 128,
            "eigen problems above 128x128 are not supported (covariance path with a large patch)");
    require(matching_only || d.max_rows <= vnlbhip::kMaxModelRows, "more than 128 model rows is not supported");
    // Candidates per anchor: the temporal range holds at most
    // min(bwd + fwd + 1, total_frames - Pt + 1) frames.
    const long long temporal = std::min<long long>(static_cast<long long>(d.search_bwd) + d.search_fwd + 1,
                                                   static_cast<long long>(max_origin) + 1);
    const long long candidates = static_cast<long long>(p.search_window) * p.search_window * temporal;
    require(candidates <= vnlbhip::kMaxCandidates,
            "search window times temporal slots above 32768 candidates is not supported");
    require(d.search_bwd + d.search_fwd + 1 <= vnlbhip::kMaxTemporal,
            "temporal search radius too large");
    require(matching_only || vnlbhip::prepare_shared_bytes(d) <= 96 * 1024,
            "model rows times patch dimension exceed the prepare kernel's LDS budget");
    require(matching_only || vnlbhip::filter_shared_bytes(d) <= 112 * 1024,
            "rank and model rows times patch dimension exceed the filter kernel's LDS budget");
    // (the budget includes the selection scratch, 3 x max(cap, 256) words)
    require(vnlbhip::match_shared_bytes(d) <= 108 * 1024,
            "search window too large for the matcher's LDS budget");
    aggregation_window(p, d.window);
    return d;
}

// reference_processing_mask_selected / initialize_frame_processing_mask
// (src/core/core.cpp:934-967): anchors in raster order plus, per row, the
// first grid index, the column phase and the anchor count.
struct Grid {
    std::vector<uint32_t> anchors;
    std::vector<int> offsets;  // per processed frame, into anchors
    std::vector<int> ids;      // anchor frame of each processed frame
    std::vector<int> rows;     // [frame][3][height]
};

// anchor_frame is local; the phases follow its absolute index and the
// clip's last patch origin, as in the reference.
void append_grid_frame(const StageDesc& d, int anchor_frame, Grid& grid) {
    const int step = d.step;
    const int max_x = d.width - d.patch;
    const int max_y = d.height - d.patch;
    const int max_frame = d.total_frames - d.patch_time;
    const int absolute_frame = anchor_frame + d.first_frame;
    const int phase_y = absolute_frame == max_frame ? 0 : absolute_frame;
    const size_t row_base = grid.rows.size();
    grid.rows.resize(row_base + static_cast<size_t>(3) * d.height, -1);
    int* first = grid.rows.data() + row_base;
    int* phase = first + d.height;
    int* count = phase + d.height;
    const int frame_begin = static_cast<int>(grid.anchors.size());
    grid.offsets.push_back(frame_begin);
    grid.ids.push_back(anchor_frame);
    for (int y = 0; y <= max_y; ++y) {
        if (y != 0 && y != max_y && y % step != phase_y % step) {
            count[y] = 0;
            continue;
        }
        const int phase_x = y == max_y ? 0 : phase_y + (y / step);
        const int ph = phase_x % step;
        first[y] = static_cast<int>(grid.anchors.size()) - frame_begin;
        phase[y] = ph;
        int n = 0;
        for (int x = 0; x <= max_x; ++x) {
            if (x == 0 || x == max_x || x % step == ph) {
                grid.anchors.push_back(vnlbhip::pack_position(x, y, anchor_frame));
                ++n;
            }
        }
        count[y] = n;
    }
    for (int y = max_y + 1; y < d.height; ++y) {
        count[y] = 0;
    }
}

// Anchor tiles of the MFMA matcher: `rows` consecutive grid rows x `cols`
// consecutive anchors of each row, of one frame (rows, cols <= 4; slot
// 4 * rr + ii, -1 where a row is shorter or the shape narrower: the matrix
// products take the 16 slots as their anchor rows and the matcher's
// epilogue takes each group of 4 slots as one grid row), and the largest
// union of the tiles' search windows.  The union per anchor is what the
// shape trades: 4 x 4 is smallest for 19-wide windows; wider windows may
// need fewer rows to stay within the kernel's limits.
struct Tiles {
    std::vector<int> slots;
    int max_cols = 0;
    int max_rows = 0;
};

struct TileShape {
    int rows;
    int cols;
};

// The shapes the autotuner tries, the default first.
constexpr TileShape kTileShapes[] = {{4, 4}, {3, 4}, {2, 4}, {4, 3}, {4, 2}, {1, 4}};
constexpr int kTileShapeCount = static_cast<int>(sizeof(kTileShapes) / sizeof(kTileShapes[0]));

Tiles build_tiles(const StageDesc& d, const Grid& grid, int height, TileShape shape) {
    Tiles tiles;
    const int rows = shape.rows;
    const int cols = shape.cols;
    const int frame_count = static_cast<int>(grid.ids.size());
    for (int fi = 0; fi < frame_count; ++fi) {
        const int frame_begin = grid.offsets[static_cast<size_t>(fi)];
        const int* first = grid.rows.data() + static_cast<size_t>(fi) * 3 * height;
        const int* count = first + 2 * height;
        std::vector<int> ys;
        for (int y = 0; y < height; ++y) {
            if (count[y] > 0) {
                ys.push_back(y);
            }
        }
        for (size_t r0 = 0; r0 < ys.size(); r0 += static_cast<size_t>(rows)) {
            int widest = 0;
            for (size_t rr = r0; rr < r0 + static_cast<size_t>(rows) && rr < ys.size(); ++rr) {
                widest = std::max(widest, count[ys[rr]]);
            }
            for (int i0 = 0; i0 < widest; i0 += cols) {
                int x0 = INT_MAX;
                int x1 = INT_MIN;
                int y0 = INT_MAX;
                int y1 = INT_MIN;
                for (int slot = 0; slot < 16; ++slot) {
                    const int rr = slot / 4;
                    const int ii = slot - rr * 4;
                    int index = -1;
                    const size_t row = r0 + static_cast<size_t>(rr);
                    if (rr < rows && ii < cols && row < ys.size() && i0 + ii < count[ys[row]]) {
                        index = frame_begin + first[ys[row]] + i0 + ii;
                        const uint32_t p = grid.anchors[static_cast<size_t>(index)];
                        int lo = 0;
                        int hi = 0;
                        vnlbhip::spatial_range(d, vnlbhip::position_x(p), d.width, lo, hi);
                        x0 = std::min(x0, lo);
                        x1 = std::max(x1, hi);
                        vnlbhip::spatial_range(d, vnlbhip::position_y(p), d.height, lo, hi);
                        y0 = std::min(y0, lo);
                        y1 = std::max(y1, hi);
                    }
                    tiles.slots.push_back(index);
                }
                tiles.max_cols = std::max(tiles.max_cols, x1 - x0 + 1);
                tiles.max_rows = std::max(tiles.max_rows, y1 - y0 + 1);
            }
        }
    }
    return tiles;
}

// One batch of anchor frames: its grid and, once the MFMA matcher has
// needed them, its anchor tiles.
struct BatchPlan {
    Grid grid;  // offsets end with the batch's anchor count
    int max_anchors = 0;
    bool tiles_built = false;
    TileShape tile_shape{4, 4};  // the shape requested when the tiles were built
    TileShape used_shape{4, 4};  // the shape built (2 x 4 where the requested one does not fit)
    Tiles tiles;
};

// Reallocation count of one context.  Captured stage graphs hold raw
// addresses and plan data, so a graph captured at an older generation is
// not replayed; each context counts its own buffers and plans.
// A stage call that must not allocate (vnlbhip_stage_scratch) carves every
// buffer, in call order, from the device's fixed scratch array instead.
struct Arena {
    unsigned char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
    bool active = false;
};

// Every device buffer of a context, so a scratch call can set the owned
// allocations aside (park) and bring them back afterwards (unpark).
struct BufferSlot {
    virtual ~BufferSlot() = default;
    virtual void park() = 0;
    virtual void drop_carved() = 0;
    virtual void unpark() = 0;
};

struct Generation {
    unsigned long long value = 0;   // bumped on any reallocation or plan rebuild
    int64_t reallocations = 0;      // device buffers (re)allocated so far
    Arena arena;
    std::vector<BufferSlot*> buffers;
};

template <class T> struct DeviceBuffer final : BufferSlot {
    T* ptr = nullptr;
    size_t capacity = 0;
    T* owned_ptr = nullptr;  // the owned allocation while parked
    size_t owned_capacity = 0;
    bool parked = false;
    Generation* generation;
    explicit DeviceBuffer(Generation* owner) : generation(owner) { owner->buffers.push_back(this); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() override {
        unpark();
        release();
    }
    void park() override {
        if (!parked) {
            owned_ptr = ptr;
            owned_capacity = capacity;
            ptr = nullptr;
            capacity = 0;
            parked = true;
        }
    }
    void drop_carved() override {
        if (parked) {
            ptr = nullptr;
            capacity = 0;
        }
    }
    void unpark() override {
        if (parked) {
            ptr = owned_ptr;
            capacity = owned_capacity;
            owned_ptr = nullptr;
            owned_capacity = 0;
            parked = false;
        }
    }
    void release() {
        if (ptr != nullptr && !parked) {
            (void)hipFree(ptr);
        }
        ptr = nullptr;
        capacity = 0;
    }
    T* reserve(size_t count) {
        if (count == 0) {
            count = 1;
        }
        Arena& arena = generation->arena;
        if (arena.active) {
            if (count > capacity) {
                // 256-byte aligned pieces of the scratch array, never freed
                // within the call (earlier pieces may still be read by
                // queued kernels).
                const size_t left = arena.size - arena.used;
                if (count > left / sizeof(T)) {
                    throw InvalidArgument(
                        "the stage needs more device memory than the scratch array of "
                        "vnlbhip_stage_scratch holds (VNLBHIP_SCRATCH_BYTES); use vnlbhip_stage");
                }
                const size_t bytes = (count * sizeof(T) + 255) & ~static_cast<size_t>(255);
                ptr = reinterpret_cast<T*>(arena.base + arena.used);
                capacity = count;
                arena.used += std::min(bytes, left);
            }
            return ptr;
        }
        if (count > capacity) {
            // Work queued on any stream may still use the old allocation.
            // vnlbhip_reserve sizes every buffer up front so that later
            // stage calls of that shape never get here.
            check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            release();
            check(hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T)), "hipMalloc");
            capacity = count;
            ++generation->value;
            ++generation->reallocations;
        }
        return ptr;
    }
};

struct DebugOutput {
    int capacity_anchors = 0;
    int* num_anchors = nullptr;
    int* anchor_x = nullptr;
    int* anchor_y = nullptr;
    int* counts = nullptr;
    int* positions = nullptr;
    float* distances = nullptr;
    unsigned char* active = nullptr;
    float* weights = nullptr;
    float* eigenvalues = nullptr;
    int* group_flags = nullptr;
    int gram_dim = 0;  // grams: [capacity_anchors][gram_dim][gram_dim]
    float* grams = nullptr;
};

}  // namespace

// Anchor-frame parts a batch is pipelined in (matching, paste scan, model):
// at most kMaxParts, kDefaultParts unless the autotuner chose another count.
constexpr int kMaxParts = 4;
constexpr int kDefaultParts = 2;

struct vnlbhip_context {
    int device = 0;
    Generation generation;
    std::string error;
    vnlbhip_stats stats{};
    DeviceBuffer<uint32_t> anchors{&generation};
    DeviceBuffer<int> offsets{&generation};
    DeviceBuffer<int> ids{&generation};
    DeviceBuffer<int> rows{&generation};
    DeviceBuffer<int> counts{&generation};
    DeviceBuffer<uint32_t> positions{&generation};
    DeviceBuffer<float> distances{&generation};
    DeviceBuffer<unsigned char> active{&generation};
    DeviceBuffer<int> group_list{&generation};
    DeviceBuffer<unsigned long long> paste_dep{&generation};
    DeviceBuffer<int> paste_target_counts{&generation};
    DeviceBuffer<int> paste_targets{&generation};
    DeviceBuffer<int> group_count{&generation};
    DeviceBuffer<int> group_flags{&generation};
    DeviceBuffer<int> problem_dims{&generation};
    DeviceBuffer<float> means{&generation};
    DeviceBuffer<float> problems{&generation};
    DeviceBuffer<float> eigenvalues{&generation};
    DeviceBuffer<float> vectors{&generation};
    DeviceBuffer<int> status{&generation};
    DeviceBuffer<int> refine_list{&generation};
    DeviceBuffer<int> work_counters{&generation};  // prepare, filter: per-XCD group queues
    // Search-centre offsets (device, [frames][temporal][H][W][2] int16) or
    // null: every window centred on its anchor.
    const int16_t* search_offsets = nullptr;
    int search_offsets_temporal = 0;
    int search_offsets_frames = 0;  // the field's clip shape, checked by every stage
    int search_offsets_height = 0;
    int search_offsets_width = 0;
    float refine_gap = 1e-3F;  // FAST form: relative eigen gap below which groups are re-solved in fp64
    bool fused_selection = true;  // FAST form: k-nearest selection inside the MFMA matcher
    // Learned candidate proposal (vnlbhip_set_proposer): the proposer, its
    // top_k, score margin and inference precision.
    const vnlbhip_proposer* proposer = nullptr;
    int proposer_top_k = 48;
    float proposer_margin = 0.0F;
    int proposer_precision = VNLBHIP_PRECISION_BF16;
    // Jacobi sweep limit (JacobiEigenOptions::max_sweeps, 64 by default) and
    // whether unconverged problems fail the stage (the CUDA plugin) or are
    // only counted (the CPU path).
    int max_sweeps = 64;
    bool fail_unconverged = false;
    DeviceBuffer<float> weights{&generation};
    DeviceBuffer<double> accumulators{&generation};
    DeviceBuffer<int> summary{&generation};
    DeviceBuffer<float> host_in{&generation};
    DeviceBuffer<float> host_ref{&generation};
    DeviceBuffer<float> host_out{&generation};
    size_t batch_entry_budget = static_cast<size_t>(1) << 28;  // anchors * cap
    size_t fast_stage_budget = static_cast<size_t>(1) << 30;  // anchors * window slots
    int form = VNLBHIP_FORM_PLAIN;
    DeviceBuffer<float> fast_offsets{&generation};
    DeviceBuffer<float> fast_split{&generation};  // packed bf16 hi / lo planes (row sums on the way)
    DeviceBuffer<float> fast_energy{&generation};
    DeviceBuffer<int> fast_tiles{&generation};
    DeviceBuffer<float> fast_staged{&generation};
    std::vector<long long> plan_key;  // see run_stage
    std::vector<BatchPlan> plans;
    bool profiling = false;
    vnlbhip_kernel_times times{};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    // Side stream of the paste scans (default priority), its events and
    // pinned host words (kMaxParts group counts, then the 8-word summary);
    // created on first use.
    hipStream_t side = nullptr;
    hipEvent_t ready_event = nullptr;
    hipEvent_t match_event[kMaxParts] = {};
    hipEvent_t paste_event[kMaxParts] = {};
    int parts = kDefaultParts;  // pipeline parts of the next stage
    TileShape tile_shape{4, 4};        // MFMA matcher anchor tiles of a vnlbhip_stage call
    TileShape fixed_tile_shape{4, 4};  // vnlbhip_set_tile_shape: every call without autotuning
    bool stage_shape = false;          // run_stage takes tile_shape (vnlbhip_stage) not fixed_tile_shape
    bool autotune = false;
    // Stage graphs: the launch sequence of a stage captured once and
    // replayed (same stage, form, parts, geometry, parameters and pointers).
    struct StageGraph {
        hipGraphExec_t exec = nullptr;
        unsigned long long generation = 0;
        vnlbhip_stats stats{};
    };
    bool use_graph = false;
    int64_t graph_replays = 0;
    bool capturing = false;  // run_stage enqueues only (no uploads, no read-back)
    bool sizing = false;     // run_stage only sizes the buffers (vnlbhip_reserve)
    hipStream_t capture = nullptr;
    std::map<std::vector<long long>, StageGraph> graphs;
    // (stage, form, geometry, parameters) -> the part count timed fastest
    struct Tuned {
        int parts;
        TileShape shape;
    };
    std::map<std::vector<long long>, Tuned> tuned;
    int* host_counts = nullptr;
    // This device's copy of the scratch array (vnlbhip_stage_scratch), from
    // hipGetSymbolAddress at set-up.
    void* scratch = nullptr;
    ~vnlbhip_context() {
        for (auto& entry : graphs) {
            if (entry.second.exec != nullptr) {
                (void)hipGraphExecDestroy(entry.second.exec);
            }
        }
        if (capture != nullptr) {
            (void)hipStreamDestroy(capture);
        }
        if (host_counts != nullptr) {
            (void)hipHostFree(host_counts);
        }
        for (hipEvent_t e : match_event) {
            if (e != nullptr) {
                (void)hipEventDestroy(e);
            }
        }
        for (hipEvent_t e : paste_event) {
            if (e != nullptr) {
                (void)hipEventDestroy(e);
            }
        }
        if (ready_event != nullptr) {
            (void)hipEventDestroy(ready_event);
        }
        if (side != nullptr) {
            (void)hipStreamDestroy(side);
        }
        for (auto& p : pending) {
            event_pool.push_back(p.second.first);
            event_pool.push_back(p.second.second);
        }
        for (hipEvent_t e : event_pool) {
            (void)hipEventDestroy(e);
        }
    }
};

// The learned candidate proposer: fp32 master parameters, Adam moments, the
// bf16 copies the matrix products read, the FP8 inference copies with their
// scales, and the training workspace.
struct vnlbhip_proposer {
    int device = 0;
    std::string error;
    float* params = nullptr;   // kPropParams
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    __bf16* w1 = nullptr;      // [256][128]
    __bf16* w2 = nullptr;      // [256][256]
    uint8_t* w1_fp8 = nullptr;
    uint8_t* w2_fp8 = nullptr;
    float* fp8_scales = nullptr;  // kPropScales (proposer_types.h)
    bool fp8_ready = false;
    float* workspace = nullptr;
    size_t workspace_floats = 0;
    float* loss_scratch = nullptr;  // train_step without a loss buffer
    int loss_capacity = 0;
    int step = 0;
    ~vnlbhip_proposer() {
        (void)hipSetDevice(device);
        for (void* p : {static_cast<void*>(params), static_cast<void*>(grads), static_cast<void*>(adam_m),
                        static_cast<void*>(adam_v), static_cast<void*>(w1), static_cast<void*>(w2),
                        static_cast<void*>(w1_fp8), static_cast<void*>(w2_fp8), static_cast<void*>(fp8_scales),
                        static_cast<void*>(workspace), static_cast<void*>(loss_scratch)}) {
            if (p != nullptr) {
                (void)hipFree(p);
            }
        }
    }
    vnlbhip::ProposerWeights weights() const {
        vnlbhip::ProposerWeights wt{};
        wt.w1 = w1;
        wt.b1 = params + vnlbhip::kOffB1;
        wt.w2 = w2;
        wt.b2 = params + vnlbhip::kOffB2;
        wt.w3 = params + vnlbhip::kOffW3;
        wt.b3 = params + vnlbhip::kOffB3;
        wt.fp8.w1 = w1_fp8;
        wt.fp8.w2 = w2_fp8;
        wt.fp8.scales = fp8_scales;
        return wt;
    }
};

namespace vnlbhip {
hipError_t launch_summarize(const int* flags, const int* status, const int* counts,
                            const int* group_list, const int* groups_dev, int group_capacity,
                            int models, int* summary, hipStream_t stream);
}

namespace {

// The side stream of the paste scans, its events, the pinned host words and
// this device's scratch array: set up once, when the context is created, so
// that no stage call creates or allocates them.
void set_up_context(vnlbhip_context& ctx) {
    check(hipStreamCreateWithFlags(&ctx.side, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreateWithFlags(&ctx.ready_event, hipEventDisableTiming), "hipEventCreate");
    for (int h = 0; h < kMaxParts; ++h) {
        check(hipEventCreateWithFlags(&ctx.match_event[h], hipEventDisableTiming), "hipEventCreate");
        check(hipEventCreateWithFlags(&ctx.paste_event[h], hipEventDisableTiming), "hipEventCreate");
    }
    check(hipHostMalloc(reinterpret_cast<void**>(&ctx.host_counts), (kMaxParts + 8) * sizeof(int)),
          "hipHostMalloc");
    check(vnlbhip::scratch_address(ctx.device, &ctx.scratch), "hipGetSymbolAddress (scratch)");
    require(vnlbhip::scratch_bytes() == VNLBHIP_SCRATCH_BYTES, "scratch array size mismatch");
}

hipEvent_t take_event(vnlbhip_context& ctx) {
    if (!ctx.event_pool.empty()) {
        hipEvent_t e = ctx.event_pool.back();
        ctx.event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

// Records events around one launch when profiling is on.
template <class Fn>
void timed(vnlbhip_context& ctx, int kernel, hipStream_t stream, const char* what, Fn&& launch) {
    if (ctx.sizing) {
        return;
    }
    if (!ctx.profiling) {
        check(launch(), what);
        return;
    }
    hipEvent_t start = take_event(ctx);
    hipEvent_t stop = take_event(ctx);
    check(hipEventRecord(start, stream), "hipEventRecord");
    check(launch(), what);
    check(hipEventRecord(stop, stream), "hipEventRecord");
    ctx.pending.push_back({kernel, {start, stop}});
}

// A launch outside the per-kernel timings (fills, compaction, summary).
template <class Fn> void untimed(vnlbhip_context& ctx, const char* what, Fn&& launch) {
    if (!ctx.sizing) {
        check(launch(), what);
    }
}

// Folds completed launch timings into the totals (stream synchronised).
void collect_times(vnlbhip_context& ctx) {
    for (auto& p : ctx.pending) {
        float ms = 0.0F;
        check(hipEventElapsedTime(&ms, p.second.first, p.second.second), "hipEventElapsedTime");
        ctx.times.milliseconds[p.first] += ms;
        ctx.times.launches[p.first] += 1;
        ctx.event_pool.push_back(p.second.first);
        ctx.event_pool.push_back(p.second.second);
    }
    ctx.pending.clear();
}

// Waits for the stage and folds the device summary into the stats: model
// failures, unconverged problems, retained samples, fp64 re-solves, groups.
void finish_stage(vnlbhip_context& ctx, hipStream_t stream) {
    int* host_summary = ctx.host_counts + kMaxParts;  // pinned, 8 ints
    check(hipMemcpyAsync(host_summary, ctx.summary.ptr, 8 * sizeof(int), hipMemcpyDeviceToHost, stream),
          "read summary");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    collect_times(ctx);
    ctx.stats.unconverged_groups = host_summary[1];
    int64_t retained = 0;
    std::memcpy(&retained, host_summary + 2, sizeof(retained));
    ctx.stats.retained = retained;
    ctx.stats.refined = host_summary[4];
    ctx.stats.groups = host_summary[5];
    ctx.stats.exhaustive_fallbacks = host_summary[6];
    if (host_summary[0] != 0) {
        throw ModelFailure("group_size must cover the effective VNLB rank");
    }
    if (ctx.fail_unconverged && host_summary[1] != 0) {
        // The CUDA plugin's solver status check (src/vs/cuda_plugin.cpp:1831-1846).
        throw ModelFailure(("the eigen solver did not converge within " + std::to_string(ctx.max_sweeps) +
                            " sweeps for " + std::to_string(host_summary[1]) + " problems")
                               .c_str());
    }
}

// Match lists of one anchor frame copied out of the stage (vnlbhip_match_lists).
struct ListsOut {
    int capacity = 0;
    int* anchor_count = nullptr;  // host
    uint32_t* anchors = nullptr;  // device
    int* counts = nullptr;
    uint32_t* positions = nullptr;
    float* distances = nullptr;
};

// only_frame: one anchor frame (local index) or kAllFrames; stack (with
// only_frame): that anchor frame's contribution stack, packed; lists (with
// only_frame): matching only, the lists copied out.
constexpr int kAllFrames = INT_MIN;

void run_stage(vnlbhip_context& ctx, int stage, const vnlbhip_params& p, const float* noisy,
               const float* ref, float* out, const Geometry& g, hipStream_t stream,
               int only_frame, DebugOutput* debug, float* stack = nullptr, ListsOut* lists = nullptr) {
    const StageDesc d = [&]() {
        StageDesc desc = make_desc(stage, p, g, lists != nullptr);
        desc.max_sweeps = ctx.max_sweeps;
        return desc;
    }();
    require(ctx.sizing || noisy != nullptr, "video data pointer must not be null");
    if (stage == VNLBHIP_STAGE_FINAL) {
        require(ctx.sizing || ref != nullptr, "Final needs the ref (basic) clip");
    }
    check(hipSetDevice(ctx.device), "hipSetDevice");
    ctx.stats = vnlbhip_stats{};
    ctx.stats.graph_replays = ctx.graph_replays;
    const int64_t reallocations_before = ctx.generation.reallocations;

    // Anchor frames (local indices): only_frame, or every frame of the
    // window; frames past the clip's last patch origin have empty stacks
    // (process_anchor_impl, src/core/core.cpp:1839-1843).
    const int last_anchor_frame = g.source_frames() - p.patch_time - g.first_frame;
    std::vector<int> frames;
    if (only_frame != kAllFrames) {
        if (only_frame + g.first_frame >= 0 && only_frame <= last_anchor_frame) {
            frames.push_back(only_frame);
        }
    } else {
        for (int f = 0; f <= std::min(last_anchor_frame, g.frames - 1); ++f) {
            frames.push_back(f);
        }
    }
    // Each anchor frame reads the frames of its temporal range
    // (input_frame_range_for_anchor, src/core/core.cpp:2026-2034): the
    // window must hold them.
    for (const int f : frames) {
        int low = 0;
        int high = 0;
        vnlbhip::temporal_range(d, f, low, high);
        require(low >= 0 && high + p.patch_time - 1 < g.frames,
                "the frame window does not hold the anchor frame's input frames");
    }

    const size_t plane = static_cast<size_t>(g.width) * g.height;
    const size_t acc_count = plane * (g.channels + 1) * g.frames;
    // The FAST filter aggregates into fp32 accumulators (same layout, in
    // the same buffer); every other path into fp64 ones.
    const bool acc_f32 = ctx.form == VNLBHIP_FORM_FAST && vnlbhip::fast_filter_supported(d) &&
                         ctx.search_offsets == nullptr;
    double* acc = ctx.accumulators.reserve(acc_count);
    float* acc32 = reinterpret_cast<float*>(acc);
    untimed(ctx, "fill accumulators", [&]() { return vnlbhip::launch_fill_zero(acc, acc_f32 ? (acc_count + 1) / 2 : acc_count, stream); });
    int* summary = ctx.summary.reserve(8);
    untimed(ctx, "clear summary", [&]() { return vnlbhip::launch_fill_zero_int(summary, 8, stream); });

    const float* source = stage == VNLBHIP_STAGE_FINAL ? ref : noisy;
    // MFMA matcher: patch energies of the matched source, once per stage.
    float* fast_offsets = nullptr;
    unsigned* fast_flags = nullptr;
    float* fast_energy = nullptr;
    uint32_t* fast_split = nullptr;
    const bool guided = ctx.search_offsets != nullptr;
    if (guided) {
        require(static_cast<long long>(ctx.search_offsets_temporal) ==
                    static_cast<long long>(p.search_bwd) + p.search_fwd + 1 &&
                    d.search_bwd == p.search_bwd && d.search_fwd == p.search_fwd,
                "search offsets were set for another temporal window (search_bwd + search_fwd + 1)");
        require(ctx.search_offsets_frames == g.frames && ctx.search_offsets_height == g.height &&
                    ctx.search_offsets_width == g.width,
                "search offsets were set for a clip of another shape");
        require(vnlbhip::match_shared_bytes(d, true) <= 152 * 1024,
                "search window too large for guided matching's LDS budget");
    }
    // Learned candidate proposal: the proposing matcher replaces the
    // matcher of either form.
    const bool learned = ctx.proposer != nullptr && d.similar > 1;
    if (learned) {
        require(!guided && vnlbhip::learned_match_supported(d),
                "the learned proposer needs 8x8 patches, patch_time 1 or 2 and unguided search windows");
        require(ctx.proposer_top_k >= d.similar, "the proposer's top_k must be at least group_size");
        require(ctx.proposer_precision == VNLBHIP_PRECISION_BF16 || ctx.proposer->fp8_ready,
                "fp8 proposals need a calibrated proposer (vnlbhip_proposer_calibrate)");
    }
    // The MFMA matcher's anchor tiles and the fast filter's aggregation tile
    // assume windows centred on the anchor: guided stages keep the plain
    // matcher and filter.
    if (ctx.form == VNLBHIP_FORM_FAST && d.patch == 8 && d.similar > 1 && !guided && !learned) {
        const int match_channels = stage == VNLBHIP_STAGE_FINAL ? g.channels : 1;
        // Channel offsets [0, 4), the energy flag at [4], the offsets'
        // partial sums (doubles) from [8].
        fast_offsets = ctx.fast_offsets.reserve(8 + 2 * vnlbhip::kFastPartialDoubles);
        fast_flags = reinterpret_cast<unsigned*>(fast_offsets + 4);
        double* fast_partials = reinterpret_cast<double*>(fast_offsets + 8);
        fast_split = reinterpret_cast<uint32_t*>(ctx.fast_split.reserve(plane * match_channels * g.frames));
        fast_energy = ctx.fast_energy.reserve(plane * g.frames);
        timed(ctx, VNLBHIP_KERNEL_ENERGY, stream, "energy kernels", [&]() {
            return vnlbhip::launch_fast_energy(d, source, fast_offsets, fast_flags, fast_energy, fast_split,
                                               fast_partials, stream);
        });
    }
    const int cap = d.cap;
    const int n_stride = d.jacobi_dim;
    // Anchor grids of the batches (and, lazily, their matcher tiles) depend
    // only on the geometry, the grid step, the search window, the cap and the
    // frame list: Basic and Final, and repeated calls, reuse them.
    const std::vector<long long> plan_key = [&]() {
        std::vector<long long> key = {g.width, g.height, g.frames, d.first_frame, d.total_frames, d.step,
                                      d.patch, d.patch_time, d.search_window, cap,
                                      static_cast<long long>(ctx.batch_entry_budget)};
        key.insert(key.end(), frames.begin(), frames.end());
        return key;
    }();
    if (ctx.plan_key != plan_key) {
        require(!ctx.capturing, "the anchor plan changed during capture");
        ++ctx.generation.value;
        ctx.plan_key.clear();
        ctx.plans.clear();
        size_t next = 0;
        while (next < frames.size()) {
            // Gather frames into a batch within the entry budget.
            BatchPlan plan;
            Grid& grid = plan.grid;
            while (next < frames.size()) {
                const size_t before = grid.anchors.size();
                append_grid_frame(d, frames[next], grid);
                const int added = static_cast<int>(grid.anchors.size() - before);
                const bool over = grid.anchors.size() * static_cast<size_t>(cap) > ctx.batch_entry_budget;
                if (over && grid.offsets.size() > 1) {
                    // Roll back this frame; it starts the next batch.
                    grid.anchors.resize(before);
                    grid.offsets.pop_back();
                    grid.ids.pop_back();
                    grid.rows.resize(grid.rows.size() - static_cast<size_t>(3) * g.height);
                    break;
                }
                plan.max_anchors = std::max(plan.max_anchors, added);
                ++next;
            }
            grid.offsets.push_back(static_cast<int>(grid.anchors.size()));
            ctx.plans.push_back(std::move(plan));
        }
        ctx.plan_key = plan_key;
    }
    require(ctx.side != nullptr, "context set-up incomplete");
    for (BatchPlan& plan : ctx.plans) {
        const Grid& grid = plan.grid;
        const int max_anchors = plan.max_anchors;
        const int frame_count = static_cast<int>(grid.offsets.size()) - 1;
        const int anchor_count = static_cast<int>(grid.anchors.size());
        ctx.stats.anchors += anchor_count;
        ++ctx.stats.batches;
        if (learned) {
            ctx.stats.proposed += anchor_count;
        }
        if (lists != nullptr) {
            require(ctx.plans.size() == 1 && anchor_count <= lists->capacity,
                    "the list buffers hold fewer anchors than the frame has");
        }

        uint32_t* d_anchors = ctx.anchors.reserve(grid.anchors.size());
        int* d_offsets = ctx.offsets.reserve(grid.offsets.size());
        int* d_ids = ctx.ids.reserve(grid.ids.size());
        int* d_rows = ctx.rows.reserve(grid.rows.size());
        if (!ctx.capturing && !ctx.sizing) {
            // (A captured stage is single-batch and reuses the tables the
            // eager run before the capture uploaded.)
            check(hipMemcpyAsync(d_anchors, grid.anchors.data(), grid.anchors.size() * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, stream), "upload anchors");
            check(hipMemcpyAsync(d_offsets, grid.offsets.data(), grid.offsets.size() * sizeof(int),
                                 hipMemcpyHostToDevice, stream), "upload offsets");
            check(hipMemcpyAsync(d_ids, grid.ids.data(), grid.ids.size() * sizeof(int),
                                 hipMemcpyHostToDevice, stream), "upload frame ids");
            check(hipMemcpyAsync(d_rows, grid.rows.data(), grid.rows.size() * sizeof(int),
                                 hipMemcpyHostToDevice, stream), "upload row tables");
        }

        int* d_counts = ctx.counts.reserve(anchor_count);
        uint32_t* d_positions = ctx.positions.reserve(static_cast<size_t>(anchor_count) * cap);
        float* d_distances = ctx.distances.reserve(static_cast<size_t>(anchor_count) * cap);
        // 64 bytes past the batch's anchors: scratch of the paste walk.
        unsigned char* d_active = ctx.active.reserve(static_cast<size_t>(anchor_count) + 64);
        int* d_list = ctx.group_list.reserve(anchor_count);
        int* d_group_count = ctx.group_count.reserve(kMaxParts);
        unsigned long long* d_dep = ctx.paste_dep.reserve(anchor_count);
        int* d_tcount = ctx.paste_target_counts.reserve(anchor_count);
        int* d_targets = ctx.paste_targets.reserve(vnlbhip::paste_targets_size(d, anchor_count));
        untimed(ctx, "clear paste masks", [&]() { return vnlbhip::launch_fill_zero(reinterpret_cast<double*>(d_dep), anchor_count, stream); });
        untimed(ctx, "clear group counts", [&]() { return vnlbhip::launch_fill_zero_int(d_group_count, kMaxParts, stream); });

        // The batch runs as parts of consecutive anchor frames: the serial
        // paste scan of a part (a few wavefronts) runs on the side stream
        // beside the next part's matching and the previous part's model, and
        // each part's model starts once its groups are listed.  Debug runs
        // keep one part.
        const int parts = debug == nullptr ? std::max(1, std::min(ctx.parts, frame_count)) : 1;
        ctx.stats.parts = std::max(ctx.stats.parts, parts);
        int part_frame[kMaxParts + 1];
        for (int h = 0; h <= parts; ++h) {
            part_frame[h] = frame_count * h / parts;
        }
        check(hipEventRecord(ctx.ready_event, stream), "hipEventRecord");
        check(hipStreamWaitEvent(ctx.side, ctx.ready_event, 0), "hipStreamWaitEvent");

        const Tiles* tiles = nullptr;
        int* d_tiles = nullptr;
        bool fast_matcher = false;
        if (fast_energy != nullptr) {
            const TileShape want = ctx.stage_shape ? ctx.tile_shape : ctx.fixed_tile_shape;
            if (!plan.tiles_built || plan.tile_shape.rows != want.rows || plan.tile_shape.cols != want.cols) {
                // The requested tile shape where the window union fits the
                // kernel, else 2 x 4.
                plan.tiles = build_tiles(d, grid, g.height, want);
                plan.used_shape = want;
                if (!vnlbhip::fast_match_supported(d, plan.tiles.max_cols, plan.tiles.max_rows)) {
                    plan.tiles = build_tiles(d, grid, g.height, TileShape{2, 4});
                    plan.used_shape = TileShape{2, 4};
                }
                plan.tiles_built = true;
                plan.tile_shape = want;
            }
            tiles = &plan.tiles;
            fast_matcher = vnlbhip::fast_match_supported(d, tiles->max_cols, tiles->max_rows);
            if (fast_matcher) {
                ctx.stats.tile_rows = plan.used_shape.rows;
                ctx.stats.tile_cols = plan.used_shape.cols;
                // The tile list lives in the cached plan, so the upload may
                // complete asynchronously.
                d_tiles = ctx.fast_tiles.reserve(tiles->slots.size());
                if (!ctx.capturing && !ctx.sizing) {
                    check(hipMemcpyAsync(d_tiles, tiles->slots.data(), tiles->slots.size() * sizeof(int),
                                         hipMemcpyHostToDevice, stream), "upload tiles");
                }
            }
        }
        const int tile_count = fast_matcher ? static_cast<int>(tiles->slots.size() / 16) : 0;
        // First tile of the frame that starts at anchor_begin: tiles hold
        // anchors of one frame and come in frame order, so "first anchor >=
        // anchor_begin" is false and then true along the tile list.
        const auto tile_at = [&](int anchor_begin) {
            int lo = 0;
            int hi = tile_count;
            while (lo < hi) {
                const int t = lo + (hi - lo) / 2;
                int first = -1;
                for (int slot = 0; slot < 16 && first < 0; ++slot) {
                    first = tiles->slots[static_cast<size_t>(t) * 16 + slot];
                }
                if (first >= anchor_begin) {
                    hi = t;
                } else {
                    lo = t + 1;
                }
            }
            return lo;
        };

        for (int h = 0; h < parts; ++h) {
            const int fb = part_frame[h];
            const int fe = part_frame[h + 1];
            const int ab = grid.offsets[static_cast<size_t>(fb)];
            const int ae = grid.offsets[static_cast<size_t>(fe)];
            // Matching of the part's anchors (stage stream).
            if (learned) {
                timed(ctx, VNLBHIP_KERNEL_MATCH, stream, "learned match kernel", [&]() {
                    return vnlbhip::launch_learned_match(
                        d, source, d_anchors + ab, ae - ab, d_counts + ab, d_positions + static_cast<size_t>(ab) * cap,
                        d_distances + static_cast<size_t>(ab) * cap, ctx.proposer->weights(), ctx.proposer_top_k,
                        ctx.proposer_margin, ctx.proposer_precision, summary + 6, stream);
                });
            } else if (fast_matcher && ctx.fused_selection &&
                vnlbhip::fast_match_fused(d, tiles->max_cols, tiles->max_rows)) {
                const int tb = tile_at(ab);
                const int te = tile_at(ae);
                timed(ctx, VNLBHIP_KERNEL_MATCH, stream, "fast match kernel", [&]() {
                    return vnlbhip::launch_fast_match(d, tiles->max_cols, tiles->max_rows, fast_split,
                                                      fast_flags, fast_energy, d_anchors,
                                                      d_tiles + static_cast<size_t>(tb) * 16, te - tb, 0,
                                                      nullptr, true, d_counts, d_positions, d_distances,
                                                      stream);
                });
            } else if (fast_matcher) {
                // Distances are staged for a chunk of anchor frames at a
                // time (tiles and anchors are both in frame order) and
                // selected before the next chunk reuses the buffer.
                const size_t stride = vnlbhip::fast_match_stride(d);
                const int chunk_anchors = static_cast<int>(std::max<size_t>(
                    static_cast<size_t>(max_anchors), ctx.fast_stage_budget / stride));
                float* staged = ctx.fast_staged.reserve(static_cast<size_t>(
                    std::min(chunk_anchors, anchor_count)) * stride);
                int frame = fb;
                while (frame < fe) {
                    const int anchor_begin = grid.offsets[static_cast<size_t>(frame)];
                    int frame_end = frame + 1;
                    while (frame_end < fe &&
                           grid.offsets[static_cast<size_t>(frame_end + 1)] - anchor_begin <= chunk_anchors) {
                        ++frame_end;
                    }
                    const int anchor_end = grid.offsets[static_cast<size_t>(frame_end)];
                    const int tb = tile_at(anchor_begin);
                    const int te = tile_at(anchor_end);
                    timed(ctx, VNLBHIP_KERNEL_MATCH, stream, "fast match kernel", [&]() {
                        return vnlbhip::launch_fast_match(d, tiles->max_cols, tiles->max_rows, fast_split,
                                                          fast_flags, fast_energy, d_anchors,
                                                          d_tiles + static_cast<size_t>(tb) * 16, te - tb,
                                                          anchor_begin, staged, false, nullptr, nullptr,
                                                          nullptr, stream);
                    });
                    timed(ctx, VNLBHIP_KERNEL_SELECT, stream, "select kernel", [&]() {
                        return vnlbhip::launch_fast_select(
                            d, staged, d_anchors + anchor_begin, anchor_end - anchor_begin,
                            d_counts + anchor_begin, d_positions + static_cast<size_t>(anchor_begin) * cap,
                            d_distances + static_cast<size_t>(anchor_begin) * cap, stream);
                    });
                    frame = frame_end;
                }
            } else {
                timed(ctx, VNLBHIP_KERNEL_MATCH, stream, "match kernel", [&]() {
                    return vnlbhip::launch_match(d, source, d_anchors + ab, ae - ab, d_counts + ab,
                                                 d_positions + static_cast<size_t>(ab) * cap,
                                                 d_distances + static_cast<size_t>(ab) * cap,
                                                 ctx.search_offsets, ctx.search_offsets_temporal, stream);
                });
            }
            if (lists != nullptr) {
                continue;  // matching only
            }
            check(hipEventRecord(ctx.match_event[h], stream), "hipEventRecord");
            // Paste mask and group list of the part (side stream).
            check(hipStreamWaitEvent(ctx.side, ctx.match_event[h], 0), "hipStreamWaitEvent");
            timed(ctx, VNLBHIP_KERNEL_PASTE, ctx.side, "paste kernel", [&]() {
                return vnlbhip::launch_paste(d, fe - fb, d_offsets + fb, d_ids + fb,
                                             d_rows + static_cast<size_t>(fb) * 3 * g.height, d_counts,
                                             d_positions, d_active, max_anchors, ae - ab, d_dep, d_tcount,
                                             d_targets, anchor_count, ctx.side);
            });
            untimed(ctx, "compact kernel", [&]() { return vnlbhip::launch_compact(d_active, ab, ae - ab, d_list + ab, d_group_count + h, ctx.side); });
            check(hipEventRecord(ctx.paste_event[h], ctx.side), "hipEventRecord");
        }

        // Model, filter and aggregation of one list of groups (stage stream).
        // The model buffers are shared by the parts in stream order; the
        // variables below keep the last part's for the debug read-back.
        const float* basic = stage == VNLBHIP_STAGE_FINAL ? ref : noisy;
        int groups = 0;
        int problems = 0;
        int* d_flags = nullptr;
        float* d_eigen = nullptr;
        float* d_weights = nullptr;
        int* d_dims_last = nullptr;
        float* d_problems_last = nullptr;
        if (debug != nullptr) {
            d_weights = ctx.weights.reserve(static_cast<size_t>(anchor_count) * cap);
        }
        // The group count of a part is read on the device (the compaction
        // wrote it); grids and buffers are sized by the largest part's anchor
        // count, so the host never waits for the paste scans.
        int part_capacity = 0;
        for (int h = 0; h < parts; ++h) {
            part_capacity = std::max(part_capacity, grid.offsets[static_cast<size_t>(part_frame[h + 1])] -
                                                        grid.offsets[static_cast<size_t>(part_frame[h])]);
        }
        const auto run_model = [&](int* d_part_list, const int* d_count, int capacity) {
            const int problem_capacity = capacity * d.models;
            d_flags = ctx.group_flags.reserve(part_capacity);
            int* d_dims = ctx.problem_dims.reserve(part_capacity);
            float* d_means = ctx.means.reserve(static_cast<size_t>(part_capacity) * 2 * d.dim);
            float* d_problems = ctx.problems.reserve(static_cast<size_t>(part_capacity) * d.models *
                                                     n_stride * n_stride);
            d_eigen = ctx.eigenvalues.reserve(static_cast<size_t>(part_capacity) * d.models *
                                              std::max(d.rank, 1));
            float* d_vectors = ctx.vectors.reserve(static_cast<size_t>(part_capacity) * d.models *
                                                   std::max(d.rank, 1) * n_stride);
            int* d_status = ctx.status.reserve(static_cast<size_t>(part_capacity) * d.models);
            d_dims_last = d_dims;
            d_problems_last = d_problems;
            // Per-XCD group queues: prepare's, then the filter's.
            int* d_work = ctx.work_counters.reserve(2 * vnlbhip::kXcdCounterWords);
            untimed(ctx, "clear work counters", [&]() {
                return vnlbhip::launch_fill_zero_int(d_work, 2 * vnlbhip::kXcdCounterWords, stream);
            });
            timed(ctx, VNLBHIP_KERNEL_PREPARE, stream, "prepare kernel", [&]() {
                return vnlbhip::launch_prepare(d, noisy, basic, d_anchors, d_counts, d_positions,
                                               d_part_list, d_count, capacity, n_stride, d_flags, d_dims,
                                               d_means, d_problems, ctx.form == VNLBHIP_FORM_FAST, d_work,
                                               stream);
            });
            if (d.rank > 0) {
                const bool parallel_jacobi =
                    ctx.form == VNLBHIP_FORM_FAST && vnlbhip::jacobi_parallel_supported(n_stride);
                int* d_refine = nullptr;
                if (parallel_jacobi && ctx.refine_gap > 0.0F) {
                    // refine_list[0] counts, the list follows.
                    d_refine = ctx.refine_list.reserve(static_cast<size_t>(part_capacity) * d.models + 1);
                    untimed(ctx, "clear refine count", [&]() { return vnlbhip::launch_fill_zero_int(d_refine, 1, stream); });
                }
                timed(ctx, VNLBHIP_KERNEL_JACOBI, stream, "jacobi kernel", [&]() {
                    if (parallel_jacobi) {
                        return vnlbhip::launch_jacobi_parallel(
                            d, d_count, problem_capacity, n_stride, d_flags, d_dims, d_problems, d_eigen,
                            d_vectors, d_status,
                            ctx.refine_gap, d_refine != nullptr ? d_refine + 1 : nullptr, d_refine,
                            summary + 4, stream);
                    }
                    return vnlbhip::launch_jacobi(d, d_count, problem_capacity, n_stride, d_flags, d_dims,
                                                  d_problems, d_eigen, d_vectors, d_status, stream);
                });
            }
            const bool fast_filter = acc_f32;
            timed(ctx, VNLBHIP_KERNEL_FILTER, stream, "filter kernel", [&]() {
                if (fast_filter) {
                    return vnlbhip::launch_fast_filter(d, noisy, basic, d_anchors, d_counts, d_positions,
                                                       d_part_list, d_count, capacity, n_stride, d_flags,
                                                       d_means, d_eigen, d_vectors, d_problems, acc32,
                                                       d_weights, d_work + vnlbhip::kXcdCounterWords, stream);
                }
                return vnlbhip::launch_filter(d, noisy, basic, d_anchors, d_counts, d_positions,
                                              d_part_list, d_count, capacity, n_stride, d_flags, d_dims,
                                              d_means, d_eigen, d_vectors, acc, d_weights,
                                              d_work + vnlbhip::kXcdCounterWords, stream);
            });
            untimed(ctx, "summarize kernel", [&]() { return vnlbhip::launch_summarize(d_flags, d.rank > 0 ? d_status : nullptr, d_counts,
                                            d_part_list, d_count, capacity, d.models, summary, stream); });
        };
        for (int h = 0; h < parts && lists == nullptr; ++h) {
            check(hipStreamWaitEvent(stream, ctx.paste_event[h], 0), "hipStreamWaitEvent");
            const int ab = grid.offsets[static_cast<size_t>(part_frame[h])];
            const int ae = grid.offsets[static_cast<size_t>(part_frame[h + 1])];
            run_model(d_list + ab, d_group_count + h, ae - ab);
        }
        if (lists != nullptr && !ctx.sizing) {
            *lists->anchor_count = anchor_count;
            check(hipMemcpyAsync(lists->anchors, d_anchors, static_cast<size_t>(anchor_count) * sizeof(uint32_t),
                                 hipMemcpyDeviceToDevice, stream), "copy anchors");
            check(hipMemcpyAsync(lists->counts, d_counts, static_cast<size_t>(anchor_count) * sizeof(int),
                                 hipMemcpyDeviceToDevice, stream), "copy counts");
            check(hipMemcpyAsync(lists->positions, d_positions,
                                 static_cast<size_t>(anchor_count) * cap * sizeof(uint32_t),
                                 hipMemcpyDeviceToDevice, stream), "copy positions");
            check(hipMemcpyAsync(lists->distances, d_distances, static_cast<size_t>(anchor_count) * cap * sizeof(float),
                                 hipMemcpyDeviceToDevice, stream), "copy distances");
        }
        if (debug != nullptr) {
            // One part: its group count, read back for the diagnostics.
            check(hipMemcpyAsync(ctx.host_counts, d_group_count, sizeof(int), hipMemcpyDeviceToHost, stream),
                  "read group count");
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            groups = ctx.host_counts[0];
            problems = groups * d.models;
            require(anchor_count <= debug->capacity_anchors, "debug capacity too small");
            std::vector<int> counts(anchor_count);
            std::vector<uint32_t> packed(static_cast<size_t>(anchor_count) * cap);
            std::vector<float> dist(static_cast<size_t>(anchor_count) * cap);
            std::vector<unsigned char> act(anchor_count);
            std::vector<float> wts(static_cast<size_t>(anchor_count) * cap);
            std::vector<int> list(groups);
            std::vector<int> flags(groups);
            std::vector<float> eig(static_cast<size_t>(problems) * std::max(d.rank, 1));
            std::vector<int> dims(groups);
            std::vector<float> gram_rows;
            check(hipMemcpy(counts.data(), d_counts, counts.size() * sizeof(int),
                            hipMemcpyDeviceToHost), "read counts");
            check(hipMemcpy(packed.data(), d_positions, packed.size() * sizeof(uint32_t),
                            hipMemcpyDeviceToHost), "read positions");
            check(hipMemcpy(dist.data(), d_distances, dist.size() * sizeof(float),
                            hipMemcpyDeviceToHost), "read distances");
            check(hipMemcpy(act.data(), d_active, act.size(), hipMemcpyDeviceToHost), "read active");
            check(hipMemcpy(wts.data(), d_weights, wts.size() * sizeof(float),
                            hipMemcpyDeviceToHost), "read weights");
            if (groups > 0) {
                check(hipMemcpy(list.data(), d_list, list.size() * sizeof(int),
                                hipMemcpyDeviceToHost), "read group list");
                check(hipMemcpy(flags.data(), d_flags, flags.size() * sizeof(int),
                                hipMemcpyDeviceToHost), "read group flags");
                if (d.rank > 0) {
                    check(hipMemcpy(eig.data(), d_eigen, eig.size() * sizeof(float),
                                    hipMemcpyDeviceToHost), "read eigenvalues");
                }
                if (debug->grams != nullptr && d.rank > 0) {
                    // The model kernel's Gram / covariance per problem,
                    // [problem][n_stride][n_stride]; the eigen solvers only
                    // read it.
                    gram_rows.resize(static_cast<size_t>(problems) * n_stride * n_stride);
                    check(hipMemcpy(dims.data(), d_dims_last, dims.size() * sizeof(int),
                                    hipMemcpyDeviceToHost), "read problem dims");
                    check(hipMemcpy(gram_rows.data(), d_problems_last, gram_rows.size() * sizeof(float),
                                    hipMemcpyDeviceToHost), "read Grams");
                }
            }
            if (debug->num_anchors != nullptr) {
                *debug->num_anchors = anchor_count;
            }
            for (int a = 0; a < anchor_count; ++a) {
                const uint32_t anchor = grid.anchors[static_cast<size_t>(a)];
                if (debug->anchor_x != nullptr) {
                    debug->anchor_x[a] = vnlbhip::position_x(anchor);
                }
                if (debug->anchor_y != nullptr) {
                    debug->anchor_y[a] = vnlbhip::position_y(anchor);
                }
                if (debug->counts != nullptr) {
                    debug->counts[a] = counts[static_cast<size_t>(a)];
                }
                for (int i = 0; i < counts[static_cast<size_t>(a)] && i < cap; ++i) {
                    const size_t index = static_cast<size_t>(a) * cap + i;
                    const uint32_t pk = packed[index];
                    if (debug->positions != nullptr) {
                        debug->positions[index] =
                            (vnlbhip::position_frame(pk) * g.height + vnlbhip::position_y(pk)) *
                                g.width +
                            vnlbhip::position_x(pk);
                    }
                    if (debug->distances != nullptr) {
                        debug->distances[index] = dist[index];
                    }
                    if (debug->weights != nullptr && act[static_cast<size_t>(a)] != 0) {
                        debug->weights[index] = wts[index];
                    }
                }
                if (debug->active != nullptr) {
                    debug->active[a] = act[static_cast<size_t>(a)];
                }
                if (debug->group_flags != nullptr) {
                    debug->group_flags[a] = 0;
                }
            }
            for (int gi = 0; gi < groups; ++gi) {
                const int a = list[static_cast<size_t>(gi)];
                if (debug->group_flags != nullptr) {
                    debug->group_flags[a] = flags[static_cast<size_t>(gi)] & 7;
                }
                const int n = gram_rows.empty() ? 0 : dims[static_cast<size_t>(gi)];
                if (debug->grams != nullptr && (flags[static_cast<size_t>(gi)] & 4) != 0 && n > 0 &&
                    n <= debug->gram_dim && n < d.model_dim) {
                    const float* src = gram_rows.data() + static_cast<size_t>(gi) * d.models * n_stride * n_stride;
                    for (int r = 0; r < n; ++r) {
                        for (int c = 0; c < n; ++c) {
                            debug->grams[(static_cast<size_t>(a) * debug->gram_dim + r) * debug->gram_dim + c] =
                                src[r * n_stride + c];
                        }
                    }
                }
                if (debug->eigenvalues != nullptr && (flags[static_cast<size_t>(gi)] & 4) != 0) {
                    for (int j = 0; j < d.rank; ++j) {
                        debug->eigenvalues[static_cast<size_t>(a) * d.rank + j] =
                            eig[static_cast<size_t>(gi) * d.models * d.rank + j];
                    }
                }
            }
        }
    }

    if (stack != nullptr) {
        // Slot s of the stack is output frame only_frame - search_bwd + s.
        const int slots = d.slots;
        const int first = only_frame - d.search_bwd;
        check(acc_f32 ? vnlbhip::launch_pack_stack(d, acc32, first, slots, stack, stream)
                      : vnlbhip::launch_pack_stack(d, acc, first, slots, stack, stream),
              "pack kernel");
    }
    if (out != nullptr) {
        timed(ctx, VNLBHIP_KERNEL_NORMALIZE, stream, "normalize kernel",
              [&]() {
                  return acc_f32 ? vnlbhip::launch_normalize(d, acc32, noisy, out, stream)
                                 : vnlbhip::launch_normalize(d, acc, noisy, out, stream);
              });
    }
    ctx.stats.reallocations = ctx.generation.reallocations - reallocations_before;
    if (ctx.capturing || ctx.sizing) {
        return;  // the summary is read after each replay; a sizing pass launches nothing
    }
    finish_stage(ctx, stream);
}

// Captures the launch sequence enqueue(stream) enqueues on an internal
// stream (the side stream joins through its events) and keeps the
// instantiated graph under key.  Nothing is allocated or uploaded while
// capturing: the eager run before it sized every buffer.
template <class Fn>
void capture_stage(vnlbhip_context& ctx, const std::vector<long long>& key, const vnlbhip_stats& stats,
                   Fn&& enqueue) {
    if (ctx.capture == nullptr) {
        check(hipStreamCreateWithFlags(&ctx.capture, hipStreamNonBlocking), "hipStreamCreate");
    }
    auto old = ctx.graphs.find(key);
    if (old != ctx.graphs.end()) {
        check(hipGraphExecDestroy(old->second.exec), "hipGraphExecDestroy");
        ctx.graphs.erase(old);
    }
    const unsigned long long generation = ctx.generation.value;
    check(hipStreamBeginCapture(ctx.capture, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    ctx.capturing = true;
    hipGraph_t graph = nullptr;
    try {
        enqueue(ctx.capture);
    } catch (...) {
        ctx.capturing = false;
        (void)hipStreamEndCapture(ctx.capture, &graph);
        if (graph != nullptr) {
            (void)hipGraphDestroy(graph);
        }
        throw;
    }
    ctx.capturing = false;
    check(hipStreamEndCapture(ctx.capture, &graph), "hipStreamEndCapture");
    require(generation == ctx.generation.value, "device buffers changed during capture");
    hipGraphExec_t exec = nullptr;
    const hipError_t status = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    check(status, "hipGraphInstantiate");
    ctx.graphs[key] = vnlbhip_context::StageGraph{exec, generation, stats};
}

// Context settings that change what a stage computes or which kernels it
// launches: the arithmetic form, the fp64 re-solve gap (a kernel argument),
// the selection placement and the search-offset field.  Keys of captured
// graphs, tuned part counts and cached stacks include them.
void append_settings(const vnlbhip_context& ctx, std::vector<long long>& key) {
    uint32_t gap_bits = 0;
    std::memcpy(&gap_bits, &ctx.refine_gap, sizeof(gap_bits));
    key.push_back(ctx.form);
    key.push_back(gap_bits);
    key.push_back(ctx.fused_selection ? 1 : 0);
    key.push_back(reinterpret_cast<long long>(ctx.search_offsets));
    key.push_back(ctx.search_offsets_temporal);
    uint32_t margin_bits = 0;
    std::memcpy(&margin_bits, &ctx.proposer_margin, sizeof(margin_bits));
    key.push_back(reinterpret_cast<long long>(ctx.proposer));
    key.push_back(ctx.proposer_top_k);
    key.push_back(margin_bits);
    key.push_back(ctx.proposer_precision);
    key.push_back(ctx.max_sweeps);
}

int fail(vnlbhip_context* ctx, int code, const char* message) {
    if (ctx != nullptr) {
        ctx->error = message;
    }
    return code;
}

template <class Fn> int guarded(vnlbhip_context* ctx, Fn&& fn) {
    try {
        fn();
        if (ctx != nullptr) {
            ctx->error.clear();
        }
        return VNLBHIP_OK;
    } catch (const InvalidArgument& e) {
        return fail(ctx, VNLBHIP_ERROR_INVALID, e.what());
    } catch (const ModelFailure& e) {
        return fail(ctx, VNLBHIP_ERROR_MODEL, e.what());
    } catch (const DeviceFailure& e) {
        return fail(ctx, VNLBHIP_ERROR_DEVICE, e.what());
    } catch (const std::exception& e) {
        return fail(ctx, VNLBHIP_ERROR_DEVICE, e.what());
    }
}

}  // namespace

extern "C" {

int vnlbhip_default_params(int stage, vnlbhip_params* out) {
    if (out == nullptr || (stage != VNLBHIP_STAGE_BASIC && stage != VNLBHIP_STAGE_FINAL)) {
        return VNLBHIP_ERROR_INVALID;
    }
    vnlbhip_params p{};
    p.sigma = 0.0F;
    p.patch_size = 8;
    p.patch_time = 1;
    p.search_window = 19;
    p.search_bwd = 1;
    p.search_fwd = 1;
    p.similar = 8;
    p.rank = 8;
    p.similar_cap_factor = 4.0F;
    p.model_cap_factor = 1.0F;
    p.beta = 1.0F;
    p.tau = 0.0F;
    p.variance_threshold = 1.1F;
    p.sigma_basic = 0.0F;
    p.gamma = 0.95F;
    p.weight_alpha = 0.75F;
    p.weight_beta = 0.35F;
    p.weight_gamma = 1.0F;
    p.weight_epsilon = 1.0e-6F;
    p.membership_noise_floor = 0.25F;
    p.proc_step = 8;
    p.flat_areas = 0;
    p.couple_channels = 1;
    if (stage == VNLBHIP_STAGE_FINAL) {
        p.tau = 400.0F / (255.0F * 255.0F);
        p.variance_threshold = 1.7F;
        p.flat_areas = 1;
        p.weight_alpha = 1.0F;
        p.weight_beta = 0.5F;
        p.weight_gamma = 1.0F;
    }
    *out = p;
    return VNLBHIP_OK;
}

int vnlbhip_validate(int stage, const vnlbhip_params* params, int width, int height, int frames,
                     int channels, char* message, size_t message_size) {
    try {
        if (params == nullptr) {
            throw InvalidArgument("params must not be null");
        }
        (void)make_desc(stage, *params, Geometry{width, height, frames, channels});
        if (message != nullptr && message_size > 0) {
            message[0] = '\0';
        }
        return VNLBHIP_OK;
    } catch (const std::exception& e) {
        if (message != nullptr && message_size > 0) {
            std::snprintf(message, message_size, "%s", e.what());
        }
        return VNLBHIP_ERROR_INVALID;
    }
}

int vnlbhip_anchor_grid(const vnlbhip_params* params, int width, int height, int total_frames, int anchor_frame,
                        int capacity, int* xs, int* ys) {
    if (params == nullptr || params->patch_size <= 0 || params->patch_time <= 0 || params->proc_step < 0 ||
        width < params->patch_size || height < params->patch_size || width > vnlbhip::kMaxSide ||
        height > vnlbhip::kMaxSide || total_frames < params->patch_time || anchor_frame < 0 ||
        anchor_frame >= total_frames || capacity < 0) {
        return VNLBHIP_ERROR_INVALID;
    }
    StageDesc d{};
    d.width = width;
    d.height = height;
    d.frames = 1;
    d.first_frame = anchor_frame;
    d.total_frames = total_frames;
    d.patch = params->patch_size;
    d.patch_time = params->patch_time;
    d.step = params->proc_step > 0 ? params->proc_step : (params->patch_size <= 1 ? 1 : params->patch_size);
    Grid grid;
    append_grid_frame(d, 0, grid);
    const int count = static_cast<int>(grid.anchors.size());
    for (int i = 0; i < count && i < capacity; ++i) {
        if (xs != nullptr) {
            xs[i] = vnlbhip::position_x(grid.anchors[static_cast<size_t>(i)]);
        }
        if (ys != nullptr) {
            ys[i] = vnlbhip::position_y(grid.anchors[static_cast<size_t>(i)]);
        }
    }
    return count;
}

int vnlbhip_aggregation_window(const vnlbhip_params* params, float* window) {
    if (params == nullptr || window == nullptr || params->patch_size <= 0 ||
        params->patch_size > vnlbhip::kMaxPatch) {
        return VNLBHIP_ERROR_INVALID;
    }
    aggregation_window(*params, window);
    return VNLBHIP_OK;
}

int vnlbhip_retained_cap(const vnlbhip_params* params) {
    if (params == nullptr || params->similar <= 0) {
        return VNLBHIP_ERROR_INVALID;
    }
    return retained_cap(*params);
}

int vnlbhip_create(int device, vnlbhip_context** out) {
    if (out == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        return VNLBHIP_ERROR_DEVICE;
    }
    if (device < 0 || device >= count) {
        return VNLBHIP_ERROR_INVALID;
    }
    if (hipSetDevice(device) != hipSuccess) {
        return VNLBHIP_ERROR_DEVICE;
    }
    auto* ctx = new (std::nothrow) vnlbhip_context();
    if (ctx == nullptr) {
        return VNLBHIP_ERROR_DEVICE;
    }
    ctx->device = device;
    try {
        set_up_context(*ctx);
    } catch (const std::exception&) {
        delete ctx;
        return VNLBHIP_ERROR_DEVICE;
    }
    *out = ctx;
    return VNLBHIP_OK;
}

void vnlbhip_destroy(vnlbhip_context* context) {
    if (context != nullptr) {
        (void)hipSetDevice(context->device);
        delete context;
    }
}

const char* vnlbhip_last_error(const vnlbhip_context* context) {
    return context == nullptr ? "null context" : context->error.c_str();
}

int vnlbhip_get_stats(const vnlbhip_context* context, vnlbhip_stats* out) {
    if (context == nullptr || out == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    *out = context->stats;
    return VNLBHIP_OK;
}

int vnlbhip_set_profiling(vnlbhip_context* context, int enable) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    context->profiling = enable != 0;
    context->times = vnlbhip_kernel_times{};
    return VNLBHIP_OK;
}

int vnlbhip_set_form(vnlbhip_context* context, int form) {
    if (context == nullptr || (form != VNLBHIP_FORM_PLAIN && form != VNLBHIP_FORM_FAST)) {
        return VNLBHIP_ERROR_INVALID;
    }
    context->form = form;
    return VNLBHIP_OK;
}

int vnlbhip_set_eigen_refine(vnlbhip_context* context, float relative_gap) {
    if (context == nullptr || !(relative_gap >= 0.0F) || !std::isfinite(relative_gap)) {
        return VNLBHIP_ERROR_INVALID;
    }
    context->refine_gap = relative_gap;
    return VNLBHIP_OK;
}

int vnlbhip_set_search_offsets(vnlbhip_context* context, const int16_t* offsets, int temporal, int frames,
                               int height, int width) {
    if (context == nullptr || (offsets != nullptr && (temporal <= 0 || frames <= 0 || height <= 0 || width <= 0))) {
        return VNLBHIP_ERROR_INVALID;
    }
    context->search_offsets = offsets;
    context->search_offsets_temporal = offsets != nullptr ? temporal : 0;
    context->search_offsets_frames = offsets != nullptr ? frames : 0;
    context->search_offsets_height = offsets != nullptr ? height : 0;
    context->search_offsets_width = offsets != nullptr ? width : 0;
    return VNLBHIP_OK;
}

int vnlbhip_mvtools_offsets(vnlbhip_context* context, const vnlbhip_params* params,
                            const vnlbhip_mv_shape* fw_shape, const int32_t* fw_vectors,
                            const uint8_t* fw_valid, const vnlbhip_mv_shape* bw_shape,
                            const int32_t* bw_vectors, const uint8_t* bw_valid, int16_t* offsets,
                            int width, int height, int frames, void* stream) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        require(params != nullptr && offsets != nullptr, "params and offsets must not be null");
        require(reinterpret_cast<uintptr_t>(offsets) % 4 == 0, "offsets must be 4-byte aligned");
        const Geometry g{width, height, frames, 1};
        const StageDesc d = make_desc(VNLBHIP_STAGE_BASIC, *params, g);
        require(d.search_bwd == params->search_bwd && d.search_fwd == params->search_fwd,
                "search offsets need search_bwd and search_fwd within frames - patch_time");
        // MVToolsVectorGrid::usable and parse_mvtools_analysis's checks
        // (src/flow/flow.hpp:176-181,196-258) on a clip's grid shape.
        const auto shape_of = [](const vnlbhip_mv_shape* in, const int32_t* vectors, const uint8_t* valid,
                                 const char* name) {
            vnlbhip::MvShape s{1, 1, 0, 0, 1, 1, 1};
            if (vectors == nullptr) {
                return s;
            }
            require(in != nullptr && valid != nullptr,
                    (std::string(name) + ": vectors need a shape and per-frame valid flags").c_str());
            require(in->block_size_x > 0 && in->block_size_y > 0 && in->pel > 0 && in->block_count_x > 0 &&
                        in->block_count_y > 0,
                    (std::string(name) + ": block sizes, block counts and pel must be positive").c_str());
            require(in->overlap_x >= 0 && in->overlap_y >= 0 && in->overlap_x < in->block_size_x &&
                        in->overlap_y < in->block_size_y,
                    (std::string(name) + ": MVTools overlap must be smaller than block size").c_str());
            s = vnlbhip::MvShape{in->block_size_x, in->block_size_y, in->overlap_x, in->overlap_y,
                                 in->block_count_x, in->block_count_y, in->pel};
            return s;
        };
        const vnlbhip::MvShape fw = shape_of(fw_shape, fw_vectors, fw_valid, "mvfw");
        const vnlbhip::MvShape bw = shape_of(bw_shape, bw_vectors, bw_valid, "mvbw");
        check(hipSetDevice(context->device), "hipSetDevice");
        check(vnlbhip::launch_mvtools_offsets(d, fw, fw_vectors, fw_valid, bw, bw_vectors, bw_valid, offsets,
                                              reinterpret_cast<hipStream_t>(stream)),
              "mvtools offsets kernel");
    });
}

int vnlbhip_set_graph(vnlbhip_context* context, int enable) {
    if (context == nullptr || (enable != 0 && enable != 1)) {
        return VNLBHIP_ERROR_INVALID;
    }
    context->use_graph = enable != 0;
    return VNLBHIP_OK;
}

int vnlbhip_set_autotune(vnlbhip_context* context, int enable) {
    if (context == nullptr || (enable != 0 && enable != 1)) {
        return VNLBHIP_ERROR_INVALID;
    }
    context->autotune = enable != 0;
    return VNLBHIP_OK;
}

int vnlbhip_set_tile_shape(vnlbhip_context* context, int rows, int cols) {
    if (context == nullptr || rows < 1 || cols < 1 || rows > 4 || cols > 4) {
        return VNLBHIP_ERROR_INVALID;
    }
    context->fixed_tile_shape = TileShape{rows, cols};
    return VNLBHIP_OK;
}

int vnlbhip_set_fused_selection(vnlbhip_context* context, int enable) {
    if (context == nullptr || (enable != 0 && enable != 1)) {
        return VNLBHIP_ERROR_INVALID;
    }
    context->fused_selection = enable != 0;
    return VNLBHIP_OK;
}

int vnlbhip_set_jacobi(vnlbhip_context* context, int max_sweeps, int fail_unconverged) {
    if (context == nullptr || max_sweeps < 1 || max_sweeps > 10000 || (fail_unconverged != 0 && fail_unconverged != 1)) {
        return VNLBHIP_ERROR_INVALID;
    }
    context->max_sweeps = max_sweeps;
    context->fail_unconverged = fail_unconverged != 0;
    return VNLBHIP_OK;
}

int vnlbhip_get_kernel_times(const vnlbhip_context* context, vnlbhip_kernel_times* out) {
    if (context == nullptr || out == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    *out = context->times;
    return VNLBHIP_OK;
}

int vnlbhip_stage(vnlbhip_context* context, int stage, const vnlbhip_params* params,
                  const float* noisy, const float* ref, float* out, int width, int height,
                  int frames, int channels, void* stream) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        require(params != nullptr && out != nullptr, "params and out must not be null");
        const Geometry g{width, height, frames, channels};
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        context->parts = kDefaultParts;
        context->tile_shape = context->fixed_tile_shape;
        context->stage_shape = true;
        struct ShapeScope {
            vnlbhip_context* ctx;
            ~ShapeScope() { ctx->stage_shape = false; }
        } shape_scope{context};
        if (context->autotune) {
            // Key: everything the stage's kernels and grids depend on.
            std::vector<long long> key = {stage, width, height, frames, channels,
                                          context->search_offsets != nullptr};
            append_settings(*context, key);
            const unsigned char* raw = reinterpret_cast<const unsigned char*>(params);
            key.insert(key.end(), raw, raw + sizeof(vnlbhip_params));
            auto it = context->tuned.find(key);
            if (it == context->tuned.end()) {
                // One untimed run (plans, buffers), then every MFMA-matcher
                // tile shape that fits the kernel (default part count), then
                // every part count with the fastest shape; the stage blocks
                // until done, so host time is device time plus the fixed
                // launch work.
                const bool profiling = context->profiling;
                context->profiling = false;
                run_stage(*context, stage, *params, noisy, ref, out, g, s, kAllFrames, nullptr);
                const auto timed_run = [&]() {
                    const auto t0 = std::chrono::steady_clock::now();
                    run_stage(*context, stage, *params, noisy, ref, out, g, s, kAllFrames, nullptr);
                    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                };
                vnlbhip_context::Tuned best{kDefaultParts, context->fixed_tile_shape};
                double best_ms = timed_run();
                const StageDesc d = make_desc(stage, *params, g);
                const bool mfma_matcher = context->form == VNLBHIP_FORM_FAST && d.patch == 8 && d.similar > 1 &&
                                          context->search_offsets == nullptr && context->proposer == nullptr &&
                                          !context->plans.empty();
                for (int i = 0; mfma_matcher && i < kTileShapeCount; ++i) {
                    const TileShape shape = kTileShapes[i];
                    if (shape.rows == best.shape.rows && shape.cols == best.shape.cols) {
                        continue;  // timed above
                    }
                    const Tiles t = build_tiles(d, context->plans.front().grid, height, shape);
                    if (!vnlbhip::fast_match_supported(d, t.max_cols, t.max_rows)) {
                        continue;
                    }
                    context->tile_shape = shape;
                    const double ms = timed_run();
                    if (ms < best_ms) {
                        best = vnlbhip_context::Tuned{kDefaultParts, shape};
                        best_ms = ms;
                    }
                }
                context->tile_shape = best.shape;
                for (int parts = 1; parts <= kMaxParts; ++parts) {
                    if (parts == kDefaultParts) {
                        continue;
                    }
                    context->parts = parts;
                    const double ms = timed_run();
                    if (ms < best_ms) {
                        best.parts = parts;
                        best_ms = ms;
                    }
                }
                context->profiling = profiling;
                it = context->tuned.emplace(key, best).first;
            }
            context->parts = it->second.parts;
            context->tile_shape = it->second.shape;
        }
        if (context->use_graph && !context->profiling) {
            // Key: the stage's launch sequence and every address it holds.
            std::vector<long long> key = {stage, context->parts, context->tile_shape.rows,
                                          context->tile_shape.cols, width, height, frames,
                                          channels, reinterpret_cast<long long>(noisy),
                                          reinterpret_cast<long long>(ref), reinterpret_cast<long long>(out)};
            append_settings(*context, key);
            const unsigned char* raw = reinterpret_cast<const unsigned char*>(params);
            key.insert(key.end(), raw, raw + sizeof(vnlbhip_params));
            auto it = context->graphs.find(key);
            if (it != context->graphs.end() && it->second.generation == context->generation.value) {
                check(hipSetDevice(context->device), "hipSetDevice");
                check(hipGraphLaunch(it->second.exec, s), "hipGraphLaunch");
                context->stats = it->second.stats;
                finish_stage(*context, s);
                context->stats.graph_replays = ++context->graph_replays;
                return;
            }
            // Eager run (the output), then the capture for the next calls:
            // single-batch stages only (one upload of the anchor tables).
            run_stage(*context, stage, *params, noisy, ref, out, g, s, kAllFrames, nullptr);
            if (context->plans.size() == 1) {
                const vnlbhip_stats stats = context->stats;
                capture_stage(*context, key, stats, [&](hipStream_t cs) {
                    run_stage(*context, stage, *params, noisy, ref, out, g, cs, kAllFrames, nullptr);
                });
                context->stats = stats;
            }
            return;
        }
        run_stage(*context, stage, *params, noisy, ref, out, g, s, kAllFrames, nullptr);
    });
}

}  // extern "C"

namespace {

// The context state of one vnlbhip_stage_scratch call: owned buffers set
// aside, every buffer carved from the scratch array, no graphs, no
// profiling events, no autotuning runs; all restored on exit (also when the
// call fails).
struct ScratchCall {
    vnlbhip_context& ctx;
    bool profiling;
    bool use_graph;
    bool autotune;
    explicit ScratchCall(vnlbhip_context& c)
        : ctx(c), profiling(c.profiling), use_graph(c.use_graph), autotune(c.autotune) {
        ctx.profiling = false;
        ctx.use_graph = false;
        ctx.autotune = false;
        for (BufferSlot* b : ctx.generation.buffers) {
            b->park();
        }
        Arena& arena = ctx.generation.arena;
        arena.base = static_cast<unsigned char*>(ctx.scratch);
        arena.size = VNLBHIP_SCRATCH_BYTES;
        arena.used = 0;
        arena.active = true;
    }
    // Back to an empty scratch array, for the pass after the sizing pass.
    void rewind() {
        for (BufferSlot* b : ctx.generation.buffers) {
            b->drop_carved();
        }
        ctx.generation.arena.used = 0;
    }
    ~ScratchCall() {
        ctx.sizing = false;
        ctx.generation.arena = Arena{};
        for (BufferSlot* b : ctx.generation.buffers) {
            b->unpark();
        }
        ctx.profiling = profiling;
        ctx.use_graph = use_graph;
        ctx.autotune = autotune;
    }
    ScratchCall(const ScratchCall&) = delete;
    ScratchCall& operator=(const ScratchCall&) = delete;
};

}  // namespace

extern "C" {

int vnlbhip_stage_scratch(vnlbhip_context* context, int stage, const vnlbhip_params* params,
                          const float* noisy, const float* ref, float* out, int width, int height,
                          int frames, int channels, void* stream) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        require(params != nullptr && out != nullptr, "params and out must not be null");
        require(context->scratch != nullptr, "no scratch array on this device");
        const Geometry g{width, height, frames, channels};
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        ScratchCall call(*context);
        context->parts = kDefaultParts;
        // Sizing pass: carves every buffer without launching anything, so a
        // stage too large for the array is refused before any kernel runs.
        context->sizing = true;
        run_stage(*context, stage, *params, noisy, ref, out, g, s, kAllFrames, nullptr);
        context->sizing = false;
        call.rewind();
        run_stage(*context, stage, *params, noisy, ref, out, g, s, kAllFrames, nullptr);
    });
}

}  // extern "C"

namespace {

// One anchor frame's stack over a window of the clip.  The stack layout
// has search_bwd + search_fwd + patch_time slots, so radii that reach past
// the clip (which the stage itself clamps) are refused here.
void anchor_stack_window(vnlbhip_context& ctx, int stage, const vnlbhip_params& p, const float* noisy,
                         const float* ref, const Geometry& g, int anchor_frame, float* stack, hipStream_t stream) {
    require(stack != nullptr, "stack must not be null");
    const StageDesc d = make_desc(stage, p, g);
    require(d.search_bwd == p.search_bwd && d.search_fwd == p.search_fwd,
            "contribution stacks need search_bwd and search_fwd within frames - patch_time");
    run_stage(ctx, stage, p, noisy, ref, nullptr, g, stream, anchor_frame - g.first_frame, nullptr, stack);
}

}  // namespace

extern "C" {

int vnlbhip_reserve(vnlbhip_context* context, int stage, const vnlbhip_params* params, int width, int height,
                    int frames, int channels) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        require(params != nullptr, "params must not be null");
        context->sizing = true;
        try {
            run_stage(*context, stage, *params, nullptr, nullptr, nullptr,
                      Geometry{width, height, frames, channels}, nullptr, kAllFrames, nullptr);
        } catch (...) {
            context->sizing = false;
            throw;
        }
        context->sizing = false;
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

int vnlbhip_anchor_stack(vnlbhip_context* context, int stage, const vnlbhip_params* params,
                         const float* noisy, const float* ref, int width, int height, int frames,
                         int channels, int anchor_frame, float* stack, void* stream) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        require(params != nullptr, "params must not be null");
        require(anchor_frame >= 0 && anchor_frame < frames, "anchor_frame must be a frame of the clip");
        anchor_stack_window(*context, stage, *params, noisy, ref, Geometry{width, height, frames, channels},
                            anchor_frame, stack, reinterpret_cast<hipStream_t>(stream));
    });
}

int vnlbhip_anchor_stack_window(vnlbhip_context* context, int stage, const vnlbhip_params* params,
                                const float* noisy, const float* ref, int width, int height, int frames,
                                int channels, int first_frame, int total_frames, int anchor_frame, float* stack,
                                void* stream) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        require(params != nullptr, "params must not be null");
        require(total_frames > 0, "total_frames must be positive");
        require(anchor_frame >= 0 && anchor_frame < total_frames, "anchor_frame must be a frame of the clip");
        anchor_stack_window(*context, stage, *params, noisy, ref,
                            Geometry{width, height, frames, channels, first_frame, total_frames}, anchor_frame,
                            stack, reinterpret_cast<hipStream_t>(stream));
    });
}

// Per-output-frame rendering over a device cache of anchor stacks (the
// frame-request front end of the VapourSynth plugins: vnlb.Basic /
// vnlb.Final produce one anchor stack per frame, src/vs/plugin.cpp:616-758,
// and vnlb.Aggregate sums the stacks of the anchors around an output frame,
// :966-1045).  Each anchor frame's stack is computed once while it stays
// cached; the least recently used stack outside the current output frame's
// anchors is replaced.  Renders of one cache are serialised by its mutex.
struct vnlbhip_frame_cache {
    vnlbhip_context* ctx = nullptr;
    int device = 0;                   // freeing needs no context
    int stage = 0;
    vnlbhip_params params{};
    int width = 0;
    int height = 0;
    int channels = 0;
    int total_frames = 0;             // frames of the clip
    StageDesc d{};
    size_t stack_floats = 0;
    std::vector<float*> stacks;
    std::vector<hipEvent_t> read_done;  // per stack: the last render that read it
    std::vector<int> anchor;          // -1: empty
    std::vector<uint64_t> used;
    uint64_t tick = 0;
    // What the cached stacks were computed from: whole-clip renders key on
    // the clip pointers; windowed renders on frame numbers alone (a clip's
    // frames do not change, as in VapourSynth).
    int mode = -1;                    // 0: whole clip, 1: windows
    const float* noisy = nullptr;
    const float* ref = nullptr;
    std::vector<long long> settings;  // append_settings of the context
    std::mutex mutex;
    vnlbhip_frame_cache_stats stats{};

    void invalidate() { std::fill(anchor.begin(), anchor.end(), -1); }

    ~vnlbhip_frame_cache() {
        if (!stacks.empty() || !read_done.empty()) {
            (void)hipSetDevice(device);
            (void)hipDeviceSynchronize();
            for (float* p : stacks) {
                (void)hipFree(p);
            }
            for (hipEvent_t e : read_done) {
                (void)hipEventDestroy(e);
            }
        }
    }
};

int vnlbhip_frame_cache_create(vnlbhip_context* context, int stage, const vnlbhip_params* params,
                               int width, int height, int frames, int channels, int capacity,
                               vnlbhip_frame_cache** out) {
    if (context == nullptr || out == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    *out = nullptr;
    return guarded(context, [&]() {
        require(params != nullptr, "params must not be null");
        require(frames > 0, "video geometry must be non-empty");
        // Every limit of a stage over a window of the clip.
        const Geometry g{width, height, std::min(frames, vnlbhip::kMaxFrames), channels, 0, frames};
        const StageDesc d = make_desc(stage, *params, g);
        const int needed = d.slots;
        require(capacity >= needed, "frame cache capacity must cover search_bwd + search_fwd + patch_time stacks");
        require(needed <= vnlbhip::kMaxStackSet, "more than 32 anchor stacks per output frame are not supported");
        check(hipSetDevice(context->device), "hipSetDevice");
        auto cache = std::make_unique<vnlbhip_frame_cache>();
        cache->ctx = context;
        cache->device = context->device;
        cache->stage = stage;
        cache->params = *params;
        cache->width = width;
        cache->height = height;
        cache->channels = channels;
        cache->total_frames = frames;
        cache->d = d;
        cache->stack_floats = static_cast<size_t>(width) * height * channels * d.slots * 2;
        for (int i = 0; i < capacity; ++i) {
            float* p = nullptr;
            check(hipMalloc(reinterpret_cast<void**>(&p), cache->stack_floats * sizeof(float)),
                  "hipMalloc frame cache stack");
            cache->stacks.push_back(p);
            hipEvent_t e = nullptr;
            check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            cache->read_done.push_back(e);
            cache->anchor.push_back(-1);
            cache->used.push_back(0);
        }
        *out = cache.release();
    });
}

namespace {

// Renders output frame `frame` from the cached stacks, computing the missing
// ones over the window [first_frame, first_frame + window_frames) that
// noisy / ref hold.
void render_frame(vnlbhip_frame_cache& cache, int frame, const float* noisy, const float* ref,
                  int first_frame, int window_frames, float* out, hipStream_t s, int mode) {
    vnlbhip_context& ctx = *cache.ctx;
    const StageDesc& d = cache.d;
    require(noisy != nullptr && out != nullptr, "noisy and out must not be null");
    require(frame >= 0 && frame < cache.total_frames, "frame must be a frame of the clip");
    require(first_frame >= 0 && window_frames > 0 &&
                static_cast<long long>(first_frame) + window_frames <= cache.total_frames,
            "the frame window must lie inside the clip");
    require(frame >= first_frame && frame < first_frame + window_frames,
            "the frame window must hold the output frame");
    const Geometry g{cache.width, cache.height, window_frames, cache.channels, first_frame, cache.total_frames};
    std::vector<long long> settings;
    append_settings(ctx, settings);
    if (mode != cache.mode || settings != cache.settings ||
        (mode == 0 && (noisy != cache.noisy || ref != cache.ref))) {
        // Another clip, arithmetic form, search field or re-solve gap:
        // nothing cached applies.
        cache.invalidate();
        cache.mode = mode;
        cache.settings = settings;
        cache.noisy = mode == 0 ? noisy : nullptr;
        cache.ref = mode == 0 ? ref : nullptr;
    }
    // Anchors whose stacks hold this frame (anchors past the last patch
    // origin have empty stacks).
    const int first = std::max(0, frame - (d.search_fwd + d.patch_time - 1));
    const int last = std::min(cache.total_frames - d.patch_time, frame + d.search_bwd);
    ++cache.tick;
    vnlbhip::StackSet set{};
    for (int a = first; a <= last; ++a) {
        int slot = -1;
        for (size_t i = 0; i < cache.anchor.size(); ++i) {
            if (cache.anchor[i] == a) {
                slot = static_cast<int>(i);
                break;
            }
        }
        if (slot >= 0) {
            ++cache.stats.hits;
        } else {
            // Least recently used entry not holding this frame's anchors.
            for (size_t i = 0; i < cache.anchor.size(); ++i) {
                if (cache.used[i] != cache.tick &&
                    (slot < 0 || cache.anchor[i] < 0 ||
                     (cache.anchor[slot] >= 0 && cache.used[i] < cache.used[slot]))) {
                    slot = static_cast<int>(i);
                    if (cache.anchor[i] < 0) {
                        break;
                    }
                }
            }
            require(slot >= 0, "frame cache has no free entry");
            if (cache.anchor[slot] >= 0) {
                ++cache.stats.evictions;
            }
            cache.anchor[slot] = -1;
            // A render on another stream may still read the old stack.
            check(hipStreamWaitEvent(s, cache.read_done[static_cast<size_t>(slot)], 0), "hipStreamWaitEvent");
            run_stage(ctx, cache.stage, cache.params, noisy, ref, nullptr, g, s, a - first_frame, nullptr,
                      cache.stacks[static_cast<size_t>(slot)]);
            cache.anchor[slot] = a;
            ++cache.stats.misses;
        }
        cache.used[slot] = cache.tick;
        set.stack[set.count] = cache.stacks[static_cast<size_t>(slot)];
        set.slot[set.count] = frame - a + d.search_bwd;
        ++set.count;
    }
    const size_t plane = static_cast<size_t>(cache.width) * cache.height * cache.channels;
    check(vnlbhip::launch_render_frame(d, set, noisy + static_cast<size_t>(frame - first_frame) * plane, out, s),
          "render frame kernel");
    for (int a = first; a <= last; ++a) {
        for (size_t i = 0; i < cache.anchor.size(); ++i) {
            if (cache.anchor[i] == a) {
                check(hipEventRecord(cache.read_done[i], s), "hipEventRecord");
            }
        }
    }
}

}  // namespace

int vnlbhip_frame_cache_render(vnlbhip_frame_cache* cache, int frame, const float* noisy, const float* ref,
                               float* out, void* stream) {
    if (cache == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(cache->ctx, [&]() {
        std::lock_guard<std::mutex> lock(cache->mutex);
        render_frame(*cache, frame, noisy, ref, 0, cache->total_frames, out, reinterpret_cast<hipStream_t>(stream),
                     0);
    });
}

int vnlbhip_frame_cache_render_window(vnlbhip_frame_cache* cache, int frame, const float* noisy,
                                      const float* ref, int first_frame, int window_frames, float* out,
                                      void* stream) {
    if (cache == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(cache->ctx, [&]() {
        std::lock_guard<std::mutex> lock(cache->mutex);
        render_frame(*cache, frame, noisy, ref, first_frame, window_frames, out,
                     reinterpret_cast<hipStream_t>(stream), 1);
    });
}

int vnlbhip_frame_cache_invalidate(vnlbhip_frame_cache* cache) {
    if (cache == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    std::lock_guard<std::mutex> lock(cache->mutex);
    cache->invalidate();
    return VNLBHIP_OK;
}

int vnlbhip_frame_cache_get_stats(const vnlbhip_frame_cache* cache, vnlbhip_frame_cache_stats* out) {
    if (cache == nullptr || out == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    *out = cache->stats;
    return VNLBHIP_OK;
}

void vnlbhip_frame_cache_destroy(vnlbhip_frame_cache* cache) {
    delete cache;
}

int vnlbhip_stage_host(vnlbhip_context* context, int stage, const vnlbhip_params* params,
                       const float* noisy, const float* ref, float* out, int width, int height,
                       int frames, int channels) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        require(params != nullptr && noisy != nullptr && out != nullptr,
                "params, noisy and out must not be null");
        require(stage != VNLBHIP_STAGE_FINAL || ref != nullptr, "Final needs the ref (basic) clip");
        require(width > 0 && height > 0 && frames > 0 && channels > 0,
                "video geometry must be non-empty");
        check(hipSetDevice(context->device), "hipSetDevice");
        const size_t count = static_cast<size_t>(width) * height * frames * channels;
        float* d_in = context->host_in.reserve(count);
        float* d_out = context->host_out.reserve(count);
        float* d_ref = nullptr;
        check(hipMemcpy(d_in, noisy, count * sizeof(float), hipMemcpyHostToDevice), "upload clip");
        if (stage == VNLBHIP_STAGE_FINAL) {
            d_ref = context->host_ref.reserve(count);
            check(hipMemcpy(d_ref, ref, count * sizeof(float), hipMemcpyHostToDevice), "upload ref");
        }
        run_stage(*context, stage, *params, d_in, d_ref, d_out,
                  Geometry{width, height, frames, channels}, nullptr, kAllFrames, nullptr);
        check(hipMemcpy(out, d_out, count * sizeof(float), hipMemcpyDeviceToHost), "download output");
    });
}

int vnlbhip_debug_frame(vnlbhip_context* context, int stage, const vnlbhip_params* params,
                        const float* noisy, const float* ref, int width, int height, int frames,
                        int channels, int anchor_frame, int capacity_anchors, int* num_anchors,
                        int* anchor_x, int* anchor_y, int* counts, int* positions,
                        float* distances, unsigned char* active, float* weights,
                        float* eigenvalues, int* group_flags, int gram_dim, float* grams) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        require(params != nullptr, "params must not be null");
        require(anchor_frame >= 0, "anchor_frame must be non-negative");
        DebugOutput debug;
        debug.capacity_anchors = capacity_anchors;
        debug.num_anchors = num_anchors;
        debug.anchor_x = anchor_x;
        debug.anchor_y = anchor_y;
        debug.counts = counts;
        debug.positions = positions;
        debug.distances = distances;
        debug.active = active;
        debug.weights = weights;
        debug.eigenvalues = eigenvalues;
        debug.group_flags = group_flags;
        require(grams == nullptr || gram_dim > 0, "gram_dim must be positive when grams are requested");
        debug.gram_dim = gram_dim;
        debug.grams = grams;
        if (num_anchors != nullptr) {
            *num_anchors = 0;
        }
        run_stage(*context, stage, *params, noisy, ref, nullptr,
                  Geometry{width, height, frames, channels}, nullptr, anchor_frame, &debug);
    });
}

}  // extern "C"

// ---- Learned candidate proposer ----

namespace {

template <class Fn> int proposer_guarded(vnlbhip_proposer* proposer, Fn&& fn) {
    try {
        fn();
        if (proposer != nullptr) {
            proposer->error.clear();
        }
        return VNLBHIP_OK;
    } catch (const InvalidArgument& e) {
        if (proposer != nullptr) {
            proposer->error = e.what();
        }
        return VNLBHIP_ERROR_INVALID;
    } catch (const std::exception& e) {
        if (proposer != nullptr) {
            proposer->error = e.what();
        }
        return VNLBHIP_ERROR_DEVICE;
    }
}

// Training buffers for `rows` input rows.
void reserve_training(vnlbhip_proposer& p, int rows) {
    const size_t need = vnlbhip::proposer_workspace_floats(rows);
    if (need > p.workspace_floats) {
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (p.workspace != nullptr) {
            (void)hipFree(p.workspace);
            p.workspace = nullptr;
        }
        check(hipMalloc(reinterpret_cast<void**>(&p.workspace), need * sizeof(float)), "hipMalloc proposer workspace");
        p.workspace_floats = need;
    }
}

void check_rows(const void* x, int rows, int group) {
    require(x != nullptr && reinterpret_cast<uintptr_t>(x) % 16 == 0, "input rows must be 16-byte aligned device memory");
    require(rows > 0 && group > 0 && rows % group == 0, "rows must be a positive multiple of group");
    require(rows <= (1 << 22), "at most 4194304 rows per call");
}

}  // namespace

extern "C" {

int vnlbhip_proposer_create(int device, uint64_t seed, vnlbhip_proposer** out) {
    if (out == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        return VNLBHIP_ERROR_INVALID;
    }
    auto* p = new (std::nothrow) vnlbhip_proposer();
    if (p == nullptr) {
        return VNLBHIP_ERROR_DEVICE;
    }
    p->device = device;
    const int rc = proposer_guarded(p, [&]() {
        check(hipSetDevice(device), "hipSetDevice");
        const size_t n = vnlbhip::kPropParams;
        for (float** buf : {&p->params, &p->grads, &p->adam_m, &p->adam_v}) {
            check(hipMalloc(reinterpret_cast<void**>(buf), n * sizeof(float)), "hipMalloc proposer");
        }
        check(hipMalloc(reinterpret_cast<void**>(&p->w1), vnlbhip::kPropHidden * vnlbhip::kPropIn * sizeof(__bf16)),
              "hipMalloc proposer");
        check(hipMalloc(reinterpret_cast<void**>(&p->w2), vnlbhip::kPropHidden * vnlbhip::kPropHidden * sizeof(__bf16)),
              "hipMalloc proposer");
        check(hipMalloc(reinterpret_cast<void**>(&p->w1_fp8), vnlbhip::kPropHidden * vnlbhip::kPropIn), "hipMalloc proposer");
        check(hipMalloc(reinterpret_cast<void**>(&p->w2_fp8), vnlbhip::kPropHidden * vnlbhip::kPropHidden),
              "hipMalloc proposer");
        check(hipMalloc(reinterpret_cast<void**>(&p->fp8_scales), vnlbhip::kPropScales * sizeof(float)),
              "hipMalloc proposer");
        check(hipMemset(p->adam_m, 0, n * sizeof(float)), "hipMemset");
        check(hipMemset(p->adam_v, 0, n * sizeof(float)), "hipMemset");
        check(vnlbhip::launch_proposer_init(p->params, seed, nullptr), "proposer init kernel");
        check(vnlbhip::launch_proposer_refresh(p->params, p->w1, p->w2, nullptr), "proposer refresh kernel");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
    if (rc != VNLBHIP_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return VNLBHIP_OK;
}

void vnlbhip_proposer_destroy(vnlbhip_proposer* proposer) { delete proposer; }

const char* vnlbhip_proposer_last_error(const vnlbhip_proposer* proposer) {
    return proposer == nullptr ? "null proposer" : proposer->error.c_str();
}

int vnlbhip_proposer_get_params(const vnlbhip_proposer* proposer, float* params) {
    if (proposer == nullptr || params == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    if (hipSetDevice(proposer->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(params, proposer->params, vnlbhip::kPropParams * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
        return VNLBHIP_ERROR_DEVICE;
    }
    return VNLBHIP_OK;
}

int vnlbhip_proposer_set_params(vnlbhip_proposer* proposer, const float* params) {
    if (proposer == nullptr || params == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return proposer_guarded(proposer, [&]() {
        check(hipSetDevice(proposer->device), "hipSetDevice");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check(hipMemcpy(proposer->params, params, vnlbhip::kPropParams * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy");
        check(vnlbhip::launch_proposer_refresh(proposer->params, proposer->w1, proposer->w2, nullptr),
              "proposer refresh kernel");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        proposer->fp8_ready = false;
    });
}

int vnlbhip_proposer_features(vnlbhip_proposer* proposer, const float* clip, int width, int height, int frames,
                              int channels, int patch_time, const uint32_t* anchors, const uint32_t* candidates,
                              int rows, void* out, void* stream) {
    if (proposer == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return proposer_guarded(proposer, [&]() {
        require(clip != nullptr && anchors != nullptr && candidates != nullptr && out != nullptr,
                "pointers must not be null");
        require(reinterpret_cast<uintptr_t>(out) % 16 == 0, "out must be 16-byte aligned");
        require(width >= 9 && height >= 9 && width <= vnlbhip::kMaxSide && height <= vnlbhip::kMaxSide,
                "frame sides must be in [9, 4096]");
        require(frames > 0 && frames <= vnlbhip::kMaxFrames && channels > 0 && channels <= vnlbhip::kMaxChannels,
                "frames in [1, 256], channels in [1, 4]");
        require(patch_time == 1 || patch_time == 2, "patch_time must be 1 or 2");
        require(rows >= 0, "rows must be non-negative");
        check(hipSetDevice(proposer->device), "hipSetDevice");
        // Positions are trusted only after a bounds check on the host side
        // of the caller's data: the kernel reads 8x8 patches at them, so
        // they are validated here on a host copy.
        std::vector<uint32_t> a(static_cast<size_t>(rows));
        std::vector<uint32_t> c(static_cast<size_t>(rows));
        if (rows > 0) {
            check(hipMemcpy(a.data(), anchors, a.size() * sizeof(uint32_t), hipMemcpyDeviceToHost), "read anchors");
            check(hipMemcpy(c.data(), candidates, c.size() * sizeof(uint32_t), hipMemcpyDeviceToHost),
                  "read candidates");
        }
        for (int r = 0; r < rows; ++r) {
            for (const uint32_t q : {a[static_cast<size_t>(r)], c[static_cast<size_t>(r)]}) {
                require(vnlbhip::position_x(q) + 8 <= width && vnlbhip::position_y(q) + 8 <= height &&
                            vnlbhip::position_frame(q) + patch_time <= frames,
                        "a patch position lies outside the clip");
            }
        }
        check(vnlbhip::launch_proposer_features(clip, width, height, channels, patch_time, anchors, candidates, rows,
                                                static_cast<__bf16*>(out), reinterpret_cast<hipStream_t>(stream)),
              "proposer features kernel");
    });
}

int vnlbhip_proposer_gradients(vnlbhip_proposer* proposer, const void* x, const float* targets, int rows, int group,
                               float* grads, float* loss_groups, void* stream) {
    if (proposer == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return proposer_guarded(proposer, [&]() {
        check_rows(x, rows, group);
        require(targets != nullptr && grads != nullptr && loss_groups != nullptr, "pointers must not be null");
        check(hipSetDevice(proposer->device), "hipSetDevice");
        reserve_training(*proposer, rows);
        check(vnlbhip::launch_proposer_gradients(proposer->params, proposer->w1, proposer->w2,
                                                 static_cast<const __bf16*>(x), targets, rows, group,
                                                 proposer->workspace, grads, loss_groups,
                                                 reinterpret_cast<hipStream_t>(stream)),
              "proposer gradient kernels");
    });
}

int vnlbhip_proposer_train_step(vnlbhip_proposer* proposer, const void* x, const float* targets, int rows, int group,
                                float learning_rate, float* loss_groups, void* stream) {
    if (proposer == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return proposer_guarded(proposer, [&]() {
        check_rows(x, rows, group);
        require(targets != nullptr, "targets must not be null");
        require(std::isfinite(learning_rate) && learning_rate > 0.0F, "learning_rate must be positive");
        check(hipSetDevice(proposer->device), "hipSetDevice");
        reserve_training(*proposer, rows);
        float* losses = loss_groups;
        if (losses == nullptr) {
            const int groups = rows / group;
            if (groups > proposer->loss_capacity) {
                check(hipDeviceSynchronize(), "hipDeviceSynchronize");
                if (proposer->loss_scratch != nullptr) {
                    (void)hipFree(proposer->loss_scratch);
                    proposer->loss_scratch = nullptr;
                }
                check(hipMalloc(reinterpret_cast<void**>(&proposer->loss_scratch), groups * sizeof(float)),
                      "hipMalloc proposer losses");
                proposer->loss_capacity = groups;
            }
            losses = proposer->loss_scratch;
        }
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        check(vnlbhip::launch_proposer_gradients(proposer->params, proposer->w1, proposer->w2,
                                                 static_cast<const __bf16*>(x), targets, rows, group,
                                                 proposer->workspace, proposer->grads, losses, s),
              "proposer gradient kernels");
        ++proposer->step;
        check(vnlbhip::launch_proposer_adam(proposer->params, proposer->grads, proposer->adam_m, proposer->adam_v,
                                            learning_rate, proposer->step, s),
              "proposer adam kernel");
        check(vnlbhip::launch_proposer_refresh(proposer->params, proposer->w1, proposer->w2, s),
              "proposer refresh kernel");
        proposer->fp8_ready = false;
    });
}

int vnlbhip_proposer_scores(vnlbhip_proposer* proposer, const void* x, int rows, int precision, float* scores,
                            void* stream) {
    if (proposer == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return proposer_guarded(proposer, [&]() {
        check_rows(x, rows, 1);
        require(scores != nullptr, "scores must not be null");
        require(precision == VNLBHIP_PRECISION_BF16 || precision == VNLBHIP_PRECISION_FP8,
                "precision must be VNLBHIP_PRECISION_BF16 or VNLBHIP_PRECISION_FP8");
        require(precision == VNLBHIP_PRECISION_BF16 || proposer->fp8_ready,
                "fp8 scores need a calibrated proposer (vnlbhip_proposer_calibrate)");
        check(hipSetDevice(proposer->device), "hipSetDevice");
        check(vnlbhip::launch_proposer_scores(static_cast<const __bf16*>(x), rows, proposer->weights(), precision,
                                              scores, reinterpret_cast<hipStream_t>(stream)),
              "proposer scores kernel");
    });
}

int vnlbhip_proposer_calibrate(vnlbhip_proposer* proposer, const void* x, int rows, void* stream) {
    if (proposer == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return proposer_guarded(proposer, [&]() {
        check_rows(x, rows, 1);
        check(hipSetDevice(proposer->device), "hipSetDevice");
        reserve_training(*proposer, rows);
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        check(vnlbhip::launch_proposer_calibrate(proposer->params, proposer->w1, static_cast<const __bf16*>(x), rows,
                                                 proposer->workspace, proposer->w1_fp8, proposer->w2_fp8,
                                                 proposer->fp8_scales, s),
              "proposer calibration kernels");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
        proposer->fp8_ready = true;
    });
}

int vnlbhip_set_proposer(vnlbhip_context* context, const vnlbhip_proposer* proposer, int top_k, float margin,
                         int precision) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        if (proposer == nullptr) {
            context->proposer = nullptr;
            return;
        }
        require(proposer->device == context->device, "the proposer lives on another device");
        require(top_k >= 1 && top_k < vnlbhip::kPropMaxProposed, "top_k must be in [1, 255]");
        require(std::isfinite(margin), "margin must be finite");
        require(precision == VNLBHIP_PRECISION_BF16 || precision == VNLBHIP_PRECISION_FP8,
                "precision must be VNLBHIP_PRECISION_BF16 or VNLBHIP_PRECISION_FP8");
        require(precision == VNLBHIP_PRECISION_BF16 || proposer->fp8_ready,
                "fp8 proposals need a calibrated proposer (vnlbhip_proposer_calibrate)");
        context->proposer = proposer;
        context->proposer_top_k = top_k;
        context->proposer_margin = margin;
        context->proposer_precision = precision;
    });
}

int vnlbhip_match_lists(vnlbhip_context* context, int stage, const vnlbhip_params* params, const float* noisy,
                        const float* ref, int width, int height, int frames, int channels, int anchor_frame,
                        int capacity, int* anchor_count, uint32_t* anchors, int* counts, uint32_t* positions,
                        float* distances, void* stream) {
    if (context == nullptr) {
        return VNLBHIP_ERROR_INVALID;
    }
    return guarded(context, [&]() {
        require(params != nullptr && anchor_count != nullptr && anchors != nullptr && counts != nullptr &&
                    positions != nullptr && distances != nullptr,
                "pointers must not be null");
        require(anchor_frame >= 0 && anchor_frame < frames, "anchor_frame must be a frame of the clip");
        ListsOut lists;
        lists.capacity = capacity;
        lists.anchor_count = anchor_count;
        lists.anchors = anchors;
        lists.counts = counts;
        lists.positions = positions;
        lists.distances = distances;
        *anchor_count = 0;
        run_stage(*context, stage, *params, noisy, ref, nullptr, Geometry{width, height, frames, channels},
                  reinterpret_cast<hipStream_t>(stream), anchor_frame, nullptr, nullptr, &lists);
    });
}

int vnlbhip_search_windows(const vnlbhip_params* params, int width, int height, int total_frames, int count,
                           const uint32_t* anchors, int32_t* windows) {
    if (params == nullptr || anchors == nullptr || windows == nullptr || count < 0) {
        return VNLBHIP_ERROR_INVALID;
    }
    try {
        const Geometry g{width, height, total_frames, 1};
        const vnlbhip::StageDesc d = make_desc(VNLBHIP_STAGE_BASIC, *params, g);
        for (int i = 0; i < count; ++i) {
            const uint32_t a = anchors[i];
            int t_low = 0;
            int t_high = 0;
            int x_low = 0;
            int x_high = 0;
            int y_low = 0;
            int y_high = 0;
            vnlbhip::temporal_range(d, vnlbhip::position_frame(a), t_low, t_high);
            vnlbhip::spatial_range(d, vnlbhip::position_x(a), width, x_low, x_high);
            vnlbhip::spatial_range(d, vnlbhip::position_y(a), height, y_low, y_high);
            int32_t* w = windows + static_cast<size_t>(i) * 6;
            w[0] = x_low;
            w[1] = y_low;
            w[2] = x_high - x_low + 1;
            w[3] = y_high - y_low + 1;
            w[4] = t_low;
            w[5] = t_high;
        }
    } catch (const std::exception&) {
        return VNLBHIP_ERROR_INVALID;
    }
    return VNLBHIP_OK;
}

}  // extern "C"
