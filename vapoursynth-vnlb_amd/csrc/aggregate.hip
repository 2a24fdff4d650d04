// Accumulator clear and normalisation.
//
// The stage accumulates every weighted contribution straight into one
// double-precision (numerator per channel, weight) set per output frame.
// That is the sum the reference forms in two steps -- per anchor-frame
// contribution stacks (src/aggregate/aggregate.hpp:18-71) summed by
// vnlb.Aggregate over the anchors [t - fwd - (Pt - 1), t + bwd]
// (src/vs/plugin.cpp:940-1045) -- with the slot filter applied at scatter
// time.  Normalisation: numerator / weight, or the source sample where no
// weight landed.
#include "common.h"
#include "kernels.h"

namespace vnlbhip {
namespace {

__global__ __launch_bounds__(256) void fill_zero_kernel(double* __restrict__ data, size_t count) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += stride) {
        data[i] = 0.0;
    }
}

__global__ __launch_bounds__(256) void fill_zero_int_kernel(int* __restrict__ data, size_t count) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += stride) {
        data[i] = 0;
    }
}

template <class T>
__global__ __launch_bounds__(256) void normalize_kernel(StageDesc d,
                                                        const T* __restrict__ acc,
                                                        const float* __restrict__ source,
                                                        float* __restrict__ out) {
    const size_t plane = static_cast<size_t>(d.width) * d.height;
    const size_t total = plane * d.frames;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += stride) {
        const size_t frame = i / plane;
        const size_t pixel = i - frame * plane;
        const T* base = acc + frame * (d.channels + 1) * plane;
        const float weight = static_cast<float>(base[static_cast<size_t>(d.channels) * plane + pixel]);
        for (int c = 0; c < d.channels; ++c) {
            const size_t index = (frame * d.channels + c) * plane + pixel;
            const float numerator = static_cast<float>(base[static_cast<size_t>(c) * plane + pixel]);
            out[index] = weight > 0.0F ? numerator / weight : source[index];
        }
    }
}


// One anchor frame's contribution stack in the vnlb.Basic / vnlb.Final
// output-frame layout (src/vs/plugin.cpp:264-271,625-641; pack_kernel,
// src/cuda/aggregation.cu:508-528): plane c * slots + s holds the
// numerators of channel c for output frame first_frame + s, then that
// frame's weights.  The accumulators hold only this anchor frame's groups;
// slots outside the clip are zero.
template <class T>
__global__ __launch_bounds__(256) void pack_stack_kernel(StageDesc d, const T* __restrict__ acc,
                                                         int first_frame, int slots,
                                                         float* __restrict__ stack) {
    const size_t pixels = static_cast<size_t>(d.width) * d.height;
    const size_t total = pixels * d.channels * slots;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        const size_t plane = i / pixels;
        const size_t pixel = i - plane * pixels;
        const int c = static_cast<int>(plane / slots);
        const int slot = static_cast<int>(plane - static_cast<size_t>(c) * slots);
        const int frame = first_frame + slot;
        float numerator = 0.0F;
        float weight = 0.0F;
        if (frame >= 0 && frame < d.frames) {
            const T* base = acc + static_cast<size_t>(frame) * (d.channels + 1) * pixels;
            numerator = static_cast<float>(base[static_cast<size_t>(c) * pixels + pixel]);
            weight = static_cast<float>(base[static_cast<size_t>(d.channels) * pixels + pixel]);
        }
        stack[plane * 2 * pixels + pixel] = numerator;
        stack[(plane * 2 + 1) * pixels + pixel] = weight;
    }
}

// vnlb.Aggregate of one output frame from anchor stacks
// (render_aggregate_frame, src/vs/plugin.cpp:966-1045; aggregate_frame,
// src/aggregate/aggregate.cpp:95-163): the channel-c numerators and the
// weights at each contributing anchor's slot, summed in double, then
// numerator / weight, or the source sample where no weight landed.
__global__ __launch_bounds__(256) void render_frame_kernel(StageDesc d, StackSet set,
                                                           const float* __restrict__ source,
                                                           float* __restrict__ out) {
    const size_t pixels = static_cast<size_t>(d.width) * d.height;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t pixel = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; pixel < pixels;
         pixel += stride) {
        double weight = 0.0;
        for (int k = 0; k < set.count; ++k) {
            weight += set.stack[k][(static_cast<size_t>(set.slot[k]) * 2 + 1) * pixels + pixel];
        }
        const float w = static_cast<float>(weight);
        for (int c = 0; c < d.channels; ++c) {
            double numerator = 0.0;
            for (int k = 0; k < set.count; ++k) {
                const size_t plane = static_cast<size_t>(c) * d.slots + set.slot[k];
                numerator += set.stack[k][plane * 2 * pixels + pixel];
            }
            const size_t index = static_cast<size_t>(c) * pixels + pixel;
            out[index] = w > 0.0F ? static_cast<float>(numerator) / w : source[index];
        }
    }
}

// Stage health counters: summary[0] counts groups whose model rows cannot
// cover the rank (the reference throws for those, core.cpp:645-648),
// summary[1] eigen problems that reached the sweep limit without meeting the
// convergence test, summary[2..3] retained samples over the groups (64-bit),
// summary[4] problems re-solved in fp64, summary[5] groups.  The group count
// of the list is read on the device (grids are sized by the capacity).
// A grid-stride loop over the problems (a few hundred workgroups: one
// problem per thread would launch a workgroup per 256 problems, whose
// dispatch costs more than their work); each wavefront adds its counts with
// one atomic per counter.
__global__ __launch_bounds__(256) void summarize_kernel(const int* __restrict__ flags,
                                                        const int* __restrict__ status,
                                                        const int* __restrict__ counts,
                                                        const int* __restrict__ group_list,
                                                        const int* __restrict__ groups_dev, int models,
                                                        int* __restrict__ summary) {
    const int groups = *groups_dev;
    const int first = blockIdx.x * blockDim.x + threadIdx.x;
    if (first == 0) {
        atomicAdd(&summary[5], groups);
    }
    unsigned long long retained = 0;
    int short_count = 0;
    int unconverged_count = 0;
    const int problems = groups * models;
    for (int base = blockIdx.x * blockDim.x; base < problems; base += gridDim.x * blockDim.x) {
        const int problem = base + threadIdx.x;
        bool short_rows = false;
        bool unconverged = false;
        if (problem < problems) {
            const int group = problem / models;
            const int flag = flags[group];
            if ((problem % models) == 0) {
                short_rows = (flag & 8) != 0;
                retained += static_cast<unsigned long long>(counts[group_list[group]]);
            }
            unconverged = status != nullptr && (flag & kFlagModelled) != 0 && (status[problem] >> 16) != 0;
        }
        short_count += __popcll(__ballot(short_rows));
        unconverged_count += __popcll(__ballot(unconverged));
    }
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
        retained += __shfl_xor(retained, offset, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (short_count != 0) {
            atomicAdd(&summary[0], short_count);
        }
        if (unconverged_count != 0) {
            atomicAdd(&summary[1], unconverged_count);
        }
        if (retained != 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(summary + 2), retained);
        }
    }
}

unsigned grid_for(size_t count) {
    size_t blocks = (count + 255) / 256;
    if (blocks > 65536) {
        blocks = 65536;
    }
    return static_cast<unsigned>(blocks == 0 ? 1 : blocks);
}

}  // namespace

hipError_t launch_summarize(const int* flags, const int* status, const int* counts,
                            const int* group_list, const int* groups_dev, int group_capacity,
                            int models, int* summary, hipStream_t stream) {
    const int problems = group_capacity * models;
    if (problems <= 0) {
        return hipSuccess;
    }
    const int blocks = (problems + 255) / 256 < 512 ? (problems + 255) / 256 : 512;
    hipLaunchKernelGGL(summarize_kernel, dim3(static_cast<unsigned>(blocks)),
                       dim3(256), 0, stream, flags, status, counts, group_list, groups_dev, models,
                       summary);
    return hipGetLastError();
}

hipError_t launch_fill_zero(double* data, size_t count, hipStream_t stream) {
    if (count == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(count)), dim3(256), 0, stream, data, count);
    return hipGetLastError();
}

hipError_t launch_fill_zero_int(int* data, size_t count, hipStream_t stream) {
    if (count == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(fill_zero_int_kernel, dim3(grid_for(count)), dim3(256), 0, stream, data, count);
    return hipGetLastError();
}

template <class T>
hipError_t launch_pack(const StageDesc& d, const T* acc, int first_frame, int slots, float* stack,
                       hipStream_t stream) {
    const size_t total = static_cast<size_t>(d.width) * d.height * d.channels * slots;
    hipLaunchKernelGGL(pack_stack_kernel<T>, dim3(grid_for(total)), dim3(256), 0, stream, d, acc,
                       first_frame, slots, stack);
    return hipGetLastError();
}

hipError_t launch_pack_stack(const StageDesc& d, const double* acc, int first_frame, int slots,
                             float* stack, hipStream_t stream) {
    return launch_pack(d, acc, first_frame, slots, stack, stream);
}

hipError_t launch_pack_stack(const StageDesc& d, const float* acc, int first_frame, int slots,
                             float* stack, hipStream_t stream) {
    return launch_pack(d, acc, first_frame, slots, stack, stream);
}

hipError_t launch_render_frame(const StageDesc& d, const StackSet& set, const float* source, float* out,
                               hipStream_t stream) {
    const size_t pixels = static_cast<size_t>(d.width) * d.height;
    hipLaunchKernelGGL(render_frame_kernel, dim3(grid_for(pixels)), dim3(256), 0, stream, d, set, source,
                       out);
    return hipGetLastError();
}

hipError_t launch_normalize(const StageDesc& d, const double* acc, const float* source,
                            float* out, hipStream_t stream) {
    const size_t total = static_cast<size_t>(d.width) * d.height * d.frames;
    hipLaunchKernelGGL(normalize_kernel<double>, dim3(grid_for(total)), dim3(256), 0, stream, d, acc,
                       source, out);
    return hipGetLastError();
}

hipError_t launch_normalize(const StageDesc& d, const float* acc, const float* source,
                            float* out, hipStream_t stream) {
    const size_t total = static_cast<size_t>(d.width) * d.height * d.frames;
    hipLaunchKernelGGL(normalize_kernel<float>, dim3(grid_for(total)), dim3(256), 0, stream, d, acc,
                       source, out);
    return hipGetLastError();
}

}  // namespace vnlbhip
