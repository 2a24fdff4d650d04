// Device scratch of vnlbhip_stage_scratch: one fixed array per GPU, so a
// stage call can run without allocating device memory.
//
// The array lives in this code object; each device holds its own copy (at
// its own address) from the moment the library is first set up on it until
// the process ends.  Host code must never pass `scratch` itself to a kernel:
// in host code the name is a host-side stand-in, not a device address.  The
// device address comes from hipGetSymbolAddress, looked up once per device
// with that device current, and only that address is used afterwards.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

#include "kernels.h"

namespace {

constexpr size_t kScratchBytes = static_cast<size_t>(128) << 20;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace vnlbhip {

size_t scratch_bytes() { return kScratchBytes; }

hipError_t scratch_address(int device, void** out) {
    if (out == nullptr || device < 0 || device >= kMaxDevices) {
        return hipErrorInvalidValue;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_address[device] == nullptr) {
        int current = -1;
        hipError_t status = hipGetDevice(&current);
        if (status != hipSuccess) {
            return status;
        }
        if (current != device) {
            return hipErrorInvalidDevice;  // the caller makes `device` current first
        }
        void* ptr = nullptr;
        status = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (status != hipSuccess) {
            return status;
        }
        if (ptr == nullptr) {
            return hipErrorInvalidSymbol;
        }
        g_scratch_address[device] = ptr;
    }
    *out = g_scratch_address[device];
    return hipSuccess;
}

}  // namespace vnlbhip
