// scratch.hip — the device memory of uc_run_lle_noalloc, the entry point
// that allocates nothing (for callers that cannot call hipMalloc: code under
// graph capture, tight serving loops).
//
// A launch's input tables (axis values nx + ny, 64 parameter slots, 32
// initial-state slots, in doubles) live in `scratch`, a file-scope device
// array that each GPU holds from the first engine created on it until the
// process ends. In host code `scratch` names a host-side stand-in, not device
// memory: host code uses only the address hipGetSymbolAddress returns for the
// current device, kept per device by uc_scratch_setup. Every call on one GPU
// shares that GPU's array: calls that overlap on one GPU (different streams,
// or different threads) race on it.
//
// Built by hipcc for gfx950 into libucuda.so; the kernels themselves are
// hipRTC modules and receive the address as an ordinary pointer.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "scratch.hpp"

static __device__ __align__(256) double scratch[UC_SCRATCH_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];
bool g_tried[kMaxDevices];
hipError_t g_status[kMaxDevices];
}  // namespace

// Once per device, with `device` current: the device address of `scratch`.
int uc_scratch_setup(int device, char* msg, size_t cap) {
    if (device < 0 || device >= kMaxDevices) {
        if (msg && cap) snprintf(msg, cap, "scratch: device %d outside 0..%d", device, kMaxDevices - 1);
        return 0;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_tried[device]) {
        g_tried[device] = true;
        void* p = nullptr;
        g_status[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[device] = g_status[device] == hipSuccess ? p : nullptr;
    }
    if (!g_ptr[device]) {
        if (msg && cap)
            snprintf(msg, cap, "scratch: hipGetSymbolAddress failed on device %d: %s", device,
                     hipGetErrorString(g_status[device]));
        return 0;
    }
    return 1;
}

void* uc_scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}
