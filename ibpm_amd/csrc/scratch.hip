// scratch.hip -- the fixed device scratch array of the allocation-free entry
// point (ibpm_elliptic_solve_noalloc).
//
// The array is a file-scope __device__ variable: every GPU the library runs
// on holds its own copy, at its own address, from the time the code object
// is loaded until the process ends.  Host code must never pass the array's
// name to HIP or to a kernel (on the host it names a stand-in, not device
// memory): scratch_attach() looks up the device address once per device, with
// that device current, and scratch_device_ptr() hands that address out.
#include <hip/hip_runtime.h>

#include <mutex>

#include "scratch.h"

namespace ibpmg {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_mu;
void* g_ptr[kScratchMaxDevices] = {};
}  // namespace

int scratch_attach(int device) {
    if (device < 0 || device >= kScratchMaxDevices) {
        set_error("scratch: device %d out of range (at most %d devices)", device, kScratchMaxDevices);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return 0;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) {
        set_error("scratch: device %d is not current (current %d)", device, cur);
        return -1;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error("scratch: hipGetSymbolAddress failed on device %d: %s", device, hipGetErrorString(e));
        return -1;
    }
    g_ptr[device] = p;
    return 0;
}

int scratch_device_ptr(int device, void** out) {
    if (device < 0 || device >= kScratchMaxDevices) {
        set_error("scratch: device %d out of range", device);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        set_error("scratch: not attached on device %d", device);
        return -1;
    }
    *out = g_ptr[device];
    return 0;
}

}  // namespace ibpmg
