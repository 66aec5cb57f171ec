// scratch.h -- the fixed device scratch of the allocation-free entry point.
#pragma once
#include <cstddef>

namespace ibpmg {

// Bytes of the per-device scratch array (32 MiB).
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kScratchMaxDevices = 64;

// Looks up the device address of the scratch array on `device`, which must be
// the current device; once per device (later calls return at once).
int scratch_attach(int device);
// The address scratch_attach found on `device`.
int scratch_device_ptr(int device, void** out);

void set_error(const char* fmt, ...);

}  // namespace ibpmg
