// rd_scratch.h — the fixed device work area of rd_render_device_noalloc.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/radiantium_hip.h"

namespace rd {

// Size of the work area (RD_NOALLOC_SCRATCH_BYTES of the C ABI).
constexpr size_t kScratchBytes = (size_t)RD_NOALLOC_SCRATCH_BYTES;

// Device address of `device`'s copy of the work area.  Call with `device`
// current; the lookup runs once per device and is cached.
hipError_t scratch_address(int device, void** out);

}  // namespace rd
