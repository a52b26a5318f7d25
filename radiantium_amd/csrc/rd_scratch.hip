// rd_scratch.hip — the fixed device work area of rd_render_device_noalloc.
//
// A render needs device memory beyond its inputs and outputs: the path pool,
// its ray queues, per-sample results, per-pixel sums and the segment table.
// rd_render_device allocates them (and keeps them on its context);
// rd_render_device_noalloc takes them from this array instead, so a caller
// that cannot allocate device memory can still render.  The array is part of
// the library's code object: every device the library runs on holds its own
// copy, at its own address, until the process ends.
//
// Host code never passes `scratch` itself to a kernel or a copy (on the host
// the name is only a stand-in, not a device address): rd_create looks up the
// device address once per device with hipGetSymbolAddress and the render
// carves its buffers from that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rd_scratch.h"

static __device__ __align__(256) unsigned char scratch[rd::kScratchBytes];

namespace rd {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[device]) {
        // the caller has made `device` current: the lookup loads the code
        // object there and returns that device's copy of the array
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_addr[device] = p;
    }
    *out = g_addr[device];
    return hipSuccess;
}

}  // namespace rd
