// sl_noalloc.hip -- the static device scratch of the allocation-free frame
// (sl_rasterize_noalloc in include/scanline.h).
//
// `scratch` is one file-scope device array; every GPU the library runs on holds its own
// copy, at its own address, for the life of the process.  In host code the array's name
// is only a host-side stand-in: its device address on a GPU comes from
// hipGetSymbolAddress with that GPU current, looked up once per device (at sl_create) and
// kept here.  Host code passes that address, never the name, to kernels and copies.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "sl_internal.h"

namespace {

constexpr size_t kScratchBytes = (size_t)256 << 20;  // 256 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

}  // namespace

namespace sl {

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// Device address of `scratch` on `device`, which must be the current device.
hipError_t noalloc_scratch(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return hipSuccess;
}

}  // namespace sl
