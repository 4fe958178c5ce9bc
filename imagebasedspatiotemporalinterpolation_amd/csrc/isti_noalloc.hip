// isti_noalloc.hip - isti_render_noalloc: Renderer::render for callers that
// cannot allocate device memory (code under graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// Every scratch plane of the frame (per keyframe slot and view: the u64
// splat cells, the cam_ff and full_ff flow records; for timewarp the row
// sums and the mean z) is carved from `scratch`, one static device array of
// ISTI_NOALLOC_SCRATCH_BYTES per GPU, loaded with this library's code object
// and held until the process ends. Its device address is looked up once per
// device with hipGetSymbolAddress; host code passes that address, never the
// array's name, to the kernels.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/isti.h"
#include "isti_internal.hpp"

static __device__ __align__(256) unsigned char scratch[ISTI_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on `device` (which is current), looked up on
// the first call for that device.
hipError_t scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" isti_status isti_render_noalloc(isti_renderer *r, const isti_render_info *info, uint8_t *out,
                                           size_t out_pitch, void *stream) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    if (!info || !out) return isti_internal_fail(r, ISTI_ERR_INVALID_ARG, "info and out must be non-null");
    // look the array up with the renderer's device current
    const int dev = isti_internal_device(r);
    if (dev < 0 || dev >= MAX_DEVICES) return isti_internal_fail(r, ISTI_ERR_DEVICE, "device index out of range");
    hipError_t e = hipSetDevice(dev);
    void *base = nullptr;
    if (e == hipSuccess) e = scratch_address(dev, &base);
    if (e != hipSuccess)
        return isti_internal_fail(r, ISTI_ERR_DEVICE, std::string("static scratch: ") + hipGetErrorString(e));
    return isti_internal_render(r, info, out, out_pitch, stream, base, ISTI_NOALLOC_SCRATCH_BYTES);
}
