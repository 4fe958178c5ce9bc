// resolve_kernels.hip - the per-render resolve pass (K2, renderer.cu:148-465)
// for every search window / validity variant, aux-camera count and tile
// shape; the tile shape is picked per render size by isti_autotune.

#include "resolve_device.hpp"

namespace isti {

template <int W, bool MAV, int NAUX, int TW, int TH>
__global__ __launch_bounds__(TW * TH) __attribute__((amdgpu_waves_per_eu(resolve_min_waves(W, NAUX))))
void resolve_kernel(const ResolveArgs args) {
    // Read the argument block in place (constant address space) instead of
    // through a by-value copy, which the compiler may spill to scratch.
#if defined(__HIP_DEVICE_COMPILE__)
    const KernArg<ResolveArgs> &a = *kernarg_ptr<ResolveArgs>();
#else
    const ResolveArgs &a = args;  // host pass only parses the body
#endif
    __shared__ ResolveLds<W, NAUX, TW, TH> lds;
    resolve_tile<W, MAV, NAUX, false, TW, TH, true>(a, xcd_stripe_tile(blockIdx.x, a.n_tiles, a.tiles_x), lds);
}

template <int W, bool MAV, int TW, int TH>
static hipError_t launch_resolve_n(const ResolveArgs &a, hipStream_t s) {
    const dim3 grid(a.n_tiles), block(TW * TH);
    switch (a.n_aux) {
        case 0: hipLaunchKernelGGL((resolve_kernel<W, MAV, 0, TW, TH>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((resolve_kernel<W, MAV, 1, TW, TH>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((resolve_kernel<W, MAV, 2, TW, TH>), grid, block, 0, s, a); break;
        case 3: hipLaunchKernelGGL((resolve_kernel<W, MAV, 3, TW, TH>), grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL((resolve_kernel<W, MAV, 4, TW, TH>), grid, block, 0, s, a); break;
    }
    return hipGetLastError();
}

template <int W, bool MAV>
static hipError_t launch_resolve_shape(const ResolveArgs &a, int shape, hipStream_t s) {
    switch (shape) {
        case 0: return launch_resolve_n<W, MAV, 16, 16>(a, s);
        case 1: return launch_resolve_n<W, MAV, 32, 8>(a, s);
        case 2: return launch_resolve_n<W, MAV, 8, 8>(a, s);
        case 3: return launch_resolve_n<W, MAV, 8, 32>(a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_resolve(const ResolveArgs &a, int window, bool always_valid, int shape, hipStream_t s) {
    if (a.n_tiles == 0) return hipSuccess;
    if (a.n_aux < 0 || a.n_aux > ISTI_MAX_AUX) return hipErrorInvalidValue;
    // the reference pairs window 1 with plain validity and window 3 with
    // main-always-valid (renderer.cu:678-689); those are the two variants built
    if (window == 1 && !always_valid) return launch_resolve_shape<1, false>(a, shape, s);
    if (window == 3 && always_valid) return launch_resolve_shape<3, true>(a, shape, s);
    return hipErrorInvalidValue;
}

}  // namespace isti
