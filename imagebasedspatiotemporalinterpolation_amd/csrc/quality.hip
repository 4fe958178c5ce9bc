// quality.hip - PSNR and SSIM of two RGBA8 device images (include/isti_quality.h).
//
// SSIM (scripts/loss_utils.py:17-57): per channel, the five Gaussian-filtered
// quantities mu1, mu2, E[x^2], E[y^2], E[xy] (11x11 window, sigma 1.5, zero
// padding), then the SSIM map, averaged over every RGB sample. The 2D window
// is the outer product of the normalised 1D Gaussian, so the filter runs as a
// horizontal then a vertical pass over a tile staged in LDS: a 32x8 output
// tile reads a 42x18 input region once. PSNR (scripts/eval.py:57) shares the
// staged samples. Per-block sums are added in double precision.

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/isti_quality.h"

namespace isti_q {

constexpr int TX = 32, TY = 8;  // output tile (one 256-thread block)
constexpr int R = 5, K = 2 * R + 1;
constexpr int IX = TX + 2 * R, IY = TY + 2 * R;
constexpr float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;

struct Gauss {
    float g[K];
};

__global__ void clear_acc_kernel(double *acc) {
    if (threadIdx.x < 2) acc[threadIdx.x] = 0.0;
}

__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < 256 / 64; i++) s += red[i];
    return s;
}

__global__ __launch_bounds__(256) void quality_kernel(const uint8_t *__restrict__ a, uint32_t pa,
                                                      const uint8_t *__restrict__ b, uint32_t pb, uint32_t w,
                                                      uint32_t h, const Gauss gw, double *acc) {
    __shared__ float s_in[2][3][IY][IX];  // both images, RGB / 255, zero outside
    __shared__ float s_h[5][IY][TX];      // horizontal pass of one channel
    __shared__ double s_red[2][4];
    const int tid = (int)threadIdx.x;
    const int x0 = (int)(blockIdx.x * TX) - R, y0 = (int)(blockIdx.y * TY) - R;
    for (int i = tid; i < IX * IY; i += 256) {
        const int ly = i / IX, lx = i - ly * IX;
        const int gx = x0 + lx, gy = y0 + ly;
        const bool in = gx >= 0 && gx < (int)w && gy >= 0 && gy < (int)h;
        uchar4 ca = make_uchar4(0, 0, 0, 0), cb = make_uchar4(0, 0, 0, 0);
        if (in) {
            ca = *reinterpret_cast<const uchar4 *>(a + (size_t)gy * pa + 4u * (uint32_t)gx);
            cb = *reinterpret_cast<const uchar4 *>(b + (size_t)gy * pb + 4u * (uint32_t)gx);
        }
        s_in[0][0][ly][lx] = in ? (float)ca.x / 255.0f : 0.0f;
        s_in[0][1][ly][lx] = in ? (float)ca.y / 255.0f : 0.0f;
        s_in[0][2][ly][lx] = in ? (float)ca.z / 255.0f : 0.0f;
        s_in[1][0][ly][lx] = in ? (float)cb.x / 255.0f : 0.0f;
        s_in[1][1][ly][lx] = in ? (float)cb.y / 255.0f : 0.0f;
        s_in[1][2][ly][lx] = in ? (float)cb.z / 255.0f : 0.0f;
    }
    __syncthreads();
    const int ox = tid % TX, oy = tid / TX;
    const bool out = (int)(blockIdx.x * TX) + ox < (int)w && (int)(blockIdx.y * TY) + oy < (int)h;
    double ssim_sum = 0.0, sse = 0.0;
    for (int c = 0; c < 3; c++) {
        for (int i = tid; i < IY * TX; i += 256) {
            const int ly = i / TX, lx = i - ly * TX;
            float m1 = 0.f, m2 = 0.f, s11 = 0.f, s22 = 0.f, s12 = 0.f;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const float u = s_in[0][c][ly][lx + k], v = s_in[1][c][ly][lx + k], g = gw.g[k];
                m1 += g * u;
                m2 += g * v;
                s11 += g * (u * u);
                s22 += g * (v * v);
                s12 += g * (u * v);
            }
            s_h[0][ly][lx] = m1;
            s_h[1][ly][lx] = m2;
            s_h[2][ly][lx] = s11;
            s_h[3][ly][lx] = s22;
            s_h[4][ly][lx] = s12;
        }
        __syncthreads();
        if (out) {
            float q[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < K; k++) {
                const float g = gw.g[k];
#pragma unroll
                for (int j = 0; j < 5; j++) q[j] += g * s_h[j][oy + k][ox];
            }
            const float mu1 = q[0], mu2 = q[1];
            const float mu1_sq = mu1 * mu1, mu2_sq = mu2 * mu2, mu1_mu2 = mu1 * mu2;
            const float sigma1_sq = q[2] - mu1_sq, sigma2_sq = q[3] - mu2_sq, sigma12 = q[4] - mu1_mu2;
            const float num = (2.0f * mu1_mu2 + C1) * (2.0f * sigma12 + C2);
            const float den = (mu1_sq + mu2_sq + C1) * (sigma1_sq + sigma2_sq + C2);
            ssim_sum += (double)(num / den);
            const float d = s_in[0][c][oy + R][ox + R] - s_in[1][c][oy + R][ox + R];
            sse += (double)(d * d);
        }
        __syncthreads();  // s_h is rewritten for the next channel
    }
    const double t_sse = block_sum(sse, s_red[0]);
    const double t_ssim = block_sum(ssim_sum, s_red[1]);
    if (tid == 0) {
        atomicAdd(&acc[0], t_sse);
        atomicAdd(&acc[1], t_ssim);
    }
}

// The reference's window (loss_utils.py:17-25): exp(-(x - 5)^2 / (2 sigma^2))
// for x = 0..10, normalised to sum 1, in single precision.
Gauss gaussian_window() {
    Gauss gw;
    float s = 0.0f;
    for (int x = 0; x < K; x++) {
        gw.g[x] = (float)std::exp(-(double)((x - R) * (x - R)) / (2.0 * 1.5 * 1.5));
        s += gw.g[x];
    }
    for (int x = 0; x < K; x++) gw.g[x] /= s;
    return gw;
}

}  // namespace isti_q

extern "C" isti_status isti_image_quality(const uint8_t *a, size_t pitch_a, const uint8_t *b, size_t pitch_b,
                                          uint32_t w, uint32_t h, double *acc, void *stream) {
    using namespace isti_q;
    if (!a || !b || !acc || w == 0 || h == 0 || w > ISTI_MAX_DIM || h > ISTI_MAX_DIM || pitch_a < 4ull * w ||
        pitch_b < 4ull * w || pitch_a % 4 || pitch_b % 4 || pitch_a > 0xFFFFFFFFull || pitch_b > 0xFFFFFFFFull)
        return ISTI_ERR_INVALID_ARG;
    static const Gauss gw = gaussian_window();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(clear_acc_kernel, dim3(1), dim3(64), 0, s, acc);
    const dim3 grid((w + TX - 1) / TX, (h + TY - 1) / TY);
    hipLaunchKernelGGL(quality_kernel, grid, dim3(256), 0, s, a, (uint32_t)pitch_a, b, (uint32_t)pitch_b, w, h, gw,
                       acc);
    return hipGetLastError() == hipSuccess ? ISTI_OK : ISTI_ERR_DEVICE;
}
