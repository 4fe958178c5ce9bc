// unet.hip - the learned search initialiser's network on gfx950 (see
// include/isti_learned.h for the architecture and the C ABI).
//
// Convolutions are implicit GEMMs on MFMA (v_mfma_f32_16x16x16_bf16):
//   forward / input gradient: M = 16 pixels of a row, N = output channels,
//     K = (ky, kx, input channel); a wave owns 16 pixels x all outputs;
//   weight gradient: M = 16 output channels, N = 16 (ky, kx, cin) columns,
//     K = pixels; partial sums of a wave are added atomically.
// A 16x16x16 MFMA takes lane l's A[l%16][4(l/16)..+3] and B[4(l/16)..+3][l%16]
// and returns D[4(l/16)+i][l%16] in element i.
// Activations are NHWC bf16 with channel counts that are multiples of 16.
// The input of a decoder convolution is concat(upsample2x(low), skip), read
// on the fly (no concatenated or upsampled tensor is materialised).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/isti_learned.h"

namespace isti_nn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Input of a convolution: channels [0, ca) from tensor a (at half resolution,
// read at (y/2, x/2), when up_a), channels [ca, ca+cb) from tensor b.
struct Src {
    const uint16_t *a;
    const uint16_t *b;
    int ca, cb, up_a;
};

__device__ __forceinline__ bf16x4 load_src4(const Src &s, int n, int y, int x, int H, int W, int ch) {
    const uint16_t *p;
    if (ch < s.ca) {
        if (s.up_a) {
            const int h2 = H >> 1, w2 = W >> 1;
            p = s.a + (((size_t)n * h2 + (y >> 1)) * w2 + (x >> 1)) * s.ca + ch;
        } else {
            p = s.a + (((size_t)n * H + y) * W + x) * s.ca + ch;
        }
    } else {
        p = s.b + (((size_t)n * H + y) * W + x) * s.cb + (ch - s.ca);
    }
    const uint2 v = *reinterpret_cast<const uint2 *>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xffffu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xffffu);
    r[3] = (short)(v.y >> 16);
    return r;
}

__device__ __forceinline__ uint16_t load_src1(const Src &s, int n, int y, int x, int H, int W, int ch) {
    if (ch < s.ca) {
        if (s.up_a) {
            const int h2 = H >> 1, w2 = W >> 1;
            return s.a[(((size_t)n * h2 + (y >> 1)) * w2 + (x >> 1)) * s.ca + ch];
        }
        return s.a[(((size_t)n * H + y) * W + x) * s.ca + ch];
    }
    return s.b[(((size_t)n * H + y) * W + x) * s.cb + (ch - s.ca)];
}

// out[n][y][x][co] = act(bias[co] + sum_{ky,kx,ci} w[co][ky][kx][ci] * in[n][y+ky-p][x+kx-p][ci])
// A wave computes MT groups of 16 consecutive pixels of a row x NT*16 of the
// cout outputs, the group blockIdx.y: each weight fragment loaded for a K step
// feeds MT MFMAs and each input fragment NT MFMAs (register blocking; the K
// order per output, and so the result, depends on neither MT nor NT).
template <int NT, int MT>
__global__ __launch_bounds__(256) void conv_kernel(Src s, const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                                   int cin, int ks, int N, int H, int W, int cout, uint16_t *out_bf,
                                                   float *out_f, int relu) {
    const int lane = threadIdx.x & 63;
    constexpr int PXW = 16 * MT;
    const int chunks = (W + PXW - 1) / PXW;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= (long)N * H * chunks) return;
    const int n = (int)(wave / ((long)H * chunks));
    const int rem = (int)(wave - (long)n * H * chunks);
    const int y = rem / chunks, x0 = (rem - y * chunks) * PXW;
    const int pad = ks >> 1;
    const int kq = (lane >> 4) * 4;
    const int co_base = (int)blockIdx.y * NT * 16;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int ky = 0; ky < ks; ky++) {
        const int yy = y + ky - pad;
        for (int kx = 0; kx < ks; kx++) {
            bool in[MT];
            int xx[MT];
#pragma unroll
            for (int m = 0; m < MT; m++) {
                const int px = x0 + m * 16 + (lane & 15);
                xx[m] = px + kx - pad;
                in[m] = yy >= 0 && yy < H && xx[m] >= 0 && xx[m] < W && px < W;
            }
            const uint16_t *wp0 = w + (size_t)((co_base + (lane & 15)) * ks * ks + ky * ks + kx) * cin + kq;
            for (int c0 = 0; c0 < cin; c0 += 16) {
                bf16x4 av[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) {
                    av[m] = bf16x4{0, 0, 0, 0};
                    if (in[m]) av[m] = load_src4(s, n, yy, xx[m], H, W, c0 + kq);
                }
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const uint2 bv2 = *reinterpret_cast<const uint2 *>(wp0 + c0 + (size_t)t * 16 * ks * ks * cin);
                    bf16x4 bv;
                    bv[0] = (short)(bv2.x & 0xffffu);
                    bv[1] = (short)(bv2.x >> 16);
                    bv[2] = (short)(bv2.y & 0xffffu);
                    bv[3] = (short)(bv2.y >> 16);
#pragma unroll
                    for (int m = 0; m < MT; m++)
                        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[m], bv, acc[m][t], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co = co_base + t * 16 + (lane & 15);
        const float bsv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; m++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int x = x0 + m * 16 + kq + i;
                if (x >= W) continue;
                float v = acc[m][t][i] + bsv;
                if (relu) v = v > 0.0f ? v : 0.0f;
                const size_t o = (((size_t)n * H + y) * W + x) * cout + co;
                if (out_f) out_f[o] = v;
                else out_bf[o] = f2bf(v);
            }
        }
    }
}

// 3x3 convolution with the input staged in LDS: a block computes 4 rows x
// 16*MT pixels x NT*16 outputs (one row per wave) and walks the input
// channels in chunks of 16. Each chunk's 6-row halo is loaded once into LDS
// (zero outside the image) - the next chunk's into registers while the
// current one is multiplied - and the nine taps read their A fragments from
// LDS, so a chunk costs one round trip to memory instead of nine, overlapped
// with the previous chunk's MFMAs. The K order is chunk-major (c0, then ky,
// kx), so sums are rounded in a different order than conv_kernel's.
template <int NT, int MT>
__global__ __launch_bounds__(256) void conv3_lds_kernel(Src s, const uint16_t *__restrict__ w,
                                                        const float *__restrict__ bias, int cin, int N, int H, int W,
                                                        int cout, uint16_t *out_bf, float *out_f, int relu) {
    constexpr int PXW = 16 * MT, COLS = PXW + 2, QUADS = 4;  // 4 channels per 8-byte quad
    __shared__ uint2 tile[6][COLS][QUADS];
    const int lane = threadIdx.x & 63, wv = (int)(threadIdx.x >> 6);
    const int chunks = (W + PXW - 1) / PXW, rgroups = (H + 3) / 4;
    const int b = (int)blockIdx.x;
    const int n = b / (rgroups * chunks);
    const int rem = b - n * rgroups * chunks;
    const int ry = rem / chunks, x0 = (rem - ry * chunks) * PXW;
    const int y0 = ry * 4, y = y0 + wv;
    const int kq = (lane >> 4) * 4;
    const int co_base = (int)blockIdx.y * NT * 16;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    constexpr int ITEMS = 6 * COLS * QUADS, PER = (ITEMS + 255) / 256;
    uint2 pre[PER];
    auto load_chunk = [&](int c) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = (int)threadIdx.x + k * 256;
            pre[k] = make_uint2(0u, 0u);
            if (i < ITEMS) {
                const int r = i / (COLS * QUADS), rc = i - r * COLS * QUADS;
                const int col = rc / QUADS, q = rc - col * QUADS;
                const int yy = y0 - 1 + r, xx = x0 - 1 + col;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                    const bf16x4 v = load_src4(s, n, yy, xx, H, W, c + q * 4);
                    pre[k].x = (uint32_t)(uint16_t)v[0] | ((uint32_t)(uint16_t)v[1] << 16);
                    pre[k].y = (uint32_t)(uint16_t)v[2] | ((uint32_t)(uint16_t)v[3] << 16);
                }
            }
        }
    };
    auto store_chunk = [&]() {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = (int)threadIdx.x + k * 256;
            if (i < ITEMS) (&tile[0][0][0])[i] = pre[k];
        }
    };
    load_chunk(0);
    store_chunk();
    __syncthreads();
    for (int c0 = 0; c0 < cin; c0 += 16) {
        const bool more = c0 + 16 < cin;
        if (more) load_chunk(c0 + 16);  // in flight during this chunk's MFMAs
#pragma unroll
        for (int ky = 0; ky < 3; ky++) {
#pragma unroll
            for (int kx = 0; kx < 3; kx++) {
                bf16x4 av[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) {
                    const uint2 u = tile[wv + ky][m * 16 + (lane & 15) + kx][kq >> 2];
                    av[m][0] = (short)(u.x & 0xffffu);
                    av[m][1] = (short)(u.x >> 16);
                    av[m][2] = (short)(u.y & 0xffffu);
                    av[m][3] = (short)(u.y >> 16);
                }
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const uint2 bv2 = *reinterpret_cast<const uint2 *>(
                        w + (size_t)((co_base + t * 16 + (lane & 15)) * 9 + ky * 3 + kx) * cin + c0 + kq);
                    bf16x4 bv;
                    bv[0] = (short)(bv2.x & 0xffffu);
                    bv[1] = (short)(bv2.x >> 16);
                    bv[2] = (short)(bv2.y & 0xffffu);
                    bv[3] = (short)(bv2.y >> 16);
#pragma unroll
                    for (int m = 0; m < MT; m++)
                        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[m], bv, acc[m][t], 0, 0, 0);
                }
            }
        }
        __syncthreads();  // the next chunk overwrites the tile
        if (more) {
            store_chunk();
            __syncthreads();
        }
    }
    if (y >= H) return;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co = co_base + t * 16 + (lane & 15);
        const float bsv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; m++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int x = x0 + m * 16 + kq + i;
                if (x >= W) continue;
                float v = acc[m][t][i] + bsv;
                if (relu) v = v > 0.0f ? v : 0.0f;
                const size_t o = (((size_t)n * H + y) * W + x) * cout + co;
                if (out_f) out_f[o] = v;
                else out_bf[o] = f2bf(v);
            }
        }
    }
}

// dw[co][kk] += sum_p g[p][co] * in[p + off(kk)][ci(kk)], kk = (ky*ks+kx)*cin+ci
__global__ __launch_bounds__(256) void wgrad_kernel(Src s, const uint16_t *__restrict__ g, int cout, int cin, int ks,
                                                    int N, int H, int W, float *dw) {
    const int lane = threadIdx.x & 63;
    const int kk_tiles = (ks * ks * cin) / 16;
    const int co0 = (blockIdx.x / kk_tiles) * 16, kk0 = (blockIdx.x % kk_tiles) * 16;
    const int kk = kk0 + (lane & 15);
    const int kpos = kk / cin, ci = kk - kpos * cin;
    const int ky = kpos / ks, kx = kpos - ky * ks, pad = ks >> 1;
    const int co = co0 + (lane & 15), jq = (lane >> 4) * 4;
    const int chunks = (W + 15) >> 4;
    const long total = (long)N * H * chunks;
    const long stride = (long)gridDim.y * 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long c = (long)blockIdx.y * 4 + (threadIdx.x >> 6); c < total; c += stride) {
        const int n = (int)(c / ((long)H * chunks));
        const int rem = (int)(c - (long)n * H * chunks);
        const int y = rem / chunks, x0 = (rem - y * chunks) * 16;
        bf16x4 av, bv;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int x = x0 + jq + j;
            const bool inx = x < W;
            av[j] = inx ? (short)g[(((size_t)n * H + y) * W + x) * cout + co] : (short)0;
            const int yy = y + ky - pad, xx = x + kx - pad;
            const bool in = inx && yy >= 0 && yy < H && xx >= 0 && xx < W;
            bv[j] = in ? (short)load_src1(s, n, yy, xx, H, W, ci) : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    const int total_k = ks * ks * cin;
#pragma unroll
    for (int i = 0; i < 4; i++) atomicAdd(&dw[(size_t)(co0 + jq + i) * total_k + kk], acc[i]);
}

// db[co] += sum_p g[p][co]
__global__ __launch_bounds__(256) void bgrad_kernel(const uint16_t *__restrict__ g, int cout, long npx, float *db) {
    const int per = 256 / cout;  // threads per channel group
    const int co = threadIdx.x % cout, grp = threadIdx.x / cout;
    if (grp >= per) return;
    float acc = 0.f;
    for (long p = (long)blockIdx.x * per + grp; p < npx; p += (long)gridDim.x * per) acc += bf2f(g[p * cout + co]);
    atomicAdd(&db[co], acc);
}

// g = (y > 0) ? bf16(dy) : 0 (ReLU backward); `relu` 0 passes dy through
__global__ void mask_kernel(const uint16_t *__restrict__ y, const float *__restrict__ dy, uint16_t *g, long n,
                            int relu) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = dy[i];
    g[i] = (!relu || bf2f(y[i]) > 0.0f) ? f2bf(d) : (uint16_t)0;
}

__global__ void pool_fwd_kernel(const uint16_t *__restrict__ x, uint16_t *y, int N, int H, int W, int C) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int h2 = H >> 1, w2 = W >> 1;
    if (i >= (long)N * h2 * w2 * C) return;
    const int c = (int)(i % C);
    long r = i / C;
    const int xo = (int)(r % w2);
    r /= w2;
    const int yo = (int)(r % h2);
    const int n = (int)(r / h2);
    auto at = [&](int yy, int xx) { return bf2f(x[(((size_t)n * H + yy) * W + xx) * C + c]); };
    const float s = (at(2 * yo, 2 * xo) + at(2 * yo, 2 * xo + 1)) + (at(2 * yo + 1, 2 * xo) + at(2 * yo + 1, 2 * xo + 1));
    y[i] = f2bf(s * 0.25f);
}

// dx[n][y][x][c] += dy[n][y/2][x/2][c] / 4
__global__ void pool_bwd_kernel(const float *__restrict__ dy, float *dx, int N, int H, int W, int C) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * H * W * C) return;
    const int c = (int)(i % C);
    long r = i / C;
    const int x = (int)(r % W);
    r /= W;
    const int y = (int)(r % H);
    const int n = (int)(r / H);
    dx[i] += 0.25f * dy[(((size_t)n * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * C + c];
}

// Splits the gradient of concat(up2x(low), skip): channels [0, ca) are summed
// over each 2x2 block into d_low, channels [ca, ca+cb) are added to d_skip.
__global__ void split_kernel(const float *__restrict__ dcat, float *d_low, float *d_skip, int N, int H, int W, int ca,
                             int cb) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = ca + cb;
    if (i >= (long)N * H * W * C) return;
    const int c = (int)(i % C);
    const long p = i / C;
    if (c >= ca) {
        d_skip[p * cb + (c - ca)] += dcat[i];
        return;
    }
    // low-resolution cell: only the top-left child sums its block (no atomics)
    const int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((long)W * H));
    if ((x & 1) || (y & 1)) return;
    const size_t r0 = ((size_t)n * H + y) * W + x;
    const float s = (dcat[r0 * C + c] + dcat[(r0 + 1) * C + c]) + (dcat[(r0 + W) * C + c] + dcat[(r0 + W + 1) * C + c]);
    d_low[(((size_t)n * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * ca + c] += s;
}

// Adam on fp32 master weights, refreshing the bf16 working copies
__global__ void adam_kernel(float *p, const float *__restrict__ g, float *m, float *v, long n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of one layer's weights: w (forward) and wt[ci][ky'][kx'][co] =
// w[co][ks-1-ky'][ks-1-kx'][ci] (input gradient)
__global__ void cast_weights_kernel(const float *__restrict__ wf, uint16_t *w, uint16_t *wt, int cout, int cin, int ks) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)cout * ks * ks * cin;
    if (i >= n) return;
    const int ci = (int)(i % cin);
    long r = i / cin;
    const int kx = (int)(r % ks);
    r /= ks;
    const int ky = (int)(r % ks);
    const int co = (int)(r / ks);
    const uint16_t b = f2bf(wf[i]);
    w[i] = b;
    wt[(((size_t)ci * ks + (ks - 1 - ky)) * ks + (ks - 1 - kx)) * cout + co] = b;
}

}  // namespace isti_nn

using namespace isti_nn;

namespace {

struct LayerDef {
    int ca, cb, cout, ks, up_a, relu;
};
// e0a e0b | e1a e1b | e2a e2b | e3a e3b | d2a d2b | d1a d1b | d0a d0b | head
const LayerDef LAYERS[15] = {
    {16, 0, 16, 3, 0, 1},   {16, 0, 16, 3, 0, 1},   {16, 0, 32, 3, 0, 1},   {32, 0, 32, 3, 0, 1},
    {32, 0, 64, 3, 0, 1},   {64, 0, 64, 3, 0, 1},   {64, 0, 128, 3, 0, 1},  {128, 0, 128, 3, 0, 1},
    {128, 64, 64, 3, 1, 1}, {64, 0, 64, 3, 0, 1},   {64, 32, 32, 3, 1, 1},  {32, 0, 32, 3, 0, 1},
    {32, 16, 16, 3, 1, 1},  {16, 0, 16, 3, 0, 1},   {16, 0, 16, 1, 0, 0},
};
constexpr int NL = 15;
const int LEVEL[NL] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};

size_t w_count(int l) { return (size_t)LAYERS[l].cout * LAYERS[l].ks * LAYERS[l].ks * (LAYERS[l].ca + LAYERS[l].cb); }
size_t total_params() {
    size_t n = 0;
    for (int l = 0; l < NL; l++) n += w_count(l) + LAYERS[l].cout;
    return n;
}

thread_local std::string g_err;

}  // namespace

struct isti_unet {
    int device = 0;
    int N = 0, H = 0, W = 0;
    size_t np = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // fp32 master / grads / Adam
    size_t w_off[NL], b_off[NL];
    uint16_t *wb[NL] = {}, *wt[NL] = {};
    uint16_t *y[NL] = {};     // post-activation outputs (bf16); the head writes fp32 to the caller
    uint16_t *pool[4] = {};   // pooled inputs of levels 1..3
    float *dy[NL] = {};       // gradient of each layer's output
    float *dpool[4] = {};     // gradient of pooled tensors
    float *dcat = nullptr;    // gradient of a layer input (largest size)
    uint16_t *gm = nullptr;   // masked bf16 output gradient (largest size)
    const uint16_t *x = nullptr;
    int adam_t = 0;
    std::vector<void *> allocs;
};

namespace {

size_t px_at(const isti_unet *n, int level) { return (size_t)n->N * (n->H >> level) * (n->W >> level); }

hipError_t dalloc(isti_unet *n, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) n->allocs.push_back(*p);
    return e;
}

template <typename K>
unsigned blocks_for(long n, K) {
    return (unsigned)((n + 255) / 256);
}

template <int NT, int MT>
hipError_t launch_conv_tiles(const Src &s, const uint16_t *w, const float *bias, int cin, int cout, int ks, int N, int H,
                             int W, uint16_t *out_bf, float *out_f, int relu, hipStream_t st) {
    if (ks == 3) {  // LDS-staged input, one block per 4 rows x 16*MT pixels
        const long blocks = (long)N * ((H + 3) / 4) * ((W + 16 * MT - 1) / (16 * MT));
        const dim3 grid((unsigned)blocks, (unsigned)(cout / (16 * NT))), block(256);
        hipLaunchKernelGGL((conv3_lds_kernel<NT, MT>), grid, block, 0, st, s, w, bias, cin, N, H, W, cout, out_bf, out_f,
                           relu);
        return hipGetLastError();
    }
    const long waves = (long)N * H * ((W + 16 * MT - 1) / (16 * MT));
    const dim3 grid((unsigned)((waves + 3) / 4), (unsigned)(cout / (16 * NT))), block(256);
    hipLaunchKernelGGL((conv_kernel<NT, MT>), grid, block, 0, st, s, w, bias, cin, ks, N, H, W, cout, out_bf, out_f,
                       relu);
    return hipGetLastError();
}

// Wave shape (NT output tiles x MT pixel groups): the most register
// blocking that still gives the launch CONV_MIN_WAVES waves (2 per SIMD of
// the chip), else the most waves. Deep levels are small (60 x 34 pixels),
// so their output channels are split over the grid's y dimension.
constexpr long CONV_MIN_WAVES = 8192;

hipError_t launch_conv(const Src &s, const uint16_t *w, const float *bias, int cin, int cout, int ks, int N, int H,
                       int W, uint16_t *out_bf, float *out_f, int relu, hipStream_t st) {
    const int tiles = cout / 16;
    if (tiles < 1 || tiles * 16 != cout) return hipErrorInvalidValue;
    static const int shapes[][2] = {{4, 4}, {4, 2}, {2, 4}, {3, 2}, {4, 1}, {2, 2}, {1, 4}, {3, 1},
                                    {2, 1}, {1, 2}, {1, 1}};
    int pick = -1, most = -1;
    long most_waves = -1;
    for (int i = 0; i < (int)(sizeof(shapes) / sizeof(shapes[0])); i++) {
        const int nt = shapes[i][0], mt = shapes[i][1];
        if (tiles % nt != 0 || (mt == 4 && W < 192) || (mt == 2 && W < 64)) continue;
        const long waves = (long)N * H * ((W + 16 * mt - 1) / (16 * mt)) * (tiles / nt);
        if (waves >= CONV_MIN_WAVES) {
            pick = i;
            break;
        }
        if (waves > most_waves) most_waves = waves, most = i;
    }
    if (pick < 0) pick = most;
    switch (shapes[pick][0] * 8 + shapes[pick][1]) {
        case 4 * 8 + 4: return launch_conv_tiles<4, 4>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        case 4 * 8 + 2: return launch_conv_tiles<4, 2>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        case 2 * 8 + 4: return launch_conv_tiles<2, 4>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        case 3 * 8 + 2: return launch_conv_tiles<3, 2>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        case 4 * 8 + 1: return launch_conv_tiles<4, 1>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        case 2 * 8 + 2: return launch_conv_tiles<2, 2>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        case 1 * 8 + 4: return launch_conv_tiles<1, 4>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        case 3 * 8 + 1: return launch_conv_tiles<3, 1>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        case 2 * 8 + 1: return launch_conv_tiles<2, 1>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        case 1 * 8 + 2: return launch_conv_tiles<1, 2>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
        default: return launch_conv_tiles<1, 1>(s, w, bias, cin, cout, ks, N, H, W, out_bf, out_f, relu, st);
    }
}

// input of layer l in the forward pass
Src layer_src(const isti_unet *n, int l) {
    Src s{nullptr, nullptr, LAYERS[l].ca, LAYERS[l].cb, LAYERS[l].up_a};
    switch (l) {
        case 0: s.a = n->x; break;
        case 2: s.a = n->pool[1]; break;
        case 4: s.a = n->pool[2]; break;
        case 6: s.a = n->pool[3]; break;
        case 8: s.a = n->y[7]; s.b = n->y[5]; break;   // up(e3), e2
        case 10: s.a = n->y[9]; s.b = n->y[3]; break;  // up(d2), e1
        case 12: s.a = n->y[11]; s.b = n->y[1]; break; // up(d1), e0
        default: s.a = n->y[l - 1]; break;
    }
    return s;
}

isti_status fail(hipError_t e, const char *where) {
    g_err = std::string(where) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? ISTI_ERR_OOM : ISTI_ERR_DEVICE;
}

hipError_t refresh_weights(isti_unet *n, hipStream_t st) {
    for (int l = 0; l < NL; l++) {
        const long c = (long)w_count(l);
        hipLaunchKernelGGL(cast_weights_kernel, dim3(blocks_for(c, 0)), dim3(256), 0, st, n->p + n->w_off[l], n->wb[l],
                           n->wt[l], LAYERS[l].cout, LAYERS[l].ca + LAYERS[l].cb, LAYERS[l].ks);
    }
    return hipGetLastError();
}

}  // namespace

extern "C" {

size_t isti_unet_param_count(void) { return total_params(); }

isti_status isti_unet_create(int32_t device, uint32_t width, uint32_t height, int32_t batch, uint64_t seed,
                             isti_unet **out) {
    if (!out || width == 0 || height == 0 || width % 8 || height % 8 || batch < 1 || width > 16384 || height > 16384)
        return ISTI_ERR_INVALID_ARG;
    *out = nullptr;
    isti_unet *n = new (std::nothrow) isti_unet();
    if (!n) return ISTI_ERR_OOM;
    n->device = device;
    n->N = batch;
    n->H = (int)height;
    n->W = (int)width;
    n->np = total_params();
    hipError_t e = hipSetDevice(device);
    size_t off = 0;
    for (int l = 0; l < NL; l++) {
        n->w_off[l] = off;
        off += w_count(l);
        n->b_off[l] = off;
        off += LAYERS[l].cout;
    }
    const size_t pb = n->np * sizeof(float);
    if (e == hipSuccess) e = dalloc(n, (void **)&n->p, pb);
    if (e == hipSuccess) e = dalloc(n, (void **)&n->g, pb);
    if (e == hipSuccess) e = dalloc(n, (void **)&n->m, pb);
    if (e == hipSuccess) e = dalloc(n, (void **)&n->v, pb);
    size_t max_cat = 0, max_out = 0;
    for (int l = 0; l < NL && e == hipSuccess; l++) {
        const size_t px = px_at(n, LEVEL[l]);
        e = dalloc(n, (void **)&n->wb[l], w_count(l) * 2);
        if (e == hipSuccess) e = dalloc(n, (void **)&n->wt[l], w_count(l) * 2);
        if (e == hipSuccess && l < NL - 1) e = dalloc(n, (void **)&n->y[l], px * LAYERS[l].cout * 2);
        if (e == hipSuccess && l < NL - 1) e = dalloc(n, (void **)&n->dy[l], px * LAYERS[l].cout * 4);
        max_cat = std::max(max_cat, px * (size_t)(LAYERS[l].ca + LAYERS[l].cb));
        max_out = std::max(max_out, px * (size_t)LAYERS[l].cout);
    }
    for (int lv = 1; lv <= 3 && e == hipSuccess; lv++) {
        const int c = LAYERS[2 * lv].ca;
        e = dalloc(n, (void **)&n->pool[lv], px_at(n, lv) * c * 2);
        if (e == hipSuccess) e = dalloc(n, (void **)&n->dpool[lv], px_at(n, lv) * c * 4);
    }
    if (e == hipSuccess) e = dalloc(n, (void **)&n->dcat, max_cat * 4);
    if (e == hipSuccess) e = dalloc(n, (void **)&n->gm, max_out * 2);
    if (e != hipSuccess) {
        isti_status st = fail(e, "isti_unet_create");
        isti_unet_destroy(n);
        return st;
    }
    // He initialisation (weights ~ N(0, 2/fan_in)), zero biases; the head's
    // padding outputs 3..15 stay zero.
    std::vector<float> h(n->np, 0.0f);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto unif = [&s]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return ((s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < NL; l++) {
        const int cin = LAYERS[l].ca + LAYERS[l].cb, ks = LAYERS[l].ks;
        const double sd = std::sqrt(2.0 / (cin * ks * ks)) * (l == NL - 1 ? 0.1 : 1.0);
        const int live = l == NL - 1 ? 3 : LAYERS[l].cout;
        for (int co = 0; co < live; co++)
            for (int k = 0; k < ks * ks * cin; k++) {
                const double u1 = unif(), u2 = unif();
                h[n->w_off[l] + (size_t)co * ks * ks * cin + k] =
                    (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
    }
    isti_status st = isti_unet_set_params(n, h.data());
    if (st != ISTI_OK) {
        isti_unet_destroy(n);
        return st;
    }
    e = hipMemset(n->m, 0, pb);
    if (e == hipSuccess) e = hipMemset(n->v, 0, pb);
    if (e == hipSuccess) e = hipMemset(n->g, 0, pb);
    if (e != hipSuccess) {
        st = fail(e, "isti_unet_create");
        isti_unet_destroy(n);
        return st;
    }
    *out = n;
    return ISTI_OK;
}

isti_status isti_unet_destroy(isti_unet *n) {
    if (!n) return ISTI_ERR_INVALID_ARG;
    (void)hipSetDevice(n->device);
    for (void *p : n->allocs) (void)hipFree(p);
    delete n;
    return ISTI_OK;
}

isti_status isti_unet_get_params(isti_unet *n, float *host) {
    if (!n || !host) return ISTI_ERR_INVALID_ARG;
    hipError_t e = hipMemcpy(host, n->p, n->np * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? ISTI_OK : fail(e, "get_params");
}

isti_status isti_unet_set_params(isti_unet *n, const float *host) {
    if (!n || !host) return ISTI_ERR_INVALID_ARG;
    hipError_t e = hipSetDevice(n->device);
    if (e == hipSuccess) e = hipMemcpy(n->p, host, n->np * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = refresh_weights(n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? ISTI_OK : fail(e, "set_params");
}

isti_status isti_unet_get_grads(isti_unet *n, float *host) {
    if (!n || !host) return ISTI_ERR_INVALID_ARG;
    hipError_t e = hipMemcpy(host, n->g, n->np * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? ISTI_OK : fail(e, "get_grads");
}

isti_status isti_unet_forward(isti_unet *n, const uint16_t *x, float *out, void *stream) {
    if (!n || !x || !out) return ISTI_ERR_INVALID_ARG;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = hipSetDevice(n->device);
    n->x = x;
    for (int l = 0; l < NL && e == hipSuccess; l++) {
        const int lv = LEVEL[l];
        const int H = n->H >> lv, W = n->W >> lv;
        if (l == 2 || l == 4 || l == 6) {  // pool the previous level's output
            const int c = LAYERS[l].ca;
            const long cnt = (long)px_at(n, lv) * c;
            hipLaunchKernelGGL(pool_fwd_kernel, dim3(blocks_for(cnt, 0)), dim3(256), 0, st, n->y[l - 1], n->pool[lv],
                               n->N, H * 2, W * 2, c);
        }
        const Src s = layer_src(n, l);
        e = launch_conv(s, n->wb[l], n->p + n->b_off[l], LAYERS[l].ca + LAYERS[l].cb, LAYERS[l].cout, LAYERS[l].ks,
                        n->N, H, W, l < NL - 1 ? n->y[l] : nullptr, l < NL - 1 ? nullptr : out, LAYERS[l].relu, st);
    }
    return e == hipSuccess ? ISTI_OK : fail(e, "forward");
}

isti_status isti_unet_backward(isti_unet *n, const float *dout, void *stream) {
    if (!n || !dout || !n->x) return ISTI_ERR_INVALID_ARG;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = hipSetDevice(n->device);
    for (int l = 0; l < NL - 1 && e == hipSuccess; l++)
        e = hipMemsetAsync(n->dy[l], 0, px_at(n, LEVEL[l]) * LAYERS[l].cout * 4, st);
    for (int lv = 1; lv <= 3 && e == hipSuccess; lv++)
        e = hipMemsetAsync(n->dpool[lv], 0, px_at(n, lv) * LAYERS[2 * lv].ca * 4, st);
    for (int l = NL - 1; l >= 0 && e == hipSuccess; l--) {
        const LayerDef &L = LAYERS[l];
        const int lv = LEVEL[l], H = n->H >> lv, W = n->W >> lv, cin = L.ca + L.cb;
        const long cnt_out = (long)px_at(n, lv) * L.cout;
        // masked output gradient (bf16)
        const float *dyl = l == NL - 1 ? dout : n->dy[l];
        hipLaunchKernelGGL(mask_kernel, dim3(blocks_for(cnt_out, 0)), dim3(256), 0, st,
                           l == NL - 1 ? nullptr : n->y[l], dyl, n->gm, cnt_out, L.relu);
        const Src s = layer_src(n, l);
        // weight and bias gradients
        hipLaunchKernelGGL(wgrad_kernel, dim3((L.cout / 16) * (L.ks * L.ks * cin / 16), 16), dim3(256), 0, st, s,
                           (const uint16_t *)n->gm, L.cout, cin, L.ks, n->N, H, W, n->g + n->w_off[l]);
        hipLaunchKernelGGL(bgrad_kernel, dim3(64), dim3(256), 0, st, (const uint16_t *)n->gm, L.cout,
                           (long)px_at(n, lv), n->g + n->b_off[l]);
        if (l == 0) break;  // no gradient for the network input
        // input gradient: conv of g with flipped/transposed weights
        Src gs{n->gm, nullptr, L.cout, 0, 0};
        e = launch_conv(gs, n->wt[l], nullptr, L.cout, cin, L.ks, n->N, H, W, nullptr, n->dcat, 0, st);
        if (e != hipSuccess) break;
        const long cnt_in = (long)px_at(n, lv) * cin;
        if (L.up_a) {  // decoder: split into the low-resolution input and the skip
            float *d_low = n->dy[l - 1];
            float *d_skip = n->dy[l == 8 ? 5 : (l == 10 ? 3 : 1)];
            hipLaunchKernelGGL(split_kernel, dim3(blocks_for(cnt_in, 0)), dim3(256), 0, st, (const float *)n->dcat, d_low,
                               d_skip, n->N, H, W, L.ca, L.cb);
        } else if (l == 2 || l == 4 || l == 6) {  // pooled input: route to the previous level
            hipLaunchKernelGGL(pool_bwd_kernel, dim3(blocks_for((long)px_at(n, lv - 1) * cin, 0)), dim3(256), 0, st,
                               (const float *)n->dcat, n->dy[l - 1], n->N, H * 2, W * 2, cin);
        } else {  // plain chain: add to the previous layer's output gradient
            e = hipMemcpyAsync(n->dy[l - 1], n->dcat, cnt_in * 4, hipMemcpyDeviceToDevice, st);
        }
        if (e == hipSuccess) e = hipGetLastError();
    }
    return e == hipSuccess ? ISTI_OK : fail(e, "backward");
}

isti_status isti_unet_zero_grads(isti_unet *n, void *stream) {
    if (!n) return ISTI_ERR_INVALID_ARG;
    hipError_t e = hipMemsetAsync(n->g, 0, n->np * 4, reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? ISTI_OK : fail(e, "zero_grads");
}

isti_status isti_unet_adam(isti_unet *n, float lr, float b1, float b2, float eps, void *stream) {
    if (!n) return ISTI_ERR_INVALID_ARG;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    n->adam_t++;
    const float c1 = 1.0f - std::pow(b1, (float)n->adam_t), c2 = 1.0f - std::pow(b2, (float)n->adam_t);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for((long)n->np, 0)), dim3(256), 0, st, n->p, (const float *)n->g, n->m,
                       n->v, (long)n->np, lr, b1, b2, eps, c1, c2);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = refresh_weights(n, st);
    return e == hipSuccess ? ISTI_OK : fail(e, "adam");
}

}  // extern "C"
