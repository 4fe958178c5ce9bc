// codec.hip - byte and quantisation kernels of the keyframe wire codec
// (include/isti_codec.h; the reference's compression.py).
//
// Run-length coding is data-parallel over runs: a byte starts a run when it
// differs from its left neighbour; an exclusive scan of those flags numbers
// the runs and scatters their start positions; each run emits
// ceil(length / 255) pieces, and a second scan of the piece counts gives
// every run its output offset. Decoding scans the piece lengths and expands
// each piece. The scans are three-pass (block sums, one block scanning the
// block sums, block-local scans plus offsets) over 1024 elements per block.
// All passes are HBM-bound streams of bytes and 32-bit counts.

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "../../include/isti_codec.h"

namespace isti_codec_k {

constexpr uint32_t THREADS = 256;
constexpr uint32_t PER = 4;
constexpr uint32_t SCAN_BLOCK = THREADS * PER;  // elements per scan block

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[THREADS / 64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if ((int)lane >= off) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t w = 0; w < THREADS / 64; w++) {
        if (w < wv) base += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return base + x - v;
}

// per block: sum of its SCAN_BLOCK elements
__global__ __launch_bounds__(THREADS) void scan_reduce_kernel(const uint32_t *in, uint32_t n, uint32_t *sums) {
    const uint32_t b0 = blockIdx.x * SCAN_BLOCK;
    uint32_t s = 0;
    for (uint32_t k = 0; k < PER; k++) {
        const uint32_t i = b0 + k * THREADS + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t total;
    (void)block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// one block: exclusive scan of nb block sums in place; total[0] = their sum
__global__ __launch_bounds__(THREADS) void scan_sums_kernel(uint32_t *sums, uint32_t nb, uint32_t *total) {
    uint32_t carry = 0;
    for (uint32_t c = 0; c < nb; c += THREADS) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t t;
        const uint32_t ex = block_exclusive_scan(v, &t);
        if (i < nb) sums[i] = carry + ex;
        carry += t;
    }
    if (threadIdx.x == 0) *total = carry;
}

// exclusive scan of each block's elements plus its block offset
__global__ __launch_bounds__(THREADS) void scan_final_kernel(const uint32_t *in, uint32_t n, const uint32_t *sums,
                                                             uint32_t *out) {
    const uint32_t b0 = blockIdx.x * SCAN_BLOCK + threadIdx.x * PER;  // PER consecutive elements per thread
    uint32_t v[PER], s = 0;
    for (uint32_t k = 0; k < PER; k++) {
        v[k] = b0 + k < n ? in[b0 + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = sums[blockIdx.x] + block_exclusive_scan(s, &total);
    for (uint32_t k = 0; k < PER; k++) {
        if (b0 + k < n) out[b0 + k] = run;
        run += v[k];
    }
}

__global__ void run_flags_kernel(const uint8_t *a, uint32_t n, uint32_t *f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = (i == 0 || a[i] != a[i - 1]) ? 1u : 0u;
}

// run r starts at S[r]; S[R] = n closes the last run
__global__ void run_starts_kernel(const uint32_t *f, const uint32_t *pos, uint32_t n, const uint32_t *n_runs,
                                  uint32_t *S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && f[i]) S[pos[i]] = i;
    if (i == 0) S[*n_runs] = n;
}

__global__ void run_pieces_kernel(const uint32_t *S, const uint32_t *n_runs, uint32_t *P) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < *n_runs) P[r] = (S[r + 1] - S[r] + 254u) / 255u;
}

// compression.py:186-202: pieces of 255 (the last shorter); count 0 for 1
__global__ void run_write_kernel(const uint8_t *a, const uint32_t *S, const uint32_t *n_runs, const uint32_t *off,
                                 uint8_t *out) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= *n_runs) return;
    const uint32_t s = S[r], len = S[r + 1] - s;
    const uint8_t v = a[s];
    uint8_t *o = out + 2ull * off[r];
    for (uint32_t done = 0, j = 0; done < len; j++) {
        const uint32_t c = min(255u, len - done);
        o[2 * j] = c > 1 ? (uint8_t)c : (uint8_t)0;
        o[2 * j + 1] = v;
        done += c;
    }
}

// compression.py:206-218: a count of 0 stands for one byte
__global__ void piece_len_kernel(const uint8_t *in, uint32_t n_pieces, uint32_t *len) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n_pieces) len[j] = in[2 * j] ? in[2 * j] : 1u;
}

__global__ void piece_expand_kernel(const uint8_t *in, uint32_t n_pieces, const uint32_t *len, const uint32_t *off,
                                    uint8_t *out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_pieces) return;
    const uint8_t v = in[2 * j + 1];
    uint8_t *o = out + off[j];
    for (uint32_t k = 0; k < len[j]; k++) o[k] = v;
}

__global__ void xor_kernel(const uint8_t *a, const uint8_t *b, uint64_t n, uint8_t *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] ^ b[i];
}

__global__ void extract_kernel(const uint8_t *rgba, uint64_t n_px, int ch, uint8_t *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_px) out[i] = rgba[4 * i + ch];
}

__global__ void expand_kernel(const uint8_t *plane, uint64_t n_px, uint8_t *rgba) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_px) reinterpret_cast<uchar4 *>(rgba)[i] = make_uchar4(0, 0, plane[i], 255);
}

// load_exr's clean-up (compression.py:63): |v| < 1e-4 reads as 0
__device__ __forceinline__ float cleaned(float v, int zero_small) {
    return (zero_small && fabsf(v) < 1e-4f) ? 0.0f : v;
}

// NaN-propagating min / max, as numpy's np.min / np.max
__device__ __forceinline__ float nmin(float a, float b) { return (a != a || b != b) ? __builtin_nanf("") : fminf(a, b); }
__device__ __forceinline__ float nmax(float a, float b) { return (a != a || b != b) ? __builtin_nanf("") : fmaxf(a, b); }

__global__ __launch_bounds__(THREADS) void minmax_kernel(const float *in, uint64_t n, int zero_small, float2 *part) {
    __shared__ float smin[THREADS], smax[THREADS];
    float lo = INFINITY, hi = -INFINITY;
    for (uint64_t i = (uint64_t)blockIdx.x * THREADS + threadIdx.x; i < n; i += (uint64_t)gridDim.x * THREADS) {
        const float v = cleaned(in[i], zero_small);
        lo = nmin(lo, v);
        hi = nmax(hi, v);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (uint32_t off = THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            smin[threadIdx.x] = nmin(smin[threadIdx.x], smin[threadIdx.x + off]);
            smax[threadIdx.x] = nmax(smax[threadIdx.x], smax[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = make_float2(smin[0], smax[0]);
}

// compression.py:88: np.round(np.clip((v - min) / (max - min), 0, 1) * 65535)
// in binary32, round half to even; a NaN converts to 0
__global__ void quantize_kernel(const float *in, uint64_t n, int zero_small, float mn, float range, int degenerate,
                                uint16_t *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (degenerate) {
        out[i] = 0;
        return;
    }
    float t = (cleaned(in[i], zero_small) - mn) / range;
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);  // NaN stays NaN
    const float q = rintf(t * 65535.0f);
    out[i] = q == q ? (uint16_t)q : (uint16_t)0;
}

// compression.py:97: ((q / 65535) * (max - min) + min).astype(float16)
__global__ void dequantize_kernel(const uint16_t *in, uint64_t n, float mn, float range, int degenerate, float *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (degenerate) {
        out[i] = 0.0f;
        return;
    }
    const float v = ((float)in[i] / 65535.0f) * range + mn;
    out[i] = __half2float(__float2half_rn(v));
}

}  // namespace isti_codec_k

using namespace isti_codec_k;

struct isti_codec {
    int device = 0;
    uint64_t cap = 0;
    uint32_t *flags = nullptr, *pos = nullptr, *starts = nullptr, *pieces = nullptr, *offs = nullptr;
    uint32_t *sums = nullptr, *totals = nullptr;  // totals[0]: runs, [1]: pieces / bytes
    float2 *parts = nullptr;
    std::string err;
};

namespace {

constexpr uint32_t MINMAX_BLOCKS = 1024;

isti_status cfail(isti_codec *c, isti_status st, const std::string &m) {
    if (c) c->err = m;
    return st;
}
isti_status cdev(isti_codec *c, hipError_t e, const char *where) {
    return cfail(c, e == hipErrorOutOfMemory ? ISTI_ERR_OOM : ISTI_ERR_DEVICE,
                 std::string(where) + ": " + hipGetErrorString(e));
}
uint32_t grid_for(uint64_t n) { return (uint32_t)((n + THREADS - 1) / THREADS); }

// exclusive scan of n u32 (n >= 1) into out; *total (device) = their sum
hipError_t scan(isti_codec *c, const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *total, hipStream_t s) {
    const uint32_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(THREADS), 0, s, in, n, c->sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(THREADS), 0, s, c->sums, nb, total);
    hipLaunchKernelGGL(scan_final_kernel, dim3(nb), dim3(THREADS), 0, s, in, n, c->sums, out);
    return hipGetLastError();
}

void free_all(isti_codec *c) {
    for (void *p : {(void *)c->flags, (void *)c->pos, (void *)c->starts, (void *)c->pieces, (void *)c->offs,
                    (void *)c->sums, (void *)c->totals, (void *)c->parts})
        if (p) (void)hipFree(p);
}

}  // namespace

extern "C" {

isti_status isti_codec_create(int32_t device, uint64_t max_bytes, isti_codec **out) {
    if (!out || max_bytes == 0 || max_bytes > 0xFFFFFFF0ull) return ISTI_ERR_INVALID_ARG;
    *out = nullptr;
    isti_codec *c = new (std::nothrow) isti_codec();
    if (!c) return ISTI_ERR_OOM;
    c->device = device;
    c->cap = max_bytes;
    hipError_t e = hipSetDevice(device);
    const size_t n = (size_t)max_bytes + 1;
    const size_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK + 1;
    if (e == hipSuccess) e = hipMalloc(&c->flags, n * 4);
    if (e == hipSuccess) e = hipMalloc(&c->pos, n * 4);
    if (e == hipSuccess) e = hipMalloc(&c->starts, (n + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&c->pieces, n * 4);
    if (e == hipSuccess) e = hipMalloc(&c->offs, n * 4);
    if (e == hipSuccess) e = hipMalloc(&c->sums, nb * 4);
    if (e == hipSuccess) e = hipMalloc(&c->totals, 16);
    if (e == hipSuccess) e = hipMalloc(&c->parts, MINMAX_BLOCKS * sizeof(float2));
    if (e != hipSuccess) {
        free_all(c);
        delete c;
        return e == hipErrorOutOfMemory ? ISTI_ERR_OOM : ISTI_ERR_DEVICE;
    }
    *out = c;
    return ISTI_OK;
}

isti_status isti_codec_destroy(isti_codec *c) {
    if (!c) return ISTI_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    free_all(c);
    delete c;
    return ISTI_OK;
}

const char *isti_codec_last_error(const isti_codec *c) { return c ? c->err.c_str() : "null codec"; }

isti_status isti_codec_rle_encode(isti_codec *c, const uint8_t *in, uint64_t n, uint8_t *out, uint64_t out_cap,
                                  uint64_t *out_len, void *stream) {
    if (!c) return ISTI_ERR_INVALID_ARG;
    if (!out_len || (n && (!in || !out))) return cfail(c, ISTI_ERR_INVALID_ARG, "null argument");
    if (n > c->cap) return cfail(c, ISTI_ERR_UNSUPPORTED_SHAPE, "input larger than the codec's max_bytes");
    *out_len = 0;
    if (n == 0) return ISTI_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = hipSetDevice(c->device);
    const uint32_t nn = (uint32_t)n;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(run_flags_kernel, dim3(grid_for(nn)), dim3(THREADS), 0, s, in, nn, c->flags);
        e = scan(c, c->flags, nn, c->pos, c->totals, s);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(run_starts_kernel, dim3(grid_for(nn)), dim3(THREADS), 0, s, c->flags, c->pos, nn,
                           c->totals, c->starts);
        hipLaunchKernelGGL(run_pieces_kernel, dim3(grid_for(nn)), dim3(THREADS), 0, s, c->starts, c->totals,
                           c->pieces);
        e = hipGetLastError();
    }
    uint32_t runs = 0, pieces = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&runs, c->totals, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = scan(c, c->pieces, runs, c->offs, c->totals + 1, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&pieces, c->totals + 1, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return cdev(c, e, "rle encode");
    if (2ull * pieces > out_cap) return cfail(c, ISTI_ERR_INVALID_ARG, "rle output does not fit out_cap");
    hipLaunchKernelGGL(run_write_kernel, dim3(grid_for(runs)), dim3(THREADS), 0, s, in, c->starts, c->totals,
                       c->offs, out);
    if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return cdev(c, e, "rle write");
    *out_len = 2ull * pieces;
    return ISTI_OK;
}

isti_status isti_codec_rle_decode(isti_codec *c, const uint8_t *in, uint64_t n, uint8_t *out, uint64_t out_cap,
                                  uint64_t *out_len, void *stream) {
    if (!c) return ISTI_ERR_INVALID_ARG;
    if (!out_len || (n && !in)) return cfail(c, ISTI_ERR_INVALID_ARG, "null argument");
    if (n & 1) return cfail(c, ISTI_ERR_INVALID_ARG, "an rle stream holds (count, value) pairs: even length");
    if (n / 2 > c->cap) return cfail(c, ISTI_ERR_UNSUPPORTED_SHAPE, "input larger than the codec's max_bytes");
    if ((n / 2) * 255ull > 0xFFFFFFFFull)  // the expanded length must stay a 32-bit count
        return cfail(c, ISTI_ERR_UNSUPPORTED_SHAPE, "rle stream too long (at most 16843009 pieces per call)");
    *out_len = 0;
    if (n == 0) return ISTI_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = hipSetDevice(c->device);
    const uint32_t np = (uint32_t)(n / 2);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(piece_len_kernel, dim3(grid_for(np)), dim3(THREADS), 0, s, in, np, c->pieces);
        e = scan(c, c->pieces, np, c->offs, c->totals + 1, s);
    }
    uint32_t bytes = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bytes, c->totals + 1, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return cdev(c, e, "rle decode");
    if (bytes > out_cap || !out) return cfail(c, ISTI_ERR_INVALID_ARG, "decoded bytes do not fit out_cap");
    hipLaunchKernelGGL(piece_expand_kernel, dim3(grid_for(np)), dim3(THREADS), 0, s, in, np, c->pieces, c->offs, out);
    if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return cdev(c, e, "rle expand");
    *out_len = bytes;
    return ISTI_OK;
}

isti_status isti_codec_xor(const uint8_t *a, const uint8_t *b, uint64_t n, uint8_t *out, void *stream) {
    if (n == 0) return ISTI_OK;
    if (!a || !b || !out) return ISTI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(xor_kernel, dim3(grid_for(n)), dim3(THREADS), 0, reinterpret_cast<hipStream_t>(stream), a, b,
                       n, out);
    return hipGetLastError() == hipSuccess ? ISTI_OK : ISTI_ERR_DEVICE;
}

isti_status isti_codec_extract_channel(const uint8_t *rgba, uint64_t n_px, int32_t channel, uint8_t *out,
                                       void *stream) {
    if (n_px == 0) return ISTI_OK;
    if (!rgba || !out || channel < 0 || channel > 3) return ISTI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(extract_kernel, dim3(grid_for(n_px)), dim3(THREADS), 0, reinterpret_cast<hipStream_t>(stream),
                       rgba, n_px, channel, out);
    return hipGetLastError() == hipSuccess ? ISTI_OK : ISTI_ERR_DEVICE;
}

isti_status isti_codec_expand_channel(const uint8_t *plane, uint64_t n_px, uint8_t *rgba, void *stream) {
    if (n_px == 0) return ISTI_OK;
    if (!plane || !rgba) return ISTI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(expand_kernel, dim3(grid_for(n_px)), dim3(THREADS), 0, reinterpret_cast<hipStream_t>(stream),
                       plane, n_px, rgba);
    return hipGetLastError() == hipSuccess ? ISTI_OK : ISTI_ERR_DEVICE;
}

isti_status isti_codec_normalize(isti_codec *c, const float *in, uint64_t n, int32_t zero_small, uint16_t *out,
                                 float *min_max, void *stream) {
    if (!c) return ISTI_ERR_INVALID_ARG;
    if (!in || !out || !min_max || n == 0) return cfail(c, ISTI_ERR_INVALID_ARG, "null argument or empty input");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = hipSetDevice(c->device);
    const uint32_t nb = (uint32_t)std::min<uint64_t>(MINMAX_BLOCKS, (n + THREADS - 1) / THREADS);
    std::vector<float2> part(nb);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(minmax_kernel, dim3(nb), dim3(THREADS), 0, s, in, n, zero_small, c->parts);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(part.data(), c->parts, nb * sizeof(float2), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return cdev(c, e, "normalize");
    float mn = part[0].x, mx = part[0].y;
    for (uint32_t b = 1; b < nb; b++) {
        const float a = part[b].x, z = part[b].y;
        mn = (mn != mn || a != a) ? NAN : std::min(mn, a);
        mx = (mx != mx || z != z) ? NAN : std::max(mx, z);
    }
    const float range = mx - mn;
    const int degenerate = std::fabs(range) <= 1e-6f ? 1 : 0;  // NaN is not <= 1e-6
    hipLaunchKernelGGL(quantize_kernel, dim3(grid_for(n)), dim3(THREADS), 0, s, in, n, zero_small, mn, range,
                       degenerate, out);
    if ((e = hipGetLastError()) != hipSuccess) return cdev(c, e, "quantize");
    min_max[0] = degenerate ? 0.0f : mn;
    min_max[1] = degenerate ? 0.0f : mx;
    return ISTI_OK;
}

isti_status isti_codec_denormalize(const uint16_t *in, uint64_t n, float min, float max, float *out, void *stream) {
    if (n == 0) return ISTI_OK;
    if (!in || !out) return ISTI_ERR_INVALID_ARG;
    const float range = max - min;
    const int degenerate = std::fabs(range) <= 1e-6f ? 1 : 0;
    hipLaunchKernelGGL(dequantize_kernel, dim3(grid_for(n)), dim3(THREADS), 0, reinterpret_cast<hipStream_t>(stream),
                       in, n, min, range, degenerate, out);
    return hipGetLastError() == hipSuccess ? ISTI_OK : ISTI_ERR_DEVICE;
}

}  // extern "C"
