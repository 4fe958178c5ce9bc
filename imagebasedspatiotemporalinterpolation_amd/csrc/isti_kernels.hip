// isti_kernels.hip - gfx950 kernels of the per-frame interpolation path.
//
//   clear_splat_kernel      splat buffers <- 0xfe.. (renderer.cu:631-638)
//   reproject_splat_kernel  K1, every source view of the frame in ONE launch
//                           (renderer.cu:76-146, launched 2+2*n_aux times
//                           by the reference, :643-649)
//   resolve_*_kernel        K2 (renderer.cu:148-465) over a batch work table
//                           and with learned starts; the per-render resolve
//                           kernels are in resolve_kernels.hip, the shared
//                           device code of K2 in resolve_device.hpp
//   row_sum_kernel,
//   mean_kernel             deterministic mean z (replaces the float
//                           atomicAdd of image_buffer.cu:27-36,80-101)
//   timewarp_kernel         mode 4 baseline (renderer.cu:467-522)
//
// Launch geometry is wave64-first: 256-thread blocks (4 waves). Block ids are
// remapped so that consecutive tiles land on the same XCD (the hardware
// deals workgroups round-robin over the 8 XCDs, each with its own L2), which
// keeps the neighbourhood gathers of adjacent tiles in one L2.

#include "resolve_device.hpp"

namespace isti {

// ---------------------------------------------------------------------------
// Splat clear: 16-byte stores over every buffer of the frame.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void clear_pairs(unsigned long long *buf, uint32_t n, uint32_t pair) {
    const uint32_t i0 = pair * 2u;  // 2 cells per thread
    if (i0 + 1u < n) {
        reinterpret_cast<ulonglong2 *>(buf)[pair] = make_ulonglong2(SPLAT_EMPTY, SPLAT_EMPTY);
    } else if (i0 < n) {
        buf[i0] = SPLAT_EMPTY;
    }
}

__global__ __launch_bounds__(256) void clear_splat_kernel(const ClearArgs a) {
    const uint32_t b = blockIdx.x;
    int j = 0;
    for (int i = 1; i < a.n_bufs; i++)
        if (b >= a.block_begin[i]) j = i;
    clear_pairs(a.buf[j], a.n[j], (b - a.block_begin[j]) * 256u + threadIdx.x);
}

// ---------------------------------------------------------------------------
// K1: reprojection + depth-sorted splat (renderer.cu:76-146).
// One thread per source pixel, row-major, so z / motion loads and the two
// float4 flow stores are fully coalesced; the u64 atomicMin is the only
// scattered access.
// ---------------------------------------------------------------------------
// What a source pixel of a keyframe view contributes to every frame of its
// keyframe pair: its clip-space position, its world position at the
// keyframe and moved to the other keyframe, and its visible-in-main flag.
struct PairPoint {
    float4 cs_from;
    float4 ws_from, ws_to;
    float vis;
};

template <typename J_t, typename M_t>
__device__ __forceinline__ PairPoint pair_point(const J_t &J, const M_t &vp_main, uint32_t x, uint32_t y, uint32_t p) {
    PairPoint q;
    const float z_from = J.z[p];
    float4 om = make_float4(0.f, 0.f, 0.f, 0.f);
    if (J.mx != nullptr) om = make_float4(J.mx[p], J.my[p], J.mz[p], 0.f);

    q.cs_from = make_float4(cs_coord(x, J.inv_w), cs_coord(y, J.inv_h), z_from, 1.0f);
    q.ws_from = project_unit(J.inv_vp_from, q.cs_from);
    const float4 cs_moved = make_float4(q.cs_from.x + om.x, q.cs_from.y + om.y, q.cs_from.z + om.z, q.cs_from.w + 0.0f);
    q.ws_to = project_unit(J.inv_vp_from, cs_moved);

    // visible in the main view: -lim < RN32(c / w) < lim for c = x, y, z, with
    // lim = RN32(1 - RN32(1e-3)) = 16760439 * 2^-24. RN32 is monotone and lim's
    // significand is odd, so RN32(q) < lim exactly when q <= LIM_MID, the
    // midpoint below lim (a tie rounds to the even neighbour below). For w
    // finite and non-zero that is |c| <= LIM_MID |w|, and LIM_MID (25 bits)
    // times w (24 bits) is exact in binary64: no division is needed. A zero,
    // infinite or NaN w takes the IEEE quotients. Only an aux view's flag is
    // ever read (the aux candidate's, renderer.cu:455; the main camera's
    // result never consults its own), so a main view (main_slot < 0, uniform
    // per job) skips the test and stores 0.
    bool visible = false;
    if (J.main_slot >= 0) {
        const float4 rm = mat_vec(vp_main, q.ws_from);
        if (__builtin_amdgcn_classf(rm.w, CLASS_FINITE_NONZERO)) {
            const double bound = LIM_MID * __builtin_fabs((double)rm.w);
            visible = __builtin_fabs((double)rm.x) <= bound && __builtin_fabs((double)rm.y) <= bound &&
                      __builtin_fabs((double)rm.z) <= bound;
        } else {
            const float lim = 1.0f - 1e-3f;
            const float qx = rm.x / rm.w, qy = rm.y / rm.w, qz = rm.z / rm.w;
            visible = qx > -lim && qx < lim && qy > -lim && qy < lim && qz > -lim && qz < lim;
        }
    }
    q.vis = visible ? 1.0f : 0.0f;
    return q;
}

// The flow toward the other keyframe (renderer.cu:119-127 of the splat):
// cs_to - cs_from.
template <typename J_t>
__device__ __forceinline__ float3 full_flow(const J_t &J, const PairPoint &q) {
    const float4 cs_to = project(J.vp_to, q.ws_to);
    return make_float3(cs_to.x - q.cs_from.x, cs_to.y - q.cs_from.y, cs_to.z - q.cs_from.z);
}

// One frame of the pair (weight wt toward the other keyframe, render camera
// vp_cam): the cam_ff record of the source pixel, and whether it splats, with
// its target pixel and key.
template <typename M_t>
__device__ __forceinline__ bool frame_point(const PairPoint &q, const M_t &vp_cam, float wt, uint32_t x, uint32_t y,
                                            uint32_t w, uint32_t h, float4 &cam, uint32_t &tx, uint32_t &ty,
                                            unsigned long long &key) {
    // lerp of the two world points; both w are 1, so the lerped w is
    // 1 + wt * 0: 1 for a finite weight and NaN (carried in x) otherwise
    const float sx = q.ws_from.x + wt * (q.ws_to.x - q.ws_from.x);
    const float4 ws_scaled = make_float4(__builtin_isfinite(wt) ? sx : __builtin_nanf(""),
                                         q.ws_from.y + wt * (q.ws_to.y - q.ws_from.y),
                                         q.ws_from.z + wt * (q.ws_to.z - q.ws_from.z), 1.0f);
    const float4 cs_cam = project(vp_cam, ws_scaled);
    cam = make_float4(cs_cam.x - q.cs_from.x, cs_cam.y - q.cs_from.y, q.cs_from.z + (cs_cam.z - q.cs_from.z), q.vis);
    const float depth_cam = z_to_depth(cs_cam.z);
    const int2 pc = cs2pixel(cs_cam.x, cs_cam.y, w, h);
    if (!in_range(pc, w, h) || depth_cam < 0.0f) return false;
    tx = (uint32_t)pc.x;
    ty = (uint32_t)pc.y;
    key = encode_key(depth_cam, x, y);
    return true;
}

// Reprojection of source pixel (x, y) (p = y w + x): writes full_ff and
// cam_ff and returns whether it splats, with its target pixel and key.
template <typename J_t, typename M_t>
__device__ __forceinline__ bool project_pixel(const J_t &J, const M_t &vp_cam, const M_t &vp_main, bool init_linear,
                                              uint32_t x, uint32_t y, uint32_t p, uint32_t &tx, uint32_t &ty,
                                              unsigned long long &key) {
    const PairPoint q = pair_point(J, vp_main, x, y, p);
    const float wt = J.weight;
    // The flow toward the other keyframe depends on the keyframe pair alone
    // (not on the render camera or alpha): the first render of a pair writes
    // it and later renders keep it (write_full is uniform per job). The
    // linear initialisation scales it into cam_ff, so it computes it anyway.
    float3 f = make_float3(0.0f, 0.0f, 0.0f);
    if (J.write_full || init_linear) {
        f = full_flow(J, q);
        if (J.write_full) J.full_ff[p] = Flow3{f.x, f.y, f.z};
    }
    if (!init_linear) {
        float4 cam;
        const bool hit = frame_point(q, vp_cam, wt, x, y, J.w, J.h, cam, tx, ty, key);
        J.cam_ff[p] = cam;
        return hit;
    }
    J.cam_ff[p] = make_float4(f.x * wt, f.y * wt, q.cs_from.z + f.z * wt, q.vis);
    tx = x;
    ty = y;
    key = encode_key(z_to_depth(q.cs_from.z), x, y);
    return true;
}

// Direct splat: one 64-bit atomicMin into HBM per splatting source pixel.
template <typename J_t, typename M_t>
__device__ __forceinline__ void splat_pixel(const J_t &J, const M_t &vp_cam, const M_t &vp_main, bool init_linear,
                                            uint32_t p) {
    if (p >= J.n_px) return;
    const uint32_t y = div_px(p, J.w, J.rcp_w);
    const uint32_t x = p - __umul24(y, J.w);
    uint32_t tx, ty;
    unsigned long long key;
    if (project_pixel(J, vp_cam, vp_main, init_linear, x, y, p, tx, ty, key))
        atomicMin(&J.splat[ty * J.w + tx], key);
}

__global__ __launch_bounds__(256) void reproject_splat_kernel(const SplatArgs a) {
    const uint32_t b = xcd_remap(blockIdx.x, a.n_blocks);
    int j = 0;
    for (int i = 1; i < a.n_jobs; i++)
        if (b >= a.job[i].block_begin) j = i;
    const SplatJob &J = a.job[j];
    splat_pixel(J, a.vp_cam, a.vp_main[J.main_slot < 0 ? 0 : J.main_slot], a.init_linear != 0, (b - J.block_begin) * SPLAT_BLOCK + threadIdx.x);
}

// LDS-binned splat: a block reprojects a 32x8 source tile and resolves the
// atomicMin of its targets in an LDS bin (64-bit ds_min) before they reach
// HBM; each non-empty bin cell then costs one global atomicMin, and targets
// outside the bin go straight to HBM. The minimum is order-independent, so
// the buffer is the same as the direct splat's.
// The bin follows the targets, not the sources: a moving camera shifts a
// whole tile's targets by tens to hundreds of pixels (45 px on the 1080p
// workload, 200 px on the fast stereo camera), while within a tile they
// stay within a few pixels of one common shift. The anchor is the shift of
// the first splatting source pixel of the block (in the waves' order), and
// the bin covers the source tile moved by it, plus BIN_M cells a side,
// which holds 95-98 % of the targets on those workloads.
constexpr int BIN_TW = 32, BIN_TH = 8, BIN_M = 4;
constexpr int BIN_W = BIN_TW + 2 * BIN_M, BIN_H = BIN_TH + 2 * BIN_M;
constexpr int BIN_CELLS = BIN_W * BIN_H;
constexpr int32_t BIN_NO_ANCHOR = INT32_MIN;

struct SplatBin {
    unsigned long long cell[BIN_CELLS];
    int32_t anchor[2][256 / 64];  // per wave: bin origin x, y (before the margin)
};

// Source tile t (row-major over 32x8 tiles) of job J through the bin. Every
// thread of the block calls it; it returns with the bin free for the next
// tile only after a block barrier by the caller.
template <typename J_t, typename M_t>
__device__ __forceinline__ void splat_tile_binned(const J_t &J, const M_t &vp_cam, const M_t &vp_main,
                                                  bool init_linear, uint32_t t, SplatBin &sb) {
    const uint32_t tiles_x = (J.w + BIN_TW - 1) / BIN_TW;
    const uint32_t tyi = t / tiles_x;
    const int x0 = (int)((t - tyi * tiles_x) * BIN_TW), y0 = (int)(tyi * BIN_TH);
    for (int i = (int)threadIdx.x; i < BIN_CELLS; i += 256) sb.cell[i] = SPLAT_EMPTY;
    const int lx = (int)(threadIdx.x % BIN_TW), ly = (int)(threadIdx.x / BIN_TW);
    const uint32_t x = (uint32_t)(x0 + lx), y = (uint32_t)(y0 + ly);
    uint32_t tx = 0, ty = 0;
    unsigned long long key = 0;
    bool hit = false;
    if (x < J.w && y < J.h) hit = project_pixel(J, vp_cam, vp_main, init_linear, x, y, y * J.w + x, tx, ty, key);
    // the wave's first splatting lane proposes the bin origin: its target
    // minus its place in the tile
    const unsigned long long hits = __ballot(hit);
    const uint32_t wave = threadIdx.x / 64u;
    if (hits != 0ull) {
        const int src = __ffsll((long long)hits) - 1;
        const int ox = __builtin_amdgcn_readlane((int)tx - lx, src);
        const int oy = __builtin_amdgcn_readlane((int)ty - ly, src);
        if ((threadIdx.x & 63u) == 0u) {
            sb.anchor[0][wave] = ox;
            sb.anchor[1][wave] = oy;
        }
    } else if ((threadIdx.x & 63u) == 0u) {
        sb.anchor[0][wave] = BIN_NO_ANCHOR;
    }
    __syncthreads();
    int bx0 = BIN_NO_ANCHOR, by0 = 0;
#pragma unroll
    for (int w = 256 / 64 - 1; w >= 0; w--) {  // the first wave with a splatting lane
        if (sb.anchor[0][w] != BIN_NO_ANCHOR) {
            bx0 = sb.anchor[0][w] - BIN_M;
            by0 = sb.anchor[1][w] - BIN_M;
        }
    }
    if (bx0 == BIN_NO_ANCHOR) return;  // nothing in this tile splats (block-uniform)
    if (hit) {
        const uint32_t bx = tx - (uint32_t)bx0, by = ty - (uint32_t)by0;  // wraps when left of / above the bin
        if (bx < (uint32_t)BIN_W && by < (uint32_t)BIN_H) atomicMin(&sb.cell[by * BIN_W + bx], key);
        else atomicMin(&J.splat[ty * J.w + tx], key);
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < BIN_CELLS; i += 256) {
        const unsigned long long k = sb.cell[i];
        if (k != SPLAT_EMPTY) {  // only in-view targets were binned
            const int gx = bx0 + i % BIN_W, gy = by0 + i / BIN_W;
            atomicMin(&J.splat[(uint32_t)gy * J.w + (uint32_t)gx], k);
        }
    }
}

__global__ __launch_bounds__(256) void reproject_splat_binned_kernel(const SplatArgs a) {
    __shared__ SplatBin sb;
    const uint32_t b = xcd_remap(blockIdx.x, a.n_blocks);
    int j = 0;
    for (int i = 1; i < a.n_jobs; i++)
        if (b >= a.job[i].block_begin) j = i;
    const SplatJob &J = a.job[j];
    splat_tile_binned(J, a.vp_cam, a.vp_main[J.main_slot < 0 ? 0 : J.main_slot], a.init_linear != 0,
                      b - J.block_begin, sb);
}

// Learned search starts: the same resolve, with the main camera's searches
// starting where the network says (see resolve_tile, LEARNED).
template <int NAUX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(resolve_min_waves_wide(1, NAUX))))
void resolve_learned_kernel(const ResolveArgs args) {
#if defined(__HIP_DEVICE_COMPILE__)
    const KernArg<ResolveArgs> &a = *kernarg_ptr<ResolveArgs>();
#else
    const ResolveArgs &a = args;
#endif
    __shared__ ResolveLds<1, NAUX> lds;
    resolve_tile<1, false, NAUX, true>(a, xcd_stripe_tile(blockIdx.x, a.n_tiles, a.tiles_x), lds);
}

// ---------------------------------------------------------------------------
// Network input: 15 features per cell and direction (see isti_learned.h),
// sampled at the cell's top-left full-resolution pixel; channel 15 and cells
// outside the image are zero.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void features_kernel(const FeatureArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t cells = a.net_w * a.net_h;
    if (i >= 2u * cells) return;
    const uint32_t d = i / cells, c = i - d * cells;
    const uint32_t cy = c / a.net_w, cx = c - cy * a.net_w;
    const uint32_t x = cx << a.net_shift, y = cy << a.net_shift;
    float f[16];
#pragma unroll
    for (int k = 0; k < 16; k++) f[k] = 0.0f;
    if (x < a.w && y < a.h) {
        const ResolveView &own = a.v[d], &oth = a.v[1 - d];
        const uint32_t p = y * a.w + x;
        const uchar4 co = own.img[p], cq = oth.img[p];
        const float4 mo = own.cam_ff[p], mq = oth.cam_ff[p];
        // displacements in network cells (2^net_shift render pixels), so a
        // network sees the same numbers for the same scene at any render size
        const float cell = 1.0f / (float)(1u << a.net_shift);
        const float hw = 0.5f * (float)a.w * cell, hh = 0.5f * (float)a.h * cell;
        f[0] = co.x / 255.0f; f[1] = co.y / 255.0f; f[2] = co.z / 255.0f;
        f[3] = cq.x / 255.0f; f[4] = cq.y / 255.0f; f[5] = cq.z / 255.0f;
        f[6] = own.z[p]; f[7] = oth.z[p];
        // z motion back from the moved z the records hold
        f[8] = mo.x * hw; f[9] = mo.y * hh; f[10] = mo.z - f[6];
        f[11] = mq.x * hw; f[12] = mq.y * hh; f[13] = mq.z - f[7];
        f[14] = d == 0 ? a.alpha : 1.0f - a.alpha;
    }
    uint16_t *o = a.x + (size_t)i * 16;
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
        uint32_t u0 = __float_as_uint(f[k]), u1 = __float_as_uint(f[k + 1]);
        u0 = (u0 + 0x7fffu + ((u0 >> 16) & 1u)) >> 16;  // round to nearest even (finite inputs)
        u1 = (u1 + 0x7fffu + ((u1 >> 16) & 1u)) >> 16;
        reinterpret_cast<uint32_t *>(o)[k / 2] = u0 | (u1 << 16);
    }
}

// Training loss against captured plain-search results (see isti_learned.h).
__global__ __launch_bounds__(256) void learned_loss_kernel(const float *__restrict__ out, const uint32_t *__restrict__ q,
                                                           const float4 *__restrict__ ff_fwd,
                                                           const float4 *__restrict__ ff_bwd, uint32_t w, uint32_t h,
                                                           uint32_t net_w, uint32_t net_h, uint32_t shift,
                                                           float conf_weight, float *dout, float *loss_acc) {
    __shared__ float red[256];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t cells = net_w * net_h;
    float loss = 0.0f;
    if (i < 2u * cells) {
        const uint32_t d = i / cells, c = i - d * cells;
        const uint32_t cy = c / net_w, cx = c - cy * net_w;
        const uint32_t x = cx << shift, y = cy << shift;
        const float *o = out + (size_t)i * 16;
        float *g = dout + (size_t)i * 16;
#pragma unroll
        for (int k = 0; k < 16; k++) g[k] = 0.0f;
        if (x < w && y < h) {
            const float norm = 1.0f / (2.0f * (float)cells);
            const float cell = 1.0f / (float)(1u << shift);
            const uint32_t t = q[2 * (y * w + x) + d];
            const bool valid = t != 0xffffffffu;
            bool close = false;  // confidence target: a valid target within one cell
            if (valid) {
                // target residual in cells: the start (x + 0.5 - m.x w/2) + o cell
                // lands on the centre of the found pixel
                const float4 m = (d ? ff_bwd : ff_fwd)[y * w + x];
                const float tx = ((float)(t & 0xFFFFu) - (float)x + m.x * (0.5f * (float)w)) * cell;
                const float ty = ((float)(t >> 16) - (float)y + m.y * (0.5f * (float)h)) * cell;
                const float ex = o[0] - tx, ey = o[1] - ty;
                const float ax = fabsf(ex), ay = fabsf(ey);
                loss += (ax < 1.0f ? 0.5f * ex * ex : ax - 0.5f) + (ay < 1.0f ? 0.5f * ey * ey : ay - 0.5f);
                g[0] = norm * fmaxf(-1.0f, fminf(1.0f, ex));
                g[1] = norm * fmaxf(-1.0f, fminf(1.0f, ey));
                close = ax < 1.0f && ay < 1.0f;
            }
            const float z = o[2], c01 = close ? 1.0f : 0.0f;
            const float sp = z > 0.0f ? z + log1pf(expf(-z)) : log1pf(expf(z));  // softplus
            loss += conf_weight * (sp - c01 * z);
            g[2] = norm * conf_weight * (1.0f / (1.0f + expf(-z)) - c01);
            loss *= norm;
        }
    }
    red[threadIdx.x] = loss;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_acc, red[0]);
}

// ---------------------------------------------------------------------------
// Batched, persistent variants: a fixed grid walks a device-resident work
// table (many renders of uneven size, one launch per stage). The grid and
// the arguments never change between batches, so the launch sequence can be
// replayed as a hipGraph; only the table contents differ.
// Work item k of T is owned by XCD x = (its block id) % 8, which walks a
// contiguous range of items, so neighbouring tiles share an L2.
// ---------------------------------------------------------------------------
struct XcdRange {
    uint32_t begin, end, stride, first;
};
__device__ __forceinline__ XcdRange xcd_range(uint32_t total) {
    constexpr uint32_t NX = 8;
    const uint32_t x = blockIdx.x % NX, j = blockIdx.x / NX, per = gridDim.x / NX;
    const uint32_t q = total / NX, r = total % NX;
    const uint32_t b = x * q + (x < r ? x : r);
    return XcdRange{b, b + q + (x < r ? 1u : 0u), per, b + j};
}

// largest i in [0, n) with begin[i] <= k (begin[0] == 0, ascending)
__device__ __forceinline__ uint32_t find_owner(const uint32_t *begin, uint32_t n, uint32_t k) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (begin[mid] <= k) lo = mid; else hi = mid;
    }
    return lo;
}

// The work table is read-only while a batch runs: reading it through the
// constant address space lets the compiler keep descriptors in scalar
// registers, as it does for kernel arguments.
template <typename T>
__device__ __forceinline__ KernArg<T> *as_const(const T *p) {
    return (KernArg<T> *)p;
}
#if defined(__HIP_DEVICE_COMPILE__)
#define ISTI_CONST_TABLE(t, tg) const KernArg<BatchTable> *t = as_const(tg)
#else
#define ISTI_CONST_TABLE(t, tg) const BatchTable *t = tg  // host pass only parses the body
#endif

// A block's work items increase by a fixed stride, so their owners (jobs or
// requests, in prefix order) only move forward: the owner is found once by
// binary search and then advanced, and its descriptor is re-read only when
// it changes (each lookup is a dependent scalar load).
#if defined(__HIP_DEVICE_COMPILE__)
#define ISTI_CONST_REF(T, name, ptr) const KernArg<T> &name = *as_const(ptr)
#else
#define ISTI_CONST_REF(T, name, ptr) const T &name = *(ptr)
#endif

__global__ __launch_bounds__(256) void clear_batch_kernel(const BatchTable *__restrict__ tg) {
    ISTI_CONST_TABLE(t, tg);
    const XcdRange r = xcd_range(t->clear_blocks);
    if (r.first >= r.end) return;
    const auto *begin = as_const(t->clear_begin);
    uint32_t j = find_owner(t->clear_begin, t->n_jobs, r.first);
    uint32_t jb = begin[j], je = begin[j + 1];
    for (uint32_t k = r.first; k < r.end; k += r.stride) {
        while (k >= je) {
            ++j;
            jb = je;
            je = begin[j + 1];
        }
        ISTI_CONST_REF(SplatJob, J, &t->jobs[j]);
        clear_pairs(J.splat, J.n_px, (k - jb) * 256u + threadIdx.x);
    }
}

__global__ __launch_bounds__(256) void splat_batch_kernel(const BatchTable *__restrict__ tg) {
    __shared__ SplatBin sb;
    ISTI_CONST_TABLE(t, tg);
    const XcdRange r = xcd_range(t->splat_blocks);
    if (r.first >= r.end) return;
    const auto *begin = as_const(t->splat_begin);
    uint32_t j = find_owner(t->splat_begin, t->n_jobs, r.first);
    uint32_t jb = begin[j], je = begin[j + 1];
    for (uint32_t k = r.first; k < r.end; k += r.stride) {
        while (k >= je) {
            ++j;
            jb = je;
            je = begin[j + 1];
        }
        ISTI_CONST_REF(SplatJob, J, &t->jobs[j]);
        ISTI_CONST_REF(BatchRequest, R, &t->req[J.req]);
        if (J.binned) {  // uniform per job: this block's item is a 32x8 tile
            splat_tile_binned(J, R.vp_cam, R.vp_main[J.main_slot < 0 ? 0 : J.main_slot], R.init_linear != 0, k - jb,
                              sb);
            __syncthreads();  // the next item re-initialises the bin
        } else {
            splat_pixel(J, R.vp_cam, R.vp_main[J.main_slot < 0 ? 0 : J.main_slot], R.init_linear != 0,
                        (k - jb) * SPLAT_BLOCK + threadIdx.x);
        }
    }
}

template <int W, bool MAV, int NAUX>
__global__ __launch_bounds__(BATCH_TILE_W * BATCH_TILE_H) __attribute__((amdgpu_waves_per_eu(resolve_min_waves_wide(W, NAUX))))
void resolve_batch_kernel(const BatchTable *__restrict__ tg, int variant) {
    __shared__ ResolveLds<W, NAUX, BATCH_TILE_W, BATCH_TILE_H> lds;
    ISTI_CONST_TABLE(t, tg);
    const uint32_t r0 = t->var_req_begin[variant], r1 = t->var_req_begin[variant + 1];
    const uint32_t base = t->tile_begin[r0];
    const XcdRange r = xcd_range(t->tile_begin[r1] - base);
    if (r.first >= r.end) return;
    const auto *tile_begin = as_const(t->tile_begin);
    uint32_t q = r0 + find_owner(t->tile_begin + r0, r1 - r0, base + r.first);
    uint32_t qe = tile_begin[q + 1];
    for (uint32_t k = r.first; k < r.end; k += r.stride) {
        while (base + k >= qe) qe = tile_begin[++q + 1];
#if defined(__HIP_DEVICE_COMPILE__)
        const KernArg<ResolveArgs> &a = *as_const(&t->req[q].args);
#else
        const ResolveArgs &a = t->req[q].args;
#endif
        // row-major tiles: an XCD's contiguous range of items is a band of
        // rows (the stripe-major order of the per-render resolve measured
        // 270 -> 280 us per batched stereo frame here)
        resolve_tile<W, MAV, NAUX, false, BATCH_TILE_W, BATCH_TILE_H>(a, base + k - tile_begin[q], lds);
        __syncthreads();  // the next tile restages LDS
    }
}

// ---------------------------------------------------------------------------
// Deterministic mean z. Row sums: one wave per row, lane l accumulates
// columns l, l+64, ... in double, lanes folded by halving. Mean: one block,
// lane l accumulates rows l, l+256, ..., lanes folded by halving.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void row_sum_kernel(const float *z, uint32_t w, uint32_t h, double *rows) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t y = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (y >= h) return;  // whole wave exits together
    double acc = 0.0;
    const float *row = z + (size_t)y * w;
    for (uint32_t x = lane; x < w; x += 64u) acc += (double)row[x];
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) rows[y] = acc;
}

__global__ __launch_bounds__(256) void mean_kernel(const double *rows, uint32_t w, uint32_t h, float *mean) {
    __shared__ double s[256];
    const uint32_t l = threadIdx.x;
    double acc = 0.0;
    for (uint32_t y = l; y < h; y += 256u) acc += rows[y];
    s[l] = acc;
    __syncthreads();
    for (uint32_t off = 128; off >= 1; off >>= 1) {
        if (l < off) s[l] += s[l + off];
        __syncthreads();
    }
    if (l == 0) *mean = (float)(s[0] / ((double)w * (double)h));
}

// ---------------------------------------------------------------------------
// Timewarp (mode 4), renderer.cu:467-522. glm's dot(vec3) = (x+y)+z,
// dot(vec4) = (x+y)+(z+w), normalize(v) = v * (1/sqrt(dot(v,v))).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    const float tx = ax * bx, ty = ay * by, tz = az * bz;
    return (tx + ty) + tz;
}

__global__ __launch_bounds__(256) void timewarp_kernel(const TimewarpArgs a) {
    const uint32_t tile = xcd_remap(blockIdx.x, a.n_tiles);
    const uint32_t ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const uint32_t x = tx * TILE_W + (threadIdx.x & (TILE_W - 1));
    const uint32_t y = ty * TILE_H + (threadIdx.x / TILE_W);
    if (x >= a.w || y >= a.h) return;

    const float2 pcs = pixel2cs(x, y, a.w, a.h);
    const float4 near_ws = project(a.inv_vp_from, make_float4(pcs.x, pcs.y, 0.0f, 1.0f));
    const float4 far_ws = project(a.inv_vp_from, make_float4(pcs.x, pcs.y, 1.0f, 1.0f));
    const float rdx0 = far_ws.x - near_ws.x, rdy0 = far_ws.y - near_ws.y, rdz0 = far_ws.z - near_ws.z;
    const float rs = 1.0f / sqrtf(dot3(rdx0, rdy0, rdz0, rdx0, rdy0, rdz0));
    const float rdx = rdx0 * rs, rdy = rdy0 * rs, rdz = rdz0 * rs;

    const float4 center = project(a.inv_vp_to, make_float4(0.0f, 0.0f, *a.depth_to, 1.0f));
    const float cpx = a.inv_vp_to.m[12], cpy = a.inv_vp_to.m[13], cpz = a.inv_vp_to.m[14], cpw = a.inv_vp_to.m[15];
    const float vx0 = center.x - cpx, vy0 = center.y - cpy, vz0 = center.z - cpz, vw0 = center.w - cpw;
    const float d4 = (vx0 * vx0 + vy0 * vy0) + (vz0 * vz0 + vw0 * vw0);
    const float vs = 1.0f / sqrtf(d4);
    const float nx = vx0 * vs, ny = vy0 * vs, nz = vz0 * vs;

    const float denom = dot3(rdx, rdy, rdz, nx, ny, nz);
    if (fabsf(denom) < 1e-6f) return;
    const float t = dot3(cpx - near_ws.x, cpy - near_ws.y, cpz - near_ws.z, nx, ny, nz) / denom;
    const float4 inter = make_float4(near_ws.x + t * rdx, near_ws.y + t * rdy, near_ws.z + t * rdz, 1.0f);

    const float4 pc = project(a.vp_to, inter);
    const int2 pix = cs2pixel(pc.x, pc.y, a.w_to, a.h_to);
    uchar4 c;
    if (in_range(pix, a.w_to, a.h_to) || !a.aux_set) {
        const int cx = max(0, min(pix.x, (int)a.w_to - 1)), cy = max(0, min(pix.y, (int)a.h_to - 1));
        c = fetch(a.img_to, a.w_to, a.h_to, (uint32_t)cx, (uint32_t)cy);
    } else {
        const float4 pa = project(a.vp_aux, inter);
        const int2 pixa = cs2pixel(pa.x, pa.y, a.w_aux, a.h_aux);
        const int cx = max(0, min(pixa.x, (int)a.w_aux - 1)), cy = max(0, min(pixa.y, (int)a.h_aux - 1));
        c = fetch(a.img_aux, a.w_aux, a.h_aux, (uint32_t)cx, (uint32_t)cy);
    }
    a.out[(size_t)y * a.out_pitch_px + x] = c;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_clear(const ClearArgs &a, hipStream_t s) {
    if (a.n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(clear_splat_kernel, dim3(a.n_blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_splat(const SplatArgs &a, bool binned, hipStream_t s) {
    if (a.n_blocks == 0) return hipSuccess;
    if (binned)
        hipLaunchKernelGGL(reproject_splat_binned_kernel, dim3(a.n_blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(reproject_splat_kernel, dim3(a.n_blocks), dim3(SPLAT_BLOCK), 0, s, a);
    return hipGetLastError();
}

uint32_t splat_blocks(uint32_t w, uint32_t h, bool binned) {
    return binned ? ((w + BIN_TW - 1) / BIN_TW) * ((h + BIN_TH - 1) / BIN_TH) : (w * h + SPLAT_BLOCK - 1) / SPLAT_BLOCK;
}

hipError_t launch_features(const FeatureArgs &a, hipStream_t s) {
    const uint32_t n = 2u * a.net_w * a.net_h;
    hipLaunchKernelGGL(features_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_learned_loss(const float *out, const uint32_t *q, const float4 *ff_fwd, const float4 *ff_bwd,
                               uint32_t w, uint32_t h, uint32_t net_w, uint32_t net_h, uint32_t shift,
                               float conf_weight, float *dout, float *loss_acc, hipStream_t s) {
    const uint32_t n = 2u * net_w * net_h;
    hipLaunchKernelGGL(learned_loss_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, out, q, ff_fwd, ff_bwd, w, h,
                       net_w, net_h, shift, conf_weight, dout, loss_acc);
    return hipGetLastError();
}

hipError_t launch_resolve_learned(const ResolveArgs &a, hipStream_t s) {
    if (a.n_tiles == 0) return hipSuccess;
    const dim3 grid(a.n_tiles), block(TILE_W * TILE_H);
    switch (a.n_aux) {
        case 0: hipLaunchKernelGGL(resolve_learned_kernel<0>, grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL(resolve_learned_kernel<1>, grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL(resolve_learned_kernel<2>, grid, block, 0, s, a); break;
        case 3: hipLaunchKernelGGL(resolve_learned_kernel<3>, grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL(resolve_learned_kernel<4>, grid, block, 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_clear_batch(const BatchTable *t, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(clear_batch_kernel, dim3(grid), dim3(256), 0, s, t);
    return hipGetLastError();
}

hipError_t launch_splat_batch(const BatchTable *t, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(splat_batch_kernel, dim3(grid), dim3(SPLAT_BLOCK), 0, s, t);
    return hipGetLastError();
}

// Kernel of each resolve variant: (window 1, plain) or (window 3, main always valid) x n_aux.
template <int V>
static void *resolve_batch_fn() {
    constexpr bool mav = V > ISTI_MAX_AUX;
    constexpr int n_aux = mav ? V - 1 - ISTI_MAX_AUX : V;
    return reinterpret_cast<void *>(&resolve_batch_kernel<mav ? 3 : 1, mav, n_aux>);
}
static void *resolve_batch_ptr(int v) {
    static void *const table[N_VARIANTS] = {resolve_batch_fn<0>(), resolve_batch_fn<1>(), resolve_batch_fn<2>(),
                                            resolve_batch_fn<3>(), resolve_batch_fn<4>(), resolve_batch_fn<5>(),
                                            resolve_batch_fn<6>(), resolve_batch_fn<7>(), resolve_batch_fn<8>(),
                                            resolve_batch_fn<9>()};
    return (v >= 0 && v < N_VARIANTS) ? table[v] : nullptr;
}

hipError_t launch_resolve_batch(const BatchTable *t, int variant, uint32_t grid, hipStream_t s) {
    void *fn = resolve_batch_ptr(variant);
    if (!fn) return hipErrorInvalidValue;
    void *args[] = {&t, &variant};
    return hipLaunchKernel(fn, dim3(grid), dim3(BATCH_TILE_W * BATCH_TILE_H), args, 0, s);
}

int resolve_batch_blocks_per_cu(int variant) {
    void *fn = resolve_batch_ptr(variant);
    int n = 0;
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, BATCH_TILE_W * BATCH_TILE_H, 0) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

hipError_t launch_mean_z(const float *z, uint32_t w, uint32_t h, double *row_sums, float *mean, hipStream_t s) {
    hipLaunchKernelGGL(row_sum_kernel, dim3((h + 3u) / 4u), dim3(256), 0, s, z, w, h, row_sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mean_kernel, dim3(1), dim3(256), 0, s, (const double *)row_sums, w, h, mean);
    return hipGetLastError();
}

// One block: thread t takes slot row t, the rows are summed through LDS into
// the caller's counters and zeroed.
__global__ __launch_bounds__(STAT_SLOTS) void stats_reduce_kernel(uint32_t *slots, uint32_t *out) {
    __shared__ uint32_t sum[STAT_COUNTERS];
    if (threadIdx.x < STAT_COUNTERS) sum[threadIdx.x] = 0u;
    __syncthreads();
    uint32_t *row = slots + threadIdx.x * STAT_STRIDE;
    for (uint32_t k = 0; k < STAT_COUNTERS; k++) {
        const uint32_t v = row[k];
        row[k] = 0u;
        if (v != 0u) atomicAdd(&sum[k], v);
    }
    __syncthreads();
    if (threadIdx.x < STAT_COUNTERS) atomicAdd(&out[threadIdx.x], sum[threadIdx.x]);
}

hipError_t launch_stats_reduce(uint32_t *slots, uint32_t *out, hipStream_t s) {
    hipLaunchKernelGGL(stats_reduce_kernel, dim3(1), dim3(STAT_SLOTS), 0, s, slots, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Self-test of div3_rn against the IEEE division. Operand families (i mod 4):
// 0 random bit patterns (every class), 1 moderate exponents [-40, 40],
// 2 quotients near rounding midpoints (n = RN32((q + ulp(q)/2) w)),
// 3 every exponent field including subnormals. Mismatching bit patterns are
// counted per family (counts[0..3]), and so are the sets whose quotients came
// from the reciprocal path (counts[4..7]).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t v) {  // splitmix64 finaliser
    v ^= v >> 30;
    v *= 0xbf58476d1ce4e5b9ull;
    v ^= v >> 27;
    v *= 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}
__device__ __forceinline__ float moderate(uint32_t r, int span) {
    const uint32_t e = 127u + (uint32_t)((int)((r >> 23) % (uint32_t)(2 * span + 1)) - span);
    return __uint_as_float((r & 0x80000000u) | (e << 23) | (r & 0x7fffffu));
}
__device__ __forceinline__ float near_midpoint(uint32_t r, float w) {
    const float q = moderate(r, 20);
    const double m = (double)q + ldexp((double)(q < 0.0f ? -1.0 : 1.0), ilogbf(q) - 24);
    return (float)(m * (double)w);
}
__global__ __launch_bounds__(256) void div_selftest_kernel(uint64_t n, uint64_t seed, uint32_t *counts) {
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += stride) {
        const uint64_t h0 = mix64(seed + 0x9e3779b97f4a7c15ull * (i + 1)), h1 = mix64(h0);
        const uint32_t r0 = (uint32_t)h0, r1 = (uint32_t)(h0 >> 32), r2 = (uint32_t)h1, r3 = (uint32_t)(h1 >> 32);
        const int fam = (int)(i & 3u);
        float x, y, z, w;
        if (fam == 0) {
            x = __uint_as_float(r0), y = __uint_as_float(r1), z = __uint_as_float(r2), w = __uint_as_float(r3);
        } else if (fam == 1) {
            x = moderate(r0, 40), y = moderate(r1, 40), z = moderate(r2, 40), w = moderate(r3, 40);
        } else if (fam == 2) {
            w = moderate(r3, 20);
            x = near_midpoint(r0, w), y = near_midpoint(r1, w), z = near_midpoint(r2, w);
        } else {
            x = __uint_as_float(r0 & 0x807fffffu | (((r0 >> 8) % 255u) << 23));
            y = __uint_as_float(r1 & 0x807fffffu | (((r1 >> 8) % 255u) << 23));
            z = __uint_as_float(r2 & 0x807fffffu | (((r2 >> 8) % 255u) << 23));
            w = __uint_as_float(r3 & 0x807fffffu | (((r3 >> 8) % 255u) << 23));
        }
        const float3 f = div3_rn(x, y, z, w);
        const bool bad = __float_as_uint(f.x) != __float_as_uint(x / w) ||
                         __float_as_uint(f.y) != __float_as_uint(y / w) ||
                         __float_as_uint(f.z) != __float_as_uint(z / w);
        if (bad) atomicAdd(&counts[fam], 1u);
        // sets the reciprocal path answered (the rest fell back to the division)
        const bool fast = __builtin_amdgcn_classf(w, CLASS_NORMAL) &&
                          __builtin_amdgcn_classf(f.x, CLASS_NORMAL_OR_INF) &&
                          __builtin_amdgcn_classf(f.y, CLASS_NORMAL_OR_INF) &&
                          __builtin_amdgcn_classf(f.z, CLASS_NORMAL_OR_INF);
        const bool leader = (int)(threadIdx.x & 63u) == __ffsll((long long)__ballot(1)) - 1;
        for (int k = 0; k < 4; k++) {  // lanes of one wave mix the families
            const unsigned long long lanes = __ballot(fast && fam == k);
            if (leader && lanes != 0) atomicAdd(&counts[4 + k], (uint32_t)__popcll(lanes));
        }
    }
}

hipError_t launch_div_selftest(uint64_t n, uint64_t seed, uint32_t *counts, hipStream_t s) {
    const uint64_t blocks = (n + 255u) / 256u;
    const uint32_t grid = (uint32_t)(blocks < 8192u ? (blocks > 0 ? blocks : 1u) : 8192u);
    hipLaunchKernelGGL(div_selftest_kernel, dim3(grid), dim3(256), 0, s, n, seed, counts);
    return hipGetLastError();
}

hipError_t launch_timewarp(const TimewarpArgs &a, hipStream_t s) {
    if (a.n_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(timewarp_kernel, dim3(a.n_tiles), dim3(TILE_W * TILE_H), 0, s, a);
    return hipGetLastError();
}

}  // namespace isti
