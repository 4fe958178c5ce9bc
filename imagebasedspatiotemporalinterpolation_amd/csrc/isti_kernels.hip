// isti_kernels.hip - gfx950 kernels of the per-frame interpolation path.
//
//   clear_splat_kernel      splat buffers <- 0xfe.. (renderer.cu:631-638)
//   reproject_splat_kernel  K1, every source view of the frame in ONE launch
//                           (renderer.cu:76-146, launched 2+2*n_aux times
//                           by the reference, :643-649)
//   resolve_kernel<W, AV>   K2: window init, fixed-point search, decision,
//                           second chance, soft shadows, aux fallback, hole
//                           retention (renderer.cu:148-465)
//   row_sum_kernel,
//   mean_kernel             deterministic mean z (replaces the float
//                           atomicAdd of image_buffer.cu:27-36,80-101)
//   timewarp_kernel         mode 4 baseline (renderer.cu:467-522)
//
// Launch geometry is wave64-first: 256-thread blocks (4 waves). Block ids are
// remapped so that consecutive tiles land on the same XCD (the hardware
// deals workgroups round-robin over the 8 XCDs, each with its own L2), which
// keeps the neighbourhood gathers of adjacent tiles in one L2.

#include "isti_kernels.hpp"

namespace isti {

// The explicit kernel arguments start at offset 0 of the kernarg segment.
template <typename T>
using KernArg = __attribute__((address_space(4))) const T;
template <typename T>
__device__ __forceinline__ KernArg<T> *kernarg_ptr() {
    return (KernArg<T> *)__builtin_amdgcn_kernarg_segment_ptr();
}

// Block id -> logical tile id such that each XCD owns one contiguous range.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    constexpr uint32_t NX = 8;
    const uint32_t q = nb / NX, r = nb % NX;
    const uint32_t x = b % NX, k = b / NX;
    return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------------------
// Splat clear: 16-byte stores over every buffer of the frame.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void clear_splat_kernel(const ClearArgs a) {
    const uint32_t b = blockIdx.x;
    int j = 0;
    for (int i = 1; i < a.n_bufs; i++)
        if (b >= a.block_begin[i]) j = i;
    unsigned long long *buf = a.buf[j];
    const uint32_t n = a.n[j];
    const uint32_t pair = (b - a.block_begin[j]) * 256u + threadIdx.x;  // 2 cells per thread
    const uint32_t i0 = pair * 2u;
    if (i0 + 1u < n) {
        reinterpret_cast<ulonglong2 *>(buf)[pair] = make_ulonglong2(SPLAT_EMPTY, SPLAT_EMPTY);
    } else if (i0 < n) {
        buf[i0] = SPLAT_EMPTY;
    }
}

// ---------------------------------------------------------------------------
// K1: reprojection + depth-sorted splat (renderer.cu:76-146).
// One thread per source pixel, row-major, so z / motion loads and the two
// float4 flow stores are fully coalesced; the u64 atomicMin is the only
// scattered access.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reproject_splat_kernel(const SplatArgs a) {
    const uint32_t b = xcd_remap(blockIdx.x, a.n_blocks);
    int j = 0;
    for (int i = 1; i < a.n_jobs; i++)
        if (b >= a.job[i].block_begin) j = i;
    const SplatJob &J = a.job[j];
    const uint32_t p = (b - J.block_begin) * SPLAT_BLOCK + threadIdx.x;
    if (p >= J.n_px) return;
    const uint32_t y = p / J.w;
    const uint32_t x = p - y * J.w;

    const float z_from = J.z[p];
    float4 om = make_float4(0.f, 0.f, 0.f, 0.f);
    if (J.mx != nullptr) om = make_float4(J.mx[p], J.my[p], J.mz[p], 0.f);

    const float2 pcs = pixel2cs(x, y, J.w, J.h);
    const float4 cs_from = make_float4(pcs.x, pcs.y, z_from, 1.0f);
    const float4 ws_from = project(J.inv_vp_from, cs_from);
    const float4 cs_moved = make_float4(cs_from.x + om.x, cs_from.y + om.y, cs_from.z + om.z, cs_from.w + 0.0f);
    const float4 ws_to = project(J.inv_vp_from, cs_moved);

    const float4 cs_main = project(a.vp_main[J.main_slot], ws_from);
    const float lim = 1.0f - 1e-3f;
    const bool visible = cs_main.x > -lim && cs_main.x < lim && cs_main.y > -lim && cs_main.y < lim &&
                         cs_main.z > -lim && cs_main.z < lim;
    const float vis = visible ? 1.0f : 0.0f;

    const float wt = J.weight;
    const float4 ws_scaled = make_float4(ws_from.x + wt * (ws_to.x - ws_from.x), ws_from.y + wt * (ws_to.y - ws_from.y),
                                         ws_from.z + wt * (ws_to.z - ws_from.z), ws_from.w + wt * (ws_to.w - ws_from.w));
    const float4 cs_to = project(J.vp_to, ws_to);
    const float fx = cs_to.x - cs_from.x, fy = cs_to.y - cs_from.y, fz = cs_to.z - cs_from.z;
    J.full_ff[p] = make_float4(fx, fy, fz, vis);

    float depth = z_to_depth(cs_from.z);
    uint32_t tx = x, ty = y;
    if (!a.init_linear) {
        const float4 cs_cam = project(a.vp_cam, ws_scaled);
        J.cam_ff[p] = make_float4(cs_cam.x - cs_from.x, cs_cam.y - cs_from.y, cs_cam.z - cs_from.z, vis);
        const float depth_cam = z_to_depth(cs_cam.z);
        const int2 pc = cs2pixel(cs_cam.x, cs_cam.y, J.w, J.h);
        if (!in_range(pc, J.w, J.h) || depth_cam < 0.0f) return;
        depth = depth_cam;
        tx = (uint32_t)pc.x;
        ty = (uint32_t)pc.y;
    } else {
        J.cam_ff[p] = make_float4(fx * wt, fy * wt, fz * wt, vis);
    }
    atomicMin(&J.splat[(size_t)ty * J.w + tx], encode_key(depth, x, y));
}

// ---------------------------------------------------------------------------
// K2: resolve (renderer.cu:148-465)
// ---------------------------------------------------------------------------
struct SplatInfo {
    float depth;
    uint32_t x, y;
};

// initializeSearchFromSplatting<W>, renderer.cu:148-175
template <int W>
__device__ __forceinline__ SplatInfo init_search(uint32_t ix, uint32_t iy, const ResolveView &v) {
    const unsigned long long c = v.splat[(size_t)iy * v.w + ix];
    SplatInfo best{key_depth(c), key_x(c), key_y(c)};
#pragma unroll
    for (int dy = -W; dy <= W; dy++) {
#pragma unroll
        for (int dx = -W; dx <= W; dx++) {
            const int2 n = make_int2((int)ix + dx, (int)iy + dy);
            if (!in_range(n, v.w, v.h)) continue;
            const unsigned long long k = v.splat[(size_t)n.y * v.w + (uint32_t)n.x];
            const float d = key_depth(k);
            if (d > 0.0f && (!__builtin_isfinite(best.depth) || best.depth < 0.0f || d < best.depth))
                best = SplatInfo{d, key_x(k), key_y(k)};
        }
    }
    return best;
}

// iterativeSearch<N>, renderer.cu:177-202. The reference always runs N
// steps; an unchanged or out-of-range candidate is a fixed point (every later
// step repeats it), so stopping there is bit-identical. Returns the cam_ff
// sample at the final pixel in `m`.
__device__ __forceinline__ uint2 search(uint2 start, float2 pcs_full, const ResolveView &v, int n_iter, float4 &m) {
    uint2 r = start;
    m = fetch(v.cam_ff, v.w, v.h, r.x, r.y);
    for (int i = 0; i < n_iter; i++) {
        const int2 t = cs2pixel(pcs_full.x - m.x, pcs_full.y - m.y, v.w, v.h);
        if (!in_range(t, v.w, v.h)) break;
        const uint2 nt = make_uint2((uint32_t)t.x, (uint32_t)t.y);
        if (nt.x == r.x && nt.y == r.y) break;
        r = nt;
        m = fetch(v.cam_ff, v.w, v.h, r.x, r.y);
    }
    return r;
}

// computeReprojError, renderer.cu:204-215
__device__ __forceinline__ float reproj_error(uint2 q, const ResolveView &v, float2 pcs_cam, float4 m) {
    const float2 pf = pixel2cs(q.x, q.y, v.w, v.h);
    float ex = pf.x + m.x - pcs_cam.x;
    float ey = pf.y + m.y - pcs_cam.y;
    const float mn = (float)(v.w < v.h ? v.w : v.h);
    const float sx = (float)v.w / mn, sy = (float)v.h / mn;
    ex *= sx;
    ey *= sy;
    return sqrtf(ex * ex + ey * ey);
}

// colorFromShadow(idx, ...), renderer.cu:217-268
__device__ uchar4 color_from_shadow_1(uint32_t vidx, int sfpa, const ResolveView &v, uint32_t px, uint32_t py) {
    const uchar4 main_color = fetch(v.img, v.w, v.h, px, py);
    if (vidx < (uint32_t)sfpa && vidx > 0) {
        const uchar4 si = fetch(v.sh_info, v.w, v.h, px, py);
        const uint32_t bits = ((uint32_t)si.x << 16) | ((uint32_t)si.y << 8) | (uint32_t)si.z;
        if (v.sh_b != nullptr && bits != 0) {
            const int init_val = (int)(bits & SOFTSHADOW_VAL_MASK);
            const int my_val = (int)(bits >> (SOFTSHADOW_BITS * vidx) & SOFTSHADOW_VAL_MASK);
            int min_val = (int)SOFTSHADOW_VAL_MASK, max_val = 0;
            for (int i = 0; i < sfpa; i++) {
                const int c = (int)(bits >> (SOFTSHADOW_BITS * (uint32_t)i) & SOFTSHADOW_VAL_MASK);
                min_val = c < min_val ? c : min_val;
                max_val = c > max_val ? c : max_val;
            }
            const float t = (float)(my_val - min_val) / (float)(max_val - min_val);
            const uchar4 cmin = min_val == init_val ? main_color : fetch(v.sh_a, v.w, v.h, px, py);
            const uchar4 cmax = max_val == init_val ? main_color : fetch(v.sh_b, v.w, v.h, px, py);
            const float omt = (float)(1.0 - (double)t);
            const float rx = (float)cmin.x * omt + (float)cmax.x * t;
            const float ry = (float)cmin.y * omt + (float)cmax.y * t;
            const float rz = (float)cmin.z * omt + (float)cmax.z * t;
            return make_uchar4(cvt_u8(rx), cvt_u8(ry), cvt_u8(rz), 255);
        }
    }
    return main_color;
}

// colorFromShadow(idcs, t, ...), renderer.cu:270-295 (frameIdx2Idx :227-230)
__device__ uchar4 color_from_shadow(uint2 idcs, float st, int sfpa, bool no_shadows, const ResolveView &v,
                                    uint32_t px, uint32_t py) {
    if (no_shadows || idcs.x >= (uint32_t)sfpa || idcs.y == 0 || v.sh_a == nullptr)
        return fetch(v.img, v.w, v.h, px, py);
    const uint32_t vx = (uint32_t)(v.forward ? (int)idcs.x : sfpa - (int)idcs.x);
    const uint32_t vy = (uint32_t)(v.forward ? (int)idcs.y : sfpa - (int)idcs.y);
    if (vx == vy || vy >= (uint32_t)sfpa) return color_from_shadow_1(vx, sfpa, v, px, py);
    if (vx >= (uint32_t)sfpa) return color_from_shadow_1(vy, sfpa, v, px, py);
    const uchar4 ca = color_from_shadow_1(vx, sfpa, v, px, py);
    const uchar4 cb = color_from_shadow_1(vy, sfpa, v, px, py);
    return lerp_to_u8(ca, cb, st);
}

struct FrameConst {
    float alpha;
    int fpa, sfpa;
    bool no_shadows;
    uint32_t frame_idx;
    uint2 sidx;
    float st;
};

// secondChanceAndBlend_kernel<EPS_2>, renderer.cu:297-327
__device__ uchar4 second_chance(float weight, const FrameConst &fc, SplatInfo si, const ResolveView &from,
                                const ResolveView &to, float eps2) {
    const float4 m_ft = fetch(from.full_ff, from.w, from.h, si.x, si.y);
    const float2 spcs = pixel2cs(si.x, si.y, from.w, from.h);
    const float sz = fetch(from.z, from.w, from.h, si.x, si.y);
    const float proj_x = spcs.x + m_ft.x, proj_y = spcs.y + m_ft.y, proj_z = sz + m_ft.z;
    const int2 pp = cs2pixel(proj_x, proj_y, to.w, to.h);

    uchar4 out = color_from_shadow(fc.sidx, fc.st, fc.sfpa, fc.no_shadows, from, si.x, si.y);
    if (fc.frame_idx >= (uint32_t)fc.fpa) return out;
    if (in_range(pp, to.w, to.h)) {
        const uint32_t ux = (uint32_t)pp.x, uy = (uint32_t)pp.y;
        const float zt = fetch(to.z, to.w, to.h, ux, uy);
        const float4 m_tf = fetch(to.full_ff, to.w, to.h, ux, uy);
        const float dx = m_ft.x - (-m_tf.x), dy = m_ft.y - (-m_tf.y), dz = m_ft.z - (-m_tf.z);
        const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
        if (dist < EPS_3 && fabsf(zt - proj_z) < eps2)
            out = lerp_to_u8(color_from_shadow(fc.sidx, fc.st, fc.sfpa, fc.no_shadows, to, ux, uy), out, weight);
    }
    return out;
}

struct Result {
    bool valid;
    bool visible_in_main;
    int priority;
    float error;
    float depth;
    uchar4 color;
};

// bidirectionalReprojection<SW, SI, ALWAYS_VALID, EPS_1, EPS_2>, renderer.cu:329-398
template <int W, bool ALWAYS_VALID>
__device__ Result bidir(uint32_t ix, uint32_t iy, float2 pcs_full, float2 pcs, const FrameConst &fc, int n_iter,
                        const ResolveView &fwd, const ResolveView &bwd, float eps1, float eps2) {
    float fwd_e = FLT_MAX;
    bool fwd_vis = false;
    SplatInfo fi = init_search<W>(ix, iy, fwd);
    if (fi.depth > 0.0f && fc.alpha < 1.0f) {
        float4 m;
        const uint2 q = search(make_uint2(fi.x, fi.y), pcs_full, fwd, n_iter, m);
        fi.x = q.x;
        fi.y = q.y;
        fi.depth = z_to_depth(fetch(fwd.z, fwd.w, fwd.h, q.x, q.y) + m.z);
        fwd_e = reproj_error(q, fwd, pcs, m);
        fwd_vis = m.w > 0.0f;
    }
    float bwd_e = FLT_MAX;
    bool bwd_vis = false;
    SplatInfo bi = init_search<W>(ix, iy, bwd);
    if (bi.depth > 0.0f) {
        float4 m;
        const uint2 q = search(make_uint2(bi.x, bi.y), pcs_full, bwd, n_iter, m);
        bi.x = q.x;
        bi.y = q.y;
        bi.depth = z_to_depth(fetch(bwd.z, bwd.w, bwd.h, q.x, q.y) + m.z);
        bwd_e = reproj_error(q, bwd, pcs, m);
        bwd_vis = m.w > 0.0f;
    }

    Result r{false, false, 0, 0.f, 0.f, make_uchar4(0, 0, 0, 0)};
    if (fwd_e < eps1 || bwd_e < eps1) {
        const bool both = fwd_e < eps1 && bwd_e < eps1;
        const bool similar = fabsf(depth_to_z(fi.depth) - depth_to_z(bi.depth)) < eps2;
        const bool choose_fwd = (both && !similar && fi.depth < bi.depth) || ((!both || similar) && fwd_e < bwd_e);
        r.valid = true;
        if (choose_fwd) {
            r.visible_in_main = fwd_vis || (both && similar && bwd_vis);
            r.priority = fwd.priority;
            r.error = fwd_e;
            r.depth = fi.depth;
            r.color = second_chance(1.0f - fc.alpha, fc, fi, fwd, bwd, eps2);
        } else {
            r.visible_in_main = bwd_vis || (both && similar && fwd_vis);
            r.priority = bwd.priority;
            r.error = bwd_e;
            r.depth = bi.depth;
            r.color = second_chance(fc.alpha, fc, bi, bwd, fwd, eps2);
        }
    } else if (ALWAYS_VALID) {
        r.valid = true;
        const uchar4 cf = second_chance(1.0f - fc.alpha, fc, fi, fwd, bwd, eps2);
        const uchar4 cb = second_chance(fc.alpha, fc, bi, bwd, fwd, eps2);
        r.color = lerp_to_u8(cf, cb, 0.5f);
    }
    return r;
}

// bidirectionalReprojection_kernel<SW, SI, MAV>, renderer.cu:400-465
template <int W, bool MAV>
__global__ __launch_bounds__(256) void resolve_kernel(const ResolveArgs args) {
    // Read the argument block in place (constant address space) instead of
    // through the by-value copy: the aux loop indexes it dynamically, which
    // would otherwise spill the whole block to per-lane scratch.
#if defined(__HIP_DEVICE_COMPILE__)
    const KernArg<ResolveArgs> &a = *kernarg_ptr<ResolveArgs>();
#else
    const ResolveArgs &a = args;  // host pass only parses the body
#endif
    const uint32_t tile = xcd_remap(blockIdx.x, a.n_tiles);
    const uint32_t ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const uint32_t x = tx * TILE_W + (threadIdx.x & (TILE_W - 1));
    const uint32_t y = ty * TILE_H + (threadIdx.x / TILE_W);
    if (x >= a.w || y >= a.h) return;

    FrameConst fc;
    fc.alpha = a.alpha;
    fc.fpa = a.fpa;
    fc.sfpa = a.sfpa;
    fc.no_shadows = a.no_shadows != 0;
    fc.frame_idx = f2u_rn((float)a.fpa * a.alpha);
    fc.sidx = make_uint2(f2u_rd((float)a.sfpa * a.alpha + 1e-9f), f2u_ru((float)a.sfpa * a.alpha - 1e-9f));
    fc.st = (float)a.sfpa * a.alpha - (float)fc.sidx.x;

    const float2 pcs = pixel2cs(x, y, a.w, a.h);
    uchar4 *out = a.out + (size_t)y * a.out_pitch_px + x;
    const Result main_r = bidir<W, MAV>(x, y, pcs, pcs, fc, a.max_iter, a.fwd[0], a.bwd[0], EPS_1_FULL_RES, EPS_2_FULL_RES);
    if (main_r.valid) *out = main_r.color;
    if (a.n_aux <= 0) return;

    Result best{false, false, 0, 0.f, 0.f, make_uchar4(0, 0, 0, 0)};
    for (int i = 1; i <= a.n_aux; i++) {
        const ResolveView &fa = a.fwd[i];
        const uint32_t ax = x / (a.w / fa.w), ay = y / (a.h / fa.h);
        const float2 apcs = pixel2cs(ax, ay, fa.w, fa.h);
        const Result ar = bidir<W, false>(ax, ay, pcs, apcs, fc, a.max_iter, fa, a.bwd[i], EPS_1_LOW_RES, EPS_2_LOW_RES);
        if (!best.valid) {
            best = ar;
        } else if (ar.valid) {
            const bool similar = fabsf(depth_to_z(best.depth) - depth_to_z(ar.depth)) < EPS_2_LOW_RES;
            const bool choose_new = (!similar && ar.depth < best.depth) ||
                                    (similar && ar.priority >= best.priority && ar.error < best.error);
            if (choose_new) best = ar;
        }
    }
    if (best.valid) {
        bool choose_aux = !main_r.valid;
        if (main_r.valid && !best.visible_in_main) {
            const bool similar = fabsf(depth_to_z(best.depth) - depth_to_z(main_r.depth)) < EPS_2_LOW_RES;
            choose_aux = !similar && best.depth < main_r.depth;
        }
        if (choose_aux) *out = best.color;
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

hipError_t launch_splat(const SplatArgs &a, hipStream_t s) {
    if (a.n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(reproject_splat_kernel, dim3(a.n_blocks), dim3(SPLAT_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_resolve(const ResolveArgs &a, int window, bool always_valid, hipStream_t s) {
    if (a.n_tiles == 0) return hipSuccess;
    const dim3 grid(a.n_tiles), block(TILE_W * TILE_H);
    if (window == 1 && !always_valid)
        hipLaunchKernelGGL((resolve_kernel<1, false>), grid, block, 0, s, a);
    else if (window == 3 && always_valid)
        hipLaunchKernelGGL((resolve_kernel<3, true>), grid, block, 0, s, a);
    else if (window == 1 && always_valid)
        hipLaunchKernelGGL((resolve_kernel<1, true>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((resolve_kernel<3, false>), grid, block, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mean_z(const float *z, uint32_t w, uint32_t h, double *row_sums, float *mean, hipStream_t s) {
    hipLaunchKernelGGL(row_sum_kernel, dim3((h + 3u) / 4u), dim3(256), 0, s, z, w, h, row_sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mean_kernel, dim3(1), dim3(256), 0, s, (const double *)row_sums, w, h, mean);
    return hipGetLastError();
}

hipError_t launch_timewarp(const TimewarpArgs &a, hipStream_t s) {
    if (a.n_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(timewarp_kernel, dim3(a.n_tiles), dim3(TILE_W * TILE_H), 0, s, a);
    return hipGetLastError();
}

}  // namespace isti
