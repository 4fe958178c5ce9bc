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
template <typename J_t, typename M_t>
__device__ __forceinline__ void splat_pixel(const J_t &J, const M_t &vp_cam, const M_t &vp_main, bool init_linear,
                                            uint32_t p) {
    if (p >= J.n_px) return;
    const uint32_t y = p / J.w;
    const uint32_t x = p - y * J.w;

    const float z_from = J.z[p];
    float4 om = make_float4(0.f, 0.f, 0.f, 0.f);
    if (J.mx != nullptr) om = make_float4(J.mx[p], J.my[p], J.mz[p], 0.f);

    const float4 cs_from = make_float4(cs_coord(x, J.inv_w), cs_coord(y, J.inv_h), z_from, 1.0f);
    const float4 ws_from = project(J.inv_vp_from, cs_from);
    const float4 cs_moved = make_float4(cs_from.x + om.x, cs_from.y + om.y, cs_from.z + om.z, cs_from.w + 0.0f);
    const float4 ws_to = project(J.inv_vp_from, cs_moved);

    const float4 cs_main = project(vp_main, ws_from);
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
    if (!init_linear) {
        const float4 cs_cam = project(vp_cam, ws_scaled);
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
    atomicMin(&J.splat[ty * J.w + tx], encode_key(depth, x, y));
}

__global__ __launch_bounds__(256) void reproject_splat_kernel(const SplatArgs a) {
    const uint32_t b = xcd_remap(blockIdx.x, a.n_blocks);
    int j = 0;
    for (int i = 1; i < a.n_jobs; i++)
        if (b >= a.job[i].block_begin) j = i;
    const SplatJob &J = a.job[j];
    splat_pixel(J, a.vp_cam, a.vp_main[J.main_slot], a.init_linear != 0, (b - J.block_begin) * SPLAT_BLOCK + threadIdx.x);
}

// ---------------------------------------------------------------------------
// K2: resolve (renderer.cu:148-465)
// ---------------------------------------------------------------------------
struct SplatInfo {
    float depth;
    uint32_t x, y;
};

// initializeSearchFromSplatting<W>, renderer.cu:148-175, over splat cells
// staged in LDS (row length rw; (lx, ly) is the centre). Cells outside the
// view were staged as the empty key, whose negative depth never wins, which
// is the reference's `continue` for out-of-range cells.
template <int W>
__device__ __forceinline__ SplatInfo init_search_lds(const unsigned long long *t, int rw, int lx, int ly) {
    const unsigned long long c = t[ly * rw + lx];
    SplatInfo best{key_depth(c), key_x(c), key_y(c)};
#pragma unroll
    for (int dy = -W; dy <= W; dy++) {
#pragma unroll
        for (int dx = -W; dx <= W; dx++) {
            const unsigned long long k = t[(ly + dy) * rw + lx + dx];
            const float d = key_depth(k);
            if (d > 0.0f && (!__builtin_isfinite(best.depth) || best.depth < 0.0f || d < best.depth))
                best = SplatInfo{d, key_x(k), key_y(k)};
        }
    }
    return best;
}

// Stages the splat cells [ox, ox+rw) x [oy, oy+rh) of view v into LDS rows
// of constant stride SPAN (rw <= SPAN).
template <int SPAN>
__device__ __forceinline__ void stage_splat(unsigned long long *dst, const ResolveView &v, int ox, int oy, int rw,
                                            int rh) {
    const int n = SPAN * rh;
    for (int i = (int)threadIdx.x; i < n; i += (int)(TILE_W * TILE_H)) {
        const int yy = i / SPAN, xx = i - yy * SPAN;
        const int gx = ox + xx, gy = oy + yy;
        const bool in = gx >= 0 && gx < (int)v.w && gy >= 0 && gy < (int)v.h;
        if (xx < rw) dst[i] = in ? v.splat[(uint32_t)gy * v.w + (uint32_t)gx] : SPLAT_EMPTY;
    }
}

// iterativeSearch<N>, renderer.cu:177-202. The reference always runs N
// steps of q <- step(q). step() depends on q alone (the target pixel is
// fixed), so once the sequence revisits a pixel it is periodic from there
// on, and q_N can be read off the cycle without running the remaining steps:
//   - an out-of-range candidate keeps q (the reference's `continue`), and an
//     unchanged candidate is a fixed point: period 1, stop;
//   - q_{i+1} == q_{i+1-p} for p = 2..4: with c_0 = q_{i+1} .. c_{p-1} = q_i,
//     q_N = c_{(N-i-1) mod p}, taken from the last four positions.
// This is bit-identical to the N-step loop. (At low resolution ~40% of the
// aux searches settle into 2-cycles, which would otherwise run all N steps.)
// Returns the cam_ff sample at the final pixel in `m`.
__device__ __forceinline__ uint32_t pack_px(uint32_t x, uint32_t y) { return (y << 16) | x; }

// One fixed-point search in flight. All searches of a pixel (fwd/bwd
// keyframe of the main and every aux camera) advance in lock-step, so their
// dependent gathers overlap instead of running one chain after another.
struct Chain {
    uint32_t h0, h1, h2, h3;  // q_i, q_{i-1}, q_{i-2}, q_{i-3} (packed)
    float4 m;                 // cam_ff at h0
    bool active;
};
constexpr uint32_t PX_NONE = 0xffffffffu;  // x, y <= 65534: never a packed pixel

// A chain that does not run (no splat candidate, or the skipped fwd search)
// parks at pixel (0, 0), so the unconditional reloads in chain_step stay
// inside the view; its result is never used.
__device__ __forceinline__ void chain_start(Chain &c, bool run, uint32_t x, uint32_t y, const ResolveView &v) {
    c.h0 = run ? pack_px(x, y) : 0u;
    c.h1 = c.h2 = c.h3 = PX_NONE;
    c.active = run;
    c.m = run ? at(v.cam_ff, v.w, x, y) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Step i (0-based) of n_iter: q_{i+1} = step(q_i), with the exits described above.
__device__ __forceinline__ void chain_step(Chain &c, int i, int n_iter, float2 pcs_full, const ResolveView &v) {
    if (!c.active) return;
    const int2 t = cs2pixel(pcs_full.x - c.m.x, pcs_full.y - c.m.y, v.w, v.h);
    if (!in_range(t, v.w, v.h)) {
        c.active = false;
        return;
    }
    const uint32_t nt = pack_px((uint32_t)t.x, (uint32_t)t.y);
    if (nt == c.h0) {
        c.active = false;
        return;
    }
    const int p = nt == c.h1 ? 2 : (nt == c.h2 ? 3 : (nt == c.h3 ? 4 : 0));
    if (p != 0) {
        const int idx = p - 1 - (n_iter - (i + 1)) % p;  // c_r = h_{p-1-r}
        const uint32_t q = idx == 0 ? c.h0 : (idx == 1 ? c.h1 : (idx == 2 ? c.h2 : c.h3));
        c.h0 = q;
        c.m = at(v.cam_ff, v.w, q & 0xFFFFu, q >> 16);
        c.active = false;
        return;
    }
    c.h3 = c.h2;
    c.h2 = c.h1;
    c.h1 = c.h0;
    c.h0 = nt;
    c.m = at(v.cam_ff, v.w, (uint32_t)t.x, (uint32_t)t.y);
}

// computeReprojError, renderer.cu:204-215
__device__ __forceinline__ float reproj_error(uint2 q, const ResolveView &v, float2 pcs_cam, float4 m) {
    float ex = cs_coord(q.x, v.inv_w) + m.x - pcs_cam.x;
    float ey = cs_coord(q.y, v.inv_h) + m.y - pcs_cam.y;
    ex *= v.rs_x;
    ey *= v.rs_y;
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
    const float sz = fetch(from.z, from.w, from.h, si.x, si.y);
    const float proj_x = cs_coord(si.x, from.inv_w) + m_ft.x, proj_y = cs_coord(si.y, from.inv_h) + m_ft.y;
    const float proj_z = sz + m_ft.z;
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

// ReprojResult (renderer.cu:49-57) without its colour: the colour depends
// only on the chosen candidate, so it is computed once, for the candidate
// that is finally written (the reference computes it for the main camera and
// for every aux camera).
struct Cand {
    bool valid;
    bool visible_in_main;
    int priority;
    float error;
    float depth;
    int dir;  // 0: fwd keyframe, 1: bwd keyframe, 2: 0.5 blend of both (ALWAYS_VALID)
};

// Forward/backward decision of one camera, renderer.cu:366-396.
template <bool ALWAYS_VALID>
__device__ __forceinline__ Cand decide(float fe, float be, float fd, float bd, bool fv, bool bv, int fprio, int bprio,
                                       float eps1, float eps2) {
    Cand k{false, false, 0, 0.f, 0.f, 0};
    if (fe < eps1 || be < eps1) {
        const bool both = fe < eps1 && be < eps1;
        const bool similar = fabsf(depth_to_z(fd) - depth_to_z(bd)) < eps2;
        const bool choose_fwd = (both && !similar && fd < bd) || ((!both || similar) && fe < be);
        k.valid = true;
        k.visible_in_main = choose_fwd ? (fv || (both && similar && bv)) : (bv || (both && similar && fv));
        k.priority = choose_fwd ? fprio : bprio;
        k.error = choose_fwd ? fe : be;
        k.depth = choose_fwd ? fd : bd;
        k.dir = choose_fwd ? 0 : 1;
    } else if (ALWAYS_VALID) {
        k.valid = true;
        k.dir = 2;
    }
    return k;
}

// Colour of a candidate of camera (fwd, bwd): renderer.cu:378-396.
__device__ __forceinline__ uchar4 cand_colour(const Cand &k, const FrameConst &fc, SplatInfo fi, SplatInfo bi,
                                              const ResolveView &fwd, const ResolveView &bwd, float eps2) {
    if (k.dir == 0) return second_chance(1.0f - fc.alpha, fc, fi, fwd, bwd, eps2);
    if (k.dir == 1) return second_chance(fc.alpha, fc, bi, bwd, fwd, eps2);
    const uchar4 cf = second_chance(1.0f - fc.alpha, fc, fi, fwd, bwd, eps2);
    const uchar4 cb = second_chance(fc.alpha, fc, bi, bwd, fwd, eps2);
    return lerp_to_u8(cf, cb, 0.5f);
}

// bidirectionalReprojection_kernel<SW, SI, MAV>, renderer.cu:400-465, with
// bidirectionalReprojection (:329-398) evaluated for the main camera and
// NAUX aux cameras at once:
//   1. splat-window initialisation of every (camera, keyframe) search;
//   2. all searches in lock-step (see Chain);
//   3. per-camera forward/backward decision (:366-396);
//   4. aux selection and main/aux choice (:429-464);
//   5. second chance + blend + soft shadows for the written candidate only.
// Pixels without a valid result are not written (hole retention).
template <int W, int NAUX>
struct ResolveLds {
    static constexpr int NC = 1 + NAUX;
    static constexpr int SPAN = TILE_W + 2 * W;  // staged cells per axis (TILE_W == TILE_H)
    unsigned long long splat[NC][2][SPAN * SPAN];
};

// One 16x16 tile of one render. Every thread of the block must call it (the
// LDS staging ends in a barrier); pixels outside the image return early.
template <int W, bool MAV, int NAUX, bool LEARNED = false, typename A>
__device__ __forceinline__ void resolve_tile(const A &a, uint32_t tile, ResolveLds<W, NAUX> &lds) {
    constexpr int NC = 1 + NAUX;
    constexpr int SPAN = TILE_W + 2 * W;
    auto &s_splat = lds.splat;
    const uint32_t ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const uint32_t x0 = tx * TILE_W, y0 = ty * TILE_H;
    const uint32_t x = x0 + (threadIdx.x & (TILE_W - 1));
    const uint32_t y = y0 + (threadIdx.x / TILE_W);

    // Stage every splat cell the tile's windows read: camera c covers camera
    // pixels [x0/rx, xl/rx] (xl = last in-image column of the tile) plus a
    // W-cell border, at most SPAN x SPAN cells.
    int org_x[NC], org_y[NC], span_x[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const uint32_t rx = a.w / a.fwd[c].w, ry = a.h / a.fwd[c].h;
        const uint32_t xl = min(x0 + TILE_W - 1, a.w - 1), yl = min(y0 + TILE_H - 1, a.h - 1);
        org_x[c] = (int)(x0 / rx) - W;
        org_y[c] = (int)(y0 / ry) - W;
        span_x[c] = (int)(xl / rx - x0 / rx) + 1 + 2 * W;
        const int span_y = (int)(yl / ry - y0 / ry) + 1 + 2 * W;
        stage_splat<SPAN>(s_splat[c][0], a.fwd[c], org_x[c], org_y[c], span_x[c], span_y);
        stage_splat<SPAN>(s_splat[c][1], a.bwd[c], org_x[c], org_y[c], span_x[c], span_y);
    }
    __syncthreads();
    if (x >= a.w || y >= a.h) return;

    FrameConst fc;
    fc.alpha = a.alpha;
    fc.fpa = a.fpa;
    fc.sfpa = a.sfpa;
    fc.no_shadows = a.no_shadows != 0;
    fc.frame_idx = f2u_rn((float)a.fpa * a.alpha);
    fc.sidx = make_uint2(f2u_rd((float)a.sfpa * a.alpha + 1e-9f), f2u_ru((float)a.sfpa * a.alpha - 1e-9f));
    fc.st = (float)a.sfpa * a.alpha - (float)fc.sidx.x;

    const float2 pcs = make_float2(cs_coord(x, a.fwd[0].inv_w), cs_coord(y, a.fwd[0].inv_h));
    // per camera: pixel in the camera's own resolution (renderer.cu:432-434)
    uint32_t cx[NC], cy[NC];
    float2 cpcs[NC];
    cx[0] = x;
    cy[0] = y;
    cpcs[0] = pcs;
#pragma unroll
    for (int c = 1; c < NC; c++) {
        cx[c] = x / (a.w / a.fwd[c].w);
        cy[c] = y / (a.h / a.fwd[c].h);
        cpcs[c] = make_float2(cs_coord(cx[c], a.fwd[c].inv_w), cs_coord(cy[c], a.fwd[c].inv_h));
    }

    // 1. window initialisation from LDS
    SplatInfo si[NC][2];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int lx = (int)cx[c] - org_x[c], ly = (int)cy[c] - org_y[c];
        si[c][0] = init_search_lds<W>(s_splat[c][0], SPAN, lx, ly);
        si[c][1] = init_search_lds<W>(s_splat[c][1], SPAN, lx, ly);
    }
    // 2. searches in lock-step; the fwd search is skipped for alpha >= 1 (:335)
    Chain ch[NC][2];
    bool ran[NC][2];
    const SplatInfo window0[2] = {si[0][0], si[0][1]};  // main camera's window starts
    bool learned[2] = {false, false};
#pragma unroll
    for (int c = 0; c < NC; c++) {
        ran[c][0] = si[c][0].depth > 0.0f && fc.alpha < 1.0f;
        ran[c][1] = si[c][1].depth > 0.0f;
        if (LEARNED && c == 0) {
            // learned start: the network's offset from this pixel, where it is confident
#pragma unroll
            for (int d = 0; d < 2; d++) {
                if (d == 0 && !(fc.alpha < 1.0f)) continue;
                const ResolveView &v = d ? a.bwd[0] : a.fwd[0];
                const float *o = a.net_out + (((size_t)d * a.net_h + (y >> a.net_shift)) * a.net_w + (x >> a.net_shift)) * 16;
                if (o[2] >= a.conf_logit) {
                    // one analytic search step from the pixel itself (p - cam_ff(p)),
                    // corrected by the network's residual in cells of 2^net_shift pixels
                    const float cell = (float)(1u << a.net_shift);
                    const float4 m = at(v.cam_ff, v.w, x, y);
                    float fx = ((float)x + 0.5f - m.x * (0.5f * (float)v.w)) + o[0] * cell;
                    float fy = ((float)y + 0.5f - m.y * (0.5f * (float)v.h)) + o[1] * cell;
                    fx = fx >= 0.0f ? (fx <= (float)(v.w - 1) ? fx : (float)(v.w - 1)) : 0.0f;  // NaN -> 0
                    fy = fy >= 0.0f ? (fy <= (float)(v.h - 1) ? fy : (float)(v.h - 1)) : 0.0f;
                    si[0][d].x = (uint32_t)fx;
                    si[0][d].y = (uint32_t)fy;
                    ran[0][d] = true;
                    learned[d] = true;
                }
            }
        }
        chain_start(ch[c][0], ran[c][0], si[c][0].x, si[c][0].y, a.fwd[c]);
        chain_start(ch[c][1], ran[c][1], si[c][1].x, si[c][1].y, a.bwd[c]);
    }
    uint32_t steps = 0;  // main-camera search steps taken (learned statistics)
    for (int i = 0; i < a.max_iter; i++) {
        bool any = false;
#pragma unroll
        for (int c = 0; c < NC; c++) any = any || ch[c][0].active || ch[c][1].active;
        if (!any) break;
        if (LEARNED) steps += (uint32_t)ch[0][0].active + (uint32_t)ch[0][1].active;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            chain_step(ch[c][0], i, a.max_iter, pcs, a.fwd[c]);
            chain_step(ch[c][1], i, a.max_iter, pcs, a.bwd[c]);
        }
    }
    // depth, reprojection error and visibility at the found pixels (:338-341)
    float err[NC][2];
    bool vis[NC][2];
#pragma unroll
    for (int c = 0; c < NC; c++) {
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const ResolveView &v = d ? a.bwd[c] : a.fwd[c];
            err[c][d] = FLT_MAX;
            vis[c][d] = false;
            if (ran[c][d]) {
                const uint32_t qx = ch[c][d].h0 & 0xFFFFu, qy = ch[c][d].h0 >> 16;
                const float4 m = ch[c][d].m;
                si[c][d].x = qx;
                si[c][d].y = qy;
                si[c][d].depth = z_to_depth(at(v.z, v.w, qx, qy) + m.z);
                err[c][d] = reproj_error(make_uint2(qx, qy), v, cpcs[c], m);
                vis[c][d] = m.w > 0.0f;
            }
        }
    }
    if (LEARNED) {
        // the plain search judges the network: a learned start is redone from
        // the window start unless its search ends below EPS_1 on a surface
        // not behind the window's nearest splat candidate (the splat's depth
        // test; a fixed point on an occluded surface is valid but hidden)
        bool redo[2];
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const bool hidden = window0[d].depth > 0.0f &&
                                !(depth_to_z(si[0][d].depth) - depth_to_z(window0[d].depth) < EPS_2_FULL_RES);
            redo[d] = learned[d] && !(err[0][d] < EPS_1_FULL_RES && !hidden);
            if (redo[d]) {
                si[0][d] = window0[d];
                ran[0][d] = window0[d].depth > 0.0f && (d == 1 || fc.alpha < 1.0f);
            }
            chain_start(ch[0][d], redo[d] && ran[0][d], si[0][d].x, si[0][d].y, d ? a.bwd[0] : a.fwd[0]);
        }
        for (int i = 0; i < a.max_iter; i++) {
            if (!(ch[0][0].active || ch[0][1].active)) break;
            steps += (uint32_t)ch[0][0].active + (uint32_t)ch[0][1].active;
            chain_step(ch[0][0], i, a.max_iter, pcs, a.fwd[0]);
            chain_step(ch[0][1], i, a.max_iter, pcs, a.bwd[0]);
        }
#pragma unroll
        for (int d = 0; d < 2; d++) {
            if (!redo[d]) continue;
            const ResolveView &v = d ? a.bwd[0] : a.fwd[0];
            err[0][d] = FLT_MAX;
            vis[0][d] = false;
            if (ran[0][d]) {
                const uint32_t qx = ch[0][d].h0 & 0xFFFFu, qy = ch[0][d].h0 >> 16;
                const float4 m = ch[0][d].m;
                si[0][d].x = qx;
                si[0][d].y = qy;
                si[0][d].depth = z_to_depth(at(v.z, v.w, qx, qy) + m.z);
                err[0][d] = reproj_error(make_uint2(qx, qy), v, cpcs[0], m);
                vis[0][d] = m.w > 0.0f;
            }
        }
        if (a.learned_stats) {  // one atomic per wave and counter (ballots see active lanes only)
            const bool fwd_ok = fc.alpha < 1.0f;
            const uint32_t kept = (uint32_t)(__popcll(__ballot(learned[0] && !redo[0])) +
                                             __popcll(__ballot(learned[1] && !redo[1])));
            const uint32_t rej = (uint32_t)(__popcll(__ballot(redo[0])) + __popcll(__ballot(redo[1])));
            const uint32_t low = (uint32_t)(__popcll(__ballot(fwd_ok && !learned[0])) + __popcll(__ballot(!learned[1])));
            const uint32_t tot = (uint32_t)(__popcll(__ballot(fwd_ok)) + __popcll(__ballot(1)));
            uint32_t wave_steps = 0;  // sum over the active lanes, one ballot per bit
#pragma unroll
            for (int b = 0; b < 16; b++) wave_steps += (uint32_t)__popcll(__ballot((steps >> b) & 1u)) << b;
            const unsigned long long act = __ballot(1);
            if ((int)(threadIdx.x & 63) == __ffsll((long long)act) - 1) {
                atomicAdd(&a.learned_stats[0], kept);
                atomicAdd(&a.learned_stats[1], rej);
                atomicAdd(&a.learned_stats[2], low);
                atomicAdd(&a.learned_stats[3], tot);
                atomicAdd(&a.learned_stats[4], wave_steps);
            }
        }
    }
    if (a.q_out) {  // capture of the main camera's search results
#pragma unroll
        for (int d = 0; d < 2; d++)
            a.q_out[2 * (y * a.w + x) + d] = err[0][d] < EPS_1_FULL_RES ? pack_px(si[0][d].x, si[0][d].y) : 0xffffffffu;
    }
    // 3.-4. per-camera decision (:366-396), folded at once into the running
    // aux selection (:429-451) so no per-camera array outlives the loop
    // (a dynamically indexed array would live in scratch memory).
    Cand main_k = decide<MAV>(err[0][0], err[0][1], si[0][0].depth, si[0][1].depth, vis[0][0], vis[0][1],
                              a.fwd[0].priority, a.bwd[0].priority, EPS_1_FULL_RES, EPS_2_FULL_RES);
    bool aux_valid = false;
    Cand aux_k = Cand{false, false, 0, 0.f, 0.f, 0};
    SplatInfo aux_f = si[0][0], aux_b = si[0][1];
    int aux_c = 0;
#pragma unroll
    for (int c = 1; c < NC; c++) {
        const Cand kc = decide<false>(err[c][0], err[c][1], si[c][0].depth, si[c][1].depth, vis[c][0], vis[c][1],
                                      a.fwd[c].priority, a.bwd[c].priority, EPS_1_LOW_RES, EPS_2_LOW_RES);
        bool take = false;
        if (!aux_valid) {
            take = kc.valid;  // an invalid first result leaves "no valid best"
        } else if (kc.valid) {
            const bool similar = fabsf(depth_to_z(aux_k.depth) - depth_to_z(kc.depth)) < EPS_2_LOW_RES;
            take = (!similar && kc.depth < aux_k.depth) ||
                   (similar && kc.priority >= aux_k.priority && kc.error < aux_k.error);
        }
        if (take) {
            aux_valid = true;
            aux_k = kc;
            aux_f = si[c][0];
            aux_b = si[c][1];
            aux_c = c;
        }
    }
    // main vs aux (:453-464)
    bool use_aux = false;
    if (aux_valid) {
        use_aux = !main_k.valid;
        if (main_k.valid && !aux_k.visible_in_main) {
            const bool similar = fabsf(depth_to_z(aux_k.depth) - depth_to_z(main_k.depth)) < EPS_2_LOW_RES;
            use_aux = !similar && aux_k.depth < main_k.depth;
        }
    }
    if (!use_aux && !main_k.valid) return;  // hole: keep the previous frame's pixel
    // 5. colour of the written candidate
    const int wc = use_aux ? aux_c : 0;
    const Cand wk = use_aux ? aux_k : main_k;
    const SplatInfo wf = use_aux ? aux_f : si[0][0];
    const SplatInfo wb = use_aux ? aux_b : si[0][1];
    const uchar4 col = cand_colour(wk, fc, wf, wb, a.fwd[wc], a.bwd[wc], wc == 0 ? EPS_2_FULL_RES : EPS_2_LOW_RES);
    a.out[y * a.out_pitch_px + x] = col;
}

template <int W, bool MAV, int NAUX>
__global__ __launch_bounds__(256) void resolve_kernel(const ResolveArgs args) {
    // Read the argument block in place (constant address space) instead of
    // through a by-value copy, which the compiler may spill to scratch.
#if defined(__HIP_DEVICE_COMPILE__)
    const KernArg<ResolveArgs> &a = *kernarg_ptr<ResolveArgs>();
#else
    const ResolveArgs &a = args;  // host pass only parses the body
#endif
    __shared__ ResolveLds<W, NAUX> lds;
    resolve_tile<W, MAV, NAUX>(a, xcd_remap(blockIdx.x, a.n_tiles), lds);
}

// Learned search starts: the same resolve, with the main camera's searches
// starting where the network says (see resolve_tile, LEARNED).
template <int NAUX>
__global__ __launch_bounds__(256) void resolve_learned_kernel(const ResolveArgs args) {
#if defined(__HIP_DEVICE_COMPILE__)
    const KernArg<ResolveArgs> &a = *kernarg_ptr<ResolveArgs>();
#else
    const ResolveArgs &a = args;
#endif
    __shared__ ResolveLds<1, NAUX> lds;
    resolve_tile<1, false, NAUX, true>(a, xcd_remap(blockIdx.x, a.n_tiles), lds);
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
        f[8] = mo.x * hw; f[9] = mo.y * hh; f[10] = mo.z;
        f[11] = mq.x * hw; f[12] = mq.y * hh; f[13] = mq.z;
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

__global__ __launch_bounds__(256) void clear_batch_kernel(const BatchTable *__restrict__ tg) {
    ISTI_CONST_TABLE(t, tg);
    const XcdRange r = xcd_range(t->clear_blocks);
    for (uint32_t k = r.first; k < r.end; k += r.stride) {
        const uint32_t j = find_owner(t->clear_begin, t->n_jobs, k);
        clear_pairs(t->jobs[j].splat, t->jobs[j].n_px, (k - t->clear_begin[j]) * 256u + threadIdx.x);
    }
}

__global__ __launch_bounds__(256) void splat_batch_kernel(const BatchTable *__restrict__ tg) {
    ISTI_CONST_TABLE(t, tg);
    const XcdRange r = xcd_range(t->splat_blocks);
    for (uint32_t k = r.first; k < r.end; k += r.stride) {
        const uint32_t j = find_owner(t->splat_begin, t->n_jobs, k);
#if defined(__HIP_DEVICE_COMPILE__)
        const KernArg<SplatJob> &J = *as_const(&t->jobs[j]);
        const KernArg<BatchRequest> &R = *as_const(&t->req[J.req]);
#else
        const SplatJob &J = t->jobs[j];
        const BatchRequest &R = t->req[J.req];
#endif
        splat_pixel(J, R.vp_cam, R.vp_main[J.main_slot], R.init_linear != 0,
                    (k - t->splat_begin[j]) * SPLAT_BLOCK + threadIdx.x);
    }
}

template <int W, bool MAV, int NAUX>
__global__ __launch_bounds__(256) void resolve_batch_kernel(const BatchTable *__restrict__ tg, int variant) {
    __shared__ ResolveLds<W, NAUX> lds;
    ISTI_CONST_TABLE(t, tg);
    const uint32_t r0 = t->var_req_begin[variant], r1 = t->var_req_begin[variant + 1];
    const uint32_t base = t->tile_begin[r0];
    const XcdRange r = xcd_range(t->tile_begin[r1] - base);
    for (uint32_t k = r.first; k < r.end; k += r.stride) {
        const uint32_t q = r0 + find_owner(t->tile_begin + r0, r1 - r0, base + k);
#if defined(__HIP_DEVICE_COMPILE__)
        const KernArg<ResolveArgs> &a = *as_const(&t->req[q].args);
#else
        const ResolveArgs &a = t->req[q].args;
#endif
        resolve_tile<W, MAV, NAUX>(a, base + k - t->tile_begin[q], lds);
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

hipError_t launch_splat(const SplatArgs &a, hipStream_t s) {
    if (a.n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(reproject_splat_kernel, dim3(a.n_blocks), dim3(SPLAT_BLOCK), 0, s, a);
    return hipGetLastError();
}

template <int W, bool MAV>
static hipError_t launch_resolve_n(const ResolveArgs &a, hipStream_t s) {
    const dim3 grid(a.n_tiles), block(TILE_W * TILE_H);
    switch (a.n_aux) {
        case 0: hipLaunchKernelGGL((resolve_kernel<W, MAV, 0>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((resolve_kernel<W, MAV, 1>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((resolve_kernel<W, MAV, 2>), grid, block, 0, s, a); break;
        case 3: hipLaunchKernelGGL((resolve_kernel<W, MAV, 3>), grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL((resolve_kernel<W, MAV, 4>), grid, block, 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_resolve(const ResolveArgs &a, int window, bool always_valid, hipStream_t s) {
    if (a.n_tiles == 0) return hipSuccess;
    if (a.n_aux < 0 || a.n_aux > ISTI_MAX_AUX) return hipErrorInvalidValue;
    // the reference pairs window 1 with plain validity and window 3 with
    // main-always-valid (renderer.cu:678-689); those are the two variants built
    if (window == 1 && !always_valid) return launch_resolve_n<1, false>(a, s);
    if (window == 3 && always_valid) return launch_resolve_n<3, true>(a, s);
    return hipErrorInvalidValue;
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
    return hipLaunchKernel(fn, dim3(grid), dim3(TILE_W * TILE_H), args, 0, s);
}

int resolve_batch_blocks_per_cu(int variant) {
    void *fn = resolve_batch_ptr(variant);
    int n = 0;
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, TILE_W * TILE_H, 0) != hipSuccess) return 1;
    return n > 0 ? n : 1;
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
