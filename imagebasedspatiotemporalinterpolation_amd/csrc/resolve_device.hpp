// resolve_device.hpp - device code of the resolve pass (K2), shared by the
// per-render kernels (resolve_kernels.hip) and the batched / learned
// variants (isti_kernels.hip). Only templates, inline and file-local
// functions: every translation unit gets its own copy.
#pragma once

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

// Block id -> tile of a grid tiles_x tiles wide (row-major ids), for the
// resolve: the grid is cut into 8 vertical stripes of whole tile columns and
// XCD x (blocks x, x + 8, ...) walks stripe x row by row, so the blocks an
// XCD has in flight cover a compact patch of the frame (neighbouring tiles
// share its L2), and every XCD gets the same mix of top and bottom rows (a
// frame's cost varies far more from top to bottom - sky, ground, shadows -
// than across; per-XCD bands of whole rows measured 10 % slower on the
// 1080p + 2 aux workload). Stripes hold the contiguous ranges xcd_remap
// deals, in stripe-major order, so every tile is taken exactly once.
// stripe_major_tile: tile id of position L in that stripe-major order.
__device__ __forceinline__ uint32_t stripe_major_tile(uint32_t L, uint32_t n_tiles, uint32_t tiles_x) {
    const uint32_t rows = n_tiles / tiles_x;
    uint32_t c0 = 0, c1 = 0, acc = 0;
    for (uint32_t s = 0; s < 8; s++) {
        c0 = c1;
        c1 = ((s + 1) * tiles_x) / 8;
        const uint32_t n = rows * (c1 - c0);
        if (L < acc + n) break;
        acc += n;
    }
    const uint32_t l = L - acc, w = c1 - c0;  // w > 0: the stripe holding L is not empty
    return (l / w) * tiles_x + c0 + l % w;
}
__device__ __forceinline__ uint32_t xcd_stripe_tile(uint32_t b, uint32_t n_tiles, uint32_t tiles_x) {
    return stripe_major_tile(xcd_remap(b, n_tiles), n_tiles, tiles_x);
}

// Occupancy floor of the resolve variants (waves per SIMD). The pass spends
// most of its time issuing VALU instructions and the rest waiting on
// dependent gathers, so waves in flight pay: the 2-aux window-1 variant fits
// 64 VGPRs without spilling and runs at 8 waves, the 3-aux one at 6 (79
// VGPRs). A floor that forces spills costs more than the waves gain. 1 = no
// floor.
constexpr int resolve_min_waves(int w, int naux) { return w != 1 ? 1 : (naux == 2 ? 8 : (naux == 3 ? 6 : 1)); }
// The same floor for the batched and the learned resolve kernels, whose extra
// state (the work-table walk, the learned starts and their fallback) needs
// more registers: for the 2-aux window-1 variant 7 waves (72 VGPRs, an 8-byte
// spill) measured faster than 8 (44-byte spill) and than no floor (83-90
// VGPRs, 5 waves): batched 1080p + 2 aux frames 0.276 -> 0.253 ms, learned
// starts on that workload 7.1 -> 7.9 Gpx/s.
constexpr int resolve_min_waves_wide(int w, int naux) { return w == 1 && naux == 2 ? 7 : 1; }

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
    const unsigned long long *centre = t + (__umul24((uint32_t)ly, (uint32_t)rw) + (uint32_t)lx);  // ly, lx >= W
    const unsigned long long c = centre[0];
    SplatInfo best{key_depth(c), key_x(c), key_y(c)};
#pragma unroll
    for (int dy = -W; dy <= W; dy++) {
#pragma unroll
        for (int dx = -W; dx <= W; dx++) {
            const unsigned long long k = centre[dy * rw + dx];
            const float d = key_depth(k);
            if (d > 0.0f && (!__builtin_isfinite(best.depth) || best.depth < 0.0f || d < best.depth))
                best = SplatInfo{d, key_x(k), key_y(k)};
        }
    }
    return best;
}

// Stages the splat cells [ox, ox+rw) x [oy, oy+rh) of view v into LDS rows
// of constant stride SPAN (rw <= SPAN).
// The row of cell i is i / SPAN, taken as (i * M) >> 20 with M =
// ceil(2^20 / SPAN) on the 24-bit multiplier: for i < 2^10 the excess
// i (M - 2^20 / SPAN) / 2^20 < 2^-10 stays below 1 / SPAN (SPAN < 2^10), so
// the quotient is exact.
template <int SPAN, int NTH>
__device__ __forceinline__ void stage_splat(unsigned long long *dst, const ResolveView &v, int ox, int oy, int rw,
                                            int rh) {
    constexpr uint32_t M = ((1u << 20) + SPAN - 1) / SPAN;
    static_assert(SPAN < 1024, "staged row too long for the 24-bit row split");
    const int n = SPAN * rh;  // rh <= ROWS, and SPAN * ROWS < 2^10 (ResolveLds)
    for (int i = (int)threadIdx.x; i < n; i += NTH) {
        const int yy = (int)(__umul24((uint32_t)i, M) >> 20), xx = i - yy * SPAN;
        const int gx = ox + xx, gy = oy + yy;
        const bool in = gx >= 0 && gx < (int)v.w && gy >= 0 && gy < (int)v.h;
        if (xx < rw) dst[i] = in ? plane_ref(v.splat, px_index(v.w, (uint32_t)gx, (uint32_t)gy)) : SPLAT_EMPTY;
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
// Only cam_ff.xy is live during the search (8-byte gathers, fewer registers
// across the lock-step chains); callers fetch the full record at the end.
__device__ __forceinline__ uint32_t pack_px(uint32_t x, uint32_t y) { return (y << 16) | x; }

// One fixed-point search in flight. All searches of a pixel (fwd/bwd
// keyframe of the main and every aux camera) advance in lock-step, so their
// dependent gathers overlap instead of running one chain after another.
struct Chain {
    uint32_t h0, h1, h2, h3;  // q_i, q_{i-1}, q_{i-2}, q_{i-3} (packed)
    float2 m;                 // cam_ff.xy at h0 while the chain is active
    bool active;
};
constexpr uint32_t PX_NONE = 0xffffffffu;  // x, y <= 65534: never a packed pixel

// A chain that does not run (no splat candidate, or the skipped fwd search)
// stays inactive at its window pixel (which the always-valid blend reads).
__device__ __forceinline__ void chain_start(Chain &c, bool run, uint32_t x, uint32_t y, const ResolveView &v) {
    c.h0 = pack_px(x, y);  // a chain that does not run keeps its window pixel
    c.h1 = c.h2 = c.h3 = PX_NONE;
    c.active = run;
    c.m = run ? at_xy(v.cam_ff, v.w, x, y) : make_float2(0.f, 0.f);
}

// Step i (0-based) of n_iter: q_{i+1} = step(q_i), with the exits described above.
__device__ __forceinline__ void chain_step(Chain &c, int i, int n_iter, float2 pcs_full, const ResolveView &v) {
    if (!c.active) return;
    // The exits are evaluated side by side and applied with selects, so the
    // common paths (advance, or stop) take no nested branches; only the rare
    // cycle jump and the gather of an advancing chain are branched.
    const int2 t = cs2pixel(pcs_full.x - c.m.x, pcs_full.y - c.m.y, v.w, v.h);
    const bool inr = in_range_u(t, v.w, v.h);
    const uint32_t nt = pack_px((uint32_t)t.x, (uint32_t)t.y);  // meaningful where inr
    const bool moved = inr && nt != c.h0;
    const bool cyc = moved && ((int)(nt == c.h1) | (int)(nt == c.h2) | (int)(nt == c.h3));
    if (cyc) {
        const int p = nt == c.h1 ? 2 : (nt == c.h2 ? 3 : 4);
        const int idx = p - 1 - (n_iter - (i + 1)) % p;  // c_r = h_{p-1-r}
        c.h0 = idx == 0 ? c.h0 : (idx == 1 ? c.h1 : (idx == 2 ? c.h2 : c.h3));  // the chain stops
    }
    const bool adv = moved && !cyc;
    c.active = adv;
    c.h3 = adv ? c.h2 : c.h3;
    c.h2 = adv ? c.h1 : c.h2;
    c.h1 = adv ? c.h0 : c.h1;
    c.h0 = adv ? nt : c.h0;
    if (adv) c.m = at_xy(v.cam_ff, v.w, (uint32_t)t.x, (uint32_t)t.y);
}

// computeReprojError, renderer.cu:204-215, without its final square root:
// the binary32 sum ex*ex + ey*ey the reference takes sqrtf of. The errors
// are only ever compared, with each other (sqrt_less) and with EPS_1
// (against sqrt_threshold squares), which the squared values decide exactly.
__device__ __forceinline__ float reproj_error2(uint2 q, const ResolveView &v, float2 pcs_cam, float4 m) {
    float ex = cs_coord(q.x, v.inv_w) + m.x - pcs_cam.x;
    float ey = cs_coord(q.y, v.inv_h) + m.y - pcs_cam.y;
    ex *= v.rs_x;
    ey *= v.rs_y;
    return ex * ex + ey * ey;
}
// The error of a search that did not run: the reference's FLT_MAX. It is
// only ever compared with the error of a valid candidate (below EPS_1), and
// +inf orders against those exactly as FLT_MAX does.
constexpr float ERR2_NONE = __builtin_inff();

// RN32(sqrt(a)) < RN32(sqrt(b)) for squared errors a, b (non-negative, +inf
// or NaN). RN32 o sqrt is non-decreasing, so a >= b (or a NaN) decides
// "false". For a >= 2^-100 and b > RN32(a (1 + 2^-18)) the square roots
// differ by a factor above 1 + 2^-20, far more than the 2^-23 relative width
// of one binary32 rounding interval, so they round apart: "true". Only
// near-ties (and tiny a) take the square roots.
__device__ __forceinline__ bool sqrt_less(float a, float b) {
    const bool lt = a < b;
    bool r = lt && a >= 0x1p-100f && b > a * 0x1.00004p0f;
    if (lt && !r) {
        // The empty volatile asm keeps the square roots inside this rarely
        // taken branch (the compiler would otherwise compute them up front
        // and select).
        float ra = a, rb = b;
        asm volatile("" : "+v"(ra), "+v"(rb));
        r = sqrtf(ra) < sqrtf(rb);
    }
    return r;
}

// The texels colorFromShadow reads at one pixel of a view: colour, and, when
// the frame uses soft shadows, ShadowInfo and both shadow colours. They are
// loaded together (the shadow colours speculatively, before ShadowInfo says
// whether they are needed) so the colour costs one round trip to memory.
struct Texels {
    uchar4 img, info, sa, sb;
};

// The fields of a view the colour stage reads, taken per lane from one of two
// views (the keyframe a lane's candidate came from, or the other one). Lanes
// of a wave that chose different keyframes then share one pass through the
// colour stage instead of running it once per keyframe.
struct LaneView {
    const uchar4 *img, *sh_info, *sh_a, *sh_b;
    const float *z;
    const Flow3 *full_ff;
    uint32_t w, h;
    double inv_w, inv_h;
    int32_t forward;
};
__device__ __forceinline__ LaneView lane_view(bool second, const ResolveView &a, const ResolveView &b) {
    LaneView v;
    v.img = second ? b.img : a.img;
    v.sh_info = second ? b.sh_info : a.sh_info;
    v.sh_a = second ? b.sh_a : a.sh_a;
    v.sh_b = second ? b.sh_b : a.sh_b;
    v.z = second ? b.z : a.z;
    v.full_ff = second ? b.full_ff : a.full_ff;
    v.w = second ? b.w : a.w;
    v.h = second ? b.h : a.h;
    v.inv_w = second ? b.inv_w : a.inv_w;
    v.inv_h = second ? b.inv_h : a.inv_h;
    v.forward = second ? b.forward : a.forward;
    return v;
}

template <typename V>
__device__ __forceinline__ Texels load_texels(const V &v, uint32_t px, uint32_t py, bool shadows) {
    Texels t;
    t.img = fetch(v.img, v.w, v.h, px, py);
    t.info = t.sa = t.sb = make_uchar4(0, 0, 0, 0);
    if (shadows) {
        t.info = fetch(v.sh_info, v.w, v.h, px, py);
        t.sa = fetch(v.sh_a, v.w, v.h, px, py);
        t.sb = fetch(v.sh_b, v.w, v.h, px, py);
    }
    return t;
}

// colorFromShadow(idcs, t, ...) reads shadows only for these frames and views
// (renderer.cu:272-273): otherwise it returns the plain colour.
template <typename V>
__device__ __forceinline__ bool uses_shadows(uint2 idcs, int sfpa, bool no_shadows, const V &v) {
    return !(no_shadows || idcs.x >= (uint32_t)sfpa || idcs.y == 0 || v.sh_a == nullptr);
}

// colorFromShadow(idx, ...), renderer.cu:217-268, on loaded texels
template <typename V>
__device__ __forceinline__ uchar4 color_from_shadow_1(uint32_t vidx, int sfpa, const V &v, const Texels &t) {
    const uchar4 main_color = t.img;
    if (vidx < (uint32_t)sfpa && vidx > 0) {
        const uchar4 si = t.info;
        const uint32_t bits = ((uint32_t)si.x << 16) | ((uint32_t)si.y << 8) | (uint32_t)si.z;
        if (v.sh_b != nullptr && bits != 0) {
            const int init_val = (int)(bits & SOFTSHADOW_VAL_MASK);
            const int my_val = (int)(bits >> (SOFTSHADOW_BITS * vidx) & SOFTSHADOW_VAL_MASK);
            // min / max of the 2-bit codes of sub-frames 0..sfpa-1 (:220-224),
            // bit-parallel: which code values occur among those fields
            const uint32_t used = ((1u << (SOFTSHADOW_BITS * (uint32_t)sfpa)) - 1u) & 0x555555u;
            const uint32_t lo = bits & used, hi = (bits >> 1) & used;
            const bool has0 = (~lo & ~hi & used) != 0, has1 = (lo & ~hi) != 0, has2 = (hi & ~lo) != 0,
                       has3 = (lo & hi) != 0;
            const int min_val = has0 ? 0 : (has1 ? 1 : (has2 ? 2 : 3));
            const int max_val = has3 ? 3 : (has2 ? 2 : (has1 ? 1 : 0));
            // (my - min) / (max - min) with a quotient of small integers: RN32 of
            // RN64(a) * RN64(1/b) is the IEEE quotient (see cs_coord); 0/0 is NaN
            const int den = max_val - min_val;
            const double inv = den == 1 ? 1.0 : (den == 2 ? 0.5 : (den == 3 ? 1.0 / 3.0 : __builtin_nan("")));
            const float t01 = (float)((double)(my_val - min_val) * inv);
            const uchar4 cmin = min_val == init_val ? main_color : t.sa;
            const uchar4 cmax = max_val == init_val ? main_color : t.sb;
            const float omt = (float)(1.0 - (double)t01);
            const float rx = (float)cmin.x * omt + (float)cmax.x * t01;
            const float ry = (float)cmin.y * omt + (float)cmax.y * t01;
            const float rz = (float)cmin.z * omt + (float)cmax.z * t01;
            return make_uchar4(cvt_u8(rx), cvt_u8(ry), cvt_u8(rz), 255);
        }
    }
    return main_color;
}

// colorFromShadow(idcs, t, ...), renderer.cu:270-295 (frameIdx2Idx :227-230),
// on texels loaded with load_texels(..., uses_shadows(...))
template <typename V>
__device__ __forceinline__ uchar4 color_from_shadow(uint2 idcs, float st, int sfpa, bool no_shadows, const V &v,
                                                    const Texels &t) {
    if (!uses_shadows(idcs, sfpa, no_shadows, v)) return t.img;
    const uint32_t vx = (uint32_t)(v.forward ? (int)idcs.x : sfpa - (int)idcs.x);
    const uint32_t vy = (uint32_t)(v.forward ? (int)idcs.y : sfpa - (int)idcs.y);
    if (vx == vy || vy >= (uint32_t)sfpa) return color_from_shadow_1(vx, sfpa, v, t);
    if (vx >= (uint32_t)sfpa) return color_from_shadow_1(vy, sfpa, v, t);
    const uchar4 ca = color_from_shadow_1(vx, sfpa, v, t);
    const uchar4 cb = color_from_shadow_1(vy, sfpa, v, t);
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

// secondChanceAndBlend_kernel<EPS_2>, renderer.cu:297-327. The other
// keyframe's depth, flow and texels at the projected pixel are loaded in one
// round trip, before the similarity test decides whether they are used.
template <typename V>
__device__ __forceinline__ uchar4 second_chance(float weight, const FrameConst &fc, SplatInfo si, const V &from,
                                                const V &to, float eps2) {
    const bool sh_from = uses_shadows(fc.sidx, fc.sfpa, fc.no_shadows, from);
    const bool sh_to = uses_shadows(fc.sidx, fc.sfpa, fc.no_shadows, to);
    const Flow3 m_ft = fetch(from.full_ff, from.w, from.h, si.x, si.y);
    const float sz = fetch(from.z, from.w, from.h, si.x, si.y);
    const Texels tq = load_texels(from, si.x, si.y, sh_from);
    const float proj_x = cs_coord(si.x, from.inv_w) + m_ft.x, proj_y = cs_coord(si.y, from.inv_h) + m_ft.y;
    const float proj_z = sz + m_ft.z;
    const int2 pp = cs2pixel(proj_x, proj_y, to.w, to.h);

    uchar4 out = color_from_shadow(fc.sidx, fc.st, fc.sfpa, fc.no_shadows, from, tq);
    if (fc.frame_idx >= (uint32_t)fc.fpa) return out;
    if (in_range(pp, to.w, to.h)) {
        const uint32_t ux = (uint32_t)pp.x, uy = (uint32_t)pp.y;
        const float zt = at(to.z, to.w, ux, uy);
        const Flow3 m_tf = at(to.full_ff, to.w, ux, uy);
        const Texels tp = load_texels(to, ux, uy, sh_to);
        const float dx = m_ft.x - (-m_tf.x), dy = m_ft.y - (-m_tf.y), dz = m_ft.z - (-m_tf.z);
        const float dist2 = dx * dx + dy * dy + dz * dz;  // sqrtf(dist2) < EPS_3 exactly when dist2 < EPS_3_SQ
        if (dist2 < EPS_3_SQ && fabsf(zt - proj_z) < eps2)
            out = lerp_to_u8(color_from_shadow(fc.sidx, fc.st, fc.sfpa, fc.no_shadows, to, tp), out, weight);
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
    float error2;  // squared reprojection error (reproj_error2)
    float depth;
    int dir;  // 0: fwd keyframe, 1: bwd keyframe, 2: 0.5 blend of both (ALWAYS_VALID)
};

// Forward/backward decision of one camera, renderer.cu:366-396, on squared
// errors fe, be: e < EPS_1 is fe < eps1_sq (eps1_sq = sqrt_threshold(EPS_1)).
// With exactly one error below EPS_1 the reference's fe < be is decided by
// which one it is (fe < eps1 <= be or the reverse; a NaN error compares
// false either way, as in the reference), so the square roots of sqrt_less
// are only looked at when both are below.
template <bool ALWAYS_VALID>
__device__ __forceinline__ Cand decide(float fe, float be, float fd, float bd, bool fv, bool bv, int fprio, int bprio,
                                       float eps1_sq, float eps2) {
    Cand k{false, false, 0, 0.f, 0.f, 0};
    const bool fok = fe < eps1_sq, bok = be < eps1_sq;
    if (fok || bok) {
        const bool both = fok && bok;
        const bool similar = fabsf(depth_to_z(fd) - depth_to_z(bd)) < eps2;
        const bool f_less = both ? sqrt_less(fe, be) : (fok && !__builtin_isnan(be));
        const bool choose_fwd = (both && !similar && fd < bd) || ((!both || similar) && f_less);
        k.valid = true;
        k.visible_in_main = choose_fwd ? (fv || (both && similar && bv)) : (bv || (both && similar && fv));
        k.priority = choose_fwd ? fprio : bprio;
        k.error2 = choose_fwd ? fe : be;
        k.depth = choose_fwd ? fd : bd;
        k.dir = choose_fwd ? 0 : 1;
    } else if (ALWAYS_VALID) {
        k.valid = true;
        k.dir = 2;
    }
    return k;
}

// Colour of a candidate of camera (fwd, bwd): renderer.cu:378-396. With
// LANE_VIEWS the keyframe is picked per lane (one pass through the colour
// stage for a wave whose lanes chose different keyframes, at the price of
// per-lane addresses); without, each keyframe's pass runs with scalar
// addresses for the lanes that chose it.
template <bool LANE_VIEWS>
__device__ __forceinline__ uchar4 cand_colour(const Cand &k, const FrameConst &fc, SplatInfo fi, SplatInfo bi,
                                              const ResolveView &fwd, const ResolveView &bwd, float eps2) {
    if (!LANE_VIEWS) {
        if (k.dir == 0) return second_chance(1.0f - fc.alpha, fc, fi, fwd, bwd, eps2);
        if (k.dir == 1) return second_chance(fc.alpha, fc, bi, bwd, fwd, eps2);
    } else if (k.dir != 2) {
        const bool b = k.dir == 1;
        return second_chance(b ? fc.alpha : 1.0f - fc.alpha, fc, b ? bi : fi, lane_view(b, fwd, bwd),
                             lane_view(b, bwd, fwd), eps2);
    }
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
template <int W, int NAUX, int TW = TILE_W, int TH = TILE_H>
struct ResolveLds {
    static constexpr int NC = 1 + NAUX;
    static constexpr int SPAN = TW + 2 * W;  // staged row stride (cells)
    static constexpr int ROWS = TH + 2 * W;  // staged rows
    static_assert(SPAN * ROWS < 1024, "stage_splat splits cell indices below 2^10");
    unsigned long long splat[NC][2][SPAN * ROWS];
};

// The clear a resolve workgroup of NTH threads does for the next frame
// (ClearAhead): one 16-byte store per cell pair, non-temporal because the
// buffer is not read before the next frame's splat. resolve_tile issues it
// right after its staging loads, so the stores' acknowledgements overlap the
// latency of those loads, which the workgroup waits for at its first barrier
// anyway (issued at the end of the tile instead, the workgroup would wait for
// them before it retires; at the start, they would delay the staging loads).
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// b is the tile's index in its render (each index runs once per launch).
template <int NTH, typename A>
__device__ __forceinline__ void clear_ahead(const A &a, uint32_t b) {
    if (b >= a.ahead.n_blocks) return;
    int j = 0;
    for (int i = 1; i < a.ahead.n_bufs; i++)
        if (b >= a.ahead.block_begin[i]) j = i;
    unsigned long long *buf = a.ahead.buf[j];
    const uint32_t n = a.ahead.n[j];
    uint32_t pair = (b - a.ahead.block_begin[j]) * a.ahead.rounds * NTH + threadIdx.x;
    for (uint32_t r = 0; r < a.ahead.rounds; r++, pair += NTH) {
        const uint32_t i0 = pair * 2u;
        if (i0 + 1u < n) {
            const u64x2 e = {SPLAT_EMPTY, SPLAT_EMPTY};
            __builtin_nontemporal_store(e, reinterpret_cast<u64x2 *>(buf) + pair);
        } else if (i0 < n) {
            __builtin_nontemporal_store(SPLAT_EMPTY, buf + i0);
        }
    }
}

// One TW x TH tile of one render, one thread per pixel (TW * TH threads: 64,
// one wave, or up to 256). Every thread of the block must call it (the LDS
// staging ends in a barrier); pixels outside the image return early. AHEAD:
// the kernel clears its share of a.ahead (only the per-render kernel; the
// batched and learned-start kernels are built without that code).
template <int W, bool MAV, int NAUX, bool LEARNED = false, int TW = TILE_W, int TH = TILE_H, bool AHEAD = false,
          typename A>
__device__ __forceinline__ void resolve_tile(const A &a, uint32_t tile, ResolveLds<W, NAUX, TW, TH> &lds) {
    constexpr int NTH = TW * TH;  // threads of the block
    static_assert(NTH % 64 == 0 && NTH <= 256 && (TW & (TW - 1)) == 0, "tile shape");
    constexpr int NC = 1 + NAUX;
    constexpr int SPAN = TW + 2 * W;
    auto &s_splat = lds.splat;
    const uint32_t ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const uint32_t x0 = tx * TW, y0 = ty * TH;
    const uint32_t x = x0 + (threadIdx.x & (TW - 1));
    const uint32_t y = y0 + (threadIdx.x / TW);
    // Learned starts: the network's output at this pixel's cell and cam_ff.xy
    // at the pixel depend on the pixel alone, so they are loaded before the
    // staging and its barriers, which hide their latency (loaded after them,
    // every wave waited for them at the start of its searches).
    float net_o[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    float2 net_m[2] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    if (LEARNED && x < a.w && y < a.h) {
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const ResolveView &v = d ? a.bwd[0] : a.fwd[0];
            const float *o = a.net_out + (((size_t)d * a.net_h + (y >> a.net_shift)) * a.net_w + (x >> a.net_shift)) * 16;
            net_o[d][0] = o[0];
            net_o[d][1] = o[1];
            net_o[d][2] = o[2];
            net_m[d] = at_xy(v.cam_ff, v.w, x, y);
        }
    }

    // Stage every splat cell the tile's windows read: camera c covers camera
    // pixels [x0/rx, xl/rx] (xl = last in-image column of the tile) plus a
    // W-cell border, at most SPAN x SPAN cells.
    int org_x[NC], org_y[NC], span_x[NC], span_y[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const ResolveView &v = a.fwd[c];
        const uint32_t xl = min(x0 + TW - 1, a.w - 1), yl = min(y0 + TH - 1, a.h - 1);
        const uint32_t cx0 = div_small(x0, v.ratio_x, v.magic_x), cy0 = div_small(y0, v.ratio_y, v.magic_y);
        org_x[c] = (int)cx0 - W;
        org_y[c] = (int)cy0 - W;
        span_x[c] = (int)(div_small(xl, v.ratio_x, v.magic_x) - cx0) + 1 + 2 * W;
        span_y[c] = (int)(div_small(yl, v.ratio_y, v.magic_y) - cy0) + 1 + 2 * W;
        stage_splat<SPAN, NTH>(s_splat[c][0], a.fwd[c], org_x[c], org_y[c], span_x[c], span_y[c]);
        stage_splat<SPAN, NTH>(s_splat[c][1], a.bwd[c], org_x[c], org_y[c], span_x[c], span_y[c]);
    }
    if constexpr (AHEAD) clear_ahead<NTH>(a, tile);
    __syncthreads();
    // 1a. aux cameras: a window start depends only on the aux pixel, which
    // ratio_x * ratio_y render pixels share, so each distinct aux pixel of the
    // tile is initialised once and its start replaces the staged cells (packed
    // like a splat key) for the pixels to read.
    if constexpr (NAUX > 0) {
        constexpr int PER = (NAUX * 2 * TW * TH + NTH - 1) / NTH;
        unsigned long long res[PER];
        int nx[NC], cnt[NC], first[NC + 1];
        float rcp_nx[NC];  // RN32(1/nx): cell rows by div_px instead of an integer divide
        first[1] = 0;
#pragma unroll
        for (int c = 1; c < NC; c++) {
            nx[c] = span_x[c] - 2 * W;
            cnt[c] = nx[c] * (span_y[c] - 2 * W);
            first[c + 1] = first[c] + 2 * cnt[c];
            rcp_nx[c] = 1.0f / (float)nx[c];
        }
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int k = (int)threadIdx.x + r * NTH;
            res[r] = SPLAT_EMPTY;
#pragma unroll
            for (int c = 1; c < NC; c++) {
                if (k >= first[c] && k < first[c + 1]) {
                    const int j = k - first[c], d = j >= cnt[c] ? 1 : 0, p = j - d * cnt[c];
                    const int py = (int)div_px((uint32_t)p, (uint32_t)nx[c], rcp_nx[c]), px = p - py * nx[c];
                    const SplatInfo w = init_search_lds<W>(s_splat[c][d], SPAN, px + W, py + W);
                    res[r] = encode_key(w.depth, w.x, w.y);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int k = (int)threadIdx.x + r * NTH;
#pragma unroll
            for (int c = 1; c < NC; c++) {
                if (k >= first[c] && k < first[c + 1]) {
                    const int j = k - first[c], d = j >= cnt[c] ? 1 : 0;
                    s_splat[c][d][j - d * cnt[c]] = res[r];
                }
            }
        }
        __syncthreads();
    }
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
        cx[c] = div_small(x, a.fwd[c].ratio_x, a.fwd[c].magic_x);
        cy[c] = div_small(y, a.fwd[c].ratio_y, a.fwd[c].magic_y);
        cpcs[c] = make_float2(cs_coord(cx[c], a.fwd[c].inv_w), cs_coord(cy[c], a.fwd[c].inv_h));
    }

    // 1. window initialisation from LDS
    SplatInfo si[NC][2];
    si[0][0] = init_search_lds<W>(s_splat[0][0], SPAN, (int)x - org_x[0], (int)y - org_y[0]);
    si[0][1] = init_search_lds<W>(s_splat[0][1], SPAN, (int)x - org_x[0], (int)y - org_y[0]);
#pragma unroll
    for (int c = 1; c < NC; c++) {
        const int p = ((int)cy[c] - org_y[c] - W) * (span_x[c] - 2 * W) + ((int)cx[c] - org_x[c] - W);
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const unsigned long long k = s_splat[c][d][p];
            si[c][d] = SplatInfo{key_depth(k), key_x(k), key_y(k)};
        }
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
                const float *o = net_o[d];
                if (o[2] >= a.conf_logit) {
                    // one analytic search step from the pixel itself (p - cam_ff(p)),
                    // corrected by the network's residual in cells of 2^net_shift pixels
                    const float cell = (float)(1u << a.net_shift);
                    const float2 m = net_m[d];
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
    uint32_t steps = 0;  // main-camera search steps taken (learned statistics only)
    // a learned start gets at most LEARNED_STEPS steps: one still moving by
    // then is judged at that point (and falls back to the window start if
    // rejected), so no wave waits for a wandering learned chain
    const int n_learned = a.max_iter < LEARNED_STEPS ? a.max_iter : LEARNED_STEPS;
    for (int i = 0; i < a.max_iter; i++) {
        if (LEARNED && i == n_learned) {  // (a stopped chain stays stopped)
#pragma unroll
            for (int d = 0; d < 2; d++)
                if (learned[d]) ch[0][d].active = false;
        }
        bool any = false;
#pragma unroll
        for (int c = 0; c < NC; c++) any = any || ch[c][0].active || ch[c][1].active;
        if (!any) break;
        if (LEARNED && a.learned_stats) steps += (uint32_t)ch[0][0].active + (uint32_t)ch[0][1].active;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            chain_step(ch[c][0], i, (LEARNED && c == 0 && learned[0]) ? n_learned : a.max_iter, pcs, a.fwd[c]);
            chain_step(ch[c][1], i, (LEARNED && c == 0 && learned[1]) ? n_learned : a.max_iter, pcs, a.bwd[c]);
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
            err[c][d] = ERR2_NONE;
            vis[c][d] = false;
            // the search result, or the window pixel of a search that did not
            // run (its depth is never read: its error is the reference's FLT_MAX)
            const uint32_t qx = ch[c][d].h0 & 0xFFFFu, qy = ch[c][d].h0 >> 16;
            si[c][d] = SplatInfo{0.0f, qx, qy};
            if (ran[c][d]) {
                const float4 m = at(v.cam_ff, v.w, qx, qy);
                si[c][d].depth = z_to_depth(m.z);  // m.z = z + flow z (the splat's sum)
                err[c][d] = reproj_error2(make_uint2(qx, qy), v, cpcs[c], m);
                vis[c][d] = m.w > 0.0f;
            }
        }
    }
    if (LEARNED) {
        // the plain search judges the network: a learned start is redone from
        // the window start unless its search ends below EPS_1 on a surface
        // not behind the window's nearest splat candidate (the splat's depth
        // test; a fixed point on an occluded surface is valid but hidden)
        bool redo[2] = {false, false};
        if (__ballot(learned[0] || learned[1])) {  // waves without a learned start skip the check
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const bool hidden = window0[d].depth > 0.0f &&
                                    !(depth_to_z(si[0][d].depth) - depth_to_z(window0[d].depth) < EPS_2_FULL_RES);
                redo[d] = learned[d] && !(err[0][d] < EPS_1_FULL_RES_SQ && !hidden);
                if (redo[d]) {
                    si[0][d] = window0[d];
                    ran[0][d] = window0[d].depth > 0.0f && (d == 1 || fc.alpha < 1.0f);
                }
                chain_start(ch[0][d], redo[d] && ran[0][d], si[0][d].x, si[0][d].y, d ? a.bwd[0] : a.fwd[0]);
            }
            for (int i = 0; i < a.max_iter; i++) {
                if (!(ch[0][0].active || ch[0][1].active)) break;
                if (a.learned_stats) steps += (uint32_t)ch[0][0].active + (uint32_t)ch[0][1].active;
                chain_step(ch[0][0], i, a.max_iter, pcs, a.fwd[0]);
                chain_step(ch[0][1], i, a.max_iter, pcs, a.bwd[0]);
            }
#pragma unroll
            for (int d = 0; d < 2; d++) {
                if (!redo[d]) continue;
                const ResolveView &v = d ? a.bwd[0] : a.fwd[0];
                err[0][d] = ERR2_NONE;
                vis[0][d] = false;
                if (ran[0][d]) {
                    const uint32_t qx = ch[0][d].h0 & 0xFFFFu, qy = ch[0][d].h0 >> 16;
                    const float4 m = at(v.cam_ff, v.w, qx, qy);
                    si[0][d].x = qx;
                    si[0][d].y = qy;
                    si[0][d].depth = z_to_depth(m.z);
                    err[0][d] = reproj_error2(make_uint2(qx, qy), v, cpcs[0], m);
                    vis[0][d] = m.w > 0.0f;
                }
            }
        }
        if (a.learned_stats) {  // one atomic per wave and counter into the block's slot row (ballots see active lanes only)
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
                uint32_t *row = a.learned_stats + (blockIdx.x % STAT_SLOTS) * STAT_STRIDE;
                atomicAdd(&row[0], kept);
                atomicAdd(&row[1], rej);
                atomicAdd(&row[2], low);
                atomicAdd(&row[3], tot);
                atomicAdd(&row[4], wave_steps);
            }
        }
    }
    if (a.q_out) {  // capture of the main camera's search results
#pragma unroll
        for (int d = 0; d < 2; d++)
            a.q_out[2 * (y * a.w + x) + d] = err[0][d] < EPS_1_FULL_RES_SQ ? pack_px(si[0][d].x, si[0][d].y) : 0xffffffffu;
    }
    // 3.-4. per-camera decision (:366-396), folded at once into the running
    // aux selection (:429-451) so no per-camera array outlives the loop
    // (a dynamically indexed array would live in scratch memory).
    Cand main_k = decide<MAV>(err[0][0], err[0][1], si[0][0].depth, si[0][1].depth, vis[0][0], vis[0][1],
                              a.fwd[0].priority, a.bwd[0].priority, EPS_1_FULL_RES_SQ, EPS_2_FULL_RES);
    bool aux_valid = false;
    Cand aux_k = Cand{false, false, 0, 0.f, 0.f, 0};
    SplatInfo aux_f = si[0][0], aux_b = si[0][1];
    int aux_c = 0;
#pragma unroll
    for (int c = 1; c < NC; c++) {
        const Cand kc = decide<false>(err[c][0], err[c][1], si[c][0].depth, si[c][1].depth, vis[c][0], vis[c][1],
                                      a.fwd[c].priority, a.bwd[c].priority, EPS_1_LOW_RES_SQ, EPS_2_LOW_RES);
        bool take = false;
        if (!aux_valid) {
            take = kc.valid;  // an invalid first result leaves "no valid best"
        } else if (kc.valid) {
            const bool similar = fabsf(depth_to_z(aux_k.depth) - depth_to_z(kc.depth)) < EPS_2_LOW_RES;
            take = (!similar && kc.depth < aux_k.depth) ||
                   (similar && kc.priority >= aux_k.priority && sqrt_less(kc.error2, aux_k.error2));
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
    // One uniform branch per camera, so each camera's view descriptors stay
    // scalar (a per-lane camera index would select every descriptor field
    // per lane); a wave runs the branches its lanes chose.
    uchar4 col = make_uchar4(0, 0, 0, 0);
    if constexpr (NC == 1) {
        col = cand_colour<false>(wk, fc, wf, wb, a.fwd[0], a.bwd[0], EPS_2_FULL_RES);
    } else {
#pragma unroll
        for (int c = 0; c < NC; c++)
            if (wc == c)
                col = cand_colour<true>(wk, fc, wf, wb, a.fwd[c], a.bwd[c], c == 0 ? EPS_2_FULL_RES : EPS_2_LOW_RES);
    }
    a.out[y * a.out_pitch_px + x] = col;
}

}  // namespace isti
