// oracle.cpp - CPU restatement of the reference's per-frame interpolation
// path (DerThomy/ImageBasedSpatioTemporalInterpolation, src/renderer.cu).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP renderer. Only the test
// suite, the smoke check and the benchmark's CPU-baseline leg load it; the
// product path never does.
//
// Every function below names the reference lines it restates. The code is
// scalar and literal on purpose: one IEEE binary32 operation per operator,
// in the order the reference writes it, compiled without FMA contraction.
// Parity pins (closed forms, invariants, brute force, known-answer tests)
// live in tests/test_oracle_*.py.

#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// src/renderer.cu:16-22, 27-28
constexpr float EPS_1_FULL_RES = 0.002f;
constexpr float EPS_1_LOW_RES = EPS_1_FULL_RES * 4.0f;
constexpr float EPS_2_FULL_RES = 0.001f;
constexpr float EPS_2_LOW_RES = 0.004f;
constexpr float EPS_3 = 0.01f;
constexpr uint64_t SPLAT_EMPTY = 0xfefefefefefefefeull;  // memset 0xfe
constexpr uint32_t SOFTSHADOW_BITS = 2;
constexpr uint32_t SOFTSHADOW_VAL_MASK = (1u << SOFTSHADOW_BITS) - 1u;

struct F2 { float x, y; };
struct F4 { float x, y, z, w; };
struct U2 { uint32_t x, y; };
struct I2 { int32_t x, y; };
struct C4 { uint8_t x, y, z, w; };

// ---------------------------------------------------------------------------
// Conversions. C++ leaves float->int of NaN / out-of-range values undefined;
// the reference ran them as GPU conversions (NaN -> 0, saturating).
// helper_math.h:117-120 make_int2(float2) truncates.
// ---------------------------------------------------------------------------
inline int32_t cvt_s32(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}
// (unsigned char)(float) as in make_uchar4(float4), helper_math_ext.h:19-22
inline uint8_t cvt_u8(float f) {
    if (f != f) return 0;
    if (f <= 0.0f) return 0;
    if (f >= 255.0f) return 255;
    return (uint8_t)f;
}
inline uint32_t sat_u32(float r) {
    if (r != r) return 0;
    if (r <= 0.0f) return 0;
    if (r >= 4294967296.0f) return UINT32_MAX;
    return (uint32_t)r;
}
// __float2uint_rn / _rd / _ru (src/renderer.cu:359-362)
inline uint32_t f2u_rn(float f) { return sat_u32(std::nearbyint(f)); }
inline uint32_t f2u_rd(float f) { return sat_u32(std::floor(f)); }
inline uint32_t f2u_ru(float f) { return sat_u32(std::ceil(f)); }

// ---------------------------------------------------------------------------
// helper_common.h:8-15 inRange / to1D
// ---------------------------------------------------------------------------
inline bool in_range(I2 p, U2 d) {
    return p.x >= 0 && (uint32_t)p.x < d.x && p.y >= 0 && (uint32_t)p.y < d.y;
}
inline size_t to1d(U2 p, U2 d) { return (size_t)p.y * d.x + p.x; }

// helper_common.h:66-69 pixel2CS: 2*(p+0.5)/res - 1
inline F2 pixel2cs(U2 p, U2 res) {
    F2 r;
    r.x = 2.0f * ((float)p.x + 0.5f) / (float)res.x - 1.0f;
    r.y = 2.0f * ((float)p.y + 0.5f) / (float)res.y - 1.0f;
    return r;
}
// helper_common.h:71-75 CS2pixel: floor((cs+1)*res*0.5), truncated to int
inline I2 cs2pixel(F2 c, U2 res) {
    I2 r;
    r.x = cvt_s32(std::floor((c.x + 1.0f) * (float)res.x * 0.5f));
    r.y = cvt_s32(std::floor((c.y + 1.0f) * (float)res.y * 0.5f));
    return r;
}
// glm mat4 * vec4 (column-major), association (m0 v.x + m1 v.y) + (m2 v.z + m3 v.w)
inline F4 mat_vec(const float *m, F4 v) {
    F4 r;
    r.x = (m[0] * v.x + m[4] * v.y) + (m[8] * v.z + m[12] * v.w);
    r.y = (m[1] * v.x + m[5] * v.y) + (m[9] * v.z + m[13] * v.w);
    r.z = (m[2] * v.x + m[6] * v.y) + (m[10] * v.z + m[14] * v.w);
    r.w = (m[3] * v.x + m[7] * v.y) + (m[11] * v.z + m[15] * v.w);
    return r;
}
inline F4 div_w(F4 p) {
    const float w = p.w;
    return F4{p.x / w, p.y / w, p.z / w, p.w / w};
}
// helper_common.h:77-87 CS2WS / WS2CS: homogeneous product then divide by w
inline F4 cs2ws(F4 p, const float *inv_vp) { return div_w(mat_vec(inv_vp, p)); }
inline F4 ws2cs(F4 p, const float *vp) { return div_w(mat_vec(vp, p)); }
// helper_common.h:89-95
inline float z_to_depth(float z) { return (z + 1.0f) * 0.5f; }
inline float depth_to_z(float d) { return d * 2.0f - 1.0f; }

// helper_common.h:22-38 lerp_color / to_8bit / color_to_uchar4
inline int32_t to_8bit(float f) { return std::min(255, std::max(0, cvt_s32(f * 255.f))); }
inline C4 lerp_to_u8(C4 a, C4 b, float t) {
    // lerp(float4(a), float4(b), t) / 255 = (a + t*(b-a)) / 255, then to_8bit
    auto ch = [t](uint8_t ua, uint8_t ub) -> uint8_t {
        const float fa = (float)ua, fb = (float)ub;
        const float l = (fa + t * (fb - fa)) / 255.f;
        return (uint8_t)to_8bit(l);
    };
    return C4{ch(a.x, b.x), ch(a.y, b.y), ch(a.z, b.z), ch(a.w, b.w)};
}

// renderer.cu:59-74 splat key codec: depth bits << 32 | x << 16 | y. The
// reference runs on NVIDIA GPUs, whose binary32 arithmetic returns the
// canonical NaN 0x7fffffff; a NaN depth is encoded with those bits (an x86
// NaN is negative and would sort after the empty-cell value).
inline uint64_t encode_key(float depth, U2 src) {
    uint32_t bits = 0x7fffffffu;
    if (depth == depth) std::memcpy(&bits, &depth, 4);
    return ((uint64_t)bits << 32) | (uint64_t)((src.x << 16) | src.y);
}
struct SplatInfo { float depth; U2 src; };
inline SplatInfo decode_key(uint64_t key) {
    const uint32_t bits = (uint32_t)(key >> 32);
    SplatInfo s;
    std::memcpy(&s.depth, &bits, 4);
    s.src.x = (uint32_t)key >> 16;
    s.src.y = (uint32_t)(key & 0xFFFFull);
    return s;
}

inline void atomic_min_u64(uint64_t *p, uint64_t v) {
    uint64_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < cur) {
        if (__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
            break;
    }
}

// CamReprojInfo (include/renderer.h:37-50) over linear host planes. Fetches
// are point samples with border addressing (out of range -> 0), as the
// reference's textures (src/util/image_buffer.cu:227-249). A missing plane
// (texture handle -1 in the reference) reads as 0.
struct View {
    U2 res{0, 0};
    bool forward = true;
    int priority = 0;
    const C4 *img = nullptr;
    const float *z = nullptr;
    const C4 *shadow_a = nullptr;
    const C4 *shadow_b = nullptr;
    const C4 *shadow_info = nullptr;
    const F4 *cam_ff = nullptr;
    const F4 *full_ff = nullptr;
    const uint64_t *splat = nullptr;
};
template <typename T>
inline T fetch(const T *plane, U2 res, U2 p) {
    if (plane == nullptr || p.x >= res.x || p.y >= res.y) return T{};
    return plane[to1d(p, res)];
}

struct Settings {  // AdditionalRenderInfo, include/renderer.h:18-24
    int fpa;
    int sfpa;
    bool no_shadows;
};

// ---------------------------------------------------------------------------
// K1: computeFlowFields_sortedSplatting_kernel, src/renderer.cu:76-146
// ---------------------------------------------------------------------------
struct SplatJob {
    U2 res;
    float weight;
    bool init_linear;
    const float *z;
    const float *mx, *my, *mz;  // all null: no object motion
    const float *inv_vp_from, *vp_to, *vp_cam, *vp_from_main;
    F4 *full_ff, *cam_ff;
    uint64_t *splat;
};

void splat_pixel(const SplatJob &j, U2 idx) {
    const size_t i = to1d(idx, j.res);
    const float z_from = j.z[i];
    F4 om{0.f, 0.f, 0.f, 0.f};
    if (j.mx != nullptr) om = F4{j.mx[i], j.my[i], j.mz[i], 0.f};

    const F2 pcs = pixel2cs(idx, j.res);
    const F4 pos_cs_from{pcs.x, pcs.y, z_from, 1.0f};
    const F4 pos_ws_from = cs2ws(pos_cs_from, j.inv_vp_from);
    const F4 pos_cs_moved{pos_cs_from.x + om.x, pos_cs_from.y + om.y, pos_cs_from.z + om.z,
                          pos_cs_from.w + 0.0f};
    const F4 pos_ws_to = cs2ws(pos_cs_moved, j.inv_vp_from);

    const F4 pos_cs_main = ws2cs(pos_ws_from, j.vp_from_main);
    const float lim = 1.0f - 1e-3f;
    const bool visible = pos_cs_main.x > -lim && pos_cs_main.x < lim && pos_cs_main.y > -lim &&
                         pos_cs_main.y < lim && pos_cs_main.z > -lim && pos_cs_main.z < lim;
    const float vis = visible ? 1.0f : 0.0f;

    const F4 pos_ws_scaled{pos_ws_from.x + j.weight * (pos_ws_to.x - pos_ws_from.x),
                           pos_ws_from.y + j.weight * (pos_ws_to.y - pos_ws_from.y),
                           pos_ws_from.z + j.weight * (pos_ws_to.z - pos_ws_from.z),
                           pos_ws_from.w + j.weight * (pos_ws_to.w - pos_ws_from.w)};
    const F4 pos_cs_to = ws2cs(pos_ws_to, j.vp_to);
    const F4 full{pos_cs_to.x - pos_cs_from.x, pos_cs_to.y - pos_cs_from.y,
                  pos_cs_to.z - pos_cs_from.z, pos_cs_to.w - pos_cs_from.w};
    j.full_ff[i] = F4{full.x, full.y, full.z, vis};

    float depth = z_to_depth(pos_cs_from.z);
    U2 pix = idx;
    if (!j.init_linear) {
        const F4 pos_cs_cam = ws2cs(pos_ws_scaled, j.vp_cam);
        j.cam_ff[i] = F4{pos_cs_cam.x - pos_cs_from.x, pos_cs_cam.y - pos_cs_from.y,
                         pos_cs_cam.z - pos_cs_from.z, vis};
        const float depth_cam = z_to_depth(pos_cs_cam.z);
        const I2 pc = cs2pixel(F2{pos_cs_cam.x, pos_cs_cam.y}, j.res);
        if (!in_range(pc, j.res) || depth_cam < 0.0f) return;
        depth = depth_cam;
        pix = U2{(uint32_t)pc.x, (uint32_t)pc.y};
    } else {
        j.cam_ff[i] = F4{full.x * j.weight, full.y * j.weight, full.z * j.weight, vis};
    }
    atomic_min_u64(&j.splat[to1d(pix, j.res)], encode_key(depth, idx));
}

// ---------------------------------------------------------------------------
// K2 helpers, src/renderer.cu:148-398
// ---------------------------------------------------------------------------
// initializeSearchFromSplatting<W>, :148-175
SplatInfo init_search(U2 idx, U2 res, const uint64_t *splat, int W) {
    SplatInfo best = decode_key(splat[to1d(idx, res)]);
    for (int dy = -W; dy <= W; dy++) {
        for (int dx = -W; dx <= W; dx++) {
            const I2 n{(int32_t)idx.x + dx, (int32_t)idx.y + dy};
            if (!in_range(n, res)) continue;
            const SplatInfo cur = decode_key(splat[to1d(U2{(uint32_t)n.x, (uint32_t)n.y}, res)]);
            if (cur.depth > 0.0f &&
                (!std::isfinite(best.depth) || best.depth < 0.0f || cur.depth < best.depth))
                best = cur;
        }
    }
    return best;
}

// iterativeSearch<N>, :177-202 (N full steps, no early exit)
U2 search(U2 start, U2 idx, U2 res, U2 low_res, const F4 *ff, int N) {
    U2 result = start;
    for (int i = 0; i < N; i++) {
        const F4 m = fetch(ff, low_res, result);
        const F2 p = pixel2cs(idx, res);
        const I2 t = cs2pixel(F2{p.x - m.x, p.y - m.y}, low_res);
        if (!in_range(t, low_res)) continue;
        result = U2{(uint32_t)t.x, (uint32_t)t.y};
    }
    return result;
}

// computeReprojError, :204-215
float reproj_error(U2 q, U2 res_from, F2 pcs_cam, F4 m) {
    const F2 pf = pixel2cs(q, res_from);
    F2 e{pf.x + m.x - pcs_cam.x, pf.y + m.y - pcs_cam.y};
    const float mn = (float)std::min(res_from.x, res_from.y);
    const F2 s{(float)res_from.x / mn, (float)res_from.y / mn};
    e.x *= s.x;
    e.y *= s.y;
    return std::sqrt(e.x * e.x + e.y * e.y);
}

// colorFromShadow(idx, ...), :217-268 (getMinMaxShadowValues :217-225)
C4 color_from_shadow_1(uint32_t vidx, int sfpa, const View &v, U2 pixel) {
    const C4 main_color = fetch(v.img, v.res, pixel);
    if (vidx < (uint32_t)sfpa && vidx > 0) {
        const C4 si = fetch(v.shadow_info, v.res, pixel);
        const uint32_t bits = ((uint32_t)si.x << 16) | ((uint32_t)si.y << 8) | (uint32_t)si.z;
        if (v.shadow_b != nullptr) {
            if (bits != 0) {
                const int init_val = (int)(bits & SOFTSHADOW_VAL_MASK);
                const int my_val = (int)(bits >> (SOFTSHADOW_BITS * vidx) & SOFTSHADOW_VAL_MASK);
                int min_val = (int)SOFTSHADOW_VAL_MASK, max_val = 0;
                for (int i = 0; i < sfpa; i++) {
                    const int c = (int)(bits >> (SOFTSHADOW_BITS * (uint32_t)i) & SOFTSHADOW_VAL_MASK);
                    min_val = std::min(min_val, c);
                    max_val = std::max(max_val, c);
                }
                const float t = (float)(my_val - min_val) / (float)(max_val - min_val);
                const C4 cmin = min_val == init_val ? main_color : fetch(v.shadow_a, v.res, pixel);
                const C4 cmax = max_val == init_val ? main_color : fetch(v.shadow_b, v.res, pixel);
                // float4 * (1.0 - t): the double difference is narrowed to float
                const float omt = (float)(1.0 - (double)t);
                const float rx = (float)cmin.x * omt + (float)cmax.x * t;
                const float ry = (float)cmin.y * omt + (float)cmax.y * t;
                const float rz = (float)cmin.z * omt + (float)cmax.z * t;
                return C4{cvt_u8(rx), cvt_u8(ry), cvt_u8(rz), 255};
            }
        }
    }
    return main_color;
}

// frameIdx2Idx, :227-230
inline uint32_t frame_idx_to_idx(bool forward, uint32_t frame_idx, int fpa) {
    const int r = forward ? (int)frame_idx : (fpa - (int)frame_idx);
    return (uint32_t)r;
}

// colorFromShadow(idcs, t, ...), :270-295
C4 color_from_shadow(U2 idcs, float t, const Settings &s, const View &v, U2 pixel) {
    if (s.no_shadows || idcs.x >= (uint32_t)s.sfpa || idcs.y == 0 || v.shadow_a == nullptr)
        return fetch(v.img, v.res, pixel);
    const U2 vi{frame_idx_to_idx(v.forward, idcs.x, s.sfpa), frame_idx_to_idx(v.forward, idcs.y, s.sfpa)};
    if (vi.x == vi.y || vi.y >= (uint32_t)s.sfpa) return color_from_shadow_1(vi.x, s.sfpa, v, pixel);
    if (vi.x >= (uint32_t)s.sfpa) return color_from_shadow_1(vi.y, s.sfpa, v, pixel);
    const C4 a = color_from_shadow_1(vi.x, s.sfpa, v, pixel);
    const C4 b = color_from_shadow_1(vi.y, s.sfpa, v, pixel);
    return lerp_to_u8(a, b, t);
}

// secondChanceAndBlend_kernel<EPS_2>, :297-327
C4 second_chance(float weight, uint32_t frame_idx, U2 sidx, float st, SplatInfo si, const Settings &s,
                 const View &from, const View &to, float eps2) {
    const F4 m_ft = fetch(from.full_ff, from.res, si.src);
    const F2 spcs = pixel2cs(si.src, from.res);
    const float sz = fetch(from.z, from.res, si.src);
    const float proj_x = spcs.x + m_ft.x, proj_y = spcs.y + m_ft.y, proj_z = sz + m_ft.z;
    const I2 pp = cs2pixel(F2{proj_x, proj_y}, to.res);

    C4 out = color_from_shadow(sidx, st, s, from, si.src);
    if (frame_idx >= (uint32_t)s.fpa) return out;

    if (in_range(pp, to.res)) {
        const U2 up{(uint32_t)pp.x, (uint32_t)pp.y};
        const float zt = fetch(to.z, to.res, up);
        const F4 m_tf = fetch(to.full_ff, to.res, up);
        const float dx = m_ft.x - (-m_tf.x), dy = m_ft.y - (-m_tf.y), dz = m_ft.z - (-m_tf.z);
        const float dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (dist < EPS_3 && std::fabs(zt - proj_z) < eps2)
            out = lerp_to_u8(color_from_shadow(sidx, st, s, to, up), out, weight);
    }
    return out;
}

struct Result {  // ReprojResult, :49-57
    bool valid = false;
    bool visible_in_main = false;
    int priority = 0;
    float error = 0.f;
    float depth = 0.f;
    C4 color{0, 0, 0, 0};
};

// bidirectionalReprojection<SW, SI, ALWAYS_VALID, EPS_1, EPS_2>, :329-398
Result bidir(U2 idx, U2 full_idx, U2 full_res, float alpha, F2 pcs, const Settings &s, const View &fwd,
             const View &bwd, int SW, int SI, bool always_valid, float eps1, float eps2) {
    float fwd_e = FLT_MAX;
    bool fwd_vis = false;
    SplatInfo fi = init_search(idx, fwd.res, fwd.splat, SW);
    if (fi.depth > 0.0f && alpha < 1.0f) {
        fi.src = search(fi.src, full_idx, full_res, fwd.res, fwd.cam_ff, SI);
        const F4 m = fetch(fwd.cam_ff, fwd.res, fi.src);
        fi.depth = z_to_depth(fetch(fwd.z, fwd.res, fi.src) + m.z);
        fwd_e = reproj_error(fi.src, fwd.res, pcs, m);
        fwd_vis = m.w > 0.0f;
    }
    float bwd_e = FLT_MAX;
    bool bwd_vis = false;
    SplatInfo bi = init_search(idx, bwd.res, bwd.splat, SW);
    if (bi.depth > 0.0f) {
        bi.src = search(bi.src, full_idx, full_res, bwd.res, bwd.cam_ff, SI);
        const F4 m = fetch(bwd.cam_ff, bwd.res, bi.src);
        bi.depth = z_to_depth(fetch(bwd.z, bwd.res, bi.src) + m.z);
        bwd_e = reproj_error(bi.src, bwd.res, pcs, m);
        bwd_vis = m.w > 0.0f;
    }

    const uint32_t frame_idx = f2u_rn((float)s.fpa * alpha);
    const U2 sidx{f2u_rd((float)s.sfpa * alpha + 1e-9f), f2u_ru((float)s.sfpa * alpha - 1e-9f)};
    const float st = (float)s.sfpa * alpha - (float)sidx.x;

    Result r;
    if (fwd_e < eps1 || bwd_e < eps1) {
        const bool both = fwd_e < eps1 && bwd_e < eps1;
        const bool similar = std::fabs(depth_to_z(fi.depth) - depth_to_z(bi.depth)) < eps2;
        const bool choose_fwd = (both && !similar && fi.depth < bi.depth) || ((!both || similar) && fwd_e < bwd_e);
        if (choose_fwd) {
            r.valid = true;
            r.visible_in_main = fwd_vis || (both && similar && bwd_vis);
            r.priority = fwd.priority;
            r.error = fwd_e;
            r.depth = fi.depth;
            r.color = second_chance(1.0f - alpha, frame_idx, sidx, st, fi, s, fwd, bwd, eps2);
        } else {
            r.valid = true;
            r.visible_in_main = bwd_vis || (both && similar && fwd_vis);
            r.priority = bwd.priority;
            r.error = bwd_e;
            r.depth = bi.depth;
            r.color = second_chance(alpha, frame_idx, sidx, st, bi, s, bwd, fwd, eps2);
        }
    } else if (always_valid) {
        r.valid = true;
        const C4 cf = second_chance(1.0f - alpha, frame_idx, sidx, st, fi, s, fwd, bwd, eps2);
        const C4 cb = second_chance(alpha, frame_idx, sidx, st, bi, s, bwd, fwd, eps2);
        r.color = lerp_to_u8(cf, cb, 0.5f);
    }
    return r;
}

// bidirectionalReprojection_kernel<SW, SI, MAV>, :400-465, one pixel
void resolve_pixel(U2 idx, U2 res, float alpha, const Settings &s, const View *fwd, const View *bwd,
                   int n_aux, int SW, int SI, bool mav, C4 *out) {
    const F2 pcs = pixel2cs(idx, res);
    const Result main_r = bidir(idx, idx, res, alpha, pcs, s, fwd[0], bwd[0], SW, SI, mav,
                                EPS_1_FULL_RES, EPS_2_FULL_RES);
    if (main_r.valid) out[to1d(idx, res)] = main_r.color;
    if (n_aux <= 0) return;

    Result best;
    for (int i = 0; i < n_aux; i++) {
        const U2 ares = fwd[1 + i].res;
        const U2 aidx{idx.x / (res.x / ares.x), idx.y / (res.y / ares.y)};
        const F2 apcs = pixel2cs(aidx, ares);
        const Result ar = bidir(aidx, idx, res, alpha, apcs, s, fwd[1 + i], bwd[1 + i], SW, SI, false,
                                EPS_1_LOW_RES, EPS_2_LOW_RES);
        if (!best.valid) {
            best = ar;
        } else if (ar.valid) {
            const bool similar = std::fabs(depth_to_z(best.depth) - depth_to_z(ar.depth)) < EPS_2_LOW_RES;
            const bool choose_new = (!similar && ar.depth < best.depth) ||
                                    (similar && ar.priority >= best.priority && ar.error < best.error);
            if (choose_new) best = ar;
        }
    }
    if (best.valid) {
        bool choose_aux = !main_r.valid;
        if (main_r.valid && !best.visible_in_main) {
            const bool similar = std::fabs(depth_to_z(best.depth) - depth_to_z(main_r.depth)) < EPS_2_LOW_RES;
            choose_aux = !similar && best.depth < main_r.depth;
        }
        if (choose_aux) out[to1d(idx, res)] = best.color;
    }
}

// ---------------------------------------------------------------------------
// Timewarp (mode 4): timewarp_kernel, src/renderer.cu:467-522
// ---------------------------------------------------------------------------
struct V3 { float x, y, z; };
inline float dot3(V3 a, V3 b) {  // glm dot(vec3): (ax*bx + ay*by) + az*bz
    const float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
    return (tx + ty) + tz;
}
inline V3 normalize3(V3 v) {  // glm normalize: v * (1 / sqrt(dot(v, v)))
    const float s = 1.0f / std::sqrt(dot3(v, v));
    return V3{v.x * s, v.y * s, v.z * s};
}
inline F4 normalize4(F4 v) {  // dot(vec4): (x*x + y*y) + (z*z + w*w)
    const float tx = v.x * v.x, ty = v.y * v.y, tz = v.z * v.z, tw = v.w * v.w;
    const float d = (tx + ty) + (tz + tw);
    const float s = 1.0f / std::sqrt(d);
    return F4{v.x * s, v.y * s, v.z * s, v.w * s};
}

struct TwInfo {  // TimewarpInfo, :37-47
    bool isset = false;
    const float *inv_vp_from = nullptr;
    const float *inv_vp_to = nullptr;
    const float *vp_to = nullptr;
    float depth_to = 0.f;
    U2 res_to{0, 0};
    const C4 *img_to = nullptr;
};

inline I2 clamp2(I2 v, I2 lo, I2 hi) {
    return I2{std::max(lo.x, std::min(v.x, hi.x)), std::max(lo.y, std::min(v.y, hi.y))};
}

void timewarp_pixel(U2 idx, U2 res, const TwInfo &tw, const TwInfo &aux, C4 *out) {
    const F2 pcs = pixel2cs(idx, res);
    const F4 near_ws = cs2ws(F4{pcs.x, pcs.y, 0.0f, 1.0f}, tw.inv_vp_from);
    const F4 far_ws = cs2ws(F4{pcs.x, pcs.y, 1.0f, 1.0f}, tw.inv_vp_from);
    const V3 ro{near_ws.x, near_ws.y, near_ws.z};
    const V3 rd = normalize3(V3{far_ws.x - near_ws.x, far_ws.y - near_ws.y, far_ws.z - near_ws.z});

    const F4 center_ws = cs2ws(F4{0.0f, 0.0f, tw.depth_to, 1.0f}, tw.inv_vp_to);
    const F4 campos{tw.inv_vp_to[12], tw.inv_vp_to[13], tw.inv_vp_to[14], tw.inv_vp_to[15]};
    const F4 vd4 = normalize4(F4{center_ws.x - campos.x, center_ws.y - campos.y, center_ws.z - campos.z,
                                 center_ws.w - campos.w});
    const V3 n{vd4.x, vd4.y, vd4.z};
    const V3 pp{campos.x, campos.y, campos.z};

    const float denom = dot3(rd, n);
    if (std::fabs(denom) < 1e-6f) return;
    const float t = dot3(V3{pp.x - ro.x, pp.y - ro.y, pp.z - ro.z}, n) / denom;
    const F4 inter{ro.x + t * rd.x, ro.y + t * rd.y, ro.z + t * rd.z, 1.0f};

    const F4 pc = ws2cs(inter, tw.vp_to);
    const I2 pix = cs2pixel(F2{pc.x, pc.y}, tw.res_to);
    C4 c;
    if (in_range(pix, tw.res_to) || !aux.isset) {
        const I2 cl = clamp2(pix, I2{0, 0}, I2{(int32_t)tw.res_to.x - 1, (int32_t)tw.res_to.y - 1});
        c = fetch(tw.img_to, tw.res_to, U2{(uint32_t)cl.x, (uint32_t)cl.y});
    } else {
        const F4 pa = ws2cs(inter, aux.vp_to);
        const I2 pixa = cs2pixel(F2{pa.x, pa.y}, aux.res_to);
        const I2 cl = clamp2(pixa, I2{0, 0}, I2{(int32_t)aux.res_to.x - 1, (int32_t)aux.res_to.y - 1});
        c = fetch(aux.img_to, aux.res_to, U2{(uint32_t)cl.x, (uint32_t)cl.y});
    }
    out[to1d(idx, res)] = c;
}

float mean_z(const float *z, uint32_t w, uint32_t h) {
    std::vector<double> rows(h);
    for (uint32_t y = 0; y < h; y++) {
        double lane[64];
        for (int l = 0; l < 64; l++) {
            double acc = 0.0;
            for (uint32_t x = (uint32_t)l; x < w; x += 64) acc += (double)z[(size_t)y * w + x];
            lane[l] = acc;
        }
        for (int off = 32; off >= 1; off >>= 1)
            for (int l = 0; l < off; l++) lane[l] += lane[l + off];
        rows[y] = lane[0];
    }
    double lane[256];
    for (int l = 0; l < 256; l++) {
        double acc = 0.0;
        for (uint32_t y = (uint32_t)l; y < h; y += 256) acc += rows[y];
        lane[l] = acc;
    }
    for (int off = 128; off >= 1; off >>= 1)
        for (int l = 0; l < off; l++) lane[l] += lane[l + off];
    return (float)(lane[0] / ((double)w * (double)h));
}

bool has_motion(const isti_view_desc &d) {
    return d.motion_x != nullptr && d.motion_y != nullptr && d.motion_z != nullptr;
}

}  // namespace

// ===========================================================================
// C entry points
// ===========================================================================
extern "C" int oracle_render(const isti_view_desc *fwd, const isti_view_desc *bwd, int n_aux,
                             const isti_render_info *info, uint8_t *out_rgba8,
                             oracle_scratch *scratch_fwd, oracle_scratch *scratch_bwd, int n_threads) {
    if (!fwd || !bwd || !info || !out_rgba8 || n_aux < 0 || n_aux > ISTI_MAX_AUX) return ISTI_ERR_INVALID_ARG;
    const U2 res{info->width, info->height};
    if (res.x == 0 || res.y == 0 || res.x > ISTI_MAX_DIM || res.y > ISTI_MAX_DIM) return ISTI_ERR_UNSUPPORTED_SHAPE;
    if (info->mode < 0 || info->mode > 4) return ISTI_ERR_INVALID_ARG;
#ifdef _OPENMP
    const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#else
    const int nt = 1;
    (void)n_threads;
#endif
    C4 *out = reinterpret_cast<C4 *>(out_rgba8);

    if (info->mode == ISTI_MODE_TIMEWARP) {
        // src/renderer.cu:615-629
        const isti_view_desc &f = fwd[0];
        TwInfo tw;
        tw.isset = true;
        tw.inv_vp_from = info->inv_viewproj;
        tw.inv_vp_to = f.inv_viewproj;
        tw.vp_to = f.viewproj;
        tw.depth_to = mean_z(f.z, f.width, f.height);
        tw.res_to = U2{f.width, f.height};
        tw.img_to = reinterpret_cast<const C4 *>(f.color);
        TwInfo ta;
        if (n_aux > 0) {
            const isti_view_desc &a = fwd[1];
            ta.isset = true;
            ta.inv_vp_from = info->inv_viewproj;
            ta.inv_vp_to = a.inv_viewproj;
            ta.vp_to = a.viewproj;
            ta.depth_to = tw.depth_to;
            ta.res_to = U2{a.width, a.height};
            ta.img_to = reinterpret_cast<const C4 *>(a.color);
        }
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t y = 0; y < (int64_t)res.y; y++)
            for (uint32_t x = 0; x < res.x; x++) timewarp_pixel(U2{x, (uint32_t)y}, res, tw, ta, out);
        return ISTI_OK;
    }

    // Preconditions the reference leaves implicit (renderer.cu:145,153,433).
    if (fwd[0].width != res.x || fwd[0].height != res.y || bwd[0].width != res.x || bwd[0].height != res.y)
        return ISTI_ERR_UNSUPPORTED_SHAPE;
    if (info->shadow_frames_per_anchor < 0 || info->shadow_frames_per_anchor > ISTI_MAX_SHADOW_FRAMES)
        return ISTI_ERR_INVALID_ARG;
    for (int i = 1; i <= n_aux; i++) {
        if (fwd[i].width == 0 || fwd[i].height == 0 || fwd[i].width != bwd[i].width || fwd[i].height != bwd[i].height)
            return ISTI_ERR_UNSUPPORTED_SHAPE;
        if (fwd[i].width > res.x || fwd[i].height > res.y || res.x % fwd[i].width != 0 || res.y % fwd[i].height != 0)
            return ISTI_ERR_UNSUPPORTED_SHAPE;
    }

    const int nv = 1 + n_aux;
    // scratch (caller-provided or internal)
    std::vector<std::vector<uint64_t>> own_splat(2 * nv);
    std::vector<std::vector<F4>> own_cam(2 * nv), own_full(2 * nv);
    std::vector<uint64_t *> splat(2 * nv);
    std::vector<F4 *> cam(2 * nv), full(2 * nv);
    for (int s = 0; s < 2; s++) {
        const isti_view_desc *views = s == 0 ? fwd : bwd;
        oracle_scratch *sc = s == 0 ? scratch_fwd : scratch_bwd;
        for (int v = 0; v < nv; v++) {
            const size_t n = (size_t)views[v].width * views[v].height;
            const int k = s * nv + v;
            if (sc && sc[v].splat) splat[k] = sc[v].splat; else { own_splat[k].resize(n); splat[k] = own_splat[k].data(); }
            if (sc && sc[v].cam_ff) cam[k] = reinterpret_cast<F4 *>(sc[v].cam_ff); else { own_cam[k].resize(n); cam[k] = own_cam[k].data(); }
            if (sc && sc[v].full_ff) full[k] = reinterpret_cast<F4 *>(sc[v].full_ff); else { own_full[k].resize(n); full[k] = own_full[k].data(); }
            std::fill(splat[k], splat[k] + n, SPLAT_EMPTY);  // src/renderer.cu:631-638
        }
    }

    // src/renderer.cu:640-649: K1 for fwd, bwd, aux curr, aux next
    const bool init_linear = info->mode == ISTI_MODE_BACKWARD_BIDIRECTIONAL;
    std::vector<SplatJob> jobs;
    for (int v = 0; v < nv; v++) {
        for (int s = 0; s < 2; s++) {
            const isti_view_desc &d = s == 0 ? fwd[v] : bwd[v];
            const isti_view_desc &other = s == 0 ? bwd[v] : fwd[v];
            const isti_view_desc &main_from = s == 0 ? fwd[0] : bwd[0];
            SplatJob j;
            j.res = U2{d.width, d.height};
            j.weight = s == 0 ? info->alpha : 1.0f - info->alpha;
            j.init_linear = init_linear;
            j.z = d.z;
            const bool hm = has_motion(d);
            j.mx = hm ? d.motion_x : nullptr;
            j.my = hm ? d.motion_y : nullptr;
            j.mz = hm ? d.motion_z : nullptr;
            j.inv_vp_from = d.inv_viewproj;
            j.vp_to = other.viewproj;
            j.vp_cam = info->viewproj;
            j.vp_from_main = main_from.viewproj;
            const int k = s * nv + v;
            j.full_ff = full[k];
            j.cam_ff = cam[k];
            j.splat = splat[k];
            jobs.push_back(j);
        }
    }
    for (const SplatJob &j : jobs) {
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t y = 0; y < (int64_t)j.res.y; y++)
            for (uint32_t x = 0; x < j.res.x; x++) splat_pixel(j, U2{x, (uint32_t)y});
    }

    // src/renderer.cu:652-693: K2
    std::vector<View> vf(nv), vb(nv);
    for (int s = 0; s < 2; s++) {
        for (int v = 0; v < nv; v++) {
            const isti_view_desc &d = s == 0 ? fwd[v] : bwd[v];
            View &w = s == 0 ? vf[v] : vb[v];
            const int k = s * nv + v;
            w.res = U2{d.width, d.height};
            w.forward = s == 0;
            w.priority = v == 0 ? 0 : d.priority;
            w.img = reinterpret_cast<const C4 *>(d.color);
            w.z = d.z;
            w.shadow_a = reinterpret_cast<const C4 *>(d.shadow_color_a);
            w.shadow_b = reinterpret_cast<const C4 *>(d.shadow_color_b);
            w.shadow_info = reinterpret_cast<const C4 *>(d.shadow_info);
            w.cam_ff = cam[k];
            w.full_ff = full[k];
            w.splat = splat[k];
        }
    }
    int SW = 1, SI = 3;
    bool mav = false;
    switch (info->mode) {
        case ISTI_MODE_OURS:
        case ISTI_MODE_BACKWARD_UNIDIRECTIONAL: SW = 1; SI = 3; mav = false; break;
        case ISTI_MODE_SPLATTING: SW = 1; SI = 0; mav = false; break;
        case ISTI_MODE_BACKWARD_BIDIRECTIONAL: SW = 3; SI = 5; mav = true; break;
        default: return ISTI_ERR_INVALID_ARG;
    }
    if (info->max_search_iterations >= 0) SI = info->max_search_iterations;
    const Settings st{info->frames_per_anchor, info->shadow_frames_per_anchor, info->no_shadows != 0};
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (int64_t y = 0; y < (int64_t)res.y; y++)
        for (uint32_t x = 0; x < res.x; x++)
            resolve_pixel(U2{x, (uint32_t)y}, res, info->alpha, st, vf.data(), vb.data(), n_aux, SW, SI, mav, out);
    return ISTI_OK;
}

extern "C" void oracle_pixel2cs(uint32_t x, uint32_t y, uint32_t w, uint32_t h, float out[2]) {
    const F2 r = pixel2cs(U2{x, y}, U2{w, h});
    out[0] = r.x;
    out[1] = r.y;
}
extern "C" void oracle_cs2pixel(float cx, float cy, uint32_t w, uint32_t h, int32_t out[2]) {
    const I2 r = cs2pixel(F2{cx, cy}, U2{w, h});
    out[0] = r.x;
    out[1] = r.y;
}
extern "C" void oracle_mat_vec(const float m[16], const float v[4], float out[4]) {
    const F4 r = mat_vec(m, F4{v[0], v[1], v[2], v[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
extern "C" void oracle_cs2ws(const float v[4], const float inv_vp[16], float out[4]) {
    const F4 r = cs2ws(F4{v[0], v[1], v[2], v[3]}, inv_vp);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
extern "C" void oracle_ws2cs(const float v[4], const float vp[16], float out[4]) {
    const F4 r = ws2cs(F4{v[0], v[1], v[2], v[3]}, vp);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
extern "C" uint64_t oracle_encode_key(float depth, uint32_t x, uint32_t y) { return encode_key(depth, U2{x, y}); }
extern "C" void oracle_decode_key(uint64_t key, float *depth, uint32_t *x, uint32_t *y) {
    const SplatInfo s = decode_key(key);
    *depth = s.depth;
    *x = s.src.x;
    *y = s.src.y;
}
extern "C" float oracle_reproj_error(uint32_t qx, uint32_t qy, uint32_t w, uint32_t h, float pcs_x, float pcs_y,
                                     float mx, float my) {
    return reproj_error(U2{qx, qy}, U2{w, h}, F2{pcs_x, pcs_y}, F4{mx, my, 0.f, 0.f});
}
extern "C" void oracle_lerp_color(const uint8_t a[4], const uint8_t b[4], float t, uint8_t out[4]) {
    const C4 r = lerp_to_u8(C4{a[0], a[1], a[2], a[3]}, C4{b[0], b[1], b[2], b[3]}, t);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
extern "C" int32_t oracle_to_8bit(float f) { return to_8bit(f); }
extern "C" uint32_t oracle_float2uint_rn(float f) { return f2u_rn(f); }
extern "C" uint32_t oracle_float2uint_rd(float f) { return f2u_rd(f); }
extern "C" uint32_t oracle_float2uint_ru(float f) { return f2u_ru(f); }
extern "C" void oracle_init_search(const uint64_t *splat, uint32_t w, uint32_t h, uint32_t x, uint32_t y, int window,
                                   float *depth, uint32_t *sx, uint32_t *sy) {
    const SplatInfo s = init_search(U2{x, y}, U2{w, h}, splat, window);
    *depth = s.depth;
    *sx = s.src.x;
    *sy = s.src.y;
}
extern "C" void oracle_search(const float *cam_ff, uint32_t low_w, uint32_t low_h, uint32_t start_x, uint32_t start_y,
                              uint32_t px, uint32_t py, uint32_t full_w, uint32_t full_h, int iterations,
                              uint32_t *qx, uint32_t *qy) {
    const U2 r = search(U2{start_x, start_y}, U2{px, py}, U2{full_w, full_h}, U2{low_w, low_h},
                        reinterpret_cast<const F4 *>(cam_ff), iterations);
    *qx = r.x;
    *qy = r.y;
}
/* Whole-render main-camera searches (checks the learned starts): for every
 * target pixel p of a w x h render whose keyframe view has the render
 * resolution, the search (src/renderer.cu:177-202) from start[p] (packed
 * y << 16 | x; 0xffffffff = no search) and its reprojection error
 * (src/renderer.cu:204-215; FLT_MAX when no search ran). */
extern "C" void oracle_search_field(const float *cam_ff, uint32_t w, uint32_t h, const uint32_t *start, int iterations,
                                    uint32_t *q, float *err) {
    const F4 *ff = reinterpret_cast<const F4 *>(cam_ff);
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const size_t i = (size_t)y * w + x;
            if (start[i] == 0xffffffffu) {
                q[i] = 0xffffffffu;
                err[i] = FLT_MAX;
                continue;
            }
            const U2 r = search(U2{start[i] & 0xFFFFu, start[i] >> 16}, U2{x, y}, U2{w, h}, U2{w, h}, ff, iterations);
            err[i] = reproj_error(r, U2{w, h}, pixel2cs(U2{x, y}, U2{w, h}), ff[(size_t)r.y * w + r.x]);
            q[i] = r.y << 16 | r.x;
        }
}
/* The window starts (src/renderer.cu:148-175) of every pixel of a w x h
 * splat buffer, packed as above (no candidate, depth <= 0: 0xffffffff), and,
 * if depth is non-null, the candidate's decoded depth. */
extern "C" void oracle_window_starts(const uint64_t *splat, uint32_t w, uint32_t h, int window, uint32_t *start,
                                     float *depth) {
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const SplatInfo s = init_search(U2{x, y}, U2{w, h}, splat, window);
            start[(size_t)y * w + x] = s.depth > 0.0f ? (s.src.y << 16 | s.src.x) : 0xffffffffu;
            if (depth) depth[(size_t)y * w + x] = s.depth;
        }
}
extern "C" void oracle_color_from_shadow(const isti_view_desc *d, int forward, uint32_t idx_x, uint32_t idx_y,
                                         float t, int sfpa, int no_shadows, uint32_t px, uint32_t py,
                                         uint8_t out[4]) {
    View v;
    v.res = U2{d->width, d->height};
    v.forward = forward != 0;
    v.img = reinterpret_cast<const C4 *>(d->color);
    v.z = d->z;
    v.shadow_a = reinterpret_cast<const C4 *>(d->shadow_color_a);
    v.shadow_b = reinterpret_cast<const C4 *>(d->shadow_color_b);
    v.shadow_info = reinterpret_cast<const C4 *>(d->shadow_info);
    const Settings s{0, sfpa, no_shadows != 0};
    const C4 r = color_from_shadow(U2{idx_x, idx_y}, t, s, v, U2{px, py});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
extern "C" float oracle_mean_z(const float *z, uint32_t w, uint32_t h) { return mean_z(z, w, h); }
