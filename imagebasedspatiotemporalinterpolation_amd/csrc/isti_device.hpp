// isti_device.hpp - gfx950 device math for the interpolation path.
//
// Every helper states the reference lines whose semantics it implements.
// Numerics contract (shared with the CPU oracle, checked bit-for-bit by the
// parity tests): compiled with -ffp-contract=off and IEEE-correct fp32
// division and square root, one binary32 operation per operator in the
// order the reference writes it; 4x4 matrix-vector products associate as
// (m0*v.x + m1*v.y) + (m2*v.z + m3*v.w) (glm's mat4*vec4 order).
// float->int conversions follow the GPU semantics the reference ran with:
// truncation toward zero, NaN -> 0, saturation at the integer range.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>

namespace isti {

struct Mat4 {
    float m[16];  // column-major, (row r, col c) at m[c*4+r]
};

// src/renderer.cu:16-22, 27-28
// Search steps granted to a learned start (isti_render_learned).
constexpr int LEARNED_STEPS = 4;
constexpr float EPS_1_FULL_RES = 0.002f;
constexpr float EPS_1_LOW_RES = EPS_1_FULL_RES * 4.0f;
constexpr float EPS_2_FULL_RES = 0.001f;
constexpr float EPS_2_LOW_RES = 0.004f;
constexpr float EPS_3 = 0.01f;
// The smallest binary32 a with RN32(sqrt(a)) >= eps, for each threshold eps
// above: RN32 o sqrt is non-decreasing, so RN32(sqrt(a)) < eps exactly when
// a < that value, and an error or distance compares with eps without its
// square root (the values are checked against IEEE sqrtf on the CPU).
constexpr float EPS_1_FULL_RES_SQ = 0x1.0c6f7ap-18f;  // 0x368637bd
constexpr float EPS_1_LOW_RES_SQ = 0x1.0c6f7ap-14f;   // 0x388637bd
constexpr float EPS_3_SQ = 0x1.a36e2cp-14f;           // 0x38d1b716
constexpr unsigned long long SPLAT_EMPTY = 0xfefefefefefefefeull;  // memset(0xfe), renderer.cu:22
constexpr uint32_t SOFTSHADOW_BITS = 2;
constexpr uint32_t SOFTSHADOW_VAL_MASK = (1u << SOFTSHADOW_BITS) - 1u;

// float -> int32 with truncation, NaN -> 0, saturation: exactly what
// v_cvt_i32_f32 does, written as the instruction so the compiler cannot
// treat out-of-range values as undefined.
__device__ __forceinline__ int32_t cvt_s32(float f) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
}
// float -> uint32 with truncation, NaN and negatives -> 0, saturation (v_cvt_u32_f32)
__device__ __forceinline__ uint32_t sat_u32(float f) {
    uint32_t r;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
}
// (unsigned char)(float) of make_uchar4(float4), helper_math_ext.h:19-22
__device__ __forceinline__ uint8_t cvt_u8(float f) {
    const uint32_t u = sat_u32(f);
    return (uint8_t)(u < 255u ? u : 255u);
}
// __float2uint_rn / _rd / _ru, renderer.cu:359-362
__device__ __forceinline__ uint32_t f2u_rn(float f) { return sat_u32(__builtin_rintf(f)); }
__device__ __forceinline__ uint32_t f2u_rd(float f) { return sat_u32(floorf(f)); }
__device__ __forceinline__ uint32_t f2u_ru(float f) { return sat_u32(ceilf(f)); }

// helper_common.h:8-11
__device__ __forceinline__ bool in_range(int2 p, uint32_t w, uint32_t h) {
    return p.x >= 0 && (uint32_t)p.x < w && p.y >= 0 && (uint32_t)p.y < h;
}
// helper_common.h:66-69
__device__ __forceinline__ float2 pixel2cs(uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    float2 r;
    r.x = 2.0f * ((float)x + 0.5f) / (float)w - 1.0f;
    r.y = 2.0f * ((float)y + 0.5f) / (float)h - 1.0f;
    return r;
}
// One coordinate of pixel2CS, RN32(RN32(2*(p+0.5)) / n) - 1, with the
// division evaluated as RN32(RN64((2p+1) * RN64(1/n))). For 2p+1 < 2^17 and
// n < 2^16 the exact quotient is never a binary32 rounding midpoint and lies
// at least 2^-41 (relative) from every midpoint, while the double product is
// within 2^-52 of it, so the result is bit-identical to the IEEE division.
// inv_n = RN64(1/n) is computed once per view on the host.
__device__ __forceinline__ float cs_coord(uint32_t p, double inv_n) {
    return (float)((double)(2u * p + 1u) * inv_n) - 1.0f;
}

// helper_common.h:71-75
__device__ __forceinline__ int2 cs2pixel(float cx, float cy, uint32_t w, uint32_t h) {
    int2 r;
    r.x = cvt_s32(floorf((cx + 1.0f) * (float)w * 0.5f));
    r.y = cvt_s32(floorf((cy + 1.0f) * (float)h * 0.5f));
    return r;
}
// in_range for coordinates that may be negative: one unsigned compare per axis.
__device__ __forceinline__ bool in_range_u(int2 p, uint32_t w, uint32_t h) {
    return (uint32_t)p.x < w && (uint32_t)p.y < h;
}
__device__ __forceinline__ float4 mat_vec(const Mat4 &M, float4 v) {
    const float *m = M.m;
    float4 r;
    r.x = (m[0] * v.x + m[4] * v.y) + (m[8] * v.z + m[12] * v.w);
    r.y = (m[1] * v.x + m[5] * v.y) + (m[9] * v.z + m[13] * v.w);
    r.z = (m[2] * v.x + m[6] * v.y) + (m[10] * v.z + m[14] * v.w);
    r.w = (m[3] * v.x + m[7] * v.y) + (m[11] * v.z + m[15] * v.w);
    return r;
}
// Class masks of v_cmp_class_f32 (bit 2 -inf, 3 -normal, 8 +normal, 9 +inf).
constexpr int CLASS_NORMAL = (1 << 3) | (1 << 8);
constexpr int CLASS_NORMAL_OR_INF = CLASS_NORMAL | (1 << 2) | (1 << 9);
constexpr int CLASS_FINITE_NONZERO = CLASS_NORMAL | (1 << 4) | (1 << 7);  // + -subnormal
// Midpoint between RN32(1 - RN32(1e-3)) = 16760439 * 2^-24 (the main-view
// visibility limit of K1, renderer.cu:76-146) and its predecessor.
constexpr double LIM_MID = 33520877.0 / 33554432.0;

// RN32(n / w) for the three numerators of a projective division, equal to
// the IEEE quotients. For a normal w, R = 1/w comes from the hardware
// reciprocal (error < 2^-22) refined by two Newton steps in binary64
// (error < 2^-52), and q = RN32(RN64(n R)) is within 2^-51 |n / w| of the
// exact quotient. In the normal range the exact quotient of two binary32
// values is never a rounding midpoint M 2^f (M odd with 25 bits: n = w M 2^f
// would need a 25-bit significand) and lies at least 2^f / B > 2^-49 |n / w|
// from every midpoint (B < 2^24 the significand of w), so q rounds the same
// way; the one subnormal midpoint that rounds up to 2^-126 is an exact
// quotient only for w a power of two, where R and q are exact. So q is the
// correctly rounded quotient whenever it is normal or infinite. A zero,
// subnormal or NaN q (exact midpoints occur there, or the reciprocal
// flushed) and a w that is not normal take the IEEE division.
__device__ __forceinline__ float3 div3_rn(float x, float y, float z, float w) {
    if (__builtin_amdgcn_classf(w, CLASS_NORMAL)) {
        const double wd = (double)w;
        double r = (double)__builtin_amdgcn_rcpf(w);
        r = __builtin_fma(r, __builtin_fma(-wd, r, 1.0), r);
        r = __builtin_fma(r, __builtin_fma(-wd, r, 1.0), r);
        const float qx = (float)((double)x * r), qy = (float)((double)y * r), qz = (float)((double)z * r);
        if (__builtin_amdgcn_classf(qx, CLASS_NORMAL_OR_INF) && __builtin_amdgcn_classf(qy, CLASS_NORMAL_OR_INF) &&
            __builtin_amdgcn_classf(qz, CLASS_NORMAL_OR_INF))
            return make_float3(qx, qy, qz);
    }
    return make_float3(x / w, y / w, z / w);
}

// helper_common.h:77-87 CS2WS / WS2CS. w/w is 1 for finite non-zero w and
// NaN otherwise, so that quotient is a select instead of a division.
__device__ __forceinline__ float4 project(const Mat4 &M, float4 v) {
    const float4 r = mat_vec(M, v);
    const float w = r.w;
    const float ww = (__builtin_isfinite(w) && w != 0.0f) ? 1.0f : __builtin_nanf("");
    const float3 q = div3_rn(r.x, r.y, r.z, w);
    return make_float4(q.x, q.y, q.z, ww);
}
// project() for a point that further mat_vec calls consume: w is returned as
// the literal 1 (so their m[12+i] * w products fold away) and a bad w (not
// finite, or zero) is carried as NaN in x instead. Every row of a later
// mat_vec holds an m * x term, so it is NaN exactly where project()'s NaN w
// would have made it NaN.
__device__ __forceinline__ float4 project_unit(const Mat4 &M, float4 v) {
    const float4 r = mat_vec(M, v);
    const float w = r.w;
    const float3 q = div3_rn(r.x, r.y, r.z, w);
    return make_float4((__builtin_isfinite(w) && w != 0.0f) ? q.x : __builtin_nanf(""), q.y, q.z, 1.0f);
}
// helper_common.h:89-95
__device__ __forceinline__ float z_to_depth(float z) { return (z + 1.0f) * 0.5f; }
__device__ __forceinline__ float depth_to_z(float d) { return d * 2.0f - 1.0f; }

// helper_common.h:22-38: lerp_color then color_to_uchar4 (to_8bit truncates).
// The division v / 255 is evaluated as RN32(RN64(v) * RN64(1/255)): with
// v = M 2^e (M < 2^24) and q = v / 255 in [2^E, 2^(E+1)), q - m for any
// binary32 rounding midpoint m is a non-zero multiple of 2^(E-24) / 255
// (v and 255 m are multiples of 2^(E-24); v / 255 = m would need 255 | M and
// then q is representable), so |q - m| >= 2^(E-32), while the double product
// is within 2^(E-51) of q: both round to the same binary32 (subnormal q
// likewise, on the fixed 2^-149 grid).
constexpr double INV_255 = 1.0 / 255.0;
__device__ __forceinline__ uint8_t lerp_ch(uint32_t ua, uint32_t ub, float t) {
    const float fa = (float)ua, fb = (float)ub;
    const float l = (float)((double)(fa + t * (fb - fa)) * INV_255);
    const int32_t v = cvt_s32(l * 255.f);
    return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
}

// floor(x / r) for x, r < 2^16 as a multiply-high by m = floor(2^32 / r) + 1
// (m unused for r = 1): with m r = 2^32 + e, 0 < e <= r, the error term
// x e / 2^32 < 1 / r never reaches the next integer.
__device__ __forceinline__ uint32_t div_small(uint32_t x, uint32_t r, uint32_t m) {
    return r == 1u ? x : __umulhi(x, m);
}
__device__ __forceinline__ uchar4 lerp_to_u8(uchar4 a, uchar4 b, float t) {
    return make_uchar4(lerp_ch(a.x, b.x, t), lerp_ch(a.y, b.y, t), lerp_ch(a.z, b.z, t), lerp_ch(a.w, b.w, t));
}

// renderer.cu:59-74 splat key: depth bits << 32 | x << 16 | y. A NaN depth
// is encoded as 0x7fffffff, the canonical NaN that NVIDIA GPUs return from
// binary32 arithmetic, so a NaN key sorts after every number and before the
// empty-cell value whatever sign and payload the NaN had here.
__device__ __forceinline__ unsigned long long encode_key(float depth, uint32_t x, uint32_t y) {
    const uint32_t bits = __builtin_isnan(depth) ? 0x7fffffffu : __float_as_uint(depth);
    return ((unsigned long long)bits << 32) | (unsigned long long)((x << 16) | y);
}
__device__ __forceinline__ float key_depth(unsigned long long k) { return __uint_as_float((uint32_t)(k >> 32)); }
__device__ __forceinline__ uint32_t key_x(unsigned long long k) { return (uint32_t)k >> 16; }
__device__ __forceinline__ uint32_t key_y(unsigned long long k) { return (uint32_t)(k & 0xFFFFull); }

// Linear index of pixel (x, y) in a plane with row length w. Views are at
// most 65535 pixels per axis, so y * w has 24-bit operands and a product
// below 2^32: the full-rate 24-bit multiply (v_mad_u32_u24) is exact where
// a 32-bit multiply would take four passes.
__device__ __forceinline__ uint32_t px_index(uint32_t w, uint32_t x, uint32_t y) { return __umul24(y, w) + x; }

// Row of linear pixel p < w h (w, h <= 65535) without an integer divide:
// RN32(RN32(p) RN32(1/w)) errs by less than 3 2^-24 p / w < 3 2^-8 (p / w <
// h < 2^16), so its truncation is the row or one off, and one remainder
// check corrects it.
__device__ __forceinline__ uint32_t div_px(uint32_t p, uint32_t w, float rcp_w) {
    uint32_t q = (uint32_t)((float)p * rcp_w);
    const int32_t r = (int32_t)(p - __umul24(q, w));
    if (r < 0) q -= 1u;
    else if (r >= (int32_t)w) q += 1u;
    return q;
}

// Element idx of a plane through a 32-bit byte offset: views hold fewer than
// 2^28 pixels (ISTI_MAX_VIEW_PIXELS), so even float4 planes stay below 4 GiB,
// and the load takes the scalar base + 32-bit vector offset form (no 64-bit
// address arithmetic, one VGPR per address).
template <typename T>
__device__ __forceinline__ const T &plane_ref(const T *plane, uint32_t idx) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(plane) + idx * (uint32_t)sizeof(T));
}

// Unchecked load at pixel (x, y) of a plane with row length w.
template <typename T>
__device__ __forceinline__ T at(const T *plane, uint32_t w, uint32_t x, uint32_t y) {
    return plane_ref(plane, px_index(w, x, y));
}

// The xy half of a float4 plane's sample (one 8-byte load).
__device__ __forceinline__ float2 at_xy(const float4 *plane, uint32_t w, uint32_t x, uint32_t y) {
    return plane_ref(reinterpret_cast<const float2 *>(plane), 2u * px_index(w, x, y));
}

// Point sample with border addressing (src/util/image_buffer.cu:227-249):
// out of range, or a missing plane, reads as zero.
template <typename T>
__device__ __forceinline__ T fetch(const T *plane, uint32_t w, uint32_t h, uint32_t x, uint32_t y) {
    if (plane == nullptr || x >= w || y >= h) return T{};
    return plane_ref(plane, px_index(w, x, y));
}

}  // namespace isti
