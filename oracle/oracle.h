/*
 * oracle.h - CPU restatement of the reference's per-frame interpolation path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP renderer:
 * only the test suite, the smoke check and the benchmark's CPU-baseline leg
 * may load it. The product path never links or calls it.
 *
 * It restates, scalar and literally, src/renderer.cu:16-465 and :524-694
 * (plus include/util/helper_common.h:8-95, helper_math.h operators and
 * src/util/image_buffer.cu:27-36,80-101) of DerThomy/
 * ImageBasedSpatioTemporalInterpolation. Built with `g++ -O2
 * -ffp-contract=off` (no -ffast-math) so every float operation is a single
 * IEEE-754 binary32 operation in the order the reference writes it; the
 * 4x4 matrix-vector product uses the association order
 * (m0*v.x + m1*v.y) + (m2*v.z + m3*v.w) of glm's mat4*vec4 (glm is an
 * un-vendored submodule of the reference; its pinned commit is not
 * recoverable, so this order is the one the oracle and the HIP kernels share).
 *
 * Where the reference is undefined in C++ (float -> int conversion of NaN or
 * out-of-range values) the oracle uses the GPU conversion semantics the
 * reference ran with: NaN -> 0, saturation at the integer range.
 */
#ifndef ISTI_ORACLE_H
#define ISTI_ORACLE_H

#include <stdint.h>
#include "../include/isti.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host scratch of one (slot, view): u64 splat buffer (W*H), cam_ff and
 * full_ff float4 planes (W*H*4 floats). Any pointer may be NULL, in which
 * case the oracle allocates it internally for the call. */
typedef struct oracle_scratch {
    uint64_t *splat;
    float *cam_ff;
    float *full_ff;
} oracle_scratch;

/* Whole-frame render: Renderer::render (src/renderer.cu:607-694).
 * fwd[0]/bwd[0] are the main view, fwd[1..n_aux]/bwd[1..n_aux] the aux
 * cameras; the descriptors carry HOST pointers. out_rgba8 (W*H*4, packed)
 * is read-modify-write: only valid pixels are written. n_threads <= 0 uses
 * all OpenMP threads. Returns 0 on success, an isti_status otherwise. */
int oracle_render(const isti_view_desc *fwd, const isti_view_desc *bwd, int n_aux,
                  const isti_render_info *info, uint8_t *out_rgba8,
                  oracle_scratch *scratch_fwd, oracle_scratch *scratch_bwd,
                  int n_threads);

/* --- unit-level entry points (known-answer tests) --- */
void oracle_pixel2cs(uint32_t x, uint32_t y, uint32_t w, uint32_t h, float out[2]);
void oracle_cs2pixel(float cx, float cy, uint32_t w, uint32_t h, int32_t out[2]);
void oracle_mat_vec(const float m[16], const float v[4], float out[4]);
void oracle_cs2ws(const float v[4], const float inv_vp[16], float out[4]);
void oracle_ws2cs(const float v[4], const float vp[16], float out[4]);
uint64_t oracle_encode_key(float depth, uint32_t x, uint32_t y);
void oracle_decode_key(uint64_t key, float *depth, uint32_t *x, uint32_t *y);
float oracle_reproj_error(uint32_t qx, uint32_t qy, uint32_t w, uint32_t h,
                          float pcs_x, float pcs_y, float mx, float my);
void oracle_lerp_color(const uint8_t a[4], const uint8_t b[4], float t, uint8_t out[4]);
int32_t oracle_to_8bit(float f);
uint32_t oracle_float2uint_rn(float f);
uint32_t oracle_float2uint_rd(float f);
uint32_t oracle_float2uint_ru(float f);
/* Splat-window initialisation (src/renderer.cu:148-175) at pixel (x, y). */
void oracle_init_search(const uint64_t *splat, uint32_t w, uint32_t h, uint32_t x, uint32_t y,
                        int window, float *depth, uint32_t *sx, uint32_t *sy);
/* Fixed-point search (src/renderer.cu:177-202); cam_ff is W*H float4. */
void oracle_search(const float *cam_ff, uint32_t low_w, uint32_t low_h, uint32_t start_x,
                   uint32_t start_y, uint32_t px, uint32_t py, uint32_t full_w, uint32_t full_h,
                   int iterations, uint32_t *qx, uint32_t *qy);
/* Every main-camera search of a w x h render from packed starts
 * (y << 16 | x, 0xffffffff = none) with its reprojection error. */
void oracle_search_field(const float *cam_ff, uint32_t w, uint32_t h, const uint32_t *start, int iterations,
                         uint32_t *q, float *err);
/* Window starts of every pixel of a splat buffer, packed as above, and
 * (if depth is non-null) the candidates' decoded depths. */
void oracle_window_starts(const uint64_t *splat, uint32_t w, uint32_t h, int window, uint32_t *start,
                          float *depth);
/* colorFromShadow(idcs, t, ...) (src/renderer.cu:270-295) for one pixel of
 * a view whose planes are host pointers. */
void oracle_color_from_shadow(const isti_view_desc *view, int forward, uint32_t idx_x,
                              uint32_t idx_y, float t, int sfpa, int no_shadows,
                              uint32_t px, uint32_t py, uint8_t out[4]);
/* Deterministic mean of a z plane: rows summed by 64 strided lanes in
 * double, lanes folded by halving, rows folded the same way over 256
 * lanes; replaces the reference's float atomicAdd sum
 * (src/util/image_buffer.cu:27-36, 80-101). */
float oracle_mean_z(const float *z, uint32_t w, uint32_t h);

#ifdef __cplusplus
}
#endif

#endif
