/*
 * isti_learned.h - learned initialiser for the fixed-point flow search.
 *
 * No counterpart in the reference: there, the search starts from the
 * nearest splat candidate in a small window (src/renderer.cu:148-175). Here
 * a small U-Net predicts, per target pixel and keyframe direction, where the
 * search should start, plus a confidence. The plain path stays the judge:
 * a learned start is only kept when its fixed-point search ends with a
 * reprojection error below the reference's EPS_1 (src/renderer.cu:366);
 * otherwise the pixel falls back to the window initialisation.
 *
 * Network: 4 levels with 16-32-64-128 channels, two 3x3 convolutions + ReLU
 * per level, 2x2 average pooling down, nearest-neighbour upsampling and skip
 * concatenation up, a 1x1 head with 3 outputs (flow offset x/y in target
 * pixels, confidence logit). Activations are NHWC bf16 with 16-channel
 * granularity (the 15 input features are padded to 16; the head's 3 outputs
 * are channels 0..2 of a 16-channel fp32 output), arithmetic on MFMA with
 * fp32 accumulation, fp32 master weights with Adam.
 *
 * Input features (15, at net resolution, sampled at the top-left full-
 * resolution pixel of each cell), for direction d with "own" keyframe d and
 * "other" keyframe 1-d: own colour RGB/255, other colour RGB/255, own z,
 * other z, own cam_ff x/y in pixels and z, other cam_ff x/y in pixels and z,
 * the direction's time (alpha for fwd, 1-alpha for bwd).
 */
#ifndef ISTI_LEARNED_H
#define ISTI_LEARNED_H

#include <stddef.h>
#include <stdint.h>

#include "isti.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ISTI_UNET_CH 16 /* channel count of the input and output tensors */

typedef struct isti_unet isti_unet;

/* A network for `batch` images of width x height (multiples of 8), with
 * He-initialised weights from `seed`. Allocates activations and gradients. */
isti_status isti_unet_create(int32_t device, uint32_t width, uint32_t height, int32_t batch, uint64_t seed,
                             isti_unet **out);
isti_status isti_unet_destroy(isti_unet *net);

/* Number of fp32 parameters (weights [cout][ky][kx][cin] then bias, per layer). */
size_t isti_unet_param_count(void);
isti_status isti_unet_get_params(isti_unet *net, float *host);
isti_status isti_unet_set_params(isti_unet *net, const float *host);
isti_status isti_unet_get_grads(isti_unet *net, float *host);

/* x: device bf16 [batch][height][width][16]; out: device fp32
 * [batch][height][width][16] (channels 0..2 meaningful). Keeps the
 * activations for isti_unet_backward. */
isti_status isti_unet_forward(isti_unet *net, const uint16_t *x, float *out, void *stream);

/* Times every wave shape of each forward 3x3 convolution on this network's
 * size (reps launches each, on `stream`, synchronising it) and keeps the
 * fastest for later forward passes. Every shape computes the same bits; the
 * activation buffers hold garbage afterwards (run a forward pass before a
 * backward pass). */
isti_status isti_unet_autotune(isti_unet *net, int32_t reps, void *stream);

/* Accumulates parameter gradients for dout (device fp32, same shape as out)
 * of the last forward pass. */
isti_status isti_unet_backward(isti_unet *net, const float *dout, void *stream);
isti_status isti_unet_zero_grads(isti_unet *net, void *stream);
/* One Adam step (bias-corrected) on the accumulated gradients. */
isti_status isti_unet_adam(isti_unet *net, float lr, float beta1, float beta2, float eps, void *stream);

/* --- integration with the renderer --- */

/* When q_out (device, 2 * W * H uint32 for the render resolution) is set,
 * every render of r writes the main camera's search results per pixel:
 * q_out[2*(y*W+x) + d] = (qy << 16 | qx) of direction d (0 fwd, 1 bwd) if
 * its reprojection error is below EPS_1, else 0xffffffff. NULL disables. */
isti_status isti_set_search_capture(isti_renderer *r, uint32_t *q_out);

/* First half of a learned-start render: runs the frame's splat stage (M1
 * clear and K1 reprojection + splat of every view, as isti_render does) and
 * builds the network input of both directions (batch 2: fwd, bwd) at
 * net_w x net_h (render size / 2^k for k in 0..3, each rounded up to a
 * multiple of 8) from the main keyframes and the fresh cam_ff planes.
 * x: device bf16 [2][net_h][net_w][16]. Displacement features are in
 * network cells (2^k render pixels). x NULL runs the splat stage alone: the
 * frame is prepared for isti_render_learned with a network output computed
 * for an earlier frame (the network runs once per keyframe pair, on its
 * first frame, and later frames of the pair reuse its output; the learned
 * start stays relative to each frame's own one-step start, and the plain
 * search still judges it). */
isti_status isti_learned_features(isti_renderer *r, const isti_render_info *info, uint16_t *x, uint32_t net_w,
                                  uint32_t net_h, void *stream);

/* Training loss of a network output for the frame prepared by the last
 * isti_learned_features call, against its captured plain-search results
 * (q_target as written by isti_set_search_capture): Huber (delta 1 cell) on
 * the start residual of pixels with a valid target (the residual that puts
 * the learned start, see isti_render_learned, on the centre of the found
 * pixel), plus conf_weight times the binary cross-entropy of the confidence
 * logit against "the target is valid and the predicted residual is within
 * one cell of it" (the network learns where its own start is right). Each
 * cell is scored at its top-left render pixel. Writes dout (same shape as
 * out) and adds the mean loss to *loss_acc (device float). */
isti_status isti_learned_loss(isti_renderer *r, const float *out, const uint32_t *q_target, uint32_t net_w,
                              uint32_t net_h, float conf_weight, float *dout, float *loss_acc, void *stream);

/* Second half: the resolve of the frame prepared by the last
 * isti_learned_features call (same info, keyframes unchanged since), with
 * learned starts for the main camera. Where the network's confidence is
 * >= conf_threshold, the fwd/bwd search of pixel (x, y) starts at
 * floor((x + 0.5 - m.x * W/2) + o.x * 2^k), likewise in y, clamped to the
 * view: one analytic search step from the pixel itself (m = the direction's
 * cam_ff at the pixel, in NDC) corrected by the network's residual o
 * (channels 0/1, in cells of 2^k pixels). A learned start's search runs at
 * most 4 steps (fewer if the frame's cap is lower). A learned result is kept
 * when that search ends with an error below EPS_1 on a surface not behind the window
 * start's candidate (depth_to_z difference < EPS_2, the splat's depth test);
 * a rejected learned start, and every low-confidence pixel, uses the
 * reference's window start instead. conf_threshold 1 disables the learned starts (the
 * plain resolve, with statistics). net_out: fp32 [2][net_h][net_w][16]. If
 * stats is non-null it receives (device, uint32[5], accumulated): learned
 * starts kept, learned starts rejected (fallback), low-confidence searches,
 * total searches, main-camera search steps taken. */
isti_status isti_render_learned(isti_renderer *r, const isti_render_info *info, const float *net_out,
                                uint32_t net_w, uint32_t net_h, float conf_threshold, uint32_t *stats,
                                uint8_t *out_rgba8, size_t out_pitch_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif
