/*
 * isti_quality.h - image-quality measures of the reference's evaluation,
 * computed on the GPU over RGBA8 device images (RGB channels, each /255):
 *
 *   PSNR = -10 log10(MSE), MSE over all RGB samples       (scripts/eval.py:57)
 *   SSIM = mean of the SSIM map with an 11x11 Gaussian window (sigma 1.5),
 *          zero padding, per channel, C1 = 0.01^2, C2 = 0.03^2
 *                                                   (scripts/loss_utils.py:17-57)
 *
 * The reference runs these in PyTorch on decoded PNG files; here they run on
 * the images a render leaves in device memory, so a serving or training loop
 * can score frames without copying them to the host.
 */
#ifndef ISTI_QUALITY_H
#define ISTI_QUALITY_H

#include <stddef.h>
#include <stdint.h>

#include "isti.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Accumulates into acc (device, 2 doubles, cleared first by this call):
 * acc[0] = sum of squared RGB differences, acc[1] = sum of the SSIM map over
 * RGB samples; both over w x h pixels of images a and b (row pitches in
 * bytes). Stream-ordered; no host synchronisation. Then
 * PSNR = -10 log10(acc[0] / (3 w h)) and SSIM = acc[1] / (3 w h). */
isti_status isti_image_quality(const uint8_t *a, size_t pitch_a, const uint8_t *b, size_t pitch_b, uint32_t w,
                               uint32_t h, double *acc, void *stream);

#ifdef __cplusplus
}
#endif

#endif
