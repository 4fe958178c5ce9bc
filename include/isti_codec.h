/*
 * isti_codec.h - the keyframe wire codec of the reference's compression
 * round trip (compression.py), on the GPU.
 *
 * What is restated (compression.py line numbers):
 *   - the lossless ShadowInfo path (compress_raw_images_lossless :223-279,
 *     decompress_raw_images_lossless :281-325): per frame the third colour
 *     channel of the ShadowInfo image, optionally XOR-ed with the anchor
 *     frame taken every server_fps frames (:249-252, :316-319), optionally
 *     run-length coded (rle_encode :176-204, rle_decode :206-218), framed as
 *     [u16 name length][name][u32 width][u32 height][u32 length][bytes]
 *     (little endian, :262-268) inside one zstd stream. The byte kernels are
 *     here; the framing and zstd are host code (codec.py, libzstd);
 *   - the 16-bit normalisation of depth and flow images before the video
 *     encoder (normalize_exr :84-91, denormalize_exr :93-98, with
 *     load_exr's zeroing of |v| < 1e-4, :60-64).
 * Not restated: the HEVC encode/decode of colour, depth and flow (ffmpeg,
 * absent here) and the morphological clean-up of shadow images (OpenCV).
 *
 * Device pointers everywhere; stream-ordered. Calls that return a
 * data-dependent length synchronise the stream.
 */
#ifndef ISTI_CODEC_H
#define ISTI_CODEC_H

#include <stddef.h>
#include <stdint.h>

#include "isti.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct isti_codec isti_codec;

/* Codec working memory for inputs of up to max_bytes bytes / max_bytes
 * floats per call (scan and run tables, allocated here). */
isti_status isti_codec_create(int32_t device, uint64_t max_bytes, isti_codec **out);
isti_status isti_codec_destroy(isti_codec *c);

/* rle_encode (compression.py:176-204) of n bytes: each run of equal bytes
 * becomes (count, value) pairs of at most 255 bytes each, the count written
 * as 0 for a piece of length 1. Writes the pairs to out (capacity out_cap
 * bytes; 2n suffices) and their byte count to *out_len. Synchronises. */
isti_status isti_codec_rle_encode(isti_codec *c, const uint8_t *in, uint64_t n, uint8_t *out, uint64_t out_cap,
                                  uint64_t *out_len, void *stream);

/* rle_decode (compression.py:206-218) of n bytes of (count, value) pairs
 * (n even; count 0 stands for 1). Writes the bytes to out (capacity out_cap)
 * and their count to *out_len; a stream that expands beyond out_cap is an
 * error and writes nothing. Synchronises. */
isti_status isti_codec_rle_decode(isti_codec *c, const uint8_t *in, uint64_t n, uint8_t *out, uint64_t out_cap,
                                  uint64_t *out_len, void *stream);

/* out[i] = a[i] ^ b[i] (the XOR delta against the anchor frame). out may
 * alias a or b. */
isti_status isti_codec_xor(const uint8_t *a, const uint8_t *b, uint64_t n, uint8_t *out, void *stream);

/* Byte k of every 4-byte pixel: channel extraction of an RGBA8 image
 * (channel 2 is what the ShadowInfo codec keeps), and its inverse, which
 * writes RGBA (0, 0, plane, 255) pixels as the decoder does. */
isti_status isti_codec_extract_channel(const uint8_t *rgba, uint64_t n_px, int32_t channel, uint8_t *out,
                                       void *stream);
isti_status isti_codec_expand_channel(const uint8_t *plane, uint64_t n_px, uint8_t *rgba, void *stream);

/* normalize_exr (compression.py:84-91) of n floats: after load_exr's zeroing
 * of |v| < 1e-4 (when zero_small), min and max; if |max - min| <= 1e-6 every
 * output is 0 and min = max = 0, else out = RN(clip((v - min) / (max - min),
 * 0, 1) * 65535) (round half to even) as uint16, in binary32 arithmetic.
 * min_max (host, 2 floats) receives (min, max). Synchronises. */
isti_status isti_codec_normalize(isti_codec *c, const float *in, uint64_t n, int32_t zero_small, uint16_t *out,
                                 float *min_max, void *stream);

/* denormalize_exr (compression.py:93-98): out = half(q / 65535 * (max - min)
 * + min) in binary32, rounded to binary16 (returned as the float it
 * represents); 0 everywhere when |max - min| <= 1e-6. */
isti_status isti_codec_denormalize(const uint16_t *in, uint64_t n, float min, float max, float *out, void *stream);

const char *isti_codec_last_error(const isti_codec *c);

#ifdef __cplusplus
}
#endif

#endif
