/*
 * isti_io.h - reader of the reference's keyframe recordings (host code).
 *
 * Replaces FrameFileReader (src/file_reader.cu:17-230, include/file_reader.h:9-67)
 * and the image decoding of ImageBuffer::readFromFile (src/util/image_buffer.cu:104-178).
 * A recording directory holds one directory per camera ("main" and the aux
 * camera names), each with:
 *   cam.json                     frames[] {idx, w2c, proj}, resolution {x, y},
 *                                clientFps (60), serverFps (10), shadowFps (60),
 *                                objectMotion (true), shadowGradients (true)
 *   color/NNNN.png               RGBA8 colour of frame idx NNNN
 *   depth/NNNN.exr               depth (first channel), reversed-Z in [0, 1]
 *   forwardFlowField/NNNN.exr    clip-space motion of frame i toward i+1 (RGB)
 *   backwardFlowField/NNNN.exr   clip-space motion of frame i+1 toward i, stored
 *                                under frame i's index
 *   forward|backwardShadowInfo|ShadowColorA|ShadowColorB/NNNN.png (optional)
 * Images are decoded here (PNG: 8-bit grey/grey-alpha/RGB/RGBA, any filter,
 * not interlaced; EXR: scanline, NONE/ZIPS/ZIP compression, HALF/FLOAT
 * channels) into the layouts isti_keyframe_set takes: RGBA8 colour, NDC z
 * (z = 2 (1 - d) - 1, file_reader.cu:17-37), three motion planes, and the
 * shadow triple; matrices are normalised as FrameFileReader::normalizeProjMat
 * does (Unity 2022.3, reversed Z, [0, 1] depth; file_reader.cu:219-230).
 */
#ifndef ISTI_IO_H
#define ISTI_IO_H

#include <stddef.h>
#include <stdint.h>

#include "isti.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FrameFileReader::CameraInfo (include/file_reader.h:14-24). */
typedef struct isti_io_camera {
    uint32_t width, height;
    int32_t client_fps, server_fps, shadow_fps;
    int32_t priority; /* objectMotion * 2 + shadowGradients (file_reader.cu:102-104) */
    int32_t n_frames;
} isti_io_camera;

/* One camera frame: image index, world-to-camera and (raw) projection
 * matrices, column-major (glm layout; cam.json "eRC" is row R, column C). */
typedef struct isti_io_frame {
    int32_t idx;
    float w2c[16];
    float proj[16];
} isti_io_frame;

/* Reads <cam_dir>/cam.json. frames (optional) receives up to max_frames
 * entries. */
isti_status isti_io_read_camera(const char *cam_dir, isti_io_camera *cam, isti_io_frame *frames,
                                int32_t max_frames);

/* One keyframe view in host memory, owned by the library (isti_io_free_view).
 * Absent optional planes are NULL. */
typedef struct isti_io_view {
    uint32_t width, height;
    uint8_t *color;                          /* RGBA8, w*h*4 */
    float *z;                                /* NDC z, w*h */
    float *motion_x, *motion_y, *motion_z;   /* clip-space motion, w*h each */
    uint8_t *shadow_info, *shadow_color_a, *shadow_color_b; /* RGBA8 */
    float view[16];                          /* w2c */
    float proj[16];                          /* normalised projection */
    float viewproj[16];                      /* proj * w2c */
    float inv_viewproj[16];                  /* its inverse */
} isti_io_view;

/* Keyframe view `which` (0: frame `pair` with its forward data, 1: frame
 * pair+1 with its backward data) of the keyframe pair (pair, pair+1) of the
 * camera in cam_dir (FrameFileReader::readColorDepthCam + readNextFrame,
 * file_reader.cu:138-217). */
isti_status isti_io_read_view(const char *cam_dir, int32_t pair, int32_t which, isti_io_view *out);
void isti_io_free_view(isti_io_view *v);

/* The client camera between two keyframes (src/main.cu:26-29, 158-161):
 * view_t = glm::interpolate(view_fwd, view_bwd, t) - the rotation part
 * rotated about the axis of view_bwd * transpose(rot(view_fwd)) by t times
 * its angle, the translation column lerped - and viewproj = proj_fwd *
 * view_t with its inverse. Restates glm 0.9.9 gtx/matrix_interpolation
 * (interpolate, axisAngle, axisAngleMatrix, extractMatrixRotation) in
 * single precision. Any output may be NULL. */
isti_status isti_io_camera_at(const float view_fwd[16], const float view_bwd[16], const float proj_fwd[16], float t,
                              float view_t[16], float viewproj[16], float inv_viewproj[16]);

/* The client camera of a camera-path file at one client frame
 * (src/main.cu:158-161 with -c): viewproj = normalizeProjMat(proj) * w2c
 * (file_reader.cu:219-230) and its inverse, from a frame's raw matrices as
 * isti_io_read_camera returns them. Either output may be NULL. */
isti_status isti_io_client_camera(const float w2c[16], const float proj[16], float viewproj[16],
                                  float inv_viewproj[16]);

/* Image decoders (exposed for tests and tools). PNG -> RGBA8; EXR -> the
 * listed channels (names, comma separated, e.g. "R,G,B"; a missing channel
 * reads 0) as interleaved float. *w and *h receive the size; out may be NULL
 * to query it. Otherwise out holds out_bytes bytes / out_floats floats, and a
 * decoded image larger than that (w*h*4 bytes, w*h*n_channels floats; the
 * file may change between the query and the fill) is an error that writes
 * nothing. Images of ISTI_MAX_VIEW_PIXELS pixels or more are rejected. */
isti_status isti_io_decode_png(const char *path, uint32_t *w, uint32_t *h, uint8_t *out, size_t out_bytes);
isti_status isti_io_decode_exr(const char *path, const char *channels, uint32_t *w, uint32_t *h, float *out,
                               size_t out_floats);

/* Message of the calling thread's last failure. Never NULL. */
const char *isti_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
