/*
 * isti_scene.h - seeded synthetic keyframe generator (analytic ray cast).
 *
 * Stands in for the reference's on-disk keyframe reader
 * (src/file_reader.cu:109-217) and emits keyframes in the reference's
 * formats: RGBA8 colour, OpenGL NDC z (src/file_reader.cu:17-37), object
 * motion as a clip-space delta in the keyframe's own camera toward the other
 * keyframe (consumed at src/renderer.cu:101-102), and the soft-shadow triple
 * ShadowInfo / ShadowColorA / ShadowColorB (consumed at
 * src/renderer.cu:217-295).
 *
 * Scene: a checker-textured ground plane (y = 0) plus spheres moving at
 * constant velocity (or on sinusoidal paths), one camera path that translates
 * and yaws at a constant rate per frame, perspective projection with
 * near/far planes, optional aux cameras (offset along the camera's right axis)
 * or stereo eyes, optional soft shadows from one disk area light.
 * Time is measured in client frames; keyframe k is at time k * keyframe_spacing.
 * Matrices are column-major float[16]; rows of every image are top-down
 * (the projection flips y, as the reference's Unity normalisation does,
 * include/file_reader.h:64).
 */
#ifndef ISTI_SCENE_H
#define ISTI_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct isti_scene_config {
    uint64_t seed;
    uint32_t width, height;         /* main view (each eye when n_eyes == 2) */
    float vfov_deg;                 /* main vertical field of view */
    float near_m, far_m;            /* projection planes; linear depth range */
    int32_t n_aux;                  /* aux cameras: 0 or 2 (left / right) */
    uint32_t aux_width, aux_height;
    float aux_offset_m;             /* aux camera offset along camera right */
    float aux_vfov_deg;
    int32_t n_eyes;                 /* 1 = mono, 2 = stereo (eye = view index) */
    float ipd_m;
    int32_t n_spheres;
    float radius_min, radius_max;
    float speed_max;                /* |v| per axis, m/frame, U[-speed_max, speed_max] */
    int32_t sinusoidal;             /* 1: nonlinear sphere paths */
    float cam_speed;                /* m/frame */
    float cam_yaw_deg;              /* deg/frame */
    int32_t keyframe_spacing;       /* client frames between keyframes */
    int32_t shadows;                /* 1: emit the soft-shadow triple */
    int32_t shadow_samples;         /* area-light samples per pixel */
    float light_radius_m;
    int32_t shadow_frames;          /* shadow sub-frames per keyframe pair (<= 12) */
} isti_scene_config;

/* Fills cfg with the defaults of a named workload:
 *   1: 640x360, 16 spheres, no aux, no shadows, spacing 4;
 *   2: 1920x1080, 2 aux cameras 960x540 at +-0.3 m, soft shadows
 *      (radius 0.5 m, 16 samples, 8 sub-frames), spacing 8;
 *   3: stereo 2064x2208 per eye, IPD 63 mm, 64 spheres on sinusoidal paths,
 *      0.15 m/frame, 2 deg/frame, spacing 6.
 * Returns 0, or -1 for an unknown workload. */
int isti_scene_defaults(int workload, uint64_t seed, isti_scene_config *cfg);

/* Camera of view `view` at time `t` (frames): view 0 main (or left eye),
 * 1..n_aux aux cameras (or 1 = right eye). Writes the view-projection
 * matrix, its inverse, and the view's width/height. */
int isti_scene_camera(const isti_scene_config *cfg, double t, int view, float viewproj[16],
                      float inv_viewproj[16], uint32_t *width, uint32_t *height);

/* The view (world -> camera) and projection matrices of `view` at time t,
 * column-major; viewproj of isti_scene_camera is proj * view. */
int isti_scene_camera_parts(const isti_scene_config *cfg, double t, int view, float view_m[16], float proj[16]);

/* Keyframe `k` of view `view`. direction +1: motion (and shadow sub-frames)
 * toward keyframe k+1 (the reference's forward flow field / forward shadow
 * dirs); -1: toward keyframe k-1 (backward). Outputs (any may be NULL):
 *   color RGBA8 [W*H*4], z NDC [W*H], motion x/y/z [W*H each],
 *   shadow_info / shadow_a / shadow_b RGBA8 [W*H*4] (zero when shadows off),
 *   viewproj / inv_viewproj [16].
 * n_threads <= 0 uses all OpenMP threads. Returns 0 on success. */
int isti_scene_keyframe(const isti_scene_config *cfg, int k, int view, int direction,
                        uint8_t *color, float *z, float *motion_x, float *motion_y, float *motion_z,
                        uint8_t *shadow_info, uint8_t *shadow_a, uint8_t *shadow_b,
                        float viewproj[16], float inv_viewproj[16], int n_threads);

/* Analytic ground-truth colour of view `view` at time `t` (frames), with the
 * same shading (and quantised shadows) as the keyframes. */
int isti_scene_ground_truth(const isti_scene_config *cfg, double t, int view, uint8_t *color,
                            int n_threads);

#ifdef __cplusplus
}
#endif

#endif
