/*
 * isti.h - C ABI of the MI355X (gfx950) per-frame interpolation renderer.
 *
 * This is the drop-in boundary for the reference's `Renderer::render` path
 * (include/renderer.h:52-56, src/renderer.cu:607-694). The reference binds a
 * C++ class directly; here the same call shape is exposed as plain C so a
 * C++ wrapper (isti_renderer.hpp), ctypes, or any FFI can bind it. Nothing in
 * these signatures is a torch or HIP C++ type: device buffers are plain
 * pointers, the stream is an opaque `void*` (a hipStream_t, or NULL for the
 * default stream).
 *
 * Conventions (fixed by the reference, restated here so this header stands on
 * its own):
 *   - matrices are 4x4 float, column-major, element (row r, col c) at m[c*4+r]
 *     (glm::mat4 layout of the FrameImage matrices, include/data.h:22-25);
 *   - pixel (x, y) maps to NDC (2(x+.5)/W-1, 2(y+.5)/H-1)
 *     (include/util/helper_common.h:66-75);
 *   - depth planes hold OpenGL NDC z in [-1, 1] (src/file_reader.cu:17-37);
 *   - object motion is a clip-space delta in the keyframe's own camera that
 *     moves the pixel's surface point to its position at the other keyframe
 *     (src/renderer.cu:101-102);
 *   - colour planes are RGBA8, rows packed (pitch = width * 4 bytes);
 *   - shadow planes follow the reference's soft-shadow encoding
 *     (src/renderer.cu:217-295): `shadow_info` RGB is a 24-bit field
 *     (R<<16 | G<<8 | B) holding a 2-bit shadow code per sub-frame,
 *     `shadow_color_a` / `shadow_color_b` the colours at the minimum / maximum
 *     code.
 *
 * Errors: every entry point returns an isti_status; no C++ exception crosses
 * this boundary. `isti_last_error` returns the message of the last failure of
 * a renderer (or of isti_create when passed NULL). The reference prints and
 * swallows device errors (include/util/cuda_helper_host.h:19-42) and throws
 * std::runtime_error on shape changes (src/util/image_buffer.cu:44); both
 * become status codes here.
 */
#ifndef ISTI_H
#define ISTI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Largest number of auxiliary cameras one renderer accepts. */
#define ISTI_MAX_AUX 4
/* Largest view width/height: source pixels are packed as 16-bit coordinates
 * into the splat keys (src/renderer.cu:65-68). */
#define ISTI_MAX_DIM 65535
/* Largest width * height of a view (2^28 - 1): byte offsets into its float4
 * flow planes fit 32 bits. */
#define ISTI_MAX_VIEW_PIXELS 268435455u
/* Largest shadow_frames_per_anchor: 2-bit codes in a 24-bit field. */
#define ISTI_MAX_SHADOW_FRAMES 12
/* Upper bound of isti_render_info.max_search_iterations. */
#define ISTI_MAX_SEARCH_ITERATIONS 4096

typedef enum isti_status {
    ISTI_OK = 0,
    ISTI_ERR_INVALID_ARG = 1,
    ISTI_ERR_OOM = 2,
    ISTI_ERR_DEVICE = 3,
    ISTI_ERR_UNSUPPORTED_SHAPE = 4
} isti_status;

/* RenderMode, include/renderer.h:9-16. */
typedef enum isti_render_mode {
    ISTI_MODE_OURS = 0,
    ISTI_MODE_BACKWARD_BIDIRECTIONAL = 1,
    ISTI_MODE_BACKWARD_UNIDIRECTIONAL = 2,
    ISTI_MODE_SPLATTING = 3,
    ISTI_MODE_TIMEWARP = 4
} isti_render_mode;

/* Keyframe slots: FrameData::fwd / FrameData::bwd (include/data.h:87-88),
 * and for aux cameras AuxCamData::curr / ::next (include/data.h:75). */
typedef enum isti_slot {
    ISTI_SLOT_FWD = 0,
    ISTI_SLOT_BWD = 1
} isti_slot;

/* One keyframe view (MotionZBufferImage, include/data.h:12-52). All plane
 * pointers are device pointers owned by the caller and must stay valid while
 * the renderer may read them (until the slot is set again or the renderer is
 * destroyed). View 0 is the main camera, views 1..n_aux the aux cameras.
 *   motion_x/y/z: all three NULL means "no object motion" (the reference's
 *                 std::nullopt object_motion, src/renderer.cu:539-546);
 *   shadow_*:     each may be NULL (the reference's texture handle -1,
 *                 src/renderer.cu:652). */
typedef struct isti_view_desc {
    uint32_t width;
    uint32_t height;
    const uint8_t *color;          /* RGBA8 */
    const float *z;                /* NDC z */
    const float *motion_x;         /* clip-space motion, x */
    const float *motion_y;         /* clip-space motion, y */
    const float *motion_z;         /* clip-space motion, z */
    const uint8_t *shadow_info;    /* RGBA8, RGB = 24-bit code field */
    const uint8_t *shadow_color_a; /* RGBA8, colour at the minimum code */
    const uint8_t *shadow_color_b; /* RGBA8, colour at the maximum code */
    float viewproj[16];
    float inv_viewproj[16];
    int32_t priority;              /* aux camera priority (src/file_reader.cu:102-104) */
} isti_view_desc;

/* RenderInfo + AdditionalRenderInfo (include/renderer.h:18-35).
 * max_search_iterations extends the reference: < 0 selects the mode's own
 * count (modes 0/2: 3, mode 1: 5, mode 3: 0, src/renderer.cu:678-689);
 * a value in [0, ISTI_MAX_SEARCH_ITERATIONS] caps the fixed-point search at
 * that many steps (it stops early at a fixed point, which is bit-identical
 * to running all steps). */
typedef struct isti_render_info {
    uint32_t width;
    uint32_t height;
    float alpha;
    float viewproj[16];
    float inv_viewproj[16];
    int32_t frames_per_anchor;
    int32_t shadow_frames_per_anchor;
    int32_t no_shadows;
    int32_t mode;
    int32_t max_search_iterations;
} isti_render_info;

typedef struct isti_config {
    int32_t device; /* HIP device ordinal */
    int32_t n_aux;  /* number of aux cameras, 0..ISTI_MAX_AUX */
} isti_config;

typedef struct isti_renderer isti_renderer;

/* Device scratch of one (slot, view) after the last render, for inspection:
 * the u64 splat buffer and the two flow planes, all device pointers, rows
 * packed. Each view keeps two splat buffers (8 bytes per source pixel each):
 * isti_render's resolve clears the one the frame did not use (the
 * reference's clear, renderer.cu:631-638, done a frame early), and the next
 * frame splats into it without a clear, so `splat` alternates between them
 * from render to render. cam_ff (source -> target camera) holds 4 floats per pixel: flow x,
 * y, the moved NDC z = RN32(z + flow z) (the reference keeps the flow z and
 * adds the keyframe z where it reads it, renderer.cu:338-341; the sum is
 * stored instead, so a resolve needs no z fetch) and the visible-in-main flag
 * (computed for aux views only: a main view's flag is never read, renderer.cu:455,
 * and is 0);
 * full_ff (source -> other keyframe) 3 floats (x, y, z; the reference's
 * fourth component repeats the flag and is never read).
 * Reference: Renderer members, include/renderer.h:64-92. */
typedef struct isti_scratch_view {
    uint32_t width;
    uint32_t height;
    const uint64_t *splat;
    const float *cam_ff;
    const float *full_ff;
} isti_scratch_view;

/* Renderer::Renderer (include/renderer.h:54). Creates a renderer bound to
 * config->device. Allocates nothing on the device until the first render. */
isti_status isti_create(const isti_config *config, isti_renderer **out);

/* Frees all device scratch (the reference has no destructor and leaks). */
isti_status isti_destroy(isti_renderer *r);

/* Sets one keyframe view of a slot (the reference loads FrameData through
 * FrameFileReader, src/file_reader.cu:138-217; here the caller hands over
 * device planes). view: 0 = main, 1..n_aux = aux camera view-1.
 * Every render reads the planes afresh, except the flow of a view toward its
 * partner keyframe (the other slot's same view), which depends on the pair
 * alone: the first render after this view or its partner was set (or after a
 * swap) computes it, and later renders reuse it. A caller that rewrites a
 * keyframe's planes in place sets the view again before the next render (the
 * reference's render takes its FrameData on every call; setting every view
 * before every render reproduces that exactly). */
isti_status isti_keyframe_set(isti_renderer *r, int32_t slot, int32_t view,
                              const isti_view_desc *desc);

/* FrameData::swap (include/data.h:110-116): exchanges the fwd and bwd
 * descriptors of the main view and of every aux camera. */
isti_status isti_keyframe_swap(isti_renderer *r);

/* Renderer::render (src/renderer.cu:607-694). Writes RGBA8 pixels into the
 * caller-owned device image `out_rgba8` (row pitch `out_pitch_bytes`) with the
 * reference's contract: only pixels with a valid result are written, so holes
 * keep the previous content (the reference's hole fill). Stream-ordered on
 * `stream`; no host synchronisation inside. */
isti_status isti_render(isti_renderer *r, const isti_render_info *info,
                        uint8_t *out_rgba8, size_t out_pitch_bytes, void *stream);

/* The frames of the current keyframe pair, infos[0..n-1], into outs[i]
 * (row pitch pitches[i]): the same bytes as n isti_render calls in that
 * order on `stream`, with consecutive frames on two internal streams and
 * two sets of splat and cam_ff buffers, so one frame's clear and splat run
 * beside another frame's resolve (the pair's flow toward the other keyframe
 * is shared: the first frame computes it when the pair is new, and the
 * other stream's frames wait for it). Ordered after all work enqueued on
 * `stream` before the call; work enqueued on `stream` after it sees every
 * frame. Outputs must be distinct images; modes 0-3 (timewarp frames use
 * isti_render); 1 <= n <= ISTI_MAX_FRAMES. No counterpart in the reference,
 * whose frame loop renders one client frame per call (src/main.cu:149-191);
 * it is for callers that know several client cameras of a pair at once.
 * Stage timing (isti_set_stage_timing) does not record these frames, and
 * isti_scratch shows the buffers of the frames at even positions. */
#define ISTI_MAX_FRAMES 64
isti_status isti_render_frames(isti_renderer *r, const isti_render_info *infos, uint32_t n,
                               uint8_t *const *outs, const size_t *pitches, void *stream);

/* Size of the static device scratch of isti_render_noalloc, per GPU. */
#define ISTI_NOALLOC_SCRATCH_BYTES (64u << 20)

/* isti_render that allocates nothing. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, contract and results as isti_render, but every scratch plane of
 * the frame is carved from one static device array of
 * ISTI_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of
 * the renderer's own buffers. A frame needs 36 bytes per source pixel of
 * every keyframe view (rounded up to 256 bytes per plane; about 64 MiB holds
 * a 1280x720 main view pair), timewarp 8 bytes per keyframe row; a larger
 * frame returns ISTI_ERR_UNSUPPORTED_SHAPE before any device work. Every call
 * recomputes the keyframe pair's flow (the array keeps only the last call's
 * contents). All calls on one GPU share the array: calls whose work overlaps
 * on that GPU (other streams, other threads) race and must be serialised by
 * the caller. */
isti_status isti_render_noalloc(isti_renderer *r, const isti_render_info *info,
                                uint8_t *out_rgba8, size_t out_pitch_bytes, void *stream);

/* Resolve tile shapes (render pixels per workgroup, one thread each):
 * 0 = 16x16, 1 = 32x8, 2 = 8x8 (one wave per workgroup), 3 = 8x32. Every
 * shape gives identical results; they differ in speed (halo staging, gather
 * locality, wave shape, and how long a workgroup's finished waves wait for
 * its slowest one). No counterpart in the reference (16x16 blocks,
 * src/renderer.cu:676-693). */
#define ISTI_N_TILE_SHAPES 4

/* Splat variants: 0 = direct (one 64-bit atomicMin into HBM per splatting
 * source pixel), 1 = LDS-binned (32x8 source tiles resolve their atomicMin
 * in an LDS bin first - the source tile moved by its first splatting
 * pixel's displacement, plus 4 cells a side - then one HBM atomicMin per
 * non-empty bin cell; targets outside the bin go to HBM).
 * Both leave identical splat buffers. No counterpart in the reference
 * (direct atomicMin, src/renderer.cu:144-145). */
#define ISTI_AUTOTUNE_TIMES (ISTI_N_TILE_SHAPES + 2)

/* Autotuner: times frame `info`'s splat stage (clear + reprojection/splat)
 * with each splat variant and its resolve pass with every tile shape
 * (`reps` launches each, HIP events) and keeps the fastest of each for every
 * later isti_render of the same render size, aux count and search window,
 * in this process. times_ms (optional, ISTI_AUTOTUNE_TIMES floats) receives
 * the average resolve time per tile shape, then the splat-stage time of the
 * direct and the binned variant; *best / *best_splat (optional) the choices.
 * `out_rgba8` receives the frame. Synchronises `stream`. */
isti_status isti_autotune(isti_renderer *r, const isti_render_info *info, uint8_t *out_rgba8,
                          size_t out_pitch_bytes, int32_t reps, float *times_ms, int32_t *best,
                          int32_t *best_splat, void *stream);

/* Forces r's splat variant (0 direct, 1 LDS-binned); -1 returns to the
 * tuned one, or direct where nothing was tuned. */
isti_status isti_set_splat_binning(isti_renderer *r, int32_t mode);

/* Forces the resolve tile shape of r's renders (0..ISTI_N_TILE_SHAPES-1);
 * -1 returns to the tuned shape, or 16x16 where nothing was tuned. */
isti_status isti_set_tile_shape(isti_renderer *r, int32_t shape);

/* Device scratch of (slot, view) from the last render (see above). */
isti_status isti_scratch(const isti_renderer *r, int32_t slot, int32_t view,
                         isti_scratch_view *out);

/* Per-stage GPU timing (the reference has no timers): when enabled,
 * isti_render records HIP events between its stages on the render stream,
 * into a ring of event sets that is harvested lazily (no extra host sync in
 * steady state). Enabling also resets the totals. Off by default. */
isti_status isti_set_stage_timing(isti_renderer *r, int32_t enable);

/* Stage durations of the last timed render, in milliseconds:
 * ms[0] splat clear (mode 4: mean z), ms[1] reproject + splat (mode 4: 0),
 * ms[2] resolve (mode 4: timewarp). Waits for that render to finish. */
isti_status isti_stage_times(isti_renderer *r, float ms[3]);

/* Sums of the stage durations (ms, same order) over every timed render
 * since timing was enabled, and their count. Waits for pending renders. */
isti_status isti_stage_totals(isti_renderer *r, double total_ms[3], int64_t *count);

/* ---- Batched rendering (no counterpart in the reference, which renders one
 * view per call with a host sync inside, src/renderer.cu:650). Serves many
 * renders at once -- client sessions, stereo eyes, views of uneven size --
 * with ONE persistent launch per stage: splat clear, reproject + splat, and
 * one resolve launch per distinct (mode group, n_aux) present. Work
 * descriptors go to a device table (double-buffered, one async copy per
 * batch), so the kernel grids and arguments are the same for every batch and
 * the stage sequence can be replayed from a captured hipGraph. */
typedef struct isti_batch isti_batch;

/* A batch context for up to max_requests renders per call on `device`. */
isti_status isti_batch_create(int32_t device, int32_t max_requests, isti_batch **out);
isti_status isti_batch_destroy(isti_batch *b);

/* enable != 0: launch the stages through captured hipGraphs (one per set of
 * resolve variants in a batch, captured on first use). */
isti_status isti_batch_set_graph(isti_batch *b, int32_t enable);

/* Renders request i = (renderers[i], infos[i]) into outs[i] (row pitch
 * out_pitches[i] bytes) for i < n, stream-ordered on `stream`, with the same
 * results as n separate isti_render calls. Modes 0-3 (timewarp is not
 * batched). A renderer may appear at most once per call; its keyframes must
 * be set. Waits on the host only when both table buffers are still in use by
 * earlier batches. */
isti_status isti_render_batch(isti_batch *b, isti_renderer *const *renderers, const isti_render_info *infos,
                              uint8_t *const *outs, const size_t *out_pitches, int32_t n, void *stream);

/* Message of the last failure of a batch call. Never NULL. */
const char *isti_batch_last_error(const isti_batch *b);

/* Message of the last failure of `r`, or of the calling thread's last
 * isti_create failure when r is NULL. Never NULL. */
const char *isti_last_error(const isti_renderer *r);

/* Library version string. */
const char *isti_version(void);

/* Diagnostic with no counterpart in the reference: checks the projective
 * division of the reprojection (RN32(n / w) of three numerators through one
 * refined reciprocal of w, exact by construction) against the IEEE division
 * on n generated operand sets of four families: random bit patterns, moderate
 * exponents, quotients next to rounding midpoints, every exponent including
 * subnormals. counts: device uint32[8], accumulated (zero it first):
 * [0..3] mismatches per family, which stay 0; [4..7] the sets per family
 * answered by the reciprocal path rather than the IEEE fallback. */
isti_status isti_selftest_division(int32_t device, uint64_t n, uint64_t seed, uint32_t *counts, void *stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* ISTI_H */
