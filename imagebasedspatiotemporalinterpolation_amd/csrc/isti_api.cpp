// isti_api.cpp - the C ABI of include/isti.h: renderer state, device scratch
// and the per-frame stage sequence of Renderer::render
// (src/renderer.cu:549-694).
//
// Per frame (modes 0-3) the host issues three stream-ordered launches and
// nothing else: splat clear, one reproject+splat launch covering every
// source view, one resolve launch. No host synchronisation and no
// host->device copies happen inside isti_render, so a frame can be captured
// into a hipGraph. Scratch is (re)allocated only when a view's size changes.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <string>
#include <utility>
#include <vector>

#include "../../include/isti.h"
#include "../../include/isti_learned.h"
#include "isti_internal.hpp"
#include "isti_kernels.hpp"

using namespace isti;

namespace {

thread_local std::string g_create_error;

// Tile shapes and splat variants chosen by isti_autotune, per (width,
// height, n_aux, window).
using TuneKey = std::tuple<uint32_t, uint32_t, int, int>;
std::mutex g_tuned_mutex;
std::map<TuneKey, int> g_tuned;
std::map<TuneKey, int> g_tuned_splat;

struct Scratch {
    uint32_t w = 0, h = 0;
    unsigned long long *splat = nullptr;
    float4 *cam_ff = nullptr;
    Flow3 *full_ff = nullptr;
    // full_ff holds the flow of the keyframe pair currently set (it depends
    // on the pair alone); cleared when either view of the pair is set, on a
    // swap and when the buffers are reallocated
    bool full_ok = false;
    // second splat buffer of the view: the plain resolve of one frame clears
    // it (ClearAhead) while it reads `splat`, and the next frame planned on
    // the view swaps it in and skips its clear
    unsigned long long *spare = nullptr;
    bool spare_clean = false;
};

// Ring of per-render event sets for stage timing. A slot is harvested (its
// durations added to the totals) when it is about to be reused, by which
// time its render has long finished.
constexpr int RING = 64;
struct TimingRing {
    hipEvent_t ev[RING][4] = {};
    bool pending[RING] = {};
    int head = 0;
    int last = -1;
    double total[3] = {0.0, 0.0, 0.0};
    int64_t count = 0;
};

}  // namespace

struct isti_renderer {
    int device = 0;
    int n_aux = 0;
    isti_view_desc views[2][MAX_VIEWS];
    bool view_set[2][MAX_VIEWS];
    Scratch scratch[2][MAX_VIEWS];
    double *row_sums = nullptr;
    uint32_t row_sums_cap = 0;
    float *mean_z = nullptr;
    bool timing = false;
    TimingRing ring;
    uint32_t *q_out = nullptr;  // optional search-result capture (isti_set_search_capture)
    uint32_t *stat_slots = nullptr;  // learned-start statistic rows (zero between frames)
    // frame whose splat stage isti_learned_features ran last (isti_render_learned resolves it)
    isti_render_info prepared{};
    bool prepared_valid = false;
    // isti_render_frames: a second set of splat and cam_ff buffers (its full_ff
    // is the first set's), the two streams consecutive frames alternate on,
    // and the events that fork them from and join them to the caller's stream
    Scratch scratch2[2][MAX_VIEWS];
    hipStream_t pipe[2] = {nullptr, nullptr};
    hipEvent_t pipe_ev[4] = {};  // fork, flow written, join 0, join 1
    int tile_shape = -1;  // forced resolve tile shape, or -1 (tuned / default)
    int splat_mode = -1;  // forced splat variant (0 direct, 1 LDS-binned), or -1 (tuned / direct)
    std::string err;
};

namespace {

isti_status fail(isti_renderer *r, isti_status st, const std::string &msg) {
    if (r) r->err = msg;
    return st;
}

isti_status dev_fail(isti_renderer *r, hipError_t e, const char *where) {
    return fail(r, e == hipErrorOutOfMemory ? ISTI_ERR_OOM : ISTI_ERR_DEVICE,
                std::string(where) + ": " + hipGetErrorString(e));
}

void free_scratch(Scratch &s) {
    if (s.splat) (void)hipFree(s.splat);
    if (s.spare) (void)hipFree(s.spare);
    if (s.cam_ff) (void)hipFree(s.cam_ff);
    if (s.full_ff) (void)hipFree(s.full_ff);
    s = Scratch{};
}

isti_status ensure_scratch(isti_renderer *r, Scratch &s, uint32_t w, uint32_t h) {
    if (s.w == w && s.h == h && s.splat) return ISTI_OK;
    free_scratch(s);
    const size_t n = (size_t)w * h;
    hipError_t e = hipMalloc(&s.splat, n * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&s.cam_ff, n * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&s.full_ff, n * sizeof(Flow3));
    if (e != hipSuccess) {
        free_scratch(s);
        return dev_fail(r, e, "hipMalloc(scratch)");
    }
    s.w = w;
    s.h = h;
    return ISTI_OK;
}

// The second buffer set of isti_render_frames: splat and cam_ff only.
void free_scratch2(Scratch &s) {
    if (s.splat) (void)hipFree(s.splat);
    if (s.spare) (void)hipFree(s.spare);
    if (s.cam_ff) (void)hipFree(s.cam_ff);
    s = Scratch{};
}

isti_status ensure_scratch2(isti_renderer *r, Scratch &s, uint32_t w, uint32_t h) {
    if (s.w == w && s.h == h && s.splat) return ISTI_OK;
    free_scratch2(s);
    const size_t n = (size_t)w * h;
    hipError_t e = hipMalloc(&s.splat, n * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&s.cam_ff, n * sizeof(float4));
    if (e != hipSuccess) {
        free_scratch2(s);
        return dev_fail(r, e, "hipMalloc(scratch)");
    }
    s.w = w;
    s.h = h;
    return ISTI_OK;
}

Mat4 to_mat(const float *m) {
    Mat4 M;
    std::memcpy(M.m, m, sizeof(M.m));
    return M;
}

bool has_motion(const isti_view_desc &d) { return d.motion_x && d.motion_y && d.motion_z; }

ResolveView resolve_view(const isti_view_desc &d, const Scratch &s, bool forward, int priority) {
    ResolveView v;
    v.img = reinterpret_cast<const uchar4 *>(d.color);
    v.z = d.z;
    v.sh_info = reinterpret_cast<const uchar4 *>(d.shadow_info);
    v.sh_a = reinterpret_cast<const uchar4 *>(d.shadow_color_a);
    v.sh_b = reinterpret_cast<const uchar4 *>(d.shadow_color_b);
    v.cam_ff = s.cam_ff;
    v.full_ff = s.full_ff;
    v.splat = s.splat;
    v.w = d.width;
    v.h = d.height;
    v.priority = priority;
    v.forward = forward ? 1 : 0;
    v.inv_w = 1.0 / (double)d.width;
    v.inv_h = 1.0 / (double)d.height;
    const float mn = (float)(d.width < d.height ? d.width : d.height);
    v.rs_x = (float)d.width / mn;
    v.rs_y = (float)d.height / mn;
    v.ratio_x = v.ratio_y = 1;
    v.magic_x = v.magic_y = 0;
    return v;
}

// Multiply-high divisor of div_small: floor(2^32 / r) + 1 (unused for r = 1).
uint32_t div_magic(uint32_t r) { return r <= 1 ? 0u : (uint32_t)((1ull << 32) / r + 1ull); }

// Render-to-view pixel ratio of a resolve view (renderer.cu:433).
void set_ratio(ResolveView &v, uint32_t render_w, uint32_t render_h) {
    v.ratio_x = render_w / v.w;
    v.ratio_y = render_h / v.h;
    v.magic_x = div_magic(v.ratio_x);
    v.magic_y = div_magic(v.ratio_y);
}

isti_status validate(isti_renderer *r, const isti_render_info *info, size_t out_pitch) {
    if (info->mode < ISTI_MODE_OURS || info->mode > ISTI_MODE_TIMEWARP)
        return fail(r, ISTI_ERR_INVALID_ARG, "mode must be 0..4");
    if (info->width == 0 || info->height == 0 || info->width > ISTI_MAX_DIM || info->height > ISTI_MAX_DIM)
        return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE, "render resolution must be 1..65535 per axis");
    if (info->max_search_iterations > ISTI_MAX_SEARCH_ITERATIONS)
        return fail(r, ISTI_ERR_INVALID_ARG, "max_search_iterations must be <= 4096");
    if (out_pitch < (size_t)info->width * 4 || out_pitch % 4 != 0)
        return fail(r, ISTI_ERR_INVALID_ARG, "out_pitch_bytes must be a multiple of 4 and >= width*4");
    const int nv = info->mode == ISTI_MODE_TIMEWARP ? 1 : 1 + r->n_aux;
    for (int s = 0; s < (info->mode == ISTI_MODE_TIMEWARP ? 1 : 2); s++)
        for (int v = 0; v < nv; v++)
            if (!r->view_set[s][v]) return fail(r, ISTI_ERR_INVALID_ARG, "keyframe view not set");
    if (info->mode == ISTI_MODE_TIMEWARP) return ISTI_OK;
    const isti_view_desc &f = r->views[0][0], &b = r->views[1][0];
    // renderer.cu:145,153: splat buffers are indexed by keyframe dims but
    // sized by the render resolution, so the two must agree.
    if (f.width != info->width || f.height != info->height || b.width != info->width || b.height != info->height)
        return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE, "main keyframes must have the render resolution");
    if (info->shadow_frames_per_anchor < 0 || info->shadow_frames_per_anchor > ISTI_MAX_SHADOW_FRAMES)
        return fail(r, ISTI_ERR_INVALID_ARG, "shadow_frames_per_anchor must be 0..12");
    for (int v = 1; v <= r->n_aux; v++) {
        const isti_view_desc &af = r->views[0][v], &ab = r->views[1][v];
        if (af.width != ab.width || af.height != ab.height)
            return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE, "aux keyframes of one camera must share a resolution");
        // renderer.cu:433: aux pixel = render pixel / (render res / aux res)
        if (af.width > info->width || af.height > info->height || info->width % af.width != 0 ||
            info->height % af.height != 0)
            return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE, "render resolution must be an integer multiple of each aux resolution");
    }
    return ISTI_OK;
}

hipError_t harvest(TimingRing &t, int slot) {
    if (!t.pending[slot]) return hipSuccess;
    hipError_t e = hipEventSynchronize(t.ev[slot][3]);
    float ms[3];
    for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipEventElapsedTime(&ms[i], t.ev[slot][i], t.ev[slot][i + 1]);
    if (e != hipSuccess) return e;
    for (int i = 0; i < 3; i++) t.total[i] += ms[i];
    t.count++;
    t.pending[slot] = false;
    return hipSuccess;
}

// Records stage boundary `i` of the current render when timing is on.
// Boundary 0 claims the next ring slot (harvesting it first if needed).
hipError_t mark(isti_renderer *r, int i, hipStream_t s) {
    if (!r->timing) return hipSuccess;
    TimingRing &t = r->ring;
    if (i == 0) {
        hipError_t e = harvest(t, t.head);
        if (e != hipSuccess) return e;
        t.last = t.head;
        t.head = (t.head + 1) % RING;
    }
    hipError_t e = hipEventRecord(t.ev[t.last][i], s);
    if (e == hipSuccess && i == 3) t.pending[t.last] = true;
    return e;
}

// ext_rows / ext_mean: device scratch for the row sums (height doubles) and
// the mean, or null for the renderer's own (allocated on first use).
isti_status render_timewarp(isti_renderer *r, const isti_render_info *info, uint8_t *out, size_t pitch,
                            hipStream_t s, double *ext_rows = nullptr, float *ext_mean = nullptr) {
    // src/renderer.cu:615-629
    const isti_view_desc &f = r->views[0][0];
    if (ext_rows) {
        // caller-provided scratch: nothing is allocated
    } else if (r->row_sums_cap < f.height) {
        if (r->row_sums) (void)hipFree(r->row_sums);
        r->row_sums = nullptr;
        r->row_sums_cap = 0;
        hipError_t e = hipMalloc(&r->row_sums, sizeof(double) * f.height);
        if (e != hipSuccess) return dev_fail(r, e, "hipMalloc(row sums)");
        r->row_sums_cap = f.height;
    }
    if (!ext_mean && !r->mean_z) {
        hipError_t e = hipMalloc(&r->mean_z, sizeof(float));
        if (e != hipSuccess) return dev_fail(r, e, "hipMalloc(mean z)");
    }
    double *rows = ext_rows ? ext_rows : r->row_sums;
    float *mean = ext_mean ? ext_mean : r->mean_z;
    hipError_t e = mark(r, 0, s);
    if (e == hipSuccess) e = launch_mean_z(f.z, f.width, f.height, rows, mean, s);
    if (e != hipSuccess) return dev_fail(r, e, "mean z launch");
    if ((e = mark(r, 1, s)) != hipSuccess || (e = mark(r, 2, s)) != hipSuccess) return dev_fail(r, e, "event");

    TimewarpArgs a;
    std::memset(&a, 0, sizeof(a));
    a.inv_vp_from = to_mat(info->inv_viewproj);
    a.inv_vp_to = to_mat(f.inv_viewproj);
    a.vp_to = to_mat(f.viewproj);
    a.depth_to = mean;
    a.img_to = reinterpret_cast<const uchar4 *>(f.color);
    a.w_to = f.width;
    a.h_to = f.height;
    if (r->n_aux > 0 && r->view_set[0][1]) {
        const isti_view_desc &x = r->views[0][1];
        a.aux_set = 1;
        a.vp_aux = to_mat(x.viewproj);
        a.img_aux = reinterpret_cast<const uchar4 *>(x.color);
        a.w_aux = x.width;
        a.h_aux = x.height;
    }
    a.out = reinterpret_cast<uchar4 *>(out);
    a.out_pitch_px = (uint32_t)(pitch / 4);
    a.w = info->width;
    a.h = info->height;
    a.tiles_x = (a.w + TILE_W - 1) / TILE_W;
    a.n_tiles = a.tiles_x * ((a.h + TILE_H - 1) / TILE_H);
    e = launch_timewarp(a, s);
    if (e != hipSuccess) return dev_fail(r, e, "timewarp launch");
    if ((e = mark(r, 3, s)) != hipSuccess) return dev_fail(r, e, "event");
    return ISTI_OK;
}

// Everything one frame (modes 0-3) launches, before block offsets are
// assigned: splat jobs ordered fwd main, bwd main, then per aux camera
// curr / next with the reference's weights and matrices
// (src/renderer.cu:643-649), and the resolve descriptor (:652-693).
struct FramePlan {
    SplatJob jobs[MAX_SPLAT_JOBS];
    int nj = 0;
    Mat4 vp_cam;
    Mat4 vp_main[2];
    int32_t init_linear = 0;
    ResolveArgs ra;
    int window = 1;
    bool always_valid = false;
    bool own_scratch = true;  // false: planes carved from caller-provided scratch
    // job j's splat buffer already holds the empty key (a resolve cleared it
    // ahead), so its clear is skipped; reset once the buffer is splatted
    bool clean[MAX_SPLAT_JOBS] = {};
};

// The view's spare splat buffer, once a resolve has cleared it (ClearAhead),
// becomes the buffer of the next frame planned on the view; returns whether
// that happened (the frame's clear of this buffer is then skipped).
bool take_clean(Scratch &s) {
    if (!s.spare || !s.spare_clean) return false;
    std::swap(s.splat, s.spare);
    s.spare_clean = false;
    return true;
}

// ext: per (slot, view) scratch planes carved from caller-provided memory, or
// null for the renderer's own. Carved planes may hold another call's data, so
// the flow toward the other keyframe is always recomputed, unless ext_full
// says that ext's full_ff planes (and full_ok flags) are the renderer's own.
// The renderer's own views swap in a cleared spare splat buffer (take_clean)
// unless take_spare is false (a frame already splatted, resolved again);
// bit (slot * MAX_VIEWS + view) of ext_clean says the same of ext's buffers.
isti_status plan_frame(isti_renderer *r, const isti_render_info *info, uint8_t *out, size_t out_pitch,
                       FramePlan &P, const Scratch (*ext)[MAX_VIEWS] = nullptr, bool ext_full = false,
                       uint32_t ext_clean = 0, bool take_spare = true) {
    const int nv = 1 + r->n_aux;
    Scratch (*scratch)[MAX_VIEWS] = r->scratch;
    if (ext) {
        scratch = const_cast<Scratch (*)[MAX_VIEWS]>(ext);
        P.own_scratch = false;
    }
    for (int sl = 0; sl < 2 && !ext; sl++)
        for (int v = 0; v < nv; v++) {
            const isti_status st = ensure_scratch(r, r->scratch[sl][v], r->views[sl][v].width, r->views[sl][v].height);
            if (st != ISTI_OK) return st;
        }
    std::memset(P.jobs, 0, sizeof(P.jobs));
    P.vp_cam = to_mat(info->viewproj);
    P.vp_main[0] = to_mat(r->views[0][0].viewproj);
    P.vp_main[1] = to_mat(r->views[1][0].viewproj);
    P.init_linear = info->mode == ISTI_MODE_BACKWARD_BIDIRECTIONAL ? 1 : 0;
    P.nj = 0;
    for (int v = 0; v < nv; v++) {
        for (int sl = 0; sl < 2; sl++) {
            const isti_view_desc &d = r->views[sl][v];
            const isti_view_desc &other = r->views[1 - sl][v];
            Scratch &sc = scratch[sl][v];
            P.clean[P.nj] = ext ? ((ext_clean >> (sl * MAX_VIEWS + v)) & 1u) != 0 : take_spare && take_clean(sc);
            SplatJob &J = P.jobs[P.nj++];
            const bool hm = has_motion(d);
            J.z = d.z;
            J.mx = hm ? d.motion_x : nullptr;
            J.my = hm ? d.motion_y : nullptr;
            J.mz = hm ? d.motion_z : nullptr;
            J.full_ff = sc.full_ff;
            J.cam_ff = sc.cam_ff;
            J.splat = sc.splat;
            J.w = d.width;
            J.h = d.height;
            J.n_px = d.width * d.height;
            J.weight = sl == 0 ? info->alpha : 1.0f - info->alpha;
            J.main_slot = v == 0 ? -1 : sl;  // only aux views' visible-in-main flags are read
            J.inv_w = 1.0 / (double)d.width;
            J.rcp_w = (float)(1.0 / (double)d.width);
            J.write_full = (sc.full_ok && (!ext || ext_full)) ? 0 : 1;
            J.inv_h = 1.0 / (double)d.height;
            J.inv_vp_from = to_mat(d.inv_viewproj);
            J.vp_to = to_mat(other.viewproj);
        }
    }
    ResolveArgs &ra = P.ra;
    std::memset(&ra, 0, sizeof(ra));
    for (int v = 0; v < nv; v++) {
        const int prio_f = v == 0 ? 0 : r->views[0][v].priority;  // main priority is 0 (renderer.cu:654)
        const int prio_b = v == 0 ? 0 : r->views[1][v].priority;
        ra.fwd[v] = resolve_view(r->views[0][v], scratch[0][v], true, prio_f);
        ra.bwd[v] = resolve_view(r->views[1][v], scratch[1][v], false, prio_b);
        set_ratio(ra.fwd[v], info->width, info->height);
        set_ratio(ra.bwd[v], info->width, info->height);
    }
    ra.out = reinterpret_cast<uchar4 *>(out);
    ra.out_pitch_px = (uint32_t)(out_pitch / 4);
    ra.w = info->width;
    ra.h = info->height;
    ra.alpha = info->alpha;
    ra.fpa = info->frames_per_anchor;
    ra.sfpa = info->shadow_frames_per_anchor;
    ra.no_shadows = info->no_shadows;
    ra.n_aux = r->n_aux;
    int iters = 3;
    P.window = 1;
    P.always_valid = false;
    switch (info->mode) {  // src/renderer.cu:676-693
        case ISTI_MODE_OURS:
        case ISTI_MODE_BACKWARD_UNIDIRECTIONAL: iters = 3; break;
        case ISTI_MODE_SPLATTING: iters = 0; break;
        case ISTI_MODE_BACKWARD_BIDIRECTIONAL: P.window = 3; iters = 5; P.always_valid = true; break;
        default: break;
    }
    ra.max_iter = info->max_search_iterations >= 0 ? info->max_search_iterations : iters;
    ra.tiles_x = (ra.w + TILE_W - 1) / TILE_W;
    ra.n_tiles = ra.tiles_x * ((ra.h + TILE_H - 1) / TILE_H);
    ra.q_out = r->q_out;
    return ISTI_OK;
}

// Tile grid of a resolve launch for tile shape `shape`.
void set_tiles(ResolveArgs &ra, int shape) {
    const uint32_t tw = TILE_SHAPE_W[shape], th = TILE_SHAPE_H[shape];
    ra.tiles_x = (ra.w + tw - 1) / tw;
    ra.n_tiles = ra.tiles_x * ((ra.h + th - 1) / th);
}

int tile_shape_for(const isti_renderer *r, const FramePlan &P) {
    if (r->tile_shape >= 0) return r->tile_shape;
    std::lock_guard<std::mutex> lock(g_tuned_mutex);
    const auto it = g_tuned.find(TuneKey{P.ra.w, P.ra.h, P.ra.n_aux, P.window});
    return it == g_tuned.end() ? 0 : it->second;
}

bool splat_binned_for(const isti_renderer *r, const FramePlan &P) {
    if (r->splat_mode >= 0) return r->splat_mode == 1;
    std::lock_guard<std::mutex> lock(g_tuned_mutex);
    const auto it = g_tuned_splat.find(TuneKey{P.ra.w, P.ra.h, P.ra.n_aux, P.window});
    return it != g_tuned_splat.end() && it->second == 1;
}

// M1 clear and K1 reprojection + splat of every view of a planned frame
// (src/renderer.cu:631-649), with the stage-timing marks 0, 1 and 2.
// After a splat launch: full_ff of every view holds the current pair's flow.
void mark_full_written(isti_renderer *r) {
    for (int sl = 0; sl < 2; sl++)
        for (int v = 0; v <= r->n_aux; v++) r->scratch[sl][v].full_ok = true;
}

isti_status splat_stage(isti_renderer *r, FramePlan &P, hipStream_t s, bool binned) {
    ClearArgs ca;
    std::memset(&ca, 0, sizeof(ca));
    SplatArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.vp_cam = P.vp_cam;
    sa.vp_main[0] = P.vp_main[0];
    sa.vp_main[1] = P.vp_main[1];
    sa.init_linear = P.init_linear;
    uint32_t clear_blocks = 0, n_splat_blocks = 0;
    int nc = 0;
    for (int j = 0; j < P.nj; j++) {
        SplatJob &J = sa.job[j];
        J = P.jobs[j];
        if (!P.clean[j]) {
            ca.buf[nc] = J.splat;
            ca.n[nc] = J.n_px;
            ca.block_begin[nc] = clear_blocks;
            clear_blocks += (J.n_px + 511u) / 512u;
            nc++;
        }
        J.block_begin = n_splat_blocks;
        J.binned = binned ? 1 : 0;
        n_splat_blocks += splat_blocks(J.w, J.h, binned);
    }
    ca.n_bufs = nc;
    ca.n_blocks = clear_blocks;
    sa.n_jobs = P.nj;
    sa.n_blocks = n_splat_blocks;
    hipError_t e;
    if ((e = mark(r, 0, s)) != hipSuccess) return dev_fail(r, e, "event");
    if (clear_blocks && (e = launch_clear(ca, s)) != hipSuccess) return dev_fail(r, e, "clear launch");
    if ((e = mark(r, 1, s)) != hipSuccess) return dev_fail(r, e, "event");
    if ((e = launch_splat(sa, binned, s)) != hipSuccess) return dev_fail(r, e, "splat launch");
    for (int j = 0; j < P.nj; j++) P.clean[j] = false;
    if (P.own_scratch) mark_full_written(r);
    if ((e = mark(r, 2, s)) != hipSuccess) return dev_fail(r, e, "event");
    return ISTI_OK;
}

// Points a plain resolve launch of tile shape `shape` at the spare splat
// buffers of the views in `set` that are not clear yet, allocating them on
// first use. The launch clears them after its own work (ClearAhead); call
// ahead_cleared once it is enqueued. The spare of a view was last splatted by
// an earlier frame on the same stream, whose resolve precedes this one.
isti_status plan_ahead(isti_renderer *r, Scratch (*set)[MAX_VIEWS], ResolveArgs &ra, int shape) {
    ClearAhead &c = ra.ahead;
    std::memset(&c, 0, sizeof(c));
    for (int sl = 0; sl < 2; sl++)
        for (int v = 0; v <= r->n_aux; v++) {
            Scratch &sc = set[sl][v];
            if (sc.spare_clean || !sc.splat) continue;
            const size_t n = (size_t)sc.w * sc.h;
            if (!sc.spare) {
                const hipError_t e = hipMalloc(&sc.spare, n * sizeof(unsigned long long));
                if (e != hipSuccess) {
                    sc.spare = nullptr;
                    return dev_fail(r, e, "hipMalloc(spare splat buffer)");
                }
            }
            c.buf[c.n_bufs] = sc.spare;
            c.n[c.n_bufs] = (uint32_t)n;
            c.n_bufs++;
        }
    // fewest runs per workgroup that fit the clear into the resolve grid; a
    // grid with fewer workgroups than buffers (each buffer takes at least
    // one), or one that would need more than MAX_ROUNDS runs per workgroup
    // (a small render of large keyframes), leaves the clear to the next
    // frame's clear kernel
    constexpr uint64_t MAX_ROUNDS = 16;
    const uint64_t per_round = 2ull * TILE_SHAPE_W[shape] * TILE_SHAPE_H[shape];  // cells
    uint64_t cells = 0;
    for (int j = 0; j < c.n_bufs; j++) cells += c.n[j];
    uint64_t rounds = c.n_bufs && (uint64_t)c.n_bufs <= ra.n_tiles
                          ? (cells + per_round * ra.n_tiles - 1) / (per_round * ra.n_tiles)
                          : MAX_ROUNDS + 1;
    for (; rounds <= MAX_ROUNDS; rounds++) {
        uint64_t blocks = 0;
        for (int j = 0; j < c.n_bufs; j++) blocks += (c.n[j] + per_round * rounds - 1) / (per_round * rounds);
        if (blocks <= ra.n_tiles) break;
    }
    if (rounds > MAX_ROUNDS) {
        std::memset(&c, 0, sizeof(c));
        return ISTI_OK;
    }
    uint32_t b = 0;
    for (int j = 0; j < c.n_bufs; j++) {
        c.block_begin[j] = b;
        b += (uint32_t)((c.n[j] + per_round * rounds - 1) / (per_round * rounds));
    }
    c.n_blocks = b;
    c.rounds = (uint32_t)rounds;
    return ISTI_OK;
}

void ahead_cleared(isti_renderer *r, Scratch (*set)[MAX_VIEWS], const ResolveArgs &ra) {
    if (ra.ahead.n_bufs == 0) return;
    for (int sl = 0; sl < 2; sl++)
        for (int v = 0; v <= r->n_aux; v++)
            if (set[sl][v].spare) set[sl][v].spare_clean = true;
}

// Smallest power-of-two reduction of the render size that fits the network grid.
bool net_shift_for(uint32_t w, uint32_t h, uint32_t net_w, uint32_t net_h, uint32_t &shift) {
    for (uint32_t k = 0; k <= 4; k++) {
        const uint32_t cw = (w + (1u << k) - 1) >> k, ch = (h + (1u << k) - 1) >> k;
        if (cw <= net_w && ch <= net_h) {
            shift = k;
            return true;
        }
    }
    return false;
}

constexpr size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

// Double-buffered device work tables for batched renders (see isti.h).
struct isti_batch {
    int device = 0;
    int cap_req = 0, cap_jobs = 0;
    size_t bytes = 0;
    uint8_t *host[2] = {nullptr, nullptr};  // pinned staging
    uint8_t *dev[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool in_use[2] = {false, false};
    int slot = 0;
    bool graph = false;
    uint32_t grid_light = 0;
    uint32_t grid_resolve[N_VARIANTS] = {};
    std::map<std::pair<int, uint32_t>, hipGraphExec_t> graphs;  // (slot, variant mask)
    hipStream_t capture = nullptr;
    std::vector<FramePlan> plans;
    std::string err;
};

namespace {

isti_status bfail(isti_batch *b, isti_status st, const std::string &msg) {
    if (b) b->err = msg;
    return st;
}

// Launches the stage sequence of one table slot (or records it into a capture).
hipError_t launch_batch_stages(isti_batch *b, const BatchTable *t, uint32_t mask, hipStream_t s) {
    hipError_t e = launch_clear_batch(t, b->grid_light, s);
    if (e == hipSuccess) e = launch_splat_batch(t, b->grid_light, s);
    for (int v = 0; v < N_VARIANTS && e == hipSuccess; v++)
        if (mask & (1u << v)) e = launch_resolve_batch(t, v, b->grid_resolve[v], s);
    return e;
}

}  // namespace

extern "C" {

isti_status isti_batch_create(int32_t device, int32_t max_requests, isti_batch **out) {
    if (!out || max_requests <= 0 || max_requests > 4096) {
        g_create_error = "out must be non-null and max_requests in 1..4096";
        return ISTI_ERR_INVALID_ARG;
    }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        g_create_error = "no such HIP device";
        return ISTI_ERR_DEVICE;
    }
    isti_batch *b = new (std::nothrow) isti_batch();
    if (!b) return ISTI_ERR_OOM;
    b->device = device;
    b->cap_req = max_requests;
    b->cap_jobs = max_requests * MAX_SPLAT_JOBS;
    b->bytes = align_up(sizeof(BatchTable), 256) + align_up(sizeof(BatchRequest) * b->cap_req, 256) +
               align_up(sizeof(SplatJob) * b->cap_jobs, 256) + 3 * align_up(sizeof(uint32_t) * (b->cap_jobs + 1), 256);
    hipError_t e = hipSetDevice(device);
    for (int i = 0; i < 2 && e == hipSuccess; i++) {
        e = hipHostMalloc(&b->host[i], b->bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&b->dev[i], b->bytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        g_create_error = std::string("batch setup: ") + hipGetErrorString(e);
        isti_batch_destroy(b);
        return e == hipErrorOutOfMemory ? ISTI_ERR_OOM : ISTI_ERR_DEVICE;
    }
    const uint32_t ncu = (uint32_t)(cus > 0 ? cus : 1);
    b->grid_light = std::max<uint32_t>(8, ncu * 8 / 8 * 8);
    for (int v = 0; v < N_VARIANTS; v++)
        b->grid_resolve[v] = std::max<uint32_t>(8, ncu * (uint32_t)resolve_batch_blocks_per_cu(v) / 8 * 8);
    b->plans.resize(b->cap_req);
    *out = b;
    return ISTI_OK;
}

isti_status isti_batch_destroy(isti_batch *b) {
    if (!b) return ISTI_ERR_INVALID_ARG;
    (void)hipSetDevice(b->device);
    for (int i = 0; i < 2; i++) {
        if (b->done[i]) (void)hipEventSynchronize(b->done[i]);
    }
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
    for (int i = 0; i < 2; i++) {
        if (b->host[i]) (void)hipHostFree(b->host[i]);
        if (b->dev[i]) (void)hipFree(b->dev[i]);
        if (b->done[i]) (void)hipEventDestroy(b->done[i]);
    }
    if (b->capture) (void)hipStreamDestroy(b->capture);
    delete b;
    return ISTI_OK;
}

isti_status isti_batch_set_graph(isti_batch *b, int32_t enable) {
    if (!b) return ISTI_ERR_INVALID_ARG;
    b->graph = enable != 0;
    return ISTI_OK;
}

isti_status isti_render_batch(isti_batch *b, isti_renderer *const *rs, const isti_render_info *infos,
                              uint8_t *const *outs, const size_t *pitches, int32_t n, void *stream) {
    if (!b) return ISTI_ERR_INVALID_ARG;
    if (n == 0) return ISTI_OK;
    if (n < 0 || n > b->cap_req || !rs || !infos || !outs || !pitches)
        return bfail(b, ISTI_ERR_INVALID_ARG, "need 0 <= n <= max_requests and non-null arrays");
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return bfail(b, ISTI_ERR_DEVICE, hipGetErrorString(e));
    // validate and plan every request
    std::vector<int> var(n);
    for (int i = 0; i < n; i++) {
        isti_renderer *r = rs[i];
        if (!r || !outs[i]) return bfail(b, ISTI_ERR_INVALID_ARG, "null renderer or output in request " + std::to_string(i));
        if (r->device != b->device) return bfail(b, ISTI_ERR_INVALID_ARG, "renderer on another device");
        for (int k = 0; k < i; k++)
            if (rs[k] == r) return bfail(b, ISTI_ERR_INVALID_ARG, "a renderer appears twice in one batch");
        if (infos[i].mode == ISTI_MODE_TIMEWARP) return bfail(b, ISTI_ERR_INVALID_ARG, "timewarp is not batched");
        isti_status st = validate(r, &infos[i], pitches[i]);
        if (st == ISTI_OK) st = plan_frame(r, &infos[i], outs[i], pitches[i], b->plans[i]);
        if (st != ISTI_OK) return bfail(b, st, "request " + std::to_string(i) + ": " + r->err);
        var[i] = resolve_variant(b->plans[i].always_valid, r->n_aux);
        // each request splats with its renderer's variant (forced or tuned for its size)
        const int32_t binned = splat_binned_for(r, b->plans[i]) ? 1 : 0;
        for (int k = 0; k < b->plans[i].nj; k++) b->plans[i].jobs[k].binned = binned;
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return var[x] < var[y]; });

    // the staging buffer of this slot may still feed an earlier batch's copy
    const int sl = b->slot;
    if (b->in_use[sl]) {
        e = hipEventSynchronize(b->done[sl]);
        if (e != hipSuccess) return bfail(b, ISTI_ERR_DEVICE, hipGetErrorString(e));
    }
    uint8_t *h = b->host[sl];
    uint8_t *d = b->dev[sl];
    int nj = 0;
    for (int i = 0; i < n; i++) nj += b->plans[i].nj;
    size_t off = align_up(sizeof(BatchTable), 256);
    const size_t off_req = off;
    off = align_up(off + sizeof(BatchRequest) * n, 256);
    const size_t off_jobs = off;
    off = align_up(off + sizeof(SplatJob) * nj, 256);
    const size_t off_clear = off;
    off = align_up(off + sizeof(uint32_t) * (nj + 1), 256);
    const size_t off_splat = off;
    off = align_up(off + sizeof(uint32_t) * (nj + 1), 256);
    const size_t off_tile = off;
    const size_t used = off + sizeof(uint32_t) * (n + 1);

    BatchTable *T = reinterpret_cast<BatchTable *>(h);
    std::memset(T, 0, sizeof(BatchTable));
    BatchRequest *req = reinterpret_cast<BatchRequest *>(h + off_req);
    SplatJob *jobs = reinterpret_cast<SplatJob *>(h + off_jobs);
    uint32_t *clear_begin = reinterpret_cast<uint32_t *>(h + off_clear);
    uint32_t *splat_begin = reinterpret_cast<uint32_t *>(h + off_splat);
    uint32_t *tile_begin = reinterpret_cast<uint32_t *>(h + off_tile);
    uint32_t mask = 0, cb = 0, sb = 0, tb = 0;
    int j = 0;
    for (int q = 0; q < n; q++) {
        const FramePlan &P = b->plans[order[q]];
        BatchRequest &R = req[q];
        std::memset(&R, 0, sizeof(R));
        R.args = P.ra;
        set_tiles(R.args, BATCH_TILE_SHAPE);  // the batched resolve's tile grid
        // (no clear ahead here: clearing the next batch's buffers in the
        // batched resolve measured slower, 18.3 -> 15.6 Gpx/s on the stereo
        // frames, and so did the clear's code alone compiled into that
        // kernel, unused; the batched resolve is built without it)
        R.vp_cam = P.vp_cam;
        R.vp_main[0] = P.vp_main[0];
        R.vp_main[1] = P.vp_main[1];
        R.init_linear = P.init_linear;
        tile_begin[q] = tb;
        tb += R.args.n_tiles;
        mask |= 1u << var[order[q]];
        for (int k = 0; k < P.nj; k++, j++) {
            jobs[j] = P.jobs[k];
            jobs[j].req = q;
            clear_begin[j] = cb;
            if (!P.clean[k]) cb += (P.jobs[k].n_px + 511u) / 512u;  // cleared ahead: no blocks
            splat_begin[j] = sb;
            sb += splat_blocks(P.jobs[k].w, P.jobs[k].h, P.jobs[k].binned != 0);
        }
    }
    clear_begin[nj] = cb;
    splat_begin[nj] = sb;
    tile_begin[n] = tb;
    T->n_req = (uint32_t)n;
    T->n_jobs = (uint32_t)nj;
    T->clear_blocks = cb;
    T->splat_blocks = sb;
    for (int v = 0, q = 0; v <= N_VARIANTS; v++) {
        while (q < n && var[order[q]] < v) q++;
        T->var_req_begin[v] = (uint32_t)q;
    }
    T->req = reinterpret_cast<const BatchRequest *>(d + off_req);
    T->jobs = reinterpret_cast<const SplatJob *>(d + off_jobs);
    T->clear_begin = reinterpret_cast<const uint32_t *>(d + off_clear);
    T->splat_begin = reinterpret_cast<const uint32_t *>(d + off_splat);
    T->tile_begin = reinterpret_cast<const uint32_t *>(d + off_tile);

    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    e = hipMemcpyAsync(d, h, used, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return bfail(b, ISTI_ERR_DEVICE, std::string("table copy: ") + hipGetErrorString(e));
    const BatchTable *td = reinterpret_cast<const BatchTable *>(d);
    if (b->graph) {
        const auto key = std::make_pair(sl, mask);
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ex = nullptr;
            e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                const hipError_t le = launch_batch_stages(b, td, mask, b->capture);
                e = hipStreamEndCapture(b->capture, &g);
                if (e == hipSuccess) e = le;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) return bfail(b, ISTI_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
            it = b->graphs.emplace(key, ex).first;
        }
        e = hipGraphLaunch(it->second, s);
    } else {
        e = launch_batch_stages(b, td, mask, s);
    }
    if (e != hipSuccess) return bfail(b, ISTI_ERR_DEVICE, std::string("batch launch: ") + hipGetErrorString(e));
    for (int i = 0; i < n; i++) mark_full_written(rs[i]);
    e = hipEventRecord(b->done[sl], s);
    if (e != hipSuccess) return bfail(b, ISTI_ERR_DEVICE, hipGetErrorString(e));
    b->in_use[sl] = true;
    b->slot = 1 - sl;
    return ISTI_OK;
}

isti_status isti_set_search_capture(isti_renderer *r, uint32_t *q_out) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    r->q_out = q_out;
    return ISTI_OK;
}

isti_status isti_learned_features(isti_renderer *r, const isti_render_info *info, uint16_t *x, uint32_t net_w,
                                  uint32_t net_h, void *stream) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    if (!info || net_w == 0 || net_h == 0) return fail(r, ISTI_ERR_INVALID_ARG, "bad features arguments");
    isti_status st = validate(r, info, (size_t)info->width * 4);
    if (st != ISTI_OK) return st;
    if (info->mode == ISTI_MODE_TIMEWARP || info->mode == ISTI_MODE_BACKWARD_BIDIRECTIONAL)
        return fail(r, ISTI_ERR_INVALID_ARG, "learned starts apply to modes 0, 2 and 3");
    FeatureArgs fa;
    std::memset(&fa, 0, sizeof(fa));
    if (!net_shift_for(info->width, info->height, net_w, net_h, fa.net_shift))
        return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE, "network grid too small for the render size");
    hipError_t e = hipSetDevice(r->device);
    if (e != hipSuccess) return dev_fail(r, e, "hipSetDevice");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    FramePlan P;
    st = plan_frame(r, info, nullptr, (size_t)info->width * 4, P);
    if (st != ISTI_OK) return st;
    r->prepared_valid = false;
    st = splat_stage(r, P, s, splat_binned_for(r, P));
    if (st != ISTI_OK) return st;
    if (x) {  // no x: the splat stage alone (a network output of an earlier frame is reused)
        for (int sl = 0; sl < 2; sl++) fa.v[sl] = sl == 0 ? P.ra.fwd[0] : P.ra.bwd[0];
        fa.x = x;
        fa.w = info->width;
        fa.h = info->height;
        fa.net_w = net_w;
        fa.net_h = net_h;
        fa.alpha = info->alpha;
        e = launch_features(fa, s);
        if (e != hipSuccess) return dev_fail(r, e, "features launch");
    }
    r->prepared = *info;
    r->prepared_valid = true;
    return ISTI_OK;
}

isti_status isti_learned_loss(isti_renderer *r, const float *out, const uint32_t *q_target, uint32_t net_w,
                              uint32_t net_h, float conf_weight, float *dout, float *loss_acc, void *stream) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    if (!out || !q_target || !dout || !loss_acc) return fail(r, ISTI_ERR_INVALID_ARG, "bad loss arguments");
    if (!r->prepared_valid) return fail(r, ISTI_ERR_INVALID_ARG, "call isti_learned_features for this frame first");
    const uint32_t w = r->prepared.width, h = r->prepared.height;
    uint32_t shift = 0;
    if (!net_shift_for(w, h, net_w, net_h, shift))
        return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE, "network grid does not match the render size");
    const hipError_t e = launch_learned_loss(out, q_target, r->scratch[0][0].cam_ff, r->scratch[1][0].cam_ff, w, h,
                                             net_w, net_h, shift, conf_weight, dout, loss_acc,
                                             reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? ISTI_OK : dev_fail(r, e, "loss launch");
}

isti_status isti_render_learned(isti_renderer *r, const isti_render_info *info, const float *net_out,
                                uint32_t net_w, uint32_t net_h, float conf_threshold, uint32_t *stats,
                                uint8_t *out, size_t out_pitch, void *stream) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    if (!info || !out || !net_out) return fail(r, ISTI_ERR_INVALID_ARG, "info, out and net_out must be non-null");
    if (!(conf_threshold > 0.0f && conf_threshold <= 1.0f))
        return fail(r, ISTI_ERR_INVALID_ARG, "conf_threshold must be in (0, 1]");
    isti_status st = validate(r, info, out_pitch);
    if (st != ISTI_OK) return st;
    if (!r->prepared_valid || std::memcmp(&r->prepared, info, sizeof(*info)) != 0)
        return fail(r, ISTI_ERR_INVALID_ARG, "call isti_learned_features for this frame first");
    hipError_t e = hipSetDevice(r->device);
    if (e != hipSuccess) return dev_fail(r, e, "hipSetDevice");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    FramePlan P;
    // the buffers isti_learned_features splatted (a cleared spare stays the
    // spare: this frame may be resolved again)
    st = plan_frame(r, info, out, out_pitch, P, nullptr, false, 0, false);
    if (st != ISTI_OK) return st;
    if (!net_shift_for(info->width, info->height, net_w, net_h, P.ra.net_shift))
        return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE, "network grid too small for the render size");
    P.ra.net_out = net_out;
    P.ra.net_w = net_w;
    P.ra.net_h = net_h;
    // confidence threshold as a logit; 1 disables every learned start
    P.ra.conf_logit = conf_threshold < 1.0f ? std::log(conf_threshold / (1.0f - conf_threshold))
                                            : std::numeric_limits<float>::infinity();
    if (stats && !r->stat_slots) {  // first frame with statistics: zeroed slot rows
        const size_t bytes = (size_t)STAT_SLOTS * STAT_STRIDE * sizeof(uint32_t);
        if ((e = hipMalloc(&r->stat_slots, bytes)) != hipSuccess) return dev_fail(r, e, "hipMalloc(stat slots)");
        if ((e = hipMemsetAsync(r->stat_slots, 0, bytes, s)) != hipSuccess) return dev_fail(r, e, "hipMemsetAsync(stat slots)");
    }
    P.ra.learned_stats = stats ? r->stat_slots : nullptr;
    // (no clear ahead: in the learned-start kernel it measured no faster, C4
    // learned path 0.550x -> 0.543x of plain, headline workload 8.05 -> 8.07 Gpx/s)
    if ((e = mark(r, 2, s)) == hipSuccess) e = launch_resolve_learned(P.ra, s);
    if (e == hipSuccess && stats) e = launch_stats_reduce(r->stat_slots, stats, s);
    if (e == hipSuccess) e = mark(r, 3, s);
    return e == hipSuccess ? ISTI_OK : dev_fail(r, e, "learned render launch");
}

const char *isti_batch_last_error(const isti_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }



isti_status isti_create(const isti_config *config, isti_renderer **out) {
    if (!config || !out) {
        g_create_error = "config and out must be non-null";
        return ISTI_ERR_INVALID_ARG;
    }
    *out = nullptr;
    if (config->n_aux < 0 || config->n_aux > ISTI_MAX_AUX) {
        g_create_error = "n_aux must be 0..4";
        return ISTI_ERR_INVALID_ARG;
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        g_create_error = std::string("no HIP device: ") + hipGetErrorString(e);
        return ISTI_ERR_DEVICE;
    }
    if (config->device < 0 || config->device >= count) {
        g_create_error = "device ordinal out of range";
        return ISTI_ERR_INVALID_ARG;
    }
    isti_renderer *r = new (std::nothrow) isti_renderer();
    if (!r) {
        g_create_error = "out of host memory";
        return ISTI_ERR_OOM;
    }
    r->device = config->device;
    r->n_aux = config->n_aux;
    std::memset(r->views, 0, sizeof(r->views));
    std::memset(r->view_set, 0, sizeof(r->view_set));
    *out = r;
    return ISTI_OK;
}

isti_status isti_destroy(isti_renderer *r) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    (void)hipSetDevice(r->device);
    for (hipStream_t &ps : r->pipe)
        if (ps) {
            (void)hipStreamSynchronize(ps);
            (void)hipStreamDestroy(ps);
        }
    for (hipEvent_t &ev : r->pipe_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto &slot : r->scratch2)
        for (auto &s : slot) free_scratch2(s);
    for (auto &slot : r->scratch)
        for (auto &s : slot) free_scratch(s);
    if (r->row_sums) (void)hipFree(r->row_sums);
    if (r->mean_z) (void)hipFree(r->mean_z);
    if (r->stat_slots) (void)hipFree(r->stat_slots);
    for (auto &set : r->ring.ev)
        for (auto &ev : set)
            if (ev) (void)hipEventDestroy(ev);
    delete r;
    return ISTI_OK;
}

isti_status isti_keyframe_set(isti_renderer *r, int32_t slot, int32_t view, const isti_view_desc *d) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    r->prepared_valid = false;
    if (!d || slot < 0 || slot > 1 || view < 0 || view > r->n_aux)
        return fail(r, ISTI_ERR_INVALID_ARG, "slot must be 0/1 and view 0..n_aux, desc non-null");
    if (d->width == 0 || d->height == 0 || d->width > ISTI_MAX_DIM || d->height > ISTI_MAX_DIM)
        return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE, "view size must be 1..65535 per axis");
    if ((uint64_t)d->width * d->height > ISTI_MAX_VIEW_PIXELS)
        return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE, "a view holds at most 2^28 - 1 pixels");
    if (!d->color || !d->z) return fail(r, ISTI_ERR_INVALID_ARG, "color and z planes are required");
    const int nm = (d->motion_x != nullptr) + (d->motion_y != nullptr) + (d->motion_z != nullptr);
    if (nm != 0 && nm != 3) return fail(r, ISTI_ERR_INVALID_ARG, "motion planes: all three or none");
    r->views[slot][view] = *d;
    r->view_set[slot][view] = true;
    // the flows of this view and of its partner in the other slot (whose
    // target this view is) must be recomputed
    r->scratch[0][view].full_ok = false;
    r->scratch[1][view].full_ok = false;
    return ISTI_OK;
}

isti_status isti_keyframe_swap(isti_renderer *r) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    r->prepared_valid = false;
    for (int v = 0; v < MAX_VIEWS; v++) {
        std::swap(r->views[0][v], r->views[1][v]);
        std::swap(r->view_set[0][v], r->view_set[1][v]);
        r->scratch[0][v].full_ok = r->scratch[1][v].full_ok = false;
    }
    return ISTI_OK;
}

}  // extern "C"

namespace {
size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }
}  // namespace

int isti_internal_device(const isti_renderer *r) { return r->device; }

isti_status isti_internal_fail(isti_renderer *r, isti_status st, const std::string &msg) { return fail(r, st, msg); }

// Scratch bytes a frame of `info` needs when its planes are carved from one
// caller-provided block (each plane 256-byte aligned).
size_t isti_internal_scratch_bytes(const isti_renderer *r, const isti_render_info *info) {
    if (info->mode == ISTI_MODE_TIMEWARP) return align256(sizeof(double) * r->views[0][0].height) + 256;
    size_t bytes = 0;
    for (int sl = 0; sl < 2; sl++)
        for (int v = 0; v <= r->n_aux; v++) {
            const size_t n = (size_t)r->views[sl][v].width * r->views[sl][v].height;
            bytes += align256(n * sizeof(unsigned long long)) + align256(n * sizeof(float4)) +
                     align256(n * sizeof(Flow3));
        }
    return bytes;
}

// Renderer::render with the renderer's own scratch (ext_base null) or with
// every scratch plane carved from ext_bytes of device memory at ext_base, in
// which case nothing is allocated.
isti_status isti_internal_render(isti_renderer *r, const isti_render_info *info, uint8_t *out, size_t out_pitch,
                                 void *stream, void *ext_base, size_t ext_bytes) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    if (!info || !out) return fail(r, ISTI_ERR_INVALID_ARG, "info and out must be non-null");
    isti_status st = validate(r, info, out_pitch);
    if (st != ISTI_OK) return st;
    if (ext_base) {
        const size_t need = isti_internal_scratch_bytes(r, info);
        if (need > ext_bytes)
            return fail(r, ISTI_ERR_UNSUPPORTED_SHAPE,
                        "frame needs " + std::to_string(need) + " bytes of scratch, more than the " +
                            std::to_string(ext_bytes) + " available");
    }
    hipError_t e = hipSetDevice(r->device);
    if (e != hipSuccess) return dev_fail(r, e, "hipSetDevice");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char *base = static_cast<char *>(ext_base);

    if (info->mode == ISTI_MODE_TIMEWARP) {
        if (!base) return render_timewarp(r, info, out, out_pitch, s);
        double *rows = reinterpret_cast<double *>(base);
        float *mean = reinterpret_cast<float *>(base + align256(sizeof(double) * r->views[0][0].height));
        return render_timewarp(r, info, out, out_pitch, s, rows, mean);
    }

    FramePlan P;
    Scratch carved[2][MAX_VIEWS];
    if (base) {
        size_t off = 0;
        for (int sl = 0; sl < 2; sl++)
            for (int v = 0; v <= r->n_aux; v++) {
                Scratch &c = carved[sl][v];
                c.w = r->views[sl][v].width;
                c.h = r->views[sl][v].height;
                const size_t n = (size_t)c.w * c.h;
                c.splat = reinterpret_cast<unsigned long long *>(base + off);
                off += align256(n * sizeof(unsigned long long));
                c.cam_ff = reinterpret_cast<float4 *>(base + off);
                off += align256(n * sizeof(float4));
                c.full_ff = reinterpret_cast<Flow3 *>(base + off);
                off += align256(n * sizeof(Flow3));
            }
    }
    st = plan_frame(r, info, out, out_pitch, P, base ? carved : nullptr);
    if (st != ISTI_OK) return st;
    const int shape = tile_shape_for(r, P);
    set_tiles(P.ra, shape);
    r->prepared_valid = false;
    st = splat_stage(r, P, s, splat_binned_for(r, P));
    if (st != ISTI_OK) return st;
    if (!base && (st = plan_ahead(r, r->scratch, P.ra, shape)) != ISTI_OK) return st;
    e = launch_resolve(P.ra, P.window, P.always_valid, shape, s);
    if (e != hipSuccess) return dev_fail(r, e, "resolve launch");
    if (!base) ahead_cleared(r, r->scratch, P.ra);
    if ((e = mark(r, 3, s)) != hipSuccess) return dev_fail(r, e, "event");
    return ISTI_OK;
}

extern "C" {

isti_status isti_render(isti_renderer *r, const isti_render_info *info, uint8_t *out, size_t out_pitch,
                        void *stream) {
    return isti_internal_render(r, info, out, out_pitch, stream, nullptr, 0);
}

isti_status isti_render_frames(isti_renderer *r, const isti_render_info *infos, uint32_t n, uint8_t *const *outs,
                               const size_t *pitches, void *stream) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    if (!infos || !outs || !pitches || n == 0 || n > ISTI_MAX_FRAMES)
        return fail(r, ISTI_ERR_INVALID_ARG, "infos, outs and pitches must be non-null and n in 1..ISTI_MAX_FRAMES");
    for (uint32_t i = 0; i < n; i++) {
        if (!outs[i]) return fail(r, ISTI_ERR_INVALID_ARG, "outs[i] must be non-null");
        if (infos[i].mode == ISTI_MODE_TIMEWARP)
            return fail(r, ISTI_ERR_INVALID_ARG, "timewarp frames render with isti_render");
        const isti_status st = validate(r, &infos[i], pitches[i]);
        if (st != ISTI_OK) return st;
        for (uint32_t j = 0; j < i; j++)
            if (outs[j] == outs[i]) return fail(r, ISTI_ERR_INVALID_ARG, "the frames' outputs must be distinct");
    }
    if (n == 1) return isti_internal_render(r, &infos[0], outs[0], pitches[0], stream, nullptr, 0);
    hipError_t e = hipSetDevice(r->device);
    if (e != hipSuccess) return dev_fail(r, e, "hipSetDevice");
    for (hipStream_t &ps : r->pipe)  // default priority, like every stream here
        if (!ps && (e = hipStreamCreateWithFlags(&ps, hipStreamNonBlocking)) != hipSuccess)
            return dev_fail(r, e, "hipStreamCreateWithFlags");
    for (hipEvent_t &ev : r->pipe_ev)
        if (!ev && (e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
            return dev_fail(r, e, "hipEventCreateWithFlags");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if ((e = hipEventRecord(r->pipe_ev[0], s)) != hipSuccess) return dev_fail(r, e, "hipEventRecord");
    for (hipStream_t ps : r->pipe)
        if ((e = hipStreamWaitEvent(ps, r->pipe_ev[0], 0)) != hipSuccess) return dev_fail(r, e, "hipStreamWaitEvent");
    const bool timing = r->timing;
    r->timing = false;
    r->prepared_valid = false;
    isti_status st = ISTI_OK;
    bool flow_new = false;  // the first frame writes the pair's flow toward the other keyframe
    for (uint32_t i = 0; i < n && st == ISTI_OK; i++) {
        const int k = (int)(i & 1u);
        FramePlan P;
        if (k == 0) {
            st = plan_frame(r, &infos[i], outs[i], pitches[i], P);
        } else {
            Scratch set2[2][MAX_VIEWS];
            uint32_t clean = 0;
            for (int sl = 0; sl < 2 && st == ISTI_OK; sl++)
                for (int v = 0; v <= r->n_aux && st == ISTI_OK; v++) {
                    Scratch &own = r->scratch[sl][v], &b = r->scratch2[sl][v];
                    st = ensure_scratch2(r, b, own.w, own.h);
                    if (st == ISTI_OK && take_clean(b)) clean |= 1u << (sl * MAX_VIEWS + v);
                    set2[sl][v] = b;
                    set2[sl][v].full_ff = own.full_ff;
                    set2[sl][v].full_ok = own.full_ok;
                }
            if (st == ISTI_OK) st = plan_frame(r, &infos[i], outs[i], pitches[i], P, set2, true, clean);
        }
        if (st != ISTI_OK) break;
        if (i == 0)
            for (int j = 0; j < P.nj; j++) flow_new = flow_new || P.jobs[j].write_full != 0;
        const int shape = tile_shape_for(r, P);
        set_tiles(P.ra, shape);
        st = splat_stage(r, P, r->pipe[k], splat_binned_for(r, P));
        if (st != ISTI_OK) break;
        if (i == 0 && flow_new && (e = hipEventRecord(r->pipe_ev[1], r->pipe[0])) != hipSuccess) {
            st = dev_fail(r, e, "hipEventRecord");
            break;
        }
        // the other stream's first resolve reads the flow the first frame wrote
        if (i == 1 && flow_new && (e = hipStreamWaitEvent(r->pipe[1], r->pipe_ev[1], 0)) != hipSuccess) {
            st = dev_fail(r, e, "hipStreamWaitEvent");
            break;
        }
        Scratch (*set)[MAX_VIEWS] = k == 0 ? r->scratch : r->scratch2;
        if ((st = plan_ahead(r, set, P.ra, shape)) != ISTI_OK) break;
        if ((e = launch_resolve(P.ra, P.window, P.always_valid, shape, r->pipe[k])) != hipSuccess)
            st = dev_fail(r, e, "resolve launch");
        else
            ahead_cleared(r, set, P.ra);
    }
    r->timing = timing;
    // join (also after an error, so the caller's stream never runs ahead of enqueued work)
    for (int k = 0; k < 2; k++) {
        if ((e = hipEventRecord(r->pipe_ev[2 + k], r->pipe[k])) == hipSuccess) e = hipStreamWaitEvent(s, r->pipe_ev[2 + k], 0);
        if (e != hipSuccess && st == ISTI_OK) st = dev_fail(r, e, "join");
    }
    return st;
}

isti_status isti_scratch(const isti_renderer *r, int32_t slot, int32_t view, isti_scratch_view *out) {
    if (!r || !out || slot < 0 || slot > 1 || view < 0 || view > r->n_aux) return ISTI_ERR_INVALID_ARG;
    const Scratch &s = r->scratch[slot][view];
    out->width = s.w;
    out->height = s.h;
    out->splat = reinterpret_cast<const uint64_t *>(s.splat);
    out->cam_ff = reinterpret_cast<const float *>(s.cam_ff);
    out->full_ff = reinterpret_cast<const float *>(s.full_ff);
    return ISTI_OK;
}

isti_status isti_set_tile_shape(isti_renderer *r, int32_t shape) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    if (shape < -1 || shape >= N_TILE_SHAPES) return fail(r, ISTI_ERR_INVALID_ARG, "tile shape must be -1..3");
    r->tile_shape = shape;
    return ISTI_OK;
}

isti_status isti_autotune(isti_renderer *r, const isti_render_info *info, uint8_t *out, size_t out_pitch,
                          int32_t reps, float *times_ms, int32_t *best, int32_t *best_splat, void *stream) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    if (!info || !out || reps < 1 || reps > 1000)
        return fail(r, ISTI_ERR_INVALID_ARG, "info and out must be non-null and reps in 1..1000");
    isti_status st = validate(r, info, out_pitch);
    if (st != ISTI_OK) return st;
    if (info->mode == ISTI_MODE_TIMEWARP) return fail(r, ISTI_ERR_INVALID_ARG, "timewarp has no resolve pass");
    hipError_t e = hipSetDevice(r->device);
    if (e != hipSuccess) return dev_fail(r, e, "hipSetDevice");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    FramePlan P;
    st = plan_frame(r, info, out, out_pitch, P);
    if (st != ISTI_OK) return st;
    r->prepared_valid = false;
    const bool timing = r->timing;
    r->timing = false;  // the tuning launches stay out of the stage statistics
    hipEvent_t ev[2] = {nullptr, nullptr};
    e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    // splat stage (clear + reprojection/splat) with each splat variant. The
    // first warm-up launch writes the pair's flow toward the other keyframe;
    // the timed launches then skip that write, as the pair's later frames do.
    float splat_ms[2] = {};
    for (int v = 0; v < 2 && e == hipSuccess; v++) {
        if ((st = splat_stage(r, P, s, v == 1)) != ISTI_OK) break;  // warm-up
        for (int j = 0; j < P.nj; j++) P.jobs[j].write_full = 0;
        if ((e = hipEventRecord(ev[0], s)) != hipSuccess) break;
        for (int i = 0; i < reps && st == ISTI_OK; i++) st = splat_stage(r, P, s, v == 1);
        if (st != ISTI_OK) break;
        if ((e = hipEventRecord(ev[1], s)) == hipSuccess) e = hipEventSynchronize(ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&splat_ms[v], ev[0], ev[1]);
        splat_ms[v] /= (float)reps;
    }
    const int pick_splat = splat_ms[1] < splat_ms[0] ? 1 : 0;
    if (st == ISTI_OK && e == hipSuccess) st = splat_stage(r, P, s, pick_splat == 1);  // the frame's buffers
    r->timing = timing;
    if (st != ISTI_OK) {
        for (hipEvent_t x : ev)
            if (x) (void)hipEventDestroy(x);
        return st;
    }
    // each shape is timed with the clear of the spare splat buffers it does
    // in a frame (they are cleared over and over here, and left cleared)
    float ms[N_TILE_SHAPES] = {};
    int pick = 0;
    bool cleared = false;
    for (int k = 0; k < N_TILE_SHAPES && e == hipSuccess; k++) {
        set_tiles(P.ra, k);
        if ((st = plan_ahead(r, r->scratch, P.ra, k)) != ISTI_OK) break;
        cleared = cleared || P.ra.ahead.n_bufs > 0;
        e = launch_resolve(P.ra, P.window, P.always_valid, k, s);  // warm-up
        if (e == hipSuccess) e = hipEventRecord(ev[0], s);
        for (int i = 0; i < reps && e == hipSuccess; i++) e = launch_resolve(P.ra, P.window, P.always_valid, k, s);
        if (e == hipSuccess) e = hipEventRecord(ev[1], s);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms[k], ev[0], ev[1]);
        ms[k] /= (float)reps;
        if (ms[k] < ms[pick]) pick = k;
    }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (st != ISTI_OK) return st;
    if (e != hipSuccess) return dev_fail(r, e, "autotune");
    if (cleared) ahead_cleared(r, r->scratch, P.ra);
    {
        std::lock_guard<std::mutex> lock(g_tuned_mutex);
        g_tuned[TuneKey{P.ra.w, P.ra.h, P.ra.n_aux, P.window}] = pick;
        g_tuned_splat[TuneKey{P.ra.w, P.ra.h, P.ra.n_aux, P.window}] = pick_splat;
    }
    if (times_ms) {
        std::memcpy(times_ms, ms, sizeof(ms));
        times_ms[N_TILE_SHAPES] = splat_ms[0];
        times_ms[N_TILE_SHAPES + 1] = splat_ms[1];
    }
    if (best) *best = pick;
    if (best_splat) *best_splat = pick_splat;
    return ISTI_OK;
}

isti_status isti_set_splat_binning(isti_renderer *r, int32_t mode) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    if (mode < -1 || mode > 1) return fail(r, ISTI_ERR_INVALID_ARG, "splat mode must be -1, 0 or 1");
    r->splat_mode = mode;
    return ISTI_OK;
}

isti_status isti_set_stage_timing(isti_renderer *r, int32_t enable) {
    if (!r) return ISTI_ERR_INVALID_ARG;
    hipError_t e = hipSetDevice(r->device);
    if (e != hipSuccess) return dev_fail(r, e, "hipSetDevice");
    TimingRing &t = r->ring;
    if (enable && !t.ev[0][0]) {
        for (auto &set : t.ev)
            for (auto &ev : set)
                if ((e = hipEventCreate(&ev)) != hipSuccess) return dev_fail(r, e, "hipEventCreate");
    }
    if (enable) {
        for (int i = 0; i < RING; i++) t.pending[i] = false;
        t.head = 0;
        t.last = -1;
        t.total[0] = t.total[1] = t.total[2] = 0.0;
        t.count = 0;
    }
    r->timing = enable != 0;
    return ISTI_OK;
}

isti_status isti_stage_times(isti_renderer *r, float ms[3]) {
    if (!r || !ms) return ISTI_ERR_INVALID_ARG;
    TimingRing &t = r->ring;
    if (t.last < 0 || !t.pending[t.last])
        return fail(r, ISTI_ERR_INVALID_ARG, "no timed render pending (enable isti_set_stage_timing)");
    hipError_t e = hipEventSynchronize(t.ev[t.last][3]);
    for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipEventElapsedTime(&ms[i], t.ev[t.last][i], t.ev[t.last][i + 1]);
    if (e != hipSuccess) return dev_fail(r, e, "hipEventElapsedTime");
    return ISTI_OK;
}

isti_status isti_stage_totals(isti_renderer *r, double total_ms[3], int64_t *count) {
    if (!r || !total_ms || !count) return ISTI_ERR_INVALID_ARG;
    TimingRing &t = r->ring;
    for (int k = 0; k < RING; k++) {
        const int slot = (t.head + k) % RING;  // oldest first
        hipError_t e = harvest(t, slot);
        if (e != hipSuccess) return dev_fail(r, e, "hipEventElapsedTime");
    }
    for (int i = 0; i < 3; i++) total_ms[i] = t.total[i];
    *count = t.count;
    return ISTI_OK;
}

const char *isti_last_error(const isti_renderer *r) {
    if (!r) return g_create_error.c_str();
    return r->err.c_str();
}

isti_status isti_selftest_division(int32_t device, uint64_t n, uint64_t seed, uint32_t *counts, void *stream) {
    if (!counts) return ISTI_ERR_INVALID_ARG;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = launch_div_selftest(n, seed, counts, reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? ISTI_OK : ISTI_ERR_DEVICE;
}

const char *isti_version(void) { return "isti-gfx950 0.1.0"; }

}  // extern "C"
