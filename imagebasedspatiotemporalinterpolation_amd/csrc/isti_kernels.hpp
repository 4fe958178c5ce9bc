// isti_kernels.hpp - kernel argument blocks and launchers of the gfx950
// interpolation path. Every per-frame descriptor travels as a by-value
// kernel argument, so a frame needs no host->device copy and no host sync
// (the reference copies aux descriptors with cudaMemcpy and syncs the host
// between its two kernels, src/renderer.cu:650,670-671).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/isti.h"
#include "isti_device.hpp"

namespace isti {

constexpr int MAX_VIEWS = 1 + ISTI_MAX_AUX;        // main + aux per slot
constexpr int MAX_SPLAT_JOBS = 2 * MAX_VIEWS;       // (slot, view) pairs
constexpr uint32_t SPLAT_BLOCK = 256;               // source pixels per K1 block
constexpr uint32_t TILE_W = 16, TILE_H = 16;        // default resolve tile (one block)
// Resolve tile shapes the single-render path is built for (index = shape id;
// one thread per pixel; every shape gives identical results).
constexpr int N_TILE_SHAPES = 4;
constexpr uint32_t TILE_SHAPE_W[N_TILE_SHAPES] = {16, 32, 8, 8};
constexpr uint32_t TILE_SHAPE_H[N_TILE_SHAPES] = {16, 8, 8, 32};
// Tile shape of the persistent batched resolve (16x16; one-wave 8x8 tiles
// were slower there: 17.6 against 18.3 Gpx/s on the batched stereo frames).
constexpr int BATCH_TILE_SHAPE = 0;
constexpr uint32_t BATCH_TILE_W = TILE_SHAPE_W[BATCH_TILE_SHAPE], BATCH_TILE_H = TILE_SHAPE_H[BATCH_TILE_SHAPE];

// Flow toward the other keyframe (full_ff): x, y, z only. Its fourth
// component (the visibility flag, also in cam_ff.w) is never read, so the
// plane stores 12 bytes per pixel instead of 16.
struct Flow3 {
    float x, y, z;
};

// One source view of the reproject+splat stage (K1, renderer.cu:76-146).
struct SplatJob {
    const float *z;
    const float *mx, *my, *mz;  // null: no object motion
    Flow3 *full_ff;
    float4 *cam_ff;
    unsigned long long *splat;
    uint32_t w, h, n_px, block_begin;
    float weight;
    int32_t main_slot;  // which main view's VP decides visible_in_main; < 0: a main view (no flag)
    int32_t req;        // owning request (batched launches)
    float rcp_w;        // 1/w rounded to binary32, for div_px
    int32_t write_full; // 0: full_ff already holds this keyframe pair's flow
    int32_t binned;     // 1: 32x8 source tiles through an LDS bin (splat_tile_binned), 0: direct
    double inv_w, inv_h;  // RN64(1/w), RN64(1/h) for pixel2CS
    Mat4 inv_vp_from;
    Mat4 vp_to;
};

struct SplatArgs {
    SplatJob job[MAX_SPLAT_JOBS];
    Mat4 vp_cam;
    Mat4 vp_main[2];
    int32_t n_jobs;
    int32_t init_linear;
    uint32_t n_blocks;
    int32_t pad;
};

// Buffers of one (slot, view) read by the resolve stage (CamReprojInfo,
// include/renderer.h:37-50).
struct ResolveView {
    const uchar4 *img;
    const float *z;
    const uchar4 *sh_info;
    const uchar4 *sh_a;
    const uchar4 *sh_b;
    const float4 *cam_ff;
    const Flow3 *full_ff;
    const unsigned long long *splat;
    uint32_t w, h;
    int32_t priority;
    int32_t forward;
    double inv_w, inv_h;  // RN64(1/w), RN64(1/h) for pixel2CS
    float rs_x, rs_y;     // w / min(w, h), h / min(w, h) (renderer.cu:213)
    // render size / view size per axis (renderer.cu:433) and their
    // multiply-high divisors (div_small)
    uint32_t ratio_x, ratio_y, magic_x, magic_y;
};

// Splat buffers a resolve launch fills with the empty key beside its own work
// (the spare buffers of the frame's views, which the next frame of the same
// stream splats into instead of clearing): the workgroup resolving tile
// b < n_blocks of the render clears `rounds` runs of one cell pair per thread
// of buffer j, the last j with block_begin[j] <= b.
struct ClearAhead {
    unsigned long long *buf[MAX_SPLAT_JOBS];
    uint32_t n[MAX_SPLAT_JOBS];
    uint32_t block_begin[MAX_SPLAT_JOBS];
    int32_t n_bufs;
    uint32_t n_blocks;
    uint32_t rounds;
};

struct ResolveArgs {
    ResolveView fwd[MAX_VIEWS];
    ResolveView bwd[MAX_VIEWS];
    uchar4 *out;
    uint32_t out_pitch_px;
    uint32_t w, h;
    float alpha;
    int32_t fpa, sfpa, no_shadows, n_aux, max_iter;
    uint32_t tiles_x, n_tiles;
    uint32_t *q_out;          // optional capture of the main camera's search results
    // learned search starts (main camera): null when unused
    const float *net_out;     // [2][net_h][net_w][16]
    uint32_t net_w, net_h, net_shift;
    float conf_logit;         // starts are used where the confidence logit >= this
    uint32_t *learned_stats;  // kept, rejected, low confidence, total
    ClearAhead ahead;         // per-render resolves only; n_blocks 0 = none
};

struct ClearArgs {
    unsigned long long *buf[MAX_SPLAT_JOBS];
    uint32_t n[MAX_SPLAT_JOBS];
    uint32_t block_begin[MAX_SPLAT_JOBS];
    int32_t n_bufs;
    uint32_t n_blocks;
};

struct TimewarpArgs {
    Mat4 inv_vp_from;   // render camera
    Mat4 inv_vp_to;     // fwd keyframe
    Mat4 vp_to;
    Mat4 vp_aux;
    const float *depth_to;  // device scalar: mean z of the fwd keyframe
    const uchar4 *img_to;
    const uchar4 *img_aux;
    uchar4 *out;
    uint32_t out_pitch_px;
    uint32_t w, h, w_to, h_to, w_aux, h_aux;
    int32_t aux_set;
    uint32_t tiles_x, n_tiles;
};

// ---------------------------------------------------------------------------
// Batched renders: a device-resident work table read by persistent kernels.
// ---------------------------------------------------------------------------
constexpr int N_VARIANTS = 2 * (1 + ISTI_MAX_AUX);  // (window 1 | window 3 + always-valid) x n_aux
__host__ __device__ constexpr int resolve_variant(bool mav, int n_aux) { return (mav ? 1 + ISTI_MAX_AUX : 0) + n_aux; }

struct BatchRequest {
    ResolveArgs args;
    Mat4 vp_cam;
    Mat4 vp_main[2];
    int32_t init_linear;
    int32_t pad[3];
};

// Header of one batch; the arrays it points to live in the same device
// allocation, right behind it.
struct BatchTable {
    uint32_t n_req, n_jobs;
    uint32_t clear_blocks, splat_blocks;
    uint32_t var_req_begin[N_VARIANTS + 1];  // requests are ordered by resolve variant
    uint32_t pad;
    const BatchRequest *req;       // [n_req]
    const SplatJob *jobs;          // [n_jobs]
    const uint32_t *clear_begin;   // [n_jobs + 1] block prefix of the clear stage
    const uint32_t *splat_begin;   // [n_jobs + 1] block prefix of the splat stage
    const uint32_t *tile_begin;    // [n_req + 1] tile prefix of the resolve stage
};

hipError_t launch_clear_batch(const BatchTable *t, uint32_t grid, hipStream_t s);
hipError_t launch_splat_batch(const BatchTable *t, uint32_t grid, hipStream_t s);
hipError_t launch_resolve_batch(const BatchTable *t, int variant, uint32_t grid, hipStream_t s);
// blocks of one resolve variant that fit on one CU at once
int resolve_batch_blocks_per_cu(int variant);

struct FeatureArgs {
    ResolveView v[2];   // main fwd, main bwd (with this frame's cam_ff)
    uint16_t *x;        // [2][net_h][net_w][16] bf16
    uint32_t w, h, net_w, net_h, net_shift;
    float alpha;
};
hipError_t launch_features(const FeatureArgs &a, hipStream_t s);
hipError_t launch_learned_loss(const float *out, const uint32_t *q, const float4 *ff_fwd, const float4 *ff_bwd,
                               uint32_t w, uint32_t h, uint32_t net_w, uint32_t net_h, uint32_t shift,
                               float conf_weight, float *dout, float *loss_acc, hipStream_t s);
hipError_t launch_resolve_learned(const ResolveArgs &a, hipStream_t s);
// Learned-start statistics: the resolve adds into STAT_SLOTS spread counter
// rows (a row of STAT_STRIDE uint32 per slot; same-address atomics from
// every wave would serialise at the memory side), then one block sums the
// rows into the caller's five counters and zeroes them for the next frame.
constexpr uint32_t STAT_SLOTS = 256, STAT_STRIDE = 8, STAT_COUNTERS = 5;
hipError_t launch_stats_reduce(uint32_t *slots, uint32_t *out, hipStream_t s);

hipError_t launch_clear(const ClearArgs &a, hipStream_t s);
// binned: the LDS-binned variant; its blocks are 32x8 source tiles, so a
// job's block_begin prefix must come from splat_blocks(..., true)
hipError_t launch_splat(const SplatArgs &a, bool binned, hipStream_t s);
uint32_t splat_blocks(uint32_t w, uint32_t h, bool binned);
hipError_t launch_resolve(const ResolveArgs &a, int window, bool always_valid, int shape, hipStream_t s);
hipError_t launch_mean_z(const float *z, uint32_t w, uint32_t h, double *row_sums, float *mean, hipStream_t s);
hipError_t launch_timewarp(const TimewarpArgs &a, hipStream_t s);
hipError_t launch_div_selftest(uint64_t n, uint64_t seed, uint32_t *counts, hipStream_t s);

}  // namespace isti
