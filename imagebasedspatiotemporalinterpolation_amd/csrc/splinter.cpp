// splinter.cpp - host driver, the counterpart of the reference's
// src/main.cu (CLI :66-80, frame loop :139-191).
//
//   splinter <input> <output-dir> [-c CAMDIR] [-w] [-r W H] [-a NAME...] [-l N]
//            [-m 0..4] [-lh] [-ns] [-hb] [--frames N] [--iters N] [--seed S]
//
// <input> is either a keyframe recording in the reference's on-disk format
// (a directory with main/ and one directory per -a NAME; read through
// include/isti_io.h, as FrameFileReader does) or a seeded synthetic scene:
// "scene:1", "scene:2" or "scene:3" (optionally "scene:2:SEED"; -a names
// select up to 2 aux cameras, left/right). With a recording the client runs
// at 120 Hz, frames_per_anchor = 120 / serverFps and the target camera is the
// interpolation of the keyframe views (main.cu:139-161); with a synthetic
// scene the target camera follows the scene's analytic camera path at
// client-frame times. -c CAMDIR reads a client camera path (CAMDIR/cam.json,
// file_reader.cu:78-107): client frame n renders with frames[n].w2c and the
// normalised frames[n].proj, the file's frame count is the run length and, for
// a recording, its clientFps sets frames_per_anchor (main.cu:128-161). -w
// writes <output-dir>/NNNN.png per frame. Per-frame GPU times (hipEvents) are
// summarised at the end as Mpix/s and p50/p99 ms.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "isti_io.h"
#include "isti_renderer.hpp"
#include "isti_scene.h"

namespace fs = std::filesystem;

#define HIP_OK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));          \
            std::exit(2);                                                                \
        }                                                                                \
    } while (0)

namespace {

// --dry-run: print the frame schedule, touch no device
bool g_dry_run = false;

// ---- minimal PNG writer (stored deflate blocks) ----
uint32_t crc_table[256];
void crc_init() {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
}
uint32_t crc(const uint8_t *p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}
void be32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back(x >> 24); v.push_back(x >> 16); v.push_back(x >> 8); v.push_back(x);
}
void chunk(std::ofstream &f, const char *type, const std::vector<uint8_t> &data) {
    std::vector<uint8_t> b;
    be32(b, (uint32_t)data.size());
    b.insert(b.end(), type, type + 4);
    b.insert(b.end(), data.begin(), data.end());
    const uint32_t c = crc(b.data() + 4, b.size() - 4) ^ 0xFFFFFFFFu;
    be32(b, c);
    f.write(reinterpret_cast<const char *>(b.data()), (std::streamsize)b.size());
}
bool write_png(const fs::path &path, const uint8_t *rgba, uint32_t w, uint32_t h) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    f.write(reinterpret_cast<const char *>(sig), 8);
    std::vector<uint8_t> ihdr;
    be32(ihdr, w); be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
    chunk(f, "IHDR", ihdr);
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (w * 4 + 1));
    for (uint32_t y = 0; y < h; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba + (size_t)y * w * 4, rgba + (size_t)(y + 1) * w * 4);
    }
    std::vector<uint8_t> z{0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) { a = (a + c) % 65521; b = (b + a) % 65521; }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + n >= raw.size() ? 1 : 0);
        z.push_back(n & 0xFF); z.push_back(n >> 8);
        z.push_back(~n & 0xFF); z.push_back((~n >> 8) & 0xFF);
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        if (off + n >= raw.size()) break;
    }
    be32(z, (b << 16) | a);
    chunk(f, "IDAT", z);
    chunk(f, "IEND", {});
    return true;
}

// Device planes of one keyframe view, in the reference's formats.
struct DevView {
    uint32_t w = 0, h = 0;
    uint8_t *color = nullptr, *sh_info = nullptr, *sh_a = nullptr, *sh_b = nullptr;
    float *z = nullptr, *mx = nullptr, *my = nullptr, *mz = nullptr;
    float vp[16], ivp[16];
    float view[16], proj[16];         // recordings: w2c and normalised projection
    bool has_motion = true, has_shadows = true;
    void alloc(uint32_t W, uint32_t H, bool shadows) {
        w = W; h = H;
        const size_t n = (size_t)W * H;
        HIP_OK(hipMalloc(&color, n * 4));
        HIP_OK(hipMalloc(&z, n * 4));
        HIP_OK(hipMalloc(&mx, n * 4));
        HIP_OK(hipMalloc(&my, n * 4));
        HIP_OK(hipMalloc(&mz, n * 4));
        if (shadows) {
            HIP_OK(hipMalloc(&sh_info, n * 4));
            HIP_OK(hipMalloc(&sh_a, n * 4));
            HIP_OK(hipMalloc(&sh_b, n * 4));
        }
    }
    isti_view_desc desc(int priority) const {
        isti_view_desc d;
        std::memset(&d, 0, sizeof(d));
        d.width = w; d.height = h;
        d.color = color; d.z = z;
        if (has_motion) { d.motion_x = mx; d.motion_y = my; d.motion_z = mz; }
        if (has_shadows) { d.shadow_info = sh_info; d.shadow_color_a = sh_a; d.shadow_color_b = sh_b; }
        std::memcpy(d.viewproj, vp, sizeof(vp));
        std::memcpy(d.inv_viewproj, ivp, sizeof(ivp));
        d.priority = priority;
        return d;
    }
};

// Generates keyframe k of one view in direction dir and uploads it.
void load_keyframe(const isti_scene_config &cfg, int k, int view, int dir, DevView &dv, bool colour_depth) {
    if (g_dry_run) return;
    const size_t n = (size_t)dv.w * dv.h;
    std::vector<uint8_t> col(n * 4), si(n * 4), sa(n * 4), sb(n * 4);
    std::vector<float> z(n), mx(n), my(n), mz(n);
    if (isti_scene_keyframe(&cfg, k, view, dir, col.data(), z.data(), mx.data(), my.data(), mz.data(),
                            si.data(), sa.data(), sb.data(), dv.vp, dv.ivp, 0) != 0) {
        std::fprintf(stderr, "scene generation failed\n");
        std::exit(2);
    }
    if (colour_depth) {
        HIP_OK(hipMemcpy(dv.color, col.data(), n * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dv.z, z.data(), n * 4, hipMemcpyHostToDevice));
    }
    HIP_OK(hipMemcpy(dv.mx, mx.data(), n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dv.my, my.data(), n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dv.mz, mz.data(), n * 4, hipMemcpyHostToDevice));
    if (dv.sh_info) {
        HIP_OK(hipMemcpy(dv.sh_info, si.data(), n * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dv.sh_a, sa.data(), n * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dv.sh_b, sb.data(), n * 4, hipMemcpyHostToDevice));
    }
}

// Reads keyframe view `which` of pair `pair` of a recorded camera and uploads it.
void load_recorded(const std::string &cam_dir, int pair, int which, DevView &dv) {
    isti_io_view v;
    if (isti_io_read_view(cam_dir.c_str(), pair, which, &v) != ISTI_OK) {
        std::fprintf(stderr, "%s\n", isti_io_last_error());
        std::exit(2);
    }
    if (v.width != dv.w || v.height != dv.h) {
        std::fprintf(stderr, "%s: keyframe size changed within the recording\n", cam_dir.c_str());
        std::exit(2);
    }
    const size_t n = (size_t)dv.w * dv.h;
    std::memcpy(dv.vp, v.viewproj, sizeof(dv.vp));
    std::memcpy(dv.ivp, v.inv_viewproj, sizeof(dv.ivp));
    std::memcpy(dv.view, v.view, sizeof(dv.view));
    std::memcpy(dv.proj, v.proj, sizeof(dv.proj));
    if (g_dry_run) {
        isti_io_free_view(&v);
        return;
    }
    HIP_OK(hipMemcpy(dv.color, v.color, n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dv.z, v.z, n * 4, hipMemcpyHostToDevice));
    dv.has_motion = v.motion_x != nullptr;
    if (dv.has_motion) {
        HIP_OK(hipMemcpy(dv.mx, v.motion_x, n * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dv.my, v.motion_y, n * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dv.mz, v.motion_z, n * 4, hipMemcpyHostToDevice));
    }
    dv.has_shadows = v.shadow_info && v.shadow_color_a && v.shadow_color_b && dv.sh_info;
    if (dv.has_shadows) {
        HIP_OK(hipMemcpy(dv.sh_info, v.shadow_info, n * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dv.sh_a, v.shadow_color_a, n * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dv.sh_b, v.shadow_color_b, n * 4, hipMemcpyHostToDevice));
    }
    isti_io_free_view(&v);
}

void usage() {
    std::puts(
        "Splinter - Split-Rendering Frame Interpolation (MI355X)\n"
        "usage: splinter <recording-dir | scene:N[:SEED]> <output-dir> [options]\n"
        "  -c, --camera-path DIR      client camera path (DIR/cam.json)\n"
        "  -w, --write-images         write PNG frames to output-dir\n"
        "  -r, --resolution W H       render resolution (sets the scene's main view size)\n"
        "  -a, --aux-dirs NAME...     aux camera directories (synthetic scene: up to 2 names)\n"
        "  -l, --latency N            delay keyframes by N client frames\n"
        "  -m, --mode 0..4            render mode (RenderMode)\n"
        "  -lh, --latency-hold-last   hold the last frame while a keyframe is late\n"
        "  -ns, --no-shadows          disable dynamic shadows\n"
        "  -hb, --holes-black         clear the output every frame (no hole filling)\n"
        "  --frames N                 client frames to render (default: scene 2 keyframe pairs,\n"
        "                             recording 15 s at 120 Hz or its length)\n"
        "  --iters N                  fixed-point search cap (default: the mode's own)\n"
        "  --seed S                   scene seed\n"
        "  --dry-run                  print the frame schedule (keyframe pair, alpha, whether the\n"
        "                             renderer gets a new pair, target view-projection); no GPU");
}

}  // namespace

int main(int argc, char **argv) {
    crc_init();
    std::vector<std::string> pos, aux_names;
    std::string camera_path;
    bool write_images = false, hold_last = false, no_shadows = false, holes_black = false;
    int latency = 0, mode = 0, frames = -1, iters = -1, res_w = 0, res_h = 0;
    long long seed = -1;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto need = [&](int n) {
            if (i + n >= argc) { usage(); std::exit(1); }
        };
        if (a == "-h" || a == "--help") { usage(); return 0; }
        else if (a == "-w" || a == "--write-images") write_images = true;
        else if (a == "-lh" || a == "--latency-hold-last") hold_last = true;
        else if (a == "-ns" || a == "--no-shadows") no_shadows = true;
        else if (a == "-hb" || a == "--holes-black") holes_black = true;
        else if (a == "--dry-run") g_dry_run = true;
        else if (a == "-l" || a == "--latency") { need(1); latency = std::atoi(argv[++i]); }
        else if (a == "-m" || a == "--mode") { need(1); mode = std::atoi(argv[++i]); }
        else if (a == "--frames") { need(1); frames = std::atoi(argv[++i]); }
        else if (a == "--iters") { need(1); iters = std::atoi(argv[++i]); }
        else if (a == "--seed") { need(1); seed = std::atoll(argv[++i]); }
        else if (a == "-r" || a == "--resolution") { need(2); res_w = std::atoi(argv[++i]); res_h = std::atoi(argv[++i]); }
        else if (a == "-a" || a == "--aux-dirs") {
            while (i + 1 < argc && argv[i + 1][0] != '-') aux_names.push_back(argv[++i]);
        } else if (a == "-c" || a == "--camera-path") {
            need(1);
            camera_path = argv[++i];
        } else if (!a.empty() && a[0] == '-') { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 1; }
        else pos.push_back(a);
    }
    if (pos.size() != 2 || mode < 0 || mode > 4) { usage(); return 1; }
    const bool synthetic = pos[0].rfind("scene:", 0) == 0;
    isti_scene_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    uint32_t W = 0, H = 0, AW = 0, AH = 0;
    int n_eyes = 1, n_aux = 0, fpa = 1, sfpa = 0, total = 0, n_keyframes = 0;
    bool shadows = false;
    std::vector<int> aux_prio;
    std::string main_dir;
    std::vector<std::string> aux_dirs;
    if (synthetic) {
        int workload = 1;
        uint64_t scene_seed = 1;
        std::sscanf(pos[0].c_str() + 6, "%d", &workload);
        const size_t c2 = pos[0].find(':', 6);
        if (c2 != std::string::npos) scene_seed = std::strtoull(pos[0].c_str() + c2 + 1, nullptr, 10);
        if (seed >= 0) scene_seed = (uint64_t)seed;
        if (isti_scene_defaults(workload, scene_seed, &cfg) != 0) { std::fprintf(stderr, "unknown scene\n"); return 1; }
        if (res_w > 0 && res_h > 0) {
            const uint32_t ow = cfg.width, oh = cfg.height;
            cfg.width = (uint32_t)res_w;
            cfg.height = (uint32_t)res_h;
            if (cfg.n_aux > 0) {  // keep the aux/main ratio
                cfg.aux_width = cfg.width * cfg.aux_width / ow;
                cfg.aux_height = cfg.height * cfg.aux_height / oh;
            }
        }
        W = cfg.width;
        H = cfg.height;
        AW = cfg.aux_width;
        AH = cfg.aux_height;
        n_eyes = cfg.n_eyes;
        n_aux = n_eyes == 2 ? 0 : std::min<int>((int)aux_names.size(), cfg.n_aux);
        if (n_eyes == 1 && (int)aux_names.size() > cfg.n_aux)
            std::fprintf(stderr, "note: scene has %d aux cameras; extra -a names ignored\n", cfg.n_aux);
        fpa = cfg.keyframe_spacing;                 // main.cu:144
        sfpa = cfg.shadows ? cfg.shadow_frames : 0;  // main.cu:145
        total = frames > 0 ? frames : 2 * fpa;
        shadows = cfg.shadows != 0;
        aux_prio.assign((size_t)n_aux, 3);  // objectMotion*2 + shadowGradients
    } else {
        // FrameFileReader::init (file_reader.cu:109-136) and main.cu:139-145
        main_dir = pos[0] + "/main";
        isti_io_camera mc;
        if (isti_io_read_camera(main_dir.c_str(), &mc, nullptr, 0) != ISTI_OK) {
            std::fprintf(stderr, "%s\n", isti_io_last_error());
            return 1;
        }
        if (mc.server_fps <= 0 || mc.n_frames < 2) { std::fprintf(stderr, "recording needs serverFps > 0 and 2 frames\n"); return 1; }
        W = mc.width;
        H = mc.height;
        n_keyframes = mc.n_frames;
        fpa = 120 / mc.server_fps;  // the fixed 120 Hz client of main.cu:139-144
        sfpa = mc.shadow_fps / mc.server_fps;
        n_aux = std::min<int>((int)aux_names.size(), ISTI_MAX_AUX);
        for (int a = 0; a < n_aux; a++) {
            aux_dirs.push_back(pos[0] + "/" + aux_names[(size_t)a]);
            isti_io_camera ac;
            if (isti_io_read_camera(aux_dirs.back().c_str(), &ac, nullptr, 0) != ISTI_OK) {
                std::fprintf(stderr, "%s\n", isti_io_last_error());
                return 1;
            }
            if (a > 0 && (ac.width != AW || ac.height != AH)) { std::fprintf(stderr, "aux cameras differ in size\n"); return 1; }
            AW = ac.width;
            AH = ac.height;
            aux_prio.push_back(ac.priority);
            n_keyframes = std::min(n_keyframes, ac.n_frames);
        }
        if (res_w > 0) std::fprintf(stderr, "note: -r is ignored for recordings (render size = main keyframe size)\n");
        shadows = true;  // planes are used only where the recording has them
        total = frames > 0 ? frames : std::min(120 * 15, (n_keyframes - 1) * fpa);
    }
    // main.cu:128-142: a client camera file fixes the frame count and, with a
    // recording, the client rate
    std::vector<isti_io_frame> client_frames;
    if (!camera_path.empty()) {
        isti_io_camera cc;
        if (isti_io_read_camera(camera_path.c_str(), &cc, nullptr, 0) != ISTI_OK) {
            std::fprintf(stderr, "%s\n", isti_io_last_error());
            return 1;
        }
        client_frames.resize((size_t)cc.n_frames);
        if (isti_io_read_camera(camera_path.c_str(), &cc, client_frames.data(), cc.n_frames) != ISTI_OK) {
            std::fprintf(stderr, "%s\n", isti_io_last_error());
            return 1;
        }
        if (!synthetic) {
            isti_io_camera mc;
            isti_io_read_camera(main_dir.c_str(), &mc, nullptr, 0);
            if (cc.client_fps < mc.server_fps) {
                std::fprintf(stderr, "camera path clientFps %d is below the recording's serverFps %d\n", cc.client_fps,
                             mc.server_fps);
                return 1;
            }
            fpa = cc.client_fps / mc.server_fps;
        }
        total = frames > 0 ? std::min(frames, (int)client_frames.size()) : (int)client_frames.size();
    }
    const fs::path out_dir = pos[1];
    if (!fs::exists(out_dir) && !fs::create_directories(out_dir)) {
        std::fprintf(stderr, "Output directory could not be created\n");
        return 1;
    }

    // per eye: fwd/bwd main views + aux curr/next
    struct Eye {
        DevView slot[2];
        std::vector<DevView> aux[2];
        isti::FrameData fd;
        uint8_t *out = nullptr;
        isti::Renderer *renderer = nullptr;
    };
    std::vector<Eye> eyes(n_eyes);
    for (int e = 0; e < n_eyes; e++) {
        Eye &E = eyes[e];
        for (int s = 0; s < 2; s++) {
            E.aux[s].resize(n_aux);
            if (g_dry_run) {  // sizes only: the loads check them
                E.slot[s].w = W;
                E.slot[s].h = H;
                for (int a = 0; a < n_aux; a++) E.aux[s][a].w = AW, E.aux[s][a].h = AH;
                continue;
            }
            E.slot[s].alloc(W, H, shadows);
            for (int a = 0; a < n_aux; a++) E.aux[s][a].alloc(AW, AH, shadows);
        }
        E.fd = isti::FrameData(n_aux);
        if (g_dry_run) continue;
        HIP_OK(hipMalloc(&E.out, (size_t)W * H * 4));
        HIP_OK(hipMemset(E.out, 0, (size_t)W * H * 4));
        E.renderer = new isti::Renderer(n_aux);
    }
    int next_key = 0;  // keyframe index that will be read next
    int fwd_slot = 0;  // which DevView slot is currently fwd
    auto view_of = [&](int e, int a) { return n_eyes == 2 ? e : a; };
    auto refresh_fd = [&]() {
        for (int e = 0; e < n_eyes; e++) {
            Eye &E = eyes[e];
            E.fd.fwd = E.slot[fwd_slot].desc(0);
            E.fd.bwd = E.slot[1 - fwd_slot].desc(0);
            for (int a = 0; a < n_aux; a++) {
                E.fd.aux_cam_images[a].curr = E.aux[fwd_slot][a].desc(aux_prio[(size_t)a]);
                E.fd.aux_cam_images[a].next = E.aux[1 - fwd_slot][a].desc(aux_prio[(size_t)a]);
                E.fd.aux_cam_images[a].priority = aux_prio[(size_t)a];
            }
        }
    };
    // file_reader.cu:109-136: keyframe 0 goes into the bwd slot
    for (int e = 0; e < n_eyes; e++) {
        if (synthetic) {
            load_keyframe(cfg, 0, view_of(e, 0), -1, eyes[e].slot[1 - fwd_slot], true);
            for (int a = 0; a < n_aux; a++) load_keyframe(cfg, 0, 1 + a, -1, eyes[e].aux[1 - fwd_slot][a], true);
        } else {
            load_recorded(main_dir, 0, 0, eyes[e].slot[1 - fwd_slot]);
            for (int a = 0; a < n_aux; a++) load_recorded(aux_dirs[(size_t)a], 0, 0, eyes[e].aux[1 - fwd_slot][a]);
        }
    }
    next_key = 1;
    // a renderer is handed the keyframe pair once per pair (its planes are
    // rewritten in place by read_next), so the pair's flow toward the other
    // keyframe is computed by its first render and reused by the others;
    // every read of a new pair clears these flags
    std::vector<char> pair_set((size_t)n_eyes, 0);
    // file_reader.cu:167-217: swap, read the next keyframe into bwd, re-read
    // the (new) fwd keyframe's forward motion / shadow data
    // With -l N the reference renders its first N client frames before any
    // keyframe pair has been read, from a fwd image that was never filled
    // (main.cu:152-177, file_reader.cu:127-128); those frames are not rendered
    // here, so the output keeps its initial zeros until the first pair arrives.
    bool have_pair = false;
    auto read_next = [&]() {
        std::fill(pair_set.begin(), pair_set.end(), 0);
        have_pair = true;
        fwd_slot = 1 - fwd_slot;
        if (!synthetic && next_key >= n_keyframes) return;  // end of the recording: swap only (:169-172)
        for (int e = 0; e < n_eyes; e++) {
            if (synthetic) {
                load_keyframe(cfg, next_key - 1, view_of(e, 0), +1, eyes[e].slot[fwd_slot], false);
                load_keyframe(cfg, next_key, view_of(e, 0), -1, eyes[e].slot[1 - fwd_slot], true);
                for (int a = 0; a < n_aux; a++) {
                    load_keyframe(cfg, next_key - 1, 1 + a, +1, eyes[e].aux[fwd_slot][a], false);
                    load_keyframe(cfg, next_key, 1 + a, -1, eyes[e].aux[1 - fwd_slot][a], true);
                }
            } else {
                load_recorded(main_dir, next_key - 1, 0, eyes[e].slot[fwd_slot]);
                load_recorded(main_dir, next_key - 1, 1, eyes[e].slot[1 - fwd_slot]);
                for (int a = 0; a < n_aux; a++) {
                    load_recorded(aux_dirs[(size_t)a], next_key - 1, 0, eyes[e].aux[fwd_slot][a]);
                    load_recorded(aux_dirs[(size_t)a], next_key - 1, 1, eyes[e].aux[1 - fwd_slot][a]);
                }
            }
        }
        next_key++;
    };

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (!g_dry_run) {
        HIP_OK(hipEventCreate(&ev0));
        HIP_OK(hipEventCreate(&ev1));
    }
    std::vector<float> frame_ms;
    std::vector<uint8_t> host_img((size_t)W * H * 4);
    for (int i = 0; i < total; i += fpa) {          // main.cu:149
        for (int b = 0; b < fpa; b++) {              // main.cu:152
            if (latency == b) read_next();
            refresh_fd();
            const float t = float(latency > b ? fpa + b : b) / float(fpa);  // main.cu:157
            const int frame_number = i + b + 1;
            if (frame_number > total) break;
            float ms_total = 0.f;
            bool rendered = false;
            for (int e = 0; e < n_eyes; e++) {
                Eye &E = eyes[e];
                if (holes_black && !g_dry_run) HIP_OK(hipMemset(E.out, 0, (size_t)W * H * 4));
                if (g_dry_run && !(have_pair && !(b < latency && hold_last)))
                    std::printf("frame %d eye %d skip\n", frame_number - 1, e);
                if (have_pair && !(b < latency && hold_last)) {
                    isti::RenderInfo ri;
                    ri.resolution[0] = W;
                    ri.resolution[1] = H;
                    ri.alpha = t;
                    if (!client_frames.empty()) {  // main.cu:158-161 with -c
                        const isti_io_frame &cf = client_frames[(size_t)(i + b)];
                        if (isti_io_client_camera(cf.w2c, cf.proj, ri.viewproj_mat, ri.inv_viewproj_mat) != ISTI_OK) {
                            std::fprintf(stderr, "%s\n", isti_io_last_error());
                            return 2;
                        }
                    } else if (synthetic) {  // client camera at absolute client-frame time
                        uint32_t w, h;
                        isti_scene_camera(&cfg, (double)(i + b), view_of(e, 0), ri.viewproj_mat, ri.inv_viewproj_mat, &w, &h);
                    } else {  // getIntermediateViewMatrix (main.cu:26-29, 158-161)
                        const DevView &F = E.slot[fwd_slot], &B = E.slot[1 - fwd_slot];
                        if (isti_io_camera_at(F.view, B.view, F.proj, t, nullptr, ri.viewproj_mat, ri.inv_viewproj_mat) != ISTI_OK) {
                            std::fprintf(stderr, "%s\n", isti_io_last_error());
                            return 2;
                        }
                    }
                    ri.additional_info = {fpa, sfpa, no_shadows, (isti::RenderMode)mode};
                    ri.max_search_iterations = iters;
                    if (g_dry_run) {
                        std::printf("frame %d eye %d pair %d t %.9g set %d vp", frame_number - 1, e, next_key - 2,
                                    (double)t, pair_set[(size_t)e] ? 0 : 1);
                        for (int k = 0; k < 16; k++) std::printf(" %.9g", (double)ri.viewproj_mat[k]);
                        std::printf("\n");
                        pair_set[(size_t)e] = 1;
                        continue;
                    }
                    isti::RenderBuffer rb{E.out, (size_t)W * 4, W, H};
                    HIP_OK(hipEventRecord(ev0, nullptr));
                    try {
                        if (!pair_set[(size_t)e]) {
                            E.renderer->set_frame(E.fd);
                            pair_set[(size_t)e] = 1;
                        }
                        E.renderer->render(rb, ri);
                    } catch (const isti::Error &err) {
                        std::fprintf(stderr, "render failed: %s\n", err.what());
                        return 2;
                    }
                    HIP_OK(hipEventRecord(ev1, nullptr));
                    HIP_OK(hipEventSynchronize(ev1));
                    float ms = 0.f;
                    HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
                    ms_total += ms;
                    rendered = true;
                }
                if (write_images && !g_dry_run) {
                    HIP_OK(hipMemcpy(host_img.data(), E.out, host_img.size(), hipMemcpyDeviceToHost));
                    char name[64];
                    if (n_eyes == 2) std::snprintf(name, sizeof(name), "%04d_%s.png", frame_number - 1, e == 0 ? "left" : "right");
                    else std::snprintf(name, sizeof(name), "%04d.png", frame_number - 1);
                    write_png(out_dir / name, host_img.data(), W, H);
                }
            }
            if (rendered) frame_ms.push_back(ms_total);
            if (!g_dry_run) {
                std::printf("Frame: %3d / %3d  (%.3f ms)\r", frame_number, total, ms_total);
                std::fflush(stdout);
            }
            if (latency > b && b == fpa - 1) read_next();  // main.cu:188-189
        }
    }
    std::printf("\n");
    if (!frame_ms.empty()) {
        std::vector<float> s = frame_ms;
        std::sort(s.begin(), s.end());
        double sum = 0;
        for (float v : s) sum += v;
        const double mpix = (double)W * H * n_eyes * s.size() / (sum * 1e-3) / 1e6;
        std::printf("frames %zu  mean %.3f ms  p50 %.3f ms  p99 %.3f ms  %.1f Mpix/s (GPU time per frame)\n",
                    s.size(), sum / s.size(), s[s.size() / 2], s[std::min(s.size() - 1, (size_t)(0.99 * s.size()))], mpix);
    }
    for (auto &E : eyes) delete E.renderer;
    return 0;
}
