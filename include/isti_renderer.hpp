// isti_renderer.hpp - C++ mirror of the reference's Renderer interface over
// the C ABI of isti.h (header-only).
//
// Reference interface: include/renderer.h:9-56 (RenderMode,
// AdditionalRenderInfo, RenderInfo, Renderer::render) and include/data.h
// (FrameData with fwd/bwd and aux cameras with curr/next and swap()).
// A host written against the reference keeps its call shape:
//     renderer.render(render_buffer, render_info, frame_data);
// Differences: keyframe images are SoA device planes described by
// isti_view_desc (instead of cudaArray-backed ImageBuffers), the render
// buffer is a plain device RGBA8 image, and failures throw isti::Error
// (the reference prints and swallows device errors).
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "isti.h"

namespace isti {

enum RenderMode {  // include/renderer.h:9-16
    OURS = ISTI_MODE_OURS,
    BACKWARD_BIDIRECTIONAL = ISTI_MODE_BACKWARD_BIDIRECTIONAL,
    BACKWARD_UNIDIRECTIONAL = ISTI_MODE_BACKWARD_UNIDIRECTIONAL,
    SPLATTING = ISTI_MODE_SPLATTING,
    TIMEWARP = ISTI_MODE_TIMEWARP
};

struct AdditionalRenderInfo {  // include/renderer.h:18-24
    int frames_per_anchor = 1;
    int shadow_frames_per_anchor = 0;
    bool no_shadows = false;
    RenderMode mode = OURS;
};

struct RenderInfo {  // include/renderer.h:26-35
    uint32_t resolution[2] = {0, 0};
    float alpha = 0.f;
    float viewproj_mat[16] = {0};
    float inv_viewproj_mat[16] = {0};
    AdditionalRenderInfo additional_info;
    int max_search_iterations = -1;  // < 0: the mode's own count
};

struct Error : std::runtime_error {
    isti_status status;
    Error(isti_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

// AuxCamData, include/data.h:64-77
struct AuxCamData {
    isti_view_desc curr{};
    isti_view_desc next{};
    int priority = 0;
    void swap() { std::swap(curr, next); }
};

// FrameData, include/data.h:79-116 (views hold device planes)
struct FrameData {
    explicit FrameData(size_t number_of_aux_cams = 0) : aux_cam_images(number_of_aux_cams) {}
    isti_view_desc fwd{};
    isti_view_desc bwd{};
    std::vector<AuxCamData> aux_cam_images;
    void swap() {
        std::swap(fwd, bwd);
        for (auto &a : aux_cam_images) a.swap();
    }
};

// The caller-owned persistent output (the reference's ImageBuffer<uchar4,4>).
struct RenderBuffer {
    uint8_t *data = nullptr;  // device RGBA8
    size_t pitch_bytes = 0;
    uint32_t width = 0, height = 0;
};

class Renderer {
public:
    explicit Renderer(size_t n_aux = 0, int device = 0) {
        isti_config c{device, (int32_t)n_aux};
        check(isti_create(&c, &r_), nullptr);
    }
    ~Renderer() {
        if (r_) isti_destroy(r_);
    }
    Renderer(const Renderer &) = delete;
    Renderer &operator=(const Renderer &) = delete;

    // Renderer::render, include/renderer.h:56 / src/renderer.cu:607-694.
    // Stream-ordered; returns without waiting for the GPU. Sets the keyframes
    // on every call, as the reference takes its FrameData on every call.
    void render(RenderBuffer &render_buffer, const RenderInfo &info, const FrameData &frame_data,
                void *stream = nullptr) {
        set_frame(frame_data);
        render(render_buffer, info, stream);
    }

    // Sets the keyframe pair once for the renders that follow; those reuse
    // the pair's flow toward the other keyframe (computed by the first of
    // them). Call again when the pair, or a plane's contents, change.
    void set_frame(const FrameData &frame_data) {
        set(ISTI_SLOT_FWD, 0, frame_data.fwd, 0);
        set(ISTI_SLOT_BWD, 0, frame_data.bwd, 0);
        for (size_t i = 0; i < frame_data.aux_cam_images.size(); i++) {
            const AuxCamData &a = frame_data.aux_cam_images[i];
            set(ISTI_SLOT_FWD, (int)i + 1, a.curr, a.priority);
            set(ISTI_SLOT_BWD, (int)i + 1, a.next, a.priority);
        }
    }

    // A render of the pair last passed to set_frame.
    void render(RenderBuffer &render_buffer, const RenderInfo &info, void *stream = nullptr) {
        isti_render_info ri;
        std::memset(&ri, 0, sizeof(ri));
        ri.width = info.resolution[0];
        ri.height = info.resolution[1];
        ri.alpha = info.alpha;
        std::memcpy(ri.viewproj, info.viewproj_mat, sizeof(ri.viewproj));
        std::memcpy(ri.inv_viewproj, info.inv_viewproj_mat, sizeof(ri.inv_viewproj));
        ri.frames_per_anchor = info.additional_info.frames_per_anchor;
        ri.shadow_frames_per_anchor = info.additional_info.shadow_frames_per_anchor;
        ri.no_shadows = info.additional_info.no_shadows ? 1 : 0;
        ri.mode = (int32_t)info.additional_info.mode;
        ri.max_search_iterations = info.max_search_iterations;
        check(isti_render(r_, &ri, render_buffer.data, render_buffer.pitch_bytes, stream), r_);
    }

    isti_renderer *handle() { return r_; }

private:
    void set(int slot, int view, isti_view_desc d, int priority) {
        d.priority = priority;
        check(isti_keyframe_set(r_, slot, view, &d), r_);
    }
    static void check(isti_status s, const isti_renderer *r) {
        if (s != ISTI_OK) throw Error(s, isti_last_error(r));
    }
    isti_renderer *r_ = nullptr;
};

}  // namespace isti
