// scene.cpp - seeded synthetic keyframe generator (analytic ray cast).
//
// Produces keyframes in the reference's formats (see include/isti_scene.h):
// the reference reads Unity recordings from disk (src/file_reader.cu:109-217);
// this generator replaces that input so the interpolation path can be driven
// and measured without a dataset. All geometry is evaluated in double and
// narrowed to float at the outputs. Deterministic for a given config/seed
// (no per-pixel randomness; OpenMP only splits rows).

#include "../../include/isti_scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 normalize(D3 a) { return (1.0 / std::sqrt(dot(a, a))) * a; }

struct Rng {  // splitmix64
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform(double lo, double hi) { return lo + (hi - lo) * ((next() >> 11) * (1.0 / 9007199254740992.0)); }
};

struct Sphere {
    D3 c0, v, amp, phase;
    double omega;
    double r;
    D3 albedo;
};

struct Scene {
    isti_scene_config cfg;
    std::vector<Sphere> spheres;
    D3 light{4.0, 10.0, -8.0};
    std::vector<D3> light_samples;

    D3 center(const Sphere &s, double t) const {
        D3 c = s.c0 + t * s.v;
        if (cfg.sinusoidal) {
            c.x += s.amp.x * std::sin(s.omega * t + s.phase.x);
            c.z += s.amp.z * std::sin(s.omega * t + s.phase.z);
        }
        return c;
    }
};

Scene make_scene(const isti_scene_config &cfg) {
    Scene sc;
    sc.cfg = cfg;
    Rng rng{cfg.seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
    for (int i = 0; i < cfg.n_spheres; i++) {
        Sphere s;
        s.r = rng.uniform(cfg.radius_min, cfg.radius_max);
        s.c0 = D3{rng.uniform(-7.0, 7.0), s.r + rng.uniform(0.0, 0.8), rng.uniform(-22.0, -3.5)};
        s.v = D3{rng.uniform(-cfg.speed_max, cfg.speed_max), 0.0, rng.uniform(-cfg.speed_max, cfg.speed_max)};
        s.amp = D3{rng.uniform(0.2, 0.8), 0.0, rng.uniform(0.2, 0.8)};
        s.phase = D3{rng.uniform(0.0, 6.283185307179586), 0.0, rng.uniform(0.0, 6.283185307179586)};
        s.omega = 6.283185307179586 / rng.uniform(12.0, 36.0);
        s.albedo = D3{rng.uniform(0.25, 0.95), rng.uniform(0.25, 0.95), rng.uniform(0.25, 0.95)};
        sc.spheres.push_back(s);
    }
    const int n = std::max(1, cfg.shadow_samples);
    for (int k = 0; k < n; k++) {  // Vogel spiral on a horizontal disk
        const double rr = cfg.light_radius_m * std::sqrt((k + 0.5) / n);
        const double th = k * 2.399963229728653;
        sc.light_samples.push_back(sc.light + D3{rr * std::cos(th), 0.0, rr * std::sin(th)});
    }
    return sc;
}

struct Camera {
    D3 pos, right, up, fwd;
    double tan_half, aspect;
    uint32_t w, h;
    double view[16];  // world -> camera, column-major
    double proj[16];  // OpenGL projection (y flipped), column-major
    double vp[16];    // proj * view
    double ivp[16];
};

bool invert4(const double *m, double *inv) {
    // cofactor expansion (column-major in, column-major out)
    double a[16];
    a[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    a[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    a[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    a[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    a[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    a[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    a[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    a[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    a[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    a[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    a[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    a[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    a[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    a[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    a[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    a[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * a[0] + m[1] * a[4] + m[2] * a[8] + m[3] * a[12];
    if (det == 0.0) return false;
    for (int i = 0; i < 16; i++) inv[i] = a[i] / det;
    return true;
}

Camera make_camera(const Scene &sc, double t, int view) {
    const isti_scene_config &c = sc.cfg;
    const double deg = 3.141592653589793 / 180.0;
    const double move_dir = 30.0 * deg;
    const D3 p0{0.0, 1.6, 0.0};
    const D3 dmove{std::sin(move_dir), 0.0, -std::cos(move_dir)};
    const double yaw = c.cam_yaw_deg * t * deg;
    const double pitch = -8.0 * deg;
    Camera cam;
    cam.fwd = D3{std::sin(yaw) * std::cos(pitch), std::sin(pitch), -std::cos(yaw) * std::cos(pitch)};
    cam.right = D3{std::cos(yaw), 0.0, std::sin(yaw)};
    cam.up = cross(cam.right, cam.fwd);
    cam.pos = p0 + (c.cam_speed * t) * dmove;
    double vfov = c.vfov_deg;
    cam.w = c.width;
    cam.h = c.height;
    if (c.n_eyes == 2) {
        cam.pos = cam.pos + ((view == 0 ? -0.5 : 0.5) * c.ipd_m) * cam.right;
    } else if (view >= 1) {
        cam.pos = cam.pos + ((view == 1 ? -1.0 : 1.0) * c.aux_offset_m) * cam.right;
        vfov = c.aux_vfov_deg;
        cam.w = c.aux_width;
        cam.h = c.aux_height;
    }
    cam.tan_half = std::tan(0.5 * vfov * deg);
    cam.aspect = (double)cam.w / (double)cam.h;
    // view matrix (world -> camera), column-major
    double V[16] = {0};
    V[0] = cam.right.x; V[4] = cam.right.y; V[8] = cam.right.z;
    V[1] = cam.up.x;    V[5] = cam.up.y;    V[9] = cam.up.z;
    V[2] = -cam.fwd.x;  V[6] = -cam.fwd.y;  V[10] = -cam.fwd.z;
    V[12] = -dot(cam.right, cam.pos);
    V[13] = -dot(cam.up, cam.pos);
    V[14] = dot(cam.fwd, cam.pos);
    V[15] = 1.0;
    // OpenGL projection, y flipped so that image row 0 is the top of the view
    const double f = 1.0 / cam.tan_half, N = c.near_m, F = c.far_m;
    double P[16] = {0};
    P[0] = f / cam.aspect;
    P[5] = -f;
    P[10] = (F + N) / (N - F);
    P[11] = -1.0;
    P[14] = 2.0 * F * N / (N - F);
    for (int col = 0; col < 4; col++)
        for (int row = 0; row < 4; row++) {
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += P[k * 4 + row] * V[col * 4 + k];
            cam.vp[col * 4 + row] = s;
        }
    invert4(cam.vp, cam.ivp);
    std::memcpy(cam.view, V, sizeof(V));
    std::memcpy(cam.proj, P, sizeof(P));
    return cam;
}

struct Hit {
    int obj;  // -1 sky, 0 ground, 1 + i sphere i
    double t;
    D3 p, n;
};

Hit trace(const Scene &sc, D3 o, D3 d, double time, const Camera &cam) {
    Hit h{-1, 1e30, {0, 0, 0}, {0, 0, 0}};
    const double cosf = dot(d, cam.fwd);
    auto accept = [&](double t) {
        const double depth = t * cosf;
        return depth >= sc.cfg.near_m && depth <= sc.cfg.far_m;
    };
    if (d.y < 0.0) {
        const double t = -o.y / d.y;
        if (t > 0.0 && t < h.t && accept(t)) {
            h.obj = 0;
            h.t = t;
        }
    }
    for (size_t i = 0; i < sc.spheres.size(); i++) {
        const Sphere &s = sc.spheres[i];
        const D3 oc = o - sc.center(s, time);
        const double b = dot(oc, d);
        const double cc = dot(oc, oc) - s.r * s.r;
        const double disc = b * b - cc;
        if (disc < 0.0) continue;
        const double sq = std::sqrt(disc);
        double t = -b - sq;
        if (!(t > 0.0 && accept(t))) t = -b + sq;
        if (t > 0.0 && t < h.t && accept(t)) {
            h.obj = 1 + (int)i;
            h.t = t;
        }
    }
    if (h.obj >= 0) {
        h.p = o + h.t * d;
        if (h.obj == 0) h.n = D3{0.0, 1.0, 0.0};
        else h.n = normalize(h.p - sc.center(sc.spheres[h.obj - 1], time));
    }
    return h;
}

// point on object `obj` that is at `p` at time t0, evaluated at time t1
D3 move_point(const Scene &sc, int obj, D3 p, double t0, double t1) {
    if (obj <= 0) return p;
    const Sphere &s = sc.spheres[obj - 1];
    return p + (sc.center(s, t1) - sc.center(s, t0));
}

// fraction of the area light visible from p (normal n) at time t
double visibility(const Scene &sc, int obj, D3 p, D3 n, double t) {
    const D3 po = p + 1e-3 * n;
    const D3 seg = sc.light - po;
    const double seg_len2 = dot(seg, seg);
    int cand[64];
    int nc = 0;
    D3 cent[64];
    for (size_t i = 0; i < sc.spheres.size() && nc < 64; i++) {
        if ((int)i + 1 == obj) continue;
        const Sphere &s = sc.spheres[i];
        const D3 c = sc.center(s, t);
        const double u = std::clamp(dot(c - po, seg) / seg_len2, 0.0, 1.0);
        const D3 q = po + u * seg;
        const D3 dq = c - q;
        const double lim = s.r + sc.cfg.light_radius_m;
        if (dot(dq, dq) <= lim * lim) {
            cand[nc] = (int)i;
            cent[nc] = c;
            nc++;
        }
    }
    if (nc == 0) return 1.0;
    int free_n = 0;
    for (const D3 &ls : sc.light_samples) {
        const D3 d = ls - po;
        bool blocked = false;
        for (int k = 0; k < nc && !blocked; k++) {
            const Sphere &s = sc.spheres[cand[k]];
            const D3 oc = po - cent[k];
            const double a = dot(d, d);
            const double b = dot(oc, d);
            const double cc = dot(oc, oc) - s.r * s.r;
            const double disc = b * b - a * cc;
            if (disc < 0.0) continue;
            const double sq = std::sqrt(disc);
            const double t0 = (-b - sq) / a, t1 = (-b + sq) / a;
            if ((t0 > 0.0 && t0 < 1.0) || (t1 > 0.0 && t1 < 1.0)) blocked = true;
        }
        if (!blocked) free_n++;
    }
    return (double)free_n / (double)sc.light_samples.size();
}

inline int shadow_code(double vis) {
    int c = (int)std::lround((1.0 - vis) * 3.0);
    return std::clamp(c, 0, 3);
}

D3 albedo_of(const Scene &sc, const Hit &h, double t) {
    if (h.obj == 0) {
        const int cx = (int)std::floor(h.p.x), cz = (int)std::floor(h.p.z);
        D3 a = ((cx + cz) & 1) ? D3{0.85, 0.82, 0.75} : D3{0.25, 0.28, 0.33};
        const int fx = (int)std::floor(h.p.x * 4.0), fz = (int)std::floor(h.p.z * 4.0);
        const double m = ((fx + fz) & 1) ? 1.0 : 0.88;
        return m * a;
    }
    const Sphere &s = sc.spheres[h.obj - 1];
    const D3 l = h.p - sc.center(s, t);
    const double lon = std::atan2(l.z, l.x);
    const double lat = std::acos(std::clamp(l.y / s.r, -1.0, 1.0));
    const int iu = (int)std::floor((lon + 3.141592653589793) / 6.283185307179586 * 12.0);
    const int iv = (int)std::floor(lat / 3.141592653589793 * 6.0);
    const double m = ((iu + iv) & 1) ? 1.0 : 0.55;
    return m * s.albedo;
}

D3 shade(const Scene &sc, const Hit &h, double t, double vis) {
    const D3 a = albedo_of(sc, h, t);
    const double diff = std::max(0.0, dot(h.n, normalize(sc.light - h.p)));
    const double k = 0.3 + 0.7 * diff * vis;
    return k * a;
}

D3 sky(D3 d) {
    const double u = std::clamp(d.y * 2.0 + 0.2, 0.0, 1.0);
    return D3{0.75 + (0.35 - 0.75) * u, 0.85 + (0.55 - 0.85) * u, 0.95 + (0.85 - 0.95) * u};
}

inline uint8_t to_u8(double v) { return (uint8_t)std::clamp((int)std::lround(v * 255.0), 0, 255); }
inline void put(uint8_t *p, D3 c, uint8_t a = 255) {
    p[0] = to_u8(c.x);
    p[1] = to_u8(c.y);
    p[2] = to_u8(c.z);
    p[3] = a;
}

D3 pixel_dir(const Camera &cam, uint32_t x, uint32_t y) {
    const double nx = 2.0 * (x + 0.5) / cam.w - 1.0;
    const double ny = 2.0 * (y + 0.5) / cam.h - 1.0;
    const double u = nx * cam.aspect * cam.tan_half;
    const double v = -ny * cam.tan_half;  // y flipped: row 0 at the top
    return normalize(u * cam.right + v * cam.up + cam.fwd);
}

void clip_of(const Camera &cam, D3 p, double out[4]) {
    const double *m = cam.vp;
    double r[4];
    for (int row = 0; row < 4; row++) r[row] = m[row] * p.x + m[4 + row] * p.y + m[8 + row] * p.z + m[12 + row];
    out[0] = r[0] / r[3];
    out[1] = r[1] / r[3];
    out[2] = r[2] / r[3];
    out[3] = 1.0;
}

int threads_of(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

bool valid_view(const isti_scene_config *cfg, int view) {
    if (cfg->n_eyes == 2) return view == 0 || view == 1;
    return view >= 0 && view <= cfg->n_aux;
}

}  // namespace

extern "C" int isti_scene_defaults(int workload, uint64_t seed, isti_scene_config *c) {
    if (!c) return -1;
    std::memset(c, 0, sizeof(*c));
    c->seed = seed;
    c->vfov_deg = 60.0f;
    c->near_m = 0.5f;
    c->far_m = 50.0f;
    c->n_eyes = 1;
    c->radius_min = 0.3f;
    c->radius_max = 1.5f;
    c->speed_max = 0.3f;
    c->cam_speed = 0.05f;
    c->cam_yaw_deg = 0.5f;
    c->shadow_samples = 16;
    c->light_radius_m = 0.5f;
    c->aux_offset_m = 0.3f;
    c->aux_vfov_deg = 75.0f;
    switch (workload) {
        case 1:
            c->width = 640; c->height = 360; c->n_spheres = 16; c->keyframe_spacing = 4;
            return 0;
        case 2:
            c->width = 1920; c->height = 1080; c->n_spheres = 16; c->keyframe_spacing = 8;
            c->n_aux = 2; c->aux_width = 960; c->aux_height = 540;
            c->shadows = 1; c->shadow_frames = 8;
            return 0;
        case 3:
            c->width = 2064; c->height = 2208; c->n_spheres = 64; c->keyframe_spacing = 6;
            c->n_eyes = 2; c->ipd_m = 0.063f; c->sinusoidal = 1;
            c->cam_speed = 0.15f; c->cam_yaw_deg = 2.0f; c->vfov_deg = 90.0f;
            return 0;
        default:
            return -1;
    }
}

extern "C" int isti_scene_camera(const isti_scene_config *cfg, double t, int view, float viewproj[16],
                                 float inv_viewproj[16], uint32_t *width, uint32_t *height) {
    if (!cfg || !valid_view(cfg, view)) return -1;
    const Scene sc = make_scene(*cfg);
    const Camera cam = make_camera(sc, t, view);
    for (int i = 0; i < 16; i++) {
        if (viewproj) viewproj[i] = (float)cam.vp[i];
        if (inv_viewproj) inv_viewproj[i] = (float)cam.ivp[i];
    }
    if (width) *width = cam.w;
    if (height) *height = cam.h;
    return 0;
}

extern "C" int isti_scene_camera_parts(const isti_scene_config *cfg, double t, int view, float view_m[16],
                                       float proj[16]) {
    if (!cfg || !valid_view(cfg, view)) return -1;
    const Scene sc = make_scene(*cfg);
    const Camera cam = make_camera(sc, t, view);
    for (int i = 0; i < 16; i++) {
        if (view_m) view_m[i] = (float)cam.view[i];
        if (proj) proj[i] = (float)cam.proj[i];
    }
    return 0;
}

extern "C" int isti_scene_keyframe(const isti_scene_config *cfg, int k, int view, int direction, uint8_t *color,
                                   float *z, float *mx, float *my, float *mz, uint8_t *shadow_info,
                                   uint8_t *shadow_a, uint8_t *shadow_b, float viewproj[16],
                                   float inv_viewproj[16], int n_threads) {
    if (!cfg || !valid_view(cfg, view) || (direction != 1 && direction != -1)) return -1;
    if (cfg->shadows && (cfg->shadow_frames < 1 || cfg->shadow_frames > 12)) return -1;
    const Scene sc = make_scene(*cfg);
    const double tk = (double)k * cfg->keyframe_spacing;
    const double t_other = tk + direction * (double)cfg->keyframe_spacing;
    const Camera cam = make_camera(sc, tk, view);
    for (int i = 0; i < 16; i++) {
        if (viewproj) viewproj[i] = (float)cam.vp[i];
        if (inv_viewproj) inv_viewproj[i] = (float)cam.ivp[i];
    }
    const uint32_t W = cam.w, H = cam.h;
    const int S = cfg->shadows ? cfg->shadow_frames : 1;
    const int nt = threads_of(n_threads);
#pragma omp parallel for schedule(dynamic, 2) num_threads(nt)
    for (int64_t yy = 0; yy < (int64_t)H; yy++) {
        const uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < W; x++) {
            const size_t i = (size_t)y * W + x;
            const D3 d = pixel_dir(cam, x, y);
            const Hit h = trace(sc, cam.pos, d, tk, cam);
            float zf = 1.0f;
            float om[3] = {0.f, 0.f, 0.f};
            uint8_t col[4], sa[4], sb[4], si[4] = {0, 0, 0, 0};
            if (h.obj < 0) {
                put(col, sky(d));
                std::memcpy(sa, col, 4);
                std::memcpy(sb, col, 4);
            } else {
                double clip[4];
                clip_of(cam, h.p, clip);
                zf = (float)clip[2];
                // the renderer unprojects (pixel2CS(x,y), z, 1): use its float pixel centre
                const float pcx = 2.0f * ((float)x + 0.5f) / (float)W - 1.0f;
                const float pcy = 2.0f * ((float)y + 0.5f) / (float)H - 1.0f;
                const D3 po = move_point(sc, h.obj, h.p, tk, t_other);
                double co[4];
                clip_of(cam, po, co);
                om[0] = (float)(co[0] - (double)pcx);
                om[1] = (float)(co[1] - (double)pcy);
                om[2] = (float)(co[2] - (double)zf);
                if (cfg->shadows) {
                    int codes[12];
                    int cmin = 3, cmax = 0;
                    for (int s = 0; s < S; s++) {
                        const double ts = tk + direction * (double)s * cfg->keyframe_spacing / (double)S;
                        const D3 ps = move_point(sc, h.obj, h.p, tk, ts);
                        codes[s] = shadow_code(visibility(sc, h.obj, ps, h.n, ts));
                        cmin = std::min(cmin, codes[s]);
                        cmax = std::max(cmax, codes[s]);
                    }
                    put(col, shade(sc, h, tk, 1.0 - codes[0] / 3.0));
                    if (cmin == cmax) {
                        std::memcpy(sa, col, 4);
                        std::memcpy(sb, col, 4);
                    } else {
                        uint32_t bits = 0;
                        for (int s = 0; s < S; s++) bits |= (uint32_t)codes[s] << (2 * s);
                        si[0] = (uint8_t)(bits >> 16);
                        si[1] = (uint8_t)(bits >> 8);
                        si[2] = (uint8_t)bits;
                        si[3] = 255;
                        put(sa, shade(sc, h, tk, 1.0 - cmin / 3.0));
                        put(sb, shade(sc, h, tk, 1.0 - cmax / 3.0));
                    }
                } else {
                    put(col, shade(sc, h, tk, 1.0));
                    std::memcpy(sa, col, 4);
                    std::memcpy(sb, col, 4);
                }
            }
            if (color) std::memcpy(color + 4 * i, col, 4);
            if (z) z[i] = zf;
            if (mx) mx[i] = om[0];
            if (my) my[i] = om[1];
            if (mz) mz[i] = om[2];
            if (shadow_info) std::memcpy(shadow_info + 4 * i, si, 4);
            if (shadow_a) std::memcpy(shadow_a + 4 * i, cfg->shadows ? sa : si, 4);
            if (shadow_b) std::memcpy(shadow_b + 4 * i, cfg->shadows ? sb : si, 4);
        }
    }
    return 0;
}

extern "C" int isti_scene_ground_truth(const isti_scene_config *cfg, double t, int view, uint8_t *color,
                                       int n_threads) {
    if (!cfg || !color || !valid_view(cfg, view)) return -1;
    const Scene sc = make_scene(*cfg);
    const Camera cam = make_camera(sc, t, view);
    const int nt = threads_of(n_threads);
#pragma omp parallel for schedule(dynamic, 2) num_threads(nt)
    for (int64_t yy = 0; yy < (int64_t)cam.h; yy++) {
        const uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < cam.w; x++) {
            const D3 d = pixel_dir(cam, x, y);
            const Hit h = trace(sc, cam.pos, d, t, cam);
            uint8_t *p = color + 4 * ((size_t)y * cam.w + x);
            if (h.obj < 0) {
                put(p, sky(d));
            } else {
                const double vis = cfg->shadows ? 1.0 - shadow_code(visibility(sc, h.obj, h.p, h.n, t)) / 3.0 : 1.0;
                put(p, shade(sc, h, t, vis));
            }
        }
    }
    return 0;
}
