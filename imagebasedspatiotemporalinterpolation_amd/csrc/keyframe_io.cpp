// keyframe_io.cpp - reader of the reference's keyframe recordings (include/isti_io.h).
//
// Host code: cam.json (a small JSON reader), PNG and OpenEXR decoding on top
// of zlib, and the conversions FrameFileReader applies on the GPU in the
// reference (depth -> NDC z, projection normalisation), here on the host
// since they are per-keyframe, not per-frame, work.

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/isti_io.h"

namespace {

thread_local std::string g_err;

isti_status fail(isti_status st, const std::string &msg) {
    g_err = msg;
    return st;
}

bool read_file(const std::string &path, std::vector<uint8_t> &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

bool is_dir(const std::string &path) {
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

// zlib stream -> exactly `size` bytes
bool inflate_exact(const uint8_t *src, size_t n, uint8_t *dst, size_t size) {
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = const_cast<Bytef *>(src);
    zs.avail_in = (uInt)n;
    zs.next_out = dst;
    zs.avail_out = (uInt)size;
    const int r = inflate(&zs, Z_FINISH);
    const bool ok = (r == Z_STREAM_END || r == Z_OK || r == Z_BUF_ERROR) && zs.total_out == size;
    inflateEnd(&zs);
    return ok;
}

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// ---------------------------------------------------------------------------
// PNG (8-bit grey, grey+alpha, RGB, RGBA; not interlaced) -> RGBA8
// ---------------------------------------------------------------------------
isti_status decode_png(const std::string &path, uint32_t &w, uint32_t &h, std::vector<uint8_t> *rgba) {
    std::vector<uint8_t> f;
    if (!read_file(path, f)) return fail(ISTI_ERR_INVALID_ARG, "cannot read " + path);
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return fail(ISTI_ERR_INVALID_ARG, path + ": not a PNG");
    size_t p = 8;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat;
    w = h = 0;
    while (p + 12 <= f.size()) {
        const uint32_t len = be32(&f[p]);
        const char *type = reinterpret_cast<const char *>(&f[p + 4]);
        if (p + 12 + (size_t)len > f.size()) return fail(ISTI_ERR_INVALID_ARG, path + ": truncated chunk");
        const uint8_t *d = &f[p + 8];
        if (std::memcmp(type, "IHDR", 4) == 0 && len >= 13) {
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), d, d + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            break;
        }
        p += 12 + (size_t)len;
    }
    const int nch = ctype == 0 ? 1 : ctype == 4 ? 2 : ctype == 2 ? 3 : ctype == 6 ? 4 : 0;
    if (w == 0 || h == 0 || w > ISTI_MAX_DIM || h > ISTI_MAX_DIM || (uint64_t)w * h >= ISTI_MAX_VIEW_PIXELS)
        return fail(ISTI_ERR_UNSUPPORTED_SHAPE, path + ": bad size");
    if (depth != 8 || nch == 0 || interlace != 0)
        return fail(ISTI_ERR_INVALID_ARG, path + ": only 8-bit grey/grey-alpha/RGB/RGBA non-interlaced PNG");
    if (!rgba) return ISTI_OK;
    const size_t stride = (size_t)w * nch;
    std::vector<uint8_t> raw((stride + 1) * h);
    if (!inflate_exact(idat.data(), idat.size(), raw.data(), raw.size()))
        return fail(ISTI_ERR_INVALID_ARG, path + ": corrupt image data");
    std::vector<uint8_t> prev(stride, 0), cur(stride);
    rgba->resize((size_t)w * h * 4);
    for (uint32_t y = 0; y < h; y++) {
        const uint8_t filt = raw[y * (stride + 1)];
        const uint8_t *s = &raw[y * (stride + 1) + 1];
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= (size_t)nch ? cur[i - nch] : 0, b = prev[i], c = i >= (size_t)nch ? prev[i - nch] : 0;
            int v = s[i];
            switch (filt) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: {  // Paeth
                    const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: return fail(ISTI_ERR_INVALID_ARG, path + ": bad filter type");
            }
            cur[i] = (uint8_t)v;
        }
        uint8_t *o = &(*rgba)[(size_t)y * w * 4];
        for (uint32_t x = 0; x < w; x++) {
            const uint8_t *px = &cur[(size_t)x * nch];
            switch (nch) {
                case 1: o[4 * x] = o[4 * x + 1] = o[4 * x + 2] = px[0]; o[4 * x + 3] = 255; break;
                case 2: o[4 * x] = o[4 * x + 1] = o[4 * x + 2] = px[0]; o[4 * x + 3] = px[1]; break;
                case 3: o[4 * x] = px[0]; o[4 * x + 1] = px[1]; o[4 * x + 2] = px[2]; o[4 * x + 3] = 255; break;
                default: std::memcpy(&o[4 * x], px, 4); break;
            }
        }
        std::swap(prev, cur);
    }
    return ISTI_OK;
}

// ---------------------------------------------------------------------------
// OpenEXR, single-part scanline files: NONE / ZIPS / ZIP, HALF / FLOAT / UINT
// ---------------------------------------------------------------------------
float half_to_float(uint16_t hv) {
    const uint32_t s = (uint32_t)(hv & 0x8000u) << 16, e = (hv >> 10) & 0x1Fu, m = hv & 0x3FFu;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal half -> normal float
            int ex = -1;
            uint32_t mm = m;
            do {
                ex++;
                mm <<= 1;
            } while ((mm & 0x400u) == 0);
            bits = s | (uint32_t)(127 - 15 - ex) << 23 | (mm & 0x3FFu) << 13;
        }
    } else if (e == 31) {
        bits = s | 0x7F800000u | m << 13;
    } else {
        bits = s | (e + 112u) << 23 | m << 13;
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

struct ExrChannel {
    std::string name;
    int type;  // 0 UINT, 1 HALF, 2 FLOAT
};

isti_status decode_exr(const std::string &path, const std::vector<std::string> &want, uint32_t &w, uint32_t &h,
                       std::vector<float> *out, std::vector<std::string> *names = nullptr) {
    std::vector<uint8_t> f;
    if (!read_file(path, f)) return fail(ISTI_ERR_INVALID_ARG, "cannot read " + path);
    if (f.size() < 8 || f[0] != 0x76 || f[1] != 0x2f || f[2] != 0x31 || f[3] != 0x01)
        return fail(ISTI_ERR_INVALID_ARG, path + ": not an OpenEXR file");
    const uint32_t ver = f[4] | (uint32_t)f[5] << 8 | (uint32_t)f[6] << 16 | (uint32_t)f[7] << 24;
    if ((ver & 0xFF) != 2 || (ver & (0x200u | 0x800u | 0x1000u)))
        return fail(ISTI_ERR_INVALID_ARG, path + ": only single-part scanline OpenEXR files");
    size_t p = 8;
    auto cstr = [&](std::string &s) {
        const size_t e = std::find(f.begin() + (long)p, f.end(), 0) - f.begin();
        if (e >= f.size()) return false;
        s.assign(reinterpret_cast<const char *>(&f[p]), e - p);
        p = e + 1;
        return true;
    };
    auto rd32 = [&](size_t at) {
        int32_t v;
        std::memcpy(&v, &f[at], 4);
        return v;
    };
    std::vector<ExrChannel> chans;
    int compression = -1;
    int32_t xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    for (;;) {
        std::string name, type;
        if (!cstr(name)) return fail(ISTI_ERR_INVALID_ARG, path + ": truncated header");
        if (name.empty()) break;
        if (!cstr(type) || p + 4 > f.size()) return fail(ISTI_ERR_INVALID_ARG, path + ": truncated header");
        const int32_t size = rd32(p);
        p += 4;
        if (size < 0 || (size_t)size > f.size() - p) return fail(ISTI_ERR_INVALID_ARG, path + ": bad attribute");
        const size_t end = p + (size_t)size;
        if (name == "channels" && type == "chlist") {
            size_t q = p;
            while (q < end && f[q] != 0) {
                const size_t e = std::find(f.begin() + (long)q, f.begin() + (long)end, 0) - f.begin();
                if (e + 17 > end) return fail(ISTI_ERR_INVALID_ARG, path + ": bad channel list");
                ExrChannel c;
                c.name.assign(reinterpret_cast<const char *>(&f[q]), e - q);
                c.type = rd32(e + 1);
                if (rd32(e + 9) != 1 || rd32(e + 13) != 1)
                    return fail(ISTI_ERR_INVALID_ARG, path + ": subsampled channels are not supported");
                chans.push_back(c);
                q = e + 17;
            }
        } else if (name == "compression") {
            if (size < 1) return fail(ISTI_ERR_INVALID_ARG, path + ": bad compression attribute");
            compression = f[p];
        } else if (name == "dataWindow" && size == 16) {
            xmin = rd32(p);
            ymin = rd32(p + 4);
            xmax = rd32(p + 8);
            ymax = rd32(p + 12);
        }
        p = end;
    }
    if (xmax < xmin || ymax < ymin || chans.empty()) return fail(ISTI_ERR_INVALID_ARG, path + ": bad header");
    if (names) {
        names->clear();
        for (const ExrChannel &c : chans) names->push_back(c.name);
    }
    const int64_t w64 = (int64_t)xmax - xmin + 1, h64 = (int64_t)ymax - ymin + 1;
    if (w64 > ISTI_MAX_DIM || h64 > ISTI_MAX_DIM || w64 * h64 >= (int64_t)ISTI_MAX_VIEW_PIXELS)
        return fail(ISTI_ERR_UNSUPPORTED_SHAPE, path + ": too large");
    w = (uint32_t)w64;
    h = (uint32_t)h64;
    int lines;
    switch (compression) {
        case 0:
        case 2: lines = 1; break;
        case 3: lines = 16; break;
        default: return fail(ISTI_ERR_INVALID_ARG, path + ": compression must be NONE, ZIPS or ZIP");
    }
    if (!out) return ISTI_OK;
    const size_t nw = want.size();
    std::vector<int> src(nw, -1);
    for (size_t k = 0; k < nw; k++)
        for (size_t c = 0; c < chans.size(); c++)
            if (chans[c].name == want[k]) src[k] = (int)c;
    for (const ExrChannel &c : chans)
        if (c.type < 0 || c.type > 2) return fail(ISTI_ERR_INVALID_ARG, path + ": bad pixel type");
    size_t line_bytes = 0;
    std::vector<size_t> ch_off(chans.size());
    for (size_t c = 0; c < chans.size(); c++) {
        ch_off[c] = line_bytes;
        line_bytes += (size_t)w * (chans[c].type == 1 ? 2 : 4);
    }
    out->assign((size_t)w * h * nw, 0.0f);
    const size_t n_chunks = (h + (uint32_t)lines - 1) / (uint32_t)lines;
    if (8 * n_chunks > f.size() - p) return fail(ISTI_ERR_INVALID_ARG, path + ": truncated offsets");
    std::vector<uint8_t> raw, tmp;
    for (size_t k = 0; k < n_chunks; k++) {
        uint64_t off;
        std::memcpy(&off, &f[p + 8 * k], 8);
        // offsets come from the file: compare without forming off + 8 (it may wrap)
        if (f.size() < 8 || off > f.size() - 8) return fail(ISTI_ERR_INVALID_ARG, path + ": bad chunk offset");
        const int64_t y0_64 = (int64_t)rd32((size_t)off) - ymin;
        const int32_t n = rd32((size_t)off + 4);
        if (y0_64 < 0 || y0_64 >= (int64_t)h || n < 0 || (uint64_t)n > f.size() - 8 - off)
            return fail(ISTI_ERR_INVALID_ARG, path + ": bad chunk");
        const int32_t y0 = (int32_t)y0_64;
        const uint32_t nl = std::min<uint32_t>((uint32_t)lines, h - (uint32_t)y0);
        const size_t size = line_bytes * nl;
        const uint8_t *data = &f[(size_t)off + 8];
        if ((size_t)n == size) {  // stored uncompressed
            raw.assign(data, data + size);
        } else {
            if (compression == 0) return fail(ISTI_ERR_INVALID_ARG, path + ": bad chunk size");
            tmp.resize(size);
            if (!inflate_exact(data, (size_t)n, tmp.data(), size))
                return fail(ISTI_ERR_INVALID_ARG, path + ": corrupt chunk");
            for (size_t i = 1; i < size; i++) tmp[i] = (uint8_t)(tmp[i - 1] + tmp[i] - 128);  // predictor
            raw.resize(size);
            const uint8_t *t1 = tmp.data(), *t2 = tmp.data() + (size + 1) / 2;
            for (size_t i = 0; i < size; i++) raw[i] = (i & 1) ? *t2++ : *t1++;  // de-interleave
        }
        for (uint32_t l = 0; l < nl; l++) {
            const uint8_t *line = raw.data() + line_bytes * l;
            float *o = out->data() + ((size_t)(y0 + (int32_t)l) * w) * nw;
            for (size_t kk = 0; kk < nw; kk++) {
                if (src[kk] < 0) continue;
                const ExrChannel &c = chans[(size_t)src[kk]];
                const uint8_t *cp = line + ch_off[(size_t)src[kk]];
                for (uint32_t x = 0; x < w; x++) {
                    float v;
                    if (c.type == 1) {
                        uint16_t hv;
                        std::memcpy(&hv, cp + 2 * x, 2);
                        v = half_to_float(hv);
                    } else if (c.type == 2) {
                        std::memcpy(&v, cp + 4 * x, 4);
                    } else {
                        uint32_t u;
                        std::memcpy(&u, cp + 4 * x, 4);
                        v = (float)u;
                    }
                    o[x * nw + kk] = v;
                }
            }
        }
    }
    return ISTI_OK;
}

std::vector<std::string> split_names(const char *s) {
    std::vector<std::string> r;
    std::string cur;
    for (const char *c = s; *c; c++) {
        if (*c == ',') {
            r.push_back(cur);
            cur.clear();
        } else {
            cur += *c;
        }
    }
    r.push_back(cur);
    return r;
}

// ---------------------------------------------------------------------------
// JSON (objects, arrays, numbers, strings, true/false/null)
// ---------------------------------------------------------------------------
struct JVal {
    enum Type { NUL, BOOL, NUM, STR, ARR, OBJ } type = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
    const JVal *get(const std::string &k) const {
        const auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

struct JParser {
    const std::string &s;
    size_t p = 0;
    explicit JParser(const std::string &src) : s(src) {}
    void ws() {
        while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
    }
    bool parse(JVal &v, int depth = 0) {
        if (depth > 64) return false;
        ws();
        if (p >= s.size()) return false;
        const char c = s[p];
        if (c == '{') {
            v.type = JVal::OBJ;
            p++;
            ws();
            if (p < s.size() && s[p] == '}') return ++p, true;
            for (;;) {
                JVal key;
                ws();
                if (!parse_str(key.str)) return false;
                ws();
                if (p >= s.size() || s[p] != ':') return false;
                p++;
                if (!parse(v.obj[key.str], depth + 1)) return false;
                ws();
                if (p < s.size() && s[p] == ',') { p++; continue; }
                if (p < s.size() && s[p] == '}') return ++p, true;
                return false;
            }
        }
        if (c == '[') {
            v.type = JVal::ARR;
            p++;
            ws();
            if (p < s.size() && s[p] == ']') return ++p, true;
            for (;;) {
                v.arr.emplace_back();
                if (!parse(v.arr.back(), depth + 1)) return false;
                ws();
                if (p < s.size() && s[p] == ',') { p++; continue; }
                if (p < s.size() && s[p] == ']') return ++p, true;
                return false;
            }
        }
        if (c == '"') {
            v.type = JVal::STR;
            return parse_str(v.str);
        }
        if (s.compare(p, 4, "true") == 0) { v.type = JVal::BOOL; v.b = true; p += 4; return true; }
        if (s.compare(p, 5, "false") == 0) { v.type = JVal::BOOL; p += 5; return true; }
        if (s.compare(p, 4, "null") == 0) { p += 4; return true; }
        const char *b = s.c_str() + p;
        char *e = nullptr;
        v.num = std::strtod(b, &e);
        if (e == b) return false;
        v.type = JVal::NUM;
        p += (size_t)(e - b);
        return true;
    }
    bool parse_str(std::string &out) {
        if (p >= s.size() || s[p] != '"') return false;
        p++;
        out.clear();
        while (p < s.size() && s[p] != '"') {
            char c = s[p++];
            if (c == '\\' && p < s.size()) {
                const char e = s[p++];
                c = e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'b' ? '\b' : e == 'f' ? '\f' : e;
                if (e == 'u') {  // keep ASCII, replace the rest
                    if (p + 4 > s.size()) return false;
                    const unsigned cp = (unsigned)std::strtoul(s.substr(p, 4).c_str(), nullptr, 16);
                    p += 4;
                    c = cp < 128 ? (char)cp : '?';
                }
            }
            out += c;
        }
        if (p >= s.size()) return false;
        p++;
        return true;
    }
};

bool json_mat(const JVal *m, float out[16]) {  // "eRC" = row R, column C -> column-major
    if (!m || m->type != JVal::OBJ) return false;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            const JVal *e = m->get("e" + std::to_string(r) + std::to_string(c));
            if (!e || e->type != JVal::NUM) return false;
            out[c * 4 + r] = (float)e->num;
        }
    return true;
}

int json_int(const JVal &o, const char *k, int dflt) {
    const JVal *v = o.get(k);
    return v && v->type == JVal::NUM ? (int)v->num : dflt;
}
bool json_bool(const JVal &o, const char *k, bool dflt) {
    const JVal *v = o.get(k);
    return v && v->type == JVal::BOOL ? v->b : dflt;
}

// ---------------------------------------------------------------------------
// matrices (column-major, glm layout)
// ---------------------------------------------------------------------------
// A * B evaluated column by column as glm does: r[c] = ((A0 b0 + A1 b1) + A2 b2) + A3 b3
void mat_mul(const float *A, const float *B, float *R) {
    float t[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++)
            t[c * 4 + r] = ((A[r] * B[c * 4] + A[4 + r] * B[c * 4 + 1]) + A[8 + r] * B[c * 4 + 2]) + A[12 + r] * B[c * 4 + 3];
    std::memcpy(R, t, sizeof(t));
}

// Inverse in double (Gauss-Jordan with partial pivoting), rounded to float.
bool mat_inverse(const float *M, float *out) {
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = M[c * 4 + r];
            a[r][c + 4] = r == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < 8; k++) std::swap(a[c][k], a[piv][k]);
        const double d = a[c][c];
        for (int k = 0; k < 8; k++) a[c][k] /= d;
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            const double m = a[r][c];
            for (int k = 0; k < 8; k++) a[r][k] -= m * a[c][k];
        }
    }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out[c * 4 + r] = (float)a[r][c + 4];
    return true;
}

// FrameFileReader::normalizeProjMat (file_reader.cu:219-230) with the
// reference's settings: right-handed, reversed Z, [0, 1] depth, Unity 2022.3.
void normalize_proj(const float *proj, float *out) {
    static const float reverse_z[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0, 0, 0, 1, 1};
    static const float depth_01_to_11[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 2, 0, 0, 0, -1, 1};
    static const float unity_2022_3[16] = {1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float p[16];
    mat_mul(reverse_z, proj, p);
    mat_mul(depth_01_to_11, p, p);
    mat_mul(unity_2022_3, p, out);
}

// glm 0.9.9 gtx/matrix_interpolation, single precision (column-major, m[c][r] = m[4c + r]).
void axis_angle(const float *m, float axis[3], float &angle) {
    const float eps = std::numeric_limits<float>::epsilon() * 1e2f;
    auto M = [&](int c, int r) { return m[4 * c + r]; };
    const bool near_sym = std::fabs(M(1, 0) - M(0, 1)) < eps && std::fabs(M(2, 0) - M(0, 2)) < eps &&
                          std::fabs(M(2, 1) - M(1, 2)) < eps;
    if (near_sym) {
        const bool near_id = std::fabs(M(1, 0) + M(0, 1)) < eps && std::fabs(M(2, 0) + M(0, 2)) < eps &&
                             std::fabs(M(2, 1) + M(1, 2)) < eps && std::fabs(M(0, 0) + M(1, 1) + M(2, 2) - 3.0f) < eps;
        if (near_id) {
            angle = 0.0f;
            axis[0] = 1.0f;
            axis[1] = axis[2] = 0.0f;
            return;
        }
        angle = 3.14159265358979323846f;
        const float xx = (M(0, 0) + 1.0f) * 0.5f, yy = (M(1, 1) + 1.0f) * 0.5f, zz = (M(2, 2) + 1.0f) * 0.5f;
        const float xy = (M(1, 0) + M(0, 1)) * 0.25f, xz = (M(2, 0) + M(0, 2)) * 0.25f,
                    yz = (M(2, 1) + M(1, 2)) * 0.25f;
        if (xx > yy && xx > zz) {
            if (xx < eps) { axis[0] = 0.0f; axis[1] = 0.7071f; axis[2] = 0.7071f; }
            else { axis[0] = std::sqrt(xx); axis[1] = xy / axis[0]; axis[2] = xz / axis[0]; }
        } else if (yy > zz) {
            if (yy < eps) { axis[0] = 0.7071f; axis[1] = 0.0f; axis[2] = 0.7071f; }
            else { axis[1] = std::sqrt(yy); axis[0] = xy / axis[1]; axis[2] = yz / axis[1]; }
        } else {
            if (zz < eps) { axis[0] = 0.7071f; axis[1] = 0.7071f; axis[2] = 0.0f; }
            else { axis[2] = std::sqrt(zz); axis[0] = xz / axis[2]; axis[1] = yz / axis[2]; }
        }
        return;
    }
    const float c = (M(0, 0) + M(1, 1) + M(2, 2) - 1.0f) * 0.5f;
    angle = c >= 1.0f ? 0.0f : (c <= -1.0f ? 3.14159265358979323846f : std::acos(c));
    const float a0 = M(1, 2) - M(2, 1), a1 = M(2, 0) - M(0, 2), a2 = M(0, 1) - M(1, 0);
    const float inv = 1.0f / std::sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    axis[0] = a0 * inv;
    axis[1] = a1 * inv;
    axis[2] = a2 * inv;
}

void axis_angle_matrix(const float axis[3], float angle, float *R) {
    const float c = std::cos(angle), s = std::sin(angle), t = 1.0f - c;
    const float il = 1.0f / std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    const float x = axis[0] * il, y = axis[1] * il, z = axis[2] * il;
    const float r[16] = {t * x * x + c,     t * x * y + z * s, t * x * z - y * s, 0.0f,
                         t * x * y - z * s, t * y * y + c,     t * y * z + x * s, 0.0f,
                         t * x * z + y * s, t * y * z - x * s, t * z * z + c,     0.0f,
                         0.0f,              0.0f,              0.0f,              1.0f};
    std::memcpy(R, r, sizeof(r));
}

void interpolate(const float *m1, const float *m2, float delta, float *out) {
    float rot1[16] = {m1[0], m1[1], m1[2], 0.0f, m1[4], m1[5], m1[6], 0.0f,
                      m1[8], m1[9], m1[10], 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    float rot1_t[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) rot1_t[4 * c + r] = rot1[4 * r + c];
    float dlt[16];
    mat_mul(m2, rot1_t, dlt);
    float axis[3], angle;
    axis_angle(dlt, axis, angle);
    float R[16];
    axis_angle_matrix(axis, angle * delta, R);
    mat_mul(R, rot1, out);
    out[12] = m1[12] + delta * (m2[12] - m1[12]);
    out[13] = m1[13] + delta * (m2[13] - m1[13]);
    out[14] = m1[14] + delta * (m2[14] - m1[14]);
}

isti_status read_camera(const std::string &dir, isti_io_camera &cam, std::vector<isti_io_frame> &frames) {
    std::vector<uint8_t> raw;
    const std::string path = dir + "/cam.json";
    if (!read_file(path, raw)) return fail(ISTI_ERR_INVALID_ARG, "cannot read " + path);
    const std::string text(raw.begin(), raw.end());
    JVal root;
    JParser P(text);
    if (!P.parse(root) || root.type != JVal::OBJ) return fail(ISTI_ERR_INVALID_ARG, path + ": invalid JSON");
    const JVal *fr = root.get("frames"), *res = root.get("resolution");
    if (!fr || fr->type != JVal::ARR || fr->arr.empty())
        return fail(ISTI_ERR_INVALID_ARG, path + ": no frames");
    if (!res || res->type != JVal::OBJ) return fail(ISTI_ERR_INVALID_ARG, path + ": no resolution");
    cam.width = (uint32_t)json_int(*res, "x", 0);
    cam.height = (uint32_t)json_int(*res, "y", 0);
    cam.client_fps = json_int(root, "clientFps", 60);
    cam.server_fps = json_int(root, "serverFps", 10);
    cam.shadow_fps = json_int(root, "shadowFps", 60);
    cam.priority = json_bool(root, "objectMotion", true) * 2 + json_bool(root, "shadowGradients", true) * 1;
    frames.clear();
    for (const JVal &e : fr->arr) {
        isti_io_frame x;
        std::memset(&x, 0, sizeof(x));
        const JVal *idx = e.get("idx");
        if (e.type != JVal::OBJ || !idx || idx->type != JVal::NUM || !json_mat(e.get("w2c"), x.w2c) ||
            !json_mat(e.get("proj"), x.proj))
            return fail(ISTI_ERR_INVALID_ARG, path + ": bad frame entry");
        x.idx = (int32_t)idx->num;
        frames.push_back(x);
    }
    cam.n_frames = (int32_t)frames.size();
    return ISTI_OK;
}

std::string idx4(int idx) {
    char b[32];
    std::snprintf(b, sizeof(b), "%04d", idx);
    return b;
}

template <typename T>
T *dup(const std::vector<T> &v) {
    T *p = static_cast<T *>(std::malloc(v.size() * sizeof(T)));
    if (p) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace

extern "C" {

const char *isti_io_last_error(void) { return g_err.c_str(); }

isti_status isti_io_camera_at(const float view_fwd[16], const float view_bwd[16], const float proj_fwd[16], float t,
                              float view_t[16], float viewproj[16], float inv_viewproj[16]) {
    if (!view_fwd || !view_bwd || !proj_fwd) return fail(ISTI_ERR_INVALID_ARG, "matrices must be non-null");
    float v[16], vp[16], ivp[16];
    interpolate(view_fwd, view_bwd, t, v);
    mat_mul(proj_fwd, v, vp);
    if (!mat_inverse(vp, ivp)) return fail(ISTI_ERR_INVALID_ARG, "singular view-projection");
    if (view_t) std::memcpy(view_t, v, sizeof(v));
    if (viewproj) std::memcpy(viewproj, vp, sizeof(vp));
    if (inv_viewproj) std::memcpy(inv_viewproj, ivp, sizeof(ivp));
    return ISTI_OK;
}

isti_status isti_io_decode_png(const char *path, uint32_t *w, uint32_t *h, uint8_t *out, size_t out_bytes) {
    if (!path || !w || !h) return fail(ISTI_ERR_INVALID_ARG, "path, w and h must be non-null");
    std::vector<uint8_t> img;
    const isti_status st = decode_png(path, *w, *h, out ? &img : nullptr);
    if (st != ISTI_OK || !out) return st;
    if (img.size() > out_bytes)  // the file may have changed since the size query
        return fail(ISTI_ERR_INVALID_ARG, std::string(path) + ": image does not fit the output buffer");
    std::memcpy(out, img.data(), img.size());
    return ISTI_OK;
}

isti_status isti_io_decode_exr(const char *path, const char *channels, uint32_t *w, uint32_t *h, float *out,
                               size_t out_floats) {
    if (!path || !channels || !w || !h) return fail(ISTI_ERR_INVALID_ARG, "path, channels, w and h must be non-null");
    std::vector<float> img;
    const isti_status st = decode_exr(path, split_names(channels), *w, *h, out ? &img : nullptr);
    if (st != ISTI_OK || !out) return st;
    if (img.size() > out_floats)
        return fail(ISTI_ERR_INVALID_ARG, std::string(path) + ": image does not fit the output buffer");
    std::memcpy(out, img.data(), img.size() * sizeof(float));
    return ISTI_OK;
}

isti_status isti_io_client_camera(const float w2c[16], const float proj[16], float viewproj[16],
                                  float inv_viewproj[16]) {
    if (!w2c || !proj) return fail(ISTI_ERR_INVALID_ARG, "matrices must be non-null");
    float p[16], vp[16], ivp[16];
    normalize_proj(proj, p);
    mat_mul(p, w2c, vp);
    if (!mat_inverse(vp, ivp)) return fail(ISTI_ERR_INVALID_ARG, "singular view-projection");
    if (viewproj) std::memcpy(viewproj, vp, sizeof(vp));
    if (inv_viewproj) std::memcpy(inv_viewproj, ivp, sizeof(ivp));
    return ISTI_OK;
}

isti_status isti_io_read_camera(const char *cam_dir, isti_io_camera *cam, isti_io_frame *frames,
                                int32_t max_frames) {
    if (!cam_dir || !cam) return fail(ISTI_ERR_INVALID_ARG, "cam_dir and cam must be non-null");
    std::vector<isti_io_frame> fr;
    const isti_status st = read_camera(cam_dir, *cam, fr);
    if (st != ISTI_OK) return st;
    if (frames)
        for (int32_t i = 0; i < max_frames && i < (int32_t)fr.size(); i++) frames[i] = fr[(size_t)i];
    return ISTI_OK;
}

void isti_io_free_view(isti_io_view *v) {
    if (!v) return;
    std::free(v->color);
    std::free(v->z);
    std::free(v->motion_x);
    std::free(v->motion_y);
    std::free(v->motion_z);
    std::free(v->shadow_info);
    std::free(v->shadow_color_a);
    std::free(v->shadow_color_b);
    std::memset(v, 0, sizeof(*v));
}

isti_status isti_io_read_view(const char *cam_dir, int32_t pair, int32_t which, isti_io_view *out) {
    if (!cam_dir || !out || (which != 0 && which != 1) || pair < 0)
        return fail(ISTI_ERR_INVALID_ARG, "cam_dir and out must be non-null, which 0 or 1, pair >= 0");
    std::memset(out, 0, sizeof(*out));
    isti_io_camera cam;
    std::vector<isti_io_frame> fr;
    isti_status st = read_camera(cam_dir, cam, fr);
    if (st != ISTI_OK) return st;
    if (pair + 1 >= (int32_t)fr.size()) return fail(ISTI_ERR_INVALID_ARG, "keyframe pair beyond the recording");
    const std::string dir(cam_dir);
    const isti_io_frame &F = fr[(size_t)(pair + which)];
    const std::string name = idx4(F.idx), pair_name = idx4(fr[(size_t)pair].idx);
    // colour and depth of the frame itself (readColorDepthCam, file_reader.cu:138-160)
    uint32_t w = 0, h = 0;
    std::vector<uint8_t> color;
    if ((st = decode_png(dir + "/color/" + name + ".png", w, h, &color)) != ISTI_OK) return st;
    if (w != cam.width || h != cam.height)
        return fail(ISTI_ERR_UNSUPPORTED_SHAPE, "colour image size differs from cam.json resolution");
    // depth: channel "R", or the file's only channel (whatever its name)
    std::vector<float> depth;
    std::vector<std::string> names;
    uint32_t dw = 0, dh = 0;
    const std::string dpath = dir + "/depth/" + name + ".exr";
    if ((st = decode_exr(dpath, {}, dw, dh, nullptr, &names)) != ISTI_OK) return st;
    std::string dch;
    for (const std::string &n : names)
        if (n == "R") dch = n;
    if (dch.empty() && names.size() == 1) dch = names[0];
    if (dch.empty()) return fail(ISTI_ERR_INVALID_ARG, dpath + ": no R channel and more than one channel");
    if ((st = decode_exr(dpath, {dch}, dw, dh, &depth)) != ISTI_OK) return st;
    if (dw != w || dh != h) return fail(ISTI_ERR_UNSUPPORTED_SHAPE, "depth image size differs from colour");
    // convertDepthToZ (file_reader.cu:17-37): reversed Z, [0, 1] range, right-handed
    for (float &d : depth) {
        d = 1.0f - d;
        d = 2.0f * d - 1.0f;
    }
    out->width = w;
    out->height = h;
    std::memcpy(out->view, F.w2c, sizeof(out->view));
    normalize_proj(F.proj, out->proj);
    mat_mul(out->proj, F.w2c, out->viewproj);
    if (!mat_inverse(out->viewproj, out->inv_viewproj)) return fail(ISTI_ERR_INVALID_ARG, "singular view-projection");
    out->color = dup(color);
    out->z = dup(depth);
    // flow field and shadows of the pair, named by the pair's first frame (readNextFrame, :162-217)
    const std::string ff = dir + (which == 0 ? "/forwardFlowField" : "/backwardFlowField");
    if (is_dir(ff)) {
        std::vector<float> m;
        uint32_t mw = 0, mh = 0;
        if ((st = decode_exr(ff + "/" + pair_name + ".exr", {"R", "G", "B"}, mw, mh, &m)) != ISTI_OK) {
            isti_io_free_view(out);
            return st;
        }
        if (mw != w || mh != h) {
            isti_io_free_view(out);
            return fail(ISTI_ERR_UNSUPPORTED_SHAPE, "flow field size differs from colour");
        }
        std::vector<float> px((size_t)w * h), py(px.size()), pz(px.size());
        for (size_t i = 0; i < px.size(); i++) {
            px[i] = m[3 * i];
            py[i] = m[3 * i + 1];
            pz[i] = m[3 * i + 2];
        }
        out->motion_x = dup(px);
        out->motion_y = dup(py);
        out->motion_z = dup(pz);
    }
    const char *sh[3] = {"ShadowInfo", "ShadowColorA", "ShadowColorB"};
    uint8_t **dst[3] = {&out->shadow_info, &out->shadow_color_a, &out->shadow_color_b};
    for (int k = 0; k < 3; k++) {
        const std::string sd = dir + (which == 0 ? "/forward" : "/backward") + sh[k];
        if (!is_dir(sd)) continue;
        std::vector<uint8_t> img;
        uint32_t sw = 0, shh = 0;
        if ((st = decode_png(sd + "/" + pair_name + ".png", sw, shh, &img)) != ISTI_OK) {
            isti_io_free_view(out);
            return st;
        }
        if (sw != w || shh != h) {
            isti_io_free_view(out);
            return fail(ISTI_ERR_UNSUPPORTED_SHAPE, "shadow image size differs from colour");
        }
        *dst[k] = dup(img);
    }
    if (!out->color || !out->z) {
        isti_io_free_view(out);
        return fail(ISTI_ERR_OOM, "out of host memory");
    }
    return ISTI_OK;
}

}  // extern "C"
