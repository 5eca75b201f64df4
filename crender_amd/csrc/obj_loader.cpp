// obj_loader.cpp -- Wavefront OBJ/MTL ingest for the crt scene loader.
//
// Mirrors cr::asset_loader::load_model (CRender src/util/asset_loader.cpp:182-303): the
// reference parses with tinyobjloader (external/tinyobj/tinobj.h) in triangulating mode
// and keeps, per material, only the diffuse colour (Kd) and the diffuse texture (map_Kd,
// loaded flipped vertically, RGBA/255); every material is "smooth" with emission 0.
// Triangulation follows tinyobjloader's published rules (tinobj.h:1393-1714):
//   * a quad is split along its shorter diagonal: [0,1,2],[0,2,3] if |v2-v0| < |v3-v1|,
//     else [0,1,3],[1,2,3];
//   * a larger polygon is ear-clipped in the plane of its two dominant axes;
//   * faces with fewer than 3 vertices are dropped; indices may be negative (relative).
// Differences, each where the reference would read out of bounds: faces before any
// `usemtl` (material -1 in the reference) get an appended default material, corners
// without a texture coordinate get an appended (0, 0) coordinate, and unknown indices are
// reported as errors instead of being skipped silently.
//
// Textures: binary/ASCII PPM (P6/P3) and 8-bit non-interlaced PNG (grey, grey + alpha,
// RGB, RGBA, palette) through zlib.  A texture that cannot be loaded leaves the
// material untextured, as the reference does (asset_loader.cpp:276-277).
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/crt.h"

namespace {

struct ObjIndex {
    int v = -1, vt = -1, vn = -1;
};

struct Mtl {
    std::string name;
    float       kd[3] = {1.0f, 1.0f, 1.0f};
    std::string map_kd;
};

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// OBJ index: 1-based, or negative = relative to the current count (tinobj.h:771-792)
bool fix_index(int idx, int n, int *out)
{
    if (idx > 0) {
        *out = idx - 1;
        return true;
    }
    if (idx < 0) {
        *out = n + idx;
        return true;
    }
    return false;
}

bool parse_corner(const std::string &tok, int nv, int nvt, int nvn, ObjIndex *out)
{
    int a = 0, b = 0, c = 0;
    const size_t s1 = tok.find('/');
    a = std::atoi(tok.c_str());
    if (!fix_index(a, nv, &out->v)) return false;
    if (s1 == std::string::npos) return true;
    const size_t s2 = tok.find('/', s1 + 1);
    const std::string t = tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
    if (!t.empty()) {
        b = std::atoi(t.c_str());
        if (!fix_index(b, nvt, &out->vt)) return false;
    }
    if (s2 != std::string::npos) {
        const std::string n = tok.substr(s2 + 1);
        if (!n.empty()) {
            c = std::atoi(n.c_str());
            if (!fix_index(c, nvn, &out->vn)) return false;
        }
    }
    return true;
}

// point-in-triangle test in 2D (crossing rule), used by ear clipping
bool in_triangle(const double *vx, const double *vy, double tx, double ty)
{
    bool c = false;
    for (int i = 0, j = 2; i < 3; j = i++) {
        if (((vy[i] > ty) != (vy[j] > ty)) && (tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i])) c = !c;
    }
    return c;
}

// Triangulates one polygon face; appends corner triples to tris.
void triangulate(const std::vector<ObjIndex> &face, const std::vector<float> &v, std::vector<ObjIndex> &tris)
{
    const size_t n = face.size();
    if (n < 3) return;
    auto P = [&](const ObjIndex &c, int k) { return (double)v[3 * (size_t)c.v + k]; };
    if (n == 3) {
        tris.insert(tris.end(), face.begin(), face.end());
        return;
    }
    if (n == 4) {
        double d02 = 0.0, d13 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double a = P(face[2], k) - P(face[0], k), b = P(face[3], k) - P(face[1], k);
            d02 += a * a;
            d13 += b * b;
        }
        if (d02 < d13) {
            const int o[6] = {0, 1, 2, 0, 2, 3};
            for (int i : o) tris.push_back(face[i]);
        } else {
            const int o[6] = {0, 1, 3, 1, 2, 3};
            for (int i : o) tris.push_back(face[i]);
        }
        return;
    }
    // dominant plane: the first corner with a non-degenerate turn decides the two axes
    int ax0 = 1, ax1 = 2;
    for (size_t k = 0; k < n; ++k) {
        const ObjIndex &a = face[k], &b = face[(k + 1) % n], &c = face[(k + 2) % n];
        double e0[3], e1[3];
        for (int q = 0; q < 3; ++q) {
            e0[q] = P(b, q) - P(a, q);
            e1[q] = P(c, q) - P(b, q);
        }
        const double cx = std::fabs(e0[1] * e1[2] - e0[2] * e1[1]);
        const double cy = std::fabs(e0[2] * e1[0] - e0[0] * e1[2]);
        const double cz = std::fabs(e0[0] * e1[1] - e0[1] * e1[0]);
        const double eps = std::numeric_limits<float>::epsilon();
        if (cx > eps || cy > eps || cz > eps) {
            if (!(cx > cy && cx > cz)) {
                ax0 = 0;
                if (cz > cx && cz > cy) ax1 = 1;
            }
            break;
        }
    }
    double area = 0.0;
    for (size_t k = 0; k < n; ++k) {
        const ObjIndex &a = face[k], &b = face[(k + 1) % n];
        area += (P(a, ax0) * P(b, ax1) - P(a, ax1) * P(b, ax0)) * 0.5;
    }
    std::vector<ObjIndex> rem = face;
    size_t guess = 0, iters = rem.size(), prev = rem.size();
    while (rem.size() > 3 && iters > 0) {
        const size_t m = rem.size();
        if (guess >= m) guess -= m;
        if (prev != m) {
            prev  = m;
            iters = m;
        } else {
            --iters;
        }
        ObjIndex ind[3];
        double   vx[3], vy[3];
        for (int k = 0; k < 3; ++k) {
            ind[k] = rem[(guess + k) % m];
            vx[k]  = P(ind[k], ax0);
            vy[k]  = P(ind[k], ax1);
        }
        const double cross = (vx[1] - vx[0]) * (vy[2] - vy[1]) - (vy[1] - vy[0]) * (vx[2] - vx[1]);
        if (cross * area < 0.0) {   // reflex corner
            guess += 1;
            continue;
        }
        bool overlap = false;
        for (size_t o = 3; o < m; ++o) {
            const ObjIndex &c = rem[(guess + o) % m];
            if (in_triangle(vx, vy, P(c, ax0), P(c, ax1))) {
                overlap = true;
                break;
            }
        }
        if (overlap) {
            guess += 1;
            continue;
        }
        tris.push_back(ind[0]);
        tris.push_back(ind[1]);
        tris.push_back(ind[2]);
        rem.erase(rem.begin() + (long)((guess + 1) % m));
    }
    if (rem.size() == 3) tris.insert(tris.end(), rem.begin(), rem.end());
}

// ---- texture decoding ---------------------------------------------------------------

bool read_file(const std::string &path, std::vector<uint8_t> &out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

// 8-bit PNG -> RGBA8, top row first
bool decode_png(const std::vector<uint8_t> &f, std::vector<uint8_t> &rgba, uint32_t &w, uint32_t &h)
{
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return false;
    size_t pos = 8;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    while (pos + 12 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        if (pos + 12 + (size_t)len > f.size()) return false;
        const char *type = reinterpret_cast<const char *>(&f[pos + 4]);
        const uint8_t *d = &f[pos + 8];
        if (!std::strncmp(type, "IHDR", 4) && len >= 13) {
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (!std::strncmp(type, "PLTE", 4)) {
            plte.assign(d, d + len);
        } else if (!std::strncmp(type, "tRNS", 4)) {
            trns.assign(d, d + len);
        } else if (!std::strncmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::strncmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    if (depth != 8 || interlace != 0 || w == 0 || h == 0 || w > (1u << 15) || h > (1u << 15)) return false;
    int ch = 0;
    switch (ctype) {
    case 0: ch = 1; break;
    case 2: ch = 3; break;
    case 3: ch = 1; break;
    case 4: ch = 2; break;
    case 6: ch = 4; break;
    default: return false;
    }
    if (ctype == 3 && plte.size() < 3) return false;
    const size_t stride = (size_t)w * ch;
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf rl = (uLongf)raw.size();
    if (uncompress(raw.data(), &rl, idat.data(), (uLong)idat.size()) != Z_OK || rl != raw.size()) return false;
    std::vector<uint8_t> img(stride * h), prev(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t *src = &raw[y * (stride + 1) + 1];
        uint8_t *dst = &img[y * stride];
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= (size_t)ch ? dst[x - ch] : 0, b = prev[x], c = x >= (size_t)ch ? prev[x - ch] : 0;
            int v = src[x];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: v += paeth(a, b, c); break;
            default: return false;
            }
            dst[x] = (uint8_t)v;
        }
        std::memcpy(prev.data(), dst, stride);
    }
    rgba.resize((size_t)w * h * 4);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint8_t *p = &img[i * ch];
        uint8_t *q = &rgba[i * 4];
        if (ctype == 0) q[0] = q[1] = q[2] = p[0], q[3] = 255;
        else if (ctype == 4) q[0] = q[1] = q[2] = p[0], q[3] = p[1];
        else if (ctype == 2) q[0] = p[0], q[1] = p[1], q[2] = p[2], q[3] = 255;
        else if (ctype == 6) std::memcpy(q, p, 4);
        else {
            const size_t k = p[0];
            if (3 * k + 2 >= plte.size()) return false;
            q[0] = plte[3 * k], q[1] = plte[3 * k + 1], q[2] = plte[3 * k + 2];
            q[3] = k < trns.size() ? trns[k] : 255;
        }
    }
    return true;
}

// PPM (P6 / P3, maxval <= 255) -> RGBA8, top row first
bool decode_ppm(const std::vector<uint8_t> &f, std::vector<uint8_t> &rgba, uint32_t &w, uint32_t &h)
{
    if (f.size() < 2 || f[0] != 'P' || (f[1] != '6' && f[1] != '3')) return false;
    const bool ascii = f[1] == '3';
    size_t pos = 2;
    auto next_int = [&](long &out) -> bool {
        for (;;) {
            while (pos < f.size() && std::isspace(f[pos])) ++pos;
            if (pos < f.size() && f[pos] == '#') {
                while (pos < f.size() && f[pos] != '\n') ++pos;
                continue;
            }
            break;
        }
        if (pos >= f.size() || !std::isdigit(f[pos])) return false;
        out = 0;
        while (pos < f.size() && std::isdigit(f[pos])) out = out * 10 + (f[pos++] - '0');
        return true;
    };
    long W = 0, H = 0, M = 0;
    if (!next_int(W) || !next_int(H) || !next_int(M) || W <= 0 || H <= 0 || M <= 0 || M > 255) return false;
    if (W > (1 << 15) || H > (1 << 15)) return false;
    w = (uint32_t)W;
    h = (uint32_t)H;
    rgba.resize((size_t)w * h * 4);
    if (!ascii) {
        ++pos;   // single whitespace after maxval
        if (pos + (size_t)w * h * 3 > f.size()) return false;
    }
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        for (int c = 0; c < 3; ++c) {
            long v = 0;
            if (ascii) {
                if (!next_int(v)) return false;
            } else {
                v = f[pos++];
            }
            rgba[4 * i + c] = (uint8_t)std::min<long>(255, v * 255 / M);
        }
        rgba[4 * i + 3] = 255;
    }
    return true;
}

// Loads a texture as the reference does: flipped vertically, RGBA / 255 (asset_loader.cpp:240-270).
bool load_texture(const std::string &path, std::vector<float> &out, uint32_t &w, uint32_t &h)
{
    std::vector<uint8_t> f, rgba;
    if (!read_file(path, f)) return false;
    if (!decode_png(f, rgba, w, h) && !decode_ppm(f, rgba, w, h)) return false;
    out.resize((size_t)w * h * 4);
    for (uint32_t y = 0; y < h; ++y) {
        const uint32_t sy = h - 1 - y;   // stbi_set_flip_vertically_on_load(true)
        for (size_t x = 0; x < (size_t)w * 4; ++x) out[(size_t)y * w * 4 + x] = rgba[(size_t)sy * w * 4 + x] / 255.f;
    }
    return true;
}

std::vector<Mtl> parse_mtl(const std::string &path)
{
    std::vector<Mtl> out;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "newmtl") {
            out.emplace_back();
            std::getline(ss, out.back().name);
            out.back().name = trim(out.back().name);
        } else if (out.empty()) {
            continue;
        } else if (key == "Kd") {
            ss >> out.back().kd[0] >> out.back().kd[1] >> out.back().kd[2];
        } else if (key == "map_Kd") {
            std::string rest;
            std::getline(ss, rest);
            rest = trim(rest);
            // the file name is the last token (options such as -s come first)
            const size_t sp = rest.find_last_of(" \t");
            out.back().map_kd = sp == std::string::npos ? rest : rest.substr(sp + 1);
        }
    }
    return out;
}

}  // namespace

struct crt_obj {
    std::string                     name;
    std::vector<float>              vertices, texcoords;
    std::vector<uint32_t>           vidx, tidx, midx;
    std::vector<crt_material>       mats;
    std::vector<std::vector<float>> tex;
    std::vector<crt_image>          images;
    std::vector<uint32_t>           tex_dims;
    std::string                     warnings;
};

namespace crt {
int set_error(int code, const std::string &msg);   // api.cpp: the message crt_last_error returns
}

namespace {
int obj_fail(int code, const std::string &m) { return crt::set_error(code, m); }
}  // namespace

extern "C" {

int crt_obj_load(const char *file, const char *folder, crt_obj **out)
{
    if (!file || !out) return obj_fail(CRT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::ifstream f(file);
    if (!f) return obj_fail(CRT_ERR_INVALID_ARG, std::string("cannot open ") + file);
    std::string dir = folder ? folder : "";
    if (dir.empty()) {
        const std::string p(file);
        const size_t s = p.find_last_of('/');
        dir = s == std::string::npos ? "." : p.substr(0, s);
    }
    auto o = std::make_unique<crt_obj>();
    {
        std::string p(file);
        const size_t s = p.find_last_of('/');
        std::string base = s == std::string::npos ? p : p.substr(s + 1);
        const size_t dot = base.find_last_of('.');
        o->name = dot == std::string::npos ? base : base.substr(0, dot);
    }
    std::vector<float> vn;
    std::vector<Mtl>   mtls;
    std::unordered_map<std::string, int> mtl_of;
    int cur_mat = -1;
    bool need_default_mat = false, need_default_uv = false;
    std::vector<ObjIndex> tris;
    std::vector<int>      tri_mat;
    std::string line;
    size_t lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        // line continuation
        while (!line.empty() && line.back() == '\\') {
            line.pop_back();
            std::string more;
            if (!std::getline(f, more)) break;
            line += more;
            ++lineno;
        }
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        std::istringstream ss(t);
        std::string key;
        ss >> key;
        if (key == "v") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            o->vertices.insert(o->vertices.end(), {x, y, z});
        } else if (key == "vt") {
            float u = 0, v = 0;
            ss >> u >> v;
            o->texcoords.insert(o->texcoords.end(), {u, v});
        } else if (key == "vn") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            vn.insert(vn.end(), {x, y, z});
        } else if (key == "f") {
            std::vector<ObjIndex> face;
            std::string tok;
            const int nv = (int)(o->vertices.size() / 3), nvt = (int)(o->texcoords.size() / 2), nvn = (int)(vn.size() / 3);
            while (ss >> tok) {
                ObjIndex c;
                if (!parse_corner(tok, nv, nvt, nvn, &c) || c.v < 0 || c.v >= nv || c.vt >= nvt)
                    return obj_fail(CRT_ERR_INVALID_MESH, "bad face index at line " + std::to_string(lineno));
                face.push_back(c);
            }
            if (face.size() < 3) {
                o->warnings += "degenerate face at line " + std::to_string(lineno) + "\n";
                continue;
            }
            const size_t before = tris.size();
            triangulate(face, o->vertices, tris);
            for (size_t k = before; k < tris.size(); k += 3) tri_mat.push_back(cur_mat);
        } else if (key == "mtllib") {
            std::string rest;
            std::getline(ss, rest);
            rest = trim(rest);
            for (const Mtl &m : parse_mtl(dir + "/" + rest)) {
                mtl_of[m.name] = (int)mtls.size();
                mtls.push_back(m);
            }
        } else if (key == "usemtl") {
            std::string name;
            std::getline(ss, name);
            name = trim(name);
            const auto it = mtl_of.find(name);
            cur_mat = it == mtl_of.end() ? -1 : it->second;
        }
        // o, g, s, l, p and others carry nothing the renderer uses
    }
    // materials (asset_loader.cpp:227-283)
    std::unordered_map<std::string, int> tex_of;
    for (const Mtl &m : mtls) {
        crt_material cm;
        crt_material_default(&cm);
        cm.type = CRT_SMOOTH;
        cm.colour[0] = m.kd[0];
        cm.colour[1] = m.kd[1];
        cm.colour[2] = m.kd[2];
        cm.colour[3] = 1.0f;
        cm.emission  = 0.0f;
        if (!m.map_kd.empty()) {
            const auto it = tex_of.find(m.map_kd);
            if (it != tex_of.end()) {
                cm.tex = it->second;
            } else {
                std::vector<float> px;
                uint32_t w = 0, h = 0;
                if (load_texture(dir + "/" + m.map_kd, px, w, h)) {
                    o->tex.push_back(std::move(px));
                    o->tex_dims.insert(o->tex_dims.end(), {w, h});
                    cm.tex = (int32_t)o->tex.size() - 1;
                    tex_of[m.map_kd] = cm.tex;
                } else {
                    o->warnings += "failed to load texture " + m.map_kd + "\n";
                }
            }
        }
        o->mats.push_back(cm);
    }
    const size_t ntri = tris.size() / 3;
    for (size_t t = 0; t < ntri; ++t) {
        if (tri_mat[t] < 0) need_default_mat = true;
        for (int c = 0; c < 3; ++c)
            if (tris[3 * t + c].vt < 0) need_default_uv = true;
    }
    const uint32_t default_mat = (uint32_t)o->mats.size();
    if (need_default_mat) {
        crt_material cm;
        crt_material_default(&cm);
        cm.type = CRT_SMOOTH;
        o->mats.push_back(cm);
    }
    const bool any_uv = !o->texcoords.empty();
    const uint32_t default_uv = (uint32_t)(o->texcoords.size() / 2);
    if (any_uv && need_default_uv) o->texcoords.insert(o->texcoords.end(), {0.0f, 0.0f});
    o->vidx.resize(3 * ntri);
    o->midx.resize(ntri);
    if (any_uv) o->tidx.resize(3 * ntri);
    for (size_t t = 0; t < ntri; ++t) {
        for (int c = 0; c < 3; ++c) {
            const ObjIndex &x = tris[3 * t + c];
            o->vidx[3 * t + c] = (uint32_t)x.v;
            if (any_uv) o->tidx[3 * t + c] = x.vt >= 0 ? (uint32_t)x.vt : default_uv;
        }
        o->midx[t] = tri_mat[t] >= 0 ? (uint32_t)tri_mat[t] : default_mat;
    }
    for (size_t i = 0; i < o->tex.size(); ++i)
        o->images.push_back(crt_image{o->tex[i].data(), o->tex_dims[2 * i], o->tex_dims[2 * i + 1]});
    *out = o.release();
    return CRT_OK;
}

int crt_obj_model(const crt_obj *obj, crt_model_data *md)
{
    if (!obj || !md) return obj_fail(CRT_ERR_INVALID_ARG, "null argument");
    std::memset(md, 0, sizeof *md);
    md->name               = obj->name.c_str();
    md->vertices           = obj->vertices.data();
    md->num_vertices       = (uint32_t)(obj->vertices.size() / 3);
    md->texture_coords     = obj->texcoords.empty() ? nullptr : obj->texcoords.data();
    md->num_texture_coords = (uint32_t)(obj->texcoords.size() / 2);
    md->materials          = obj->mats.data();
    md->num_materials      = (uint32_t)obj->mats.size();
    md->textures           = obj->images.empty() ? nullptr : obj->images.data();
    md->num_textures       = (uint32_t)obj->images.size();
    md->vertex_indices     = obj->vidx.data();
    md->texture_indices    = obj->tidx.empty() ? nullptr : obj->tidx.data();
    md->material_indices   = obj->midx.data();
    md->num_triangles      = (uint32_t)obj->midx.size();
    return CRT_OK;
}

const char *crt_obj_warnings(const crt_obj *obj) { return obj ? obj->warnings.c_str() : ""; }

int crt_obj_free(crt_obj *obj)
{
    delete obj;
    return CRT_OK;
}

}  // extern "C"
