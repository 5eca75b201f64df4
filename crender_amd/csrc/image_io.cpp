// image_io.cpp -- image export of the crt renderer's buffers.
//
// Mirrors asset_loader::export_framebuffer (CRender src/util/asset_loader.cpp:89-178,
// 348-377), which writes the RGBA FP32 display buffer with stb_image_write / tinyexr:
//   PNG  RGBA8, min(v * 255, 255) truncated (export_png)
//   EXR  three half-float channels stored B, G, R, uncompressed scanlines (export_exr)
//   HDR  Radiance RGBE of pow(v, 2.2) (export_hdr)
// plus PFM (FP32 RGB, for comparing HDR running sums bit for bit), which the reference
// does not have.  Rows are written in memory order (row 0 first), as the reference's
// writers do.  The reference picks a fresh name under ./out/; here the caller gives the
// path.  JPEG is not written (CRT_ERR_UNSUPPORTED).
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/crt.h"

namespace crt {
int set_error(int code, const std::string &msg);
}

namespace {

int io_fail(int code, const std::string &m) { return crt::set_error(code, m); }

struct File {
    FILE *f = nullptr;
    explicit File(const char *path) : f(std::fopen(path, "wb")) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
    bool put(const void *p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
};

void be32(std::vector<uint8_t> &v, uint32_t x)
{
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}

int write_png(const char *path, const float *rgba, uint32_t w, uint32_t h)
{
    std::vector<uint8_t> raw((size_t)(w * 4 + 1) * h);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t *row = &raw[(size_t)y * (w * 4 + 1)];
        row[0] = 0;   // filter: none
        for (size_t i = 0; i < (size_t)w * 4; ++i) {
            const float v = std::fmin(rgba[(size_t)y * w * 4 + i] * 255.f, 255.f);
            row[1 + i] = (uint8_t)(v > 0.0f ? v : 0.0f);   // float -> uint8 conversion truncates
        }
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) return io_fail(CRT_ERR_HIP, "zlib");
    z.resize(zn);
    std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
    auto chunk = [&](const char *type, const std::vector<uint8_t> &d) {
        be32(out, (uint32_t)d.size());
        const size_t t0 = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), d.begin(), d.end());
        be32(out, (uint32_t)crc32(0, &out[t0], (uInt)(out.size() - t0)));
    };
    std::vector<uint8_t> ihdr;
    be32(ihdr, w);
    be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});   // 8-bit RGBA, deflate, no interlace
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
    File f(path);
    if (!f.f || !f.put(out.data(), out.size())) return io_fail(CRT_ERR_INVALID_ARG, std::string("cannot write ") + path);
    return CRT_OK;
}

// IEEE half from float, round to nearest even (overflow -> inf, NaN kept)
uint16_t to_half(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t ex   = (u >> 23) & 0xFFu;
    uint32_t       man  = u & 0x7FFFFFu;
    if (ex == 0xFF) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u : 0u));
    int e = (int)ex - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        const int      shift = 14 - e;
        const uint32_t half  = man >> shift;
        const uint32_t rem   = man & ((1u << shift) - 1u);
        const uint32_t mid   = 1u << (shift - 1);
        uint32_t       r     = half + ((rem > mid || (rem == mid && (half & 1u))) ? 1u : 0u);
        return (uint16_t)(sign | r);
    }
    uint32_t h   = sign | ((uint32_t)e << 10) | (man >> 13);
    const uint32_t rem = man & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;   // may carry into the exponent: correct
    return (uint16_t)h;
}

int write_exr(const char *path, const float *rgba, uint32_t w, uint32_t h)
{
    std::vector<uint8_t> hd;
    auto str = [&](const char *s) { hd.insert(hd.end(), s, s + std::strlen(s) + 1); };
    auto i32 = [&](int32_t x) { for (int k = 0; k < 4; ++k) hd.push_back((uint8_t)((uint32_t)x >> (8 * k))); };
    auto f32 = [&](float x) {
        uint32_t u;
        std::memcpy(&u, &x, 4);
        i32((int32_t)u);
    };
    auto attr = [&](const char *name, const char *type, int32_t size) {
        str(name);
        str(type);
        i32(size);
    };
    i32(20000630);   // magic
    i32(2);          // version 2, scanline, single part
    // channels B, G, R (sorted), HALF
    attr("channels", "chlist", 3 * (2 + 16) + 1);
    for (const char *c : {"B", "G", "R"}) {
        str(c);
        i32(1);          // HALF
        hd.insert(hd.end(), {0, 0, 0, 0});   // pLinear + reserved
        i32(1);
        i32(1);          // sampling
    }
    hd.push_back(0);
    attr("compression", "compression", 1);
    hd.push_back(0);     // NO_COMPRESSION
    attr("dataWindow", "box2i", 16);
    i32(0), i32(0), i32((int32_t)w - 1), i32((int32_t)h - 1);
    attr("displayWindow", "box2i", 16);
    i32(0), i32(0), i32((int32_t)w - 1), i32((int32_t)h - 1);
    attr("lineOrder", "lineOrder", 1);
    hd.push_back(0);     // INCREASING_Y
    attr("pixelAspectRatio", "float", 4);
    f32(1.0f);
    attr("screenWindowCenter", "v2f", 8);
    f32(0.0f), f32(0.0f);
    attr("screenWindowWidth", "float", 4);
    f32(1.0f);
    hd.push_back(0);     // end of header
    const uint64_t line_bytes = (uint64_t)w * 3 * 2;
    const uint64_t block      = 8 + line_bytes;
    const uint64_t base       = hd.size() + 8ull * h;
    for (uint32_t y = 0; y < h; ++y) {
        const uint64_t off = base + (uint64_t)y * block;
        for (int k = 0; k < 8; ++k) hd.push_back((uint8_t)(off >> (8 * k)));
    }
    File f(path);
    if (!f.f || !f.put(hd.data(), hd.size())) return io_fail(CRT_ERR_INVALID_ARG, std::string("cannot write ") + path);
    std::vector<uint8_t> line(block);
    for (uint32_t y = 0; y < h; ++y) {
        const int32_t yy = (int32_t)y, sz = (int32_t)line_bytes;
        std::memcpy(&line[0], &yy, 4);
        std::memcpy(&line[4], &sz, 4);
        size_t p = 8;
        for (int c : {2, 1, 0}) {   // B, G, R planes of this scanline
            for (uint32_t x = 0; x < w; ++x) {
                const uint16_t hv = to_half(rgba[((size_t)y * w + x) * 4 + c]);
                line[p++] = (uint8_t)hv;
                line[p++] = (uint8_t)(hv >> 8);
            }
        }
        if (!f.put(line.data(), line.size())) return io_fail(CRT_ERR_INVALID_ARG, std::string("cannot write ") + path);
    }
    return CRT_OK;
}

int write_hdr(const char *path, const float *rgba, uint32_t w, uint32_t h)
{
    File f(path);
    if (!f.f) return io_fail(CRT_ERR_INVALID_ARG, std::string("cannot write ") + path);
    const std::string head = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(h) + " +X " +
                             std::to_string(w) + "\n";
    if (!f.put(head.data(), head.size())) return io_fail(CRT_ERR_INVALID_ARG, "write failed");
    std::vector<uint8_t> row((size_t)w * 4);
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            float c[3];
            for (int k = 0; k < 3; ++k) c[k] = std::pow(rgba[((size_t)y * w + x) * 4 + k], 2.2f);
            const float m = std::fmax(c[0], std::fmax(c[1], c[2]));
            uint8_t *q = &row[(size_t)x * 4];
            if (!(m >= 1e-32f)) {
                q[0] = q[1] = q[2] = q[3] = 0;
            } else {
                int e = 0;
                const float scale = std::frexp(m, &e) * 256.0f / m;
                for (int k = 0; k < 3; ++k) q[k] = (uint8_t)(c[k] * scale);
                q[3] = (uint8_t)(e + 128);
            }
        }
        if (!f.put(row.data(), row.size())) return io_fail(CRT_ERR_INVALID_ARG, "write failed");
    }
    return CRT_OK;
}

int write_pfm(const char *path, const float *rgba, uint32_t w, uint32_t h)
{
    File f(path);
    if (!f.f) return io_fail(CRT_ERR_INVALID_ARG, std::string("cannot write ") + path);
    const std::string head = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
    if (!f.put(head.data(), head.size())) return io_fail(CRT_ERR_INVALID_ARG, "write failed");
    std::vector<float> row((size_t)w * 3);
    for (uint32_t r = 0; r < h; ++r) {
        const uint32_t y = h - 1 - r;   // PFM stores the bottom row first
        for (uint32_t x = 0; x < w; ++x)
            for (int k = 0; k < 3; ++k) row[(size_t)x * 3 + k] = rgba[((size_t)y * w + x) * 4 + k];
        if (!f.put(row.data(), row.size() * 4)) return io_fail(CRT_ERR_INVALID_ARG, "write failed");
    }
    return CRT_OK;
}

}  // namespace

extern "C" int crt_image_write(const char *path, int format, const float *rgba, uint32_t width, uint32_t height)
{
    if (!path || !rgba || width == 0 || height == 0) return io_fail(CRT_ERR_INVALID_ARG, "null or empty image");
    switch (format) {
    case CRT_IMAGE_PNG: return write_png(path, rgba, width, height);
    case CRT_IMAGE_EXR: return write_exr(path, rgba, width, height);
    case CRT_IMAGE_HDR: return write_hdr(path, rgba, width, height);
    case CRT_IMAGE_PFM: return write_pfm(path, rgba, width, height);
    case CRT_IMAGE_JPG: return io_fail(CRT_ERR_UNSUPPORTED, "JPEG export is not implemented");
    default: return io_fail(CRT_ERR_INVALID_ARG, "unknown image format");
    }
}
