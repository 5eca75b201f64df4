// crt_math.h -- device math of the crt path tracer (gfx950).
//
// Vector helpers, the Philox4x32-10 counter RNG, the reference's sampling routines and
// camera, GGX pieces and the FP32 watertight ray/triangle test.  Everything here is
// __host__ __device__ so the host side of the library (camera matrices, tests of the
// layout) uses the same definitions as the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CRT_HD __host__ __device__ __forceinline__
// Inside a function body: that body's products and sums round as written, whatever the
// translation unit's contraction setting (a fused multiply-add needs both operations to
// allow it, so no expression built from these helpers is fused either).  The estimator's
// arithmetic must round like the CPU restatement it is checked against.
#define CRT_EXACT _Pragma("clang fp contract(off)")

namespace crt {

// numbers.h:10-25 (CRender): the reference's literal constants, kept as written.
constexpr float kPi     = 3.14159265359f;
constexpr float kTau    = 6.28318530717f;
constexpr float kInvPi  = 1.0f / 3.14159265359f;
constexpr float kInvTau = 1.0f / 6.28318530717f;
constexpr uint32_t kKeyHi = 0x43524E44u;       // Philox key word 1 ("CRND")
constexpr uint32_t kJitterBounce = 0xFFFFFFFFu; // counter word 2 of the camera jitter

struct f3 {
    float x, y, z;
};

CRT_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
CRT_HD f3 operator+(f3 a, f3 b) { CRT_EXACT return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
CRT_HD f3 operator-(f3 a, f3 b) { CRT_EXACT return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
CRT_HD f3 operator*(f3 a, f3 b) { CRT_EXACT return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
CRT_HD f3 operator*(f3 a, float s) { CRT_EXACT return mk3(a.x * s, a.y * s, a.z * s); }
CRT_HD f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
CRT_HD float dot(f3 a, f3 b) { CRT_EXACT return a.x * b.x + a.y * b.y + a.z * b.z; }
CRT_HD f3 cross(f3 a, f3 b)
{
    CRT_EXACT
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// FP32 transcendental functions of the estimator, evaluated in FP64 and rounded once to
// FP32.  FP64 sin, cos, tan, atan2, asin, acos and pow are accurate to about one FP64 ulp in
// any good library, so the rounded result is the correctly rounded FP32 value except when the
// exact value lies within ~1e-16 relative of an FP32 rounding boundary: the device and a CPU
// restatement built on a different library consume the same FP32 numbers.  (The reference's
// glm calls go to the host library's FP32 functions, which are within one ulp of these.)
CRT_HD float sin_cr(float x) { return (float)sin((double)x); }
CRT_HD float cos_cr(float x) { return (float)cos((double)x); }
CRT_HD float tan_cr(float x) { return (float)tan((double)x); }
CRT_HD float atan2_cr(float y, float x) { return (float)atan2((double)y, (double)x); }
CRT_HD float asin_cr(float x) { return (float)asin((double)x); }
CRT_HD float acos_cr(float x) { return (float)acos((double)x); }
CRT_HD float pow_cr(float x, float y) { return (float)pow((double)x, (double)y); }

// glm::normalize = v * inversesqrt(dot(v, v)); 1/sqrt kept as two correctly rounded ops.
CRT_HD f3 normalize(f3 a) { CRT_EXACT return a * (1.0f / sqrtf(dot(a, a))); }
// glm::reflect(I, N) = I - 2 dot(N, I) N
CRT_HD f3 reflect(f3 i, f3 n) { CRT_EXACT return i - n * (2.0f * dot(n, i)); }
CRT_HD float maxc(f3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }

// ---- Philox4x32-10 (Salmon et al., SC'11) -------------------------------------------
struct u4 {
    uint32_t x, y, z, w;
};

CRT_HD u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

CRT_HD float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

struct rnd4 {
    float a, b, c, d;
};

// Random dimensions for (pixel, sample, bounce, block), key (seed, "CRND"):
//   bounce 0xFFFFFFFF block 0: jitter x,y
//   bounce i block 0: scatter u,v ; sun cone u,v
//   bounce i block 1: Russian roulette ; lobe choice
CRT_HD rnd4 rng4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t block)
{
    const u4 r = philox4x32_10(u4{pixel, sample, bounce, block}, seed, kKeyHi);
    return rnd4{u01(r.x), u01(r.y), u01(r.z), u01(r.w)};
}

// ---- sampling (CRender src/util/sampling.h) -------------------------------------------
// sampling.h:156-166
CRT_HD f3 sphere(float u, float v)
{
    CRT_EXACT
    const float ct  = 2.0f * u - 1.0f;
    const float st  = sqrtf(1.0f - ct * ct);
    const float phi = kTau * v;
    return mk3(st * cos_cr(phi), ct, st * sin_cr(phi));
}

// sampling.h:21-33 (Duff et al. 2017 orthonormal basis)
CRT_HD void onb(f3 n, f3 &t, f3 &b)
{
    CRT_EXACT
    const float s  = (n.z < 0.0f) ? -1.0f : 1.0f;
    const float a  = -1.0f / (s + n.z);
    const float bb = n.x * n.y * a;
    t = mk3(1.0f + s * n.x * n.x * a, s * bb, -s * n.x);
    b = mk3(bb, s + n.y * n.y * a, -n.y);
}

// ---- GGX (sampling.h:93-140) ------------------------------------------------------------
CRT_HD float ggx_d(float noh, float a)
{
    CRT_EXACT
    const float a2 = a * a;
    const float d  = (noh * a2 - noh) * noh + 1.0f;
    return a2 / (d * d * kPi);
}
CRT_HD float ggx_vis(float nov, float nol, float a)
{
    CRT_EXACT
    const float a2   = a * a;
    const float ggxv = nol * sqrtf(nov * nov * (1.0f - a2) + a2);
    const float ggxl = nov * sqrtf(nol * nol * (1.0f - a2) + a2);
    return 0.5f / (ggxv + ggxl);
}
CRT_HD f3 schlick(float u, f3 f0)
{
    CRT_EXACT
    const float f = pow_cr(1.0f - u, 5.0f);
    return mk3(f0.x + (1.0f - f0.x) * f, f0.y + (1.0f - f0.y) * f, f0.z + (1.0f - f0.z) * f);
}

// ---- camera (CRender camera.cpp:14-39) --------------------------------------------------
struct camera_dev {
    float pos[3];
    float fov;
    float m[16];   // column-major _cached_matrix
    int32_t mode;
    float scale;
};

CRT_HD void camera_ray(const camera_dev &c, float x, float y, float aspect, f3 &o, f3 &d)
{
    CRT_EXACT
    f3 dir;
    if (c.mode == 0) {
        const float u = (2.0f * x - 1.0f) * aspect;
        const float v = 2.0f * y - 1.0f;
        const float w = 1.0f / tan_cr(0.5f * (c.fov * 0.01745329251994329576923690768489f));
        dir = mk3(c.m[0] * u + c.m[4] * v + c.m[8] * w, c.m[1] * u + c.m[5] * v + c.m[9] * w,
                  c.m[2] * u + c.m[6] * v + c.m[10] * w);
        o = mk3(c.pos[0], c.pos[1], c.pos[2]);
    } else {
        const float u = 2.0f * x - 1.0f;
        const float v = 2.0f * y - 1.0f;
        const float su = c.scale * u, sv = c.scale * v;
        o   = mk3(c.m[0] * su + c.m[4] * sv + c.m[12], c.m[1] * su + c.m[5] * sv + c.m[13],
                  c.m[2] * su + c.m[6] * sv + c.m[14]);
        dir = mk3(c.m[8], c.m[9], c.m[10]);
    }
    d = normalize(dir);
}

// ---- watertight ray/triangle test -------------------------------------------------------
// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1) 2013, including
// the FP64 recomputation of zero edge functions.  FP contraction is off so that t, u, v
// round exactly as the operations are written.
struct wt_ray {
    int   kx, ky, kz;
    float sx, sy, sz;
};

CRT_HD wt_ray wt_setup(f3 d)
{
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const float dkz = kz == 0 ? d.x : (kz == 1 ? d.y : d.z);
    if (dkz < 0.0f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    const float dkx = kx == 0 ? d.x : (kx == 1 ? d.y : d.z);
    const float dky = ky == 0 ? d.x : (ky == 1 ? d.y : d.z);
    wt_ray r;
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.sx = dkx / dkz;
    r.sy = dky / dkz;
    r.sz = 1.0f / dkz;
    return r;
}

CRT_HD float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Returns true for a hit with tnear < t <= tfar.
CRT_HD bool wt_test(const wt_ray &r, f3 v0, f3 v1, f3 v2, f3 o, float tnear, float tfar, float &t,
                    float &u, float &v)
{
#pragma clang fp contract(off)
    const f3 A = v0 - o, B = v1 - o, C = v2 - o;
    const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    const float Ax = comp(A, r.kx) - r.sx * Akz;
    const float Ay = comp(A, r.ky) - r.sy * Akz;
    const float Bx = comp(B, r.kx) - r.sx * Bkz;
    const float By = comp(B, r.ky) - r.sy * Bkz;
    const float Cx = comp(C, r.kx) - r.sx * Ckz;
    const float Cy = comp(C, r.ky) - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T   = U * (r.sz * Akz) + V * (r.sz * Bkz) + W * (r.sz * Ckz);
    const float rcp = 1.0f / det;
    const float tt  = T * rcp;
    if (!(tt > tnear) || !(tt <= tfar)) return false;
    t = tt;
    u = V * rcp;
    v = W * rcp;
    return true;
}

// Ordered-int encoding of floats so atomicMin/atomicMax on int order like the floats.
CRT_HD int f2ord(float f)
{
    const int i = __builtin_bit_cast(int, f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
CRT_HD float ord2f(int i) { return __builtin_bit_cast(float, i >= 0 ? i : i ^ 0x7FFFFFFF); }

}  // namespace crt
