/*
 * rng_sampling.c -- oracle RNG, sampling and camera (TEST INFRASTRUCTURE ONLY).
 *
 * Philox4x32-10 follows the published algorithm (Salmon, Moraes, Dror, Shaw, "Parallel
 * random numbers: as easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with
 * the Weyl key schedule.  The sampling routines restate src/util/sampling.h and the
 * camera restates src/render/camera.cpp of CRender.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include "oracle.h"

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

/* numbers.h:10-25: the reference's literal constants. */
#define REF_PI_F  3.14159265359f
#define REF_TAU_F 6.28318530717f

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

void orc_philox_batch(const uint32_t *ctr, uint32_t n, const uint32_t key[2], uint32_t *out)
{
    for (uint32_t i = 0; i < n; ++i) orc_philox4x32_10(ctr + 4 * (size_t)i, key, out + 4 * (size_t)i);
}

float orc_u01(uint32_t x)
{
    return (float)(x >> 8) * (1.0f / 16777216.0f);
}

/* sampling.h:156-166 */
void orc_sphere(float u, float v, float out[3])
{
    const float cos_theta = 2.0f * u - 1.0f;
    const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    const float phi       = REF_TAU_F * v;
    out[0] = sin_theta * cosf(phi);
    out[1] = cos_theta;
    out[2] = sin_theta * sinf(phi);
}

/* sampling.h:168-172: n + uniform sphere point (normalised by the caller, renderer.cpp:97) */
void orc_hemp_cos(const float n[3], float u, float v, float out[3])
{
    float p[3];
    orc_sphere(u, v, p);
    out[0] = n[0] + p[0];
    out[1] = n[1] + p[1];
    out[2] = n[2] + p[2];
}

/* sampling.h:21-33 */
void orc_build_local(const float n[3], float tangent[3], float normal[3], float bitangent[3])
{
    const float s = (n[2] < 0.0f) ? -1.0f : 1.0f;
    const float a = -1.0f / (s + n[2]);
    const float b = n[0] * n[1] * a;
    normal[0] = n[0];
    normal[1] = n[1];
    normal[2] = n[2];
    tangent[0] = 1.0f + s * n[0] * n[0] * a;
    tangent[1] = s * b;
    tangent[2] = -s * n[0];
    bitangent[0] = b;
    bitangent[1] = s + n[1] * n[1] * a;
    bitangent[2] = -n[1];
}

/* sampling.h:35-42 */
void orc_map_to_solid_angle(float u, float v, float theta_max, float out[3])
{
    const float phi       = REF_TAU_F * u;
    const float cos_theta = 1.0f - v * (1.0f - cosf(theta_max));
    const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    out[0] = cosf(phi) * sin_theta;
    out[1] = cos_theta;
    out[2] = sinf(phi) * sin_theta;
}

/* sampling.h:44-47 */
float orc_solid_angle_pdf(float theta_max)
{
    return 1.0f / (REF_TAU_F * (1.0f - cosf(theta_max)));
}

/* camera.cpp:14-39.  glm::radians multiplies by pi/180; glm::normalize is v*inversesqrt. */
void orc_camera_ray(const orc_camera *cam, float x, float y, float aspect, float o[3], float d[3])
{
    const float *m = cam->matrix;
    float dir[3];
    if (cam->mode == 0) {
        const float u = (2.0f * x - 1.0f) * aspect;
        const float v = 2.0f * y - 1.0f;
        const float w = 1.0f / tanf(0.5f * (cam->fov * 0.01745329251994329576923690768489f));
        for (int k = 0; k < 3; ++k) dir[k] = m[0 + k] * u + m[4 + k] * v + m[8 + k] * w;
        o[0] = cam->pos[0];
        o[1] = cam->pos[1];
        o[2] = cam->pos[2];
    } else {
        const float u = 2.0f * x - 1.0f;
        const float v = 2.0f * y - 1.0f;
        for (int k = 0; k < 3; ++k)
            o[k] = m[0 + k] * (cam->scale * u) + m[4 + k] * (cam->scale * v) + m[12 + k];
        for (int k = 0; k < 3; ++k) dir[k] = m[8 + k];
    }
    const float inv = 1.0f / sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    d[0] = dir[0] * inv;
    d[1] = dir[1] * inv;
    d[2] = dir[2] * inv;
}

/* renderer.cpp:371-383 with n = _current_sample + 1 */
float orc_display(float sum, uint64_t n)
{
    float m = sum / (float)n;
    if (m < 0.0f) m = 0.0f;
    if (m > 1.0f) m = 1.0f;
    return powf(m, 1.0f / 2.2f);
}
