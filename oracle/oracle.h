/*
 * oracle.h -- CPU oracle for the crt path tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU and in plain C, the
 * algorithm of CRender's render hot path so that the HIP implementation in
 * crender_amd/ can be checked against it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links or calls it.
 *
 * What each part restates (paths are relative to the CRender source tree):
 *   orc_philox4x32_10     counter-based RNG replacing renderer.cpp:6-11 (mt19937); the
 *                         published Philox4x32-10 (Salmon et al., SC'11).
 *   orc_camera_ray        camera.cpp:14-39 (perspective + orthographic get_ray).
 *   orc_sphere/hemp_cos   sampling.h:156-172.
 *   orc_build_local       sampling.h:21-33 (Duff et al. 2017 ONB).
 *   orc_sun_sample        sampling.h:35-47,53-57,72-80 and registry.cpp:142-146.
 *   orc_closest_*         model.cpp:5-49 (rtcIntersect1 + rtcInterpolate0 semantics of
 *                         Embree 3: closest hit in (tnear, tfar], Ng=(v1-v0)x(v2-v0),
 *                         P=(1-u-v)v0+u*v1+v*v2) and model.cpp:99-126 (instance path).
 *   orc_bvh_*             the role of rtcCommitScene (model.cpp:94): an FP32 binned-SAH
 *                         BVH2 built on the host, used for large scenes and as the CPU
 *                         baseline's traversal.
 *   orc_render_f32/f64    renderer.cpp:21-102 (process_hit) and 258-384 (_sample_pixel),
 *                         scheduled as renderer.cpp:240-256 (one task per row, one
 *                         barrier per sample pass).
 *
 * Random-number layout shared with the HIP path (a fixed dimension layout per
 * (pixel, sample, bounce)): counter = {pixel, sample, bounce, block},
 * key = {seed, 0x43524E44}:
 *   bounce = 0xFFFFFFFF, block 0 : u0,u1 camera jitter
 *   bounce = i,          block 0 : u0,u1 scatter (sphere / GGX / lobe) ; u2,u3 sun cone
 *   bounce = i,          block 1 : u0 Russian roulette ; u1 lobe choice (pbr Fresnel)
 * u32 -> float in [0,1) is (x >> 8) * 2^-24.
 */
#ifndef CRT_ORACLE_H
#define CRT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_KEY_HI 0x43524E44u
#define ORC_JITTER_BOUNCE 0xFFFFFFFFu

/* ---- RNG ------------------------------------------------------------------------ */
void  orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_u01(uint32_t x);
/* Batch helper: out[4*i..] = philox(ctr_i, key) for n counters (ctr is 4*n). */
void  orc_philox_batch(const uint32_t *ctr, uint32_t n, const uint32_t key[2], uint32_t *out);

/* ---- sampling (float, reference semantics) -------------------------------------- */
void orc_sphere(float u, float v, float out[3]);
void orc_hemp_cos(const float n[3], float u, float v, float out[3]);
void orc_build_local(const float n[3], float tangent[3], float normal[3], float bitangent[3]);
void orc_map_to_solid_angle(float u, float v, float theta_max, float out[3]);
float orc_solid_angle_pdf(float theta_max);

/* ---- scene description ------------------------------------------------------------ */
enum { ORC_METAL = 0, ORC_SMOOTH = 1, ORC_GLASS = 2 };   /* material.h:14-19 */
enum { ORC_COMPAT = 0, ORC_PBR = 1 };
/* BVH64: the FP32 BVH walked with FP64 closest-hit decisions (orc_bvh_closest64) */
enum { ORC_TRAV_BRUTE64 = 0, ORC_TRAV_BRUTE32 = 1, ORC_TRAV_BVH32 = 2, ORC_TRAV_BVH64 = 3 };

typedef struct orc_material {
    int32_t type;          /* ORC_METAL / ORC_SMOOTH / ORC_GLASS */
    float   ior, roughness, reflectiveness, emission;
    float   colour[4];
    int32_t tex;           /* -1: none; else index into orc_scene.textures */
} orc_material;

typedef struct orc_image {
    const float *rgba;     /* w*h*4, row-major, (x + y*w)*4 */
    uint32_t     w, h;
} orc_image;

typedef struct orc_sun {
    float size, intensity;
    float dir[3];
    float colour[3];
} orc_sun;

typedef struct orc_camera {
    float   pos[3];
    float   fov;           /* degrees */
    float   matrix[16];    /* column-major 4x4, as glm::mat4 (camera.cpp:194-208) */
    int32_t mode;          /* 0 perspective, 1 orthographic */
    float   scale;         /* orthographic half-extent */
} orc_camera;

typedef struct orc_bvh orc_bvh;

typedef struct orc_scene {
    const float        *tri;      /* 9*ntri: expanded corners v0,v1,v2 per triangle */
    const float        *uv;       /* 6*ntri per-corner uv, or NULL (uv = 0) */
    const uint32_t     *mat_idx;  /* ntri */
    uint32_t            ntri;
    const orc_material *mats;
    uint32_t            nmat;
    const orc_image    *textures;
    uint32_t            ntex;
    orc_image           sky;      /* rgba NULL: no skybox (black) */
    float               sky_rot[2];
    orc_sun             sun;
    int32_t             sun_enabled;
    orc_camera          cam;
    const orc_bvh      *bvh;      /* required for ORC_TRAV_BVH32 / ORC_TRAV_BVH64 */
    int32_t             trav;     /* ORC_TRAV_* */
} orc_scene;

typedef struct orc_settings {
    uint32_t res_x, res_y;
    uint32_t max_bounces;
    uint32_t spp;            /* passes rendered in this call */
    uint32_t sample_offset;  /* absolute index of the first pass */
    uint32_t rr_start;       /* 0: no Russian roulette; else RR before bounce >= rr_start */
    uint32_t mode;           /* ORC_COMPAT / ORC_PBR */
    uint32_t seed;
    float    aspect;
    uint32_t row_begin, row_end;  /* rows rendered (row_end 0 = all) */
    int32_t  nthreads;
} orc_settings;

typedef struct orc_stats {
    uint64_t ext_rays;       /* extension rays traced */
    uint64_t shadow_rays;    /* sun shadow rays traced */
    uint64_t ref_segments;   /* reference-style count: sum of total_bounces */
    double   seconds;
} orc_stats;

/* ---- closest hit ------------------------------------------------------------------- */
/* Returns primID or -1.  t,u,v: Embree conventions.  tnear exclusive, tfar inclusive,
 * equal t resolved to the smaller primID. */
int32_t orc_closest_brute64(const float *tri, uint32_t ntri, const float o[3], const float d[3],
                            float tnear, float tfar, double *t, double *u, double *v);
int32_t orc_brute64d(const float *tri, uint32_t ntri, const double o[3], const double d[3],
                     double tnear, double tfar, double *t, double *u, double *v);
int32_t orc_closest_brute32(const float *tri, uint32_t ntri, const float o[3], const float d[3],
                            float tnear, float tfar, float *t, float *u, float *v);
/* FP32 watertight test of one triangle (Woop, Benthin, Wald, JCGT 2013). 1 = hit. */
int     orc_tri_watertight32(const float *v0, const float *v1, const float *v2,
                             const float o[3], const float d[3], float tnear, float tfar,
                             float *t, float *u, float *v);

/* Batched queries over SoA rays; nthreads <= 0 means all hardware threads. */
void orc_intersect_batch(const float *tri, uint32_t ntri, const orc_bvh *bvh, int32_t trav,
                         const float *ox, const float *oy, const float *oz,
                         const float *dx, const float *dy, const float *dz,
                         const float *tnear, const float *tfar, uint32_t n,
                         int32_t *prim, double *t, double *u, double *v, int nthreads);

/* ---- CPU BVH ------------------------------------------------------------------------ */
orc_bvh *orc_bvh_build(const float *tri, uint32_t ntri, uint32_t max_leaf);
void     orc_bvh_free(orc_bvh *bvh);
/* Structure checks: returns 0 if every child box contains its subtree's triangles and
 * every triangle appears exactly once, else a nonzero code. */
int      orc_bvh_validate(const orc_bvh *bvh, const float *tri, uint32_t ntri);
uint32_t orc_bvh_node_count(const orc_bvh *bvh);
double   orc_bvh_sah_cost(const orc_bvh *bvh);
int32_t  orc_bvh_closest32(const orc_bvh *bvh, const float o[3], const float d[3], float tnear,
                           float tfar, float *t, float *u, float *v);
int32_t  orc_bvh_closest64(const orc_bvh *bvh, const float o[3], const float d[3], float tnear,
                           float tfar, double *t, double *u, double *v);
int      orc_default_threads(void);

/* ---- integrator ---------------------------------------------------------------------- */
/* raw: W*H*3 running sums (accumulated into, caller zeroes), flipped in x and y as
 * renderer.cpp:358-365.  albedo/normal/depth: W*H*4 AOVs of the last pass or NULL. */
int orc_render_f32(const orc_scene *scene, const orc_settings *s, float *raw,
                   float *albedo, float *normal, float *depth, orc_stats *stats);
int orc_render_f64(const orc_scene *scene, const orc_settings *s, double *raw,
                   float *albedo, float *normal, float *depth, orc_stats *stats);

/* scene::cast_ray (scene.cpp:79-98) for n rays, FP32: distance, primID, material index,
 * interpolated uv (n x 2), normalize(Ng) and the point (n x 3); -1 / +inf on a miss. */
void orc_cast_records_f32(const orc_scene *scene, uint32_t n, const float *org, const float *dir, float *dist,
                          int32_t *prim, int32_t *mat, float *uv, float *nrm, float *pt);

/* One camera ray (camera.cpp:14-39) for jittered pixel coords x,y in [0,1]. */
void orc_camera_ray(const orc_camera *cam, float x, float y, float aspect, float o[3], float d[3]);

/* Nearest texel with wrap-around, image.h:104-109: column static_cast<uint64_t>(u * w) % w,
 * negative products taken as x86-64 code without AVX-512 converts them (orc_texel_index). */
uint64_t orc_texel_index(float u, uint32_t size);
void     orc_get_uv(const orc_image *img, float u, float v, float out[4]);

/* One material scatter (process_hit, renderer.cpp:21-102, in the given mode) of a hit
 * record on primitive `prim` of the scene, for the incoming direction in_d and the
 * uniforms u (block 0) and u2 (block 1) of the bounce.  flags: 1 alpha pass-through,
 * 2 delta lobe (pbr: no NEE), 4 absorbed (pbr GGX sample under the surface). */
typedef struct orc_scatter_out {
    float   o[3], d[3];      /* next ray */
    float   albedo[3];       /* throughput multiplier */
    float   colour[4];       /* material / texture colour */
    float   nee_n[3];        /* normal of the NEE offset and cosine */
    float   emission;
    int32_t flags;
} orc_scatter_out;
void orc_scatter_f32(const orc_scene *sc, uint32_t mode, int32_t prim, const float point[3],
                     const float normal[3], float uvx, float uvy, const float in_d[3],
                     const float u[4], const float u2[4], orc_scatter_out *out);
void orc_scatter_f64(const orc_scene *sc, uint32_t mode, int32_t prim, const float point[3],
                     const float normal[3], float uvx, float uvy, const float in_d[3],
                     const float u[4], const float u2[4], orc_scatter_out *out);
/* Radiance of a missed ray (renderer.cpp:277-289 -> scene.cpp:67-77), FP32 evaluation. */
void orc_sky_f32(const orc_scene *sc, const float d[3], float out[3]);
/* Four uniforms in [0,1) for counter {pixel, sample, bounce, block}, key {seed, ORC_KEY_HI}. */
void orc_rng4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t block, float u[4]);

/* GGX in FP64: D (sampling.h:93-99), height-correlated visibility (sampling.h:109-117) and
 * visible-normal sampling of the reflected direction in the local frame (normal = +z);
 * weight = G2/G1, the f*cos/pdf of the sample without Fresnel.  Returns 0 below the
 * surface. */
double orc_ggx_d(double noh, double alpha);
double orc_ggx_vis(double nov, double nol, double alpha);
int    orc_ggx_sample(const double wo_local[3], double alpha, double u1, double u2, double wi_local[3], double *weight);

/* Display value of one channel: pow(clamp(sum/n, 0, 1), 1/2.2) (renderer.cpp:371-383). */
float orc_display(float sum, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
