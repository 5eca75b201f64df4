/*
 * integrator.c -- oracle render driver (TEST INFRASTRUCTURE ONLY).
 *
 * Instantiates integrator_impl.h in FP32 and FP64 and schedules it as CRender does:
 * one task per image row, all rows of one sample pass finished before the next pass
 * starts (src/render/renderer.cpp:116-144,240-256; src/objects/thread_pool.cpp:40-60).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "oracle.h"


void orc_rng4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t block, float u[4])
{
    const uint32_t c[4] = {pixel, sample, bounce, block};
    const uint32_t k[2] = {seed, ORC_KEY_HI};
    uint32_t       r[4];
    orc_philox4x32_10(c, k, r);
    for (int i = 0; i < 4; ++i) u[i] = orc_u01(r[i]);
}

/* image.h:104-109: static_cast<uint64_t>(u * width) % width.  The cast truncates toward
 * zero.  For a negative product it is undefined in C++; x86-64 code without AVX-512
 * (cvttss2si on the signed range) yields the signed truncation reinterpreted as uint64,
 * which is what is restated here: (-1, 0] -> 0, -1.5 -> (2^64 - 1) % width. */
uint64_t orc_texel_index(float u, uint32_t size)
{
    return (uint64_t)(int64_t)(u * (float)size) % (uint64_t)size;
}

void orc_get_uv(const orc_image *img, float u, float v, float out[4])
{
    const uint64_t x = orc_texel_index(u, img->w), y = orc_texel_index(v, img->h);
    const float   *p = img->rgba + (size_t)(x + y * (uint64_t)img->w) * 4;
    out[0] = p[0];
    out[1] = p[1];
    out[2] = p[2];
    out[3] = p[3];
}

/* FP32 instantiation.  Transcendental functions are the FP64 ones rounded once to FP32 (the
 * correctly rounded FP32 value but for rare double-rounding cases), so the result does not
 * depend on which library's FP32 functions are linked; glibc's sinf and friends agree with
 * these to within one ulp. */
static inline float sin_cr(float x) { return (float)sin((double)x); }
static inline float cos_cr(float x) { return (float)cos((double)x); }
static inline float tan_cr(float x) { return (float)tan((double)x); }
static inline float atan2_cr(float y, float x) { return (float)atan2((double)y, (double)x); }
static inline float asin_cr(float x) { return (float)asin((double)x); }
static inline float acos_cr(float x) { return (float)acos((double)x); }
static inline float pow_cr(float x, float y) { return (float)pow((double)x, (double)y); }

#define R float
#define SFX f32
#define SQRT_ sqrtf
#define SIN_ sin_cr
#define COS_ cos_cr
#define TAN_ tan_cr
#define ATAN2_ atan2_cr
#define ASIN_ asin_cr
#define ACOS_ acos_cr
#define POW_ pow_cr
#define FABS_ fabsf
#include "integrator_impl.h"
#undef R
#undef SFX
#undef SQRT_
#undef SIN_
#undef COS_
#undef TAN_
#undef ATAN2_
#undef ASIN_
#undef ACOS_
#undef POW_
#undef FABS_
#undef PI_R
#undef TAU_R

#define R double
#define SFX f64
#define SQRT_ sqrt
#define SIN_ sin
#define COS_ cos
#define TAN_ tan
#define ATAN2_ atan2
#define ASIN_ asin
#define ACOS_ acos
#define POW_ pow
#define FABS_ fabs
#include "integrator_impl.h"

typedef struct {
    const orc_scene    *sc;
    const orc_settings *st;
    void               *raw;
    int                 is64;
    float              *alb, *nrm, *dep;
    uint32_t            sample;
    uint32_t            row_end;
    volatile uint32_t   next_row;
    uint64_t            ext, shadow, segs;
    pthread_mutex_t     mu;
} pass_job;

static void *pass_worker(void *arg)
{
    pass_job *job = (pass_job *)arg;
    cnt_f32   c32 = {0, 0, 0};
    cnt_f64   c64 = {0, 0, 0};
    for (;;) {
        const uint32_t y = __atomic_fetch_add(&job->next_row, 1u, __ATOMIC_RELAXED);
        if (y >= job->row_end) break;
        for (uint32_t x = 0; x < job->st->res_x; ++x) {
            if (job->is64)
                sample_pixel_f64(job->sc, job->st, x, y, job->sample, (double *)job->raw, job->alb, job->nrm, job->dep, &c64);
            else
                sample_pixel_f32(job->sc, job->st, x, y, job->sample, (float *)job->raw, job->alb, job->nrm, job->dep, &c32);
        }
    }
    pthread_mutex_lock(&job->mu);
    job->ext += c32.ext + c64.ext;
    job->shadow += c32.shadow + c64.shadow;
    job->segs += c32.segs + c64.segs;
    pthread_mutex_unlock(&job->mu);
    return NULL;
}

static int render_common(const orc_scene *sc, const orc_settings *st, void *raw, int is64, float *alb,
                         float *nrm, float *dep, orc_stats *stats)
{
    if (!sc || !st || !raw || st->res_x == 0 || st->res_y == 0) return -1;
    if (sc->ntri && (!sc->tri || !sc->mat_idx || !sc->mats)) return -2;
    for (uint32_t i = 0; i < sc->ntri; ++i)
        if (sc->mat_idx[i] >= sc->nmat) return -3;
    if ((sc->trav == ORC_TRAV_BVH32 || sc->trav == ORC_TRAV_BVH64) && !sc->bvh) return -4;
    int nthreads = st->nthreads > 0 ? st->nthreads : orc_default_threads();
    if (nthreads > 256) nthreads = 256;
    const uint32_t  row_begin = st->row_begin;
    const uint32_t  row_end   = st->row_end ? (st->row_end < st->res_y ? st->row_end : st->res_y) : st->res_y;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    uint64_t ext = 0, shadow = 0, segs = 0;
    for (uint32_t s = 0; s < st->spp; ++s) {
        pass_job job;
        memset(&job, 0, sizeof job);
        job.sc       = sc;
        job.st       = st;
        job.raw      = raw;
        job.is64     = is64;
        job.alb      = alb;
        job.nrm      = nrm;
        job.dep      = dep;
        job.sample   = st->sample_offset + s;
        job.row_end  = row_end;
        job.next_row = row_begin;
        pthread_mutex_init(&job.mu, NULL);
        pthread_t th[256];
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, pass_worker, &job);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL); /* barrier per pass */
        pthread_mutex_destroy(&job.mu);
        ext += job.ext;
        shadow += job.shadow;
        segs += job.segs;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (stats) {
        stats->ext_rays     = ext;
        stats->shadow_rays  = shadow;
        stats->ref_segments = segs;
        stats->seconds      = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    }
    return 0;
}

int orc_render_f32(const orc_scene *scene, const orc_settings *s, float *raw, float *albedo,
                   float *normal, float *depth, orc_stats *stats)
{
    return render_common(scene, s, raw, 0, albedo, normal, depth, stats);
}

int orc_render_f64(const orc_scene *scene, const orc_settings *s, double *raw, float *albedo,
                   float *normal, float *depth, orc_stats *stats)
{
    return render_common(scene, s, raw, 1, albedo, normal, depth, stats);
}

/* scene::cast_ray (scene.cpp:79-98) for a batch of rays in FP32, through the integrator's own
 * cast (model.cpp:27-47): org / dir / uv / nrm / pt interleaved xyz (uv: xy).  A miss gives
 * distance +inf, prim and material -1 and zeros elsewhere. */
void orc_cast_records_f32(const orc_scene *scene, uint32_t n, const float *org, const float *dir, float *dist,
                          int32_t *prim, int32_t *mat, float *uv, float *nrm, float *pt)
{
    for (uint32_t i = 0; i < n; ++i) {
        hit_f32       h;
        const vec_f32 o = mk_f32(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
        const vec_f32 d = mk_f32(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        if (!cast_f32(scene, o, d, &h)) {
            dist[i] = INFINITY;
            prim[i] = -1;
            mat[i]  = -1;
            uv[2 * i] = uv[2 * i + 1] = 0.0f;
            for (int k = 0; k < 3; ++k) nrm[3 * i + k] = pt[3 * i + k] = 0.0f;
            continue;
        }
        dist[i]       = h.t;
        prim[i]       = h.prim;
        mat[i]        = (int32_t)scene->mat_idx[h.prim];
        uv[2 * i]     = h.uvx;
        uv[2 * i + 1] = h.uvy;
        nrm[3 * i] = h.normal.x, nrm[3 * i + 1] = h.normal.y, nrm[3 * i + 2] = h.normal.z;
        pt[3 * i] = h.point.x, pt[3 * i + 1] = h.point.y, pt[3 * i + 2] = h.point.z;
    }
}

/* GGX pieces in FP64 for the sampling / quadrature checks (sampling.h:93-140). */
double orc_ggx_d(double noh, double alpha) { return ggx_d_f64(noh, alpha); }
double orc_ggx_vis(double nov, double nol, double alpha) { return ggx_vis_f64(nov, nol, alpha); }
int orc_ggx_sample(const double wo_local[3], double alpha, double u1, double u2, double wi_local[3], double *weight)
{
    vec_f64 ll;
    double  g, voh;
    if (!ggx_sample_f64(mk_f64(wo_local[0], wo_local[1], wo_local[2]), alpha, u1, u2, &ll, &g, &voh)) return 0;
    wi_local[0] = ll.x;
    wi_local[1] = ll.y;
    wi_local[2] = ll.z;
    *weight     = g;
    return 1;
}

/* Exported single steps of the bounce loop for the closed-form checks of the branches. */
#define SCATTER_API(SUFFIX, RT)                                                                       \
    void orc_scatter_##SUFFIX(const orc_scene *sc, uint32_t mode, int32_t prim, const float point[3],  \
                              const float normal[3], float uvx, float uvy, const float in_d[3],        \
                              const float u[4], const float u2[4], orc_scatter_out *out)               \
    {                                                                                                  \
        orc_settings st;                                                                               \
        memset(&st, 0, sizeof st);                                                                     \
        st.mode = mode;                                                                                \
        hit_##SUFFIX h;                                                                                \
        memset(&h, 0, sizeof h);                                                                       \
        h.prim   = prim;                                                                               \
        h.point  = mk_##SUFFIX(point[0], point[1], point[2]);                                          \
        h.normal = mk_##SUFFIX(normal[0], normal[1], normal[2]);                                       \
        h.uvx    = uvx;                                                                                \
        h.uvy    = uvy;                                                                                \
        scatter_##SUFFIX ph;                                                                           \
        process_hit_##SUFFIX(sc, &st, &h, mk_##SUFFIX(in_d[0], in_d[1], in_d[2]), u, u2, &ph);         \
        const RT v[15] = {ph.o.x, ph.o.y, ph.o.z, ph.d.x, ph.d.y, ph.d.z, ph.albedo.x, ph.albedo.y,    \
                          ph.albedo.z, ph.nee_n.x, ph.nee_n.y, ph.nee_n.z, ph.emission, 0, 0};         \
        for (int k = 0; k < 3; ++k) {                                                                  \
            out->o[k]      = (float)v[k];                                                              \
            out->d[k]      = (float)v[3 + k];                                                          \
            out->albedo[k] = (float)v[6 + k];                                                          \
            out->nee_n[k]  = (float)v[9 + k];                                                          \
        }                                                                                              \
        for (int k = 0; k < 4; ++k) out->colour[k] = ph.colour[k];                                     \
        out->emission = (float)v[12];                                                                  \
        out->flags    = (ph.alpha ? 1 : 0) | (ph.delta ? 2 : 0) | (ph.absorbed ? 4 : 0);               \
    }
SCATTER_API(f32, float)
SCATTER_API(f64, double)

void orc_sky_f32(const orc_scene *sc, const float d[3], float out[3])
{
    const vec_f32 s = sky_f32(sc, mk_f32(d[0], d[1], d[2]));
    out[0] = s.x;
    out[1] = s.y;
    out[2] = s.z;
}
