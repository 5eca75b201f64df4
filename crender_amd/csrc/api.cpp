// api.cpp -- C ABI of the crt path tracer (declared in include/crt.h).
//
// Host side of the drop-in boundary for CRender's render hot path: scene-load
// (cr::scene / cr::registry), the Embree query (rtcIntersect1 via scene::cast_ray) and
// the renderer (cr::renderer: start / pause / update / current_* / current_stats).
// Every entry point returns a status code and records a thread-local message; nothing
// here exits the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <limits>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/crt.h"
#include "crt_internal.h"
#include "nrc.h"

using namespace crt;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_CHECK(expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(e_ == hipErrorOutOfMemory ? CRT_ERR_OUT_OF_MEMORY : CRT_ERR_HIP,      \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

template <class T>
hipError_t dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void **)p, sizeof(T) * count);
}

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

uint64_t fnv1a(uint64_t h, const void *data, size_t n)
{
    const uint8_t *p = static_cast<const uint8_t *>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct ModelHost {
    std::string               name;
    std::vector<float>        corners;   // 9 per triangle
    std::vector<float>        uv;        // 6 per triangle (empty: no uv)
    std::vector<uint32_t>     mat_idx;   // local material index per triangle
    std::vector<crt_material> mats;
    std::vector<std::vector<float>> tex_data;
    std::vector<std::pair<uint32_t, uint32_t>> tex_dims;
    std::vector<float>        instances; // 16 per instance
};

}  // namespace

namespace crt {
// Error reporting for the other host sources of the library (same thread-local message).
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace crt

// Camera paths a renderer's wavefront aims at (whole sample passes of the image, at most
// kMaxWavefrontPasses of them): 16M paths put 8 passes of a 1080p image in one wavefront.
// A renderer that trains the radiance cache takes half: each wavefront's training steps
// scale with its passes, and the next wavefront's cache inference waits for them.
constexpr uint32_t kWavefrontPaths      = 1u << 24;
constexpr uint32_t kWavefrontPathsCache = 1u << 23;
constexpr uint32_t kMaxWavefrontPasses = 8;

struct crt_context {
    int         device = 0;
    crt_options opt{};
    hipStream_t stream = nullptr;
    uint32_t    grid   = 0;
    uint32_t   *spill  = nullptr;
    // a renderer keeps two wavefronts in flight, one per stream, so that one's traversal
    // fills the GPU while the other's launches drain; the second has its own spill area
    hipStream_t stream2 = nullptr;
    uint32_t   *spill2  = nullptr;
    uint32_t    stack  = kStackLds;   // LDS stack entries per traversal lane (8, 16, 32)
    uint32_t    refill = 0, park_min = 0, leaf_rounds = 0;   // traversal thresholds (0: defaults)
    uint32_t    batch_paths = 0;   // tuned wavefront size of crt_render_batch (0 = default)
    // staging for crt_intersect / crt_occluded
    uint32_t    stage_cap = 0;
    float      *st_ray[8] = {};
    float      *st_t = nullptr, *st_u = nullptr, *st_v = nullptr;
    uint32_t   *st_tri = nullptr, *st_prim = nullptr, *st_geom = nullptr;
    uint32_t   *st_inst = nullptr, *st_inst_out = nullptr;
    float      *st_ng[3] = {};
    uint8_t    *st_occ = nullptr;
    uint32_t   *st_cnt = nullptr;   // [0] count, [8..15] work counters, [16] error
    uint8_t    *noalloc = nullptr;  // this device's crt_intersect_noalloc scratch (noalloc.hip)
    std::mutex  mu;                 // serialises queries on the context stream
    std::atomic<int> live{0};       // scenes, renderers and caches created on the context
};

struct crt_scene {
    crt_context           *ctx = nullptr;
    std::vector<ModelHost> models;
    bool                   has_sky = false;
    std::vector<float>     sky;
    uint32_t               sky_w = 0, sky_h = 0;
    float                  sky_rot[2] = {0.0f, 0.0f};
    int                    sun_enabled = 1;   // scene.h:45
    crt_sun                sun{};
    crt_camera             cam{};
    bool                   committed = false;
    bool                   dirty     = true;
    // device
    SceneDev               dev{};
    void                  *d_nodes = nullptr;
    float4                *d_tri = nullptr;
    uint32_t              *d_mat_idx = nullptr;
    float2                *d_uv = nullptr;
    MaterialDev           *d_mats = nullptr;
    float4                *d_texels = nullptr;
    TextureDev            *d_texs = nullptr;
    uint32_t              *d_model_first = nullptr;
    InstDev               *d_inst = nullptr;
    uint32_t              *d_tlas_inst = nullptr;
    uint32_t               ninst = 0;
    crt_bvh_info           info{};
    uint64_t               hash = 0;
    uint32_t               users = 0;  // renderers referencing this scene
};

struct crt_nrc {
    crt_context    *ctx = nullptr;
    crt_nrc_options opt{};
    float          *params = nullptr, *m = nullptr, *v = nullptr;
    __bf16         *W = nullptr, *WT = nullptr;
    __bf16         *act = nullptr;
    float          *partial = nullptr, *partial_loss = nullptr, *grad = nullptr;
    float          *q_stage = nullptr, *t_stage = nullptr, *o_stage = nullptr;
    uint32_t        stage_cap = 0;
    uint32_t        tiles_alloc = 0;
    uint64_t        step = 0;
    int             users = 0;   // renderers the cache is attached to
    // Training after a renderer's wavefront runs on its own stream, beside the next
    // wavefront's traversal; the next inference waits for ev_trained.
    hipStream_t     ts = nullptr;
    hipEvent_t      ev_targets = nullptr, ev_trained = nullptr;
    std::atomic<bool> pending{false};  // training enqueued on ts and not yet waited for
    std::mutex      mu;
};

struct crt_renderer {
    crt_context *ctx = nullptr;
    crt_scene   *scene = nullptr;
    uint32_t     W = 0, H = 0, bounces = 0;
    float        aspect = 1.0f;      // renderer.h:80: 1 until set_resolution
    uint64_t     target = 0;         // renderer.h:81: 0 = unlimited
    crt_render_options opt{};
    std::atomic<uint64_t> samples{0};
    uint32_t     kmax = 1;           // sample passes per wavefront (path slots: kmax * W * H)
    PathBuffers  b{};
    // second set of path buffers (queues, hits, path state, counters, statistics, cache
    // queries and targets) for the wavefront on the second stream; the image (raw sum and
    // AOVs) is shared.  Not used with ray or material sorting (one sort scratch).
    PathBuffers  b2{};
    bool         two_sets = false;
    hipEvent_t   ev_accdone[2] = {};   // after a set's accumulation: the next one's waits for it
    hipEvent_t   ev_aov[2]     = {};   // after a set's bounce-0 shading (AOV writes)
    uint32_t     warm_enq = 0;         // cache warm-up passes enqueued
    RaySortScratch sort{};
    float4      *display = nullptr;
    std::vector<void *> allocs;
    // background passes
    std::thread              worker;
    std::mutex               mu;
    std::condition_variable  cv;
    bool                     running = false;
    std::atomic<bool>        stop{false};
    int                      worker_status = CRT_OK;
    std::string              worker_error;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    // phase timing: [0] start of pass, then one event between consecutive phases (after
    // generate; per bounce i < kTimedBounces after extend, after shade and, with the sun, after
    // shadow), [kEvAcc] before accumulate when the cache's inference precedes it, [kEvEnd]
    // end of pass.  Two sets, so that a pass can be enqueued while the previous one runs.
    static constexpr int kTimedBounces = 15, kEvAcc = 62, kEvEnd = 63;
    hipEvent_t ev[2][64] = {};
    uint32_t  *err_host = nullptr;     // pinned: per event set, the traversal error flag [set], the
                                       // diverged-cache flag [4 + set] and
                                       // the cache-training target count [2 + set]
    double ms_phase[6] = {0, 0, 0, 0, 0, 0};  // generate, extend, shade, shadow, accumulate, render
    uint64_t extend_launches = 0;
    // closest-hit launch intervals of the current render_passes call (ms after ev_origin,
    // recorded on the first stream before anything else of the call), merged into
    // ms_extend_busy when the call ends
    hipEvent_t                           ev_origin = nullptr;
    std::vector<std::pair<float, float>> ext_iv;
    double                               ms_extend_busy = 0.0;
    // neural radiance cache
    crt_nrc               *nrc = nullptr;
    crt_nrc_render_options nopt{};
    double                 ms_nrc_train = 0.0;
    uint32_t               loss_tiles = 0;   // tiles of the last training step (loss partials)
    uint32_t               nrc_warm_done = 0;   // warm-up passes run since set_nrc
    // the cache's predictions became non-finite: it is not used again until set_nrc, and the
    // sample passes that used it are dropped and rendered again
    bool                   nrc_failed = false;
    uint32_t               nrc_discarded = 0;
    // GPU time of the cache's training steps: events around a training group on the cache's
    // stream, read once the group is known to be done
    hipEvent_t             ev_train[2] = {};
    bool                   train_timed = false;   // ev_train holds an unread group
    // A failed pass may have added samples to the accumulation without counting them (or the
    // reverse): passes are refused until reset_accum (start / reset / resize) clears it.
    bool                   needs_reset = false;
};

namespace {

TraceConfig trace_cfg(const crt_context *ctx)
{
    return TraceConfig{ctx->grid, ctx->stack, ctx->refill, ctx->park_min, ctx->leaf_rounds};
}

int ensure_spill(crt_context *ctx, int set = 0)
{
    uint32_t *&sp = set ? ctx->spill2 : ctx->spill;
    if (sp) return CRT_OK;
    HIP_CHECK(hipMalloc((void **)&sp, trace_spill_bytes(ctx->grid)));
    return CRT_OK;
}

void identity16(float *m)
{
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
}

bool is_identity(const float *m)
{
    for (int i = 0; i < 16; ++i)
        if (m[i] != ((i % 5 == 0) ? 1.0f : 0.0f)) return false;
    return true;
}

// World->object transform of a column-major affine object->world matrix m, as a
// row-major 3x4 in double; false if singular.
bool affine_inverse(const float *m, double *inv)
{
    const double a[3][3] = {{m[0], m[4], m[8]}, {m[1], m[5], m[9]}, {m[2], m[6], m[10]}};
    const double t[3]    = {m[12], m[13], m[14]};
    const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1], c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2],
                 c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    if (!(std::fabs(det) > 1e-30) || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    double r[3][3];
    r[0][0] = c00 * id;
    r[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
    r[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
    r[1][0] = c01 * id;
    r[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
    r[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
    r[2][0] = c02 * id;
    r[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
    r[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) inv[4 * i + j] = r[i][j];
        inv[4 * i + 3] = -(r[i][0] * t[0] + r[i][1] * t[1] + r[i][2] * t[2]);
    }
    return true;
}

// true unless every model has exactly the identity instance
bool scene_instanced(const crt_scene *scene)
{
    for (const auto &m : scene->models)
        if (m.instances.size() != 16 || !is_identity(m.instances.data())) return true;
    return false;
}

void free_scene_device(crt_scene *s)
{
    dfree(s->d_nodes);
    dfree(s->d_tri);
    dfree(s->d_mat_idx);
    dfree(s->d_uv);
    dfree(s->d_mats);
    dfree(s->d_texels);
    dfree(s->d_texs);
    dfree(s->d_model_first);
    dfree(s->d_inst);
    dfree(s->d_tlas_inst);
    s->d_inst      = nullptr;
    s->d_tlas_inst = nullptr;
    s->ninst       = 0;
    s->d_nodes = nullptr;
    s->d_tri = nullptr;
    s->d_mat_idx = nullptr;
    s->d_uv = nullptr;
    s->d_mats = nullptr;
    s->d_texels = nullptr;
    s->d_texs = nullptr;
    s->d_model_first = nullptr;
    s->dev       = SceneDev{};   // no stale device pointers survive the free
    s->committed = false;
}

void free_renderer_buffers(crt_renderer *r)
{
    for (void *p : r->allocs) dfree(p);
    r->allocs.clear();
    r->b       = PathBuffers{};
    r->b2      = PathBuffers{};
    r->display = nullptr;
}

template <class T>
int lalloc(std::vector<void *> &list, T **p, size_t count)
{
    HIP_CHECK(dalloc(p, count));
    list.push_back(*p);
    return CRT_OK;
}

template <class T>
int ralloc(crt_renderer *r, T **p, size_t count)
{
    return lalloc(r->allocs, p, count);
}

int alloc_rays_list(std::vector<void *> &list, RaySoA &q, size_t n)
{
    float **f[8] = {&q.ox, &q.oy, &q.oz, &q.dx, &q.dy, &q.dz, &q.tn, &q.tf};
    for (float **p : f) {
        const int st = lalloc(list, p, n);
        if (st) return st;
    }
    return CRT_OK;
}

int alloc_rays(crt_renderer *r, RaySoA &q, size_t n) { return alloc_rays_list(r->allocs, q, n); }

// Queue, hit, path-state and cache-query buffers of P paths (no image buffers).
int alloc_path_core(std::vector<void *> &L, PathBuffers &b, size_t P)
{
    int st;
#define LA(ptr, n)                               \
    if ((st = lalloc(L, ptr, n)) != CRT_OK) return st
    if ((st = alloc_rays_list(L, b.q[0], P))) return st;
    if ((st = alloc_rays_list(L, b.q[1], P))) return st;
    if ((st = alloc_rays_list(L, b.sq, 2 * P))) return st;
    LA(&b.qpath[0], P);
    LA(&b.qpath[1], P);
    LA(&b.hits.t, P);
    LA(&b.hits.u, P);
    LA(&b.hits.v, P);
    LA(&b.hits.tri, P);
    LA(&b.hits.inst, P);
    LA(&b.shits.inst, P);
    LA(&b.shits.t, P);
    LA(&b.shits.u, P);
    LA(&b.shits.v, P);
    LA(&b.shits.tri, P);
    LA(&b.spath, 2 * P);
    LA(&b.scontrib, 2 * P);
    LA(&b.occ, P);
    LA(&b.thr, P);
    LA(&b.rad, P);
    LA(&b.cnt, (size_t)kCntWords);
    LA(&b.stats, (size_t)4 * kStatSlots);
#undef LA
    return CRT_OK;
}

int alloc_nrc_queries(std::vector<void *> &L, PathBuffers &b, size_t P)
{
    int st;
    if ((st = lalloc(L, &b.nstate, P))) return st;
    if ((st = lalloc(L, &b.nq, P * 13))) return st;
    if ((st = lalloc(L, &b.nq_path, P))) return st;
    if ((st = lalloc(L, &b.nq_thr, P))) return st;
    if ((st = lalloc(L, &b.nq_out, P * 3))) return st;
    return CRT_OK;
}

// One set of path buffers for kmax sample passes plus the cache's training paths; the image
// (raw sum, AOVs) only when `image` (the second set shares the first one's).
int alloc_path_set(crt_renderer *r, PathBuffers &b, bool image)
{
    b.W    = r->W;
    b.H    = r->H;
    b.npix = r->W * r->H;
    b.ncam = b.npix * r->kmax;
    const size_t T = r->nrc ? (size_t)r->nopt.train_paths * r->kmax : 0u;
    const size_t S = r->nrc ? r->nopt.train_suffix : 0u;
    b.P = b.ncam + (uint32_t)T;
    const size_t P = b.P, NP = b.npix;
    if ((uint64_t)b.ncam + T > 0xFFFFFFFFull) return fail(CRT_ERR_INVALID_ARG, "too many path slots");
    int st;
#define RA(ptr, n)                          \
    if ((st = ralloc(r, ptr, n)) != CRT_OK) return st
    if ((st = alloc_rays(r, b.q[0], P))) return st;
    if ((st = alloc_rays(r, b.q[1], P))) return st;
    if ((st = alloc_rays(r, b.sq, 2 * P))) return st;
    RA(&b.qpath[0], P);
    RA(&b.qpath[1], P);
    RA(&b.hits.t, P);
    RA(&b.hits.u, P);
    RA(&b.hits.v, P);
    RA(&b.hits.tri, P);
    RA(&b.hits.inst, P);
    RA(&b.shits.inst, P);
    RA(&b.shits.t, P);
    RA(&b.shits.u, P);
    RA(&b.shits.v, P);
    RA(&b.shits.tri, P);
    RA(&b.spath, 2 * P);
    RA(&b.scontrib, 2 * P);
    RA(&b.occ, P);
    RA(&b.thr, P);
    RA(&b.rad, P);
    if (image) {
        RA(&b.raw, 3 * NP);
        RA(&b.aov_alb, NP);
        RA(&b.aov_nrm, NP);
        RA(&b.aov_dep, NP);
    }
    RA(&b.cnt, (size_t)kCntWords);
    RA(&b.stats, (size_t)4 * kStatSlots);
    if (T) {
        RA(&b.rq, T * S * 13);
        RA(&b.r_thr, T * S);
        RA(&b.r_rad, T * S);
        RA(&b.tq, T * S * 13);
        RA(&b.tt, T * S * 3);
    }
    if (r->ctx->opt.flags & CRT_FLAG_SORT_MATERIALS) {
        RA(&b.shade_perm, P);
        RA(&b.shade_bcount, (size_t)kShadeClasses * ((P + 255) / 256));
    }
    if (r->ctx->opt.flags & CRT_FLAG_TRAVERSAL_STATS) RA(&b.tstats, (size_t)8);
#undef RA
    if ((st = alloc_nrc_queries(r->allocs, b, P))) return st;
    hipStream_t s = r->ctx->stream;
    HIP_CHECK(hipMemsetAsync(b.cnt, 0, sizeof(uint32_t) * kCntWords, s));
    HIP_CHECK(hipMemsetAsync(b.stats, 0, sizeof(unsigned long long) * 4 * kStatSlots, s));
    if (b.tstats) HIP_CHECK(hipMemsetAsync(b.tstats, 0, sizeof(unsigned long long) * 8, s));
    return CRT_OK;
}

int alloc_renderer_buffers(crt_renderer *r)
{
    free_renderer_buffers(r);
    const uint32_t npix = r->W * r->H;
    // a wavefront carries up to kmax sample passes of the image (about wavefront_paths camera
    // paths: a traversal launch's fixed cost, its ramp-up and the tail of its longest rays,
    // is then spread over more rays), then train_paths cache-training paths per pass
    const uint64_t wf = r->ctx->opt.wavefront_paths ? r->ctx->opt.wavefront_paths
                                                    : (r->nrc ? kWavefrontPathsCache : kWavefrontPaths);
    r->kmax = (uint32_t)std::max<uint64_t>(1u, std::min<uint64_t>(kMaxWavefrontPasses, wf / npix));
    r->two_sets = (r->ctx->opt.flags & (CRT_FLAG_SORT_RAYS | CRT_FLAG_SORT_MATERIALS)) == 0;
    int st = alloc_path_set(r, r->b, true);
    if (st) return st;
    if (r->two_sets) {
        if ((st = alloc_path_set(r, r->b2, false))) return st;
        r->b2.raw     = r->b.raw;
        r->b2.aov_alb = r->b.aov_alb;
        r->b2.aov_nrm = r->b.aov_nrm;
        r->b2.aov_dep = r->b.aov_dep;
    }
    const size_t P = r->b.P, NP = npix;
#define RA(ptr, n)                          \
    if ((st = ralloc(r, ptr, n)) != CRT_OK) return st
    RA(&r->display, NP);
    r->sort = RaySortScratch{};
    if (r->ctx->opt.flags & CRT_FLAG_SORT_RAYS) {
        RA(&r->sort.keys, P);
        RA(&r->sort.keys_out, P);
        RA(&r->sort.vals, P);
        RA(&r->sort.vals_out, P);
        r->sort.temp_bytes = ray_sort_temp_bytes((uint32_t)P);
        RA(reinterpret_cast<uint8_t **>(&r->sort.temp), std::max<size_t>(r->sort.temp_bytes, 1));
        r->sort.capacity = (uint32_t)P;
    }
#undef RA
    const PathBuffers &b = r->b;
    hipStream_t        s = r->ctx->stream;
    HIP_CHECK(hipMemsetAsync(b.raw, 0, sizeof(float) * 3 * NP, s));
    // image.h:30-33: new images are filled with FLT_MAX
    HIP_CHECK(launch_fill(b.aov_alb, NP, 3.402823466e+38f, s));
    HIP_CHECK(launch_fill(b.aov_nrm, NP, 3.402823466e+38f, s));
    HIP_CHECK(launch_fill(b.aov_dep, NP, 3.402823466e+38f, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return CRT_OK;
}

camera_dev to_camera_dev(const crt_camera &c)
{
    camera_dev d{};
    for (int k = 0; k < 3; ++k) d.pos[k] = c.position[k];
    d.fov = c.fov;
    for (int k = 0; k < 16; ++k) d.m[k] = c.matrix[k];
    d.mode  = c.mode;
    d.scale = c.scale;
    return d;
}

// unit_dir: renderer rays are traced along their normalised direction (cast_ray), queries
// of the ray API along the direction as given (rtcIntersect1).
TraceQueue make_queue(const RaySoA &rays, const HitSoA &hits, const uint32_t *count, uint32_t *work, uint32_t *err,
                      bool unit_dir = true)
{
    TraceQueue q{};
    q.unit_dir = unit_dir ? 1u : 0u;
    q.rays     = rays;
    q.hits     = hits;
    q.count    = count;
    q.counters = work;
    q.error    = err;
    return q;
}

// A sample pass in flight: the event set it records into and what finishing it needs.
struct PassTicket {
    int      set    = 0;          // event set (two, so a wavefront can be enqueued while one runs)
    int      bset   = 0;          // path buffer set and stream (0 or 1; always 0 with one set)
    RenderParams rp{};            // of the wavefront (its tail is enqueued after its traversal)
    bool     warmup = false;
    bool     timing = false;
    uint32_t units  = 1;   // passes the wavefront counts for: its samples, or 1 warm-up pass
    uint32_t nsamp  = 0;   // sample passes of the image (0: cache warm-up)
    bool     train  = false;   // cache training targets were computed for this wavefront
    bool     discard = false;  // rendered with a diverged cache: not accumulated, not counted
    uint32_t count  = 0;       // their number (read when the wavefront is finished)
    // phase events (timing): each phase ends at its event and starts at the one recorded
    // before it; -1 where a phase did not run
    uint32_t timed  = 0;    // bounces with phase events
    int      ev_gen = -1;   // after generate
    int      ev_acc = -1;   // before accumulate
    int      ev_ext[crt_renderer::kTimedBounces]    = {};
    int      ev_shade[crt_renderer::kTimedBounces]  = {};
    int      ev_shadow[crt_renderer::kTimedBounces] = {};
};

// One set of paths: generate, then per bounce extend / shade / shadow.  Each traversal
// launch's epilogue clears the counters the next steps append to, so a bounce is two or
// three kernels.  With timing on, one event is recorded between consecutive phases (into
// `ev`, slots 1 ..) and read after the pass's final sync.
PathBuffers &path_set(crt_renderer *r, int set) { return set ? r->b2 : r->b; }
hipStream_t  set_stream(const crt_renderer *r, int set) { return set ? r->ctx->stream2 : r->ctx->stream; }

int run_paths(crt_renderer *r, const PathBuffers &b, const RenderParams &rp, hipEvent_t *ev, PassTicket &t)
{
    crt_context    *ctx = r->ctx;
    const SceneDev &sc  = r->scene->dev;
    const int       set = t.bset;
    hipStream_t     s   = set_stream(r, set);
    int             nev = 1;   // next free event slot
    auto mark = [&](int &slot) -> hipError_t {
        slot = nev++;
        return hipEventRecord(ev[slot], s);
    };
    t.timed = 0;
    if (int st = ensure_spill(ctx, set)) return st;   // the autotuner may have resized the grid
    uint32_t *spill = set ? ctx->spill2 : ctx->spill;
    // AOVs come from the last sample pass (renderer.cpp:367-369): a wavefront writes them
    // (bounce-0 shading, or generate when there are no bounces) after the previous one did
    if (rp.max_bounces == 0) HIP_CHECK(hipStreamWaitEvent(s, r->ev_aov[set ^ 1], 0));
    HIP_CHECK(launch_generate(sc, b, rp, s));
    if (rp.max_bounces == 0) HIP_CHECK(hipEventRecord(r->ev_aov[set], s));
    if (t.timing) HIP_CHECK(mark(t.ev_gen));
    uint32_t cur = 0;
    // bounces to run: camera paths stop at max_bounces, or at term + 1 at the latest when
    // they end in the cache (later only through alpha pass-through); training paths at
    // train_term + 1
    uint32_t nb = rp.max_bounces;
    if (rp.nrc.enabled) {
        uint32_t cam = sc.any_alpha ? rp.max_bounces : std::min(rp.max_bounces, rp.nrc.term + 2);
        if (rp.nsamp == 0) cam = 0;
        nb = std::max(cam, rp.nrc.train_term + 1);
    }
    const uint32_t timed = t.timing ? std::min<uint32_t>(nb, crt_renderer::kTimedBounces) : 0u;
    const bool sort = (ctx->opt.flags & CRT_FLAG_SORT_RAYS) && r->sort.keys;
    // While the cache trains beside the traversal, the traversal keeps a 16-entry LDS stack:
    // two 32-entry traversal blocks (2 x 67.6 KB) leave no room in a CU's LDS for a training
    // block (31.7 KB), two 16-entry ones (2 x 34.3 KB) do.
    TraceConfig tc = trace_cfg(ctx);
    if (rp.nrc.enabled && r->nopt.train) tc.stack = std::min<uint32_t>(tc.stack, 16u);
    for (uint32_t i = 0; i < nb; ++i) {
        const bool tm = i < timed;
        if (sort && i > 0) {
            // reorder the secondary rays (counted in the extend phase)
            uint32_t n = 0;
            HIP_CHECK(hipMemcpyAsync(&n, b.cnt + cur, sizeof n, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            n = std::min(n, b.P);
            HIP_CHECK(launch_ray_sort(b.q[cur], b.qpath[cur], b.cnt + cur, n, b.q[cur ^ 1u], b.qpath[cur ^ 1u],
                                      b.cnt + (cur ^ 1u), sc.bmin, sc.bmax, r->sort, s));
            cur ^= 1u;
        }
        // the epilogue clears the queue the shading appends to (the previous bounce's
        // input), the shadow queue, and the work counters; it counts the extension rays
        TraceQueue q = make_queue(b.q[cur], b.hits, b.cnt + cur, b.cnt + kCntWorkExt, b.cnt + kCntErr);
        q.tstats     = b.tstats;
        q.done       = b.cnt + kCntDoneExt;
        q.clear0     = b.cnt + (cur ^ 1u);
        q.clear1     = b.cnt + kCntShadow;
        q.stat       = b.stats + 0;
        HIP_CHECK(launch_trace_closest(sc, q, spill, tc, s));
        if (tm) HIP_CHECK(mark(t.ev_ext[i]));
        if (i == 0) HIP_CHECK(hipStreamWaitEvent(s, r->ev_aov[set ^ 1], 0));
        if (b.shade_perm) {
            // material-sorted shading: the queue's hits grouped by material class
            HIP_CHECK(launch_shade_sort(sc, b, cur, s));
            HIP_CHECK(launch_shade(sc, b, rp, i, cur, s, b.shade_perm));
        } else {
            HIP_CHECK(launch_shade(sc, b, rp, i, cur, s));
        }
        if (i == 0) HIP_CHECK(hipEventRecord(r->ev_aov[set], s));
        if (tm) HIP_CHECK(mark(t.ev_shade[i]));
        if (tm) t.ev_shadow[i] = -1;
        if (sc.sun_enabled) {
            if (!sc.any_alpha) {
                TraceQueue qs = make_queue(b.sq, HitSoA{}, b.cnt + kCntShadow, b.cnt + kCntWorkShadow, b.cnt + kCntErr);
                qs.tstats     = b.tstats;
                qs.done       = b.cnt + kCntDoneShadow;
                qs.stat       = b.stats + 1;
                HIP_CHECK(launch_trace_any(sc, qs, b.occ, spill, tc, s));
                HIP_CHECK(launch_shadow_resolve(b, s));
            } else {
                // closest hit + alpha pass-through, renderer.cpp:330-345 (bounded at 64 steps)
                HIP_CHECK(launch_shadow_prep(b, s));
                for (int step = 0; step < 64; ++step) {
                    TraceQueue qs = make_queue(b.sq, b.shits, b.cnt + kCntShadow, b.cnt + kCntWorkShadow, b.cnt + kCntErr);
                    HIP_CHECK(launch_trace_closest(sc, qs, spill, tc, s));
                    HIP_CHECK(launch_shadow_alpha_step(sc, b, s));
                    uint32_t left = 0;
                    HIP_CHECK(hipMemcpyAsync(&left, b.cnt + kCntShadow, sizeof left, hipMemcpyDeviceToHost, s));
                    HIP_CHECK(hipStreamSynchronize(s));
                    if (left == 0) break;
                }
            }
            if (tm) HIP_CHECK(mark(t.ev_shadow[i]));
        }
        cur ^= 1u;
    }
    t.timed = timed;
    r->extend_launches += timed;
    if (t.timing) t.ev_acc = nev - 1;   // accumulate starts at the last phase event
    return CRT_OK;
}

NrcParams nrc_params(const crt_renderer *r)
{
    NrcParams np{};
    const SceneDev &sc = r->scene->dev;
    np.bmin = sc.bmin;
    const f3 ext = sc.bmax - sc.bmin;
    np.binv = mk3(1.0f / std::max(ext.x, 1e-6f), 1.0f / std::max(ext.y, 1e-6f), 1.0f / std::max(ext.z, 1e-6f));
    if (r->nrc && r->bounces > 0 && !r->nrc_failed) {
        np.enabled     = 1;
        np.term        = r->nopt.term_bounce;
        np.footprint_c = r->nopt.footprint_c;
        np.rec_first   = r->nopt.term_bounce;
        np.rec_slots   = r->nopt.train_suffix;
        np.train_term  = r->nopt.term_bounce + r->nopt.train_suffix;
    }
    return np;
}

int nrc_train_step(crt_renderer *r, const PathBuffers &t, uint32_t mult, uint32_t count);
void harvest_train_time(crt_renderer *r, bool wait);
int nrc_wait_training(crt_nrc *n);
int nrc_wait_locked(crt_nrc *n);   // nrc_wait_training with n->mu held by the caller

// Enqueues the cache training of a finished wavefront (on the cache's stream).
int train_after(crt_renderer *r, const PassTicket &t)
{
    if (!t.train) return CRT_OK;
    return nrc_train_step(r, path_set(r, t.bset), std::max(1u, t.nsamp), t.count);
}

// Adds the GPU time of the last timed training group once it is done (wait: block for it).
void harvest_train_time(crt_renderer *r, bool wait)
{
    if (!r->train_timed) return;
    if (wait) (void)hipEventSynchronize(r->ev_train[1]);
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r->ev_train[0], r->ev_train[1]) == hipSuccess) {
        r->ms_nrc_train += ms;
        r->train_timed = false;
    }
}

// Enqueues the traversal part of one wavefront of `nsamp` sample passes (renderer.cpp:121-125,
// the pass barrier being stream order) from sample index `sample` on -- or of one cache
// warm-up pass while those are due -- on the stream of its buffer set: generate, then every
// bounce's extend / shade / shadow.  Nothing waits for the GPU here.
int enqueue_trace(crt_renderer *r, uint64_t sample, uint32_t nsamp, int set, int bset, PassTicket &t)
{
    crt_context       *ctx = r->ctx;
    const PathBuffers &b   = path_set(r, bset);
    hipStream_t        s   = set_stream(r, bset);
    hipEvent_t        *ev  = r->ev[set];
    RenderParams      &rp  = t.rp;
    rp.cam         = to_camera_dev(r->scene->cam);
    rp.aspect      = r->aspect;
    rp.max_bounces = r->bounces;
    rp.rr_start    = r->opt.rr_start;
    rp.mode        = r->opt.mode;
    rp.seed        = r->opt.seed;
    rp.sample      = (uint32_t)(r->opt.sample_offset + sample);
    rp.nsamp       = nsamp;
    rp.nrc         = nrc_params(r);
    // cache warm-up pass: only the training paths (their own Philox stream), no image
    t.set    = set;
    t.bset   = bset;
    t.warmup = rp.nrc.enabled && r->warm_enq < r->nopt.warmup_passes;
    if (t.warmup) {
        rp.sample = 0x40000000u + r->warm_enq++;
        rp.nsamp  = 0;
    }
    t.nsamp = rp.nsamp;
    t.units = t.warmup ? 1u : nsamp;
    // training paths: train_paths per sample pass (one warm-up pass counts as one)
    if (rp.nrc.enabled) rp.nrc.train_live = r->nopt.train_paths * std::max(1u, rp.nsamp);
    t.timing = (ctx->opt.flags & CRT_FLAG_PHASE_TIMING) != 0 && !t.warmup;
    HIP_CHECK(hipEventRecord(ev[0], s));
    return run_paths(r, b, rp, ev, t);
}

// Enqueues the rest of the wavefront: the cache's inference (after the training enqueued so
// far), the accumulation (after the previous wavefront's, so each pixel's samples are summed
// in sample order), the cache-training targets and the error flag.
int enqueue_tail(crt_renderer *r, PassTicket &t)
{
    const PathBuffers  &b  = path_set(r, t.bset);
    hipStream_t         s  = set_stream(r, t.bset);
    hipEvent_t         *ev = r->ev[t.set];
    const RenderParams &rp = t.rp;
    if (rp.nrc.enabled && r->nrc_failed) {
        // traced with the cache, which diverged meanwhile: nothing of this wavefront is kept
        t.discard = true;
        t.train   = false;
        HIP_CHECK(hipMemsetAsync(r->err_host + 4 + t.set, 0, sizeof(uint32_t), s));
        HIP_CHECK(hipMemcpyAsync(r->err_host + t.set, b.cnt + kCntErr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipEventRecord(ev[crt_renderer::kEvEnd], s));
        return CRT_OK;
    }
    if (rp.nrc.enabled) {
        // one inference for the camera and the training paths' queries, with the weights of
        // every training step enqueued so far
        // unconditional: a wait on an event that was never recorded is a no-op, and the
        // pending flag may have been cleared by a concurrent nrc_wait_training
        HIP_CHECK(hipStreamWaitEvent(s, r->nrc->ev_trained, 0));
        HIP_CHECK(nrc_launch_infer_dev(b.nq, b.nq_out, b.cnt + kCntNrcQ, b.P, r->nrc->WT, s));
        HIP_CHECK(launch_nrc_apply(b, s));
        if (t.timing) {
            t.ev_acc = crt_renderer::kEvAcc;
            HIP_CHECK(hipEventRecord(ev[crt_renderer::kEvAcc], s));
        }
    }
    if (!t.warmup) {
        HIP_CHECK(hipStreamWaitEvent(s, r->ev_accdone[t.bset ^ 1], 0));
        HIP_CHECK(launch_accumulate(b, rp.nsamp, s));
        HIP_CHECK(hipEventRecord(r->ev_accdone[t.bset], s));
    }
    t.train = rp.nrc.enabled && (r->nopt.train || t.warmup);
    if (t.train) {
        // training targets from the wavefront's training paths; their count goes to pinned
        // memory for the host, which sizes the training steps
        const uint32_t S = r->nopt.train_suffix, live = rp.nrc.train_live;
        HIP_CHECK(hipMemsetAsync(b.cnt + kCntTrain, 0, sizeof(uint32_t), s));
        HIP_CHECK(launch_nrc_targets(b, S, live * S, live, s));
        HIP_CHECK(hipMemcpyAsync(r->err_host + 2 + t.set, b.cnt + kCntTrain, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipEventRecord(r->nrc->ev_targets, s));
    }
    HIP_CHECK(hipMemcpyAsync(r->err_host + t.set, b.cnt + kCntErr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(r->err_host + 4 + t.set, b.cnt + kCntNrcBad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipEventRecord(ev[crt_renderer::kEvEnd], s));
    return CRT_OK;
}

// Waits for an enqueued wavefront, books its timing, checks the traversal error flag,
// records its cache-training target count (the training itself is enqueued by the caller,
// after the next wavefront's traversal) and counts its samples.
int finish_pass(crt_renderer *r, PassTicket &t)
{
    hipEvent_t *ev = r->ev[t.set];
    HIP_CHECK(hipEventSynchronize(ev[crt_renderer::kEvEnd]));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[crt_renderer::kEvEnd]));
    if (!t.warmup) r->ms_phase[5] += ms;
    if (t.timing) {
        auto add = [&](int phase, int from, int to) -> hipError_t {
            hipError_t e = hipEventElapsedTime(&ms, ev[from], ev[to]);
            if (e == hipSuccess) r->ms_phase[phase] += ms;
            return e;
        };
        HIP_CHECK(add(0, 0, t.ev_gen));
        int prev = t.ev_gen;
        for (uint32_t i = 0; i < t.timed; ++i) {
            HIP_CHECK(add(1, prev, t.ev_ext[i]));
            float a = 0, b = 0;
            if (hipEventElapsedTime(&a, r->ev_origin, ev[prev]) == hipSuccess &&
                hipEventElapsedTime(&b, r->ev_origin, ev[t.ev_ext[i]]) == hipSuccess)
                r->ext_iv.emplace_back(a, b);
            HIP_CHECK(add(2, t.ev_ext[i], t.ev_shade[i]));
            prev = t.ev_shade[i];
            if (t.ev_shadow[i] >= 0) {
                HIP_CHECK(add(3, prev, t.ev_shadow[i]));
                prev = t.ev_shadow[i];
            }
        }
        HIP_CHECK(add(4, t.ev_acc, crt_renderer::kEvEnd));
    }
    if (r->err_host[t.set]) return fail(CRT_ERR_HIP, "traversal stack overflow");
    t.count = r->err_host[2 + t.set];
    if (!t.discard && r->err_host[4 + t.set]) {
        // the cache diverged: this wavefront was not accumulated; the renderer goes on
        // without the cache (automatic fallback to the plain tracer)
        r->nrc_failed = true;
        t.discard     = true;
        t.train       = false;
    }
    if (t.discard) {
        r->nrc_discarded += t.nsamp;
        return CRT_OK;
    }
    if (t.warmup) {
        r->nrc_warm_done++;
        return CRT_OK;
    }
    r->samples.fetch_add(t.nsamp);
    return CRT_OK;
}

// Up to `n` passes under one hold of the context, as wavefronts of up to kmax sample passes
// each (cache warm-up passes one at a time).  Two wavefronts are in flight: the next one's
// traversal is enqueued -- on the other stream with its own buffers, so that it runs beside
// this one's last, draining launches -- before this one is waited for; then this one's cache
// training (which needs its target count on the host) and the next one's tail (inference,
// accumulation) follow.  Accumulation in wavefront order, and AOVs from the last one, keep
// the image identical to one pass at a time.  Stops early when `r->stop` is set; `done`
// receives the number of passes counted.
int render_passes(crt_renderer *r, uint32_t n, uint32_t &done)
{
    crt_context *ctx = r->ctx;
    // the context's streams and spill areas are shared with queries, batches and the autotuner
    std::lock_guard<std::mutex> ctx_lock(ctx->mu);
    done = 0;
    if (r->needs_reset)
        return fail(CRT_ERR_STATE, "a previous pass failed; reset (or start) the renderer before rendering again");
    r->warm_enq = r->nrc_warm_done;
    r->ext_iv.clear();
    if (ctx->opt.flags & CRT_FLAG_PHASE_TIMING) HIP_CHECK(hipEventRecord(r->ev_origin, ctx->stream));
    PassTicket inflight;
    bool       busy   = false;
    uint32_t   queued = 0;
    int        st     = CRT_OK;
    int        next_set = 0;
    // Balanced wavefronts, an even number of them when two are in flight and there is more
    // than one: 20 passes at 8 per wavefront run as 5 + 5 + 5 + 5 rather than 8 + 8 + 4, so
    // that the last wavefront does not run alone for long.
    uint32_t waves = (n + r->kmax - 1) / r->kmax;
    if (r->two_sets && waves > 1 && (waves & 1u)) ++waves;
    while (queued < n && st == CRT_OK && !r->stop.load()) {
        PassTicket     t;
        const uint32_t rem   = n - queued;
        const uint32_t left  = std::max<uint32_t>(1u, std::min(waves, rem));
        const uint32_t k     = std::min(r->kmax, (rem + left - 1) / left);
        const int      bset  = r->two_sets ? next_set : 0;
        st = enqueue_trace(r, r->samples.load() + (busy ? inflight.nsamp : 0u), k, next_set, bset, t);
        if (st == CRT_OK && busy) {
            st   = finish_pass(r, inflight);
            busy = false;
            if (st == CRT_OK && inflight.discard) {
                queued -= inflight.units;   // its passes are rendered again
            } else if (st == CRT_OK) {
                done += inflight.units;
                st = train_after(r, inflight);
            }
        }
        if (st == CRT_OK) st = enqueue_tail(r, t);
        if (st != CRT_OK) break;
        if (!t.discard) queued += t.units;
        if (waves > 1) --waves;
        inflight = t;
        busy     = true;
        next_set ^= 1;
    }
    if (busy && st == CRT_OK && (st = finish_pass(r, inflight)) == CRT_OK && !inflight.discard) {
        done += inflight.units;
        st = train_after(r, inflight);
    }
    if (st != CRT_OK) {
        // leave nothing running on failure
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamSynchronize(ctx->stream2);
        r->needs_reset = true;   // accumulation and sample count may disagree
    }
    // length of the union of the closest-hit intervals of this call
    std::sort(r->ext_iv.begin(), r->ext_iv.end());
    double span = 0.0, lo = -std::numeric_limits<double>::infinity(), hi = lo;
    for (const auto &iv : r->ext_iv) {
        if (iv.first > hi) {
            if (hi > lo) span += hi - lo;
            lo = iv.first;
            hi = iv.second;
        } else if (iv.second > hi) {
            hi = iv.second;
        }
    }
    if (hi > lo) span += hi - lo;
    r->ms_extend_busy += span;
    r->ext_iv.clear();
    if (r->nrc) {
        const int sw = nrc_wait_training(r->nrc);
        if (st == CRT_OK) st = sw;
    }
    return st;
}

// Training of the cache after a wavefront of `mult` passes: the training paths (slots
// ncam .., train_paths per pass) traced train_suffix bounces past the query vertex,
// recording every vertex from term_bounce on, and terminated into the cache (self-training)
// give targets L_v = (L_path - L_before_v) / T_v; train_steps Adam steps per pass follow.
int nrc_train_step(crt_renderer *r, const PathBuffers &t, uint32_t mult, uint32_t count)
{
    crt_nrc           *n  = r->nrc;
    const uint32_t     S  = r->nopt.train_suffix;
    count = std::min(count, r->nopt.train_paths * mult * S);
    const uint32_t steps = std::max<uint32_t>(1u, r->nopt.train_steps) * mult;
    const uint32_t batch = std::min(count / steps, n->opt.max_batch);
    if (batch < 64) return CRT_OK;
    std::lock_guard<std::mutex> g(n->mu);
    // the training stream starts after the targets; the next wavefront's inference waits
    // for ev_trained, and its targets (which overwrite tq / tt) come after that inference
    hipStream_t s = n->ts;
    HIP_CHECK(hipStreamWaitEvent(s, n->ev_targets, 0));
    // the previous group is done (this wavefront's inference waited for it)
    harvest_train_time(r, true);
    HIP_CHECK(hipEventRecord(r->ev_train[0], s));
    for (uint32_t k = 0; k < steps; ++k) {
        const float *q  = t.tq + (size_t)k * batch * 13;
        const float *tt = t.tt + (size_t)k * batch * 3;
        HIP_CHECK(nrc_launch_train(q, tt, batch, n->W, n->WT, n->act, n->partial, n->partial_loss, s));
        n->step++;
        HIP_CHECK(nrc_launch_adam(n->partial, nrc_partial_rows(batch), n->params, n->m, n->v, n->W, n->WT, n->opt.lr,
                                  n->opt.beta1, n->opt.beta2, n->opt.eps, n->step, 1, s));
    }
    HIP_CHECK(hipEventRecord(r->ev_train[1], s));
    r->train_timed = true;
    HIP_CHECK(hipEventRecord(n->ev_trained, s));
    n->pending = true;
    // the loss of the last step (sum of its per-tile partials) is read when stats ask for it
    r->loss_tiles = (batch + 63) / 64;
    return CRT_OK;
}

// Waits for the cache training enqueued by renderers (the cache API and the end of a group
// of wavefronts see finished weights).
int nrc_wait_locked(crt_nrc *n)
{
    if (n->pending) {
        HIP_CHECK(hipStreamSynchronize(n->ts));
        n->pending = false;
    }
    return CRT_OK;
}

int nrc_wait_training(crt_nrc *n)
{
    if (!n) return CRT_OK;
    // under the cache's lock, which nrc_train_step holds while it enqueues and sets the flag:
    // a step enqueued after the check is either waited for here or still flagged
    std::lock_guard<std::mutex> g(n->mu);
    return nrc_wait_locked(n);
}

int check_ready(crt_renderer *r)
{
    if (!r) return fail(CRT_ERR_INVALID_ARG, "null renderer");
    if (!r->scene->committed) return fail(CRT_ERR_STATE, "scene not committed");
    return CRT_OK;
}

// Wavefronts per hold of the context: other users of the context (queries, batches, the
// autotuner) get their turn between such groups.
constexpr uint32_t kPassGroup = 8;

void worker_loop(crt_renderer *r)
{
    int st = CRT_OK;
    for (;;) {
        if (r->stop.load()) break;
        const uint64_t have = r->samples.load();
        if (r->target != 0 && have >= r->target) break;
        const uint32_t g = kPassGroup * r->kmax;
        const uint32_t n = r->target != 0 ? (uint32_t)std::min<uint64_t>(g, r->target - have) : g;
        uint32_t done = 0;
        st = render_passes(r, n, done);
        if (st != CRT_OK) break;
    }
    std::lock_guard<std::mutex> g(r->mu);
    r->worker_status = st;
    if (st != CRT_OK) r->worker_error = g_err;
    r->running = false;
    r->cv.notify_all();
}

int join_worker(crt_renderer *r)
{
    if (r->worker.joinable()) r->worker.join();
    return CRT_OK;
}

int reset_accum(crt_renderer *r)
{
    hipStream_t s = r->ctx->stream;
    HIP_CHECK(hipStreamSynchronize(r->ctx->stream2));
    HIP_CHECK(hipMemsetAsync(r->b.raw, 0, sizeof(float) * 3 * (size_t)r->b.npix, s));
    for (PathBuffers *b : {&r->b, &r->b2}) {
        if (!b->stats) continue;
        HIP_CHECK(hipMemsetAsync(b->stats, 0, sizeof(unsigned long long) * 4 * kStatSlots, s));
        if (b->tstats) HIP_CHECK(hipMemsetAsync(b->tstats, 0, sizeof(unsigned long long) * 8, s));
        HIP_CHECK(hipMemsetAsync(b->cnt, 0, sizeof(uint32_t) * kCntWords, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));
    r->samples.store(0);
    r->needs_reset = false;
    for (double &m : r->ms_phase) m = 0.0;
    r->extend_launches = 0;
    r->ms_extend_busy  = 0.0;
    r->ms_nrc_train    = 0.0;
    r->t0 = std::chrono::steady_clock::now();
    return CRT_OK;
}

int stage_reserve(crt_context *ctx, uint32_t n)
{
    if (n <= ctx->stage_cap && ctx->st_cnt) return CRT_OK;
    for (float *&p : ctx->st_ray) dfree(p), p = nullptr;
    dfree(ctx->st_t);
    dfree(ctx->st_u);
    dfree(ctx->st_v);
    dfree(ctx->st_tri);
    dfree(ctx->st_prim);
    dfree(ctx->st_geom);
    dfree(ctx->st_inst);
    dfree(ctx->st_inst_out);
    for (float *&p : ctx->st_ng) dfree(p), p = nullptr;
    dfree(ctx->st_occ);
    dfree(ctx->st_cnt);
    const uint32_t cap = std::max<uint32_t>(n, 1024);
    for (float *&p : ctx->st_ray) HIP_CHECK(dalloc(&p, cap));
    HIP_CHECK(dalloc(&ctx->st_t, cap));
    HIP_CHECK(dalloc(&ctx->st_u, cap));
    HIP_CHECK(dalloc(&ctx->st_v, cap));
    HIP_CHECK(dalloc(&ctx->st_tri, cap));
    HIP_CHECK(dalloc(&ctx->st_prim, cap));
    HIP_CHECK(dalloc(&ctx->st_geom, cap));
    HIP_CHECK(dalloc(&ctx->st_inst, cap));
    HIP_CHECK(dalloc(&ctx->st_inst_out, cap));
    for (float *&p : ctx->st_ng) HIP_CHECK(dalloc(&p, cap));
    HIP_CHECK(dalloc(&ctx->st_occ, cap));
    HIP_CHECK(dalloc(&ctx->st_cnt, 32));
    ctx->stage_cap = cap;
    return CRT_OK;
}

int query(crt_scene *scene, const crt_rays *rays, uint32_t n, crt_hits *hits, uint8_t *occluded, int memory)
{
    if (!scene || !rays) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (!scene->committed) return fail(CRT_ERR_STATE, "scene not committed");
    if (memory != CRT_MEM_HOST && memory != CRT_MEM_DEVICE) return fail(CRT_ERR_INVALID_ARG, "bad memory kind");
    const float *in[8] = {rays->org_x, rays->org_y, rays->org_z, rays->dir_x, rays->dir_y, rays->dir_z, rays->tnear, rays->tfar};
    for (const float *p : in)
        if (!p && n) return fail(CRT_ERR_INVALID_ARG, "null ray component");
    if (hits && (!hits->t || !hits->prim_id) && n) return fail(CRT_ERR_INVALID_ARG, "hits.t and hits.prim_id are required");
    if (n == 0) return CRT_OK;
    crt_context *ctx = scene->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int st = stage_reserve(ctx, n);
    if (st) return st;
    if ((st = ensure_spill(ctx))) return st;
    hipStream_t s = ctx->stream;
    RaySoA      rs;
    float     **dst[8] = {&rs.ox, &rs.oy, &rs.oz, &rs.dx, &rs.dy, &rs.dz, &rs.tn, &rs.tf};
    for (int k = 0; k < 8; ++k) {
        if (memory == CRT_MEM_HOST) {
            HIP_CHECK(hipMemcpyAsync(ctx->st_ray[k], in[k], sizeof(float) * n, hipMemcpyHostToDevice, s));
            *dst[k] = ctx->st_ray[k];
        } else {
            *dst[k] = const_cast<float *>(in[k]);
        }
    }
    HIP_CHECK(hipMemsetAsync(ctx->st_cnt, 0, sizeof(uint32_t) * 32, s));
    HIP_CHECK(hipMemcpyAsync(ctx->st_cnt, &n, sizeof n, hipMemcpyHostToDevice, s));
    if (hits) {
        HitSoA h{};
        const bool dev = memory == CRT_MEM_DEVICE;
        h.t    = dev ? hits->t : ctx->st_t;
        h.u    = dev && hits->u ? hits->u : ctx->st_u;
        h.v    = dev && hits->v ? hits->v : ctx->st_v;
        h.tri  = ctx->st_tri;
        h.inst = ctx->st_inst;
        TraceQueue q = make_queue(rs, h, ctx->st_cnt, ctx->st_cnt + 8, ctx->st_cnt + 16, false);
        HIP_CHECK(launch_trace_closest(scene->dev, q, ctx->spill, trace_cfg(ctx), s));
        uint32_t *prim = dev ? hits->prim_id : ctx->st_prim;
        uint32_t *geom = dev ? hits->geom_id : (hits->geom_id ? ctx->st_geom : nullptr);
        float    *ng[3] = {hits->ng_x, hits->ng_y, hits->ng_z};
        float    *ngd[3];
        for (int k = 0; k < 3; ++k) ngd[k] = ng[k] ? (dev ? ng[k] : ctx->st_ng[k]) : nullptr;
        uint32_t *inst = hits->inst_id ? (dev ? hits->inst_id : ctx->st_inst_out) : nullptr;
        HIP_CHECK(launch_hits_to_api(scene->dev, h, n, prim, geom, ngd[0], ngd[1], ngd[2], inst, s));
        if (!dev) {
            HIP_CHECK(hipMemcpyAsync(hits->t, ctx->st_t, sizeof(float) * n, hipMemcpyDeviceToHost, s));
            if (hits->u) HIP_CHECK(hipMemcpyAsync(hits->u, ctx->st_u, sizeof(float) * n, hipMemcpyDeviceToHost, s));
            if (hits->v) HIP_CHECK(hipMemcpyAsync(hits->v, ctx->st_v, sizeof(float) * n, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipMemcpyAsync(hits->prim_id, ctx->st_prim, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
            if (hits->geom_id)
                HIP_CHECK(hipMemcpyAsync(hits->geom_id, ctx->st_geom, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
            if (hits->inst_id)
                HIP_CHECK(hipMemcpyAsync(hits->inst_id, ctx->st_inst_out, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
            for (int k = 0; k < 3; ++k)
                if (ng[k]) HIP_CHECK(hipMemcpyAsync(ng[k], ctx->st_ng[k], sizeof(float) * n, hipMemcpyDeviceToHost, s));
        }
    } else {
        TraceQueue q = make_queue(rs, HitSoA{}, ctx->st_cnt, ctx->st_cnt + 8, ctx->st_cnt + 16, false);
        uint8_t   *occ = memory == CRT_MEM_DEVICE ? occluded : ctx->st_occ;
        HIP_CHECK(launch_trace_any(scene->dev, q, occ, ctx->spill, trace_cfg(ctx), s));
        if (memory == CRT_MEM_HOST) HIP_CHECK(hipMemcpyAsync(occluded, ctx->st_occ, n, hipMemcpyDeviceToHost, s));
    }
    uint32_t err = 0;
    HIP_CHECK(hipMemcpyAsync(&err, ctx->st_cnt + 16, sizeof err, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (err) return fail(CRT_ERR_HIP, "traversal stack overflow");
    return CRT_OK;
}

// scene::cast_ray over a batch (crt_cast_rays).  Host-memory calls stage the rays and reuse
// the staging buffers for the records: the records kernel reads each element's inputs
// before it writes that element's outputs.
int cast_records(crt_scene *scene, const crt_rays *rays, uint32_t n, crt_records *rec, int memory)
{
    if (!scene || !rays || !rec) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (!scene->committed) return fail(CRT_ERR_STATE, "scene not committed");
    if (memory != CRT_MEM_HOST && memory != CRT_MEM_DEVICE) return fail(CRT_ERR_INVALID_ARG, "bad memory kind");
    const float *in[8] = {rays->org_x, rays->org_y, rays->org_z, rays->dir_x, rays->dir_y, rays->dir_z, rays->tnear, rays->tfar};
    for (int k = 0; k < 6; ++k)
        if (!in[k] && n) return fail(CRT_ERR_INVALID_ARG, "null ray origin or direction component");
    if (!rec->distance && n) return fail(CRT_ERR_INVALID_ARG, "records.distance is required");
    if (n == 0) return CRT_OK;
    crt_context *ctx = scene->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int st = stage_reserve(ctx, n);
    if (st) return st;
    if ((st = ensure_spill(ctx))) return st;
    hipStream_t s   = ctx->stream;
    const bool  dev = memory == CRT_MEM_DEVICE;
    RaySoA      rs;
    float     **dst[8] = {&rs.ox, &rs.oy, &rs.oz, &rs.dx, &rs.dy, &rs.dz, &rs.tn, &rs.tf};
    const float dflt[2] = {1e-5f, std::numeric_limits<float>::infinity()};   // model.cpp:22-23
    for (int k = 0; k < 8; ++k) {
        if (!in[k]) {
            HIP_CHECK(launch_fill_f32(ctx->st_ray[k], n, dflt[k - 6], s));
            *dst[k] = ctx->st_ray[k];
        } else if (!dev) {
            HIP_CHECK(hipMemcpyAsync(ctx->st_ray[k], in[k], sizeof(float) * n, hipMemcpyHostToDevice, s));
            *dst[k] = ctx->st_ray[k];
        } else {
            *dst[k] = const_cast<float *>(in[k]);
        }
    }
    HIP_CHECK(hipMemsetAsync(ctx->st_cnt, 0, sizeof(uint32_t) * 32, s));
    HIP_CHECK(hipMemcpyAsync(ctx->st_cnt, &n, sizeof n, hipMemcpyHostToDevice, s));
    HitSoA h{};
    h.t    = ctx->st_t;
    h.u    = ctx->st_u;
    h.v    = ctx->st_v;
    h.tri  = ctx->st_tri;
    h.inst = ctx->st_inst;
    TraceQueue q = make_queue(rs, h, ctx->st_cnt, ctx->st_cnt + 8, ctx->st_cnt + 16, true);
    HIP_CHECK(launch_trace_closest(scene->dev, q, ctx->spill, trace_cfg(ctx), s));
    // where each field lands on the device, and where it goes on the host
    struct Field {
        void  *user;
        void  *stage;
        size_t elem;
    };
    Field f[12] = {{rec->distance, ctx->st_t, 4},     {rec->prim_id, ctx->st_prim, 4}, {rec->geom_id, ctx->st_geom, 4},
                   {rec->material, ctx->st_inst_out, 4}, {rec->u, ctx->st_u, 4},        {rec->v, ctx->st_v, 4},
                   {rec->n_x, ctx->st_ng[0], 4},      {rec->n_y, ctx->st_ng[1], 4},    {rec->n_z, ctx->st_ng[2], 4},
                   {rec->p_x, ctx->st_ray[0], 4},     {rec->p_y, ctx->st_ray[1], 4},   {rec->p_z, ctx->st_ray[2], 4}};
    void *o[12];
    for (int k = 0; k < 12; ++k) o[k] = f[k].user ? (dev ? f[k].user : f[k].stage) : nullptr;
    RecordsDev out;
    out.dist = static_cast<float *>(o[0]);
    out.prim = static_cast<uint32_t *>(o[1]);
    out.geom = static_cast<uint32_t *>(o[2]);
    out.mat  = static_cast<uint32_t *>(o[3]);
    out.u    = static_cast<float *>(o[4]);
    out.v    = static_cast<float *>(o[5]);
    out.nx   = static_cast<float *>(o[6]);
    out.ny   = static_cast<float *>(o[7]);
    out.nz   = static_cast<float *>(o[8]);
    out.px   = static_cast<float *>(o[9]);
    out.py   = static_cast<float *>(o[10]);
    out.pz   = static_cast<float *>(o[11]);
    HIP_CHECK(launch_hit_records(scene->dev, rs, h, n, out, s));
    if (!dev)
        for (int k = 0; k < 12; ++k)
            if (f[k].user) HIP_CHECK(hipMemcpyAsync(f[k].user, f[k].stage, f[k].elem * n, hipMemcpyDeviceToHost, s));
    uint32_t err = 0;
    HIP_CHECK(hipMemcpyAsync(&err, ctx->st_cnt + 16, sizeof err, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (err) return fail(CRT_ERR_HIP, "traversal stack overflow");
    return CRT_OK;
}

}  // namespace

extern "C" {

const char *crt_last_error(void) { return g_err.c_str(); }
int         crt_abi_version(void) { return CRT_ABI_VERSION; }

void crt_options_default(crt_options *opt)
{
    if (!opt) return;
    std::memset(opt, 0, sizeof *opt);
    opt->device    = 0;
    opt->bvh_width = 0;
    opt->max_leaf  = 0;
}

int crt_context_create(const crt_options *opt, crt_context **out)
{
    if (!out) return fail(CRT_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    crt_options o;
    crt_options_default(&o);
    if (opt) o = *opt;
    if (o.bvh_width != 0 && o.bvh_width != 4 && o.bvh_width != 8)
        return fail(CRT_ERR_INVALID_ARG, "bvh_width must be 0, 4 or 8");
    if (!(o.bvh_prim_cost >= 0.0f) || o.bvh_prim_cost > 1e6f)
        return fail(CRT_ERR_INVALID_ARG, "bvh_prim_cost must be a non-negative number");
    if (o.trace_refill > 64 || o.trace_park_min > 64)
        return fail(CRT_ERR_INVALID_ARG, "trace_refill and trace_park_min must be at most 64");
    if (o.trace_stack != 0 && o.trace_stack != 8 && o.trace_stack != 16 && o.trace_stack != 32)
        return fail(CRT_ERR_INVALID_ARG, "trace_stack must be 0, 8, 16 or 32");
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (o.device < 0 || o.device >= ndev) return fail(CRT_ERR_INVALID_ARG, "no such HIP device");
    HIP_CHECK(hipSetDevice(o.device));
    auto *ctx   = new crt_context();
    ctx->device = o.device;
    ctx->opt    = o;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
    if (e != hipSuccess) {
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(CRT_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    ctx->grid  = trace_grid(o.device, o.trace_waves_per_cu ? (o.trace_waves_per_cu + 3) / 4 : 0);
    ctx->stack = o.trace_stack ? o.trace_stack : kStackLds;
    ctx->refill   = o.trace_refill;
    ctx->park_min = o.trace_park_min;
    ctx->leaf_rounds = o.trace_leaf_rounds;
    // the allocation-free query's scratch: a device array whose address is looked up once per
    // device, here, so that crt_intersect_noalloc itself makes no setup call
    void *na = nullptr;
    e = noalloc_scratch(o.device, &na);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        (void)hipStreamDestroy(ctx->stream2);
        delete ctx;
        return fail(CRT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    ctx->noalloc = static_cast<uint8_t *>(na);
    *out         = ctx;
    return CRT_OK;
}

int crt_context_destroy(crt_context *ctx)
{
    if (!ctx) return CRT_OK;
    if (ctx->live.load() > 0) return fail(CRT_ERR_STATE, "context still has scenes, renderers or caches");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream2);
    dfree(ctx->spill);
    dfree(ctx->spill2);
    for (float *p : ctx->st_ray) dfree(p);
    dfree(ctx->st_t);
    dfree(ctx->st_u);
    dfree(ctx->st_v);
    dfree(ctx->st_tri);
    dfree(ctx->st_prim);
    dfree(ctx->st_geom);
    dfree(ctx->st_inst);
    dfree(ctx->st_inst_out);
    for (float *p : ctx->st_ng) dfree(p);
    dfree(ctx->st_occ);
    dfree(ctx->st_cnt);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
    return CRT_OK;
}

int crt_context_device_name(crt_context *ctx, char *buf, size_t len)
{
    if (!ctx || !buf || len == 0) return fail(CRT_ERR_INVALID_ARG, "null argument");
    hipDeviceProp_t p;
    HIP_CHECK(hipGetDeviceProperties(&p, ctx->device));
    std::snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return CRT_OK;
}

void crt_camera_default(crt_camera *cam)
{
    if (!cam) return;
    std::memset(cam, 0, sizeof *cam);
    cam->position[0] = 5.0f;
    cam->position[1] = 5.0f;
    cam->fov         = 75.0f;
    cam->scale       = 1.0f;
    identity16(cam->matrix);
}

// glm::rotate(m, angle, axis) = m * R(angle, axis) (Rodrigues, axis normalised)
static void rot_mul(float *m, float deg, float ax, float ay, float az)
{
    const float a = deg * 0.01745329251994329576923690768489f;
    const float c = std::cos(a), s = std::sin(a);
    const float t[3] = {(1.0f - c) * ax, (1.0f - c) * ay, (1.0f - c) * az};
    float R[9];  // column-major 3x3
    R[0] = c + t[0] * ax;
    R[1] = t[0] * ay + s * az;
    R[2] = t[0] * az - s * ay;
    R[3] = t[1] * ax - s * az;
    R[4] = c + t[1] * ay;
    R[5] = t[1] * az + s * ax;
    R[6] = t[2] * ax + s * ay;
    R[7] = t[2] * ay - s * ax;
    R[8] = c + t[2] * az;
    float out[16];
    std::memcpy(out, m, sizeof out);
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 4; ++row)
            out[col * 4 + row] = m[0 * 4 + row] * R[col * 3 + 0] + m[1 * 4 + row] * R[col * 3 + 1] + m[2 * 4 + row] * R[col * 3 + 2];
    std::memcpy(m, out, sizeof out);
}

void crt_camera_update(crt_camera *cam)
{
    if (!cam) return;
    float m[16];
    identity16(m);
    // glm::translate(m, p): column 3 += m * p
    for (int row = 0; row < 4; ++row)
        m[12 + row] = m[0 + row] * cam->position[0] + m[4 + row] * cam->position[1] + m[8 + row] * cam->position[2] + m[12 + row];
    rot_mul(m, cam->rotation[0], 0.0f, 1.0f, 0.0f);
    rot_mul(m, cam->rotation[1], 1.0f, 0.0f, 0.0f);
    rot_mul(m, cam->rotation[2], 0.0f, 0.0f, 1.0f);
    std::memcpy(cam->matrix, m, sizeof m);
}

void crt_sun_default(crt_sun *sun)
{
    if (!sun) return;
    sun->size      = 3.14159265359f / 48.0f;
    sun->intensity = 100.0f;
    const float inv = 1.0f / std::sqrt(0.8f * 0.8f + 1.0f);
    sun->direction[0] = 0.8f * inv;
    sun->direction[1] = -1.0f * inv;
    sun->direction[2] = 0.0f;
    sun->colour[0] = 1.0f;
    sun->colour[1] = 0.9f;
    sun->colour[2] = 0.7f;
}

void crt_material_default(crt_material *m)
{
    if (!m) return;
    m->type           = CRT_SMOOTH;
    m->ior            = 1.5f;
    m->roughness      = 0.5f;
    m->reflectiveness = 1.0f;
    m->emission       = 0.0f;
    for (int k = 0; k < 4; ++k) m->colour[k] = 1.0f;
    m->tex = -1;
}

int crt_scene_create(crt_context *ctx, crt_scene **out)
{
    if (!ctx || !out) return fail(CRT_ERR_INVALID_ARG, "null argument");
    auto *s = new crt_scene();
    s->ctx  = ctx;
    crt_sun_default(&s->sun);
    crt_camera_default(&s->cam);
    ctx->live++;
    *out = s;
    return CRT_OK;
}

int crt_scene_destroy(crt_scene *scene)
{
    if (!scene) return CRT_OK;
    if (scene->users) return fail(CRT_ERR_STATE, "scene still used by a renderer");
    (void)hipSetDevice(scene->ctx->device);
    free_scene_device(scene);
    scene->ctx->live--;
    delete scene;
    return CRT_OK;
}

int crt_scene_add_model(crt_scene *scene, const crt_model_data *md, uint32_t *model_id)
{
    if (!scene || !md) return fail(CRT_ERR_INVALID_ARG, "null argument");
    const uint32_t n = md->num_triangles;
    if (n && (!md->vertices || !md->vertex_indices || !md->material_indices))
        return fail(CRT_ERR_INVALID_ARG, "vertices, vertex_indices and material_indices are required");
    if (md->num_materials && !md->materials) return fail(CRT_ERR_INVALID_ARG, "null materials");
    if (md->num_textures && !md->textures) return fail(CRT_ERR_INVALID_ARG, "null textures");
    ModelHost m;
    m.name = md->name ? md->name : "";
    m.mats.assign(md->materials, md->materials + md->num_materials);
    for (uint32_t i = 0; i < md->num_materials; ++i) {
        const crt_material &mt = m.mats[i];
        if (mt.type < 0 || mt.type > 2) return fail(CRT_ERR_INVALID_ARG, "material type out of range");
        if (mt.tex >= 0 && (uint32_t)mt.tex >= md->num_textures)
            return fail(CRT_ERR_INVALID_ARG, "material texture index out of range");
    }
    for (uint32_t i = 0; i < md->num_textures; ++i) {
        const crt_image &im = md->textures[i];
        if (!im.rgba || im.width == 0 || im.height == 0) return fail(CRT_ERR_INVALID_ARG, "empty texture");
        m.tex_data.emplace_back(im.rgba, im.rgba + (size_t)im.width * im.height * 4);
        m.tex_dims.emplace_back(im.width, im.height);
    }
    // registry.cpp:13-24,62-66: expand per corner, primID = triangle order
    m.corners.resize((size_t)n * 9);
    m.mat_idx.resize(n);
    const bool has_uv = md->texture_coords && md->texture_indices && md->num_texture_coords;
    if (has_uv) m.uv.resize((size_t)n * 6);
    for (uint32_t t = 0; t < n; ++t) {
        for (int c = 0; c < 3; ++c) {
            const uint32_t vi = md->vertex_indices[3 * (size_t)t + c];
            if (vi >= md->num_vertices) return fail(CRT_ERR_INVALID_MESH, "vertex index out of range at triangle " + std::to_string(t));
            for (int k = 0; k < 3; ++k) {
                const float x = md->vertices[3 * (size_t)vi + k];
                if (!std::isfinite(x)) return fail(CRT_ERR_INVALID_MESH, "non-finite vertex at triangle " + std::to_string(t));
                m.corners[9 * (size_t)t + 3 * c + k] = x;
            }
            if (has_uv) {
                const uint32_t ti = md->texture_indices[3 * (size_t)t + c];
                if (ti >= md->num_texture_coords) return fail(CRT_ERR_INVALID_MESH, "texture index out of range at triangle " + std::to_string(t));
                m.uv[6 * (size_t)t + 2 * c + 0] = md->texture_coords[2 * (size_t)ti + 0];
                m.uv[6 * (size_t)t + 2 * c + 1] = md->texture_coords[2 * (size_t)ti + 1];
            }
        }
        const uint32_t mi = md->material_indices[t];
        if (mi >= md->num_materials) return fail(CRT_ERR_INVALID_MESH, "material index out of range at triangle " + std::to_string(t));
        m.mat_idx[t] = mi;
    }
    m.instances.resize(16);
    identity16(m.instances.data());
    scene->models.push_back(std::move(m));
    scene->dirty = true;
    if (model_id) *model_id = (uint32_t)scene->models.size() - 1;
    return CRT_OK;
}

int crt_scene_set_instances(crt_scene *scene, uint32_t model_id, const float *transforms, uint32_t count)
{
    if (!scene || (!transforms && count)) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (model_id >= scene->models.size()) return fail(CRT_ERR_INVALID_ARG, "no such model");
    for (uint32_t i = 0; i < count; ++i) {
        const float *m = transforms + 16 * (size_t)i;
        for (int k = 0; k < 16; ++k)
            if (!std::isfinite(m[k])) return fail(CRT_ERR_INVALID_ARG, "non-finite instance transform");
        // column-major affine: bottom row (m[3], m[7], m[11], m[15]) = (0, 0, 0, 1)
        if (m[3] != 0.0f || m[7] != 0.0f || m[11] != 0.0f || m[15] != 1.0f)
            return fail(CRT_ERR_INVALID_ARG, "instance transforms must be affine");
        double inv[12];
        if (!affine_inverse(m, inv)) return fail(CRT_ERR_INVALID_ARG, "singular instance transform");
    }
    scene->models[model_id].instances.assign(transforms, transforms + 16 * (size_t)count);
    scene->dirty = true;
    return CRT_OK;
}

int crt_scene_set_materials(crt_scene *scene, uint32_t model_id, const crt_material *mats, uint32_t count)
{
    if (!scene || (!mats && count)) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (model_id >= scene->models.size()) return fail(CRT_ERR_INVALID_ARG, "no such model");
    ModelHost &m = scene->models[model_id];
    for (uint32_t t = 0; t < m.mat_idx.size(); ++t)
        if (m.mat_idx[t] >= count) return fail(CRT_ERR_INVALID_ARG, "material table smaller than the indices need");
    for (uint32_t i = 0; i < count; ++i) {
        if (mats[i].type < 0 || mats[i].type > 2) return fail(CRT_ERR_INVALID_ARG, "material type out of range");
        if (mats[i].tex >= 0 && (size_t)mats[i].tex >= m.tex_data.size())
            return fail(CRT_ERR_INVALID_ARG, "material texture index out of range");
    }
    m.mats.assign(mats, mats + count);
    scene->dirty = true;
    return CRT_OK;
}

int crt_scene_set_skybox(crt_scene *scene, const crt_image *sky)
{
    if (!scene) return fail(CRT_ERR_INVALID_ARG, "null scene");
    if (!sky || !sky->rgba) {
        scene->has_sky = false;
        scene->sky.clear();
    } else {
        if (sky->width == 0 || sky->height == 0) return fail(CRT_ERR_INVALID_ARG, "empty skybox");
        scene->has_sky = true;
        scene->sky.assign(sky->rgba, sky->rgba + (size_t)sky->width * sky->height * 4);
        scene->sky_w = sky->width;
        scene->sky_h = sky->height;
    }
    scene->dirty = true;
    return CRT_OK;
}

int crt_scene_set_skybox_rotation(crt_scene *scene, float x, float y)
{
    if (!scene) return fail(CRT_ERR_INVALID_ARG, "null scene");
    scene->sky_rot[0] = x;
    scene->sky_rot[1] = y;
    scene->dev.sky_rot[0] = x;
    scene->dev.sky_rot[1] = y;
    return CRT_OK;
}

int crt_scene_set_sun_enabled(crt_scene *scene, int enabled)
{
    if (!scene) return fail(CRT_ERR_INVALID_ARG, "null scene");
    scene->sun_enabled     = enabled ? 1 : 0;
    scene->dev.sun_enabled = scene->sun_enabled;
    return CRT_OK;
}

int crt_scene_set_sun(crt_scene *scene, const crt_sun *sun)
{
    if (!scene || !sun) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (!(sun->size > 0.0f)) return fail(CRT_ERR_INVALID_ARG, "sun size must be positive");
    scene->sun = *sun;
    scene->dev.sun.size      = sun->size;
    scene->dev.sun.intensity = sun->intensity;
    scene->dev.sun.dir       = mk3(sun->direction[0], sun->direction[1], sun->direction[2]);
    scene->dev.sun.colour    = mk3(sun->colour[0], sun->colour[1], sun->colour[2]);
    return CRT_OK;
}

int crt_scene_set_camera(crt_scene *scene, const crt_camera *camera)
{
    if (!scene || !camera) return fail(CRT_ERR_INVALID_ARG, "null argument");
    scene->cam = *camera;
    return CRT_OK;
}

}  // extern "C"

namespace {

// Two-level BVH of a scene with instances (scene.cpp:83-95 x model.cpp:99-126): one BVH
// per model in object space, one over the instances' world boxes, all nodes in one pool
// (instance BVH first, root 0) and all triangles in one pool (model k at its first primID).
int build_two_level(crt_scene *scene, const float *d_corners, const std::vector<uint32_t> &model_first, uint32_t width,
                    uint32_t max_leaf, hipStream_t s, BuildResult &out, float *bmin, float *bmax, uint32_t &max_stack)
{
    const size_t nm = scene->models.size();
    std::vector<BuildResult> blas(nm);
    struct Guard {
        std::vector<BuildResult> &b;
        void *extra[2] = {nullptr, nullptr};
        ~Guard()
        {
            for (auto &x : b) {
                dfree(x.nodes);
                dfree(x.tri);
            }
            for (void *p : extra) dfree(p);
        }
    } guard{blas};
    uint32_t max_depth = 0, total_nodes = 0, bvh2 = 0;
    float    ms = 0.0f;
    std::vector<InstDev> insts;
    std::vector<float>   boxes;
    for (int k = 0; k < 3; ++k) {
        bmin[k] = 3.4e38f;
        bmax[k] = -3.4e38f;
    }
    for (size_t k = 0; k < nm; ++k) {
        const ModelHost &m = scene->models[k];
        const uint32_t   n = (uint32_t)(m.corners.size() / 9);
        if (n == 0 || m.instances.empty()) continue;
        hipError_t e = build_bvh(d_corners + 9 * (size_t)model_first[k], n, width, max_leaf, s, &blas[k],
                                 scene->ctx->opt.bvh_prim_cost);
        if (e == hipSuccess) e = tri_add_prim_base(blas[k].tri, n, model_first[k], s);
        if (blas[k].invalid) return fail(CRT_ERR_HIP, "model BVH build: the builder's consistency checks failed");
        if (e != hipSuccess) return fail(CRT_ERR_HIP, std::string("model BVH build: ") + hipGetErrorString(e));
        max_depth = std::max(max_depth, blas[k].depth);
        ms += blas[k].ms;
        bvh2 += blas[k].bvh2_nodes;
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) lo[a] = hi[a] = m.corners[a];
        for (size_t i = 0; i < m.corners.size(); i += 3)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], m.corners[i + a]);
                hi[a] = std::max(hi[a], m.corners[i + a]);
            }
        for (size_t j = 0; j < m.instances.size() / 16; ++j) {
            const float *M = &m.instances[16 * j];
            double inv[12];
            if (!affine_inverse(M, inv)) return fail(CRT_ERR_INVALID_ARG, "singular instance transform");
            InstDev d{};
            for (int q = 0; q < 12; ++q) d.inv[q] = (float)inv[q];
            d.model = (uint32_t)k;
            d.local = (uint32_t)j;
            insts.push_back(d);
            // world box of the model box's 8 corners
            float wl[3] = {3.4e38f, 3.4e38f, 3.4e38f}, wh[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
            for (int c = 0; c < 8; ++c) {
                const double p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
                for (int a = 0; a < 3; ++a) {
                    const double w = M[a] * p[0] + M[4 + a] * p[1] + M[8 + a] * p[2] + M[12 + a];
                    wl[a] = std::min(wl[a], std::nextafter((float)w, -3.4e38f));
                    wh[a] = std::max(wh[a], std::nextafter((float)w, 3.4e38f));
                }
            }
            for (int a = 0; a < 3; ++a) {
                bmin[a] = std::min(bmin[a], wl[a]);
                bmax[a] = std::max(bmax[a], wh[a]);
            }
            // a degenerate "triangle" whose bounds are the world box
            boxes.insert(boxes.end(), {wl[0], wl[1], wl[2], wh[0], wh[1], wh[2], wl[0], wl[1], wl[2]});
        }
    }
    const uint32_t ninst = (uint32_t)insts.size();
    if (ninst == 0) {
        for (int a = 0; a < 3; ++a) bmin[a] = 0.0f, bmax[a] = 1.0f;
        out = BuildResult{};
        max_stack = 0;
        scene->ninst = 0;
        return CRT_OK;
    }
    if (ninst >= (1u << 27)) return fail(CRT_ERR_UNSUPPORTED, "too many instances");
    float *d_boxes = nullptr;
    HIP_CHECK(dalloc(&d_boxes, boxes.size()));
    guard.extra[0] = d_boxes;
    HIP_CHECK(hipMemcpyAsync(d_boxes, boxes.data(), sizeof(float) * boxes.size(), hipMemcpyHostToDevice, s));
    BuildResult tl{};
    hipError_t e = build_bvh(d_boxes, ninst, width, 1, s, &tl);
    if (e != hipSuccess) return fail(CRT_ERR_HIP, std::string("instance BVH build: ") + hipGetErrorString(e));
    guard.extra[1] = tl.nodes;
    // instance of each TLAS leaf position
    std::vector<float4> tl_tri((size_t)3 * ninst);
    HIP_CHECK(hipMemcpyAsync(tl_tri.data(), tl.tri, sizeof(float4) * tl_tri.size(), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    dfree(tl.tri);
    std::vector<uint32_t> tlas_inst(ninst);
    for (uint32_t i = 0; i < ninst; ++i) {
        uint32_t id;
        std::memcpy(&id, &tl_tri[3 * (size_t)i].w, 4);
        tlas_inst[i] = id;
    }
    // pools
    total_nodes = tl.node_count;
    std::vector<uint32_t> node_base(nm, 0);
    for (size_t k = 0; k < nm; ++k)
        if (blas[k].nodes) {
            node_base[k] = total_nodes;
            total_nodes += blas[k].node_count;
        }
    size_t tri_total = 0;
    for (const auto &m : scene->models) tri_total += m.corners.size() / 9;
    void   *pool = nullptr;
    float4 *tpool = nullptr;
    HIP_CHECK(hipMalloc(&pool, (size_t)128 * total_nodes));
    scene->d_nodes = pool;
    HIP_CHECK(dalloc(&tpool, 3 * std::max<size_t>(tri_total, 1)));
    scene->d_tri = tpool;
    HIP_CHECK(hipMemcpyAsync(pool, tl.nodes, (size_t)128 * tl.node_count, hipMemcpyDeviceToDevice, s));
    for (size_t k = 0; k < nm; ++k) {
        if (!blas[k].nodes) continue;
        const uint32_t n = (uint32_t)(scene->models[k].corners.size() / 9);
        void *dst = static_cast<char *>(pool) + (size_t)128 * node_base[k];
        HIP_CHECK(hipMemcpyAsync(dst, blas[k].nodes, (size_t)128 * blas[k].node_count, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(bvh_rebase(dst, blas[k].node_count, width, node_base[k], model_first[k], s));
        HIP_CHECK(hipMemcpyAsync(tpool + 3 * (size_t)model_first[k], blas[k].tri, sizeof(float4) * 3 * n,
                                 hipMemcpyDeviceToDevice, s));
    }
    for (InstDev &d : insts) d.root = node_base[d.model];
    HIP_CHECK(dalloc(&scene->d_inst, insts.size()));
    HIP_CHECK(dalloc(&scene->d_tlas_inst, tlas_inst.size()));
    HIP_CHECK(hipMemcpyAsync(scene->d_inst, insts.data(), sizeof(InstDev) * insts.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(scene->d_tlas_inst, tlas_inst.data(), sizeof(uint32_t) * tlas_inst.size(),
                             hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
    scene->ninst = ninst;
    out            = BuildResult{};
    out.node_count = total_nodes;
    out.bvh2_nodes = bvh2 + tl.bvh2_nodes;
    out.depth      = tl.depth + max_depth;
    out.sah        = tl.sah;
    out.ms         = ms + tl.ms;
    // instance path: exit marker + leftover instance leaf + both levels' pushes
    max_stack = (width - 1) * (tl.depth + max_depth) + 3;
    return CRT_OK;
}

}  // namespace

extern "C" {

int crt_scene_commit(crt_scene *scene)
{
    if (!scene) return fail(CRT_ERR_INVALID_ARG, "null scene");
    crt_context *ctx = scene->ctx;
    HIP_CHECK(hipSetDevice(ctx->device));
    // Renderer passes (also those of a background renderer of this scene), batches and
    // queries read the scene's device buffers under the context lock: the old buffers are
    // freed and the new ones built while none of them runs.
    std::lock_guard<std::mutex> ctx_lock(ctx->mu);
    free_scene_device(scene);
    // concatenate models (scene.cpp:83-95 loops over them; one BVH over all here)
    std::vector<float>       corners, uv;
    std::vector<uint32_t>    mat_idx, model_first;
    std::vector<MaterialDev> mats;
    std::vector<TextureDev>  texs;
    std::vector<float>       texels;
    bool any_uv = false;
    for (auto &m : scene->models) any_uv |= !m.uv.empty();
    bool any_alpha = false;
    uint64_t h = 1469598103934665603ull;
    for (auto &m : scene->models) {
        const uint32_t mat_base = (uint32_t)mats.size();
        const uint32_t tex_base = (uint32_t)texs.size();
        model_first.push_back((uint32_t)mat_idx.size());
        for (size_t i = 0; i < m.tex_data.size(); ++i) {
            TextureDev td;
            td.offset = texels.size() / 4;
            td.w      = m.tex_dims[i].first;
            td.h      = m.tex_dims[i].second;
            texs.push_back(td);
            texels.insert(texels.end(), m.tex_data[i].begin(), m.tex_data[i].end());
        }
        for (auto &mt : m.mats) {
            MaterialDev d;
            d.type           = mt.type;
            d.ior            = mt.ior;
            d.roughness      = mt.roughness;
            d.reflectiveness = mt.reflectiveness;
            d.emission       = mt.emission;
            for (int k = 0; k < 4; ++k) d.colour[k] = mt.colour[k];
            d.tex = mt.tex >= 0 ? (int32_t)(tex_base + (uint32_t)mt.tex) : -1;
            if (mt.tex >= 0 || mt.colour[3] == 0.0f) any_alpha = true;
            mats.push_back(d);
        }
        corners.insert(corners.end(), m.corners.begin(), m.corners.end());
        for (uint32_t t : m.mat_idx) mat_idx.push_back(mat_base + t);
        if (any_uv) {
            if (m.uv.empty()) uv.insert(uv.end(), m.corners.size() / 9 * 6, 0.0f);
            else uv.insert(uv.end(), m.uv.begin(), m.uv.end());
        }
        h = fnv1a(h, m.corners.data(), m.corners.size() * sizeof(float));
        h = fnv1a(h, m.mat_idx.data(), m.mat_idx.size() * sizeof(uint32_t));
        h = fnv1a(h, m.mats.data(), m.mats.size() * sizeof(crt_material));
    }
    int32_t sky_tex = -1;
    if (scene->has_sky) {
        TextureDev td;
        td.offset = texels.size() / 4;
        td.w      = scene->sky_w;
        td.h      = scene->sky_h;
        sky_tex   = (int32_t)texs.size();
        texs.push_back(td);
        texels.insert(texels.end(), scene->sky.begin(), scene->sky.end());
        h = fnv1a(h, scene->sky.data(), scene->sky.size() * sizeof(float));
    }
    const uint32_t ntri = (uint32_t)mat_idx.size();
    if (corners.size() / 9 >= (1u << 27)) return fail(CRT_ERR_UNSUPPORTED, "more than 2^27 triangles");
    float bmin[3] = {0, 0, 0}, bmax[3] = {1, 1, 1};
    if (!corners.empty()) {
        for (int k = 0; k < 3; ++k) {
            bmin[k] = corners[k];
            bmax[k] = corners[k];
        }
        for (size_t i = 0; i < corners.size(); i += 3)
            for (int k = 0; k < 3; ++k) {
                bmin[k] = std::min(bmin[k], corners[i + k]);
                bmax[k] = std::max(bmax[k], corners[i + k]);
            }
    }
    hipStream_t s = ctx->stream;
    float *d_corners = nullptr;
    HIP_CHECK(dalloc(&d_corners, corners.size()));
    if (!corners.empty()) HIP_CHECK(hipMemcpyAsync(d_corners, corners.data(), sizeof(float) * corners.size(), hipMemcpyHostToDevice, s));
    const uint32_t width = ctx->opt.bvh_width ? ctx->opt.bvh_width : 8u;
    const uint32_t max_leaf = ctx->opt.max_leaf ? ctx->opt.max_leaf : 4u;
    const bool     two_level = scene_instanced(scene);
    BuildResult br{};
    uint32_t    max_stack = 0;
    if (!two_level) {
        hipError_t be = build_bvh(d_corners, ntri, width, max_leaf, s, &br, ctx->opt.bvh_prim_cost);
        dfree(d_corners);
        if (br.invalid) return fail(CRT_ERR_HIP, "BVH build: the builder's consistency checks failed");
        if (be != hipSuccess) return fail(CRT_ERR_HIP, std::string("BVH build: ") + hipGetErrorString(be));
        scene->d_nodes = br.nodes;
        scene->d_tri   = br.tri;
        max_stack      = ntri ? (width - 1) * br.depth + 1 : 0;
    } else {
        const int st2 = build_two_level(scene, d_corners, model_first, width, max_leaf, s, br, bmin, bmax, max_stack);
        dfree(d_corners);
        if (st2) {
            free_scene_device(scene);
            return st2;
        }
        for (auto &m : scene->models) h = fnv1a(h, m.instances.data(), m.instances.size() * sizeof(float));
    }
    HIP_CHECK(dalloc(&scene->d_mat_idx, mat_idx.size()));
    HIP_CHECK(dalloc(&scene->d_mats, std::max<size_t>(mats.size(), 1)));
    if (!mat_idx.empty()) HIP_CHECK(hipMemcpyAsync(scene->d_mat_idx, mat_idx.data(), sizeof(uint32_t) * mat_idx.size(), hipMemcpyHostToDevice, s));
    if (!mats.empty()) HIP_CHECK(hipMemcpyAsync(scene->d_mats, mats.data(), sizeof(MaterialDev) * mats.size(), hipMemcpyHostToDevice, s));
    if (any_uv && !uv.empty()) {
        HIP_CHECK(dalloc(&scene->d_uv, uv.size() / 2));
        HIP_CHECK(hipMemcpyAsync(scene->d_uv, uv.data(), sizeof(float) * uv.size(), hipMemcpyHostToDevice, s));
    }
    if (!texs.empty()) {
        HIP_CHECK(dalloc(&scene->d_texs, texs.size()));
        HIP_CHECK(dalloc(&scene->d_texels, texels.size() / 4));
        HIP_CHECK(hipMemcpyAsync(scene->d_texs, texs.data(), sizeof(TextureDev) * texs.size(), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(scene->d_texels, texels.data(), sizeof(float) * texels.size(), hipMemcpyHostToDevice, s));
    }
    HIP_CHECK(dalloc(&scene->d_model_first, std::max<size_t>(model_first.size(), 1)));
    if (!model_first.empty())
        HIP_CHECK(hipMemcpyAsync(scene->d_model_first, model_first.data(), sizeof(uint32_t) * model_first.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));

    if (max_stack > kSpillDepth + 8) {   // capacity of the smallest LDS stack configuration
        free_scene_device(scene);
        return fail(CRT_ERR_UNSUPPORTED, "BVH too deep for the traversal stack");
    }
    SceneDev &d = scene->dev;
    d = SceneDev{};
    d.tri         = scene->d_tri;
    d.nodes       = scene->d_nodes;
    d.nnodes      = br.node_count;
    d.width       = width;
    d.ntri        = ntri;
    d.root        = 0;
    d.mat_idx     = scene->d_mat_idx;
    d.uv          = scene->d_uv;
    d.mats        = scene->d_mats;
    d.nmat        = (uint32_t)mats.size();
    d.texels      = scene->d_texels;
    d.texs        = scene->d_texs;
    d.ntex        = (uint32_t)texs.size() - (sky_tex >= 0 ? 1u : 0u);
    d.sky_tex     = sky_tex;
    d.sky_rot[0]  = scene->sky_rot[0];
    d.sky_rot[1]  = scene->sky_rot[1];
    d.sun_enabled = scene->sun_enabled;
    d.sun.size      = scene->sun.size;
    d.sun.intensity = scene->sun.intensity;
    d.sun.dir       = mk3(scene->sun.direction[0], scene->sun.direction[1], scene->sun.direction[2]);
    d.sun.colour    = mk3(scene->sun.colour[0], scene->sun.colour[1], scene->sun.colour[2]);
    d.any_alpha   = any_alpha ? 1 : 0;
    d.model_first = scene->d_model_first;
    d.nmodels     = (uint32_t)model_first.size();
    d.bmin        = mk3(bmin[0], bmin[1], bmin[2]);
    d.bmax        = mk3(bmax[0], bmax[1], bmax[2]);
    d.instanced   = two_level ? 1 : 0;
    d.inst        = scene->d_inst;
    d.tlas_inst   = scene->d_tlas_inst;
    d.ninst       = scene->ninst;
    crt_bvh_info &bi = scene->info;
    bi.width           = width;
    bi.node_count      = br.node_count;
    bi.bvh2_node_count = br.bvh2_nodes;
    bi.depth           = br.depth;
    bi.max_stack       = max_stack;
    bi.sah_cost        = br.sah;
    bi.build_ms        = br.ms;
    bi.bytes           = (uint64_t)br.node_count * 128u + (uint64_t)ntri * 48u;
    scene->hash      = h;
    scene->committed = true;
    scene->dirty     = false;
    return ensure_spill(ctx);
}

int crt_scene_bvh_info(crt_scene *scene, crt_bvh_info *info)
{
    if (!scene || !info) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (!scene->committed) return fail(CRT_ERR_STATE, "scene not committed");
    *info = scene->info;
    return CRT_OK;
}

int crt_scene_bvh_export(crt_scene *scene, void *nodes, uint64_t *node_bytes, uint32_t *prims, uint32_t *nprims)
{
    if (!scene) return fail(CRT_ERR_INVALID_ARG, "null scene");
    if (!scene->committed) return fail(CRT_ERR_STATE, "scene not committed");
    const uint64_t nb = (uint64_t)scene->info.node_count * 128u;
    const uint32_t np = scene->dev.ntri;
    if (nodes && node_bytes && *node_bytes >= nb && nb)
        HIP_CHECK(hipMemcpy(nodes, scene->d_nodes, nb, hipMemcpyDeviceToHost));
    if (prims && nprims && *nprims >= np && np) {
        std::vector<float4> tri((size_t)np * 3);
        HIP_CHECK(hipMemcpy(tri.data(), scene->d_tri, sizeof(float4) * tri.size(), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < np; ++i) {
            uint32_t id;
            std::memcpy(&id, &tri[3 * (size_t)i].w, 4);
            prims[i] = id;
        }
    }
    if (node_bytes) *node_bytes = nb;
    if (nprims) *nprims = np;
    return CRT_OK;
}

int crt_intersect(crt_scene *scene, const crt_rays *rays, uint32_t n, crt_hits *hits, int memory)
{
    if (!hits) return fail(CRT_ERR_INVALID_ARG, "null hits");
    return query(scene, rays, n, hits, nullptr, memory);
}

int crt_intersect_noalloc(crt_scene *scene, const crt_rays *rays, uint32_t n, crt_hits *hits)
{
    if (!scene || !rays || !hits) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (!scene->committed) return fail(CRT_ERR_STATE, "scene not committed");
    if (n > kNoAllocRays)
        return fail(CRT_ERR_INVALID_ARG, "crt_intersect_noalloc takes at most " + std::to_string(kNoAllocRays) +
                                             " rays per call (use crt_intersect, or split the batch)");
    const float *in[8] = {rays->org_x, rays->org_y, rays->org_z, rays->dir_x, rays->dir_y, rays->dir_z, rays->tnear, rays->tfar};
    for (const float *p : in)
        if (!p && n) return fail(CRT_ERR_INVALID_ARG, "null ray component");
    if ((!hits->t || !hits->prim_id) && n) return fail(CRT_ERR_INVALID_ARG, "hits.t and hits.prim_id are required");
    if (n == 0) return CRT_OK;
    crt_context *ctx = scene->ctx;
    if (!ctx->noalloc) return fail(CRT_ERR_STATE, "no scratch for this device");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s   = ctx->stream;
    uint8_t    *sp  = ctx->noalloc;   // the device address of the scratch array, never its name
    uint32_t   *cnt = reinterpret_cast<uint32_t *>(sp + kNaCnt);
    RaySoA rs;
    rs.ox = const_cast<float *>(rays->org_x);
    rs.oy = const_cast<float *>(rays->org_y);
    rs.oz = const_cast<float *>(rays->org_z);
    rs.dx = const_cast<float *>(rays->dir_x);
    rs.dy = const_cast<float *>(rays->dir_y);
    rs.dz = const_cast<float *>(rays->dir_z);
    rs.tn = const_cast<float *>(rays->tnear);
    rs.tf = const_cast<float *>(rays->tfar);
    HitSoA h{};
    h.t    = hits->t;
    h.u    = hits->u ? hits->u : reinterpret_cast<float *>(sp + kNaU);
    h.v    = hits->v ? hits->v : reinterpret_cast<float *>(sp + kNaV);
    h.tri  = reinterpret_cast<uint32_t *>(sp + kNaTri);
    h.inst = reinterpret_cast<uint32_t *>(sp + kNaInst);
    HIP_CHECK(launch_noalloc_init(cnt, n, s));
    TraceQueue q = make_queue(rs, h, cnt, cnt + 8, cnt + 16, false);
    HIP_CHECK(launch_trace_closest(scene->dev, q, reinterpret_cast<uint32_t *>(sp + kNaSpill),
                                   TraceConfig{kNoAllocGrid, ctx->stack}, s));
    HIP_CHECK(launch_hits_to_api(scene->dev, h, n, hits->prim_id, hits->geom_id, hits->ng_x, hits->ng_y, hits->ng_z,
                                 hits->inst_id, s));
    uint32_t err = 0;
    HIP_CHECK(hipMemcpyAsync(&err, cnt + 16, sizeof err, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (err) return fail(CRT_ERR_HIP, "traversal stack overflow");
    return CRT_OK;
}

int crt_cast_rays(crt_scene *scene, const crt_rays *rays, uint32_t n, crt_records *records, int memory)
{
    return cast_records(scene, rays, n, records, memory);
}

int crt_occluded(crt_scene *scene, const crt_rays *rays, uint32_t n, uint8_t *occluded, int memory)
{
    if (!occluded && n) return fail(CRT_ERR_INVALID_ARG, "null occluded");
    return query(scene, rays, n, nullptr, occluded, memory);
}

static void destroy_renderer_events(crt_renderer *r)
{
    for (auto &set : r->ev)
        for (hipEvent_t e : set)
            if (e) (void)hipEventDestroy(e);
    if (r->ev_origin) (void)hipEventDestroy(r->ev_origin);
    for (int k = 0; k < 2; ++k) {
        if (r->ev_train[k]) (void)hipEventDestroy(r->ev_train[k]);
        if (r->ev_accdone[k]) (void)hipEventDestroy(r->ev_accdone[k]);
        if (r->ev_aov[k]) (void)hipEventDestroy(r->ev_aov[k]);
    }
    if (r->err_host) (void)hipHostFree(r->err_host);
    r->err_host = nullptr;
}

int crt_renderer_create(crt_context *ctx, crt_scene *scene, uint32_t res_x, uint32_t res_y, uint32_t bounces,
                        crt_renderer **out)
{
    if (!ctx || !scene || !out) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (res_x == 0 || res_y == 0) return fail(CRT_ERR_INVALID_ARG, "resolution must be positive");
    if ((uint64_t)res_x * res_y > (1ull << 30)) return fail(CRT_ERR_INVALID_ARG, "resolution too large");
    if (scene->ctx != ctx) return fail(CRT_ERR_INVALID_ARG, "scene belongs to another context");
    HIP_CHECK(hipSetDevice(ctx->device));
    auto *r    = new crt_renderer();
    r->ctx     = ctx;
    r->scene   = scene;
    r->W       = res_x;
    r->H       = res_y;
    r->bounces = bounces;
    r->opt     = crt_render_options{CRT_MODE_COMPAT, 0u, 0u, 0u};
    int st = CRT_OK;
    for (auto &set : r->ev)
        for (hipEvent_t &e : set)
            if (st == CRT_OK && hipEventCreate(&e) != hipSuccess) st = fail(CRT_ERR_HIP, "hipEventCreate");
    if (st == CRT_OK && hipEventCreate(&r->ev_origin) != hipSuccess) st = fail(CRT_ERR_HIP, "hipEventCreate");
    for (int k = 0; k < 2; ++k) {
        if (st == CRT_OK && hipEventCreate(&r->ev_train[k]) != hipSuccess) st = fail(CRT_ERR_HIP, "hipEventCreate");
        if (st == CRT_OK && hipEventCreateWithFlags(&r->ev_accdone[k], hipEventDisableTiming) != hipSuccess)
            st = fail(CRT_ERR_HIP, "hipEventCreate");
        if (st == CRT_OK && hipEventCreateWithFlags(&r->ev_aov[k], hipEventDisableTiming) != hipSuccess)
            st = fail(CRT_ERR_HIP, "hipEventCreate");
    }
    if (st == CRT_OK && hipHostMalloc((void **)&r->err_host, 8 * sizeof(uint32_t)) != hipSuccess)
        st = fail(CRT_ERR_OUT_OF_MEMORY, "pinned host allocation");
    if (st == CRT_OK) st = alloc_renderer_buffers(r);
    if (st != CRT_OK) {
        free_renderer_buffers(r);
        destroy_renderer_events(r);
        delete r;
        return st;
    }
    scene->users++;
    ctx->live++;
    *out = r;
    return CRT_OK;
}

int crt_renderer_destroy(crt_renderer *r)
{
    if (!r) return CRT_OK;
    r->stop.store(true);
    join_worker(r);
    (void)hipSetDevice(r->ctx->device);
    free_renderer_buffers(r);
    destroy_renderer_events(r);
    r->scene->users--;
    if (r->nrc) r->nrc->users--;
    r->ctx->live--;
    delete r;
    return CRT_OK;
}

int crt_renderer_start(crt_renderer *r)
{
    int st = check_ready(r);
    if (st) return st;
    {
        std::lock_guard<std::mutex> g(r->mu);
        if (r->running) return CRT_OK;  // renderer.cpp:156: start() only acts when paused
    }
    join_worker(r);
    if ((st = reset_accum(r))) return st;
    r->stop.store(false);
    r->worker_status = CRT_OK;
    {
        std::lock_guard<std::mutex> g(r->mu);
        r->running = true;
    }
    r->worker = std::thread(worker_loop, r);
    return CRT_OK;
}

int crt_renderer_pause(crt_renderer *r)
{
    if (!r) return fail(CRT_ERR_INVALID_ARG, "null renderer");
    r->stop.store(true);
    join_worker(r);
    if (r->worker_status != CRT_OK) return fail(r->worker_status, r->worker_error);
    return CRT_OK;
}

int crt_renderer_cancel(crt_renderer *r)
{
    if (!r) return fail(CRT_ERR_INVALID_ARG, "null renderer");
    r->stop.store(true);
    return CRT_OK;
}

int crt_renderer_is_running(crt_renderer *r, int *running)
{
    if (!r || !running) return fail(CRT_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(r->mu);
    *running = r->running ? 1 : 0;
    return CRT_OK;
}

int crt_renderer_wait(crt_renderer *r)
{
    if (!r) return fail(CRT_ERR_INVALID_ARG, "null renderer");
    if (r->target == 0) {
        std::lock_guard<std::mutex> g(r->mu);
        if (r->running) return fail(CRT_ERR_STATE, "no target spp set: wait would never return");
    }
    join_worker(r);
    if (r->worker_status != CRT_OK) return fail(r->worker_status, r->worker_error);
    return CRT_OK;
}

static int require_paused(crt_renderer *r)
{
    if (!r) return fail(CRT_ERR_INVALID_ARG, "null renderer");
    std::lock_guard<std::mutex> g(r->mu);
    if (r->running) return fail(CRT_ERR_STATE, "renderer is running; pause it first");
    return CRT_OK;
}

int crt_renderer_set_resolution(crt_renderer *r, uint32_t x, uint32_t y)
{
    int st = require_paused(r);
    if (st) return st;
    if (x == 0 || y == 0) return fail(CRT_ERR_INVALID_ARG, "resolution must be positive");
    join_worker(r);
    r->W      = x;
    r->H      = y;
    r->aspect = (float)x / (float)y;   // renderer.cpp:199
    if ((st = alloc_renderer_buffers(r))) return st;
    r->samples.store(0);
    return CRT_OK;
}

int crt_renderer_set_max_bounces(crt_renderer *r, uint32_t bounces)
{
    int st = require_paused(r);
    if (st) return st;
    r->bounces = bounces;
    return CRT_OK;
}

int crt_renderer_set_target_spp(crt_renderer *r, uint64_t target)
{
    if (!r) return fail(CRT_ERR_INVALID_ARG, "null renderer");
    r->target = target;
    return CRT_OK;
}

int crt_renderer_set_options(crt_renderer *r, const crt_render_options *opt)
{
    int st = require_paused(r);
    if (st) return st;
    if (!opt) return fail(CRT_ERR_INVALID_ARG, "null options");
    if (opt->mode != CRT_MODE_COMPAT && opt->mode != CRT_MODE_PBR) return fail(CRT_ERR_INVALID_ARG, "bad mode");
    r->opt = *opt;
    return CRT_OK;
}

int crt_renderer_get_options(crt_renderer *r, crt_render_options *opt)
{
    if (!r || !opt) return fail(CRT_ERR_INVALID_ARG, "null argument");
    *opt = r->opt;
    return CRT_OK;
}

int crt_renderer_render(crt_renderer *r, uint32_t passes)
{
    int st = check_ready(r);
    if (st) return st;
    if ((st = require_paused(r))) return st;
    join_worker(r);
    HIP_CHECK(hipSetDevice(r->ctx->device));
    r->stop.store(false);
    for (uint32_t i = 0; i < passes;) {
        if (r->stop.load()) return fail(CRT_ERR_CANCELLED, "cancelled");
        uint32_t done = 0;
        if ((st = render_passes(r, std::min(kPassGroup * r->kmax, passes - i), done))) return st;
        i += done;
    }
    return CRT_OK;
}

int crt_renderer_reset(crt_renderer *r)
{
    int st = require_paused(r);
    if (st) return st;
    join_worker(r);
    return reset_accum(r);
}

int crt_renderer_current_sample_count(crt_renderer *r, uint64_t *count)
{
    if (!r || !count) return fail(CRT_ERR_INVALID_ARG, "null argument");
    *count = r->samples.load();
    return CRT_OK;
}

int crt_renderer_current_resolution(crt_renderer *r, uint32_t *x, uint32_t *y)
{
    if (!r || !x || !y) return fail(CRT_ERR_INVALID_ARG, "null argument");
    *x = r->W;
    *y = r->H;
    return CRT_OK;
}

static int buffer_ptr(crt_renderer *r, int buffer, void **ptr, size_t *bytes, bool refresh)
{
    const size_t P = r->b.npix;
    switch (buffer) {
    case CRT_BUFFER_PROGRESS:
        if (refresh) {
            HIP_CHECK(launch_display(r->b.raw, r->samples.load(), r->b.npix, r->display, r->ctx->stream));
        }
        *ptr   = r->display;
        *bytes = P * 16;
        return CRT_OK;
    case CRT_BUFFER_NORMALS: *ptr = r->b.aov_nrm; *bytes = P * 16; return CRT_OK;
    case CRT_BUFFER_ALBEDOS: *ptr = r->b.aov_alb; *bytes = P * 16; return CRT_OK;
    case CRT_BUFFER_DEPTHS: *ptr = r->b.aov_dep; *bytes = P * 16; return CRT_OK;
    case CRT_BUFFER_RAW: *ptr = r->b.raw; *bytes = P * 12; return CRT_OK;
    default: return fail(CRT_ERR_INVALID_ARG, "unknown buffer");
    }
}

int crt_renderer_read(crt_renderer *r, int buffer, float *out)
{
    if (!r || !out) return fail(CRT_ERR_INVALID_ARG, "null argument");
    HIP_CHECK(hipSetDevice(r->ctx->device));
    void  *p = nullptr;
    size_t bytes = 0;
    int    st = buffer_ptr(r, buffer, &p, &bytes, true);
    if (st) return st;
    HIP_CHECK(hipMemcpyAsync(out, p, bytes, hipMemcpyDeviceToHost, r->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(r->ctx->stream));
    return CRT_OK;
}

int crt_renderer_device_buffer(crt_renderer *r, int buffer, void **dev_ptr)
{
    if (!r || !dev_ptr) return fail(CRT_ERR_INVALID_ARG, "null argument");
    size_t bytes = 0;
    int    st = buffer_ptr(r, buffer, dev_ptr, &bytes, true);
    if (st) return st;
    HIP_CHECK(hipStreamSynchronize(r->ctx->stream));
    return CRT_OK;
}

int crt_renderer_current_stats(crt_renderer *r, crt_stats *stats)
{
    if (!r || !stats) return fail(CRT_ERR_INVALID_ARG, "null argument");
    std::vector<unsigned long long> slots(4 * kStatSlots);
    unsigned long long s[4] = {0, 0, 0, 0};
    for (const PathBuffers *b : {&r->b, &r->b2}) {
        if (!b->stats) continue;
        HIP_CHECK(hipMemcpy(slots.data(), b->stats, sizeof(unsigned long long) * slots.size(), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < kStatSlots; ++i)
            for (int k = 0; k < 4; ++k) s[k] += slots[4 * i + k];
    }
    std::memset(stats, 0, sizeof *stats);
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - r->t0).count();
    stats->running_time       = t;
    stats->samples            = r->samples.load();
    stats->total_rays         = s[2];
    stats->extension_rays     = s[0];
    stats->shadow_rays        = s[1];
    stats->rays_per_second    = t > 0 ? (uint64_t)((double)s[2] / t) : 0;
    stats->samples_per_second = t > 0 ? (uint64_t)((double)stats->samples / t) : 0;
    stats->ms_generate   = r->ms_phase[0];
    stats->ms_extend     = r->ms_phase[1];
    stats->ms_shade      = r->ms_phase[2];
    stats->ms_shadow     = r->ms_phase[3];
    stats->ms_accumulate = r->ms_phase[4];
    stats->ms_render     = r->ms_phase[5];
    stats->extend_launches = r->extend_launches;
    stats->ms_extend_busy  = r->ms_extend_busy;
    if (r->nrc) harvest_train_time(r, true);
    stats->ms_nrc_train    = r->ms_nrc_train;
    stats->nrc_detached    = r->nrc_failed ? 1u : 0u;
    stats->nrc_discarded_passes = r->nrc_discarded;
    if (r->b.tstats) {
        HIP_CHECK(hipStreamSynchronize(r->ctx->stream));
        HIP_CHECK(hipStreamSynchronize(r->ctx->stream2));
        for (const PathBuffers *b : {&r->b, &r->b2}) {
            if (!b->tstats) continue;
            unsigned long long ts[8];
            HIP_CHECK(hipMemcpy(ts, b->tstats, sizeof ts, hipMemcpyDeviceToHost));
            stats->node_visits     += ts[0];
            stats->leaf_visits     += ts[1];
            stats->tri_tests       += ts[2];
            stats->wave_node_steps += ts[3];
            stats->wave_leaf_steps += ts[4];
            stats->node_idle_no_ray    += ts[5];
            stats->node_idle_done      += ts[6];
            stats->node_idle_leaf_wait += ts[7];
        }
    }
    if (r->nrc && r->loss_tiles) {
        if (int wst = nrc_wait_training(r->nrc)) return wst;
        std::lock_guard<std::mutex> g(r->nrc->mu);
        std::vector<float> pl(r->loss_tiles);
        HIP_CHECK(hipStreamSynchronize(r->ctx->stream));
        HIP_CHECK(hipMemcpy(pl.data(), r->nrc->partial_loss, sizeof(float) * pl.size(), hipMemcpyDeviceToHost));
        double acc = 0.0;
        for (float x : pl) acc += x;
        stats->nrc_loss = (float)acc;
    }
    return CRT_OK;
}

struct CkptHeader {
    char     magic[8];
    uint32_t version, W, H, bounces, mode, rr_start, seed, pad;
    uint64_t samples, sample_offset, scene_hash;
};

int crt_renderer_checkpoint_size(crt_renderer *r, uint64_t *bytes)
{
    if (!r || !bytes) return fail(CRT_ERR_INVALID_ARG, "null argument");
    *bytes = sizeof(CkptHeader) + (uint64_t)r->b.npix * 3 * sizeof(float);
    return CRT_OK;
}

int crt_renderer_checkpoint_save(crt_renderer *r, void *buf, uint64_t bytes)
{
    int st = require_paused(r);
    if (st) return st;
    uint64_t need = 0;
    crt_renderer_checkpoint_size(r, &need);
    if (!buf || bytes < need) return fail(CRT_ERR_INVALID_ARG, "checkpoint buffer too small");
    CkptHeader h{};
    std::memcpy(h.magic, "CRTCKPT1", 8);
    h.version       = 1;
    h.W             = r->W;
    h.H             = r->H;
    h.bounces       = r->bounces;
    h.mode          = r->opt.mode;
    h.rr_start      = r->opt.rr_start;
    h.seed          = r->opt.seed;
    h.samples       = r->samples.load();
    h.sample_offset = r->opt.sample_offset;
    h.scene_hash    = r->scene->hash;
    std::memcpy(buf, &h, sizeof h);
    HIP_CHECK(hipMemcpy(static_cast<char *>(buf) + sizeof h, r->b.raw, (size_t)r->b.npix * 12, hipMemcpyDeviceToHost));
    return CRT_OK;
}

int crt_renderer_checkpoint_load(crt_renderer *r, const void *buf, uint64_t bytes)
{
    int st = require_paused(r);
    if (st) return st;
    if (!buf || bytes < sizeof(CkptHeader)) return fail(CRT_ERR_CHECKPOINT, "checkpoint truncated");
    CkptHeader h;
    std::memcpy(&h, buf, sizeof h);
    if (std::memcmp(h.magic, "CRTCKPT1", 8) != 0 || h.version != 1) return fail(CRT_ERR_CHECKPOINT, "not a crt checkpoint");
    if (h.W != r->W || h.H != r->H) return fail(CRT_ERR_CHECKPOINT, "checkpoint resolution differs");
    if (h.scene_hash != r->scene->hash) return fail(CRT_ERR_CHECKPOINT, "checkpoint scene differs");
    if (bytes < sizeof h + (uint64_t)r->b.npix * 12) return fail(CRT_ERR_CHECKPOINT, "checkpoint truncated");
    if ((st = reset_accum(r))) return st;
    HIP_CHECK(hipMemcpy(r->b.raw, static_cast<const char *>(buf) + sizeof h, (size_t)r->b.npix * 12, hipMemcpyHostToDevice));
    r->bounces           = h.bounces;
    r->opt.mode          = h.mode;
    r->opt.rr_start      = h.rr_start;
    r->opt.seed          = h.seed;
    r->opt.sample_offset = (uint32_t)h.sample_offset;
    r->samples.store(h.samples);
    return CRT_OK;
}

}  // extern "C"

// ---- neural radiance cache --------------------------------------------------------------

extern "C" void crt_nrc_render_options_default(crt_nrc_render_options *o)
{
    if (!o) return;
    o->term_bounce  = 1;
    o->footprint_c  = 0.01f;
    o->train_paths  = 65536;
    o->train_suffix = 4;
    o->train_steps  = 24;
    o->train        = 1;
    o->warmup_passes = 0;
}

// Attaching or detaching the cache re-allocates the path buffers (the training paths share
// the wavefront) and so clears the accumulation, like renderer::update (renderer.cpp:185-192).
extern "C" int crt_renderer_set_nrc(crt_renderer *r, crt_nrc *nrc, const crt_nrc_render_options *opt)
{
    int st = require_paused(r);
    if (st) return st;
    if (nrc && nrc->ctx != r->ctx) return fail(CRT_ERR_INVALID_ARG, "cache belongs to another context");
    crt_nrc_render_options o;
    crt_nrc_render_options_default(&o);
    if (opt) o = *opt;
    if (nrc && (o.term_bounce < 1 || o.train_suffix < 1 || o.train_suffix > 16 || o.train_paths < 64 ||
                o.train_paths > (1u << 22)))
        return fail(CRT_ERR_INVALID_ARG, "bad cache render options");
    join_worker(r);
    if (r->nrc) r->nrc->users--;
    if (nrc) nrc->users++;
    r->nrc           = nrc;
    r->nopt          = o;
    r->nrc_warm_done = 0;
    r->loss_tiles    = 0;
    r->nrc_failed    = false;
    r->nrc_discarded = 0;
    r->train_timed   = false;
    if ((st = alloc_renderer_buffers(r))) return st;
    r->samples.store(0);
    return reset_accum(r);
}




namespace {

uint32_t nrc_tiles_alloc(uint32_t max_batch) { return ((max_batch + 63) / 64 + 3) / 4 * 4; }

int nrc_stage(crt_nrc *n, uint32_t count)
{
    if (count <= n->stage_cap) return CRT_OK;
    dfree(n->q_stage);
    dfree(n->t_stage);
    dfree(n->o_stage);
    n->q_stage = n->t_stage = n->o_stage = nullptr;
    HIP_CHECK(dalloc(&n->q_stage, (size_t)count * kNrcQuery));
    HIP_CHECK(dalloc(&n->t_stage, (size_t)count * 3));
    HIP_CHECK(dalloc(&n->o_stage, (size_t)count * 3));
    n->stage_cap = count;
    return CRT_OK;
}

// Kaiming-uniform initialisation from Philox, zero for padded inputs and outputs.
void nrc_init_params(std::vector<float> &w, uint32_t seed)
{
    w.assign(kNrcParams, 0.0f);
    uint32_t ctr = 0;
    for (int l = 0; l < kNrcLayers; ++l) {
        const int K = l == 0 ? kNrcIn : kNrcHid, N = l == kNrcLayers - 1 ? kNrcOut : kNrcHid;
        const int fan_in = l == 0 ? 83 : K;
        const float bound = std::sqrt((l == kNrcLayers - 1 ? 3.0f : 6.0f) / (float)fan_in);
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < N; ++c) {
                const u4 r = philox4x32_10(u4{ctr++, 0u, 0xFFFFFFF0u, 0u}, seed, kKeyHi);
                float x = (2.0f * u01(r.x) - 1.0f) * bound;
                if ((l == 0 && k >= 83) || (l == kNrcLayers - 1 && c >= 3)) x = 0.0f;
                w[nrc_w_off(l) + k * N + c] = x;
            }
    }
}

int nrc_upload(crt_nrc *n, const float *host)
{
    hipStream_t s = n->ctx->stream;
    HIP_CHECK(hipMemcpyAsync(n->params, host, sizeof(float) * kNrcParams, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemsetAsync(n->m, 0, sizeof(float) * kNrcParams, s));
    HIP_CHECK(hipMemsetAsync(n->v, 0, sizeof(float) * kNrcParams, s));
    HIP_CHECK(nrc_launch_adam(nullptr, 0, n->params, n->m, n->v, n->W, n->WT, 0, 0, 0, 0, 1, 0, s));
    HIP_CHECK(hipStreamSynchronize(s));
    n->step = 0;
    return CRT_OK;
}

const float *nrc_input(crt_nrc *n, const float *p, size_t count, float *stage, int memory, int *st)
{
    *st = CRT_OK;
    if (memory == CRT_MEM_DEVICE) return p;
    hipError_t e = hipMemcpyAsync(stage, p, sizeof(float) * count, hipMemcpyHostToDevice, n->ctx->stream);
    if (e != hipSuccess) *st = fail(CRT_ERR_HIP, hipGetErrorString(e));
    return stage;
}

}  // namespace

extern "C" {

void crt_nrc_options_default(crt_nrc_options *opt)
{
    if (!opt) return;
    opt->max_batch = 65536;
    opt->seed      = 1;
    opt->lr        = 1e-3f;
    opt->beta1     = 0.9f;
    opt->beta2     = 0.99f;
    opt->eps       = 1e-10f;
}

int crt_nrc_param_count(void) { return kNrcParams; }

int crt_nrc_create(crt_context *ctx, const crt_nrc_options *opt, crt_nrc **out)
{
    if (!ctx || !out) return fail(CRT_ERR_INVALID_ARG, "null argument");
    crt_nrc_options o;
    crt_nrc_options_default(&o);
    if (opt) o = *opt;
    if (o.max_batch == 0 || o.max_batch > (1u << 24)) return fail(CRT_ERR_INVALID_ARG, "max_batch must be in [1, 2^24]");
    HIP_CHECK(hipSetDevice(ctx->device));
    auto *n = new crt_nrc();
    n->ctx = ctx;
    n->opt = o;
    ctx->live++;
    int st = CRT_OK;
    if (hipStreamCreateWithFlags(&n->ts, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&n->ev_targets, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&n->ev_trained, hipEventDisableTiming) != hipSuccess)
        st = fail(CRT_ERR_HIP, "cache stream / events");
    n->tiles_alloc = nrc_tiles_alloc(n->opt.max_batch);
    auto A = [&](auto **p, size_t c) {
        if (st == CRT_OK && dalloc(p, c) != hipSuccess) st = fail(CRT_ERR_OUT_OF_MEMORY, "nrc allocation");
    };
    A(&n->params, kNrcParams);
    A(&n->m, kNrcParams);
    A(&n->v, kNrcParams);
    A(&n->W, kNrcParams);
    A(&n->WT, kNrcParams);
    A(&n->grad, kNrcParams);
    A(&n->act, (size_t)n->tiles_alloc * 64 * kNrcActs);
    A(&n->partial, (size_t)n->tiles_alloc * kNrcParams);
    A(&n->partial_loss, n->tiles_alloc);
    if (st == CRT_OK) {
        std::vector<float> w;
        nrc_init_params(w, n->opt.seed);
        st = nrc_upload(n, w.data());
    }
    if (st != CRT_OK) {
        crt_nrc_destroy(n);
        return st;
    }
    *out = n;
    return CRT_OK;
}

int crt_nrc_destroy(crt_nrc *n)
{
    if (!n) return CRT_OK;
    if (n->users) return fail(CRT_ERR_STATE, "cache still attached to a renderer");
    (void)hipSetDevice(n->ctx->device);
    if (n->ts) (void)hipStreamSynchronize(n->ts);
    if (n->ev_targets) (void)hipEventDestroy(n->ev_targets);
    if (n->ev_trained) (void)hipEventDestroy(n->ev_trained);
    if (n->ts) (void)hipStreamDestroy(n->ts);
    void *ps[] = {n->params, n->m, n->v, n->W, n->WT, n->grad, n->act, n->partial, n->partial_loss,
                  n->q_stage, n->t_stage, n->o_stage};
    for (void *p : ps) dfree(p);
    n->ctx->live--;
    delete n;
    return CRT_OK;
}

int crt_nrc_get_params(crt_nrc *n, float *params)
{
    if (!n || !params) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (int wst = nrc_wait_training(n)) return wst;   // training enqueued by renderers
    HIP_CHECK(hipMemcpy(params, n->params, sizeof(float) * kNrcParams, hipMemcpyDeviceToHost));
    return CRT_OK;
}

int crt_nrc_set_params(crt_nrc *n, const float *params)
{
    if (!n || !params) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (int wst = nrc_wait_training(n)) return wst;   // training enqueued by renderers
    std::lock_guard<std::mutex> g(n->mu);
    return nrc_upload(n, params);
}

int crt_nrc_infer(crt_nrc *n, const float *queries, uint32_t count, float *out, int memory)
{
    if (!n || ((!queries || !out) && count)) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (int wst = nrc_wait_training(n)) return wst;   // training enqueued by renderers
    if (count == 0) return CRT_OK;
    std::lock_guard<std::mutex> g(n->mu);
    int st = memory == CRT_MEM_HOST ? nrc_stage(n, count) : CRT_OK;
    if (st) return st;
    const float *q = nrc_input(n, queries, (size_t)count * kNrcQuery, n->q_stage, memory, &st);
    if (st) return st;
    float *o = memory == CRT_MEM_DEVICE ? out : n->o_stage;
    hipStream_t s = n->ctx->stream;
    HIP_CHECK(nrc_launch_infer(q, o, count, n->WT, s));
    if (memory == CRT_MEM_HOST) HIP_CHECK(hipMemcpyAsync(out, o, sizeof(float) * 3 * count, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return CRT_OK;
}

static int nrc_fwd_bwd(crt_nrc *n, const float *queries, const float *targets, uint32_t count, int memory,
                       float *loss)
{
    if (count > n->opt.max_batch) return fail(CRT_ERR_INVALID_ARG, "batch larger than max_batch");
    if (int wst = nrc_wait_locked(n)) return wst;   // training enqueued by renderers (n->mu is held)
    int st = memory == CRT_MEM_HOST ? nrc_stage(n, count) : CRT_OK;
    if (st) return st;
    const float *q = nrc_input(n, queries, (size_t)count * kNrcQuery, n->q_stage, memory, &st);
    if (st) return st;
    const float *t = nrc_input(n, targets, (size_t)count * 3, n->t_stage, memory, &st);
    if (st) return st;
    hipStream_t s = n->ctx->stream;
    HIP_CHECK(nrc_launch_train(q, t, count, n->W, n->WT, n->act, n->partial, n->partial_loss, s));
    if (loss) {
        const uint32_t tiles = (count + 63) / 64;
        std::vector<float> pl(tiles);
        HIP_CHECK(hipMemcpyAsync(pl.data(), n->partial_loss, sizeof(float) * tiles, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        double acc = 0.0;
        for (float x : pl) acc += x;
        *loss = (float)acc;
    }
    return CRT_OK;
}

int crt_nrc_gradient(crt_nrc *n, const float *queries, const float *targets, uint32_t count, float *grad, float *loss,
                     int memory)
{
    if (!n || !grad || ((!queries || !targets) && count)) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (count == 0) return fail(CRT_ERR_INVALID_ARG, "empty batch");
    std::lock_guard<std::mutex> g(n->mu);
    int st = nrc_fwd_bwd(n, queries, targets, count, memory, loss);
    if (st) return st;
    HIP_CHECK(nrc_launch_sum_partials(n->partial, nrc_partial_rows(count), n->grad, n->ctx->stream));
    HIP_CHECK(hipMemcpyAsync(grad, n->grad, sizeof(float) * kNrcParams, hipMemcpyDeviceToHost, n->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(n->ctx->stream));
    return CRT_OK;
}

int crt_nrc_train_step(crt_nrc *n, const float *queries, const float *targets, uint32_t count, float *loss, int memory)
{
    if (!n || ((!queries || !targets) && count)) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (int wst = nrc_wait_training(n)) return wst;   // training enqueued by renderers
    if (count == 0) return CRT_OK;
    std::lock_guard<std::mutex> g(n->mu);
    int st = nrc_fwd_bwd(n, queries, targets, count, memory, loss);
    if (st) return st;
    n->step++;
    HIP_CHECK(nrc_launch_adam(n->partial, nrc_partial_rows(count), n->params, n->m, n->v, n->W, n->WT, n->opt.lr,
                              n->opt.beta1, n->opt.beta2, n->opt.eps, n->step, 1, n->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(n->ctx->stream));
    return CRT_OK;
}

}  // extern "C"

// ---- batched rendering ------------------------------------------------------------------

namespace {

struct BatchPlan {
    std::vector<BatchSeg>   segs;
    std::vector<BatchChunk> chunks;
    std::vector<uint32_t>   job_last;   // last chunk holding a pass of each job
};

// Packs every job's passes into chunks of at most P path slots.  A job whose image fits
// takes as many whole passes as fit into the current chunk (one segment, sample-major);
// a larger image is split by pixel ranges one pass at a time.  No chunk holds the same
// pixel of a job twice, so accumulation needs no atomics and keeps pass order.
void plan_batch(const crt_job *jobs, uint32_t njobs, uint32_t P, BatchPlan &plan)
{
    plan.job_last.assign(njobs, 0);
    BatchChunk cur{0, 0, 0, 0};
    auto close = [&]() {
        if (cur.nseg == 0) return;
        plan.chunks.push_back(cur);
        cur = BatchChunk{(uint32_t)plan.segs.size(), 0, 0, 0};
    };
    auto add = [&](uint32_t j, uint32_t s0, uint32_t ns, uint32_t pix0, uint32_t np) {
        BatchSeg sg{};
        sg.job   = j;
        sg.s0    = s0;
        sg.nsamp = ns;
        sg.pix0  = pix0;
        sg.npix  = np;
        sg.path0 = cur.npaths;
        sg.acc0  = cur.nacc;
        plan.segs.push_back(sg);
        cur.nseg++;
        cur.npaths += ns * np;
        cur.nacc += np;
        plan.job_last[j] = (uint32_t)plan.chunks.size();
    };
    for (uint32_t j = 0; j < njobs; ++j) {
        const uint32_t npix = jobs[j].width * jobs[j].height, spp = jobs[j].spp;
        if (npix <= P) {
            uint32_t s = 0;
            while (s < spp) {
                const uint32_t k = std::min(spp - s, (P - cur.npaths) / npix);
                if (k == 0) {
                    close();
                    continue;
                }
                add(j, s, k, 0, npix);
                s += k;
            }
        } else {
            for (uint32_t s = 0; s < spp; ++s) {
                uint32_t p = 0;
                while (p < npix) {
                    const uint32_t c = std::min(npix - p, P - cur.npaths);
                    if (c == 0) {
                        close();
                        continue;
                    }
                    add(j, s, 1, p, c);
                    p += c;
                }
            }
        }
    }
    close();
}

// Two sets of path buffers replay alternate chunks on two streams (set k renders chunks k,
// k + 2, ...), so one chunk's traversal overlaps the other's; each chunk's accumulation
// waits for the previous chunk's, so the job sums are added in chunk order.  Per set: a
// stream, a chunk cursor, a spill area and, with graphs, one graph for the traversal part of
// a replay and one for its accumulation.
struct BatchRun {
    std::vector<void *> allocs;
    hipStream_t         stream[2] = {nullptr, nullptr};
    hipGraph_t          graph[2][2] = {};      // [set][0: generate + bounces, 1: accumulate + advance]
    hipGraphExec_t      exec[2][2]  = {};
    std::vector<hipEvent_t> events;            // [0] start, [c + 1] chunk c accumulated
    ~BatchRun()
    {
        for (hipStream_t s : stream)
            if (s) (void)hipStreamSynchronize(s);
        for (auto &row : exec)
            for (hipGraphExec_t e : row)
                if (e) (void)hipGraphExecDestroy(e);
        for (auto &row : graph)
            for (hipGraph_t g : row)
                if (g) (void)hipGraphDestroy(g);
        for (hipEvent_t e : events)
            if (e) (void)hipEventDestroy(e);
        for (void *p : allocs) dfree(p);
        for (hipStream_t s : stream)
            if (s) (void)hipStreamDestroy(s);
    }
};

// The traversal part of one replay: generate the chunk's paths, trace / shade / shadow for
// max_bounces bounces.  Only kernel launches (capturable) unless some job's scene has alpha
// materials under the sun.
int record_trace(const PathBuffers &b, const BatchDev &bd, const TraceScene *tscenes, uint32_t width,
                 uint32_t max_bounces, bool any_sun, bool any_alpha, uint32_t *spill, const TraceConfig &tc,
                 hipStream_t s)
{
    HIP_CHECK(launch_batch_generate(bd, b, s));
    uint32_t cur = 0;
    for (uint32_t i = 0; i < max_bounces; ++i) {
        HIP_CHECK(launch_bounce_prep(b, cur, s));
        TraceQueue q = make_queue(b.q[cur], b.hits, b.cnt + cur, b.cnt + kCntWorkExt, b.cnt + kCntErr);
        q.ray_path = b.qpath[cur];
        q.path_job = b.pjob;
        q.scenes   = tscenes;
        HIP_CHECK(launch_trace_closest_multi(width, q, spill, tc, s));
        HIP_CHECK(launch_shade_multi(bd, b, i, cur, s));
        if (any_sun) {
            HIP_CHECK(launch_shadow_prep(b, s));
            if (!any_alpha) {
                TraceQueue qs = make_queue(b.sq, HitSoA{}, b.cnt + kCntShadow, b.cnt + kCntWorkShadow, b.cnt + kCntErr);
                qs.ray_path = b.spath;
                qs.path_job = b.pjob;
                qs.scenes   = tscenes;
                HIP_CHECK(launch_trace_any_multi(width, qs, b.occ, spill, tc, s));
                HIP_CHECK(launch_shadow_resolve(b, s));
            } else {
                // alpha pass-through (renderer.cpp:330-345): host-driven, never captured
                for (int step = 0; step < 64; ++step) {
                    TraceQueue qs = make_queue(b.sq, b.shits, b.cnt + kCntShadow, b.cnt + kCntWorkShadow, b.cnt + kCntErr);
                    qs.ray_path = b.spath;
                    qs.path_job = b.pjob;
                    qs.scenes   = tscenes;
                    HIP_CHECK(launch_trace_closest_multi(width, qs, spill, tc, s));
                    HIP_CHECK(launch_shadow_alpha_step_multi(bd, b, s));
                    uint32_t left = 0;
                    HIP_CHECK(hipMemcpyAsync(&left, b.cnt + kCntShadow, sizeof left, hipMemcpyDeviceToHost, s));
                    HIP_CHECK(hipStreamSynchronize(s));
                    if (left == 0) break;
                }
            }
        }
        cur ^= 1u;
    }
    return CRT_OK;
}

// The accumulation part of one replay: add the chunk's radiance to the job sums and move the
// set's cursor to its next chunk.
int record_accumulate(const PathBuffers &b, const BatchDev &bd, uint32_t *cursor, uint32_t step, hipStream_t s)
{
    HIP_CHECK(launch_batch_accumulate(bd, b, b.P, s));
    HIP_CHECK(launch_batch_advance(cursor, step, s));
    return CRT_OK;
}

}  // namespace

extern "C" {

void crt_batch_options_default(crt_batch_options *opt)
{
    if (!opt) return;
    opt->max_paths = 0;
    opt->use_graph = 1;
}

int crt_render_batch(crt_context *ctx, const crt_job *jobs, uint32_t njobs, const crt_batch_options *opt,
                     double *job_ms, crt_batch_stats *stats)
{
    if (!ctx || (!jobs && njobs)) return fail(CRT_ERR_INVALID_ARG, "null argument");
    if (njobs == 0) return CRT_OK;
    if (njobs > (1u << 20)) return fail(CRT_ERR_INVALID_ARG, "too many jobs");
    crt_batch_options o;
    crt_batch_options_default(&o);
    if (opt) o = *opt;
    uint32_t width = 0, max_bounces = 0;
    bool any_sun = false, any_alpha = false;
    uint64_t work = 0;
    for (uint32_t j = 0; j < njobs; ++j) {
        const crt_job &J = jobs[j];
        if (!J.scene || !J.rgb_sum) return fail(CRT_ERR_INVALID_ARG, "job without scene or output");
        if (J.scene->ctx != ctx) return fail(CRT_ERR_INVALID_ARG, "job scene belongs to another context");
        if (!J.scene->committed) return fail(CRT_ERR_STATE, "job scene not committed");
        if (J.width == 0 || J.height == 0 || J.spp == 0 || J.max_bounces == 0)
            return fail(CRT_ERR_INVALID_ARG, "job needs width, height, spp and max_bounces >= 1");
        if ((uint64_t)J.width * J.height > (1u << 28)) return fail(CRT_ERR_INVALID_ARG, "job image too large");
        if (J.options.mode > 1) return fail(CRT_ERR_INVALID_ARG, "bad mode");
        const uint32_t w = J.scene->dev.width;
        if (width && w != width) return fail(CRT_ERR_INVALID_ARG, "scenes of one batch must share the BVH width");
        width       = w;
        max_bounces = std::max(max_bounces, J.max_bounces);
        if (J.scene->dev.sun_enabled) {
            any_sun = true;
            any_alpha = any_alpha || J.scene->dev.any_alpha;
        }
        work += (uint64_t)J.width * J.height * J.spp;
    }
    // one stream and one spill area per context: batches, queries, autotuning and renderer
    // passes on a context run one at a time
    std::lock_guard<std::mutex> ctx_lock(ctx->mu);
    uint64_t P64 = o.max_paths ? o.max_paths : (ctx->batch_paths ? ctx->batch_paths : (4u << 20));
    P64 = std::min<uint64_t>(P64, work);
    P64 = std::min<uint64_t>(P64, 1u << 27);
    const uint32_t P = (uint32_t)((P64 + 255) / 256 * 256);
    BatchPlan plan;
    plan_batch(jobs, njobs, P, plan);
    const bool use_graph = o.use_graph && !any_alpha;

    HIP_CHECK(hipSetDevice(ctx->device));
    BatchRun run;
    const uint32_t nrep  = (uint32_t)plan.chunks.size();
    const uint32_t nsets = nrep > 1 ? 2u : 1u;
    for (uint32_t k = 0; k < nsets; ++k) HIP_CHECK(hipStreamCreateWithFlags(&run.stream[k], hipStreamNonBlocking));
    hipStream_t s = run.stream[0];
    auto &L = run.allocs;
    int st;
    PathBuffers b[2]{};
    uint32_t   *spill[2] = {nullptr, nullptr};
    for (uint32_t k = 0; k < nsets; ++k) {
        b[k].P    = P;
        b[k].npix = P;   // no cache-training paths in batches
        b[k].ncam = P;
        if ((st = alloc_path_core(L, b[k], P))) return st;
        if ((st = lalloc(L, &b[k].pjob, P))) return st;
        if ((st = lalloc(L, &b[k].ppix, P))) return st;
        if ((st = lalloc(L, &b[k].psamp, P))) return st;
        if ((st = lalloc(L, reinterpret_cast<uint8_t **>(&spill[k]), trace_spill_bytes(ctx->grid)))) return st;
    }
    // jobs, per-job traversal views, outputs
    std::vector<JobDev>     hj(njobs);
    std::vector<TraceScene> hts(njobs);
    std::vector<float *>    dev_out(njobs, nullptr);
    for (uint32_t j = 0; j < njobs; ++j) {
        const crt_job &J = jobs[j];
        JobDev &d = hj[j];
        d.sc             = J.scene->dev;
        d.rp             = RenderParams{};
        d.rp.cam         = to_camera_dev(J.scene->cam);
        d.rp.aspect      = (float)J.width / (float)J.height;   // renderer.cpp:199
        d.rp.max_bounces = J.max_bounces;
        d.rp.rr_start    = J.options.rr_start;
        d.rp.mode        = J.options.mode;
        d.rp.seed        = J.options.seed;
        d.rp.sample      = J.options.sample_offset;
        d.W              = J.width;
        d.H              = J.height;
        const size_t n3  = (size_t)J.width * J.height * 3;
        if (J.memory == CRT_MEM_DEVICE) {
            dev_out[j] = J.rgb_sum;
        } else if ((st = lalloc(L, &dev_out[j], n3))) {
            return st;
        }
        HIP_CHECK(hipMemsetAsync(dev_out[j], 0, sizeof(float) * n3, s));
        d.raw   = dev_out[j];
        hts[j]  = TraceScene{J.scene->dev.nodes, J.scene->dev.tri, J.scene->dev.root, J.scene->dev.ntri};
    }
    JobDev     *d_jobs = nullptr;
    TraceScene *d_ts = nullptr;
    BatchSeg   *d_segs = nullptr;
    BatchChunk *d_chunks = nullptr;
    uint32_t   *d_cursor = nullptr;
    if ((st = lalloc(L, &d_jobs, njobs))) return st;
    if ((st = lalloc(L, &d_ts, njobs))) return st;
    if ((st = lalloc(L, &d_segs, plan.segs.size()))) return st;
    if ((st = lalloc(L, &d_chunks, plan.chunks.size()))) return st;
    if ((st = lalloc(L, &d_cursor, 2))) return st;   // one per set: chunk k, then k + nsets, ...
    HIP_CHECK(hipMemcpyAsync(d_jobs, hj.data(), sizeof(JobDev) * njobs, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_ts, hts.data(), sizeof(TraceScene) * njobs, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_segs, plan.segs.data(), sizeof(BatchSeg) * plan.segs.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_chunks, plan.chunks.data(), sizeof(BatchChunk) * plan.chunks.size(),
                             hipMemcpyHostToDevice, s));
    const uint32_t cursor0[2] = {0u, 1u};
    HIP_CHECK(hipMemcpyAsync(d_cursor, cursor0, sizeof cursor0, hipMemcpyHostToDevice, s));
    for (uint32_t k = 0; k < nsets; ++k) {
        HIP_CHECK(hipMemsetAsync(b[k].cnt, 0, sizeof(uint32_t) * kCntWords, s));
        HIP_CHECK(hipMemsetAsync(b[k].stats, 0, sizeof(unsigned long long) * 4 * kStatSlots, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));   // host staging vectors are pageable; both streams start after this

    BatchDev bd[2];
    for (uint32_t k = 0; k < 2; ++k) bd[k] = BatchDev{d_jobs, d_segs, d_chunks, d_cursor + k};
    run.events.assign(nrep + 1, nullptr);
    for (hipEvent_t &e : run.events) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDefault));
    const TraceConfig tc = trace_cfg(ctx);
    auto trace_part = [&](uint32_t k) {
        return record_trace(b[k], bd[k], d_ts, width, max_bounces, any_sun, any_alpha, spill[k], tc, run.stream[k]);
    };
    auto acc_part = [&](uint32_t k) { return record_accumulate(b[k], bd[k], d_cursor + k, nsets, run.stream[k]); };
    if (use_graph) {
        for (uint32_t k = 0; k < nsets; ++k) {
            for (int part = 0; part < 2; ++part) {
                HIP_CHECK(hipStreamBeginCapture(run.stream[k], hipStreamCaptureModeThreadLocal));
                st = part == 0 ? trace_part(k) : acc_part(k);
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(run.stream[k], &g);
                run.graph[k][part] = g;
                if (st) return st;
                HIP_CHECK(ce);
                HIP_CHECK(hipGraphInstantiate(&run.exec[k][part], g, nullptr, nullptr, 0));
            }
        }
    }
    HIP_CHECK(hipEventRecord(run.events[0], s));
    for (uint32_t c = 0; c < nrep; ++c) {
        const uint32_t k  = c % nsets;
        hipStream_t    sk = run.stream[k];
        if (use_graph) HIP_CHECK(hipGraphLaunch(run.exec[k][0], sk));
        else if ((st = trace_part(k))) return st;
        // the job sums take the chunks in order
        if (c > 0) HIP_CHECK(hipStreamWaitEvent(sk, run.events[c], 0));
        if (use_graph) HIP_CHECK(hipGraphLaunch(run.exec[k][1], sk));
        else if ((st = acc_part(k))) return st;
        HIP_CHECK(hipEventRecord(run.events[c + 1], sk));
    }
    for (uint32_t k = 0; k < nsets; ++k) HIP_CHECK(hipStreamSynchronize(run.stream[k]));
    for (uint32_t k = 0; k < nsets; ++k) {
        uint32_t err = 0;
        HIP_CHECK(hipMemcpy(&err, b[k].cnt + kCntErr, sizeof err, hipMemcpyDeviceToHost));
        if (err) return fail(CRT_ERR_HIP, "traversal stack overflow");
    }
    for (uint32_t j = 0; j < njobs; ++j) {
        if (jobs[j].memory != CRT_MEM_DEVICE)
            HIP_CHECK(hipMemcpy(jobs[j].rgb_sum, dev_out[j], sizeof(float) * 3 * jobs[j].width * jobs[j].height,
                                hipMemcpyDeviceToHost));
        if (job_ms) {
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, run.events[0], run.events[plan.job_last[j] + 1]));
            job_ms[j] = ms;
        }
    }
    if (stats) {
        std::vector<unsigned long long> sv(4 * kStatSlots);
        uint64_t tot[4] = {0, 0, 0, 0};
        for (uint32_t set = 0; set < nsets; ++set) {
            HIP_CHECK(hipMemcpy(sv.data(), b[set].stats, sizeof(unsigned long long) * sv.size(), hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < kStatSlots; ++i)
                for (int k = 0; k < 4; ++k) tot[k] += sv[4 * i + k];
        }
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, run.events[0], run.events[nrep]));
        *stats = crt_batch_stats{};
        stats->ms_total        = ms;
        stats->extension_rays  = tot[0];
        stats->shadow_rays     = tot[1];
        stats->total_rays      = tot[2];
        stats->rays_per_second = ms > 0 ? (double)(tot[0] + tot[1]) / (ms * 1e-3) : 0.0;
        stats->replays         = nrep;
        stats->max_paths       = P;
        stats->used_graph      = use_graph ? 1u : 0u;
        stats->bvh_width       = width;
    }
    return CRT_OK;
}

int crt_autotune(crt_context *ctx, const crt_job *jobs, uint32_t njobs, uint32_t flags, crt_tune_result *out)
{
    if (!ctx || !jobs || njobs == 0 || !out) return fail(CRT_ERR_INVALID_ARG, "null argument");
    HIP_CHECK(hipSetDevice(ctx->device));
    // the sample: every job once at 1 spp into scratch device outputs
    std::vector<crt_job> sample(jobs, jobs + njobs);
    std::vector<float *> scratch(njobs, nullptr);
    uint64_t work = 0;
    struct Free {
        std::vector<float *> &v;
        ~Free() { for (float *p : v) dfree(p); }
    } guard{scratch};
    for (uint32_t j = 0; j < njobs; ++j) {
        if (!jobs[j].scene) return fail(CRT_ERR_INVALID_ARG, "job without scene");
        HIP_CHECK(dalloc(&scratch[j], (size_t)jobs[j].width * jobs[j].height * 3));
        sample[j].spp     = 1;
        sample[j].rgb_sum = scratch[j];
        sample[j].memory  = CRT_MEM_DEVICE;
        work += (uint64_t)jobs[j].width * jobs[j].height;
    }
    std::vector<crt_scene *> scenes;
    for (uint32_t j = 0; j < njobs; ++j)
        if (std::find(scenes.begin(), scenes.end(), jobs[j].scene) == scenes.end()) scenes.push_back(jobs[j].scene);

    uint32_t trials = 0;
    auto measure = [&](double *rays_per_s) -> int {
        crt_batch_options bo;
        crt_batch_options_default(&bo);
        bo.max_paths = ctx->batch_paths;
        crt_batch_stats bs{};
        double best = 0.0;
        for (int rep = 0; rep < 2; ++rep) {   // the first run pays for allocation and graph capture
            const int st = crt_render_batch(ctx, sample.data(), njobs, &bo, nullptr, &bs);
            if (st) return st;
            best = std::max(best, bs.rays_per_second);
        }
        ++trials;
        *rays_per_s = best;
        return CRT_OK;
    };
    auto set_grid = [&](uint32_t blocks) {
        std::lock_guard<std::mutex> lock(ctx->mu);   // no pass or query may use the old spill area
        dfree(ctx->spill);   // sized by the grid; renderers re-allocate it on demand
        dfree(ctx->spill2);
        ctx->spill  = nullptr;
        ctx->spill2 = nullptr;
        ctx->grid  = trace_grid(ctx->device, blocks);
        ctx->opt.trace_waves_per_cu = 4 * blocks;
    };
    auto recommit = [&](uint32_t width) -> int {
        ctx->opt.bvh_width = width;
        for (crt_scene *sc : scenes) {
            if (sc->dev.width == width) continue;
            const int st = crt_scene_commit(sc);
            if (st) return st;
        }
        return CRT_OK;
    };
    if (ctx->batch_paths == 0) ctx->batch_paths = 4u << 20;
    uint32_t blocks = ctx->opt.trace_waves_per_cu ? (ctx->opt.trace_waves_per_cu + 3) / 4 : 2;
    uint32_t width  = scenes[0]->dev.width;
    int st = recommit(width);   // one width for the batch
    if (st) return st;
    set_grid(blocks);
    double best = 0.0;
    if ((st = measure(&best))) return st;
    if (flags & CRT_TUNE_WIDTH) {
        const uint32_t other = width == 8 ? 4u : 8u;
        if ((st = recommit(other))) return st;
        double v = 0.0;
        if ((st = measure(&v))) return st;
        if (v > best) {
            best  = v;
            width = other;
        } else if ((st = recommit(width))) {
            return st;
        }
    }
    for (uint32_t cand : {2u, 3u, 4u, 5u, 6u, 8u}) {
        if (cand == blocks) continue;
        const uint32_t keep = blocks;
        set_grid(cand);
        double v = 0.0;
        if ((st = measure(&v))) return st;
        if (v > best) {
            best   = v;
            blocks = cand;
        } else {
            set_grid(keep);
        }
    }
    auto set_field = [&](uint32_t &field, uint32_t v) {
        std::lock_guard<std::mutex> lock(ctx->mu);   // read by passes running on the context
        field = v;
    };
    for (uint32_t cand : {8u, 16u, 32u}) {
        if (cand == ctx->stack) continue;
        const uint32_t keep = ctx->stack;
        set_field(ctx->stack, cand);
        double v = 0.0;
        if ((st = measure(&v))) return st;
        if (v > best) best = v;
        else set_field(ctx->stack, keep);
    }
    for (uint32_t cand : {1u << 20, 2u << 20, 4u << 20, 8u << 20}) {
        if (cand == ctx->batch_paths) continue;
        if (cand > 2 * work && cand > (1u << 20)) continue;   // larger than the sample: nothing to learn
        const uint32_t keep = ctx->batch_paths;
        set_field(ctx->batch_paths, cand);
        double v = 0.0;
        if ((st = measure(&v))) return st;
        if (v > best) best = v;
        else set_field(ctx->batch_paths, keep);
    }
    set_field(ctx->opt.trace_stack, ctx->stack);
    out->bvh_width           = width;
    out->trace_blocks_per_cu = blocks;
    out->trace_stack         = ctx->stack;
    out->max_paths           = ctx->batch_paths;
    out->rays_per_second     = best;
    out->trials              = trials;
    return CRT_OK;
}

}  // extern "C"

// ---- post-processing and denoising ------------------------------------------------------

namespace {

// Device copies of caller images (host memory) or the caller's device pointers; scratch.
struct ImageScratch {
    std::vector<void *> allocs;
    ~ImageScratch()
    {
        for (void *p : allocs) dfree(p);
    }
};

int image_in(ImageScratch &sc, const float *p, size_t n4, int memory, hipStream_t s, const float4 **out)
{
    if (memory == CRT_MEM_DEVICE) {
        *out = reinterpret_cast<const float4 *>(p);
        return CRT_OK;
    }
    float4 *d = nullptr;
    int     st = lalloc(sc.allocs, &d, n4);
    if (st) return st;
    HIP_CHECK(hipMemcpyAsync(d, p, sizeof(float4) * n4, hipMemcpyHostToDevice, s));
    *out = d;
    return CRT_OK;
}

}  // namespace

extern "C" {

void crt_post_options_default(crt_post_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->bloom_threshold = 0.7f;   // post_processor.h:22-23
    o->bloom_strength  = 1.0f;
    o->exposure        = 1.0f;   // post_processor.h:36-37
    o->gamma           = 2.2f;
}

int crt_post_process(crt_context *ctx, const crt_post_options *opt, const float *rgba_in, float *rgba_out,
                     uint32_t width, uint32_t height, int memory)
{
    if (!ctx || !rgba_in || !rgba_out || width == 0 || height == 0) return fail(CRT_ERR_INVALID_ARG, "null or empty image");
    if (memory != CRT_MEM_HOST && memory != CRT_MEM_DEVICE) return fail(CRT_ERR_INVALID_ARG, "bad memory kind");
    crt_post_options o;
    crt_post_options_default(&o);
    if (opt) o = *opt;
    if (o.tonemapping && (o.tonemapping_type < 0 || o.tonemapping_type > 3))
        return fail(CRT_ERR_INVALID_ARG, "tonemapping_type must be 0..3");
    const size_t n4 = (size_t)width * height;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if (!o.bloom && !o.gray_scale && !o.tonemapping) {   // post_processor.cpp:112-114
        const hipMemcpyKind k = memory == CRT_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToHost;
        HIP_CHECK(hipMemcpyAsync(rgba_out, rgba_in, sizeof(float4) * n4, k, s));
        HIP_CHECK(hipStreamSynchronize(s));
        return CRT_OK;
    }
    ImageScratch sc;
    const float4 *in = nullptr;
    int st = image_in(sc, rgba_in, n4, memory, s, &in);
    if (st) return st;
    float4 *out = nullptr, *t0 = nullptr, *t1 = nullptr;
    if (memory == CRT_MEM_DEVICE) out = reinterpret_cast<float4 *>(rgba_out);
    else if ((st = lalloc(sc.allocs, &out, n4))) return st;
    if ((st = lalloc(sc.allocs, &t0, n4))) return st;
    if (o.bloom && (st = lalloc(sc.allocs, &t1, n4))) return st;
    PostDev p{};
    p.w = width;
    p.h = height;
    p.cover_w = o.full_coverage ? width : width / 8u * 8u;
    p.cover_h = o.full_coverage ? height : height / 8u * 8u;
    p.bloom = o.bloom;
    p.threshold = o.bloom_threshold;
    p.strength = o.bloom_strength;
    p.gray = o.gray_scale;
    p.tonemap = o.tonemapping;
    p.tonemap_type = o.tonemapping_type;
    p.exposure = o.exposure;
    p.gamma = o.gamma;
    HIP_CHECK(launch_post(p, in, out, t0, t1 ? t1 : t0, s));
    if (memory != CRT_MEM_DEVICE) HIP_CHECK(hipMemcpyAsync(rgba_out, out, sizeof(float4) * n4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return CRT_OK;
}

void crt_denoise_options_default(crt_denoise_options *o)
{
    if (!o) return;
    o->levels       = 5;
    o->sigma_colour = 2.0f;
    o->sigma_normal = 0.2f;
    o->sigma_albedo = 0.2f;
}

int crt_denoise(crt_context *ctx, const crt_denoise_options *opt, const float *colour, const float *normal,
                const float *albedo, float *rgba_out, uint32_t width, uint32_t height, int memory)
{
    if (!ctx || !colour || !normal || !albedo || !rgba_out || width == 0 || height == 0)
        return fail(CRT_ERR_INVALID_ARG, "null or empty image");
    if (memory != CRT_MEM_HOST && memory != CRT_MEM_DEVICE) return fail(CRT_ERR_INVALID_ARG, "bad memory kind");
    crt_denoise_options o;
    crt_denoise_options_default(&o);
    if (opt) o = *opt;
    if (o.levels < 1 || o.levels > 10 || !(o.sigma_colour > 0.0f) || !(o.sigma_normal > 0.0f) ||
        !(o.sigma_albedo > 0.0f))
        return fail(CRT_ERR_INVALID_ARG, "bad denoise options");
    const size_t n4 = (size_t)width * height;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    ImageScratch sc;
    const float4 *c = nullptr, *n = nullptr, *a = nullptr;
    int st = image_in(sc, colour, n4, memory, s, &c);
    if (!st) st = image_in(sc, normal, n4, memory, s, &n);
    if (!st) st = image_in(sc, albedo, n4, memory, s, &a);
    if (st) return st;
    float4 *out = nullptr, *t0 = nullptr, *t1 = nullptr;
    if (memory == CRT_MEM_DEVICE) out = reinterpret_cast<float4 *>(rgba_out);
    else if ((st = lalloc(sc.allocs, &out, n4))) return st;
    if ((st = lalloc(sc.allocs, &t0, n4))) return st;
    if ((st = lalloc(sc.allocs, &t1, n4))) return st;
    HIP_CHECK(launch_atrous(c, n, a, out, t0, t1, width, height, (int)o.levels, o.sigma_colour, o.sigma_normal,
                            o.sigma_albedo, s));
    if (memory != CRT_MEM_DEVICE) HIP_CHECK(hipMemcpyAsync(rgba_out, out, sizeof(float4) * n4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return CRT_OK;
}

}  // extern "C"
