/*
 * crt.h -- C ABI of the crt MI355X path tracer.
 *
 * This is the drop-in boundary for CRender's render hot path (the src/render bounce loop
 * that calls Embree).  Every entry point below replaces one reference interface; the
 * reference file:line is given beside it.  All functions return CRT_OK (0) or a negative
 * CRT_ERR_* code and never exit the process; crt_last_error() gives the message of the
 * last failure on the calling thread.
 *
 * Threading: a scene is immutable after crt_scene_commit.  Calls on distinct renderers
 * are thread-safe.  Each context owns one HIP stream; every renderer and query of a
 * context is ordered on it.  Mutating a scene or renderer requires the renderer to be
 * paused, mirroring cr::renderer::update (renderer.cpp:185-192).
 *
 * Memory: buffers passed to crt_scene_* are copied at commit; the caller may free them
 * after the call returns (the reference shares them with Embree instead,
 * model.cpp:60-90).  crt_intersect/crt_occluded accept host or device pointers
 * (crt_memory).
 */
#ifndef CRT_H
#define CRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRT_ABI_VERSION 1

enum crt_status {
    CRT_OK               = 0,
    CRT_ERR_INVALID_ARG  = -1,
    CRT_ERR_INVALID_MESH = -2,
    CRT_ERR_HIP          = -3,
    CRT_ERR_OUT_OF_MEMORY = -4,
    CRT_ERR_STATE        = -5, /* e.g. mutation while rendering, query before commit */
    CRT_ERR_UNSUPPORTED  = -6,
    CRT_ERR_CANCELLED    = -7,
    CRT_ERR_CHECKPOINT   = -8
};

enum crt_memory { CRT_MEM_HOST = 0, CRT_MEM_DEVICE = 1 };

/* material.h:14-19 */
enum crt_material_type { CRT_METAL = 0, CRT_SMOOTH = 1, CRT_GLASS = 2 };

/* estimator: compat reproduces renderer.cpp:21-102,258-384 (with Philox); pbr fixes the
 * estimator (Lambert NEE with 1/pi, GGX metal, Fresnel glass, normals facing the ray). */
enum crt_mode { CRT_MODE_COMPAT = 0, CRT_MODE_PBR = 1 };

/* buffers of cr::renderer (renderer.h:62-68, renderer.cpp:85-91) */
enum crt_buffer {
    CRT_BUFFER_PROGRESS = 0, /* current_progress(): RGBA, pow(clamp(sum/n,0,1),1/2.2) */
    CRT_BUFFER_NORMALS  = 1, /* current_normals():  RGBA, n*0.5+0.5 of the last pass */
    CRT_BUFFER_ALBEDOS  = 2, /* current_albedos():  RGBA */
    CRT_BUFFER_DEPTHS   = 3, /* current_depths():   RGBA, min(d,200)/200 */
    CRT_BUFFER_RAW      = 4  /* _raw_buffer: RGB FP32 running sum (W*H*3) */
};

/* material::information (material.h:31-41) without the UI name */
typedef struct crt_material {
    int32_t type;            /* crt_material_type */
    float   ior;             /* default 1.5 */
    float   roughness;       /* default 0.5 (GGX alpha in pbr mode) */
    float   reflectiveness;  /* default 1 */
    float   emission;        /* default 0 */
    float   colour[4];       /* default (1,1,1,1); colour[3] == 0 is alpha pass-through */
    int32_t tex;             /* -1: none; else index into the model's textures */
} crt_material;

/* cr::image (image.h:13-151): RGBA FP32, (x + y*w)*4 */
typedef struct crt_image {
    const float *rgba;
    uint32_t     width, height;
} crt_image;

/* asset_loader::model_data (asset_loader.h:16-30).  Triangles only: vertex_indices,
 * texture_indices hold 3 entries per triangle, material_indices 1.  The model is
 * de-indexed ("expanded") at add time as registry.cpp:13-24,62-66 does, so
 * primID = triangle index in this order. */
typedef struct crt_model_data {
    const char         *name;
    const float        *vertices;         /* vec3 */
    uint32_t            num_vertices;
    const float        *texture_coords;   /* vec2, may be NULL */
    uint32_t            num_texture_coords;
    const crt_material *materials;
    uint32_t            num_materials;
    const crt_image    *textures;
    uint32_t            num_textures;
    const uint32_t     *vertex_indices;   /* 3 per triangle */
    const uint32_t     *texture_indices;  /* 3 per triangle, may be NULL */
    const uint32_t     *material_indices; /* 1 per triangle */
    uint32_t            num_triangles;
} crt_model_data;

/* entity::sun (components.h:23-29) */
typedef struct crt_sun {
    float size;       /* cone half-angle, default pi/48 */
    float intensity;  /* default 100 */
    float direction[3];
    float colour[3];
} crt_sun;

/* cr::camera (camera.h:18-49).  matrix is _cached_matrix, column-major; crt_camera_update
 * recomputes it from position and rotation as camera.cpp:54-68 does. */
typedef struct crt_camera {
    float   position[3];
    float   rotation[3];  /* degrees: yaw about +y, pitch about +x, roll about +z */
    float   fov;          /* degrees, default 75 */
    int32_t mode;         /* 0 perspective, 1 orthographic */
    float   scale;        /* orthographic half-extent */
    float   matrix[16];
} crt_camera;

typedef struct crt_options {
    int32_t  device;        /* HIP device ordinal */
    uint32_t bvh_width;     /* 0 = autotune, 2/4/8 */
    uint32_t max_leaf;      /* 0 = default (4) */
    uint32_t trace_waves_per_cu; /* 0 = default persistent occupancy (8 waves per CU) */
    uint32_t flags;         /* CRT_FLAG_* */
    uint32_t trace_stack;   /* LDS stack entries per traversal lane: 0 (= 16), 8, 16 or 32 */
    uint32_t wavefront_paths; /* camera paths per renderer wavefront: 0 = 2^24 (2^23 for a
                                 renderer with a radiance cache); a wavefront carries
                                 floor(wavefront_paths / (W*H)) sample passes (1 to 8).  Two
                                 wavefronts are in flight (two sets of path buffers, two streams). */
    uint32_t trace_refill;  /* traversal waves fetch new rays once this many of their 64 lanes
                               are idle: 0 = default (32), else 1 to 64 */
    uint32_t trace_park_min;/* the inner-node phase of a wave ends once fewer than this many
                               lanes can still park a leaf: 0 = default (16), else 1 to 64 */
    uint32_t trace_leaf_rounds; /* leaf-loop rounds after which a wave whose parked leaves are
                               all tested resumes the inner-node phase: 0 = default (3), 1 to 64,
                               or above 64 = until no lane holds a leaf */
    float    bvh_prim_cost;  /* cost of a triangle test relative to a wide-node test in the
                               collapse of the binary BVH into wide nodes: 0 = default (1.0) */
} crt_options;

#define CRT_FLAG_PHASE_TIMING 1u /* record HIP events around every kernel of a pass */
#define CRT_FLAG_SORT_RAYS    2u /* reorder secondary ray queues by origin / direction before tracing */
#define CRT_FLAG_SORT_MATERIALS 4u /* shade each bounce's hits grouped by material class */
#define CRT_FLAG_TRAVERSAL_STATS 8u /* renderers count node visits, leaf visits and triangle
                                       tests (an instrumented traversal kernel: slower) */

typedef struct crt_render_options {
    uint32_t mode;          /* crt_mode */
    uint32_t rr_start;      /* 0 = no Russian roulette, else RR before bounce >= rr_start */
    uint32_t seed;          /* Philox key word 0 */
    uint32_t sample_offset; /* absolute index of the next pass (resume) */
} crt_render_options;

/* renderer::renderer_stats (renderer.h:48-54) plus per-kind ray counts and phase times. */
typedef struct crt_stats {
    uint64_t rays_per_second;    /* reference counting: total_rays / running_time */
    uint64_t samples_per_second;
    uint64_t total_rays;         /* reference counting: sum of total_bounces (renderer.cpp:271-272,356) */
    double   running_time;       /* seconds since start() */
    uint64_t samples;            /* completed passes */
    uint64_t extension_rays;     /* closest-hit rays traced */
    uint64_t shadow_rays;        /* sun shadow rays traced */
    double   ms_generate, ms_extend, ms_shade, ms_shadow, ms_accumulate; /* with CRT_FLAG_PHASE_TIMING */
    double   ms_render;          /* GPU time of all passes (HIP events) */
    uint64_t extend_launches;    /* closest-hit launches covered by ms_extend */
    double   ms_nrc_train;       /* GPU time of the cache's training steps (events on its stream) */
    float    nrc_loss;           /* training loss of the last step (one tile) */
    uint64_t node_visits;        /* with CRT_FLAG_TRAVERSAL_STATS, over all traced rays: */
    uint64_t leaf_visits;        /*   inner nodes tested, leaves entered, */
    uint64_t tri_tests;          /*   ray-triangle tests, */
    uint64_t wave_node_steps;    /*   wave iterations of the inner-node loop, */
    uint64_t wave_leaf_steps;    /*   wave iterations of the leaf loop */
    uint32_t nrc_detached;       /* 1: the cache's predictions became non-finite and the renderer
                                    went on with the plain tracer (until the next set_nrc) */
    uint32_t nrc_discarded_passes; /* sample passes rendered with the diverged cache, dropped
                                      from the image and rendered again without it */
    uint64_t node_idle_no_ray;   /* with CRT_FLAG_TRAVERSAL_STATS: lanes idle in inner-node loop
                                    iterations because they hold no ray, */
    uint64_t node_idle_done;     /*   because their ray has nothing left but parked leaves, */
    uint64_t node_idle_leaf_wait;/*   because they reached a leaf with both parking slots full */
    double   ms_extend_busy;     /* with CRT_FLAG_PHASE_TIMING: time during which at least one
                                    closest-hit launch ran (the union of their intervals over both
                                    wavefronts in flight); ms_extend sums the launches' own spans,
                                    which overlap */
} crt_stats;

/* BVH summary returned by crt_scene_bvh_info */
typedef struct crt_bvh_info {
    uint32_t width;
    uint32_t node_count;       /* wide nodes */
    uint32_t bvh2_node_count;  /* binary nodes before collapse */
    uint32_t depth;
    uint32_t max_stack;        /* worst-case traversal stack entries */
    double   sah_cost;
    double   build_ms;         /* device build (binned SAH + collapse), HIP events */
    uint64_t bytes;            /* nodes + triangles in HBM */
} crt_bvh_info;

typedef struct crt_rays {     /* SoA, Embree RTCRayNp-like */
    const float *org_x, *org_y, *org_z;
    const float *dir_x, *dir_y, *dir_z;
    const float *tnear, *tfar;
} crt_rays;

typedef struct crt_hits {     /* SoA; any pointer except t/prim_id may be NULL */
    float    *t;              /* +inf on miss */
    float    *u, *v;          /* barycentrics of v1 and v2 (Embree convention) */
    uint32_t *prim_id;        /* 0xFFFFFFFF on miss (RTC_INVALID_GEOMETRY_ID) */
    uint32_t *geom_id;        /* model index, 0xFFFFFFFF on miss */
    float    *ng_x, *ng_y, *ng_z; /* unnormalised (v1-v0)x(v2-v0), object space for instances */
    uint32_t *inst_id;        /* instance of the model (crt_scene_set_instances order), 0xFFFFFFFF on miss */
} crt_hits;

/* cr::ray::intersection_record (ray.h:12-20) for a batch of rays.  SoA; any pointer except
 * distance may be NULL.  On a miss: distance +inf, the IDs 0xFFFFFFFF, the rest 0. */
typedef struct crt_records {
    float    *distance;           /* |P - origin| (model.cpp:118 for instances) */
    uint32_t *prim_id;            /* primitive within its model */
    uint32_t *geom_id;            /* model index */
    uint32_t *material;           /* index into the scene's material table (materials.indices[primID]) */
    float    *u, *v;              /* texture coordinates interpolated at the hit (rtcInterpolate0); 0 without uv */
    float    *n_x, *n_y, *n_z;    /* normalize(Ng); object space for a hit on an instance (model.cpp:36-37) */
    float    *p_x, *p_y, *p_z;    /* intersection point in world space */
} crt_records;

typedef struct crt_context  crt_context;
typedef struct crt_scene    crt_scene;
typedef struct crt_renderer crt_renderer;

/* ---- errors / version -------------------------------------------------------------- */
const char *crt_last_error(void);
int         crt_abi_version(void);

/* ---- lifecycle: rtcNewDevice (components.h:67) + the renderer's thread pool ----------- */
void crt_options_default(crt_options *opt);
int  crt_context_create(const crt_options *opt, crt_context **out);
/* CRT_ERR_STATE while scenes, renderers or caches created on the context are alive. */
int  crt_context_destroy(crt_context *ctx);
int  crt_context_device_name(crt_context *ctx, char *buf, size_t len);

/* ---- scene-load: cr::scene (scene.h:19-53) ------------------------------------------ */
int crt_scene_create(crt_context *ctx, crt_scene **out);                      /* scene() */
int crt_scene_destroy(crt_scene *scene);   /* CRT_ERR_STATE while a renderer uses it */
/* scene::add_model (scene.cpp:31-34) -> registry::register_model (registry.cpp:51-97) */
int crt_scene_add_model(crt_scene *scene, const crt_model_data *model, uint32_t *model_id);
/* entity::instances (components.h:18-21); transforms are column-major affine 4x4
 * (object -> world).  Identity by default (registry.cpp:71-72); count 0 hides the model.
 * A scene with any other instancing is committed as a two-level BVH and traced as
 * model.cpp:99-126 does (ray mapped into object space with its direction renormalised,
 * world distances compared; hit normals stay in object space in compat mode). */
int crt_scene_set_instances(crt_scene *scene, uint32_t model_id, const float *transforms, uint32_t count);
/* replace a model's material table (ui.h:924-929 edits) */
int crt_scene_set_materials(crt_scene *scene, uint32_t model_id, const crt_material *mats, uint32_t count);
int crt_scene_set_skybox(crt_scene *scene, const crt_image *sky);             /* scene.cpp:36-60 */
int crt_scene_set_skybox_rotation(crt_scene *scene, float x, float y);        /* scene.cpp:62-65 */
int crt_scene_set_sun_enabled(crt_scene *scene, int enabled);                 /* scene.cpp:115-118 */
int crt_scene_set_sun(crt_scene *scene, const crt_sun *sun);                  /* registry.cpp:248-256 */
int crt_scene_set_camera(crt_scene *scene, const crt_camera *camera);         /* registry camera entity */
/* Uploads geometry and builds the BVH on the device: rtcCommitGeometry / rtcCommitScene
 * (model.cpp:92-94). */
int crt_scene_commit(crt_scene *scene);
int crt_scene_bvh_info(crt_scene *scene, crt_bvh_info *info);
/* Copies the wide BVH out for structural checks: nodes as stored (128 B each) and the
 * leaf-order primID table.  Pass NULL to query sizes. */
int crt_scene_bvh_export(crt_scene *scene, void *nodes, uint64_t *node_bytes, uint32_t *prims, uint32_t *nprims);

/* ---- ray query: scene::cast_ray (scene.cpp:79-98) -> rtcIntersect1 (model.cpp:27) ------ */
int crt_intersect(crt_scene *scene, const crt_rays *rays, uint32_t n, crt_hits *hits, int memory);
/* No counterpart in the reference: crt_intersect for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Rays and hits are device pointers; the call allocates nothing.  Its scratch
 * (queue counters, intermediate hit indices, the traversal's spill stack) is a fixed
 * 48 MiB device array per GPU, held until the process ends, so a call takes at most
 * 1,048,576 rays (CRT_ERR_INVALID_ARG beyond) and traces with a 64-block grid (slower than
 * crt_intersect on large batches).  All calls on one GPU share that array: calls that
 * overlap on one GPU (from different contexts or threads) race.  Results equal
 * crt_intersect's with CRT_MEM_DEVICE. */
#define CRT_NOALLOC_MAX_RAYS 1048576u
int crt_intersect_noalloc(crt_scene *scene, const crt_rays *rays, uint32_t n, crt_hits *hits);
/* scene::cast_ray (scene.cpp:79-98) for a batch: the closest hit of each ray as the
 * intersection record model.cpp:27-47 fills in.  Each ray is traced along its normalised
 * direction, as the per-instance transform does (model.cpp:107-110); tnear / tfar default to
 * the reference's 1e-5 and +inf (model.cpp:22-23) where rays->tnear / rays->tfar are NULL. */
int crt_cast_rays(crt_scene *scene, const crt_rays *rays, uint32_t n, crt_records *records, int memory);
/* any hit in (tnear, tfar]; occluded[i] = 1/0 */
int crt_occluded(crt_scene *scene, const crt_rays *rays, uint32_t n, uint8_t *occluded, int memory);

/* ---- render / accumulate: cr::renderer (renderer.h:27-68) ---------------------------- */
/* renderer(res_x, res_y, bounces, pool, scene) (renderer.cpp:106-145).  Created paused;
 * crt_renderer_start begins background passes as the management thread does. */
int crt_renderer_create(crt_context *ctx, crt_scene *scene, uint32_t res_x, uint32_t res_y,
                        uint32_t bounces, crt_renderer **out);
int crt_renderer_destroy(crt_renderer *r);                                    /* ~renderer */
int crt_renderer_start(crt_renderer *r);      /* renderer.cpp:154-170 (clears accumulation) */
int crt_renderer_pause(crt_renderer *r);      /* renderer.cpp:172-183 (waits for the pass) */
int crt_renderer_is_running(crt_renderer *r, int *running);
int crt_renderer_wait(crt_renderer *r);       /* block until the target spp is reached */
int crt_renderer_set_resolution(crt_renderer *r, uint32_t x, uint32_t y);    /* renderer.cpp:194-208 */
int crt_renderer_set_max_bounces(crt_renderer *r, uint32_t bounces);         /* renderer.cpp:210-213 */
int crt_renderer_set_target_spp(crt_renderer *r, uint64_t target);           /* renderer.cpp:215-218 */
int crt_renderer_set_options(crt_renderer *r, const crt_render_options *opt);
int crt_renderer_get_options(crt_renderer *r, crt_render_options *opt);
/* Synchronous passes on the caller's thread (the renderer must be paused).  After a pass
 * fails (e.g. CRT_ERR_HIP "traversal stack overflow"), the accumulation may hold samples that
 * were not counted; further passes return CRT_ERR_STATE until crt_renderer_reset,
 * crt_renderer_start or a resolution change clears the accumulation. */
int crt_renderer_render(crt_renderer *r, uint32_t passes);
/* Clears accumulation, sample count and stats without starting (start() minus resume). */
int crt_renderer_reset(crt_renderer *r);
int crt_renderer_current_sample_count(crt_renderer *r, uint64_t *count);     /* renderer.cpp:391-394 */
int crt_renderer_current_resolution(crt_renderer *r, uint32_t *x, uint32_t *y); /* renderer.cpp:386-389 */
/* current_progress/normals/albedos/depths (renderer.cpp:220-238) and _raw_buffer:
 * copies into host memory (W*H*4 floats, or W*H*3 for CRT_BUFFER_RAW). */
int crt_renderer_read(crt_renderer *r, int buffer, float *out);
/* Device pointer of a buffer (valid until resize/destroy); PROGRESS is refreshed first. */
int crt_renderer_device_buffer(crt_renderer *r, int buffer, void **dev_ptr);
int crt_renderer_current_stats(crt_renderer *r, crt_stats *stats);           /* renderer.cpp:396-404 */
/* Stop the background passes at the next bounce boundary (pause() without waiting). */
int crt_renderer_cancel(crt_renderer *r);

/* Checkpoint = {raw sum, sample count, render options, resolution, bounces, scene hash}.
 * Resuming continues at the stored sample index; with the counter-based RNG the result
 * equals an uninterrupted run bit for bit. */
int crt_renderer_checkpoint_size(crt_renderer *r, uint64_t *bytes);
int crt_renderer_checkpoint_save(crt_renderer *r, void *buf, uint64_t bytes);
int crt_renderer_checkpoint_load(crt_renderer *r, const void *buf, uint64_t bytes);

/* ---- neural radiance cache (no counterpart in the reference: the learned part) -------- *
 * MLP over an encoded query (position normalised to the scene box, direction, normal,
 * albedo, roughness: 13 floats) -> RGB radiance; bf16 weights, FP32 accumulation, Adam. */
typedef struct crt_nrc crt_nrc;
typedef struct crt_nrc_options {
    uint32_t max_batch;   /* largest training batch per step (default 65536) */
    uint32_t seed;        /* weight initialisation (Philox) */
    float    lr, beta1, beta2, eps;   /* Adam (default 1e-3, 0.9, 0.99, 1e-10) */
} crt_nrc_options;
void crt_nrc_options_default(crt_nrc_options *opt);
int  crt_nrc_create(crt_context *ctx, const crt_nrc_options *opt, crt_nrc **out);
int  crt_nrc_destroy(crt_nrc *nrc);   /* CRT_ERR_STATE while attached to a renderer */
int  crt_nrc_param_count(void);
int  crt_nrc_get_params(crt_nrc *nrc, float *params);          /* FP32 master weights, host */
int  crt_nrc_set_params(crt_nrc *nrc, const float *params);    /* also resets the Adam state */
/* queries: n x 13 floats; out: n x 3 floats */
int  crt_nrc_infer(crt_nrc *nrc, const float *queries, uint32_t n, float *out, int memory);
/* gradient of the relative L2 loss (host output, kNrcParams floats) without updating */
int  crt_nrc_gradient(crt_nrc *nrc, const float *queries, const float *targets, uint32_t n, float *grad,
                      float *loss, int memory);
/* forward + backward + one Adam step */
int  crt_nrc_train_step(crt_nrc *nrc, const float *queries, const float *targets, uint32_t n, float *loss,
                        int memory);

/* Rendering with the cache: paths query it at bounce `term_bounce` (one bounce later when
 * the path footprint is still small, footprint_c > 0) instead of continuing; every pass
 * also traces train_paths training paths train_suffix bounces past the query vertex and
 * takes train_steps Adam steps on their radiance.  nrc = NULL turns the cache off. */
typedef struct crt_nrc_render_options {
    uint32_t term_bounce;   /* default 1 */
    float    footprint_c;   /* default 0.01; 0 = always query at term_bounce */
    uint32_t train_paths;   /* default 65536 */
    uint32_t train_suffix;  /* default 4 */
    uint32_t train_steps;   /* default 24 */
    uint32_t train;         /* 0 freezes the cache (inference only) */
    uint32_t warmup_passes; /* first passes after set_nrc only train the cache: their image is
                               discarded and they do not count as samples (default 0) */
} crt_nrc_render_options;
void crt_nrc_render_options_default(crt_nrc_render_options *opt);
int  crt_renderer_set_nrc(crt_renderer *r, crt_nrc *nrc, const crt_nrc_render_options *opt);

/* ---- scene-load format: Wavefront OBJ + MTL --------------------------------------------
 * asset_loader::load_model (asset_loader.cpp:182-303): triangulates polygons as
 * tinyobjloader does (tinobj.h:1393-1714), keeps Kd as the colour and map_Kd as the
 * texture (PNG / PPM, flipped vertically) of a smooth material with emission 0.  Faces
 * without a material and corners without a texture coordinate get an appended default
 * (the reference would index out of bounds).  folder NULL: the OBJ file's directory.
 * The model view stays valid until crt_obj_free; pass it to crt_scene_add_model. */
typedef struct crt_obj crt_obj;
int         crt_obj_load(const char *file, const char *folder, crt_obj **out);
int         crt_obj_model(const crt_obj *obj, crt_model_data *model);
const char *crt_obj_warnings(const crt_obj *obj);
int         crt_obj_free(crt_obj *obj);

/* ---- output formats: asset_loader::export_framebuffer (asset_loader.cpp:89-178,348-377) --
 * Writes a W*H RGBA FP32 image (row 0 first, as crt_renderer_read returns it): PNG =
 * RGBA8 of min(v*255, 255); EXR = half-float B, G, R, uncompressed; HDR = Radiance RGBE of
 * pow(v, 2.2); PFM = FP32 RGB (no counterpart; lossless HDR sums).  JPG is reported as
 * CRT_ERR_UNSUPPORTED.  The reference chooses a free name under ./out/; here the caller
 * gives the path. */
enum crt_image_format { CRT_IMAGE_PNG = 0, CRT_IMAGE_JPG = 1, CRT_IMAGE_EXR = 2, CRT_IMAGE_HDR = 3, CRT_IMAGE_PFM = 4 };
int crt_image_write(const char *path, int format, const float *rgba, uint32_t width, uint32_t height);

/* ---- post-processing: cr::post_processor (post_processor.h:19-42, post_processor.cpp,
 * post_process.comp, blur.comp) on the GPU.  In and out are W*H RGBA FP32 images. */
typedef struct crt_post_options {
    int32_t bloom;            /* default 0 */
    float   bloom_threshold;  /* default 0.7 */
    float   bloom_strength;   /* default 1 */
    int32_t gray_scale;       /* default 0 */
    int32_t tonemapping;      /* default 0 */
    int32_t tonemapping_type; /* 0 linear, 1 Reinhard, 2 Jim and Richard, 3 Uncharted 2 */
    float   exposure;         /* default 1 */
    float   gamma;            /* default 2.2 */
    int32_t full_coverage;    /* 0: like the reference's floor(W/8) x floor(H/8) dispatch, pixels
                                 beyond it are 0; 1: every pixel */
} crt_post_options;
void crt_post_options_default(crt_post_options *opt);
int  crt_post_process(crt_context *ctx, const crt_post_options *opt, const float *rgba_in, float *rgba_out,
                      uint32_t width, uint32_t height, int memory);

/* Denoising of the colour image guided by the normal and albedo images, the inputs
 * cr::denoise passes to Open Image Denoise (denoise.h:10-42).  OIDN's pretrained network is
 * not part of this build: this is an edge-avoiding a-trous wavelet filter (Dammertz et al.
 * 2010) on the GPU.  Images are W*H RGBA FP32; the output alpha is 1. */
typedef struct crt_denoise_options {
    uint32_t levels;         /* default 5 (footprint 2^levels * 4 + 1 pixels) */
    float    sigma_colour;   /* default 2.0 (display-range colour), halved per level */
    float    sigma_normal;   /* default 0.2 (normals stored n * 0.5 + 0.5) */
    float    sigma_albedo;   /* default 0.2 */
} crt_denoise_options;
void crt_denoise_options_default(crt_denoise_options *opt);
int  crt_denoise(crt_context *ctx, const crt_denoise_options *opt, const float *colour, const float *normal,
                 const float *albedo, float *out, uint32_t width, uint32_t height, int memory);

/* ---- batched rendering (no counterpart in the reference, whose renderer owns one scene) --
 * Many independent jobs -- a committed scene at its own resolution, sample count and depth
 * -- render together: every traversal launch carries rays of all jobs, and one pass of
 * the wavefront (generate, per bounce extend / shade / shadow, accumulate) is replayed from
 * a hipGraph until all jobs are done.  Each job's result equals, bit for bit, what a
 * crt_renderer with the same scene, resolution (set_resolution, so aspect = W/H), bounces
 * and options accumulates over the same number of passes (renderer.cpp:116-144,258-384). */
typedef struct crt_job {
    crt_scene         *scene;        /* committed; all scenes of a batch share one BVH width */
    uint32_t           width, height;
    uint32_t           spp;          /* passes, >= 1 */
    uint32_t           max_bounces;  /* >= 1 */
    crt_render_options options;      /* mode, rr_start, seed, sample_offset */
    float             *rgb_sum;      /* out: W*H*3 FP32 running sum, flipped like _raw_buffer */
    int32_t            memory;       /* crt_memory of rgb_sum */
} crt_job;

typedef struct crt_batch_options {
    uint32_t max_paths;   /* path slots per replay (wavefront size); 0 = the context's tuned value */
    uint32_t use_graph;   /* 1 (default): replay a captured hipGraph; 0: launch kernels directly */
} crt_batch_options;

typedef struct crt_batch_stats {
    double   ms_total;          /* first launch to the end of the last job */
    double   rays_per_second;   /* (extension + shadow rays) / ms_total */
    uint64_t extension_rays, shadow_rays;
    uint64_t total_rays;        /* reference-style count: 1 + loop iterations per path */
    uint32_t replays;           /* passes of the wavefront (graph launches) */
    uint32_t max_paths;
    uint32_t used_graph;
    uint32_t bvh_width;
} crt_batch_stats;

/* Autotuner of the traversal and wavefront shape for a workload: times short batches of
 * `jobs` (every job at 1 spp, outputs discarded) over BVH width {4, 8} (CRT_TUNE_WIDTH:
 * re-commits the jobs' scenes), traversal blocks per CU {2, 3, 4, 5, 6, 8}, LDS stack depth
 * {8, 16, 32} and wavefront size {1, 2, 4, 8} M paths, one parameter at a time, and keeps
 * the fastest on the context (used by later batches and renderers).  The context must be
 * idle. */
#define CRT_TUNE_WIDTH 1u
typedef struct crt_tune_result {
    uint32_t bvh_width;
    uint32_t trace_blocks_per_cu;
    uint32_t trace_stack;
    uint32_t max_paths;
    double   rays_per_second;   /* of the fastest trial */
    uint32_t trials;
} crt_tune_result;
int crt_autotune(crt_context *ctx, const crt_job *jobs, uint32_t njobs, uint32_t flags, crt_tune_result *out);

void crt_batch_options_default(crt_batch_options *opt);
/* job_ms (optional, njobs entries): ms from the batch start to the end of each job. */
int  crt_render_batch(crt_context *ctx, const crt_job *jobs, uint32_t njobs, const crt_batch_options *opt,
                      double *job_ms, crt_batch_stats *stats);

/* camera.cpp:54-68: matrix = T(position) * Ry(rot.x) * Rx(rot.y) * Rz(rot.z) */
void crt_camera_default(crt_camera *cam);   /* camera.h:21: position (5,5,0), fov 75, identity */
void crt_camera_update(crt_camera *cam);
void crt_sun_default(crt_sun *sun);         /* components.h:23-29 */
void crt_material_default(crt_material *m); /* material.h:31-41 */

#ifdef __cplusplus
}
#endif
#endif
