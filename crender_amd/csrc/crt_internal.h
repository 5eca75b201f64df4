// crt_internal.h -- device data layout and kernel entry points shared by the crt sources.
//
// HBM layout of a committed scene:
//   tri      float4[3*ntri]   triangles in BVH leaf order: {v0,primID bits}{v1,0}{v2,0}
//   nodes    128-byte wide nodes (Node4: FP32 boxes; Node8Q: 8-bit quantised boxes)
//   mat_idx  u32[ntri]        material per primID (original order)
//   uv       float2[3*ntri]   per-corner texture coordinates (original order)
// Rays and hits are SoA float/u32 arrays, one element per ray (coalesced per wave).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "crt_math.h"

namespace crt {

constexpr uint32_t kLeafBit    = 0x80000000u;
constexpr uint32_t kInvalidRef = 0xFFFFFFFFu;
constexpr int      kNumBins    = 16;
constexpr uint32_t kChunk      = 2048;   // prims per workgroup in the SAH builder

// child reference: inner node index, or leaf = kLeafBit | (count-1) << 27 | first
CRT_HD bool ref_is_leaf(uint32_t r) { return (r & kLeafBit) != 0u; }
CRT_HD uint32_t ref_leaf(uint32_t first, uint32_t count) { return kLeafBit | ((count - 1u) << 27) | first; }
CRT_HD uint32_t leaf_first(uint32_t r) { return r & 0x07FFFFFFu; }
CRT_HD uint32_t leaf_count(uint32_t r) { return ((r >> 27) & 0xFu) + 1u; }

// Binary node written by the builder; slot 2*begin for a segment starting at begin.
struct Bvh2Node {
    float    lo[3];
    uint32_t a;   // inner: left child slot; leaf: first primitive position
    float    hi[3];
    uint32_t b;   // inner: right child slot; leaf: kLeafBit | count
};

// 4-wide node, FP32 child boxes, one 128-byte line.
struct alignas(128) Node4 {
    float    lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t child[4];
    uint32_t nchild, pad0, pad1, pad2;
};
static_assert(sizeof(Node4) == 128, "Node4 is one cache line");

// 8-wide node, child boxes quantised to 8 bits per plane relative to the node origin:
// lo = origin + q * 2^(e-127), one 128-byte line.
struct alignas(128) Node8Q {
    float    ox, oy, oz;
    uint32_t meta;      // ex | ey << 8 | ez << 16 | nchild << 24 (biased exponents)
    uint32_t child[8];
    uint8_t  qlox[8], qhix[8], qloy[8], qhiy[8], qloz[8], qhiz[8];
    uint32_t pad[8];
};
static_assert(sizeof(Node8Q) == 128, "Node8Q is one cache line");

// A segment [begin, end) of the primitive order, its BVH2 slot and centroid bounds.
// A node in slot p covering n primitives owns slots [p, p + 2n - 1): the left child is
// slot p + 1, the right child slot p + 2 * nleft, so the tree layout is deterministic.
struct BuildTask {
    uint32_t begin, end, slot, pad;
    float    cblo[3], cbhi[3];
};

struct BinSet {   // per axis per bin: count, AABB (ordered ints), centroid bounds (ordered ints)
    int v[3][kNumBins][13];
};

struct SplitInfo {
    int32_t  axis;     // -1: split by position (all centroids coincide)
    int32_t  bin;
    uint32_t nleft;
    uint32_t leaf;     // 1: make a leaf
};

struct MaterialDev {
    int32_t type;
    float   ior, roughness, reflectiveness, emission;
    float   colour[4];
    int32_t tex;       // -1 or index into the texture table
};

struct TextureDev {
    uint64_t offset;   // float4 offset into the texel pool
    uint32_t w, h;
};

// Instance of a model (two-level scenes): world->object affine transform (row-major
// 3x4) and the model BVH's root in the scene's node pool.
struct InstDev {
    float    inv[12];
    uint32_t root;
    uint32_t model;
    uint32_t local;    // index among the model's instances
    uint32_t pad;
};
static_assert(sizeof(InstDev) == 64, "InstDev is 64 bytes");

struct SunDev {
    float size, intensity;
    f3    dir, colour;
};

// Everything the kernels read about a committed scene.
struct SceneDev {
    const float4      *tri;
    const void        *nodes;
    uint32_t           nnodes;     // wide nodes in `nodes`
    uint32_t           width;      // 4 or 8
    uint32_t           ntri;
    uint32_t           root;       // root child ref
    const uint32_t    *mat_idx;
    const float2      *uv;         // may be null
    const MaterialDev *mats;
    uint32_t           nmat;
    const float4      *texels;
    const TextureDev  *texs;
    uint32_t           ntex;
    int32_t            sky_tex;    // -1: black sky
    float              sky_rot[2];
    int32_t            sun_enabled;
    SunDev             sun;
    int32_t            any_alpha;  // some material can be alpha (colour.w == 0 or textured)
    const uint32_t    *model_first; // first primID of each model (for geomID)
    uint32_t           nmodels;
    f3                 bmin, bmax;  // scene bounding box
    // two-level scenes (instances): `nodes` holds the instance BVH (root 0) and every model
    // BVH; its leaves index tlas_inst, which names the instance; `tri` pools all models
    int32_t            instanced;
    const InstDev     *inst;
    const uint32_t    *tlas_inst;
    uint32_t           ninst;
};

// SoA ray buffer
struct RaySoA {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tn, *tf;
};
struct HitSoA {
    float    *t, *u, *v;
    uint32_t *tri;   // leaf-order triangle index or kInvalidRef
    uint32_t *inst;  // instance of the hit (two-level scenes; may be null)
};

// ---- builder ------------------------------------------------------------------------
struct BuildResult {
    void    *nodes;         // device
    float4  *tri;           // device, leaf order
    uint32_t node_count;
    uint32_t bvh2_nodes;
    uint32_t depth;
    uint32_t root;
    double   sah;
    float    ms;
    uint32_t invalid = 0;   // 1: the builder's own consistency checks failed (hipErrorUnknown)
};
// Builds the BVH of ntri triangles (9 floats each, device memory) on stream s.
// prim_cost: cost of a primitive test relative to a wide-node test in the collapse's cost
// model (0 = default).
hipError_t build_bvh(const float *d_tri, uint32_t ntri, uint32_t width, uint32_t max_leaf,
                     hipStream_t s, BuildResult *out, float prim_cost = 0.0f);
// Moves a wide BVH into a pool: inner child refs += node_base, leaf first += tri_base.
hipError_t bvh_rebase(void *nodes, uint32_t count, uint32_t width, uint32_t node_base, uint32_t tri_base,
                      hipStream_t s);
// Adds `base` to the primID stored in v0.w of every leaf-order triangle.
hipError_t tri_add_prim_base(float4 *tri, uint32_t ntri, uint32_t base, hipStream_t s);

// ---- traversal ----------------------------------------------------------------------
struct TraceScene;
struct TraceQueue {
    RaySoA          rays;
    HitSoA          hits;
    // 1: trace normalize(direction), as cr::scene::cast_ray does for every ray of the render
    // loop (model.cpp:109-110: glm::normalize(inv * dir) with the identity instance), so hit
    // distances are in units of the unit direction; 0: the direction as given (rtcIntersect1
    // semantics of crt_intersect / crt_occluded).
    uint32_t        unit_dir;
    const uint32_t *count;     // device: number of rays
    uint32_t       *counters;  // device: work counters (kXcds entries), zeroed by the caller's kernel
    uint32_t       *error;     // device: set to 1 if a traversal stack would overflow
    unsigned long long *tstats;  // device [5] or null: node visits, leaf visits, triangle tests,
                                 // wave iterations of the node loop and of the leaf loop
    // batched rendering: the scene of ray i is scenes[path_job[ray_path[i]]]
    const uint32_t   *ray_path;
    const uint32_t   *path_job;
    const TraceScene *scenes;
    // Epilogue (done != null): the last block to finish zeroes the work counters, the words
    // clear0 / clear1 (when set; never `count`), adds the ray count to *stat (when set) and
    // resets *done, which must be 0 at launch.  It prepares the next bounce's counters
    // without a kernel of its own.
    uint32_t           *done;
    uint32_t           *clear0, *clear1;
    unsigned long long *stat;
    // set by the launcher from TraceConfig: idle lanes that trigger a refill, and the number
    // of lanes still able to park a leaf below which a wave leaves the inner-node phase
    uint32_t refill_idle, park_min;
    uint32_t leaf_rounds;   // leaf-loop iterations per leaf phase before the wave resumes the
                            // inner-node phase (0: until no lane holds a leaf)
};
// Launch shape of the persistent traversal: blocks in the grid, LDS stack entries per lane
// (8, 16 or 32), and the refill / parking thresholds (0 = defaults).
struct TraceConfig {
    uint32_t grid;
    uint32_t stack;
    uint32_t refill = 0;
    uint32_t park_min = 0;
    uint32_t leaf_rounds = 0;
};
hipError_t launch_trace_closest(const SceneDev &sc, const TraceQueue &q, uint32_t *spill, const TraceConfig &tc,
                                hipStream_t s);
hipError_t launch_trace_any(const SceneDev &sc, const TraceQueue &q, uint8_t *occ, uint32_t *spill,
                            const TraceConfig &tc, hipStream_t s);
// batched: every ray carries its scene (q.ray_path / q.path_job / q.scenes); all scenes
// share one BVH width
hipError_t launch_trace_closest_multi(uint32_t width, const TraceQueue &q, uint32_t *spill, const TraceConfig &tc,
                                      hipStream_t s);
hipError_t launch_trace_any_multi(uint32_t width, const TraceQueue &q, uint8_t *occ, uint32_t *spill,
                                  const TraceConfig &tc, hipStream_t s);
// Ray-queue reordering (raysort.hip): keys, values and radix-sort temporary storage.
struct RaySortScratch {
    uint32_t *keys, *keys_out, *vals, *vals_out;
    void     *temp;
    size_t    temp_bytes;
    uint32_t  capacity;
};
size_t     ray_sort_temp_bytes(uint32_t capacity);
// Sorts the n rays of src by (origin Morton code in [bmin, bmax], direction octant) into
// dst (rays, path ids, count).
hipError_t launch_ray_sort(const RaySoA &src, const uint32_t *src_path, const uint32_t *src_count, uint32_t n,
                           const RaySoA &dst, uint32_t *dst_path, uint32_t *dst_count, f3 bmin, f3 bmax,
                           const RaySortScratch &w, hipStream_t s);
// Post-processing (post.hip): image size, the reference's dispatch coverage and settings.
struct PostDev {
    uint32_t w, h, cover_w, cover_h;
    int32_t  bloom;
    float    threshold, strength;
    int32_t  gray, tonemap, tonemap_type;
    float    exposure, gamma;
};
hipError_t launch_post(const PostDev &p, const float4 *in, float4 *out, float4 *t0, float4 *t1, hipStream_t s);
hipError_t launch_atrous(const float4 *colour, const float4 *normal, const float4 *albedo, float4 *out, float4 *t0,
                         float4 *t1, uint32_t w, uint32_t h, int levels, float sc, float sn, float sa, hipStream_t s);
uint32_t trace_grid(int device, uint32_t waves_per_cu);
uint64_t trace_spill_bytes(uint32_t grid);
constexpr uint32_t kTraceBlock = 256;
constexpr uint32_t kXcds       = 8;
// default LDS stack entries per lane: 16 entries (34.3 KB per 256-lane block) let the
// traversals of both wavefronts in flight hold their blocks on a CU at once (4 x 34.3 KB of
// 160 KB), where 32 entries (67.6 KB per block) admit only one traversal's two blocks
constexpr uint32_t kStackLds   = 16;
constexpr uint32_t kSpillDepth = 256;   // global spill entries per lane (multiple of kStackLds)

// ---- allocation-free queries (noalloc.hip) ---------------------------------------------
// Layout of the per-device scratch array (byte offsets): counters (count, work counters at
// +8, error flag at +16), leaf-order triangle and instance indices, u and v for callers that
// pass none, and the traversal spill stack of a kNoAllocGrid-block launch.
constexpr uint32_t kNoAllocRays     = 1u << 20;   // rays per crt_intersect_noalloc call
constexpr uint32_t kNoAllocGrid     = 64;         // traversal blocks (a multiple of kXcds = 8)
constexpr uint32_t kNoAllocCounters = 64;
constexpr size_t   kNaCnt   = 0;
constexpr size_t   kNaTri   = 256;
constexpr size_t   kNaInst  = kNaTri + 4ull * kNoAllocRays;
constexpr size_t   kNaU     = kNaInst + 4ull * kNoAllocRays;
constexpr size_t   kNaV     = kNaU + 4ull * kNoAllocRays;
constexpr size_t   kNaSpill = kNaV + 4ull * kNoAllocRays;
constexpr size_t   kNoAllocBytes = kNaSpill + (size_t)kNoAllocGrid * kTraceBlock * kSpillDepth * 2u * sizeof(uint32_t);
// Device address of this device's scratch array (looked up once per device).
hipError_t noalloc_scratch(int device, void **ptr);
hipError_t launch_noalloc_init(uint32_t *cnt, uint32_t n, hipStream_t s);

// ---- integrator ---------------------------------------------------------------------
enum : uint32_t { kModeCompat = 0, kModePbr = 1 };

struct PathBuffers {
    uint32_t  W, H, P;          // image size, path slots (camera paths, then cache-training paths)
    uint32_t  npix;             // W*H
    uint32_t  ncam;             // camera path slots: npix per sample pass a wavefront can carry;
                                // slots from here on train the cache
    RaySoA    q[2];             // extension queues (ping-pong)
    uint32_t *qpath[2];         // path slot of each queued ray
    HitSoA    hits;
    RaySoA    sq;               // shadow queue
    uint32_t *spath;
    float4   *scontrib;         // radiance added if the shadow ray escapes
    uint8_t  *occ;
    HitSoA    shits;            // closest hits of shadow rays (alpha pass-through)
    float4   *thr, *rad;
    float    *raw;              // W*H*3 running sum, flipped as renderer.cpp:358-365
    float4   *aov_alb, *aov_nrm, *aov_dep;
    uint32_t *cnt;              // see Cnt below
    unsigned long long *stats;  // [kStatSlots][4]: extension rays, shadow rays, reference segments (summed on read)
    // neural radiance cache
    float4   *nstate;           // per path: (a0 footprint, pdf of the last scatter, -, -)
    float    *nq;               // cache queries [P][13]
    uint32_t *nq_path;
    float4   *nq_thr;           // throughput arriving at the query vertex
    float    *nq_out;           // predictions [P][3]
    float    *rq;               // training records: queries [P * rec_slots][13]
    float4   *r_thr, *r_rad;    // throughput and radiance before each recorded vertex
    float    *tq, *tt;          // compacted training set: queries [n][13], targets [n][3]
    // batched rendering: job, pixel and absolute sample index of each path slot
    uint32_t *pjob, *ppix, *psamp;
    // material-sorted shading (CRT_FLAG_SORT_MATERIALS): queue order grouped by material
    // class, and per-block class counts [kShadeClasses][blocks(P)]
    uint32_t *shade_perm, *shade_bcount;
    // CRT_FLAG_TRAVERSAL_STATS: node visits, leaf visits, triangle tests, wave loop iterations
    unsigned long long *tstats;
};
constexpr uint32_t kStatSlots = 64;
// layout of PathBuffers::cnt
enum : uint32_t {
    kCntQ0 = 0, kCntQ1 = 1, kCntShadow = 2, kCntShadow2 = 3, kCntErr = 4, kCntNrcQ = 5, kCntTrain = 6,
    kCntNrcBad = 7,   // a cache prediction of the wavefront was not finite: its accumulation is skipped
    kCntWorkExt = 8, kCntWorkShadow = 16, kCntDoneExt = 24, kCntDoneShadow = 25, kCntWords = 32
};

// Neural radiance cache hooks of the integrator.  Camera paths (slots < npix) query the
// cache at `term` (footprint rule); training paths (slots >= npix) start at random pixels,
// record vertices rec_first .. rec_first + rec_slots - 1 and query at train_term.
struct NrcParams {
    int32_t  enabled;       // terminate paths into the cache
    uint32_t term;          // bounce at which camera paths query the cache
    float    footprint_c;   // > 0: query at `term` only if the path spread exceeds c * a0, else at term + 1
    uint32_t train_term;    // bounce at which training paths query the cache
    uint32_t rec_first, rec_slots;
    uint32_t train_live;    // training path slots in use this wavefront (from ncam on)
    f3       bmin, binv;    // scene box, for normalising query positions to [0,1]
};

struct RenderParams {
    camera_dev cam;
    float      aspect;
    uint32_t   max_bounces, rr_start, mode, seed, sample;
    // sample passes of the image in the wavefront (0: cache training only).  Camera slot p
    // is pixel p / nsamp, sample `sample + p % nsamp`: the passes of a pixel sit side by side.
    uint32_t   nsamp = 1;
    NrcParams  nrc;
};

// ---- batched rendering ----------------------------------------------------------------
// One job of a batch: a committed scene rendered at its own resolution, sample count and
// depth into its own accumulation buffer.
struct JobDev {
    SceneDev     sc;
    RenderParams rp;       // rp.sample: absolute sample index of the job's first pass
    uint32_t     W, H;
    float       *raw;      // W*H*3 running sum, flipped in x and y like the renderer's
};
// Traversal view of a job's scene.
struct TraceScene {
    const void   *nodes;
    const float4 *tri;
    uint32_t      root, ntri;
};
// A segment of a chunk: passes s0 .. s0+nsamp-1 of pixels pix0 .. pix0+npix-1 of one job,
// laid out sample-major from path slot path0; acc0 = first accumulation thread.
struct BatchSeg {
    uint32_t job, s0, nsamp, pix0, npix, path0, acc0, pad;
};
// One chunk = one replay of the pass graph: segments seg0 .. seg0+nseg-1.
struct BatchChunk {
    uint32_t seg0, nseg, npaths, nacc;
};
struct BatchDev {
    const JobDev     *jobs;
    const BatchSeg   *segs;
    const BatchChunk *chunks;
    const uint32_t   *cursor;    // device: index of the chunk being rendered
};

hipError_t launch_generate(const SceneDev &sc, const PathBuffers &b, const RenderParams &rp, hipStream_t s);
hipError_t launch_bounce_prep(const PathBuffers &b, uint32_t cur, hipStream_t s);
hipError_t launch_shade(const SceneDev &sc, const PathBuffers &b, const RenderParams &rp, uint32_t bounce,
                        uint32_t cur, hipStream_t s, const uint32_t *perm = nullptr);
// Material-sorted shading: classes miss / metal / smooth / glass of the hits of queue cur,
// a stable counting sort of the queue indices by class into b.shade_perm (device-side
// count, no host synchronisation); launch_shade then walks the queue in that order.
constexpr uint32_t kShadeClasses = 4;
hipError_t launch_shade_sort(const SceneDev &sc, const PathBuffers &b, uint32_t cur, hipStream_t s);
hipError_t launch_shadow_prep(const PathBuffers &b, hipStream_t s);
hipError_t launch_shadow_resolve(const PathBuffers &b, hipStream_t s);
hipError_t launch_shadow_alpha_step(const SceneDev &sc, const PathBuffers &b, hipStream_t s);
hipError_t launch_accumulate(const PathBuffers &b, uint32_t nsamp, hipStream_t s);
hipError_t launch_display(const float *raw, uint64_t n, uint32_t npix, float4 *out, hipStream_t s);
hipError_t launch_fill(float4 *p, uint64_t n, float v, hipStream_t s);
hipError_t launch_nrc_apply(const PathBuffers &b, hipStream_t s);
hipError_t launch_nrc_targets(const PathBuffers &b, uint32_t rec_slots, uint32_t max_out, uint32_t train_live,
                              hipStream_t s);
hipError_t nrc_launch_infer_dev(const float *q, float *out, const uint32_t *n_dev, uint32_t n_max, const __bf16 *WT,
                                hipStream_t s);
hipError_t launch_batch_generate(const BatchDev &bd, const PathBuffers &b, hipStream_t s);
hipError_t launch_shade_multi(const BatchDev &bd, const PathBuffers &b, uint32_t bounce, uint32_t cur, hipStream_t s);
hipError_t launch_shadow_alpha_step_multi(const BatchDev &bd, const PathBuffers &b, hipStream_t s);
hipError_t launch_batch_accumulate(const BatchDev &bd, const PathBuffers &b, uint32_t max_acc, hipStream_t s);
hipError_t launch_batch_advance(uint32_t *cursor, uint32_t step, hipStream_t s);
// Intersection records (crt_records): every output pointer may be null except dist.
struct RecordsDev {
    float    *dist;
    uint32_t *prim, *geom, *mat;
    float    *u, *v, *nx, *ny, *nz, *px, *py, *pz;
};
hipError_t launch_hit_records(const SceneDev &sc, const RaySoA &rs, const HitSoA &h, uint32_t n, const RecordsDev &out,
                              hipStream_t s);
hipError_t launch_fill_f32(float *p, uint32_t n, float v, hipStream_t s);
hipError_t launch_hits_to_api(const SceneDev &sc, const HitSoA &h, uint32_t n, uint32_t *prim, uint32_t *geom,
                              float *ngx, float *ngy, float *ngz, uint32_t *inst, hipStream_t s);

}  // namespace crt
