// integrator.hip -- wavefront path-tracing kernels of the crt path tracer (gfx950).
//
// The reference runs CRender src/render/renderer.cpp:258-384 (_sample_pixel) per pixel on
// CPU threads.  Here one sample pass is a sequence of kernels over SoA queues in HBM:
//   generate  camera rays for all W*H pixels (renderer.cpp:260-263, camera.cpp:14-39)
//   per bounce i < max_bounces:
//     extend  closest hit of every queued ray (traverse.hip)
//     shade   miss -> sky; hit -> process_hit (renderer.cpp:21-102), emission, AOVs on
//             bounce 0, sun NEE ray (renderer.cpp:315-354), Russian roulette, and the
//             next extension ray, appended with one atomic per wave
//     shadow  any hit of the sun rays; unoccluded rays add their contribution
//   accumulate  raw[flip(p)] += radiance (renderer.cpp:358-365)
// Each path owns one slot (its pixel), so no two threads ever add into the same pixel and
// the image is deterministic whatever order the queues are processed in.
#include <hip/hip_runtime.h>
#include <float.h>
#include "crt_internal.h"

namespace crt {

namespace {

constexpr float kInf = __builtin_huge_valf();

__device__ __forceinline__ uint32_t lane() { return __lane_id(); }

// Append to a queue with one atomic per 256-thread block: returns this lane's slot.
// Every thread of the block must call it (it contains barriers).
__device__ __forceinline__ uint32_t block_append(uint32_t *counter, bool want)
{
    __shared__ uint32_t wcnt[4];
    __shared__ uint32_t base;
    const uint64_t m = __ballot(want);
    const uint32_t w = threadIdx.x >> 6;
    if (lane() == 0) wcnt[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        base = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    uint32_t off = base;
    for (uint32_t j = 0; j < w; ++j) off += wcnt[j];
    const uint32_t slot = off + (uint32_t)__popcll(m & ((1ull << lane()) - 1ull));
    __syncthreads();   // wcnt / base reused by the next call
    return slot;
}

// Block reduction of a per-thread count, added to one of kStatSlots striped slots so
// that concurrent blocks rarely meet on the same address.  Every thread must call it.
__device__ __forceinline__ void block_add_stat(unsigned long long *stats, uint32_t which, uint32_t v)
{
    __shared__ uint32_t wsum[4];
    uint32_t s = v;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane() == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (tot) atomicAdd(&stats[(blockIdx.x % kStatSlots) * 4 + which], (unsigned long long)tot);
    }
    __syncthreads();
}

// image.h:104-109: nearest texel, static_cast<uint64_t>(u * width) % width.  The cast of a
// negative product is undefined in C++; this follows what x86-64 code without AVX-512 does
// (the signed truncation reinterpreted as uint64), so u * width in (-1, 0] maps to column 0
// and -1.5 to (2^64 - 1) % width.  For power-of-two sizes that is the positive modulo.
__device__ __forceinline__ uint64_t texel_index(float u, uint32_t size)
{
    return (uint64_t)(int64_t)(u * (float)size) % (uint64_t)size;
}

__device__ __forceinline__ float4 tex_uv(const SceneDev &sc, int32_t t, float u, float v)
{
    const TextureDev td = sc.texs[t];
    const uint64_t   x = texel_index(u, td.w), y = texel_index(v, td.h);
    return sc.texels[td.offset + x + y * (uint64_t)td.w];
}

// renderer.cpp:277-289, scene.cpp:67-77
__device__ __forceinline__ f3 sky(const SceneDev &sc, f3 d)
{
    if (sc.sky_tex < 0) return mk3(0.0f, 0.0f, 0.0f);
    const float dy = fminf(fmaxf(d.y, -1.0f), 1.0f);
    const float u  = 0.5f + atan2_cr(d.z, d.x) * kInvTau;
    const float v  = 0.5f - asin_cr(dy) * kInvPi;
    const float4 c = tex_uv(sc, sc.sky_tex, u + sc.sky_rot[0], v + sc.sky_rot[1]);
    return mk3(c.x, c.y, c.z);
}

__device__ __forceinline__ float4 mat_colour(const SceneDev &sc, const MaterialDev &m, float u, float v)
{
    if (m.tex >= 0 && (uint32_t)m.tex < sc.ntex) return tex_uv(sc, m.tex, u, v);
    return make_float4(m.colour[0], m.colour[1], m.colour[2], m.colour[3]);
}

// sampling.h:53-57
__device__ __forceinline__ f3 sun_radiance(const SunDev &s, f3 dir)
{
    const float c   = fminf(fmaxf(dot(dir, -s.dir), -1.0f), 1.0f);
    const float ang = acos_cr(c);
    return ang < s.size ? s.colour * s.intensity : mk3(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ void aov_write(const PathBuffers &b, uint32_t p, f3 alb, f3 nrm, float dep)
{
    if (!b.aov_alb) return;   // batched jobs keep no AOVs
    const uint32_t x = p % b.W, y = p / b.W;
    const uint32_t q = (b.W - 1 - x) + (b.H - 1 - y) * b.W;
    b.aov_alb[q] = make_float4(alb.x, alb.y, alb.z, 1.0f);
    b.aov_nrm[q] = make_float4(nrm.x * 0.5f + 0.5f, nrm.y * 0.5f + 0.5f, nrm.z * 0.5f + 0.5f, 1.0f);
    const float dd = fminf(dep, 200.0f) / 200.0f;
    b.aov_dep[q] = make_float4(dd, dd, dd, 1.0f);
}

__device__ __forceinline__ void put_ray(const RaySoA &q, uint32_t k, f3 o, f3 d)
{
    q.ox[k] = o.x;
    q.oy[k] = o.y;
    q.oz[k] = o.z;
    q.dx[k] = d.x;
    q.dy[k] = d.y;
    q.dz[k] = d.z;
    q.tn[k] = 1e-5f;              // model.cpp:22
    q.tf[k] = kInf;               // model.cpp:23
}

// Rays of the wavefront: rp.nsamp camera rays per pixel (slot p = pixel * nsamp + j, sample
// rp.sample + j), then rp.nrc.train_live cache-training paths (slots ncam ..), which start at
// random pixels.  Queue 0 holds them in that order, without the unused slots between.
__global__ __launch_bounds__(256) void k_generate(SceneDev sc, PathBuffers b, RenderParams rp)
{
    const uint32_t ns     = rp.nsamp ? rp.nsamp : 1u;
    const uint32_t ncl    = b.npix * rp.nsamp;                           // camera slots in use
    const uint32_t ntrain = rp.nrc.enabled ? rp.nrc.train_live : 0u;
    const uint32_t t      = blockIdx.x * blockDim.x + threadIdx.x;      // queue slot
    const bool     live   = t < ncl + ntrain;
    const bool     trainp = t >= ncl;
    if (live) {
        const uint32_t p = trainp ? b.ncam + (t - ncl) : t;
        // RNG key: the pixel of a camera path; npix + its index for a training path
        const uint32_t key    = trainp ? b.npix + (t - ncl) : p / ns;
        const uint32_t sample = trainp ? (rp.sample | 0x80000000u) : rp.sample + p % ns;
        const uint32_t pix = trainp ? philox4x32_10(u4{key, sample, 0xFFFFFFFEu, 0u}, rp.seed, kKeyHi).x % b.npix : key;
        const uint32_t x = pix % b.W, y = pix / b.W;
        const rnd4     j = rng4(rp.seed, key, sample, kJitterBounce, 0);
        if (trainp)
            for (uint32_t r = 0; r < rp.nrc.rec_slots; ++r)
                b.r_thr[(size_t)(p - b.ncam) * rp.nrc.rec_slots + r] = make_float4(0, 0, 0, 0);
        if (rp.nrc.enabled) b.nstate[p] = make_float4(0, 0, 0, 0);
        f3 o, d;
        // renderer.cpp:260-263
        camera_ray(rp.cam, ((float)x + j.a) / (float)b.W, ((float)y + j.b) / (float)b.H, rp.aspect, o, d);
        put_ray(b.q[0], t, o, d);
        b.qpath[0][t] = p;
        b.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        b.rad[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (rp.max_bounces == 0 && !trainp && p % ns == ns - 1u) aov_write(b, pix, mk3(0, 0, 0), mk3(0, 0, 0), 0.0f);
    }
    if (blockIdx.x == 0) {
        // the pass's counters: queue 0, the cache queries, and the traversal work counters and
        // finished-block counts (the traversal epilogues clear the rest bounce by bounce)
        if (threadIdx.x < 8) {
            b.cnt[kCntWorkExt + threadIdx.x]    = 0;
            b.cnt[kCntWorkShadow + threadIdx.x] = 0;
        }
        if (threadIdx.x == 0) {
            b.cnt[kCntQ0]          = rp.max_bounces > 0 ? ncl + ntrain : 0u;
            b.cnt[kCntNrcQ]        = 0;
            b.cnt[kCntNrcBad]      = 0;
            b.cnt[kCntDoneExt]     = 0;
            b.cnt[kCntDoneShadow]  = 0;
        }
    }
    // every path starts at total_bounces = 1 (renderer.cpp:271)
    block_add_stat(b.stats, 2, live && !trainp ? 1u : 0u);
}

// Before bounce i of a batch: clear the next queue, the shadow queue and the traversal work
// counters; count the extension rays about to be traced.  (The single-scene pass leaves this
// to the epilogue of the previous traversal launch.)
__global__ void k_bounce_prep(PathBuffers b, uint32_t cur)
{
    const uint32_t t = threadIdx.x;
    if (t < 8) {
        b.cnt[kCntWorkExt + t]    = 0;
        b.cnt[kCntWorkShadow + t] = 0;
    }
    if (t == 0) {
        b.cnt[cur ^ 1u]       = 0;
        b.cnt[kCntShadow]     = 0;
        atomicAdd(&b.stats[0], (unsigned long long)b.cnt[cur]);
    }
}

__global__ void k_shadow_prep(PathBuffers b)
{
    const uint32_t t = threadIdx.x;
    if (t < 8) b.cnt[kCntWorkShadow + t] = 0;
    if (t == 0) {
        b.cnt[kCntShadow2] = 0;
        atomicAdd(&b.stats[1], (unsigned long long)b.cnt[kCntShadow]);
    }
}

struct HitRec {
    f3       point, normal;
    float    dist;
    float    u, v;      // texture coordinates
    uint32_t prim;
};

// model.cpp:5-49 + model.cpp:99-126 for the identity instance
// inst: instance of the hit in a two-level scene (kInvalidRef otherwise).  The normal is
// the object-space one, as model.cpp:36-38 returns it for instances; world_normal maps it
// to world space (inverse transpose), which the pbr estimator uses.
__device__ __forceinline__ HitRec hit_record(const SceneDev &sc, f3 o, f3 d_in, float t, uint32_t tri, float bu,
                                             float bv, uint32_t inst = kInvalidRef, bool world_normal = false)
{
    HitRec h;
    const float4 a = sc.tri[3 * (size_t)tri + 0];
    const float4 b = sc.tri[3 * (size_t)tri + 1];
    const float4 c = sc.tri[3 * (size_t)tri + 2];
    h.prim   = __float_as_uint(a.w);
    const f3 d = normalize(d_in);
    h.point  = o + d * t;
    const f3 pd = h.point - o;
    h.dist   = sqrtf(dot(pd, pd));
    const f3 v0 = mk3(a.x, a.y, a.z), v1 = mk3(b.x, b.y, b.z), v2 = mk3(c.x, c.y, c.z);
    h.normal = normalize(cross(v1 - v0, v2 - v0));
    if (world_normal && inst != kInvalidRef) {
        const float *m = sc.inst[inst].inv;
        const f3     n = h.normal;
        h.normal = normalize(mk3(m[0] * n.x + m[4] * n.y + m[8] * n.z, m[1] * n.x + m[5] * n.y + m[9] * n.z,
                                 m[2] * n.x + m[6] * n.y + m[10] * n.z));
    }
    if (sc.uv) {
        const float2 w0 = sc.uv[3 * (size_t)h.prim + 0], w1 = sc.uv[3 * (size_t)h.prim + 1],
                     w2 = sc.uv[3 * (size_t)h.prim + 2];
        const float b0 = 1.0f - bu - bv;
        h.u = b0 * w0.x + bu * w1.x + bv * w2.x;
        h.v = b0 * w0.y + bu * w1.y + bv * w2.y;
    } else {
        h.u = h.v = 0.0f;
    }
    return h;
}

// Termination into the cache (Mueller et al. 2021, "Real-time Neural Radiance Caching
// for Path Tracing"): query at bounce `term`, unless the footprint heuristic is on and the
// path spread |x_{i-1} - x_i|^2 / (pdf |cos|) is still below c * a0, in which case the
// query moves to the next bounce.
__device__ __forceinline__ bool nrc_terminates(const NrcParams &np, const PathBuffers &b, uint32_t p, uint32_t bounce,
                                               const struct HitRec &h, f3 d)
{
    if (p >= b.ncam) return bounce >= np.train_term;   // training path
    if (bounce < np.term) return false;
    if (np.footprint_c <= 0.0f || bounce > np.term) return true;
    const float4 st  = b.nstate[p];
    const float  pdf = st.y;
    if (!(pdf > 0.0f)) return false;            // delta lobe: footprint unchanged
    const float c = fmaxf(fabsf(dot(h.normal, normalize(d))), 1e-4f);
    const float spread = h.dist * h.dist / (pdf * c);
    return spread > np.footprint_c * st.x;
}

// Cache query of a vertex: position normalised to the scene box, outgoing direction,
// normal facing the incoming ray, albedo, roughness (1 diffuse, alpha for metal, 0 glass).
__device__ __forceinline__ void nrc_query(const NrcParams &np, f3 point, f3 d, f3 n, f3 cc, const MaterialDev &m,
                                          float *q)
{
    const f3 pn = (point - np.bmin) * np.binv;
    const f3 wo = -normalize(d);
    const f3 nf = dot(d, n) > 0.0f ? -n : n;
    q[0]  = fminf(fmaxf(pn.x, 0.0f), 1.0f);
    q[1]  = fminf(fmaxf(pn.y, 0.0f), 1.0f);
    q[2]  = fminf(fmaxf(pn.z, 0.0f), 1.0f);
    q[3]  = wo.x;
    q[4]  = wo.y;
    q[5]  = wo.z;
    q[6]  = nf.x;
    q[7]  = nf.y;
    q[8]  = nf.z;
    q[9]  = cc.x;
    q[10] = cc.y;
    q[11] = cc.z;
    q[12] = m.type == 1 ? 1.0f : (m.type == 0 ? m.roughness : 0.0f);
}

// Output of shading one path vertex: the rays and cache query it pushes.
struct ShadeOut {
    bool     push_next, push_shadow, do_query;
    f3       no, nd, so, sd, scontrib, qthr;
    float    nq[13];
    uint32_t seg;
};

// True for the camera slot whose AOVs the pass writes: the wavefront's last sample of its
// pixel (the reference keeps the last pass's AOVs).
__device__ __forceinline__ bool aov_slot(const PathBuffers &b, const RenderParams &rp, uint32_t p)
{
    return p < b.ncam && (rp.nsamp <= 1u || p % rp.nsamp == rp.nsamp - 1u);
}

// Shades queue entry k (path slot p, pixel pix, absolute sample index `sample`) at `bounce`:
// renderer.cpp:273-356 (compat) or the physically based variant (pbr).
__device__ __forceinline__ void shade_path(const SceneDev &sc, const PathBuffers &b, const RenderParams &rp, uint32_t k,
                                           uint32_t p, uint32_t pix, uint32_t sample, uint32_t bounce, uint32_t cur,
                                           ShadeOut &out)
{
    const RaySoA &q = b.q[cur];
    bool &push_next = out.push_next, &push_shadow = out.push_shadow, &do_query = out.do_query;
    f3 &no = out.no, &nd = out.nd, &so = out.so, &sd = out.sd, &scontrib = out.scontrib, &qthr = out.qthr;
    float (&nq)[13] = out.nq;
    uint32_t &seg = out.seg;
    const f3       o   = mk3(q.ox[k], q.oy[k], q.oz[k]);
    const f3       d   = mk3(q.dx[k], q.dy[k], q.dz[k]);
    const uint32_t tri = b.hits.tri[k];
    float4         thr4 = b.thr[p];
    f3             thr  = mk3(thr4.x, thr4.y, thr4.z);
    float4         rad4 = b.rad[p];
    f3             rad  = mk3(rad4.x, rad4.y, rad4.z);
    const bool     can_continue = bounce + 1 < (p >= b.ncam ? rp.nrc.train_term + 1 : rp.max_bounces);
    qthr = thr;
    if (tri == kInvalidRef) {
        const f3 s = sky(sc, d);
        if (bounce == 0 && aov_slot(b, rp, p)) aov_write(b, pix, s, mk3(0, 0, 0), 0.0f);
        rad = rad + thr * s;
    } else {
        const uint32_t     inst = sc.instanced && b.hits.inst ? b.hits.inst[k] : kInvalidRef;
        const HitRec       h = hit_record(sc, o, d, b.hits.t[k], tri, b.hits.u[k], b.hits.v[k], inst, rp.mode == kModePbr);
        const MaterialDev &m = sc.mats[sc.mat_idx[h.prim]];
        const float4       col = mat_colour(sc, m, h.u, h.v);
        seg = 1;
        if (col.w == 0.0f) {
            // alpha pass-through (renderer.cpp:294-301)
            if (bounce == 0 && aov_slot(b, rp, p)) aov_write(b, pix, mk3(0, 0, 0), mk3(0, 0, 0), 0.0f);
            no = h.point + d * 0.1f;
            nd = d;
            push_next = can_continue;
        } else if (rp.nrc.enabled && nrc_terminates(rp.nrc, b, p, bounce, h, d)) {
            // the cache stands in for this vertex's emission, direct light and everything after it
            const f3 cc = mk3(col.x, col.y, col.z);
            if (bounce == 0 && aov_slot(b, rp, p)) aov_write(b, pix, cc, h.normal, h.dist);
            nrc_query(rp.nrc, h.point, d, h.normal, cc, m, nq);
            do_query = true;
        } else {
            const rnd4 r  = rng4(rp.seed, pix, sample, bounce, 0);
            const f3   n  = h.normal;
            const f3   cc = mk3(col.x, col.y, col.z);
            if (p >= b.ncam && bounce >= rp.nrc.rec_first && bounce < rp.nrc.rec_first + rp.nrc.rec_slots) {
                const size_t slot = (size_t)(p - b.ncam) * rp.nrc.rec_slots + (bounce - rp.nrc.rec_first);
                nrc_query(rp.nrc, h.point, d, n, cc, m, b.rq + slot * 13);
                b.r_thr[slot] = make_float4(thr.x, thr.y, thr.z, 1.0f);
                b.r_rad[slot] = make_float4(rad.x, rad.y, rad.z, 0.0f);
            }
            if (rp.nrc.enabled && bounce == 0) {
                // primary footprint a0 = |x1 - x0|^2 / (4 pi |cos|)
                const float c0 = fabsf(dot(n, normalize(d)));
                b.nstate[p].x = h.dist * h.dist / (4.0f * kPi * fmaxf(c0, 1e-4f));
            }
            f3   alb = cc, nee_n = n, f0 = mk3(0, 0, 0);
            bool delta = false, absorbed = false, ggx = false;
            float ar = 0.0f;
            if (rp.mode == kModeCompat) {
                // renderer.cpp:45-99
                if (m.type == 2) {
                    f3    out_n = n;
                    const f3 reflected = reflect(d, n);
                    float ni = 1.0f / m.ior;
                    if (dot(d, n) > 0.0f) {
                        out_n = -n;
                        ni    = m.ior;
                    }
                    const f3    uvd  = normalize(d);
                    const float dt   = dot(uvd, out_n);
                    const float disc = 1.0f - ni * ni * (1.0f - dt * dt);
                    no = h.point + out_n * -0.0001f;
                    nd = disc > 0.0f ? (uvd - out_n * dt) * ni - out_n * sqrtf(disc) : reflected;
                } else if (m.type == 0) {
                    no  = h.point + n * 0.0001f;
                    nd  = reflect(d, n);
                    alb = alb * m.reflectiveness;
                } else {
                    const f3 sp = sphere(r.a, r.b);
                    no = h.point + n * 0.0001f;
                    nd = normalize(n + sp);
                }
            } else {
                const f3 nf = dot(d, n) > 0.0f ? -n : n;
                if (m.type == 2) {
                    delta = true;
                    f3    out_n = n;
                    float ni    = 1.0f / m.ior;
                    if (dot(d, n) > 0.0f) {
                        out_n = -n;
                        ni    = m.ior;
                    }
                    const f3    uvd  = normalize(d);
                    const float dt   = dot(uvd, out_n);
                    const float disc = 1.0f - ni * ni * (1.0f - dt * dt);
                    bool refl = true;
                    if (disc > 0.0f) {
                        const rnd4  r2  = rng4(rp.seed, pix, sample, bounce, 1);
                        const float cf  = ni <= 1.0f ? -dt : sqrtf(disc);
                        const float r0  = (1.0f - m.ior) / (1.0f + m.ior);
                        const float f0s = r0 * r0;
                        const float fr  = f0s + (1.0f - f0s) * pow_cr(1.0f - cf, 5.0f);
                        refl = r2.b < fr;
                    }
                    if (refl) {
                        no = h.point + out_n * 0.0001f;
                        nd = reflect(uvd, out_n);
                    } else {
                        no = h.point + out_n * -0.0001f;
                        nd = (uvd - out_n * dt) * ni - out_n * sqrtf(disc);
                    }
                } else if (m.type == 0) {
                    const float a  = m.roughness;
                    f0   = cc * m.reflectiveness;
                    const f3 wo = -normalize(d);
                    no    = h.point + nf * 0.0001f;
                    nee_n = nf;
                    if (a < 1e-3f) {
                        delta = true;
                        nd    = reflect(d, nf);
                        alb   = schlick(dot(wo, nf), f0);
                    } else {
                        ggx = true;
                        ar  = a;
                        f3 tb, bb;
                        onb(nf, tb, bb);
                        const f3 lv = mk3(dot(wo, tb), dot(wo, bb), dot(wo, nf));
                        // Heitz 2018, sampling the GGX distribution of visible normals
                        const f3    vh    = normalize(mk3(a * lv.x, a * lv.y, lv.z));
                        const float lensq = vh.x * vh.x + vh.y * vh.y;
                        const f3    t1v   = lensq > 0.0f ? mk3(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(lensq)) : mk3(1, 0, 0);
                        const f3    t2v   = cross(vh, t1v);
                        const float rr    = sqrtf(r.a);
                        const float phi   = kTau * r.b;
                        const float t1    = rr * cos_cr(phi);
                        float       t2    = rr * sin_cr(phi);
                        const float s     = 0.5f * (1.0f + vh.z);
                        t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
                        const float nz2 = fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2);
                        const f3 nh = t1v * t1 + t2v * t2 + vh * sqrtf(nz2);
                        const f3 ne = normalize(mk3(a * nh.x, a * nh.y, fmaxf(nh.z, 0.0f)));
                        const float voh = dot(lv, ne);
                        const f3 ll = ne * (2.0f * voh) - lv;
                        if (ll.z <= 0.0f || lv.z <= 0.0f) {
                            absorbed = true;
                            alb = mk3(0, 0, 0);
                            nd  = nf;
                        } else {
                            nd = normalize(tb * ll.x + bb * ll.y + nf * ll.z);
                            const float a2 = a * a;
                            const float g1 = 2.0f * lv.z / (lv.z + sqrtf(a2 + (1.0f - a2) * lv.z * lv.z));
                            const float g2 = ggx_vis(lv.z, ll.z, a) * 4.0f * lv.z * ll.z;
                            alb = schlick(voh, f0) * (g2 / g1);
                        }
                    }
                } else {
                    const f3 sp = sphere(r.a, r.b);
                    no    = h.point + nf * 0.0001f;
                    nd    = normalize(nf + sp);
                    nee_n = nf;
                }
            }
            if (bounce == 0 && aov_slot(b, rp, p)) aov_write(b, pix, rp.mode == kModeCompat ? alb : cc, n, h.dist);
            const f3 thr_prev = thr;
            if (rp.mode == kModeCompat) {
                thr = thr * alb;
                rad = rad + thr * m.emission;
            } else {
                rad = rad + (thr_prev * cc) * m.emission;
                thr = thr * alb;
            }
            // sun NEE (renderer.cpp:315-354)
            if (sc.sun_enabled && !(rp.mode == kModePbr && delta)) {
                const SunDev &sun = sc.sun;
                so = h.point + nee_n * 0.001f;
                const f3 nd_sun = -sun.dir;
                f3 tb, bb;
                onb(nd_sun, tb, bb);
                const float ct  = 1.0f - r.d * (1.0f - cos_cr(sun.size));
                const float st  = sqrtf(1.0f - ct * ct);
                const float phi = kTau * r.c;
                const f3    loc = mk3(cos_cr(phi) * st, ct, sin_cr(phi) * st);
                sd = tb * loc.x + nd_sun * loc.y + bb * loc.z;
                const float pdf  = 1.0f / (kTau * (1.0f - cos_cr(sun.size)));
                const float cosv = fminf(fmaxf(dot(nee_n, sd), 0.0f), 1.0f);
                if (cosv > 0.0f) {
                    const f3 L = sun_radiance(sun, sd);
                    if (rp.mode == kModeCompat) {
                        const f3 tc = (thr * cc) * cosv * L;
                        scontrib = mk3(tc.x / pdf, tc.y / pdf, tc.z / pdf);
                    } else if (ggx) {
                        const f3    wo  = -normalize(d);
                        const f3    hv  = normalize(wo + sd);
                        const float nov = dot(nee_n, wo);
                        if (nov > 0.0f) {
                            const float dv = ggx_d(dot(nee_n, hv), ar) * ggx_vis(nov, cosv, ar);
                            scontrib = ((thr_prev * schlick(dot(wo, hv), f0)) * L) * (dv * cosv / pdf);
                        }
                    } else {
                        scontrib = (thr * L) * (cosv / (pdf * kPi));
                    }
                    push_shadow = scontrib.x != 0.0f || scontrib.y != 0.0f || scontrib.z != 0.0f ||
                                  rp.mode == kModeCompat;
                }
            }
            if (absorbed) {
                seg = 0;
            } else {
                push_next = can_continue;
            }
            if (rp.nrc.enabled && rp.nrc.footprint_c > 0.0f) {
                // pdf of the scattered direction (0 for delta lobes: the footprint does not grow)
                float pdf = 0.0f;
                if (!delta && !absorbed) {
                    if (ggx) {
                        const f3    wo = -normalize(d);
                        const f3    hv = normalize(wo + nd);
                        const float nv = fmaxf(dot(nee_n, wo), 1e-4f);
                        const float a2 = ar * ar;
                        const float g1 = 2.0f * nv / (nv + sqrtf(a2 + (1.0f - a2) * nv * nv));
                        pdf = ggx_d(dot(nee_n, hv), ar) * g1 / (4.0f * nv);
                    } else {
                        pdf = fabsf(dot(rp.mode == kModeCompat ? n : nee_n, nd)) * kInvPi;
                    }
                }
                b.nstate[p].y = pdf;
            }
        }
    }
    // Russian roulette at the top of the next iteration
    if (push_next && rp.rr_start && bounce + 1 >= rp.rr_start) {
        const rnd4 r2 = rng4(rp.seed, pix, sample, bounce + 1, 1);
        const float qv = fminf(maxc(thr), 0.95f);
        if (r2.a >= qv) push_next = false;
        else thr = thr * (1.0f / qv);
    }
    b.thr[p] = make_float4(thr.x, thr.y, thr.z, 0.0f);
    b.rad[p] = make_float4(rad.x, rad.y, rad.z, 0.0f);
}

// One thread per queued extension ray.  kMulti: paths of many jobs (batched rendering),
// each path's job, pixel and sample index come from pjob / ppix / psamp.
template <bool kMulti>
__global__ __launch_bounds__(256) void k_shade(SceneDev sc, PathBuffers b, RenderParams rp, const JobDev *jobs,
                                               uint32_t bounce, uint32_t cur, const uint32_t *perm)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = b.cnt[cur];
    const bool     live = t < n;
    // perm: queue entries in material-class order (the queue index itself otherwise)
    const uint32_t k = live && perm ? perm[t] : t;
    ShadeOut o;
    o.push_next = o.push_shadow = o.do_query = false;
    o.no = o.nd = o.so = o.sd = o.scontrib = o.qthr = mk3(0, 0, 0);
    o.seg = 0;
    uint32_t p = 0;
    if (live) {
        p = b.qpath[cur][k];
        if constexpr (kMulti) {
            const JobDev &J = jobs[b.pjob[p]];
            shade_path(J.sc, b, J.rp, k, p, b.ppix[p], b.psamp[p], bounce, cur, o);
        } else {
            const uint32_t ns     = rp.nsamp ? rp.nsamp : 1u;
            const bool     trainp = p >= b.ncam;
            const uint32_t key    = trainp ? b.npix + (p - b.ncam) : p / ns;
            shade_path(sc, b, rp, k, p, key, trainp ? (rp.sample | 0x80000000u) : rp.sample + p % ns, bounce, cur, o);
        }
    }
    const uint32_t nqs = (!kMulti && rp.nrc.enabled) ? block_append(&b.cnt[kCntNrcQ], o.do_query) : 0u;
    if (o.do_query) {
        for (int k = 0; k < 13; ++k) b.nq[(size_t)nqs * 13 + k] = o.nq[k];
        b.nq_path[nqs] = p;
        b.nq_thr[nqs]  = make_float4(o.qthr.x, o.qthr.y, o.qthr.z, 0.0f);
    }
    const uint32_t ns = block_append(&b.cnt[kCntShadow], o.push_shadow);
    if (o.push_shadow) {
        put_ray(b.sq, ns, o.so, o.sd);
        b.spath[ns]    = p;
        b.scontrib[ns] = make_float4(o.scontrib.x, o.scontrib.y, o.scontrib.z, 0.0f);
    }
    const uint32_t nn = block_append(&b.cnt[cur ^ 1u], o.push_next);
    if (o.push_next) {
        put_ray(b.q[cur ^ 1u], nn, o.no, o.nd);
        b.qpath[cur ^ 1u][nn] = p;
    }
    block_add_stat(b.stats, 2, p < b.ncam ? o.seg : 0u);
}

__global__ __launch_bounds__(256) void k_shadow_resolve(PathBuffers b)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.cnt[kCntShadow]) return;
    if (b.occ[k]) return;
    const uint32_t p = b.spath[k];
    const float4   c = b.scontrib[k];
    float4         r = b.rad[p];
    r.x += c.x;
    r.y += c.y;
    r.z += c.z;
    b.rad[p] = r;
}

// True if the shadow ray k stopped on a transparent texel (colour.w == 0); o2 = where it continues.
__device__ __forceinline__ bool alpha_passes(const SceneDev &sc, const PathBuffers &b, uint32_t k, f3 o, f3 d,
                                             uint32_t tri, f3 &o2)
{
    const HitRec       h   = hit_record(sc, o, d, b.shits.t[k], tri, b.shits.u[k], b.shits.v[k]);
    const MaterialDev &m   = sc.mats[sc.mat_idx[h.prim]];
    const float4       col = mat_colour(sc, m, h.u, h.v);
    o2 = h.point + d * 0.1f;
    return col.w == 0.0f;
}

// Alpha-aware shadow step (renderer.cpp:330-345): after a closest-hit trace of the shadow
// queue, escaped rays add their contribution, rays stopped by an alpha surface continue
// from hit + 0.1*dir (compacted into the front of the same queue), others are blocked.
template <bool kMulti>
__global__ __launch_bounds__(256) void k_shadow_alpha(SceneDev sc_one, PathBuffers b, const JobDev *jobs)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = b.cnt[kCntShadow];
    bool     cont = false;
    f3       o2 = mk3(0, 0, 0), d = mk3(0, 0, 0);
    uint32_t p = 0;
    float4   c = make_float4(0, 0, 0, 0);
    if (k < n) {
        p = b.spath[k];
        c = b.scontrib[k];
        const uint32_t tri = b.shits.tri[k];
        const f3 o = mk3(b.sq.ox[k], b.sq.oy[k], b.sq.oz[k]);
        d = mk3(b.sq.dx[k], b.sq.dy[k], b.sq.dz[k]);
        if (tri == kInvalidRef) {
            float4 r = b.rad[p];
            r.x += c.x;
            r.y += c.y;
            r.z += c.z;
            b.rad[p] = r;
        } else {
            if constexpr (kMulti) cont = alpha_passes(jobs[b.pjob[p]].sc, b, k, o, d, tri, o2);
            else cont = alpha_passes(sc_one, b, k, o, d, tri, o2);
        }
    }
    // continuation rays go to a staging count, then the host swaps them in
    const uint32_t slot = block_append(&b.cnt[kCntShadow2], cont);
    if (cont) {
        // staged after the live entries; k_shadow_compact moves them to the front
        const uint32_t dst = n + slot;
        if (dst < b.P * 2u) {
            put_ray(b.sq, dst, o2, d);
            b.spath[dst]    = p;
            b.scontrib[dst] = c;
        }
    }
}

__global__ __launch_bounds__(256) void k_shadow_compact(PathBuffers b)
{
    const uint32_t n = b.cnt[kCntShadow];
    const uint32_t m = b.cnt[kCntShadow2];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) {
        const uint32_t s = n + k;
        const RaySoA  &q = b.sq;
        const float ox = q.ox[s], oy = q.oy[s], oz = q.oz[s], dx = q.dx[s], dy = q.dy[s], dz = q.dz[s];
        const uint32_t pp = b.spath[s];
        const float4   cc = b.scontrib[s];
        // source indices >= n and destinations < m <= n never overlap
        q.ox[k] = ox; q.oy[k] = oy; q.oz[k] = oz; q.dx[k] = dx; q.dy[k] = dy; q.dz[k] = dz;
        q.tn[k] = 1e-5f;
        q.tf[k] = kInf;
        b.spath[k]    = pp;
        b.scontrib[k] = cc;
    }
}

__global__ void k_shadow_swap(PathBuffers b)
{
    if (threadIdx.x == 0) {
        const uint32_t m = b.cnt[kCntShadow2];
        b.cnt[kCntShadow]  = m;
        b.cnt[kCntShadow2] = 0;
        atomicAdd(&b.stats[1], (unsigned long long)m);
    }
    if (threadIdx.x < 8) b.cnt[kCntWorkShadow + threadIdx.x] = 0;
}

// raw[pixel] += the radiance of each of the wavefront's samples of the pixel, in sample
// order: the same sums as one accumulation per pass.
__global__ __launch_bounds__(256) void k_accumulate(PathBuffers b, uint32_t nsamp)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= b.npix || b.cnt[kCntNrcBad]) return;
    const uint32_t x = p % b.W, y = p / b.W;
    const uint32_t q = (b.W - 1 - x) + (b.H - 1 - y) * b.W;
    float sx = b.raw[3 * (size_t)q + 0], sy = b.raw[3 * (size_t)q + 1], sz = b.raw[3 * (size_t)q + 2];
    for (uint32_t j = 0; j < nsamp; ++j) {
        const float4 r = b.rad[(size_t)p * nsamp + j];
        sx += r.x;
        sy += r.y;
        sz += r.z;
    }
    b.raw[3 * (size_t)q + 0] = sx;
    b.raw[3 * (size_t)q + 1] = sy;
    b.raw[3 * (size_t)q + 2] = sz;
}

// renderer.cpp:371-383: pow(clamp(sum / n, 0, 1), 1/2.2); FLT_MAX before the first pass
// (image.h:30-37)
__global__ __launch_bounds__(256) void k_display(const float *raw, uint64_t n, uint32_t npix, float4 *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    if (n == 0) {
        out[i] = make_float4(FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX);
        return;
    }
    const float inv = (float)n;
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = pow_cr(fminf(fmaxf(raw[3 * (size_t)i + k] / inv, 0.0f), 1.0f), 1.0f / 2.2f);
    out[i] = make_float4(c[0], c[1], c[2], 1.0f);
}

__global__ void k_fill(float4 *p, uint64_t n, float v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_float4(v, v, v, v);
}

__global__ void k_hits_to_api(SceneDev sc, HitSoA h, uint32_t n, uint32_t *prim, uint32_t *geom, float *ngx,
                              float *ngy, float *ngz, uint32_t *inst)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t tri = h.tri[i];
    uint32_t pid = kInvalidRef, gid = kInvalidRef;
    f3 ng = mk3(0, 0, 0);
    if (tri != kInvalidRef) {
        const float4 a = sc.tri[3 * (size_t)tri + 0];
        const float4 b = sc.tri[3 * (size_t)tri + 1];
        const float4 c = sc.tri[3 * (size_t)tri + 2];
        pid = __float_as_uint(a.w);
        const f3 v0 = mk3(a.x, a.y, a.z);
        ng = cross(mk3(b.x, b.y, b.z) - v0, mk3(c.x, c.y, c.z) - v0);
        gid = 0;
        for (uint32_t m = 1; m < sc.nmodels; ++m)
            if (pid >= sc.model_first[m]) gid = m;
        pid -= sc.model_first[gid];
    }
    prim[i] = pid;
    if (geom) geom[i] = gid;
    if (inst) inst[i] = tri == kInvalidRef ? kInvalidRef : (sc.instanced && h.inst ? sc.inst[h.inst[i]].local : 0u);
    if (ngx) ngx[i] = ng.x;
    if (ngy) ngy[i] = ng.y;
    if (ngz) ngz[i] = ng.z;
}

// model.cpp:27-47, and model.cpp:114-118 for a hit on an instance: the intersection record
// of each traced ray.  Every input of element i is read before any output is written, so the
// outputs may alias the inputs of the same element (host-staged calls reuse the buffers).
__global__ void k_hit_records(SceneDev sc, RaySoA rs, HitSoA h, uint32_t n, RecordsDev out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t tri  = h.tri[i];
    const f3       o    = mk3(rs.ox[i], rs.oy[i], rs.oz[i]);
    const f3       d    = mk3(rs.dx[i], rs.dy[i], rs.dz[i]);
    const float    t    = h.t[i], bu = h.u[i], bv = h.v[i];
    const uint32_t inst = sc.instanced && h.inst ? h.inst[i] : kInvalidRef;
    float    dist = __builtin_huge_valf(), u = 0.0f, v = 0.0f;
    uint32_t pid = kInvalidRef, gid = kInvalidRef, mat = kInvalidRef;
    f3       nrm = mk3(0, 0, 0), pt = mk3(0, 0, 0);
    if (tri != kInvalidRef) {
        const HitRec r = hit_record(sc, o, d, t, tri, bu, bv, inst, false);
        dist = r.dist;
        pt   = r.point;
        nrm  = r.normal;
        u    = r.u;
        v    = r.v;
        mat  = sc.mat_idx[r.prim];
        gid  = 0;
        for (uint32_t m = 1; m < sc.nmodels; ++m)
            if (r.prim >= sc.model_first[m]) gid = m;
        pid = r.prim - sc.model_first[gid];
    }
    out.dist[i] = dist;
    if (out.prim) out.prim[i] = pid;
    if (out.geom) out.geom[i] = gid;
    if (out.mat) out.mat[i] = mat;
    if (out.u) out.u[i] = u;
    if (out.v) out.v[i] = v;
    if (out.nx) out.nx[i] = nrm.x;
    if (out.ny) out.ny[i] = nrm.y;
    if (out.nz) out.nz[i] = nrm.z;
    if (out.px) out.px[i] = pt.x;
    if (out.py) out.py[i] = pt.y;
    if (out.pz) out.pz[i] = pt.z;
}

__global__ void k_fill_f32(float *p, uint32_t n, float v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// rad[path] += throughput * max(prediction, 0) for every cache query
__global__ __launch_bounds__(256) void k_nrc_apply(PathBuffers b)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.cnt[kCntNrcQ]) return;
    const uint32_t p = b.nq_path[k];
    const float4   t = b.nq_thr[k];
    float4         r = b.rad[p];
    const float    y0 = b.nq_out[3 * (size_t)k + 0], y1 = b.nq_out[3 * (size_t)k + 1], y2 = b.nq_out[3 * (size_t)k + 2];
    // a diverged cache (non-finite or exploding weights) shows here first: flag the wavefront
    // (its accumulation is skipped and the renderer continues without the cache)
    if (!(fabsf(y0) < 1e30f) || !(fabsf(y1) < 1e30f) || !(fabsf(y2) < 1e30f)) atomicOr(&b.cnt[kCntNrcBad], 1u);
    r.x += t.x * fmaxf(y0, 0.0f);
    r.y += t.y * fmaxf(y1, 0.0f);
    r.z += t.z * fmaxf(y2, 0.0f);
    b.rad[p] = r;
}

// Training targets from the recorded vertices: L_v = (L_path - L_before_v) / T_v per
// channel; valid, finite samples are compacted into (tq, tt).
__global__ __launch_bounds__(256) void k_nrc_targets(PathBuffers b, uint32_t rec_slots, uint32_t max_out,
                                                     uint32_t train_live)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t p = b.ncam + i / rec_slots;   // training path slot
    bool  ok = false;
    float tg[3] = {0, 0, 0};
    if (p < b.ncam + train_live) {
        const float4 t = b.r_thr[i];
        if (t.w > 0.0f) {
            const float4 rb = b.rad[p], r0 = b.r_rad[i];
            const float  thr[3] = {t.x, t.y, t.z}, num[3] = {rb.x - r0.x, rb.y - r0.y, rb.z - r0.z};
            ok = true;
            for (int c = 0; c < 3; ++c) {
                if (!(thr[c] > 1e-6f)) {
                    ok = false;
                    break;
                }
                tg[c] = fmaxf(num[c] / thr[c], 0.0f);
                if (!(tg[c] < 1e4f)) ok = false;
            }
        }
    }
    const uint32_t slot = block_append(&b.cnt[kCntTrain], ok);
    if (ok && slot < max_out) {
        for (int k = 0; k < 13; ++k) b.tq[(size_t)slot * 13 + k] = b.rq[(size_t)i * 13 + k];
        for (int c = 0; c < 3; ++c) b.tt[(size_t)slot * 3 + c] = tg[c];
    }
}

// ---- batched rendering ----------------------------------------------------------------

// Segment of the current chunk that holds `key` (a path slot, or an accumulation thread
// when by_acc): segments are stored in increasing path0 / acc0 order.
__device__ __forceinline__ uint32_t find_seg(const BatchSeg *segs, uint32_t seg0, uint32_t nseg, uint32_t key,
                                             bool by_acc)
{
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid   = (lo + hi) / 2;
        const uint32_t start = by_acc ? segs[seg0 + mid].acc0 : segs[seg0 + mid].path0;
        if (start <= key) lo = mid;
        else hi = mid;
    }
    return seg0 + lo;
}

// Camera rays of every path of the current chunk (renderer.cpp:260-263 per job), the same
// Philox stream as the single-scene renderer: (job seed, pixel, sample, jitter).
__global__ __launch_bounds__(256) void k_batch_generate(BatchDev bd, PathBuffers b)
{
    const BatchChunk ch   = bd.chunks[*bd.cursor];
    const uint32_t   p    = blockIdx.x * blockDim.x + threadIdx.x;
    const bool       live = p < ch.npaths && p < b.P;
    if (live) {
        const BatchSeg sg     = bd.segs[find_seg(bd.segs, ch.seg0, ch.nseg, p, false)];
        const uint32_t local  = p - sg.path0;
        const uint32_t si     = local / sg.npix;
        const uint32_t pix    = sg.pix0 + (local - si * sg.npix);
        const JobDev  &J      = bd.jobs[sg.job];
        const uint32_t sample = J.rp.sample + sg.s0 + si;
        const uint32_t x = pix % J.W, y = pix / J.W;
        const rnd4     j = rng4(J.rp.seed, pix, sample, kJitterBounce, 0);
        f3 o, d;
        camera_ray(J.rp.cam, ((float)x + j.a) / (float)J.W, ((float)y + j.b) / (float)J.H, J.rp.aspect, o, d);
        put_ray(b.q[0], p, o, d);
        b.qpath[0][p] = p;
        b.thr[p]      = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        b.rad[p]      = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.pjob[p]     = sg.job;
        b.ppix[p]     = pix;
        b.psamp[p]    = sample;
    }
    if (p == 0) b.cnt[kCntQ0] = min(ch.npaths, b.P);
    block_add_stat(b.stats, 2, live ? 1u : 0u);
}

// raw[pixel] += radiance of each of the segment's passes, in pass order: the same sums,
// in the same order, as one accumulate per pass in the single-scene renderer.
__global__ __launch_bounds__(256) void k_batch_accumulate(BatchDev bd, PathBuffers b)
{
    const BatchChunk ch = bd.chunks[*bd.cursor];
    const uint32_t   t  = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ch.nacc) return;
    const BatchSeg sg    = bd.segs[find_seg(bd.segs, ch.seg0, ch.nseg, t, true)];
    const uint32_t local = t - sg.acc0;
    const uint32_t pix   = sg.pix0 + local;
    const JobDev  &J     = bd.jobs[sg.job];
    const uint32_t x = pix % J.W, y = pix / J.W;
    float *raw = J.raw + 3 * ((size_t)(J.W - 1 - x) + (size_t)(J.H - 1 - y) * J.W);
    float  r0 = raw[0], r1 = raw[1], r2 = raw[2];
    for (uint32_t s = 0; s < sg.nsamp; ++s) {
        const size_t path = (size_t)sg.path0 + (size_t)s * sg.npix + local;
        if (path >= b.P) break;
        const float4 r = b.rad[path];
        r0 += r.x;
        r1 += r.y;
        r2 += r.z;
    }
    raw[0] = r0;
    raw[1] = r1;
    raw[2] = r2;
}

__global__ void k_batch_advance(uint32_t *cursor, uint32_t step)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *cursor += step;
}

}  // namespace

static inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + 255) / 256); }

hipError_t launch_generate(const SceneDev &sc, const PathBuffers &b, const RenderParams &rp, hipStream_t s)
{
    const uint64_t n = (uint64_t)b.npix * rp.nsamp + (rp.nrc.enabled ? rp.nrc.train_live : 0u);
    k_generate<<<std::max(1u, blocks(n)), 256, 0, s>>>(sc, b, rp);
    return hipGetLastError();
}
hipError_t launch_bounce_prep(const PathBuffers &b, uint32_t cur, hipStream_t s)
{
    k_bounce_prep<<<1, 64, 0, s>>>(b, cur);
    return hipGetLastError();
}
hipError_t launch_shade(const SceneDev &sc, const PathBuffers &b, const RenderParams &rp, uint32_t bounce,
                        uint32_t cur, hipStream_t s, const uint32_t *perm)
{
    k_shade<false><<<blocks(b.P), 256, 0, s>>>(sc, b, rp, nullptr, bounce, cur, perm);
    return hipGetLastError();
}
hipError_t launch_shadow_prep(const PathBuffers &b, hipStream_t s)
{
    k_shadow_prep<<<1, 64, 0, s>>>(b);
    return hipGetLastError();
}
hipError_t launch_shadow_resolve(const PathBuffers &b, hipStream_t s)
{
    k_shadow_resolve<<<blocks(b.P), 256, 0, s>>>(b);
    return hipGetLastError();
}
hipError_t launch_shadow_alpha_step(const SceneDev &sc, const PathBuffers &b, hipStream_t s)
{
    k_shadow_alpha<false><<<blocks(b.P), 256, 0, s>>>(sc, b, nullptr);
    k_shadow_compact<<<blocks(b.P), 256, 0, s>>>(b);
    k_shadow_swap<<<1, 64, 0, s>>>(b);
    return hipGetLastError();
}
hipError_t launch_accumulate(const PathBuffers &b, uint32_t nsamp, hipStream_t s)
{
    k_accumulate<<<blocks(b.npix), 256, 0, s>>>(b, nsamp);
    return hipGetLastError();
}
hipError_t launch_display(const float *raw, uint64_t n, uint32_t npix, float4 *out, hipStream_t s)
{
    k_display<<<blocks(npix), 256, 0, s>>>(raw, n, npix, out);
    return hipGetLastError();
}
hipError_t launch_fill(float4 *p, uint64_t n, float v, hipStream_t s)
{
    const uint32_t g = (uint32_t)std::min<uint64_t>(blocks(n), 4096);
    if (n) k_fill<<<g, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}
hipError_t launch_nrc_apply(const PathBuffers &b, hipStream_t s)
{
    k_nrc_apply<<<blocks(b.P), 256, 0, s>>>(b);
    return hipGetLastError();
}
hipError_t launch_nrc_targets(const PathBuffers &b, uint32_t rec_slots, uint32_t max_out, uint32_t train_live,
                              hipStream_t s)
{
    k_nrc_targets<<<std::max(1u, blocks((uint64_t)train_live * rec_slots)), 256, 0, s>>>(b, rec_slots, max_out,
                                                                                       train_live);
    return hipGetLastError();
}
hipError_t launch_hits_to_api(const SceneDev &sc, const HitSoA &h, uint32_t n, uint32_t *prim, uint32_t *geom,
                              float *ngx, float *ngy, float *ngz, uint32_t *inst, hipStream_t s)
{
    if (n) k_hits_to_api<<<blocks(n), 256, 0, s>>>(sc, h, n, prim, geom, ngx, ngy, ngz, inst);
    return hipGetLastError();
}

hipError_t launch_hit_records(const SceneDev &sc, const RaySoA &rs, const HitSoA &h, uint32_t n, const RecordsDev &out,
                              hipStream_t s)
{
    if (n) k_hit_records<<<blocks(n), 256, 0, s>>>(sc, rs, h, n, out);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float *p, uint32_t n, float v, hipStream_t s)
{
    if (n) k_fill_f32<<<blocks(n), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_batch_generate(const BatchDev &bd, const PathBuffers &b, hipStream_t s)
{
    k_batch_generate<<<blocks(b.P), 256, 0, s>>>(bd, b);
    return hipGetLastError();
}

hipError_t launch_shade_multi(const BatchDev &bd, const PathBuffers &b, uint32_t bounce, uint32_t cur, hipStream_t s)
{
    k_shade<true><<<blocks(b.P), 256, 0, s>>>(SceneDev{}, b, RenderParams{}, bd.jobs, bounce, cur, nullptr);
    return hipGetLastError();
}

hipError_t launch_shadow_alpha_step_multi(const BatchDev &bd, const PathBuffers &b, hipStream_t s)
{
    k_shadow_alpha<true><<<blocks(b.P), 256, 0, s>>>(SceneDev{}, b, bd.jobs);
    k_shadow_compact<<<blocks(b.P), 256, 0, s>>>(b);
    k_shadow_swap<<<1, 64, 0, s>>>(b);
    return hipGetLastError();
}

hipError_t launch_batch_accumulate(const BatchDev &bd, const PathBuffers &b, uint32_t max_acc, hipStream_t s)
{
    k_batch_accumulate<<<blocks(max_acc), 256, 0, s>>>(bd, b);
    return hipGetLastError();
}

hipError_t launch_batch_advance(uint32_t *cursor, uint32_t step, hipStream_t s)
{
    k_batch_advance<<<1, 64, 0, s>>>(cursor, step);
    return hipGetLastError();
}

}  // namespace crt
