/*
 * geometry.c -- oracle closest-hit queries and CPU BVH (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the Embree 3 query made at CRender src/objects/model.cpp:10-47
 * (rtcIntersect1 with tnear = 1e-5, tfar = +inf; closest hit; no back-face culling;
 * Ng = (v1-v0) x (v2-v0); u,v barycentrics of v1,v2) with a watertight ray/triangle
 * test (Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013).
 * Hits are accepted for tnear < t <= tfar; equal t resolves to the smaller primID (Embree
 * leaves that order to its BVH, so ties are the one place the two may differ).
 *
 * The BVH plays the role of rtcCommitScene (model.cpp:94): an FP32 binned-SAH BVH2 with
 * up to max_leaf triangles per leaf, traversed near-child first.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include "oracle.h"


/* ---------------------------------------------------------------------------------- */
/* watertight triangle test, FP32 with the paper's FP64 fallback for zero edge values */

typedef struct {
    int   kx, ky, kz;
    float sx, sy, sz;
} wt_ray32;

static void wt_setup32(const float d[3], wt_ray32 *r)
{
    const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) {
        const int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    r->kx = kx;
    r->ky = ky;
    r->kz = kz;
    r->sx = d[kx] / d[kz];
    r->sy = d[ky] / d[kz];
    r->sz = 1.0f / d[kz];
}

static int wt_test32(const wt_ray32 *r, const float *v0, const float *v1, const float *v2,
                     const float o[3], float tnear, float tfar, float *t, float *u, float *v)
{
    const float A[3] = {v0[0] - o[0], v0[1] - o[1], v0[2] - o[2]};
    const float B[3] = {v1[0] - o[0], v1[1] - o[1], v1[2] - o[2]};
    const float C[3] = {v2[0] - o[0], v2[1] - o[1], v2[2] - o[2]};
    const float Ax = A[r->kx] - r->sx * A[r->kz];
    const float Ay = A[r->ky] - r->sy * A[r->kz];
    const float Bx = B[r->kx] - r->sx * B[r->kz];
    const float By = B[r->ky] - r->sy * B[r->kz];
    const float Cx = C[r->kx] - r->sx * C[r->kz];
    const float Cy = C[r->ky] - r->sy * C[r->kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    const float det = U + V + W;
    if (det == 0.0f) return 0;
    const float Az = r->sz * A[r->kz];
    const float Bz = r->sz * B[r->kz];
    const float Cz = r->sz * C[r->kz];
    const float T   = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float tt  = T * rcp;
    if (!(tt > tnear) || !(tt <= tfar)) return 0;
    *t = tt;
    *u = V * rcp;
    *v = W * rcp;
    return 1;
}

int orc_tri_watertight32(const float *v0, const float *v1, const float *v2, const float o[3],
                         const float d[3], float tnear, float tfar, float *t, float *u, float *v)
{
    wt_ray32 r;
    wt_setup32(d, &r);
    return wt_test32(&r, v0, v1, v2, o, tnear, tfar, t, u, v);
}

/* Same test carried out entirely in double: the brute-force ground truth. */
typedef struct {
    int    kx, ky, kz;
    double sx, sy, sz;
} wt_ray64;

static void wt_setup64(const double d[3], wt_ray64 *r)
{
    const double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    if (d[kz] < 0.0) {
        const int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    r->kx = kx;
    r->ky = ky;
    r->kz = kz;
    r->sx = d[kx] / d[kz];
    r->sy = d[ky] / d[kz];
    r->sz = 1.0 / d[kz];
}

static int wt_test64(const wt_ray64 *r, const float *fv0, const float *fv1, const float *fv2,
                     const double o[3], double tnear, double tfar, double *t, double *u, double *v)
{
    const double A[3] = {fv0[0] - o[0], fv0[1] - o[1], fv0[2] - o[2]};
    const double B[3] = {fv1[0] - o[0], fv1[1] - o[1], fv1[2] - o[2]};
    const double C[3] = {fv2[0] - o[0], fv2[1] - o[1], fv2[2] - o[2]};
    const double Ax = A[r->kx] - r->sx * A[r->kz];
    const double Ay = A[r->ky] - r->sy * A[r->kz];
    const double Bx = B[r->kx] - r->sx * B[r->kz];
    const double By = B[r->ky] - r->sy * B[r->kz];
    const double Cx = C[r->kx] - r->sx * C[r->kz];
    const double Cy = C[r->ky] - r->sy * C[r->kz];
    const double U = Cx * By - Cy * Bx;
    const double V = Ax * Cy - Ay * Cx;
    const double W = Bx * Ay - By * Ax;
    if ((U < 0.0 || V < 0.0 || W < 0.0) && (U > 0.0 || V > 0.0 || W > 0.0)) return 0;
    const double det = U + V + W;
    if (det == 0.0) return 0;
    const double Az = r->sz * A[r->kz];
    const double Bz = r->sz * B[r->kz];
    const double Cz = r->sz * C[r->kz];
    const double T  = U * Az + V * Bz + W * Cz;
    const double tt = T / det;
    if (!(tt > tnear) || !(tt <= tfar)) return 0;
    *t = tt;
    *u = V / det;
    *v = W / det;
    return 1;
}

int32_t orc_closest_brute64(const float *tri, uint32_t ntri, const float o[3], const float d[3],
                            float tnear, float tfar, double *t, double *u, double *v)
{
    const double od[3] = {o[0], o[1], o[2]};
    const double dd[3] = {d[0], d[1], d[2]};
    return orc_brute64d(tri, ntri, od, dd, (double)tnear, (double)tfar, t, u, v);
}

int32_t orc_brute64d(const float *tri, uint32_t ntri, const double od[3], const double dd[3],
                     double tnear, double tfar, double *t, double *u, double *v)
{
    wt_ray64 r;
    wt_setup64(dd, &r);
    int32_t best = -1;
    double  bt = tfar, bu = 0.0, bv = 0.0;
    for (uint32_t i = 0; i < ntri; ++i) {
        const float *p = tri + 9 * (size_t)i;
        double tt, uu, vv;
        if (wt_test64(&r, p, p + 3, p + 6, od, tnear, bt, &tt, &uu, &vv)) {
            if (best < 0 || tt < bt) {
                best = (int32_t)i;
                bt   = tt;
                bu   = uu;
                bv   = vv;
            }
        }
    }
    if (best >= 0) {
        *t = bt;
        *u = bu;
        *v = bv;
    }
    return best;
}

int32_t orc_closest_brute32(const float *tri, uint32_t ntri, const float o[3], const float d[3],
                            float tnear, float tfar, float *t, float *u, float *v)
{
    wt_ray32 r;
    wt_setup32(d, &r);
    int32_t best = -1;
    float   bt = tfar, bu = 0.0f, bv = 0.0f;
    for (uint32_t i = 0; i < ntri; ++i) {
        const float *p = tri + 9 * (size_t)i;
        float tt, uu, vv;
        if (wt_test32(&r, p, p + 3, p + 6, o, tnear, bt, &tt, &uu, &vv)) {
            if (best < 0 || tt < bt) {
                best = (int32_t)i;
                bt   = tt;
                bu   = uu;
                bv   = vv;
            }
        }
    }
    if (best >= 0) {
        *t = bt;
        *u = bu;
        *v = bv;
    }
    return best;
}

/* ---------------------------------------------------------------------------------- */
/* CPU BVH2: binned SAH over centroid bounds                                           */

typedef struct {
    float    lo[3], hi[3];
    uint32_t first;   /* leaf: first index into prims; inner: left child (right = left+1) */
    uint32_t count;   /* 0 = inner */
} bnode;

struct orc_bvh {
    bnode    *nodes;
    uint32_t  nnodes;
    uint32_t *prims;  /* leaf order -> primID */
    float    *tri;    /* 9 floats per triangle in leaf order */
    uint32_t  ntri;
};

#define NBINS 16

typedef struct {
    float lo[3], hi[3];
} aabb;

static void aabb_empty(aabb *b)
{
    for (int k = 0; k < 3; ++k) {
        b->lo[k] = FLT_MAX;
        b->hi[k] = -FLT_MAX;
    }
}

static void aabb_grow(aabb *b, const aabb *o)
{
    for (int k = 0; k < 3; ++k) {
        if (o->lo[k] < b->lo[k]) b->lo[k] = o->lo[k];
        if (o->hi[k] > b->hi[k]) b->hi[k] = o->hi[k];
    }
}

static float aabb_area(const aabb *b)
{
    const float dx = b->hi[0] - b->lo[0], dy = b->hi[1] - b->lo[1], dz = b->hi[2] - b->lo[2];
    if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

typedef struct {
    uint32_t node, begin, end;
} build_task;

orc_bvh *orc_bvh_build(const float *tri, uint32_t ntri, uint32_t max_leaf)
{
    if (max_leaf < 1) max_leaf = 1;
    orc_bvh *bvh = (orc_bvh *)calloc(1, sizeof(orc_bvh));
    bvh->ntri    = ntri;
    bvh->prims   = (uint32_t *)malloc(sizeof(uint32_t) * (ntri ? ntri : 1));
    bvh->nodes   = (bnode *)malloc(sizeof(bnode) * (2 * (size_t)ntri + 1));
    aabb  *boxes = (aabb *)malloc(sizeof(aabb) * (ntri ? ntri : 1));
    float *cent  = (float *)malloc(sizeof(float) * 3 * (ntri ? ntri : 1));
    for (uint32_t i = 0; i < ntri; ++i) {
        const float *p = tri + 9 * (size_t)i;
        aabb_empty(&boxes[i]);
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) {
                const float x = p[3 * c + k];
                if (x < boxes[i].lo[k]) boxes[i].lo[k] = x;
                if (x > boxes[i].hi[k]) boxes[i].hi[k] = x;
            }
        for (int k = 0; k < 3; ++k) cent[3 * i + k] = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
        bvh->prims[i] = i;
    }
    uint32_t    nnodes = 1;
    build_task *stack  = (build_task *)malloc(sizeof(build_task) * 128);
    size_t      scap = 128, sp = 0;
    stack[sp++] = (build_task){0, 0, ntri};
    while (sp) {
        const build_task tk = stack[--sp];
        bnode           *nd = &bvh->nodes[tk.node];
        aabb             nb, cb;
        aabb_empty(&nb);
        aabb_empty(&cb);
        for (uint32_t i = tk.begin; i < tk.end; ++i) {
            const uint32_t p = bvh->prims[i];
            aabb_grow(&nb, &boxes[p]);
            for (int k = 0; k < 3; ++k) {
                if (cent[3 * p + k] < cb.lo[k]) cb.lo[k] = cent[3 * p + k];
                if (cent[3 * p + k] > cb.hi[k]) cb.hi[k] = cent[3 * p + k];
            }
        }
        if (tk.end == tk.begin) aabb_empty(&nb);
        memcpy(nd->lo, nb.lo, sizeof nb.lo);
        memcpy(nd->hi, nb.hi, sizeof nb.hi);
        const uint32_t n = tk.end - tk.begin;
        /* binned SAH */
        int   best_axis = -1, best_bin = -1;
        float best_cost = FLT_MAX;
        if (n > 1) {
            for (int axis = 0; axis < 3; ++axis) {
                const float ext = cb.hi[axis] - cb.lo[axis];
                if (!(ext > 0.0f)) continue;
                aabb     bb[NBINS];
                uint32_t bc[NBINS] = {0};
                for (int b = 0; b < NBINS; ++b) aabb_empty(&bb[b]);
                const float k = (float)NBINS * (1.0f - 1e-6f) / ext;
                for (uint32_t i = tk.begin; i < tk.end; ++i) {
                    const uint32_t p = bvh->prims[i];
                    int            b = (int)((cent[3 * p + axis] - cb.lo[axis]) * k);
                    if (b < 0) b = 0;
                    if (b >= NBINS) b = NBINS - 1;
                    bc[b]++;
                    aabb_grow(&bb[b], &boxes[p]);
                }
                float    rarea[NBINS];
                uint32_t rcnt[NBINS];
                aabb     acc;
                aabb_empty(&acc);
                uint32_t c = 0;
                for (int b = NBINS - 1; b > 0; --b) {
                    aabb_grow(&acc, &bb[b]);
                    c += bc[b];
                    rarea[b] = aabb_area(&acc);
                    rcnt[b]  = c;
                }
                aabb_empty(&acc);
                c = 0;
                for (int b = 0; b < NBINS - 1; ++b) {
                    aabb_grow(&acc, &bb[b]);
                    c += bc[b];
                    if (c == 0 || rcnt[b + 1] == 0) continue;
                    const float cost = aabb_area(&acc) * (float)c + rarea[b + 1] * (float)rcnt[b + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = axis;
                        best_bin  = b;
                    }
                }
            }
        }
        const float parea     = aabb_area(&nb);
        const float leaf_cost = (float)n;
        const float split_cost =
            best_axis >= 0 ? 1.0f + (parea > 0.0f ? best_cost / parea : (float)n) : FLT_MAX;
        if (n <= 1 || (n <= max_leaf && leaf_cost <= split_cost)) {
            nd->first = tk.begin;
            nd->count = n;
            continue;
        }
        uint32_t mid;
        if (best_axis >= 0) {
            const float ext = cb.hi[best_axis] - cb.lo[best_axis];
            const float k   = (float)NBINS * (1.0f - 1e-6f) / ext;
            uint32_t    i = tk.begin, j = tk.end;
            while (i < j) {
                const uint32_t p = bvh->prims[i];
                int            b = (int)((cent[3 * p + best_axis] - cb.lo[best_axis]) * k);
                if (b < 0) b = 0;
                if (b >= NBINS) b = NBINS - 1;
                if (b <= best_bin)
                    ++i;
                else {
                    --j;
                    bvh->prims[i] = bvh->prims[j];
                    bvh->prims[j] = p;
                }
            }
            mid = i;
        } else {
            mid = tk.begin + n / 2; /* all centroids coincide: split by index */
        }
        if (mid == tk.begin || mid == tk.end) mid = tk.begin + n / 2;
        nd->first = nnodes;
        nd->count = 0;
        nnodes += 2;
        if (sp + 2 > scap) {
            scap *= 2;
            stack = (build_task *)realloc(stack, sizeof(build_task) * scap);
        }
        stack[sp++] = (build_task){nd->first + 1, mid, tk.end};
        stack[sp++] = (build_task){nd->first, tk.begin, mid};
    }
    bvh->nnodes = nnodes;
    bvh->tri    = (float *)malloc(sizeof(float) * 9 * (ntri ? ntri : 1));
    for (uint32_t i = 0; i < ntri; ++i) memcpy(bvh->tri + 9 * (size_t)i, tri + 9 * (size_t)bvh->prims[i], 9 * sizeof(float));
    free(stack);
    free(boxes);
    free(cent);
    return bvh;
}

void orc_bvh_free(orc_bvh *bvh)
{
    if (!bvh) return;
    free(bvh->nodes);
    free(bvh->prims);
    free(bvh->tri);
    free(bvh);
}

uint32_t orc_bvh_node_count(const orc_bvh *bvh) { return bvh->nnodes; }

double orc_bvh_sah_cost(const orc_bvh *bvh)
{
    if (bvh->ntri == 0) return 0.0;
    const double root = aabb_area((const aabb *)bvh->nodes[0].lo);
    double       cost = 0.0;
    for (uint32_t i = 0; i < bvh->nnodes; ++i) {
        const double a = aabb_area((const aabb *)bvh->nodes[i].lo);
        cost += bvh->nodes[i].count ? a * bvh->nodes[i].count : a;
    }
    return root > 0.0 ? cost / root : 0.0;
}

int orc_bvh_validate(const orc_bvh *bvh, const float *tri, uint32_t ntri)
{
    if (bvh->ntri != ntri) return 1;
    uint8_t *seen = (uint8_t *)calloc(ntri ? ntri : 1, 1);
    int      err  = 0;
    for (uint32_t i = 0; i < ntri && !err; ++i) {
        if (bvh->prims[i] >= ntri || seen[bvh->prims[i]]) err = 2;
        else seen[bvh->prims[i]] = 1;
    }
    /* leaf triangles lie inside the leaf box; child boxes lie inside the parent box */
    uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * (bvh->nnodes + 1));
    size_t    sp    = 0;
    if (ntri) stack[sp++] = 0;
    while (sp && !err) {
        const uint32_t ni = stack[--sp];
        const bnode   *nd = &bvh->nodes[ni];
        if (nd->count) {
            for (uint32_t i = nd->first; i < nd->first + nd->count; ++i) {
                const float *p = tri + 9 * (size_t)bvh->prims[i];
                for (int c = 0; c < 3; ++c)
                    for (int k = 0; k < 3; ++k)
                        if (p[3 * c + k] < nd->lo[k] || p[3 * c + k] > nd->hi[k]) err = 3;
            }
        } else {
            for (int ch = 0; ch < 2; ++ch) {
                const bnode *c = &bvh->nodes[nd->first + ch];
                for (int k = 0; k < 3; ++k)
                    if (c->lo[k] < nd->lo[k] || c->hi[k] > nd->hi[k]) err = 4;
                stack[sp++] = nd->first + ch;
            }
        }
    }
    free(stack);
    free(seen);
    return err;
}

/* slab test; conservative scaling of tmax by 1 + 2*gamma(3) (Ize, "Robust BVH Ray
 * Traversal", JCGT 2013) so rounding never culls a box the ray touches. */
static int box_hit(const bnode *n, const float o[3], const float inv[3], float tmax, float *tentry)
{
    float t0 = 0.0f, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        float tn = (n->lo[k] - o[k]) * inv[k];
        float tf = (n->hi[k] - o[k]) * inv[k];
        if (tn > tf) {
            const float s = tn;
            tn = tf;
            tf = s;
        }
        tf *= 1.0000003576278687f;
        if (tn > t0) t0 = tn;   /* NaN (0*inf) leaves the bound unchanged */
        if (tf < t1) t1 = tf;
    }
    *tentry = t0;
    return t0 <= t1;
}

static int32_t bvh_closest32(const orc_bvh *bvh, const float o[3], const float d[3], float tnear,
                             float tfar, float *t, float *u, float *v)
{
    if (bvh->ntri == 0) return -1;
    wt_ray32 r;
    wt_setup32(d, &r);
    const float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    uint32_t    stack[128];
    int         sp = 0;
    int32_t     best = -1;
    float       bt = tfar, bu = 0.0f, bv = 0.0f;
    float       te;
    if (!box_hit(&bvh->nodes[0], o, inv, bt, &te)) return -1;
    uint32_t ni = 0;
    for (;;) {
        const bnode *nd = &bvh->nodes[ni];
        if (nd->count) {
            for (uint32_t i = nd->first; i < nd->first + nd->count; ++i) {
                const float *p = bvh->tri + 9 * (size_t)i;
                float        tt, uu, vv;
                if (wt_test32(&r, p, p + 3, p + 6, o, tnear, bt, &tt, &uu, &vv)) {
                    const int32_t pid = (int32_t)bvh->prims[i];
                    if (best < 0 || tt < bt || pid < best) {
                        best = pid;
                        bt   = tt;
                        bu   = uu;
                        bv   = vv;
                    }
                }
            }
        } else {
            float        t0, t1;
            const int    h0 = box_hit(&bvh->nodes[nd->first], o, inv, bt, &t0);
            const int    h1 = box_hit(&bvh->nodes[nd->first + 1], o, inv, bt, &t1);
            if (h0 && h1) {
                const uint32_t nearc = t0 <= t1 ? nd->first : nd->first + 1;
                const uint32_t farc  = t0 <= t1 ? nd->first + 1 : nd->first;
                stack[sp++] = farc;
                ni          = nearc;
                continue;
            }
            if (h0) {
                ni = nd->first;
                continue;
            }
            if (h1) {
                ni = nd->first + 1;
                continue;
            }
        }
        if (sp == 0) break;
        ni = stack[--sp];
    }
    if (best >= 0) {
        *t = bt;
        *u = bu;
        *v = bv;
    }
    return best;
}

/* The same walk with FP64 closest-hit decisions: FP64 watertight test per triangle (the
 * brute-force ground truth's test) over the FP32 BVH, whose boxes are culled against the FP32
 * rounding of the current FP64 distance widened by one part in 10^5.  For classifying
 * precision-sensitive hits on scenes too large for FP64 brute force. */
static int32_t bvh_closest64(const orc_bvh *bvh, const float o[3], const float d[3], float tnear,
                             float tfar, double *t, double *u, double *v)
{
    if (bvh->ntri == 0) return -1;
    const double od[3] = {o[0], o[1], o[2]};
    const double dd[3] = {d[0], d[1], d[2]};
    wt_ray64 r;
    wt_setup64(dd, &r);
    const float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    uint32_t    stack[128];
    int         sp = 0;
    int32_t     best = -1;
    double      bt = tfar, bu = 0.0, bv = 0.0;
    float       te;
    if (!box_hit(&bvh->nodes[0], o, inv, tfar, &te)) return -1;
    uint32_t ni = 0;
    for (;;) {
        const bnode *nd = &bvh->nodes[ni];
        const float  cull = best < 0 ? tfar : (float)(bt * (1.0 + 1e-5)) + 1e-30f;
        if (nd->count) {
            for (uint32_t i = nd->first; i < nd->first + nd->count; ++i) {
                const float *p = bvh->tri + 9 * (size_t)i;
                double       tt, uu, vv;
                if (wt_test64(&r, p, p + 3, p + 6, od, tnear, bt, &tt, &uu, &vv)) {
                    const int32_t pid = (int32_t)bvh->prims[i];
                    if (best < 0 || tt < bt || pid < best) {
                        best = pid;
                        bt   = tt;
                        bu   = uu;
                        bv   = vv;
                    }
                }
            }
        } else {
            float     t0, t1;
            const int h0 = box_hit(&bvh->nodes[nd->first], o, inv, cull, &t0);
            const int h1 = box_hit(&bvh->nodes[nd->first + 1], o, inv, cull, &t1);
            if (h0 && h1) {
                stack[sp++] = t0 <= t1 ? nd->first + 1 : nd->first;
                ni          = t0 <= t1 ? nd->first : nd->first + 1;
                continue;
            }
            if (h0 || h1) {
                ni = h0 ? nd->first : nd->first + 1;
                continue;
            }
        }
        if (sp == 0) break;
        ni = stack[--sp];
    }
    if (best >= 0) {
        *t = bt;
        *u = bu;
        *v = bv;
    }
    return best;
}

int32_t orc_bvh_closest64(const orc_bvh *bvh, const float o[3], const float d[3], float tnear, float tfar,
                          double *t, double *u, double *v)
{
    return bvh_closest64(bvh, o, d, tnear, tfar, t, u, v);
}

/* Public wrapper; wt_test32 accepts t <= bt, so an equal-t triangle reaches the primID
 * comparison above and only a smaller primID replaces the current hit. */
int32_t orc_bvh_closest32(const orc_bvh *bvh, const float o[3], const float d[3], float tnear,
                          float tfar, float *t, float *u, float *v)
{
    return bvh_closest32(bvh, o, d, tnear, tfar, t, u, v);
}

/* ---------------------------------------------------------------------------------- */
/* batched queries                                                                     */

typedef struct {
    const float   *tri;
    uint32_t       ntri;
    const orc_bvh *bvh;
    int32_t        trav;
    const float   *ox, *oy, *oz, *dx, *dy, *dz, *tn, *tf;
    uint32_t       n;
    int32_t       *prim;
    double        *t, *u, *v;
    volatile uint32_t next;
} batch_job;

static void *batch_worker(void *arg)
{
    batch_job *job = (batch_job *)arg;
    for (;;) {
        const uint32_t begin = __atomic_fetch_add(&job->next, 256u, __ATOMIC_RELAXED);
        if (begin >= job->n) break;
        const uint32_t end = begin + 256u < job->n ? begin + 256u : job->n;
        for (uint32_t i = begin; i < end; ++i) {
            const float o[3] = {job->ox[i], job->oy[i], job->oz[i]};
            const float d[3] = {job->dx[i], job->dy[i], job->dz[i]};
            int32_t     p;
            if (job->trav == ORC_TRAV_BRUTE64 || job->trav == ORC_TRAV_BVH64) {
                double tt = 0, uu = 0, vv = 0;
                if (job->trav == ORC_TRAV_BVH64)
                    p = bvh_closest64(job->bvh, o, d, job->tn[i], job->tf[i], &tt, &uu, &vv);
                else
                    p = orc_closest_brute64(job->tri, job->ntri, o, d, job->tn[i], job->tf[i], &tt, &uu, &vv);
                job->t[i] = tt;
                job->u[i] = uu;
                job->v[i] = vv;
            } else {
                float tt = 0, uu = 0, vv = 0;
                if (job->trav == ORC_TRAV_BRUTE32)
                    p = orc_closest_brute32(job->tri, job->ntri, o, d, job->tn[i], job->tf[i], &tt, &uu, &vv);
                else
                    p = bvh_closest32(job->bvh, o, d, job->tn[i], job->tf[i], &tt, &uu, &vv);
                job->t[i] = tt;
                job->u[i] = uu;
                job->v[i] = vv;
            }
            job->prim[i] = p;
            if (p < 0) {
                job->t[i] = INFINITY;
                job->u[i] = 0.0;
                job->v[i] = 0.0;
            }
        }
    }
    return NULL;
}

int orc_default_threads(void);
int orc_default_threads(void)
{
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    if (n < 1) n = 1;
    if (n > 256) n = 256;
    return (int)n;
}

void orc_intersect_batch(const float *tri, uint32_t ntri, const orc_bvh *bvh, int32_t trav,
                         const float *ox, const float *oy, const float *oz, const float *dx,
                         const float *dy, const float *dz, const float *tnear, const float *tfar,
                         uint32_t n, int32_t *prim, double *t, double *u, double *v, int nthreads)
{
    batch_job job = {tri, ntri, bvh, trav, ox, oy, oz, dx, dy, dz, tnear, tfar, n, prim, t, u, v, 0};
    if (nthreads <= 0) nthreads = orc_default_threads();
    pthread_t th[256];
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, batch_worker, &job);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
}
