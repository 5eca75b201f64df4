/*
 * integrator_impl.h -- oracle bounce loop, instantiated once per precision by
 * integrator.c (TEST INFRASTRUCTURE ONLY).
 *
 * Macros supplied by the includer: R (float/double), SFX (name suffix), and the math
 * functions SQRT_ SIN_ COS_ TAN_ ATAN2_ ASIN_ ACOS_ POW_ FABS_ for that precision.
 *
 * compat mode restates CRender src/render/renderer.cpp:21-102 (process_hit) and
 * renderer.cpp:258-384 (_sample_pixel) with Philox in place of mt19937.  pbr mode keeps
 * the same loop and fixes the estimator where the configs need it: Lambert NEE applies
 * albedo once with 1/pi, normals face the incoming ray, GGX metal (sampling.h:93-140
 * formulas, Heitz 2018 visible-normal sampling), Fresnel (Schlick) dielectric, and no
 * NEE on delta lobes.
 */

#define CAT2_(a, b) a##_##b
#define CAT2(a, b) CAT2_(a, b)
#define F_(name) CAT2(name, SFX)

typedef struct {
    R x, y, z;
} F_(vec);
#define V F_(vec)

static inline V F_(mk)(R x, R y, R z)
{
    V v = {x, y, z};
    return v;
}
static inline V F_(add)(V a, V b) { return F_(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V F_(sub)(V a, V b) { return F_(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V F_(mul)(V a, V b) { return F_(mk)(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V F_(scl)(V a, R s) { return F_(mk)(a.x * s, a.y * s, a.z * s); }
static inline R F_(dot)(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V F_(cross)(V a, V b)
{
    return F_(mk)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* glm::normalize = v * inversesqrt(dot(v, v)) */
static inline V F_(norm)(V a) { return F_(scl)(a, (R)1 / SQRT_(F_(dot)(a, a))); }
/* glm::reflect(I, N) = I - 2 dot(N, I) N */
static inline V F_(reflect)(V i, V n) { return F_(sub)(i, F_(scl)(n, (R)2 * F_(dot)(n, i))); }
static inline R F_(maxc)(V a)
{
    R m = a.x > a.y ? a.x : a.y;
    return m > a.z ? m : a.z;
}

#define PI_R  ((R)3.14159265359)
#define TAU_R ((R)6.28318530717)

/* sampling.h:156-166 */
static inline V F_(sphere)(R u, R v)
{
    const R ct  = (R)2 * u - (R)1;
    const R st  = SQRT_((R)1 - ct * ct);
    const R phi = TAU_R * v;
    return F_(mk)(st * COS_(phi), ct, st * SIN_(phi));
}

/* sampling.h:21-33 */
static inline void F_(onb)(V n, V *t, V *b)
{
    const R s = (n.z < (R)0) ? (R)-1 : (R)1;
    const R a = (R)-1 / (s + n.z);
    const R bb = n.x * n.y * a;
    *t = F_(mk)((R)1 + s * n.x * n.x * a, s * bb, -s * n.x);
    *b = F_(mk)(bb, s + n.y * n.y * a, -n.y);
}

/* camera.cpp:14-39 */
static void F_(camera)(const orc_camera *cam, R x, R y, R aspect, V *o, V *d)
{
    const float *m = cam->matrix;
    V            dir;
    if (cam->mode == 0) {
        const R u = ((R)2 * x - (R)1) * aspect;
        const R v = (R)2 * y - (R)1;
        const R w = (R)1 / TAN_((R)0.5 * ((R)cam->fov * (R)0.01745329251994329576923690768489));
        dir = F_(mk)((R)m[0] * u + (R)m[4] * v + (R)m[8] * w, (R)m[1] * u + (R)m[5] * v + (R)m[9] * w,
                     (R)m[2] * u + (R)m[6] * v + (R)m[10] * w);
        *o  = F_(mk)(cam->pos[0], cam->pos[1], cam->pos[2]);
    } else {
        const R u = (R)2 * x - (R)1;
        const R v = (R)2 * y - (R)1;
        const R su = (R)cam->scale * u, sv = (R)cam->scale * v;
        *o  = F_(mk)((R)m[0] * su + (R)m[4] * sv + (R)m[12], (R)m[1] * su + (R)m[5] * sv + (R)m[13],
                     (R)m[2] * su + (R)m[6] * sv + (R)m[14]);
        dir = F_(mk)(m[8], m[9], m[10]);
    }
    *d = F_(norm)(dir);
}

typedef struct {
    int32_t prim;
    R       t;          /* distance |point - origin| (model.cpp:119-120) */
    V       point;
    V       normal;     /* normalize(Ng) */
    float   uvx, uvy;
} F_(hit);

/* scene::cast_ray (scene.cpp:79-98) -> model::intersect (model.cpp:99-126) ->
 * _intersect (model.cpp:5-49) for the identity instance: the direction is renormalised,
 * point = o + d't, distance recomputed as |point - o|, normal = normalize(Ng),
 * uv = rtcInterpolate0 of the per-corner uvs. */
static int F_(cast)(const orc_scene *sc, V o, V d_in, F_(hit) * h)
{
    const V d = F_(norm)(d_in);
    R       t = 0, bu = 0, bv = 0;
    int32_t p;
    if (sc->trav == ORC_TRAV_BRUTE64 || sc->trav == ORC_TRAV_BVH64) {
        double tt = 0, uu = 0, vv = 0;
        if (sc->trav == ORC_TRAV_BVH64) {
            const float of[3] = {(float)o.x, (float)o.y, (float)o.z};
            const float df[3] = {(float)d.x, (float)d.y, (float)d.z};
            p = orc_bvh_closest64(sc->bvh, of, df, 1e-5f, INFINITY, &tt, &uu, &vv);
        } else if (sizeof(R) == sizeof(double)) {
            const double od[3] = {(double)o.x, (double)o.y, (double)o.z};
            const double dd[3] = {(double)d.x, (double)d.y, (double)d.z};
            p = orc_brute64d(sc->tri, sc->ntri, od, dd, 1e-5, INFINITY, &tt, &uu, &vv);
        } else {
            const float of[3] = {(float)o.x, (float)o.y, (float)o.z};
            const float df[3] = {(float)d.x, (float)d.y, (float)d.z};
            p = orc_closest_brute64(sc->tri, sc->ntri, of, df, 1e-5f, INFINITY, &tt, &uu, &vv);
        }
        t  = (R)tt;
        bu = (R)uu;
        bv = (R)vv;
    } else {
        const float of[3] = {(float)o.x, (float)o.y, (float)o.z};
        const float df[3] = {(float)d.x, (float)d.y, (float)d.z};
        float       tt = 0, uu = 0, vv = 0;
        if (sc->trav == ORC_TRAV_BRUTE32 || !sc->bvh)
            p = orc_closest_brute32(sc->tri, sc->ntri, of, df, 1e-5f, INFINITY, &tt, &uu, &vv);
        else
            p = orc_bvh_closest32(sc->bvh, of, df, 1e-5f, INFINITY, &tt, &uu, &vv);
        t  = (R)tt;
        bu = (R)uu;
        bv = (R)vv;
    }
    h->prim = p;
    if (p < 0) return 0;
    const float *q  = sc->tri + 9 * (size_t)p;
    const V      v0 = F_(mk)(q[0], q[1], q[2]), v1 = F_(mk)(q[3], q[4], q[5]), v2 = F_(mk)(q[6], q[7], q[8]);
    h->point        = F_(add)(o, F_(scl)(d, t));
    h->t            = SQRT_(F_(dot)(F_(sub)(h->point, o), F_(sub)(h->point, o)));
    h->normal       = F_(norm)(F_(cross)(F_(sub)(v1, v0), F_(sub)(v2, v0)));
    if (sc->uv) {
        const float *w  = sc->uv + 6 * (size_t)p;
        const R      b0 = (R)1 - bu - bv;
        h->uvx = (float)(b0 * w[0] + bu * w[2] + bv * w[4]);
        h->uvy = (float)(b0 * w[1] + bu * w[3] + bv * w[5]);
    } else {
        h->uvx = h->uvy = 0.0f;
    }
    return 1;
}

static void F_(colour)(const orc_scene *sc, const orc_material *m, float u, float v, float out[4])
{
    if (m->tex >= 0 && (uint32_t)m->tex < sc->ntex) {
        orc_get_uv(&sc->textures[m->tex], u, v, out);
        return;
    }
    out[0] = m->colour[0];
    out[1] = m->colour[1];
    out[2] = m->colour[2];
    out[3] = m->colour[3];
}

/* renderer.cpp:277-289 + scene.cpp:67-77 */
static V F_(sky)(const orc_scene *sc, V d)
{
    if (!sc->sky.rgba) return F_(mk)(0, 0, 0);
    R dy = d.y;
    if (dy > (R)1) dy = (R)1;
    if (dy < (R)-1) dy = (R)-1;
    const float u = 0.5f + (float)ATAN2_(d.z, d.x) * (1.0f / 6.28318530717f);
    const float v = 0.5f - (float)ASIN_(dy) * (1.0f / 3.14159265359f);
    float       c[4];
    orc_get_uv(&sc->sky, u + sc->sky_rot[0], v + sc->sky_rot[1], c);
    return F_(mk)(c[0], c[1], c[2]);
}

/* sampling.h:53-57 */
static V F_(sun_radiance)(const orc_sun *sun, V dir)
{
    const V   nd = F_(mk)(-sun->dir[0], -sun->dir[1], -sun->dir[2]);
    R         c  = F_(dot)(dir, nd);
    if (c > (R)1) c = (R)1;
    if (c < (R)-1) c = (R)-1;
    const R ang = ACOS_(c);
    if (ang < (R)sun->size)
        return F_(mk)((R)sun->colour[0] * (R)sun->intensity, (R)sun->colour[1] * (R)sun->intensity,
                      (R)sun->colour[2] * (R)sun->intensity);
    return F_(mk)(0, 0, 0);
}

/* GGX pieces: sampling.h:93-140 */
static inline R F_(ggx_d)(R noh, R a)
{
    const R a2 = a * a;
    const R d  = (noh * a2 - noh) * noh + (R)1;
    return a2 / (d * d * PI_R);
}
static inline R F_(ggx_vis)(R nov, R nol, R a)
{
    const R a2   = a * a;
    const R ggxv = nol * SQRT_(nov * nov * ((R)1 - a2) + a2);
    const R ggxl = nov * SQRT_(nol * nol * ((R)1 - a2) + a2);
    return (R)0.5 / (ggxv + ggxl);
}
static inline V F_(schlick)(R u, V f0)
{
    const R f = POW_((R)1 - u, (R)5);
    return F_(mk)(f0.x + ((R)1 - f0.x) * f, f0.y + ((R)1 - f0.y) * f, f0.z + ((R)1 - f0.z) * f);
}

/* GGX visible-normal sampling (Heitz 2018, "Sampling the GGX Distribution of Visible
 * Normals") in the local frame with the normal on +z.  Returns 0 when the sampled
 * direction is below the surface; else the reflected direction, the weight G2/G1 of
 * the estimator f*cos/pdf (before Fresnel) and v.h. */
static int F_(ggx_sample)(V lv, R a, R u1, R u2, V *ll_out, R *g_out, R *voh_out)
{
    const V vh    = F_(norm)(F_(mk)(a * lv.x, a * lv.y, lv.z));
    const R lensq = vh.x * vh.x + vh.y * vh.y;
    const V t1v   = lensq > (R)0 ? F_(scl)(F_(mk)(-vh.y, vh.x, (R)0), (R)1 / SQRT_(lensq)) : F_(mk)(1, 0, 0);
    const V t2v   = F_(cross)(vh, t1v);
    const R r     = SQRT_(u1);
    const R phi   = TAU_R * u2;
    const R t1    = r * COS_(phi);
    R       t2    = r * SIN_(phi);
    const R s     = (R)0.5 * ((R)1 + vh.z);
    t2            = ((R)1 - s) * SQRT_((R)1 - t1 * t1) + s * t2;
    R       nz2   = (R)1 - t1 * t1 - t2 * t2;
    if (nz2 < (R)0) nz2 = (R)0;
    const V nh  = F_(add)(F_(add)(F_(scl)(t1v, t1), F_(scl)(t2v, t2)), F_(scl)(vh, SQRT_(nz2)));
    const V ne  = F_(norm)(F_(mk)(a * nh.x, a * nh.y, nh.z > (R)0 ? nh.z : (R)0));
    const R voh = F_(dot)(lv, ne);
    const V ll  = F_(sub)(F_(scl)(ne, (R)2 * voh), lv);
    if (ll.z <= (R)0 || lv.z <= (R)0) return 0;
    const R a2 = a * a;
    const R g1 = (R)2 * lv.z / (lv.z + SQRT_(a2 + ((R)1 - a2) * lv.z * lv.z));
    const R g2 = F_(ggx_vis)(lv.z, ll.z, a) * (R)4 * lv.z * ll.z;
    *ll_out  = ll;
    *g_out   = g2 / g1;
    *voh_out = voh;
    return 1;
}

typedef struct {
    int  alpha;
    int  delta;       /* pbr: no NEE on this lobe */
    int  absorbed;    /* pbr GGX: sampled direction under the surface */
    R    emission;
    V    albedo;      /* throughput multiplier */
    float colour[4];
    V    o, d;
    V    nee_n;       /* normal used for the NEE offset and cosine */
    int  ggx;
    R    alpha_r;
    V    f0;
} F_(scatter);

/* renderer.cpp:21-102 (compat) plus the pbr variants */
static void F_(process_hit)(const orc_scene *sc, const orc_settings *st, const F_(hit) * h, V in_d,
                            const float u[4], const float u2[4], F_(scatter) * out)
{
    const orc_material *m = &sc->mats[sc->mat_idx[h->prim]];
    memset(out, 0, sizeof *out);
    out->emission = (R)m->emission;
    F_(colour)(sc, m, h->uvx, h->uvy, out->colour);
    if (out->colour[3] == 0.0f) {
        out->alpha = 1;
        return;
    }
    out->albedo = F_(mk)(out->colour[0], out->colour[1], out->colour[2]);
    out->nee_n  = h->normal;
    const V n   = h->normal;
    if (st->mode == ORC_COMPAT) {
        switch (m->type) {
        case ORC_GLASS: {
            V       out_n      = n;
            const V reflected  = F_(reflect)(in_d, n);
            R       ni_over_nt = (R)1 / (R)m->ior;
            if (F_(dot)(in_d, n) > (R)0) {
                out_n      = F_(scl)(n, (R)-1);
                ni_over_nt = (R)m->ior;
            }
            const V uvd  = F_(norm)(in_d);
            const R dt   = F_(dot)(uvd, out_n);
            const R disc = (R)1 - ni_over_nt * ni_over_nt * ((R)1 - dt * dt);
            out->o       = F_(add)(h->point, F_(scl)(out_n, (R)-0.0001));
            if (disc > (R)0)
                out->d = F_(sub)(F_(scl)(F_(sub)(uvd, F_(scl)(out_n, dt)), ni_over_nt), F_(scl)(out_n, SQRT_(disc)));
            else
                out->d = reflected;
            break;
        }
        case ORC_METAL:
            out->o      = F_(add)(h->point, F_(scl)(n, (R)0.0001));
            out->d      = F_(reflect)(in_d, n);
            out->albedo = F_(scl)(out->albedo, (R)m->reflectiveness);
            break;
        default: {
            const V p = F_(sphere)(u[0], u[1]);
            out->o    = F_(add)(h->point, F_(scl)(n, (R)0.0001));
            out->d    = F_(norm)(F_(add)(n, p));
            break;
        }
        }
        return;
    }
    /* pbr */
    const V nf = F_(dot)(in_d, n) > (R)0 ? F_(scl)(n, (R)-1) : n; /* faces the incoming ray */
    switch (m->type) {
    case ORC_GLASS: {
        out->delta = 1;
        V       out_n      = n;
        R       ni_over_nt = (R)1 / (R)m->ior;
        if (F_(dot)(in_d, n) > (R)0) {
            out_n      = F_(scl)(n, (R)-1);
            ni_over_nt = (R)m->ior;
        }
        const V uvd  = F_(norm)(in_d);
        const R dt   = F_(dot)(uvd, out_n);
        const R disc = (R)1 - ni_over_nt * ni_over_nt * ((R)1 - dt * dt);
        int     refl = 1;
        if (disc > (R)0) {
            const R cosf_ = ni_over_nt <= (R)1 ? -dt : SQRT_(disc);
            const R r0    = ((R)1 - (R)m->ior) / ((R)1 + (R)m->ior);
            const R f0    = r0 * r0;
            const R fr    = f0 + ((R)1 - f0) * POW_((R)1 - cosf_, (R)5);
            refl          = (R)u2[1] < fr;
        }
        if (refl) {
            out->o = F_(add)(h->point, F_(scl)(out_n, (R)0.0001));
            out->d = F_(reflect)(uvd, out_n);
        } else {
            out->o = F_(add)(h->point, F_(scl)(out_n, (R)-0.0001));
            out->d = F_(sub)(F_(scl)(F_(sub)(uvd, F_(scl)(out_n, dt)), ni_over_nt), F_(scl)(out_n, SQRT_(disc)));
        }
        break;
    }
    case ORC_METAL: {
        const R a  = (R)m->roughness;
        const V f0 = F_(scl)(out->albedo, (R)m->reflectiveness);
        const V wo = F_(scl)(F_(norm)(in_d), (R)-1);
        out->o     = F_(add)(h->point, F_(scl)(nf, (R)0.0001));
        out->nee_n = nf;
        out->f0    = f0;
        if (a < (R)1e-3) {
            out->delta  = 1;
            out->d      = F_(reflect)(in_d, nf);
            out->albedo = F_(schlick)(F_(dot)(wo, nf), f0);
            break;
        }
        out->ggx     = 1;
        out->alpha_r = a;
        V tb, bb;
        F_(onb)(nf, &tb, &bb);
        const V lv = F_(mk)(F_(dot)(wo, tb), F_(dot)(wo, bb), F_(dot)(wo, nf));
        V       ll;
        R       g, voh;
        if (!F_(ggx_sample)(lv, a, (R)u[0], (R)u[1], &ll, &g, &voh)) {
            out->absorbed = 1;
            out->albedo   = F_(mk)(0, 0, 0);
            out->d        = nf;
            break;
        }
        out->d      = F_(norm)(F_(add)(F_(add)(F_(scl)(tb, ll.x), F_(scl)(bb, ll.y)), F_(scl)(nf, ll.z)));
        out->albedo = F_(scl)(F_(schlick)(voh, f0), g);
        break;
    }
    default: {
        const V p  = F_(sphere)(u[0], u[1]);
        out->o     = F_(add)(h->point, F_(scl)(nf, (R)0.0001));
        out->d     = F_(norm)(F_(add)(nf, p));
        out->nee_n = nf;
        break;
    }
    }
}

typedef struct {
    uint64_t ext, shadow, segs;
} F_(cnt);

/* Shadow ray: closest hit with the alpha pass-through loop of renderer.cpp:330-345
 * (bounded at 64 steps).  Returns 1 when the ray escapes. */
static int F_(unoccluded)(const orc_scene *sc, V o, V d, F_(cnt) * c)
{
    F_(hit) h;
    for (int step = 0; step < 64; ++step) {
        c->shadow++;
        if (!F_(cast)(sc, o, d, &h)) return 1;
        const orc_material *m = &sc->mats[sc->mat_idx[h.prim]];
        float               col[4];
        F_(colour)(sc, m, h.uvx, h.uvy, col);
        if (col[3] != 0.0f) return 0;
        o = F_(add)(h.point, F_(scl)(d, (R)0.1));
    }
    return 0;
}

static void F_(sample_pixel)(const orc_scene *sc, const orc_settings *st, uint32_t x, uint32_t y,
                             uint32_t sample, R *raw, float *aov_alb, float *aov_nrm, float *aov_dep,
                             F_(cnt) * c)
{
    const uint32_t W = st->res_x, H = st->res_y;
    const uint32_t pix = y * W + x;
    float          j[4];
    orc_rng4(st->seed, pix, sample, ORC_JITTER_BOUNCE, 0, j);
    V o, d;
    F_(camera)(&sc->cam, (R)(((float)x + j[0]) / (float)W), (R)(((float)y + j[1]) / (float)H), (R)st->aspect, &o, &d);

    V thr = F_(mk)(1, 1, 1), fin = F_(mk)(0, 0, 0), alb = F_(mk)(0, 0, 0), nrm = F_(mk)(0, 0, 0);
    R dep = 0;
    uint32_t total = 1;
    const orc_sun *sun = &sc->sun;
    for (uint32_t i = 0; i < st->max_bounces; ++i, ++total) {
        if (st->rr_start && i >= st->rr_start) {
            float r2[4];
            orc_rng4(st->seed, pix, sample, i, 1, r2);
            R q = F_(maxc)(thr);
            if (q > (R)0.95) q = (R)0.95;
            if ((R)r2[0] >= q) break;
            thr = F_(scl)(thr, (R)1 / q);
        }
        F_(hit) h;
        c->ext++;
        if (!F_(cast)(sc, o, d, &h)) {
            const V s = F_(sky)(sc, d);
            if (i == 0) alb = s;
            fin = F_(add)(fin, F_(mul)(thr, s));
            break;
        }
        float r[4], r2[4];
        orc_rng4(st->seed, pix, sample, i, 0, r);
        orc_rng4(st->seed, pix, sample, i, 1, r2);
        F_(scatter) ph;
        F_(process_hit)(sc, st, &h, d, r, r2, &ph);
        if (ph.alpha) {
            o = F_(add)(h.point, F_(scl)(d, (R)0.1));
            continue;
        }
        if (i == 0) {
            alb = F_(mk)(ph.colour[0], ph.colour[1], ph.colour[2]);
            if (st->mode == ORC_COMPAT) alb = ph.albedo;
            nrm = h.normal;
            dep = h.t;
        }
        const V thr_prev = thr;
        const V in_d     = d;
        if (st->mode == ORC_COMPAT) {
            thr = F_(mul)(thr, ph.albedo);
            fin = F_(add)(fin, F_(scl)(thr, ph.emission));
        } else {
            const V col = F_(mk)(ph.colour[0], ph.colour[1], ph.colour[2]);
            fin = F_(add)(fin, F_(scl)(F_(mul)(thr_prev, col), ph.emission));
            thr = F_(mul)(thr, ph.albedo);
        }
        o = ph.o;
        d = ph.d;
        if (sc->sun_enabled && !(st->mode == ORC_PBR && ph.delta)) {
            /* renderer.cpp:315-354 */
            const V sn = ph.nee_n;
            const V so = F_(add)(h.point, F_(scl)(sn, (R)0.001));
            V       tb, bb;
            F_(onb)(F_(mk)(-sun->dir[0], -sun->dir[1], -sun->dir[2]), &tb, &bb);
            const V nd   = F_(mk)(-sun->dir[0], -sun->dir[1], -sun->dir[2]);
            const R th   = (R)sun->size;
            const R phi  = TAU_R * (R)r[2];
            const R ct   = (R)1 - (R)r[3] * ((R)1 - COS_(th));
            const R stt  = SQRT_((R)1 - ct * ct);
            const V loc  = F_(mk)(COS_(phi) * stt, ct, SIN_(phi) * stt);
            /* mat3(tangent, normal, bitangent) * loc (registry.cpp:142-146) */
            const V sd   = F_(add)(F_(add)(F_(scl)(tb, loc.x), F_(scl)(nd, loc.y)), F_(scl)(bb, loc.z));
            const R pdf  = (R)1 / (TAU_R * ((R)1 - COS_(th)));
            R       cosv = F_(dot)(sn, sd);
            if (cosv < (R)0) cosv = (R)0;
            if (cosv > (R)1) cosv = (R)1;
            if (cosv > (R)0 && F_(unoccluded)(sc, so, sd, c)) {
                const V L = F_(sun_radiance)(sun, sd);
                if (st->mode == ORC_COMPAT) {
                    const V col = F_(mk)(ph.colour[0], ph.colour[1], ph.colour[2]);
                    const V tc = F_(mul)(F_(scl)(F_(mul)(thr, col), cosv), L);
                    fin = F_(add)(fin, F_(mk)(tc.x / pdf, tc.y / pdf, tc.z / pdf));
                } else if (ph.ggx) {
                    /* GGX: f = D * Vis * F with the sampling.h:93-140 formulas */
                    const V wo  = F_(scl)(F_(norm)(in_d), (R)-1);
                    const V hv  = F_(norm)(F_(add)(wo, sd));
                    const R nov = F_(dot)(sn, wo);
                    const R noh = F_(dot)(sn, hv);
                    const R voh = F_(dot)(wo, hv);
                    if (nov > (R)0) {
                        const R dv = F_(ggx_d)(noh, ph.alpha_r) * F_(ggx_vis)(nov, cosv, ph.alpha_r);
                        const V fr = F_(schlick)(voh, ph.f0);
                        fin = F_(add)(fin, F_(scl)(F_(mul)(F_(mul)(thr_prev, fr), L), dv * cosv / pdf));
                    }
                } else {
                    /* Lambert: albedo/pi applied once (thr already holds the albedo) */
                    fin = F_(add)(fin, F_(scl)(F_(mul)(thr, L), cosv / (pdf * PI_R)));
                }
            }
        }
        if (ph.absorbed) break; /* pbr GGX sample under the surface ends the path */
    }
    c->segs += total;
    const uint32_t fx = W - 1 - x, fy = H - 1 - y;
    const size_t   b3 = ((size_t)fx + (size_t)fy * W) * 3;
    raw[b3 + 0] += fin.x;
    raw[b3 + 1] += fin.y;
    raw[b3 + 2] += fin.z;
    const size_t b4 = ((size_t)fx + (size_t)fy * W) * 4;
    if (aov_alb) {
        aov_alb[b4 + 0] = (float)alb.x;
        aov_alb[b4 + 1] = (float)alb.y;
        aov_alb[b4 + 2] = (float)alb.z;
        aov_alb[b4 + 3] = 1.0f;
    }
    if (aov_nrm) {
        aov_nrm[b4 + 0] = (float)(nrm.x * (R)0.5 + (R)0.5);
        aov_nrm[b4 + 1] = (float)(nrm.y * (R)0.5 + (R)0.5);
        aov_nrm[b4 + 2] = (float)(nrm.z * (R)0.5 + (R)0.5);
        aov_nrm[b4 + 3] = 1.0f;
    }
    if (aov_dep) {
        const float dd = (float)((dep < (R)200 ? dep : (R)200) / (R)200);
        aov_dep[b4 + 0] = aov_dep[b4 + 1] = aov_dep[b4 + 2] = dd;
        aov_dep[b4 + 3] = 1.0f;
    }
}

#undef V
