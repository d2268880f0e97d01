/*
 * ORACLE -- test infrastructure only (see orc_trace.c for the rules). Compiled with
 * -ffp-contract=off.
 *
 * CPU restatement of Selas's unidirectional path tracer, PathTracer::EvaluatePath and
 * GenerateImage (Source/Applications/Selas/Source/PathTracer.cpp:95-348), with its helpers:
 *   RayPick / OcclusionRay        PathTracer.cpp:71-139 (rays through the oracle's SAH tree)
 *   CalculateSurfaceParams        Source/Core/Shading/SurfaceParameters.cpp:144-247 (flat-shaded
 *                                 triangles: n = normalize(Ng), no textures)
 *   OffsetRayOrigin               SurfaceParameters.cpp:403-408
 *   NextEventEstimation           Source/Core/Shading/AreaLighting.cpp:353-374
 *   SampleRectangleLightSolidAngle AreaLighting.cpp:113-153
 *   EvaluateBsdf / SampleBsdfFunction via orc_shade.c
 * Deviations, all documented: the sampler is a PCG32 stream per path (stream = path index,
 * initstate = seed ^ (path * 0x9E3779B97F4A7C15)) instead of one mt19937 per thread; samples of a
 * pixel are summed in sample order; media are vacuum (no config has a participating medium);
 * point lights are an addition. With flags & 1 the quad-light estimator drops the extra
 * Dot(normal, l) that Selas multiplies into the light sample (AreaLighting.cpp:140). A BSDF sample
 * of pdf 0 ends the path (SampleDisneyDiffuse, Disney.cpp:494-496, takes diffuse transmission when
 * its uniform is exactly 0 even at diffTrans = 0, returning pdf 0 and an infinite reflectance).
 * Also the timed CPU path of the C1 benchmark ("port" baseline).
 */
#include "oracle.h"
#include "orc_fn.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FLOAT_MAX 3.402823466e+38F

static inline ov3 v3(float x, float y, float z)
{
    ov3 r = {x, y, z};
    return r;
}
static inline ov3 add(ov3 a, ov3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline ov3 sub(ov3 a, ov3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline ov3 mul(ov3 a, ov3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline ov3 scl(ov3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline ov3 neg(ov3 a) { return v3(-a.x, -a.y, -a.z); }
static inline float dot(ov3 a, ov3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
static inline ov3 cross(ov3 a, ov3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline ov3 normalize(ov3 a)
{
    float inv = 1.f / sqrtf(dot(a, a));
    return v3(a.x * inv, a.y * inv, a.z * inv);
}
static inline float absf(float x) { return (x < 0) ? -x : x; }
static inline float maxf(float a, float b) { return (a > b) ? a : b; }
static inline float saturate(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }
static inline ov3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

/* rtcInterpolate0 (SurfaceParameters.cpp:155-190 calls it for slots 0, 1 and 2) as Embree 3
 * evaluates it: on a triangle (p0, p1, p2) fma(w, p0, fma(u, p1, v p2)) with w = 1 - u - v; on a
 * quad (p0, p1, p2, p3) in the triangle of Embree's split that holds (u, v): (p0, p1, p3) with
 * (u, v) when u + v <= 1, else (p2, p3, p1) with (1 - u, 1 - v). */
void orc_interp_attr(const float* a, int K, uint32_t corners, uint32_t prim, float u, float v, float* out)
{
    const size_t base = (size_t)prim * corners;
    size_t i0 = base, i1 = base + 1, i2 = base + 2;
    float U = u, V = v;
    if (corners == 4) {
        if (u + v <= 1.0f) {
            i2 = base + 3;
        } else {
            i0 = base + 2;
            i1 = base + 3;
            i2 = base + 1;
            U = 1.0f - u;
            V = 1.0f - v;
        }
    }
    const float W = 1.0f - U - V;
    for (int k = 0; k < K; ++k) out[k] = fmaf(W, a[i0 * K + k], fmaf(U, a[i1 * K + k], V * a[i2 * K + k]));
}

static void xform_linear(const float* m, const float* n, float* out);

void orc_tangent_frame(const float* tan4, orc_surface* s)
{
    s->t = v3(tan4[0], tan4[1], tan4[2]);
    s->b = scl(cross(s->n, s->t), tan4[3]);
}

/* CalculateSurfaceParams for a triangle hit (SurfaceParameters.cpp:144-247); `texel` (or NULL)
 * replaces the material's baseColor before the 2.2 power. */
int orc_surface_from_hit(const float* o, const float* d, float t, const float* ng, const orc_attr* ga,
                         uint32_t tri_in_geom, float hu, float hv, const float* l2w, const float* mat,
                         const float* texel, orc_surface* s, float* position, float* error, float* tan4)
{
    ov3 p = v3(o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]);
    const float kErr = 32.0f * 1.19209e-07f;
    *error = kErr * maxf(maxf(absf(p.x), absf(p.y)), maxf(absf(p.z), t));
    position[0] = p.x;
    position[1] = p.y;
    position[2] = p.z;
    /* the primitive and its own (u, v): for quads the quad and the coordinates the hit reports */
    uint32_t prim = tri_in_geom;
    float pu = hu, pv = hv;
    if (ga && ga->corners == 4) {
        prim = tri_in_geom >> 1;
        if (tri_in_geom & 1u) {
            pu = 1.0f - hu;
            pv = 1.0f - hv;
        }
    }
    ov3 n;
    if (ga && ga->n) {
        float a[3], w[3];
        orc_interp_attr(ga->n, 3, ga->corners, prim, pu, pv, a);
        if (l2w) {
            xform_linear(l2w, a, w);
            memcpy(a, w, sizeof(a));
        }
        n = normalize(v3(a[0], a[1], a[2]));
    } else {
        n = normalize(v3(ng[0], ng[1], ng[2]));
    }
    float nn[3] = {n.x, n.y, n.z};
    float block[22];
    memcpy(block, mat, sizeof(block));
    const float* bc = texel ? texel : mat;
    block[0] = pow_f(bc[0], 2.2f);
    block[1] = pow_f(bc[1], 2.2f);
    block[2] = pow_f(bc[2], 2.2f);
    block[10] = saturate(mat[10]);
    block[11] = saturate(mat[11]);
    block[12] = mat[12] * 0.5f;
    ov3 view = neg(v3(d[0], d[1], d[2]));
    uint32_t flags = (uint32_t)mat[21];
    block[15] = ((flags & 1u) && dot(view, n) < 0.0f) ? mat[19] : 1.0f / mat[19];
    orc_surface_from_block(block, nn, s);
    if (ga && ga->t) {
        float a[4], w[3];
        orc_interp_attr(ga->t, 4, ga->corners, prim, pu, pv, a);
        if (l2w) {
            xform_linear(l2w, a, w);
            memcpy(a, w, sizeof(w));
        }
        orc_tangent_frame(a, s);
        if (tan4) memcpy(tan4, a, 4 * sizeof(float));
        return 1;
    }
    return 0;
}

/* CalculateSurfaceParams' eUsesPtex branch (SurfaceParameters.cpp:196-203): the texture is
 * looked up at (primId, baryCoords); for Embree quad geometry those are the quad and its own
 * coordinates, i.e. (u, v) on the first triangle of the split and (1 - u, 1 - v) on the second;
 * for triangle geometry, the triangle and its barycentrics. */
const float* orc_hit_texel(const orc_scene* sc, uint32_t tri, uint32_t geom, const float* mat, float hu, float hv,
                           float* texel3)
{
    if (!sc->geom_first) return NULL;
    return orc_hit_texel_local(sc, tri - sc->geom_first[geom], geom, mat, hu, hv, texel3);
}

const float* orc_hit_texel_local(const orc_scene* sc, uint32_t local, uint32_t geom, const float* mat, float hu,
                                 float hv, float* texel3)
{
    if (!(((uint32_t)mat[21]) & 16u) || !sc->geom_tex) return NULL;
    const orc_ptex* tx = &sc->geom_tex[geom];
    if (tx->nfaces == 0) return NULL;
    if (tx->tri) {
        orc_ptex_tri_eval(tx, (int)local, hu, hv, texel3);
        return texel3;
    }
    float u = hu, v = hv;
    if (local & 1u) {
        u = 1.0f - hu;
        v = 1.0f - hv;
    }
    orc_ptex_eval(tx, (int)(local >> 1), u, v, texel3);
    return texel3;
}

/* OffsetRayOrigin (SurfaceParameters.cpp:403-416) */
void orc_offset_origin(const orc_surface* s, const float* pos, float error, const float* dir, float bias, float* out,
                       float* signed_bias)
{
    ov3 n = s->n;
    float off = dot(v3(dir[0], dir[1], dir[2]), n) < 0.0f ? -1.0f : 1.0f;
    float sb = off * error * bias;
    if (signed_bias) *signed_bias = sb;
    out[0] = pos[0] + sb * n.x;
    out[1] = pos[1] + sb * n.y;
    out[2] = pos[2] + sb * n.z;
}

/* NextEventEstimation + SampleRectangleLightSolidAngle; out = dir[3], radiance[3], distance, pdfW */
void orc_nee(orc_uniforms* u, const orc_light* lights, uint32_t n_lights, const float* position, const float* normal,
             int cosine_once, float* out)
{
    memset(out, 0, 8 * sizeof(float));
    out[7] = 1.0f;
    if (n_lights == 0) return;
    float per_light = 1.0f / n_lights;
    float p0 = orc_uniform(u);
    uint32_t idx = (uint32_t)(p0 * n_lights);
    if (idx >= n_lights) idx = n_lights - 1;
    const orc_light* L = &lights[idx];
    ov3 pos = ld3(position), nrm = ld3(normal);
    if (L->type == 1) {
        ov3 ul = sub(ld3(L->position), pos);
        float dist_sq = dot(ul, ul);
        float dist = sqrtf(dist_sq);
        ov3 l = scl(ul, 1.0f / dist);
        ov3 rad = scl(ld3(L->radiance), 1.0f / dist_sq);
        out[0] = l.x;
        out[1] = l.y;
        out[2] = l.z;
        out[3] = rad.x;
        out[4] = rad.y;
        out[5] = rad.z;
        out[6] = dist;
        out[7] = 1.0f * per_light;
        return;
    }
    ov3 ex = ld3(L->x), ez = ld3(L->z);
    ov3 s = sub(sub(ld3(L->position), scl(ex, 0.5f)), scl(ez, 0.5f));
    float q[12] = {s.x, s.y, s.z, ex.x, ex.y, ex.z, ez.x, ez.y, ez.z, pos.x, pos.y, pos.z};
    float uu = orc_uniform(u);
    float vv = orc_uniform(u);
    float r[5];
    orc_rect_light(q, uu, vv, r);
    float pdf = 1.0f / r[0];
    ov3 lp = v3(r[2], r[3], r[4]);
    ov3 ul = sub(lp, pos);
    float dist_sq = dot(ul, ul);
    float dist = sqrtf(dist_sq);
    ov3 l = scl(ul, 1.0f / dist);
    float am = dot(ld3(L->direction), neg(l)) / dist_sq;
    if (am > 0) {
        ov3 rad = cosine_once ? scl(ld3(L->radiance), am) : scl(scl(ld3(L->radiance), am), dot(nrm, l));
        out[0] = l.x;
        out[1] = l.y;
        out[2] = l.z;
        out[3] = rad.x;
        out[4] = rad.y;
        out[5] = rad.z;
        out[6] = dist;
        out[7] = pdf * am * per_light;
    } else {
        out[7] = 1.0f * per_light;
    }
}

typedef struct {
    const orc_scene* sc;
    const orc_pt_job* job;
    orc_camera cam;
    float* fb;
    uint32_t npix;
} pt_ctx;

static void bsdf_eval_any(const orc_surface* s, const float* v, const float* l, float* out5)
{
    if (s->shader == 2)
        orc_lambert_eval(s, v, l, out5);
    else
        orc_disney_eval(s, v, l, s->shader == 0, out5);
}

static int bsdf_sample_any(orc_uniforms* u, const orc_surface* s, const float* v, float* out14)
{
    memset(out14, 0, 14 * sizeof(float));
    if (s->shader == 2) return orc_lambert_sample(u, s, v, out14);
    return orc_disney_sample(u, s, v, s->shader == 0, out14);
}

/* ---- scene queries: a scene of geometries, or the composed instances of an instanced scene ---- */

static void tri_normal(const float* tp, float* ng)
{
    float e1[3] = {tp[3] - tp[0], tp[4] - tp[1], tp[5] - tp[2]};
    float e2[3] = {tp[6] - tp[0], tp[7] - tp[1], tp[8] - tp[2]};
    ov3 n = cross(ld3(e1), ld3(e2));
    ng[0] = n.x;
    ng[1] = n.y;
    ng[2] = n.z;
}

/* world = m n for the row-major 3x3 m, by the device's fused multiply-add chain */
static void xform_linear(const float* m, const float* n, float* out)
{
    for (int r = 0; r < 3; ++r) out[r] = fmaf(m[3 * r], n[0], fmaf(m[3 * r + 1], n[1], m[3 * r + 2] * n[2]));
}

static void inst_ray(const float* w, const float* o, const float* d, float* lo, float* ld)
{
    for (int r = 0; r < 3; ++r) {
        const float* row = w + 4 * r;
        lo[r] = fmaf(row[0], o[0], fmaf(row[1], o[1], fmaf(row[2], o[2], row[3])));
        ld[r] = fmaf(row[0], d[0], fmaf(row[1], d[1], row[2] * d[2]));
    }
}

int orc_scene_closest1(const orc_scene* sc, uint32_t* stack, const float* o, const float* d, float tnear, float tfar,
                       float* t, float* u, float* v, uint32_t* geom, float* ng, uint32_t* tri_in_geom,
                       const float** l2w)
{
    if (l2w) *l2w = NULL;
    if (!sc->inst) {
        uint32_t prim = orc_bvh_closest1(sc->bvh, stack, o, d, tnear, tfar, t, u, v);
        int hit = prim != 0xFFFFFFFFu;
        if (hit) {
            *geom = sc->tri_geom[prim];
            tri_normal(sc->tri9 + (size_t)prim * 9, ng);
            *tri_in_geom = sc->geom_first ? prim - sc->geom_first[*geom] : prim;
        }
        /* the curves after the triangles: a strictly closer curve hit replaces the triangle's */
        float tc, uc, nc[3];
        uint32_t pc;
        if (sc->n_curves && orc_curves_closest1(sc->curve_cp, sc->curve_box, sc->n_curves, o, d, tnear,
                                                hit ? *t : tfar, 0, &tc, &uc, &pc, nc) &&
            (!hit || tc < *t)) {
            *t = tc;
            *u = uc;
            *v = 0.0f;
            *geom = sc->curve_geom[pc];
            *tri_in_geom = sc->curve_prim[pc];
            for (int k = 0; k < 3; ++k) ng[k] = nc[k];
            hit = 1;
        }
        return hit;
    }
    const orc_iscene* is = sc->inst;
    float best_t = tfar, bu = 0.0f, bv = 0.0f;
    uint32_t best = 0xFFFFFFFFu, best_k = 0;
    for (uint32_t k = 0; k < is->n_inst; ++k) {
        float lo[3], ldir[3], tk, uk, vk;
        inst_ray(is->insts[k].w2l, o, d, lo, ldir);
        uint32_t p = orc_bvh_closest1(is->insts[k].blas, stack, lo, ldir, tnear, best_t, &tk, &uk, &vk);
        if (p != 0xFFFFFFFFu && (best == 0xFFFFFFFFu || tk < best_t)) {
            best_t = tk;
            best = p;
            best_k = k;
            bu = uk;
            bv = vk;
        }
    }
    /* curves of the composed instances, in their own space (t is the world ray's parameter) */
    uint32_t cbest = 0xFFFFFFFFu, cbest_k = 0;
    float cng[3] = {0.0f, 0.0f, 0.0f};
    if (is->n_curves)
        for (uint32_t k = 0; k < is->n_comp; ++k) {
            if (!is->n_curves[k]) continue;
            float lo[3], ldir[3], tk, uk, nk[3];
            uint32_t pk;
            inst_ray(is->insts[k].w2l, o, d, lo, ldir);
            const int found = best != 0xFFFFFFFFu || cbest != 0xFFFFFFFFu;
            if (orc_curves_closest1(is->curve_cp[k], is->curve_box[k], is->n_curves[k], lo, ldir, tnear, best_t, 0, &tk,
                                    &uk, &pk, nk) &&
                (!found || tk < best_t)) {
                best_t = tk;
                bu = uk;
                bv = 0.0f;
                cbest = pk;
                cbest_k = k;
                for (int q = 0; q < 3; ++q) cng[q] = nk[q];
            }
        }
    *t = best_t;
    if (cbest != 0xFFFFFFFFu) {
        *u = bu;
        *v = bv;
        *geom = is->geom_base[cbest_k] + is->curve_geom[cbest_k][cbest];
        *tri_in_geom = is->curve_prim[cbest_k][cbest];
        xform_linear(is->l2w + 9 * (size_t)cbest_k, cng, ng);
        if (l2w) *l2w = is->l2w + 9 * (size_t)cbest_k;
        return 1;
    }
    if (best == 0xFFFFFFFFu) return 0;
    *u = bu;
    *v = bv;
    const uint32_t lg = is->tri_geom[best_k][best];
    *geom = is->geom_base[best_k] + lg;
    *tri_in_geom = best - is->geom_first[best_k][lg];
    float nl[3];
    tri_normal(is->tri9[best_k] + (size_t)best * 9, nl);
    xform_linear(is->l2w + 9 * (size_t)best_k, nl, ng);
    if (l2w) *l2w = is->l2w + 9 * (size_t)best_k;
    return 1;
}

int orc_scene_occluded1(const orc_scene* sc, uint32_t* stack, const float* o, const float* d, float tnear, float tfar)
{
    if (!sc->inst)
        return orc_bvh_occluded1(sc->bvh, stack, o, d, tnear, tfar) ||
               (sc->n_curves && orc_curves_occluded1(sc->curve_cp, sc->curve_box, sc->n_curves, o, d, tnear, tfar));
    const orc_iscene* is = sc->inst;
    for (uint32_t k = 0; k < is->n_inst; ++k) {
        float lo[3], ldir[3];
        inst_ray(is->insts[k].w2l, o, d, lo, ldir);
        if (orc_bvh_occluded1(is->insts[k].blas, stack, lo, ldir, tnear, tfar)) return 1;
    }
    if (is->n_curves)
        for (uint32_t k = 0; k < is->n_comp; ++k) {
            if (!is->n_curves[k]) continue;
            float lo[3], ldir[3];
            inst_ray(is->insts[k].w2l, o, d, lo, ldir);
            if (orc_curves_occluded1(is->curve_cp[k], is->curve_box[k], is->n_curves[k], lo, ldir, tnear, tfar))
                return 1;
        }
    return 0;
}

uint32_t orc_scene_stack_size(const orc_scene* sc)
{
    if (!sc->inst) return orc_bvh_stack_size(sc->bvh);
    uint32_t m = 4;
    for (uint32_t k = 0; k < sc->inst->n_inst; ++k) {
        const uint32_t s = orc_bvh_stack_size(sc->inst->insts[k].blas);
        if (s > m) m = s;
    }
    return m;
}

/* EvaluatePath for one (pixel, sample); writes ld[6] = layer0 rgb, layer1 rgb */
static void evaluate_path(const pt_ctx* c, uint32_t* stack, uint32_t x, uint32_t y, uint64_t path, float* ld)
{
    const orc_scene* sc = c->sc;
    const orc_pt_job* job = c->job;
    uint64_t st, inc;
    orc_pcg32_seed(job->seed ^ (path * 0x9E3779B97F4A7C15ull), path, &st, &inc);
    orc_uniforms u = {NULL, 0, 0, &st, inc};
    ov3 l0 = v3(0, 0, 0), l1 = v3(0, 0, 0), thr = v3(1, 1, 1);
    float vx = (float)x + orc_uniform(&u);
    float vy = (float)y + orc_uniform(&u);
    float o[3], d[3];
    o[0] = c->cam.pos.x;
    o[1] = c->cam.pos.y;
    o[2] = c->cam.pos.z;
    orc_camera_dir(&c->cam, vx, vy, d);
    uint32_t bounce = 0;
    while (bounce < job->max_path_length) {
        float t, hu, hv, ng[3];
        uint32_t geom, tri_in_geom;
        const float* l2w = NULL;
        if (!orc_scene_closest1(sc, stack, o, d, 0.0f, FLOAT_MAX, &t, &hu, &hv, &geom, ng, &tri_in_geom, &l2w)) {
            /* EvaluateBackground (AreaLighting.cpp:415-424): the image-based light, else the
             * constant background */
            float bg[3];
            if (sc->ibl)
                (void)orc_ibl_lookup(sc->ibl, d, bg);
            else
                memcpy(bg, sc->background, sizeof(bg));
            l1 = add(l1, mul(ld3(bg), thr));
            break;
        }
        const float* mat = sc->materials + 22 * (size_t)sc->geom_material[geom];
        orc_surface s;
        float pos[3], err;
        float texel[3];
        orc_surface_from_hit(o, d, t, ng, sc->geom_attr ? &sc->geom_attr[geom] : NULL, tri_in_geom, hu, hv, l2w, mat,
                             orc_hit_texel_local(sc, tri_in_geom, geom, mat, hu, hv, texel), &s, pos, &err, NULL);
        float v[3] = {-d[0], -d[1], -d[2]};
        float nrm[3] = {s.n.x, s.n.y, s.n.z};
        float ls[8];
        orc_nee(&u, sc->lights, sc->n_lights, pos, nrm, (job->flags & 1u) != 0, ls);
        if (dot(ld3(ls + 3), v3(1, 1, 1)) > 0) {
            float f[5];
            bsdf_eval_any(&s, v, ls, f);
            if (dot(ld3(f), v3(1, 1, 1)) > 0) {
                float so[3];
                orc_offset_origin(&s, pos, err, ls, 0.1f, so, NULL);
                if (!orc_scene_occluded1(sc, stack, so, ls, err, ls[6])) {
                    ov3 smp = scl(mul(ld3(f), ld3(ls + 3)), 1.0f / ls[7]);
                    l1 = add(l1, mul(smp, thr));
                }
            }
        }
        float bs[14];
        if (!bsdf_sample_any(&u, &s, v, bs) || !(bs[8] > 0.0f)) break; /* pdf 0 ends the path */
        thr = mul(scl(thr, 1.0f), ld3(bs + 2));
        float no[3];
        orc_offset_origin(&s, pos, err, bs + 5, 1.0f, no, NULL);
        o[0] = no[0];
        o[1] = no[1];
        o[2] = no[2];
        d[0] = bs[5];
        d[1] = bs[6];
        d[2] = bs[7];
        if (bounce == 0) l0 = ld3(bs + 2);
        ++bounce;
        if (bounce > 8 && bounce < job->max_path_length) {
            float cp = maxf(maxf(thr.x, thr.y), thr.z);
            if (orc_uniform(&u) > cp) break;
            thr = scl(thr, 1.0f / cp);
        }
    }
    ld[0] = l0.x;
    ld[1] = l0.y;
    ld[2] = l0.z;
    ld[3] = l1.x;
    ld[4] = l1.y;
    ld[5] = l1.z;
}

static void pt_range(void* cv, uint32_t begin, uint32_t end)
{
    const pt_ctx* c = (const pt_ctx*)cv;
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * orc_scene_stack_size(c->sc));
    const uint32_t W = c->job->width, spp = c->job->spp;
    for (uint32_t pix = begin; pix < end; ++pix) {
        uint32_t x = pix % W, y = pix / W;
        ov3 a0 = v3(0, 0, 0), a1 = v3(0, 0, 0);
        for (uint32_t s = 0; s < spp; ++s) {
            float ld[6];
            evaluate_path(c, stack, x, y, (uint64_t)pix * spp + s, ld);
            a0 = add(a0, ld3(ld));
            a1 = add(a1, ld3(ld + 3));
        }
        float inv = 1.0f / (float)spp;
        float* f0 = c->fb + 3 * (size_t)pix;
        float* f1 = c->fb + 3 * ((size_t)c->npix + pix);
        /* zero + sum, then FrameBuffer_Scale */
        f0[0] = (0.0f + a0.x) * inv;
        f0[1] = (0.0f + a0.y) * inv;
        f0[2] = (0.0f + a0.z) * inv;
        f1[0] = (0.0f + a1.x) * inv;
        f1[1] = (0.0f + a1.y) * inv;
        f1[2] = (0.0f + a1.z) * inv;
    }
    free(stack);
}

void orc_render_pt(const orc_scene* sc, const orc_pt_job* job, float* fb, int threads)
{
    pt_ctx c;
    c.sc = sc;
    c.job = job;
    orc_camera_init(job->cam10, job->width, job->height, &c.cam);
    c.fb = fb;
    c.npix = job->width * job->height;
    orc_parallel_for(pt_range, &c, c.npix, threads);
}

/* ---- the deferred path tracer's estimator --------------------------------------------------------
 * DeferredPathTracer::ShadeHitPosition / TraceRayBatch / TraceOcclusionBatch / GeneratePrimaryRays
 * (Source/Applications/Selas/Source/DeferredPathTracer.cpp:61-351). Per path, in this order of
 * random numbers: at each hit NEE toward the quad and point lights (light pick + 2, weight 1:
 * lights are not hittable), a sky sample (2: SampleBackground, AreaLighting.cpp:293-314, 334-350)
 * weighted by the balance heuristic against the BSDF's pdf, a BSDF sample (1 + 2..3), and after 7
 * tracked bounces Russian roulette (1) on the largest throughput component. Primary rays are
 * CMJ-jittered (JitteredCameraRay with s of spp_x x spp_y, pattern = pixel index; no random
 * numbers). A ray leaving the scene adds the sky (EvaluateBackground) times the throughput, weighted
 * by the balance heuristic of the BSDF pdf it was sampled with against the sky's pdf; unweighted
 * while every scatter so far was a Dirac event, or when the last one was. Deviations, documented:
 * the MIS weight of the continuation multiplies the sky it finds, not the path's throughput (Selas
 * weights the throughput, which darkens all later lighting of the path); the constant background
 * is sampled with its pdf 1/(4 pi) (BackgroundLightSample sets 1 while BackgroundLightingPdf
 * returns 1/(4 pi)); max_path_length bounds the segments (Selas ends paths by roulette only);
 * one PCG32 stream per path as in orc_render_pt. Shadow rays start at tnear 0 from the origin
 * offset by 0.1 error, as TraceOcclusionBatch sets them. One output layer. */
typedef struct {
    const orc_scene* sc;
    const orc_dpt_job* job;
    orc_camera cam;
    float* fb;
} dpt_ctx;

#define DPT_INV_4PI 0.0795774715459476678844f
static const float kDptPi = 3.141592653589793f;

/* SampleBackground: direction, radiance, solid-angle pdf */
static void dpt_sky_sample(const orc_scene* sc, orc_uniforms* u, float* dir, float* rad, float* pdf)
{
    const float r0 = orc_uniform(u), r1 = orc_uniform(u);
    if (sc->ibl) {
        float s8[8];
        orc_ibl_direct(sc->ibl, r0, r1, s8);
        memcpy(dir, s8, 3 * sizeof(float));
        memcpy(rad, s8 + 3, 3 * sizeof(float));
        *pdf = s8[6];
        return;
    }
    const float uu = 2.0f * r1 - 1.0f;
    const float norm = sqrtf(maxf(0.0f, 1.0f - uu * uu));
    const float theta = (kDptPi * 2.0f) * r0;
    dir[0] = norm * cos_f(theta);
    dir[1] = uu;
    dir[2] = norm * sin_f(theta);
    memcpy(rad, sc->background, 3 * sizeof(float));
    *pdf = DPT_INV_4PI;
}

/* EvaluateBackground + BackgroundLightingPdf */
static float dpt_sky_lookup(const orc_scene* sc, const float* d, float* rad)
{
    if (sc->ibl) return orc_ibl_lookup(sc->ibl, d, rad);
    memcpy(rad, sc->background, 3 * sizeof(float));
    return DPT_INV_4PI;
}

static ov3 dpt_path(const dpt_ctx* c, uint32_t* stack, uint32_t x, uint32_t y, uint32_t s, uint64_t path)
{
    const orc_scene* sc = c->sc;
    const orc_dpt_job* job = c->job;
    uint64_t st, inc;
    orc_pcg32_seed(job->seed ^ (path * 0x9E3779B97F4A7C15ull), path, &st, &inc);
    orc_uniforms u = {NULL, 0, 0, &st, inc};
    ov3 l = v3(0, 0, 0), thr = v3(1, 1, 1);
    float o[3] = {c->cam.pos.x, c->cam.pos.y, c->cam.pos.z}, d[3];
    const int32_t pixel = (int32_t)(y * job->width + x);
    orc_camera_cmj_dir(&c->cam, (int32_t)x, (int32_t)y, (int32_t)s, (int32_t)job->spp_x, (int32_t)job->spp_y, pixel, d);
    int dirac_only = 1, last_dirac = 1;
    float last_pdf = 0.0f;
    for (uint32_t bounce = 0; bounce < job->max_path_length; ++bounce) {
        float t, hu, hv, ng[3];
        uint32_t geom, tri_in_geom;
        const float* l2w = NULL;
        if (!orc_scene_closest1(sc, stack, o, d, 0.0f, FLOAT_MAX, &t, &hu, &hv, &geom, ng, &tri_in_geom, &l2w)) {
            float bg[3];
            const float sky_pdf = dpt_sky_lookup(sc, d, bg);
            float w = 1.0f;
            if (!dirac_only && !last_dirac) w = last_pdf + sky_pdf > 0.0f ? last_pdf / (last_pdf + sky_pdf) : 0.0f;
            l = add(l, mul(scl(ld3(bg), w), thr));
            break;
        }
        const float* mat = sc->materials + 22 * (size_t)sc->geom_material[geom];
        orc_surface sf;
        float pos[3], err, texel[3];
        orc_surface_from_hit(o, d, t, ng, sc->geom_attr ? &sc->geom_attr[geom] : NULL, tri_in_geom, hu, hv, l2w, mat,
                             orc_hit_texel_local(sc, tri_in_geom, geom, mat, hu, hv, texel), &sf, pos, &err, NULL);
        float v[3] = {-d[0], -d[1], -d[2]};
        float nrm[3] = {sf.n.x, sf.n.y, sf.n.z};
        /* NEE toward the lights */
        float ls[8];
        orc_nee(&u, sc->lights, sc->n_lights, pos, nrm, 1, ls);
        if (dot(ld3(ls + 3), v3(1, 1, 1)) > 0) {
            float f[5];
            bsdf_eval_any(&sf, v, ls, f);
            const ov3 smp = scl(mul(ld3(f), ld3(ls + 3)), 1.0f / ls[7]);
            if (dot(smp, v3(1, 1, 1)) > 0) {
                float so[3];
                orc_offset_origin(&sf, pos, err, ls, 0.1f, so, NULL);
                if (!orc_scene_occluded1(sc, stack, so, ls, 0.0f, ls[6])) l = add(l, mul(smp, thr));
            }
        }
        /* a sky sample, MIS-weighted against the BSDF */
        float sd[3], srad[3], spdf;
        dpt_sky_sample(sc, &u, sd, srad, &spdf);
        if (dot(ld3(srad), v3(1, 1, 1)) > 0 && spdf > 0.0f) {
            float f[5];
            bsdf_eval_any(&sf, v, sd, f);
            const float w = spdf / (spdf + f[3]);
            const ov3 smp = scl(mul(scl(ld3(f), w), ld3(srad)), 1.0f / spdf);
            if (dot(smp, v3(1, 1, 1)) > 0) {
                float so[3];
                orc_offset_origin(&sf, pos, err, sd, 0.1f, so, NULL);
                if (!orc_scene_occluded1(sc, stack, so, sd, 0.0f, 1e36f)) l = add(l, mul(smp, thr));
            }
        }
        /* the BSDF sample continues the path */
        float bs[14];
        if (!bsdf_sample_any(&u, &sf, v, bs) || !(bs[8] > 0.0f)) break; /* pdf 0 ends the path */
        thr = mul(thr, ld3(bs + 2));
        if (dot(thr, thr) == 0.0f) break;
        const int dirac = (((uint32_t)bs[1]) & 4u) != 0;
        dirac_only = dirac_only && dirac;
        last_dirac = dirac;
        last_pdf = bs[8];
        if (bounce >= 7) {
            const float cp = maxf(maxf(thr.x, thr.y), thr.z);
            if (orc_uniform(&u) >= cp) break;
            thr = scl(thr, 1.0f / cp);
        }
        float no[3];
        orc_offset_origin(&sf, pos, err, bs + 5, 1.0f, no, NULL);
        memcpy(o, no, sizeof(no));
        memcpy(d, bs + 5, 3 * sizeof(float));
    }
    return l;
}

static void dpt_range(void* cv, uint32_t begin, uint32_t end)
{
    const dpt_ctx* c = (const dpt_ctx*)cv;
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * orc_scene_stack_size(c->sc));
    const uint32_t W = c->job->width, spp = c->job->spp_x * c->job->spp_y;
    for (uint32_t pix = begin; pix < end; ++pix) {
        const uint32_t x = pix % W, y = pix / W;
        ov3 a = v3(0, 0, 0);
        for (uint32_t s = 0; s < spp; ++s) a = add(a, dpt_path(c, stack, x, y, s, (uint64_t)pix * spp + s));
        const float inv = 1.0f / (float)spp;
        float* f = c->fb + 3 * (size_t)pix;
        f[0] = (0.0f + a.x) * inv;
        f[1] = (0.0f + a.y) * inv;
        f[2] = (0.0f + a.z) * inv;
    }
    free(stack);
}

void orc_render_dpt(const orc_scene* sc, const orc_dpt_job* job, float* fb, int threads)
{
    dpt_ctx c;
    c.sc = sc;
    c.job = job;
    orc_camera_init(job->cam10, job->width, job->height, &c.cam);
    c.fb = fb;
    orc_parallel_for(dpt_range, &c, job->width * job->height, threads);
}
