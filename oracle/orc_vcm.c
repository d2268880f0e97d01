/*
 * ORACLE -- test infrastructure only (see orc_trace.c for the rules). Compiled with
 * -ffp-contract=off.
 *
 * CPU restatement of Selas's vertex connection and merging integrator
 * (Source/Applications/Selas/Source/VCM.cpp, whose code is commented out and refers to symbols that
 * no longer exist, so it is restated rather than run), with its live helpers:
 *   CalculateIterationConstants / SearchRadius   VCMCommon.cpp:91-113
 *   GenerateCameraSample                         VCMCommon.cpp:66-88
 *   BuildHashGrid / SearchHashGrid               VCMHashGrid.cpp:15-157 (restated exactly:
 *                                                pinned bit-for-bit by golden vectors)
 *   EvaluateLightPaths / EvaluateCameraPaths     VCM.cpp:502-666
 *   ConnectLightPathToCamera                     VCM.cpp:190-242
 *   ConnectCameraPathToLight                     VCM.cpp:245-290
 *   ConnectPathVertices                          VCM.cpp:293-359
 *   MergeVertices                                VCM.cpp:362-409
 *   SampleBsdfScattering                         VCM.cpp:412-446
 * The MIS recurrences (dVCM, dVC, dVM) are those of the reference, which follow Georgiev et al.,
 * "Implementing Vertex Connection and Merging" (2012). Where the commented code cannot be run as
 * written, or is inconsistent with that report, the report's form is used (each is listed in the
 * repository's design notes): WorldToImage is defined by inverting ImageToWorldDirection; the
 * virtual image plane distance is W / (2 tan(hfov/2)); the light-tracing image-to-surface factor
 * includes |cos| / d^2; BSDF values are Selas's (f |cos l|), so connections and merges remove the
 * cosines the report does not have; throughput is divided by the Russian-roulette probability;
 * lights are quad and point lights (Selas's VCM emits only from the IBL) and, as in Selas's path
 * tracer, they are not intersectable, so the technique "camera path hits the light" does not exist
 * and its MIS terms are zero (dVC = dVM = 0 at the light, no light weight in NEE) -- exactly the
 * treatment the report gives delta lights; NEE samples quad lights by area (a point-independent
 * pdf, as the MIS weights require). Random numbers: one PCG32 stream per (iteration, path, pass).
 * A BSDF sample of pdf 0 ends the path (SampleDisneyDiffuse, Disney.cpp:494-496, can return one
 * with an infinite reflectance).
 * Also the timed CPU path of the VCM benchmark ("port" baseline).
 */
#include "oracle.h"
#include "orc_fn.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const float kPi = 3.141592653589793f;

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
static inline float length(ov3 a) { return sqrtf(dot(a, a)); }
static inline ov3 normalize(ov3 a)
{
    float inv = 1.f / length(a);
    return v3(a.x * inv, a.y * inv, a.z * inv);
}
static inline float absf(float x) { return (x < 0) ? -x : x; }
static inline float maxf(float a, float b) { return (a > b) ? a : b; }
static inline float saturate(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }
static inline ov3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }
static inline int is_zero(ov3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }

/* ---- iteration constants (VCMCommon.cpp:91-113) ------------------------------------------------ */

void orc_vcm_constants(uint32_t vm_count, uint32_t vc_count, float base_radius, float alpha, float iteration,
                       float* out6)
{
    float r = base_radius / pow_f(iteration, 0.5f * (1.0f - alpha));
    float r2 = r * r;
    out6[0] = r;
    out6[1] = r2;
    out6[2] = 1.0f / (kPi * r2 * vm_count);       /* vmNormalization */
    out6[3] = kPi * r2 * vm_count / vc_count;      /* vmWeight */
    out6[4] = vc_count / (kPi * r2 * vm_count);    /* vcWeight */
    out6[5] = 0.0f;
}

/* ---- hash grid (VCMHashGrid.cpp, restated exactly) ------------------------------------------- */

static uint32_t cell_hash(int32_t x, int32_t y, int32_t z, uint32_t cell_count)
{
    uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
    return (uint32_t)(int)(((ux * 73856093u) ^ (uy * 19349663u) ^ (uz * 83492791u)) % cell_count);
}

static uint32_t cell_of(const orc_hash_grid* g, const float* p)
{
    int32_t x = (int32_t)floorf(g->inv_cell * (p[0] - g->lo[0]));
    int32_t y = (int32_t)floorf(g->inv_cell * (p[1] - g->lo[1]));
    int32_t z = (int32_t)floorf(g->inv_cell * (p[2] - g->lo[2]));
    return cell_hash(x, y, z, g->cell_count);
}

void orc_hash_grid_build(orc_hash_grid* g, const float* points, uint32_t count, uint32_t cell_count, float radius)
{
    g->radius = radius;
    g->radius_sq = radius * radius;
    float cell = 2.0f * radius;
    g->inv_cell = 1.0f / cell;
    g->cell_count = cell_count;
    g->count = count;
    g->cell_ends = (uint32_t*)calloc(cell_count, sizeof(uint32_t));
    g->indices = (uint32_t*)malloc(sizeof(uint32_t) * (count ? count : 1));
    for (int k = 0; k < 3; ++k) {
        g->lo[k] = 3.402823466e+38F;
        g->hi[k] = -3.402823466e+38F;
    }
    for (uint32_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            float v = points[3 * i + k];
            /* IncludePosition (AxisAlignedBox.cpp:32): Min / Max */
            g->lo[k] = v < g->lo[k] ? v : g->lo[k];
            g->hi[k] = v > g->hi[k] ? v : g->hi[k];
        }
    for (uint32_t i = 0; i < count; ++i) ++g->cell_ends[cell_of(g, points + 3 * i)];
    uint32_t sum = 0;
    for (uint32_t c = 0; c < cell_count; ++c) {
        uint32_t n = g->cell_ends[c];
        g->cell_ends[c] = sum;
        sum += n;
    }
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t c = cell_of(g, points + 3 * i);
        g->indices[g->cell_ends[c]++] = i;
    }
}

void orc_hash_grid_free(orc_hash_grid* g)
{
    free(g->cell_ends);
    free(g->indices);
    g->cell_ends = NULL;
    g->indices = NULL;
}

/* SearchHashGrid: calls fn(ctx, index) for each point within radius of q, in the reference's order
 * (8 cells, x fastest; a point is visited twice when two of the cells hash to the same bucket). */
typedef void (*grid_visit_fn)(void* ctx, uint32_t index);

static void hash_grid_visit(const orc_hash_grid* g, const float* points, const float* q, grid_visit_fn fn, void* ctx)
{
    float dmin[3], dmax[3];
    for (int k = 0; k < 3; ++k) {
        dmin[k] = q[k] - g->lo[k];
        dmax[k] = g->hi[k] - q[k];
    }
    if (dmin[0] < 0.0f || dmin[1] < 0.0f || dmin[2] < 0.0f) return;
    if (dmax[0] < 0.0f || dmax[1] < 0.0f || dmax[2] < 0.0f) return;
    int32_t mn[3];
    for (int k = 0; k < 3; ++k) {
        float cp = g->inv_cell * dmin[k];
        float cf = floorf(cp);
        float fr = cp - cf;
        mn[k] = (int32_t)cf;
        if (fr < 0.5f) --mn[k];
    }
    for (int z = 0; z <= 1; ++z)
        for (int y = 0; y <= 1; ++y)
            for (int x = 0; x <= 1; ++x) {
                uint32_t c = cell_hash(mn[0] + x, mn[1] + y, mn[2] + z, g->cell_count);
                uint32_t b = c == 0 ? 0 : g->cell_ends[c - 1], e = g->cell_ends[c];
                for (uint32_t s = b; s < e; ++s) {
                    uint32_t pi = g->indices[s];
                    const float* p = points + 3 * (size_t)pi;
                    ov3 d = sub(ld3(q), ld3(p));
                    if (dot(d, d) <= g->radius_sq) fn(ctx, pi);
                }
            }
}

typedef struct {
    uint32_t* found;
    uint32_t max_found, n;
} found_list;

static void collect_found(void* ctx, uint32_t index)
{
    found_list* f = (found_list*)ctx;
    if (f->n < f->max_found) f->found[f->n] = index;
    ++f->n;
}

/* The points within radius of q in the reference's order. Returns the number found; writes up to
 * max_found indices. */
uint32_t orc_hash_grid_query(const orc_hash_grid* g, const float* points, const float* q, uint32_t* found,
                             uint32_t max_found)
{
    found_list f = {found, max_found, 0};
    hash_grid_visit(g, points, q, collect_found, &f);
    return f.n;
}

/* ---- camera projection ---------------------------------------------------------------------- */

/* WorldToImage: the inverse of ImageToWorldDirection (Camera.cpp:56-61); pixel coordinates as
 * floats (x right, y down), returns 0 when the point is behind the camera. */
static int world_to_image(const orc_camera* c, ov3 p, float* px, float* py)
{
    ov3 d = sub(p, c->pos);
    float zc = dot(d, c->z);
    if (!(zc > 0.0f)) return 0;
    float cx = -dot(d, c->x) / (dot(c->x, c->x) * zc);
    float cy = dot(d, c->y) / (dot(c->y, c->y) * zc);
    *px = (cx + 1.0f) * 0.5f * (float)c->width;
    *py = (1.0f - cy) * 0.5f * (float)c->height;
    return 1;
}

/* ---- light emission and direct sampling ------------------------------------------------------- */

typedef struct {
    ov3 position, direction, throughput;
    float dvcm, dvc, dvm;
    int infinite; /* the subpath starts at the image-based light */
} emit_state;

/* Light subpath start (GenerateLightSample, VCMCommon.cpp:37-63): one light picked uniformly among
 * the quad and point lights and the image-based light (light n_lights). Uses 5 uniforms for quad
 * and point lights (pick, position 2 -- quad only --, direction 2), 5 for the image-based light
 * (pick, then EmitIblLightSample's 4, AreaLighting.cpp:246-290). Returns 0 when the sample carries
 * no light (a zero emission pdf). */
static int emit_light(orc_uniforms* u, const orc_scene* sc, float vc_weight, emit_state* s)
{
    const uint32_t n_lights = sc->n_lights, n_pick = n_lights + (sc->ibl ? 1u : 0u);
    float p_pick = 1.0f / n_pick;
    uint32_t idx = (uint32_t)(orc_uniform(u) * n_pick);
    if (idx >= n_pick) idx = n_pick - 1;
    float emission_pdf_w, direct_pdf_a;
    ov3 radiance;
    s->infinite = idx == n_lights;
    if (s->infinite) {
        float r0 = orc_uniform(u), r1 = orc_uniform(u), r2 = orc_uniform(u), r3 = orc_uniform(u);
        float e[14];
        orc_ibl_emit(sc->ibl, r0, r1, r2, r3, e);
        s->position = v3(e[0], e[1], e[2]);
        s->direction = v3(e[3], e[4], e[5]);
        radiance = v3(e[6], e[7], e[8]);
        emission_pdf_w = e[9] * p_pick;
        direct_pdf_a = e[10] * p_pick;
    } else {
        const orc_light* L = &sc->lights[idx];
        if (L->type == 1) {
            float r0 = orc_uniform(u), r1 = orc_uniform(u);
            /* uniform sphere (CSampler::UniformSphere, Sampler.cpp) */
            float uu = 2.0f * r1 - 1.0f;
            float norm = sqrtf(maxf(0.0f, 1.0f - uu * uu));
            float theta = (kPi * 2.0f) * r0;
            s->direction = v3(norm * cos_f(theta), uu, norm * sin_f(theta));
            s->position = ld3(L->position);
            emission_pdf_w = (1.0f / (4.0f * kPi)) * p_pick;
            direct_pdf_a = 1.0f * p_pick;
            radiance = ld3(L->radiance);
        } else {
            ov3 ex = ld3(L->x), ez = ld3(L->z);
            float a = orc_uniform(u), b = orc_uniform(u);
            s->position = add(add(sub(sub(ld3(L->position), scl(ex, 0.5f)), scl(ez, 0.5f)), scl(ex, a)), scl(ez, b));
            float r0 = orc_uniform(u), r1 = orc_uniform(u);
            float r = sqrtf(r0);
            float th = (kPi * 2.0f) * r1;
            ov3 local = v3(r * cos_f(th), sqrtf(maxf(0.0f, 1 - r0)), r * sin_f(th));
            float t[3], bb[3];
            ov3 n = ld3(L->direction);
            float nn[3] = {n.x, n.y, n.z};
            orc_make_frame(nn, t, bb);
            s->direction = normalize(add(add(scl(ld3(t), local.x), scl(n, local.y)), scl(ld3(bb), local.z)));
            float cos_l = local.y;
            float area = length(ex) * length(ez);
            float pdf_a = 1.0f / area;
            emission_pdf_w = pdf_a * (cos_l / kPi) * p_pick;
            direct_pdf_a = pdf_a * p_pick;
            radiance = scl(ld3(L->radiance), cos_l);
        }
    }
    if (!(emission_pdf_w > 0.0f)) return 0;
    s->throughput = scl(radiance, 1.0f / emission_pdf_w);
    s->dvcm = direct_pdf_a / emission_pdf_w;
    if (s->infinite) {
        /* the image-based light can be hit by camera paths: cosThetaLight = 1 */
        s->dvc = 1.0f / emission_pdf_w;
        s->dvm = s->dvc * vc_weight;
    } else {
        /* quad and point lights are not intersectable: no "hit the light" technique */
        s->dvc = 0.0f;
        s->dvm = 0.0f;
    }
    return 1;
}

/* Direct light sample for NEE at x (ConnectCameraPathToLight's light part). Uses 3 uniforms.
 * out: dir, radiance, distance, directPdfW (incl. pick), emissionPdfW (incl. pick), cosAtLight,
 * and whether the light is the image-based one (DirectIblLightSample, AreaLighting.cpp:293-314). */
static int direct_light(orc_uniforms* u, const orc_scene* sc, ov3 x, ov3* dir, ov3* radiance, float* dist,
                        float* direct_pdf_w, float* emission_pdf_w, float* cos_at_light, int* infinite)
{
    const uint32_t n_lights = sc->n_lights, n_pick = n_lights + (sc->ibl ? 1u : 0u);
    float p_pick = 1.0f / n_pick;
    uint32_t idx = (uint32_t)(orc_uniform(u) * n_pick);
    if (idx >= n_pick) idx = n_pick - 1;
    float a = orc_uniform(u), b = orc_uniform(u);
    *infinite = idx == n_lights;
    if (*infinite) {
        float o[8];
        orc_ibl_direct(sc->ibl, a, b, o);
        *dir = v3(o[0], o[1], o[2]);
        *radiance = v3(o[3], o[4], o[5]);
        *dist = 1e36f;
        *direct_pdf_w = o[6] * p_pick;
        *emission_pdf_w = (orc_ibl_position_pdf(sc->ibl) * o[6]) * p_pick;
        *cos_at_light = 1.0f;
        return o[6] > 0.0f;
    }
    const orc_light* L = &sc->lights[idx];
    if (L->type == 1) {
        ov3 d = sub(ld3(L->position), x);
        float d2 = dot(d, d);
        *dist = sqrtf(d2);
        *dir = scl(d, 1.0f / *dist);
        *radiance = ld3(L->radiance);
        *direct_pdf_w = d2 * p_pick;
        *emission_pdf_w = (1.0f / (4.0f * kPi)) * p_pick;
        *cos_at_light = 1.0f;
        return 1;
    }
    ov3 ex = ld3(L->x), ez = ld3(L->z);
    ov3 y = add(add(sub(sub(ld3(L->position), scl(ex, 0.5f)), scl(ez, 0.5f)), scl(ex, a)), scl(ez, b));
    ov3 d = sub(y, x);
    float d2 = dot(d, d);
    *dist = sqrtf(d2);
    *dir = scl(d, 1.0f / *dist);
    float cl = dot(ld3(L->direction), neg(*dir));
    if (!(cl > 0.0f)) return 0;
    float pdf_a = 1.0f / (length(ex) * length(ez));
    *radiance = ld3(L->radiance);
    *direct_pdf_w = pdf_a * d2 / cl * p_pick;
    *emission_pdf_w = pdf_a * (cl / kPi) * p_pick;
    *cos_at_light = cl;
    return 1;
}

/* ---- per-vertex surfaces ------------------------------------------------------------------------ */

typedef struct {
    ov3 pos, n, wfix; /* wfix: direction toward the previous vertex of the subpath */
    ov3 throughput;
    ov3 base;         /* base colour after CalculateSurfaceParams (texel or material, ^2.2) */
    float dvcm, dvc, dvm;
    uint32_t path_length;
    uint32_t geom;
    float error;
    int has_tan;   /* the frame came from the geometry's tangents (t, w) */
    float tan[4];
} vcm_vertex;

typedef struct {
    const orc_scene* sc;
    const orc_vcm_job* job;
    orc_camera cam;
    float vm_weight, vc_weight, vm_norm, radius;
    uint32_t* stack;
} vcm_ctx;

/* The surface of the closest hit; *has_tan (may be NULL) tells whether its frame came from the
 * geometry's tangents, then stored in tan4. */
static int hit_surface(const vcm_ctx* c, const float* o, const float* d, float tnear, orc_surface* s, float* pos,
                       float* err, uint32_t* geom_out, int* has_tan, float* tan4)
{
    float t, hu, hv, ng[3];
    uint32_t geom, tri_in_geom;
    const float* l2w = NULL;
    if (!orc_scene_closest1(c->sc, c->stack, o, d, tnear, 3.402823466e+38F, &t, &hu, &hv, &geom, ng, &tri_in_geom,
                            &l2w))
        return 0;
    const float* mat = c->sc->materials + 22 * (size_t)c->sc->geom_material[geom];
    float texel[3];
    const int tf = orc_surface_from_hit(o, d, t, ng, c->sc->geom_attr ? &c->sc->geom_attr[geom] : NULL, tri_in_geom,
                                        hu, hv, l2w, mat,
                                        orc_hit_texel_local(c->sc, tri_in_geom, geom, mat, hu, hv, texel), s, pos, err,
                                        tan4);
    if (has_tan) *has_tan = tf;
    *geom_out = geom;
    return 1;
}

static void surface_of_vertex(const vcm_ctx* c, const vcm_vertex* v, orc_surface* s)
{
    /* the stored vertex keeps what CalculateSurfaceParams derives: position, normal, material */
    const float* mat = c->sc->materials + 22 * (size_t)c->sc->geom_material[v->geom];
    float block[22];
    memcpy(block, mat, sizeof(block));
    block[0] = v->base.x;
    block[1] = v->base.y;
    block[2] = v->base.z;
    block[10] = saturate(mat[10]);
    block[11] = saturate(mat[11]);
    block[12] = mat[12] * 0.5f;
    uint32_t flags = (uint32_t)mat[21];
    block[15] = ((flags & 1u) && dot(v->wfix, v->n) < 0.0f) ? mat[19] : 1.0f / mat[19];
    float nn[3] = {v->n.x, v->n.y, v->n.z};
    orc_surface_from_block(block, nn, s);
    if (v->has_tan) orc_tangent_frame(v->tan, s);
}

static void eval_bsdf(const orc_surface* s, ov3 v, ov3 l, ov3* f, float* fwd, float* rev)
{
    float vv[3] = {v.x, v.y, v.z}, ll[3] = {l.x, l.y, l.z}, out[5];
    if (s->shader == 2)
        orc_lambert_eval(s, vv, ll, out);
    else
        orc_disney_eval(s, vv, ll, s->shader == 0, out);
    *f = ld3(out);
    *fwd = out[3];
    *rev = out[4];
}

static float cont_prob(const orc_surface* s)
{
    return saturate(maxf(maxf(s->base_color.x, s->base_color.y), s->base_color.z));
}

/* SampleBsdfScattering (VCM.cpp:412-446) with the RR-compensated throughput. Returns 0 to stop. */
static int scatter(orc_uniforms* u, const vcm_ctx* c, const orc_surface* s, ov3 pos, ov3* dir, ov3* thr, float* dvcm,
                   float* dvc, float* dvm, uint32_t* path_length)
{
    float out[14];
    memset(out, 0, sizeof(out));
    float v[3] = {-dir->x, -dir->y, -dir->z};
    int ok = s->shader == 2 ? orc_lambert_sample(u, s, v, out) : orc_disney_sample(u, s, v, s->shader == 0, out);
    if (!ok) return 0;
    if (!(out[8] > 0.0f)) return 0; /* a sample of pdf 0 ends the path */
    ov3 refl = ld3(out + 2);
    if (is_zero(refl)) return 0;
    float cp = cont_prob(s);
    if (orc_uniform(u) > cp) return 0;
    float fwd = out[8] * cp, rev = out[9] * cp;
    ov3 wi = ld3(out + 5);
    float cos_b = absf(dot(wi, s->n));
    (void)pos;
    *thr = scl(mul(*thr, refl), 1.0f / cp);
    float k = cos_b / fwd;
    float nvc = k * (*dvc * rev + *dvcm + c->vm_weight);
    float nvm = k * (*dvm * rev + *dvcm * c->vc_weight + 1.0f);
    *dvc = nvc;
    *dvm = nvm;
    *dvcm = 1.0f / fwd;
    *dir = wi;
    ++*path_length;
    return 1;
}

/* ---- the integrator ------------------------------------------------------------------------------ */

static uint64_t stream_id(uint32_t iteration, uint32_t path, uint32_t pass)
{
    return ((uint64_t)iteration << 33) | ((uint64_t)path << 1) | pass;
}

/* Per-iteration work shared by the threads. Each light path writes its vertices and camera splats
 * into slots of its own; they are then gathered in path order, so the lists (and the sums) are
 * those of the sequential loop of EvaluateLightPaths whatever the thread count. */
typedef struct {
    vcm_ctx c;
    float vipd;
    uint32_t W, H, N, L, mlv, iteration;
    vcm_vertex* slot_verts; /* N x mlv */
    uint32_t* slot_nverts;
    ov3* splat_c;           /* N x mlv */
    uint32_t* splat_pix;
    uint32_t* slot_nsplat;
    uint32_t* path_rays;    /* rays traced by path i in the current pass */
    const vcm_vertex* verts;
    const uint32_t* path_ends;
    const float* vpos;
    uint32_t nv;
    const orc_hash_grid* grid;
    float* iter_fb;
} vcm_iter;

static void light_paths(void* ctx, uint32_t b, uint32_t e)
{
    vcm_iter* I = (vcm_iter*)ctx;
    vcm_ctx c = I->c;
    const orc_scene* sc = c.sc;
    const uint32_t W = I->W, H = I->H, N = I->N, L = I->L;
    c.stack = (uint32_t*)malloc(sizeof(uint32_t) * orc_scene_stack_size(sc));
    for (uint32_t i = b; i < e; ++i) {
        uint32_t nvert = 0, nsplat = 0, nrays = 0;
        vcm_vertex* mine = I->slot_verts + (size_t)i * I->mlv;
        uint64_t st, inc, sid = stream_id(I->iteration, i, 0);
        orc_pcg32_seed(c.job->seed ^ (sid * 0x9E3779B97F4A7C15ull), sid, &st, &inc);
        orc_uniforms u = {NULL, 0, 0, &st, inc};
        emit_state es;
        if (sc->n_lights + (sc->ibl ? 1u : 0u) > 0 && emit_light(&u, sc, c.vc_weight, &es)) {
            ov3 pos = es.position, dir = es.direction, thr = es.throughput;
            float dvcm = es.dvcm, dvc = es.dvc, dvm = es.dvm;
            uint32_t plen = 1;
            while (plen + 2 < L) {
                float o[3] = {pos.x, pos.y, pos.z}, d[3] = {dir.x, dir.y, dir.z};
                orc_surface s;
                float hp[3], err;
                uint32_t geom;
                ++nrays;
                int has_tan = 0;
                float tan4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                if (!hit_surface(&c, o, d, 0.00001f, &s, hp, &err, &geom, &has_tan, tan4)) break;
                ov3 x = ld3(hp);
                ov3 view = neg(dir);
                float d2 = dot(sub(pos, x), sub(pos, x));
                float cos_v = absf(dot(s.n, view));
                /* the first segment is converted to area measure too, except from a light at
                 * infinity */
                if (plen > 1 || !es.infinite) dvcm *= d2;
                dvcm *= (1.0f / cos_v);
                dvc *= (1.0f / cos_v);
                dvm *= (1.0f / cos_v);
                vcm_vertex* vx = &mine[nvert++];
                vx->pos = x;
                vx->n = s.n;
                vx->wfix = view;
                vx->throughput = thr;
                vx->base = s.base_color;
                vx->dvcm = dvcm;
                vx->dvc = dvc;
                vx->dvm = dvm;
                vx->path_length = plen;
                vx->geom = geom;
                vx->error = err;
                vx->has_tan = has_tan;
                memcpy(vx->tan, tan4, sizeof(tan4));
                /* ConnectLightPathToCamera (VCM.cpp:190-242) */
                ov3 to_pos = sub(x, c.cam.pos);
                float px, py;
                if (dot(c.cam.forward, to_pos) > 0.0f && world_to_image(&c.cam, x, &px, &py) && px >= 0.0f &&
                    py >= 0.0f && px < (float)W && py < (float)H) {
                    float dist = length(to_pos);
                    ov3 to_cam = scl(to_pos, -1.0f / dist);
                    float cos_cam = dot(c.cam.forward, neg(to_cam));
                    ov3 f = v3(0, 0, 0);
                    float fwd, rev;
                    /* a connection in the surface's plane carries nothing (its BSDF value holds the
                     * cosine) and the BSDF is not evaluated there (0/0 in Disney's terms) */
                    if (dot(s.n, to_cam) != 0.0f) eval_bsdf(&s, view, to_cam, &f, &fwd, &rev);
                    if (!is_zero(f)) {
                        float cp = cont_prob(&s);
                        rev *= cp;
                        float img_dist = I->vipd / cos_cam;
                        float img_to_solid = img_dist * img_dist / cos_cam;
                        float cos_to_cam = absf(dot(s.n, to_cam));
                        float img_to_surface = img_to_solid * cos_to_cam / (dist * dist);
                        float w_light = (img_to_surface / (float)N) * (c.vm_weight + dvcm + dvc * rev);
                        float mis = 1.0f / (w_light + 1.0f);
                        /* f carries |cos|: contribution = mis thr f img_to_solid / (N d^2) */
                        ov3 contrib = scl(mul(thr, f), mis * img_to_solid / ((float)N * dist * dist));
                        int ix = (int)floorf(px), iy = (int)floorf(py);
                        if (!is_zero(contrib) && ix >= 0 && iy >= 0 && ix < (int)W && iy < (int)H) {
                            float so[3];
                            float tc[3] = {to_cam.x, to_cam.y, to_cam.z};
                            orc_offset_origin(&s, hp, err, tc, 0.1f, so, NULL);
                            ++nrays;
                            if (!orc_scene_occluded1(sc, c.stack, so, tc, err, dist)) {
                                I->splat_c[(size_t)i * I->mlv + nsplat] = contrib;
                                I->splat_pix[(size_t)i * I->mlv + nsplat] = (uint32_t)iy * W + (uint32_t)ix;
                                ++nsplat;
                            }
                        }
                    }
                }
                if (!scatter(&u, &c, &s, x, &dir, &thr, &dvcm, &dvc, &dvm, &plen)) break;
                pos = x;
            }
        }
        I->slot_nverts[i] = nvert;
        I->slot_nsplat[i] = nsplat;
        I->path_rays[i] = nrays;
    }
    free(c.stack);
}

/* MergeVertices (VCM.cpp:362-409) for one light vertex found by the range search. */
typedef struct {
    const vcm_ctx* c;
    const vcm_vertex* verts;
    const orc_surface* s;
    ov3 view;
    float dvcm, dvm;
    uint32_t plen;
    float ccp;
    ov3 acc;
} merge_ctx;

static void merge_vertex(void* ctx, uint32_t index)
{
    merge_ctx* m = (merge_ctx*)ctx;
    const vcm_vertex* lv = &m->verts[index];
    if (m->plen + lv->path_length > m->c->job->max_path_length) return;
    orc_surface ls;
    surface_of_vertex(m->c, lv, &ls);
    /* the reference's transparency rule (VCM.cpp:381-383) */
    if ((m->s->flags & 1u) != (ls.flags & 1u)) return;
    /* a photon arriving in the surface's plane is not evaluated (measure zero; 0/0 below) */
    if (dot(m->s->n, lv->wfix) == 0.0f) return;
    ov3 f;
    float fwd, rev;
    eval_bsdf(m->s, m->view, lv->wfix, &f, &fwd, &rev);
    if (is_zero(f)) return;
    float cos_l = absf(dot(m->s->n, lv->wfix));
    fwd *= m->ccp;
    rev *= cont_prob(&ls);
    float w_light = lv->dvcm * m->c->vc_weight + lv->dvm * fwd;
    float w_cam = m->dvcm * m->c->vc_weight + m->dvm * rev;
    float mis = 1.0f / (w_light + 1.0f + w_cam);
    /* density estimate uses f without the cosine Selas's BSDF value carries */
    m->acc = add(m->acc, scl(mul(f, lv->throughput), mis / cos_l));
}

static void camera_paths(void* ctx, uint32_t b, uint32_t e)
{
    vcm_iter* I = (vcm_iter*)ctx;
    vcm_ctx c = I->c;
    const orc_scene* sc = c.sc;
    const uint32_t W = I->W, N = I->N, L = I->L;
    c.stack = (uint32_t*)malloc(sizeof(uint32_t) * orc_scene_stack_size(sc));
    for (uint32_t i = b; i < e; ++i) {
        uint32_t nrays = 0;
        uint64_t st, inc, sid = stream_id(I->iteration, i, 1);
        orc_pcg32_seed(c.job->seed ^ (sid * 0x9E3779B97F4A7C15ull), sid, &st, &inc);
        orc_uniforms u = {NULL, 0, 0, &st, inc};
        uint32_t px = i % W, py = i / W;
        /* GenerateCameraSample (VCMCommon.cpp:66-88) */
        float jx = (float)px + orc_uniform(&u), jy = (float)py + orc_uniform(&u);
        float dd[3];
        orc_camera_dir(&c.cam, jx, jy, dd);
        ov3 dir = ld3(dd), pos = c.cam.pos, thr = v3(1, 1, 1), color = v3(0, 0, 0);
        float cos_cam = dot(c.cam.forward, dir);
        float img_dist = I->vipd / cos_cam;
        float cam_pdf_w = img_dist * img_dist / cos_cam;
        float dvcm = (float)N * (1.0f / cam_pdf_w), dvc = 0.0f, dvm = 0.0f;
        uint32_t plen = 1;
        while (plen < L) {
            float o[3] = {pos.x, pos.y, pos.z}, d[3] = {dir.x, dir.y, dir.z};
            orc_surface s;
            float hp[3], err;
            uint32_t geom;
            ++nrays;
            if (!hit_surface(&c, o, d, 0.00001f, &s, hp, &err, &geom, NULL, NULL)) {
                if (sc->ibl) {
                    /* ConnectToSkyLight (VCM.cpp:171-186): unweighted when seen from the camera */
                    float rad[3];
                    float pdf = orc_ibl_lookup(sc->ibl, d, rad);
                    float mis = 1.0f;
                    if (plen != 1) {
                        float p_pick = 1.0f / (sc->n_lights + 1);
                        float direct_pdf_a = pdf * p_pick;
                        float emission_pdf_w = (orc_ibl_position_pdf(sc->ibl) * pdf) * p_pick;
                        mis = 1.0f / (1.0f + (direct_pdf_a * dvcm + emission_pdf_w * dvc));
                    }
                    color = add(color, mul(thr, scl(ld3(rad), mis)));
                } else if (plen == 1) {
                    color = add(color, mul(thr, ld3(sc->background)));
                }
                break;
            }
            ov3 x = ld3(hp);
            ov3 view = neg(dir);
            float d2 = dot(sub(pos, x), sub(pos, x));
            float cos_v = absf(dot(s.n, view));
            dvcm *= d2;
            dvcm /= cos_v;
            dvc /= cos_v;
            dvm /= cos_v;
            /* NEE (ConnectCameraPathToLight, VCM.cpp:245-290) */
            if (plen + 1 < L && sc->n_lights + (sc->ibl ? 1u : 0u) > 0) {
                ov3 ldir, rad;
                float ldist, dpdf, epdf, cl;
                int infinite;
                if (direct_light(&u, sc, x, &ldir, &rad, &ldist, &dpdf, &epdf, &cl, &infinite)) {
                    ov3 f = v3(0, 0, 0);
                    float fwd, rev;
                    if (dot(s.n, ldir) != 0.0f) eval_bsdf(&s, view, ldir, &f, &fwd, &rev);
                    if (!is_zero(f)) {
                        float cp = cont_prob(&s);
                        rev *= cp;
                        float cos_s = absf(dot(s.n, ldir));
                        /* only the image-based light can also be reached by the camera path */
                        float w_light = infinite ? (fwd * cp) / dpdf : 0.0f;
                        float w_cam = (epdf * cos_s / (dpdf * cl)) * (c.vm_weight + dvcm + dvc * rev);
                        float mis = 1.0f / (w_light + 1.0f + w_cam);
                        ov3 contrib = scl(mul(f, rad), mis / dpdf);
                        if (!is_zero(contrib)) {
                            float so[3], ld[3] = {ldir.x, ldir.y, ldir.z};
                            orc_offset_origin(&s, hp, err, ld, 0.1f, so, NULL);
                            ++nrays;
                            if (!orc_scene_occluded1(sc, c.stack, so, ld, err, ldist))
                                color = add(color, mul(thr, contrib));
                        }
                    }
                }
            }
            /* VC to the vertices of light path i (vcCount = 1, VCM.cpp:619-636) */
            {
                uint32_t vb = i == 0 ? 0 : I->path_ends[i - 1], ve = I->path_ends[i];
                for (uint32_t k = vb; k < ve; ++k) {
                    const vcm_vertex* lv = &I->verts[k];
                    if (lv->path_length + 1 + plen > L) break;
                    ov3 dv = sub(lv->pos, x);
                    float dist2 = dot(dv, dv);
                    float dist = sqrtf(dist2);
                    ov3 cdir = scl(dv, 1.0f / dist);
                    /* a connection in either surface's plane carries nothing (see the splats) */
                    if (dot(cdir, s.n) == 0.0f || dot(neg(cdir), lv->n) == 0.0f) continue;
                    ov3 fc;
                    float cfwd, crev;
                    eval_bsdf(&s, view, cdir, &fc, &cfwd, &crev);
                    if (is_zero(fc)) continue;
                    orc_surface ls;
                    surface_of_vertex(&c, lv, &ls);
                    ov3 fl;
                    float lfwd, lrev;
                    eval_bsdf(&ls, lv->wfix, neg(cdir), &fl, &lfwd, &lrev);
                    if (is_zero(fl)) continue;
                    float cos_c = absf(dot(cdir, s.n));
                    float cos_l = absf(dot(neg(cdir), lv->n));
                    float ccp = cont_prob(&s), lcp = cont_prob(&ls);
                    cfwd *= ccp;
                    crev *= ccp;
                    lfwd *= lcp;
                    lrev *= lcp;
                    float c_pdf_a = cfwd * cos_l / dist2;
                    float l_pdf_a = lfwd * cos_c / dist2;
                    float w_light = c_pdf_a * (c.vm_weight + lv->dvcm + lv->dvc * lrev);
                    float w_cam = l_pdf_a * (c.vm_weight + dvcm + dvc * crev);
                    float mis = 1.0f / (w_light + 1.0f + w_cam);
                    ov3 contrib = scl(mul(fc, fl), mis / dist2);
                    if (is_zero(contrib)) continue;
                    float sb;
                    float so[3], cd[3] = {cdir.x, cdir.y, cdir.z};
                    orc_offset_origin(&s, hp, err, cd, 0.1f, so, &sb);
                    ++nrays;
                    if (!orc_scene_occluded1(sc, c.stack, so, cd, err, dist - 16.0f * absf(sb)))
                        color = add(color, mul(mul(thr, lv->throughput), contrib));
                }
            }
            /* VM (SearchHashGrid + MergeVertices, VCM.cpp:362-409, 638-655) */
            if (I->nv > 0) {
                merge_ctx m = {&c, I->verts, &s, view, dvcm, dvm, plen, cont_prob(&s), v3(0, 0, 0)};
                hash_grid_visit(I->grid, I->vpos, hp, merge_vertex, &m);
                color = add(color, scl(mul(thr, m.acc), c.vm_norm));
            }
            if (!scatter(&u, &c, &s, x, &dir, &thr, &dvcm, &dvc, &dvm, &plen)) break;
            pos = x;
        }
        float* p = I->iter_fb + 3 * (size_t)i;
        p[0] += color.x;
        p[1] += color.y;
        p[2] += color.z;
        I->path_rays[i] = nrays;
    }
    free(c.stack);
}

void orc_render_vcm(const orc_scene* sc, const orc_vcm_job* job, float* fb_out, float* stats_out, int threads)
{
    const uint32_t W = job->width, H = job->height, N = W * H, L = job->max_path_length;
    const uint32_t mlv = L > 2 ? L - 2 : 1;
    vcm_iter I;
    memset(&I, 0, sizeof(I));
    I.c.sc = sc;
    I.c.job = job;
    orc_camera_init(job->cam10, W, H, &I.c.cam);
    /* virtual image plane distance: pixels per unit length at distance 1 (see the file comment) */
    float hfov = job->cam10[9] * (kPi / 180.f);
    I.vipd = (float)W / (2.0f * tanf(hfov * 0.5f));
    I.W = W;
    I.H = H;
    I.N = N;
    I.L = L;
    I.mlv = mlv;
    I.slot_verts = (vcm_vertex*)malloc(sizeof(vcm_vertex) * (size_t)N * mlv);
    I.slot_nverts = (uint32_t*)malloc(sizeof(uint32_t) * N);
    I.splat_c = (ov3*)malloc(sizeof(ov3) * (size_t)N * mlv);
    I.splat_pix = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)N * mlv);
    I.slot_nsplat = (uint32_t*)malloc(sizeof(uint32_t) * N);
    I.path_rays = (uint32_t*)malloc(sizeof(uint32_t) * N);
    vcm_vertex* verts = (vcm_vertex*)malloc(sizeof(vcm_vertex) * (size_t)N * mlv);
    uint32_t* path_ends = (uint32_t*)malloc(sizeof(uint32_t) * N);
    float* vpos = (float*)malloc(sizeof(float) * 3 * (size_t)N * mlv);
    float* acc = (float*)calloc(3 * (size_t)N, sizeof(float));
    float* iter_fb = (float*)malloc(sizeof(float) * 3 * (size_t)N);
    double n_light_verts = 0, n_rays = 0;

    for (uint32_t it = 0; it < job->iterations; ++it) {
        float consts[6];
        orc_vcm_constants(N, 1, job->base_radius, job->radius_alpha, (float)(job->iteration_base + it + 1), consts);
        I.c.radius = consts[0];
        I.c.vm_norm = consts[2];
        I.c.vm_weight = consts[3];
        I.c.vc_weight = consts[4];
        I.iteration = job->iteration_base + it;
        memset(iter_fb, 0, sizeof(float) * 3 * (size_t)N);

        /* light subpaths (EvaluateLightPaths, VCM.cpp:502-561), then gathered in path order */
        orc_parallel_for(light_paths, &I, N, threads);
        uint32_t nv = 0;
        for (uint32_t i = 0; i < N; ++i) {
            for (uint32_t k = 0; k < I.slot_nverts[i]; ++k) {
                const vcm_vertex* vx = &I.slot_verts[(size_t)i * mlv + k];
                verts[nv] = *vx;
                vpos[3 * nv] = vx->pos.x;
                vpos[3 * nv + 1] = vx->pos.y;
                vpos[3 * nv + 2] = vx->pos.z;
                ++nv;
            }
            path_ends[i] = nv;
            for (uint32_t k = 0; k < I.slot_nsplat[i]; ++k) {
                float* p = iter_fb + 3 * (size_t)I.splat_pix[(size_t)i * mlv + k];
                ov3 cc = I.splat_c[(size_t)i * mlv + k];
                p[0] += cc.x;
                p[1] += cc.y;
                p[2] += cc.z;
            }
            n_rays += I.path_rays[i];
        }
        n_light_verts += nv;

        /* photon grid (BuildHashGrid, VCMHashGrid.cpp:48) */
        orc_hash_grid grid;
        orc_hash_grid_build(&grid, vpos, nv, N, I.c.radius);

        /* camera subpaths (EvaluateCameraPaths, VCM.cpp:576-666): path i adds to pixel i only */
        I.verts = verts;
        I.path_ends = path_ends;
        I.vpos = vpos;
        I.nv = nv;
        I.grid = &grid;
        I.iter_fb = iter_fb;
        orc_parallel_for(camera_paths, &I, N, threads);
        for (uint32_t i = 0; i < N; ++i) n_rays += I.path_rays[i];
        orc_hash_grid_free(&grid);
        for (size_t k = 0; k < 3 * (size_t)N; ++k) acc[k] += iter_fb[k];
    }
    float inv = 1.0f / (float)job->iterations;
    for (size_t k = 0; k < 3 * (size_t)N; ++k) fb_out[k] = acc[k] * inv;
    if (stats_out) {
        stats_out[0] = (float)(n_light_verts / job->iterations);
        stats_out[1] = (float)n_rays;
    }
    free(I.slot_verts);
    free(I.slot_nverts);
    free(I.splat_c);
    free(I.splat_pix);
    free(I.slot_nsplat);
    free(I.path_rays);
    free(verts);
    free(path_ends);
    free(vpos);
    free(acc);
    free(iter_fb);
}
