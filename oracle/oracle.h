/*
 * ORACLE -- test infrastructure only (see orc_trace.c). Declarations of the CPU restatements
 * that tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load from liboracle.so.
 */
#ifndef SGPU_ORACLE_H
#define SGPU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tnear, *tfar;
} orc_rays;

typedef struct {
    float* t;
    uint32_t* prim; /* global triangle index, 0xFFFFFFFF on a miss */
    float* u;
    float* v;
} orc_hits;

typedef struct orc_bvh orc_bvh;

int orc_tri_test(const float* v0, const float* e1, const float* e2, const float* o, const float* d, float tnear,
                 float* t_out, float* u_out, float* v_out);
void orc_closest_brute_f32(const float* tri9, uint32_t ntri, const orc_rays* rays, uint32_t n, const orc_hits* hits,
                           int threads);
void orc_occluded_brute_f32(const float* tri9, uint32_t ntri, const orc_rays* rays, uint32_t n, uint32_t* occluded,
                            int threads);
void orc_closest_brute_f64(const float* tri9, uint32_t ntri, const orc_rays* rays, uint32_t n, uint32_t* prim,
                           double* t1, double* t2, double* margin, int threads);
orc_bvh* orc_bvh_build(const float* tri9, uint32_t ntri);
void orc_bvh_free(orc_bvh* b);
void orc_bvh_closest(const orc_bvh* b, const orc_rays* rays, uint32_t n, const orc_hits* hits, int threads);
void orc_bvh_occluded(const orc_bvh* b, const orc_rays* rays, uint32_t n, uint32_t* occluded, int threads);
uint32_t orc_bvh_node_count(const orc_bvh* b);
/* single-ray forms for the integrator restatements; stack must hold orc_bvh_stack_size entries */
uint32_t orc_bvh_stack_size(const orc_bvh* b);
uint32_t orc_bvh_closest1(const orc_bvh* b, uint32_t* stack, const float* o, const float* d, float tnear, float tfar,
                          float* t, float* u, float* v);
int orc_bvh_occluded1(const orc_bvh* b, uint32_t* stack, const float* o, const float* d, float tnear, float tfar);
void orc_parallel_for(void (*fn)(void* ctx, uint32_t begin, uint32_t end), void* ctx, uint32_t n, int threads);
/* instanced scenes (composed instances: world-to-local rows of 4 floats and the placed scene's tree);
 * closest hits give prim (triangle index in the placed scene) and inst_out (composed instance) */
typedef struct {
    float w2l[12];
    const orc_bvh* blas;
} orc_inst;
void orc_inst_closest(const orc_inst* insts, uint32_t n_inst, const orc_rays* rays, uint32_t n, const orc_hits* hits,
                      uint32_t* inst_out, int threads);
void orc_inst_occluded(const orc_inst* insts, uint32_t n_inst, const orc_rays* rays, uint32_t n, uint32_t* occluded,
                       int threads);

/* ---- shading (orc_shade.c) ---------------------------------------------------------------------- */
typedef struct {
    float x, y, z;
} ov3;

typedef struct {
    ov3 n, t, b; /* shading normal and tangent frame (tangent space is y-up: n = local y) */
    ov3 base_color, transmittance_color;
    float sheen, sheen_tint, clearcoat, clearcoat_gloss, metallic, spec_trans, diff_trans, flatness;
    float anisotropic, relative_ior, specular_tint, roughness, scatter_distance, ior;
    int shader; /* 0 Disney thin, 1 Disney solid, 2 Lambert */
    uint32_t flags;
} orc_surface;

/* A stream of uniforms consumed in order (the role of CSampler::UniformFloat): an explicit array,
 * or a PCG32 generator when pcg_state is set (O'Neill 2014, pcg32_random_r; uniform = top 24 bits
 * x 2^-24). */
typedef struct {
    const float* values;
    uint32_t count;
    uint32_t pos;
    uint64_t* pcg_state;
    uint64_t pcg_inc;
} orc_uniforms;

typedef struct {
    ov3 x, y, z, pos, forward;
    uint32_t width, height;
    float vipd; /* virtualImagePlaneDistance */
} orc_camera;

/* Transcendentals: 0 = double evaluated, rounded to float (as the device code); 1 = the
 * single-precision libm functions (as Selas's code). See orc_fn.h. */
void orc_set_libm(int single);
float orc_uniform(orc_uniforms* u);
uint32_t orc_pcg32_next(uint64_t* state, uint64_t inc);
void orc_pcg32_seed(uint64_t initstate, uint64_t initseq, uint64_t* state, uint64_t* inc);
void orc_make_frame(const float* n, float* t_out, float* b_out);
void orc_surface_from_block(const float* block22, const float* n, orc_surface* s);
float orc_schlick_weight(float u);
float orc_schlick_r0(float eta);
float orc_schlick(float r0, float c);
float orc_fresnel_dielectric(float cos_i, float ni, float nt);
void orc_ggx(const float* wo, const float* wi, const float* wm, float ax, float ay, float u1, float u2, float* out8);
void orc_disney_eval(const orc_surface* s, const float* v, const float* l, int thin, float* out5);
int orc_disney_sample(orc_uniforms* u, const orc_surface* s, const float* v, int thin, float* out14);
void orc_lambert_eval(const orc_surface* s, const float* v, const float* l, float* out5);
int orc_lambert_sample(orc_uniforms* u, const orc_surface* s, const float* v, float* out14);
void orc_cmj(int32_t s, int32_t m, int32_t n, int32_t p, float* out2);
void orc_camera_init(const float* cam10, uint32_t width, uint32_t height, orc_camera* c);
void orc_camera_dir(const orc_camera* c, float x, float y, float* out3);
void orc_camera_cmj_dir(const orc_camera* c, int32_t x, int32_t y, int32_t s, int32_t m, int32_t n, int32_t p,
                        float* out3);
void orc_rect_light(const float* q12, float u, float v, float* out5);
int orc_ray_quad(const float* q18);
void orc_concentric_disc(float r0, float r1, float* out2);
/* Batch forms over surface blocks of 22 floats (layout in orc_surface_from_block; block[20] is the
 * shader: 0 Disney thin, 1 Disney solid, 2 Lambert). out = 5 / 14 floats per entry. */
void orc_bsdf_eval_batch(const float* blocks, const float* n, const float* v, const float* l, uint32_t count,
                         float* out);
void orc_bsdf_sample_batch(const float* blocks, const float* n, const float* v, const float* uniforms,
                           uint32_t n_uniforms, uint32_t count, float* out);

/* ---- integrators (orc_pt.c, orc_vcm.c) ------------------------------------------------------- */
typedef struct {
    uint32_t type; /* 0 quad, 1 point; same layout as sgpu_light */
    float position[3], direction[3], x[3], z[3], radiance[3];
} orc_light;

/* ---- per-face textures (orc_ptex.c) ------------------------------------------------------------- */
typedef struct {
    int res_log2;           /* every face has (1 << res_log2)^2 texels */
    uint32_t nfaces;
    const uint8_t* texels;  /* nfaces x res x res x 3 (row = v texel) */
    const int32_t* adjfaces; /* 4 per face: bottom, right, top, left; -1 = border */
    const int32_t* adjedges; /* 4 per face: the neighbour's edge id of the shared edge */
    int tri;                 /* 1: a triangle mesh (3 edges per face: v = 0, w = 0, u = 0) */
} orc_ptex;

/* ---- image-based light (orc_ibl.c) -------------------------------------------------------------- */
typedef struct {
    uint32_t w, h;
    const float* rgb;  /* h x w x 3, row 0 at +y */
    const float* mdf;  /* h: marginal CDF */
    const float* cdf;  /* h x w: conditional CDFs */
    float rotation;    /* radians in [0, 2 pi) (orc_ibl_rotation) */
    float exposure;
    float center[3], radius; /* sphere around the triangles' bounding box (orc_bounding_sphere) */
} orc_ibl;

void orc_ibl_tables(uint32_t w, uint32_t h, const float* rgb, float* mdf, float* cdf);
float orc_ibl_rotation(float radians);
void orc_bounding_sphere(const float* tri9, uint32_t n_tris, float* out4);
uint32_t orc_ibl_sample_distribution(const float* d, uint32_t count, float r, float* pdf);
/* out6: theta, phi, x, y, pdf, 0 */
void orc_ibl_sample(const orc_ibl* b, float r0, float r1, float* out6);
/* radiance (rgb_out) from unit direction wi; returns the pdf of sampling wi */
float orc_ibl_lookup(const orc_ibl* b, const float* wi, float* rgb_out);
/* out14: position 3, direction 3, radiance 3, emission pdf (area x solid angle), direction pdf, 0 x 3 */
void orc_ibl_emit(const orc_ibl* b, float r0, float r1, float r2, float r3, float* out14);
/* out8: direction 3, radiance 3, pdf, 0 */
void orc_ibl_direct(const orc_ibl* b, float r0, float r1, float* out8);
float orc_ibl_position_pdf(const orc_ibl* b);

/* An instanced scene for the integrators (orc_pt.c): the composed instances with, per composed
 * instance, the placed scene's triangles (its own space), their geometry ids and each geometry's
 * first triangle, the instance's first geometry in the scene's list of geometries, and the
 * linear part of its local-to-world transform (row-major 3x3) that carries hit normals to world. */
typedef struct {
    const orc_inst* insts;
    uint32_t n_inst;
    const float* const* tri9;
    const uint32_t* const* tri_geom;
    const uint32_t* const* geom_first;
    const uint32_t* geom_base;
    const float* l2w;
    /* round B-spline curves (orc_curve.c) per composed instance, NULL for none: the placed scene's
     * segments in its own space (16 floats each), their boxes, geometries and segment indices; the
     * composed instances n_inst .. n_comp - 1 place scenes of curves only (no triangle tree) */
    uint32_t n_comp;
    const float* const* curve_cp;
    const float* const* curve_box;
    const uint32_t* const* curve_geom;
    const uint32_t* const* curve_prim;
    const uint32_t* n_curves;
} orc_iscene;

/* Vertex attributes of one geometry, as the device keeps them: per primitive its corners' values
 * in the primitive's own vertex order (3 corners per triangle, 4 per quad), normals 3, tangents 4
 * (xyz, w), UVs 2 floats per corner; NULL where the geometry has none. */
typedef struct {
    const float* n;
    const float* t;
    const float* uv;
    uint32_t corners;
} orc_attr;

/* rtcInterpolate0 of a K-float attribute at primitive prim, (u, v) (a quad's own coordinates for
 * quads), as Embree 3 evaluates it (see orc_pt.c). */
void orc_interp_attr(const float* a, int K, uint32_t corners, uint32_t prim, float u, float v, float* out);

typedef struct {
    const orc_bvh* bvh;
    const float* tri9;          /* world-space triangles, 9 floats each (global index order) */
    const uint32_t* tri_geom;   /* geometry of each triangle */
    const float* materials;     /* 22 floats per material (raw MaterialResourceData values) */
    const uint32_t* geom_material;
    uint32_t n_geoms;
    const orc_light* lights;
    uint32_t n_lights;
    float background[3];
    const uint32_t* geom_first; /* first triangle of each geometry; NULL when no textures */
    const orc_ptex* geom_tex;   /* per geometry (nfaces == 0: none); a textured geometry is a quad
                                   mesh whose triangles 2q and 2q+1 are Embree's split of quad q */
    const orc_ibl* ibl;         /* image-based light, NULL for none (replaces the background) */
    const orc_iscene* inst;     /* instanced scene, NULL for none: then bvh, tri9, tri_geom and
                                   geom_first are unused and the shading tables above are indexed
                                   by the scene's list of geometries */
    const orc_attr* geom_attr;  /* vertex attributes per geometry (scene list), NULL for none */
    /* round B-spline curves of a scene of geometries (world-space segments, 16 floats each), their
     * boxes, geometries and segment indices within them; n_curves = 0 for none */
    const float* curve_cp;
    const float* curve_box;
    const uint32_t* curve_geom;
    const uint32_t* curve_prim;
    uint32_t n_curves;
} orc_scene;

/* Closest hit in a scene of geometries or an instanced scene: 0 on a miss; else t, u, v, the
 * hit's geometry (scene list), its unnormalised world-space geometric normal and the index of the
 * triangle within its geometry. */
int orc_scene_closest1(const orc_scene* sc, uint32_t* stack, const float* o, const float* d, float tnear, float tfar,
                       float* t, float* u, float* v, uint32_t* geom, float* ng, uint32_t* tri_in_geom,
                       const float** l2w);
int orc_scene_occluded1(const orc_scene* sc, uint32_t* stack, const float* o, const float* d, float tnear, float tfar);
uint32_t orc_scene_stack_size(const orc_scene* sc);
/* the texel of a hit on triangle tri_in_geom of geometry geom (NULL: the material's colour) */
const float* orc_hit_texel_local(const orc_scene* sc, uint32_t tri_in_geom, uint32_t geom, const float* mat, float hu,
                                 float hv, float* texel3);

typedef struct {
    float cam10[10]; /* position, look_at, up, fov_degrees */
    uint32_t width, height, spp, max_path_length;
    uint64_t seed;
    uint32_t flags; /* 1: quad-light estimator without the extra Dot(normal, l) */
} orc_pt_job;

/* CalculateSurfaceParams at a hit: position and error bound; the shading frame from the
 * geometry's vertex attributes ga (NULL: none) interpolated on triangle tri_in_geom of the geometry
 * (a quad geometry's triangles 2q and 2q + 1 are quad q) at (hu, hv) and carried to world space by
 * l2w (row-major 3x3, NULL: identity), else from the geometric normal ng (world space); the
 * material with `texel` (or NULL) as base colour. Returns 1 when the frame came from tangents
 * (then tan4, if not NULL, receives (t, w)). */
int orc_surface_from_hit(const float* o, const float* d, float t, const float* ng, const orc_attr* ga,
                         uint32_t tri_in_geom, float hu, float hv, const float* l2w, const float* mat,
                         const float* texel, orc_surface* s, float* position, float* error, float* tan4);
/* The frame of a stored vertex with shading normal s->n and stored tangent (t, w). */
void orc_tangent_frame(const float* tan4, orc_surface* s);
/* The base-colour texel of a hit on triangle `tri` (barycentrics hu, hv) when the material uses a
 * per-face texture and its geometry has one; returns NULL otherwise. */
const float* orc_hit_texel(const orc_scene* sc, uint32_t tri, uint32_t geom, const float* mat, float hu, float hv,
                           float* texel3);
void orc_offset_origin(const orc_surface* s, const float* pos, float error, const float* dir, float bias, float* out,
                       float* signed_bias);
void orc_nee(orc_uniforms* u, const orc_light* lights, uint32_t n_lights, const float* position, const float* normal,
             int cosine_once, float* out8);
/* fb: 2 x height x width x 3 floats (layer 0 first-bounce reflectance, layer 1 radiance) */
void orc_render_pt(const orc_scene* sc, const orc_pt_job* job, float* fb, int threads);

typedef struct {
    float cam10[10]; /* position, look_at, up, fov_degrees */
    uint32_t width, height, spp_x, spp_y, max_path_length;
    uint64_t seed;
} orc_dpt_job;
/* the deferred path tracer's estimator; fb: height x width x 3 floats */
void orc_render_dpt(const orc_scene* sc, const orc_dpt_job* job, float* fb, int threads);

/* VCM (orc_vcm.c) */
typedef struct {
    float lo[3], hi[3];
    float radius, radius_sq, inv_cell;
    uint32_t cell_count, count;
    uint32_t* cell_ends; /* exclusive end of each cell's range */
    uint32_t* indices;   /* point index per sorted slot */
} orc_hash_grid;

typedef struct {
    float cam10[10];
    uint32_t width, height;
    uint32_t iterations;     /* VCM iterations (one light path and one camera path per pixel each) */
    uint32_t iteration_base; /* global index of the first iteration (radius schedule, streams) */
    uint32_t max_path_length;
    float base_radius, radius_alpha;
    uint64_t seed;
} orc_vcm_job;

void orc_vcm_constants(uint32_t vm_count, uint32_t vc_count, float base_radius, float alpha, float iteration,
                       float* out6);
void orc_hash_grid_build(orc_hash_grid* g, const float* points, uint32_t count, uint32_t cell_count, float radius);
void orc_hash_grid_free(orc_hash_grid* g);
uint32_t orc_hash_grid_query(const orc_hash_grid* g, const float* points, const float* q, uint32_t* found,
                             uint32_t max_found);
/* fb_out: height x width x 3 floats (radiance averaged over iterations); stats_out (2 floats, may be
 * NULL): light vertices per iteration, rays traced. Paths of one pass run on `threads` threads; the
 * result does not depend on the thread count. */
void orc_render_vcm(const orc_scene* sc, const orc_vcm_job* job, float* fb_out, float* stats_out, int threads);

void orc_ptex_eval(const orc_ptex* tx, int face, float u, float v, float* out3);
void orc_ptex_tri_eval(const orc_ptex* tx, int face, float u, float v, float* out3);
void orc_ptex_eval_batch(const orc_ptex* tx, uint32_t n, const int32_t* faces, const float* u, const float* v,
                         float* out);

/* Round B-spline curves (orc_curve.c): segments of 16 floats (4 control points x, y, z, radius) */
int orc_curve_hit_dense(const float* cp, const float* o, const float* d, float tnear, float tfar, double* t,
                        double* u, double* ng);
int orc_curve_hit(const float* cp, const float* o, const float* d, float tnear, float tfar, int any, double* t,
                  double* u, double* ng);
double orc_curve_clearance(const float* cp, const float* o, const float* d);
void orc_curve_boxes(const float* cps, uint32_t nseg, float* boxes);
int orc_curves_closest1(const float* cps, const float* boxes, uint32_t nseg, const float* o, const float* d,
                        float tnear, float tfar, int dense, float* t, float* u, uint32_t* prim, float* ng);
int orc_curves_occluded1(const float* cps, const float* boxes, uint32_t nseg, const float* o, const float* d,
                         float tnear, float tfar);
void orc_curves_closest(const float* cps, uint32_t nseg, const orc_rays* rays, uint32_t n, float* t, float* u,
                        uint32_t* prim, float* ng, int dense, int threads);
void orc_curves_occluded(const float* cps, uint32_t nseg, const orc_rays* rays, uint32_t n, uint32_t* occ,
                         int threads);

#ifdef __cplusplus
}
#endif

#endif
