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

/* A stream of uniforms consumed in order (a restatement of CSampler::UniformFloat's role). */
typedef struct {
    const float* values;
    uint32_t count;
    uint32_t pos;
} orc_uniforms;

typedef struct {
    ov3 x, y, z, pos, forward;
    uint32_t width, height;
    float vipd; /* virtualImagePlaneDistance */
} orc_camera;

float orc_uniform(orc_uniforms* u);
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

#ifdef __cplusplus
}
#endif

#endif
