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

#ifdef __cplusplus
}
#endif

#endif
