/*
 * ORACLE -- test infrastructure only. Nothing in the product (libselas_gpu) links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do, as the
 * checker. Compiled with -ffp-contract=off (see oracle/Makefile).
 *
 * CPU restatement of the ray-cast boundary Selas calls through Embree 3.2:
 *   rtcIntersect1 (RayPick, Source/Applications/Selas/Source/PathTracer.cpp:95-139) and
 *   rtcOccluded1 (OcclusionRay, Source/Applications/Selas/Source/PathTracer.cpp:71-92).
 * Embree's library is not available (only its headers ship with the reference), so its exact
 * arithmetic cannot be run. What is restated is the contract Selas relies on -- closest hit in
 * (tnear, tfar], miss = 0xFFFFFFFF ids, u/v = weights of vertices 1 and 2, Ng = (v1-v0)x(v2-v0),
 * occluded iff any hit in (tnear, tfar] -- with a triangle test whose FP32 arithmetic is fixed
 * operation by operation (orc_tri_test), identical to the device kernel's.
 *
 * Three independent checkers:
 *   orc_closest_brute_f64   double-precision Moller-Trumbore over every triangle (pins the
 *                           arithmetic itself: a wrong sign or operand changes which triangle wins);
 *   orc_closest_brute_f32   orc_tri_test over every triangle (bit-exact reference for small inputs);
 *   orc_bvh_*               orc_tri_test under this file's own top-down SAH tree (bit-exact
 *                           reference at full size; its tree shares nothing with the device LBVH).
 * Parity status: pinned by brute force (f64 and f32) and by invariants; Embree bit-exactness is
 * unpinned (no Embree binary exists for this platform).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- the FP32 triangle test ------------------------------------------------------------------ */

static inline void cross_fma(const float* a, const float* b, float* r)
{
    r[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    r[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    r[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}

static inline float dot_fma(const float* a, const float* b) { return fmaf(a[0], b[0], fmaf(a[1], b[1], a[2] * b[2])); }

int orc_tri_test(const float* v0, const float* e1, const float* e2, const float* o, const float* d, float tnear,
                 float* t_out, float* u_out, float* v_out)
{
    float pvec[3], qvec[3], tvec[3];
    cross_fma(d, e2, pvec);
    float det = dot_fma(e1, pvec);
    tvec[0] = o[0] - v0[0];
    tvec[1] = o[1] - v0[1];
    tvec[2] = o[2] - v0[2];
    float U = dot_fma(tvec, pvec);
    cross_fma(tvec, e1, qvec);
    float V = dot_fma(d, qvec);
    float T = dot_fma(e2, qvec);
    if (det == 0.0f) return 0;
    if (det < 0.0f) {
        U = -U;
        V = -V;
        T = -T;
    }
    float a = fabsf(det);
    if (!(U >= 0.0f && V >= 0.0f && U + V <= a)) return 0;
    float t = T / a;
    if (!(t > tnear)) return 0;
    *t_out = t;
    *u_out = U / a;
    *v_out = V / a;
    return 1;
}

/* triangle soup -> (v0, e1, e2) records in input order */
typedef struct {
    float v0[3], e1[3], e2[3];
} tri_rec;

static tri_rec* make_recs(const float* tri9, uint32_t n)
{
    tri_rec* r = (tri_rec*)malloc(sizeof(tri_rec) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = tri9 + (size_t)i * 9;
        for (int k = 0; k < 3; ++k) {
            r[i].v0[k] = p[k];
            r[i].e1[k] = p[3 + k] - p[k];
            r[i].e2[k] = p[6 + k] - p[k];
        }
    }
    return r;
}

/* ---- parallel-for over ray ranges ---------------------------------------------------------------- */

typedef void (*range_fn)(void* ctx, uint32_t begin, uint32_t end);
typedef struct {
    range_fn fn;
    void* ctx;
    uint32_t n;
    uint32_t chunk;
    volatile uint32_t next;
} pf_job;

static void* pf_worker(void* arg)
{
    pf_job* job = (pf_job*)arg;
    for (;;) {
        uint32_t b = __sync_fetch_and_add(&job->next, job->chunk);
        if (b >= job->n) break;
        uint32_t e = b + job->chunk < job->n ? b + job->chunk : job->n;
        job->fn(job->ctx, b, e);
    }
    return NULL;
}

void orc_parallel_for(range_fn fn, void* ctx, uint32_t n, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pf_job job = {fn, ctx, n, 256, 0};
    if (threads == 1) {
        pf_worker(&job);
        return;
    }
    pthread_t th[256];
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, pf_worker, &job);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
}

/* ---- brute force ------------------------------------------------------------------------------- */

typedef struct {
    const tri_rec* recs;
    uint32_t ntri;
    orc_rays rays;
    orc_hits hits;
    uint32_t* occluded;
} brute_ctx;

static void brute_closest_range(void* c, uint32_t b, uint32_t e)
{
    brute_ctx* x = (brute_ctx*)c;
    for (uint32_t i = b; i < e; ++i) {
        float o[3] = {x->rays.ox[i], x->rays.oy[i], x->rays.oz[i]};
        float d[3] = {x->rays.dx[i], x->rays.dy[i], x->rays.dz[i]};
        float tnear = x->rays.tnear[i];
        float best_t = x->rays.tfar[i], bu = 0.0f, bv = 0.0f;
        uint32_t best = 0xFFFFFFFFu;
        for (uint32_t k = 0; k < x->ntri; ++k) {
            float t, u, v;
            if (orc_tri_test(x->recs[k].v0, x->recs[k].e1, x->recs[k].e2, o, d, tnear, &t, &u, &v)) {
                if (t < best_t || (t == best_t && k < best)) {
                    best_t = t;
                    best = k;
                    bu = u;
                    bv = v;
                }
            }
        }
        x->hits.t[i] = best_t;
        x->hits.prim[i] = best;
        x->hits.u[i] = best == 0xFFFFFFFFu ? 0.0f : bu;
        x->hits.v[i] = best == 0xFFFFFFFFu ? 0.0f : bv;
    }
}

void orc_closest_brute_f32(const float* tri9, uint32_t ntri, const orc_rays* rays, uint32_t n, const orc_hits* hits,
                           int threads)
{
    tri_rec* recs = make_recs(tri9, ntri);
    brute_ctx c = {recs, ntri, *rays, *hits, NULL};
    orc_parallel_for(brute_closest_range, &c, n, threads);
    free(recs);
}

static void brute_occ_range(void* c, uint32_t b, uint32_t e)
{
    brute_ctx* x = (brute_ctx*)c;
    for (uint32_t i = b; i < e; ++i) {
        float o[3] = {x->rays.ox[i], x->rays.oy[i], x->rays.oz[i]};
        float d[3] = {x->rays.dx[i], x->rays.dy[i], x->rays.dz[i]};
        float tnear = x->rays.tnear[i], tfar = x->rays.tfar[i];
        uint32_t occ = 0;
        for (uint32_t k = 0; k < x->ntri && !occ; ++k) {
            float t, u, v;
            if (orc_tri_test(x->recs[k].v0, x->recs[k].e1, x->recs[k].e2, o, d, tnear, &t, &u, &v) && t <= tfar)
                occ = 1;
        }
        x->occluded[i] = occ;
    }
}

void orc_occluded_brute_f32(const float* tri9, uint32_t ntri, const orc_rays* rays, uint32_t n, uint32_t* occluded,
                            int threads)
{
    tri_rec* recs = make_recs(tri9, ntri);
    brute_ctx c = {recs, ntri, *rays, {0}, occluded};
    orc_parallel_for(brute_occ_range, &c, n, threads);
    free(recs);
}

/* Double-precision Moller-Trumbore (Moller & Trumbore 1997) on the same float vertices, no
 * epsilon. For every ray: the closest triangle, its t, the second-closest t (any other triangle)
 * and the hit's barycentric margin min(u, v, 1-u-v), so a test can tell genuine ties and edge
 * grazes from errors. */
typedef struct {
    const float* tri9;
    uint32_t ntri;
    orc_rays rays;
    double* t1;
    double* t2;
    double* margin;
    uint32_t* prim;
} f64_ctx;

static void brute_f64_range(void* c, uint32_t b, uint32_t e)
{
    f64_ctx* x = (f64_ctx*)c;
    for (uint32_t i = b; i < e; ++i) {
        double o[3] = {x->rays.ox[i], x->rays.oy[i], x->rays.oz[i]};
        double d[3] = {x->rays.dx[i], x->rays.dy[i], x->rays.dz[i]};
        double tnear = x->rays.tnear[i], tfar = x->rays.tfar[i];
        double b1 = INFINITY, b2 = INFINITY, bm = 0.0;
        uint32_t best = 0xFFFFFFFFu;
        for (uint32_t k = 0; k < x->ntri; ++k) {
            const float* p = x->tri9 + (size_t)k * 9;
            double e1[3], e2[3], pv[3], tv[3], qv[3];
            for (int q = 0; q < 3; ++q) {
                e1[q] = (double)p[3 + q] - (double)p[q];
                e2[q] = (double)p[6 + q] - (double)p[q];
                tv[q] = o[q] - (double)p[q];
            }
            pv[0] = d[1] * e2[2] - d[2] * e2[1];
            pv[1] = d[2] * e2[0] - d[0] * e2[2];
            pv[2] = d[0] * e2[1] - d[1] * e2[0];
            double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
            if (det == 0.0) continue;
            double inv = 1.0 / det;
            double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
            if (u < 0.0 || u > 1.0) continue;
            qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
            qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
            qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
            double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
            if (v < 0.0 || u + v > 1.0) continue;
            double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
            if (!(t > tnear && t <= tfar)) continue;
            if (t < b1) {
                b2 = b1;
                b1 = t;
                best = k;
                double w = 1.0 - u - v;
                bm = u < v ? (u < w ? u : w) : (v < w ? v : w);
            } else if (t < b2) {
                b2 = t;
            }
        }
        x->t1[i] = b1;
        x->t2[i] = b2;
        x->margin[i] = bm;
        x->prim[i] = best;
    }
}

void orc_closest_brute_f64(const float* tri9, uint32_t ntri, const orc_rays* rays, uint32_t n, uint32_t* prim,
                           double* t1, double* t2, double* margin, int threads)
{
    f64_ctx c = {tri9, ntri, *rays, t1, t2, margin, prim};
    orc_parallel_for(brute_f64_range, &c, n, threads);
}

/* ---- top-down SAH BVH (binned, 16 bins, leaves of <= 4 triangles) ------------------------------ */

typedef struct {
    float lo[3], hi[3];
    uint32_t left_or_first; /* internal: index of left child (right = left + 1); leaf: first prim */
    uint32_t count;         /* 0 for internal nodes */
} orc_node;

struct orc_bvh {
    tri_rec* recs;     /* in original order */
    uint32_t* prims;   /* leaf order -> original index */
    orc_node* nodes;
    uint32_t n_nodes;
    uint32_t ntri;
    uint32_t depth; /* deepest node level (root = 1) */
    float eps;
};

typedef struct {
    float lo[3], hi[3];
} box3;

static inline void box_empty(box3* b)
{
    for (int k = 0; k < 3; ++k) {
        b->lo[k] = INFINITY;
        b->hi[k] = -INFINITY;
    }
}
static inline void box_grow(box3* b, const float* lo, const float* hi)
{
    for (int k = 0; k < 3; ++k) {
        if (lo[k] < b->lo[k]) b->lo[k] = lo[k];
        if (hi[k] > b->hi[k]) b->hi[k] = hi[k];
    }
}
static inline float box_area(const box3* b)
{
    float dx = b->hi[0] - b->lo[0], dy = b->hi[1] - b->lo[1], dz = b->hi[2] - b->lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

typedef struct {
    box3* tbox;
    float* cent; /* 3 per tri */
    uint32_t* prims;
    orc_node* nodes;
    uint32_t n_nodes;
    uint32_t depth;
    float eps;
} build_ctx;

static void build_rec(build_ctx* c, uint32_t ni, uint32_t first, uint32_t count, uint32_t level)
{
    if (level > c->depth) c->depth = level;
    box3 nb, cb;
    box_empty(&nb);
    box_empty(&cb);
    for (uint32_t i = first; i < first + count; ++i) {
        uint32_t p = c->prims[i];
        box_grow(&nb, c->tbox[p].lo, c->tbox[p].hi);
        box_grow(&cb, &c->cent[3 * p], &c->cent[3 * p]);
    }
    orc_node* n = &c->nodes[ni];
    for (int k = 0; k < 3; ++k) {
        n->lo[k] = nb.lo[k] - c->eps;
        n->hi[k] = nb.hi[k] + c->eps;
    }
    if (count <= 4) {
        n->left_or_first = first;
        n->count = count;
        return;
    }
    enum { NB = 16 };
    int best_axis = -1, best_split = -1;
    float best_cost = INFINITY;
    for (int axis = 0; axis < 3; ++axis) {
        float lo = cb.lo[axis], ext = cb.hi[axis] - cb.lo[axis];
        if (!(ext > 0.0f)) continue;
        box3 bb[NB];
        uint32_t bc[NB];
        for (int b = 0; b < NB; ++b) {
            box_empty(&bb[b]);
            bc[b] = 0;
        }
        for (uint32_t i = first; i < first + count; ++i) {
            uint32_t p = c->prims[i];
            int b = (int)((c->cent[3 * p + axis] - lo) / ext * NB);
            if (b >= NB) b = NB - 1;
            if (b < 0) b = 0;
            box_grow(&bb[b], c->tbox[p].lo, c->tbox[p].hi);
            bc[b]++;
        }
        float la[NB], ra[NB];
        uint32_t lc[NB], rc[NB];
        box3 acc;
        box_empty(&acc);
        uint32_t cc = 0;
        for (int b = 0; b < NB; ++b) {
            box_grow(&acc, bb[b].lo, bb[b].hi);
            cc += bc[b];
            la[b] = box_area(&acc);
            lc[b] = cc;
        }
        box_empty(&acc);
        cc = 0;
        for (int b = NB - 1; b >= 0; --b) {
            box_grow(&acc, bb[b].lo, bb[b].hi);
            cc += bc[b];
            ra[b] = box_area(&acc);
            rc[b] = cc;
        }
        for (int s = 0; s < NB - 1; ++s) {
            if (lc[s] == 0 || rc[s + 1] == 0) continue;
            float cost = la[s] * lc[s] + ra[s + 1] * rc[s + 1];
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = axis;
                best_split = s;
            }
        }
    }
    uint32_t mid;
    if (best_axis < 0) {
        mid = first + count / 2; /* all centroids equal: split the list in half */
    } else {
        float lo = cb.lo[best_axis], ext = cb.hi[best_axis] - cb.lo[best_axis];
        uint32_t i = first, j = first + count;
        while (i < j) {
            uint32_t p = c->prims[i];
            int b = (int)((c->cent[3 * p + best_axis] - lo) / ext * NB);
            if (b >= NB) b = NB - 1;
            if (b < 0) b = 0;
            if (b <= best_split) {
                ++i;
            } else {
                --j;
                uint32_t tmp = c->prims[i];
                c->prims[i] = c->prims[j];
                c->prims[j] = tmp;
            }
        }
        mid = i;
        if (mid == first || mid == first + count) mid = first + count / 2;
    }
    uint32_t left = c->n_nodes;
    c->n_nodes += 2;
    n->left_or_first = left;
    n->count = 0;
    build_rec(c, left, first, mid - first, level + 1);
    build_rec(c, left + 1, mid, first + count - mid, level + 1);
}

orc_bvh* orc_bvh_build(const float* tri9, uint32_t ntri)
{
    orc_bvh* b = (orc_bvh*)calloc(1, sizeof(orc_bvh));
    b->ntri = ntri;
    b->recs = make_recs(tri9, ntri);
    b->prims = (uint32_t*)malloc(sizeof(uint32_t) * (ntri ? ntri : 1));
    b->nodes = (orc_node*)malloc(sizeof(orc_node) * (2 * (size_t)ntri + 1));
    if (ntri == 0) return b;
    box3* tbox = (box3*)malloc(sizeof(box3) * ntri);
    float* cent = (float*)malloc(sizeof(float) * 3 * ntri);
    float maxabs = 0.0f;
    for (uint32_t i = 0; i < ntri; ++i) {
        const float* p = tri9 + (size_t)i * 9;
        for (int k = 0; k < 3; ++k) {
            float a = p[k], bb = p[3 + k], cc = p[6 + k];
            float lo = fminf(a, fminf(bb, cc)), hi = fmaxf(a, fmaxf(bb, cc));
            tbox[i].lo[k] = lo;
            tbox[i].hi[k] = hi;
            cent[3 * i + k] = 0.5f * (lo + hi);
            if (fabsf(lo) > maxabs) maxabs = fabsf(lo);
            if (fabsf(hi) > maxabs) maxabs = fabsf(hi);
        }
        b->prims[i] = i;
    }
    build_ctx c = {tbox, cent, b->prims, b->nodes, 1, 0, maxabs * (1.0f / 262144.0f)};
    build_rec(&c, 0, 0, ntri, 1);
    b->n_nodes = c.n_nodes;
    b->depth = c.depth;
    b->eps = c.eps;
    free(tbox);
    free(cent);
    return b;
}

void orc_bvh_free(orc_bvh* b)
{
    if (!b) return;
    free(b->recs);
    free(b->prims);
    free(b->nodes);
    free(b);
}

/* slab test in float with exact reciprocals; boxes are inflated, so culling stays conservative */
static inline int box_hit(const orc_node* n, const float* o, const float* inv, float tnear, float tfar, float* tmin_out)
{
    float tmin = tnear, tmax = tfar;
    for (int k = 0; k < 3; ++k) {
        float t0 = (n->lo[k] - o[k]) * inv[k];
        float t1 = (n->hi[k] - o[k]) * inv[k];
        if (t0 > t1) {
            float s = t0;
            t0 = t1;
            t1 = s;
        }
        if (t0 > tmin) tmin = t0;
        if (t1 < tmax) tmax = t1;
    }
    *tmin_out = tmin;
    return tmin <= tmax;
}

typedef struct {
    const orc_bvh* b;
    orc_rays rays;
    orc_hits hits;
    uint32_t* occluded;
    int any;
} trav_ctx;

/* One ray through the tree. `stack` holds at least depth + 4 entries. Closest hit (any == 0):
 * returns the hit triangle's global index or 0xFFFFFFFF, with t/u/v; any hit: returns 1 / 0. */
static uint32_t trav_one(const orc_bvh* b, uint32_t* stack, const float* o, const float* d, float tnear, float tfar,
                         int any, float* t_out, float* u_out, float* v_out)
{
    float inv[3];
    for (int k = 0; k < 3; ++k) {
        float dk = fabsf(d[k]) < 1e-20f ? copysignf(1e-20f, d[k]) : d[k];
        inv[k] = 1.0f / dk;
    }
    float best_t = tfar, bu = 0.0f, bv = 0.0f;
    uint32_t best = 0xFFFFFFFFu;
    int occ = 0;
    int sp = 0;
    if (b->ntri > 0) stack[sp++] = 0;
    while (sp > 0 && !occ) {
        const orc_node* n = &b->nodes[stack[--sp]];
        float tm;
        if (!box_hit(n, o, inv, tnear, best_t, &tm)) continue;
        if (n->count > 0) {
            for (uint32_t q = n->left_or_first; q < n->left_or_first + n->count; ++q) {
                uint32_t k = b->prims[q];
                float t, u, v;
                if (!orc_tri_test(b->recs[k].v0, b->recs[k].e1, b->recs[k].e2, o, d, tnear, &t, &u, &v)) continue;
                if (any) {
                    if (t <= tfar) {
                        occ = 1;
                        break;
                    }
                } else if (t < best_t || (t == best_t && k < best)) {
                    best_t = t;
                    best = k;
                    bu = u;
                    bv = v;
                }
            }
        } else {
            uint32_t l = n->left_or_first;
            float tl, tr;
            int hl = box_hit(&b->nodes[l], o, inv, tnear, best_t, &tl);
            int hr = box_hit(&b->nodes[l + 1], o, inv, tnear, best_t, &tr);
            if (hl && hr) {
                if (tl <= tr) {
                    stack[sp++] = l + 1;
                    stack[sp++] = l;
                } else {
                    stack[sp++] = l;
                    stack[sp++] = l + 1;
                }
            } else if (hl) {
                stack[sp++] = l;
            } else if (hr) {
                stack[sp++] = l + 1;
            }
        }
    }
    if (any) return (uint32_t)occ;
    *t_out = best_t;
    *u_out = best == 0xFFFFFFFFu ? 0.0f : bu;
    *v_out = best == 0xFFFFFFFFu ? 0.0f : bv;
    return best;
}

uint32_t orc_bvh_stack_size(const orc_bvh* b) { return b->depth + 4; }

uint32_t orc_bvh_closest1(const orc_bvh* b, uint32_t* stack, const float* o, const float* d, float tnear, float tfar,
                          float* t, float* u, float* v)
{
    return trav_one(b, stack, o, d, tnear, tfar, 0, t, u, v);
}

int orc_bvh_occluded1(const orc_bvh* b, uint32_t* stack, const float* o, const float* d, float tnear, float tfar)
{
    return (int)trav_one(b, stack, o, d, tnear, tfar, 1, NULL, NULL, NULL);
}

static void trav_range(void* cv, uint32_t begin, uint32_t end)
{
    trav_ctx* c = (trav_ctx*)cv;
    const orc_bvh* b = c->b;
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (b->depth + 4));
    for (uint32_t i = begin; i < end; ++i) {
        float o[3] = {c->rays.ox[i], c->rays.oy[i], c->rays.oz[i]};
        float d[3] = {c->rays.dx[i], c->rays.dy[i], c->rays.dz[i]};
        float tnear = c->rays.tnear[i], tfar = c->rays.tfar[i];
        if (c->any) {
            c->occluded[i] = trav_one(b, stack, o, d, tnear, tfar, 1, NULL, NULL, NULL);
        } else {
            c->hits.prim[i] = trav_one(b, stack, o, d, tnear, tfar, 0, &c->hits.t[i], &c->hits.u[i], &c->hits.v[i]);
        }
    }
    free(stack);
}

void orc_bvh_closest(const orc_bvh* b, const orc_rays* rays, uint32_t n, const orc_hits* hits, int threads)
{
    trav_ctx c = {b, *rays, *hits, NULL, 0};
    orc_parallel_for(trav_range, &c, n, threads);
}

void orc_bvh_occluded(const orc_bvh* b, const orc_rays* rays, uint32_t n, uint32_t* occluded, int threads)
{
    trav_ctx c = {b, *rays, {0}, occluded, 1};
    orc_parallel_for(trav_range, &c, n, threads);
}

uint32_t orc_bvh_node_count(const orc_bvh* b) { return b->n_nodes; }

/* ---- instanced scenes --------------------------------------------------------------------------
 * An instanced scene restated as its composed instances (instance i of the scene, then its
 * instance j), each a world-to-local transform (3 rows of 4, float) of a scene of geometries with
 * its own tree. Every instance is tested (no top tree): the ray is moved into the instance's space
 * with the fused multiply-add chains the device uses (x' = fma(w0, x, fma(w1, y, fma(w2, z, w3))),
 * directions without w3), the instance's closest hit found by trav_one, and the smallest t kept,
 * ties to the earlier instance (then, inside it, the smaller triangle index). What Embree does
 * for RTC_GEOMETRY_TYPE_INSTANCE (SubsceneResource.cpp:61-78) and Selas's user-geometry instances
 * (SceneResource.cpp:87-169): t is the world ray's parameter (the direction is not renormalized). */

static void inst_ray(const float* w, const float* o, const float* d, float* lo, float* ld)
{
    for (int r = 0; r < 3; ++r) {
        const float* row = w + 4 * r;
        lo[r] = fmaf(row[0], o[0], fmaf(row[1], o[1], fmaf(row[2], o[2], row[3])));
        ld[r] = fmaf(row[0], d[0], fmaf(row[1], d[1], row[2] * d[2]));
    }
}

typedef struct {
    const orc_inst* insts;
    uint32_t n_inst;
    orc_rays rays;
    orc_hits hits;
    uint32_t* inst_out;
    uint32_t* occluded;
    int any;
} inst_ctx;

static void inst_range(void* cv, uint32_t begin, uint32_t end)
{
    inst_ctx* c = (inst_ctx*)cv;
    uint32_t depth = 0;
    for (uint32_t k = 0; k < c->n_inst; ++k)
        if (c->insts[k].blas->depth > depth) depth = c->insts[k].blas->depth;
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (depth + 4));
    for (uint32_t i = begin; i < end; ++i) {
        const float o[3] = {c->rays.ox[i], c->rays.oy[i], c->rays.oz[i]};
        const float d[3] = {c->rays.dx[i], c->rays.dy[i], c->rays.dz[i]};
        const float tnear = c->rays.tnear[i], tfar = c->rays.tfar[i];
        float best_t = tfar, bu = 0.0f, bv = 0.0f;
        uint32_t best = 0xFFFFFFFFu, best_inst = 0xFFFFFFFFu;
        int occ = 0;
        for (uint32_t k = 0; k < c->n_inst && !occ; ++k) {
            float lo[3], ld[3];
            inst_ray(c->insts[k].w2l, o, d, lo, ld);
            if (c->any) {
                occ = (int)trav_one(c->insts[k].blas, stack, lo, ld, tnear, tfar, 1, NULL, NULL, NULL);
                continue;
            }
            float t, u, v;
            uint32_t p = trav_one(c->insts[k].blas, stack, lo, ld, tnear, best_t, 0, &t, &u, &v);
            if (p != 0xFFFFFFFFu && (best == 0xFFFFFFFFu || t < best_t)) {
                best_t = t;
                best = p;
                best_inst = k;
                bu = u;
                bv = v;
            }
        }
        if (c->any) {
            c->occluded[i] = (uint32_t)occ;
        } else {
            c->hits.t[i] = best_t;
            c->hits.u[i] = bu;
            c->hits.v[i] = bv;
            c->hits.prim[i] = best;
            c->inst_out[i] = best_inst;
        }
    }
    free(stack);
}

void orc_inst_closest(const orc_inst* insts, uint32_t n_inst, const orc_rays* rays, uint32_t n, const orc_hits* hits,
                      uint32_t* inst_out, int threads)
{
    inst_ctx c = {insts, n_inst, *rays, *hits, inst_out, NULL, 0};
    orc_parallel_for(inst_range, &c, n, threads);
}

void orc_inst_occluded(const orc_inst* insts, uint32_t n_inst, const orc_rays* rays, uint32_t n, uint32_t* occluded,
                       int threads)
{
    inst_ctx c = {insts, n_inst, *rays, {0}, NULL, occluded, 1};
    orc_parallel_for(inst_range, &c, n, threads);
}
