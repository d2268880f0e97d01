/*
 * ORACLE -- test infrastructure only (see orc_trace.c for the rules). Compiled with
 * -ffp-contract=off.
 *
 * CPU restatement of the per-face texture lookup Selas makes for materials with eUsesPtex:
 * PtexFilter{f_bspline}.eval(out, 0, 3, primId, u, v, 0, 0, 0, 0) (SurfaceParameters.cpp:196-203),
 * i.e. Ptex 2.x's separable cubic B-spline filter with a zero footprint:
 *   PtexSeparableFilter::eval / splitAndApply / applyAcrossEdge / applyToCorner /
 *   applyToCornerFace / apply          Middleware/ptex/Include/PtexSeparableFilter.cpp:45-330
 *   PtexWidth4Filter::buildKernelAxis,
 *   PtexBicubicFilter (sharpness 0)    Middleware/ptex/Include/PtexFilters.cpp:116-282
 *   PtexSeparableKernel (split, merge,
 *   rotate, makeSymmetric, apply)      Middleware/ptex/Include/PtexSeparableKernel.h:48-480,
 *                                      PtexSeparableKernel.cpp:40-75
 * for what the renderer's textures are: quad meshes, 8-bit RGB, every face at one resolution (no
 * subfaces, so no kernel up/down-resolution), clamp border mode, no edge filter mode. Same
 * operations in the same order as the reference, so results match its compiled code bit for bit
 * (pinned by tests against vectors recorded from it on Ptex files with borders, rotated
 * neighbours and valence 3, 4 and 5 corners).
 */
#include "oracle.h"
#include "orc_fn.h"

#include <math.h>
#include <stddef.h>
#include <string.h>

#define KMAX 10

typedef struct {
    int res_log2;
    int u, v, uw, vw;
    float ku[KMAX], kv[KMAX]; /* active windows start at index 0 */
    int rot;
} pkernel;

typedef struct {
    const orc_ptex* tx;
    float result[3];
    float weight;
} pfilter;

static float accumulate(const float* p, int n)
{
    float r = 0;
    for (int i = 0; i < n; ++i) r += p[i];
    return r;
}

static float kweight(const pkernel* k) { return accumulate(k->ku, k->uw) * accumulate(k->kv, k->vw); }

static int kres(const pkernel* k) { return 1 << k->res_log2; }

static void drop_front(float* a, int n, int w) { memmove(a, a + w, sizeof(float) * (size_t)(n - w)); }

/* clamp border mode: weights beyond the border fold onto the border texel */
static void merge_l(pkernel* k)
{
    int w = -k->u;
    k->ku[w] += accumulate(k->ku, w);
    drop_front(k->ku, k->uw, w);
    k->uw -= w;
    k->u = 0;
}
static void merge_r(pkernel* k)
{
    int w = k->uw + k->u - kres(k);
    float* kp = k->ku + k->uw - w;
    kp[-1] += accumulate(kp, w);
    k->uw -= w;
}
static void merge_b(pkernel* k)
{
    int w = -k->v;
    k->kv[w] += accumulate(k->kv, w);
    drop_front(k->kv, k->vw, w);
    k->vw -= w;
    k->v = 0;
}
static void merge_t(pkernel* k)
{
    int w = k->vw + k->v - kres(k);
    float* kp = k->kv + k->vw - w;
    kp[-1] += accumulate(kp, w);
    k->vw -= w;
}

static void kset(pkernel* d, int res_log2, int u, int v, int uw, int vw, const float* ku, const float* kv, int rot)
{
    float tu[KMAX], tv[KMAX];
    memcpy(tu, ku, sizeof(float) * (size_t)uw);
    memcpy(tv, kv, sizeof(float) * (size_t)vw);
    d->res_log2 = res_log2;
    d->u = u;
    d->v = v;
    d->uw = uw;
    d->vw = vw;
    memcpy(d->ku, tu, sizeof(float) * (size_t)uw);
    memcpy(d->kv, tv, sizeof(float) * (size_t)vw);
    d->rot = rot;
}

static void split_l(pkernel* k, pkernel* out)
{
    int w = -k->u;
    if (w < k->uw) {
        kset(out, k->res_log2, kres(k) - w, k->v, w, k->vw, k->ku, k->kv, 0);
        k->u = 0;
        drop_front(k->ku, k->uw, w);
        k->uw -= w;
    } else {
        *out = *k;
        out->u += kres(k);
        k->u = 0;
        k->uw = 0;
    }
}
static void split_r(pkernel* k, pkernel* out)
{
    int w = k->u + k->uw - kres(k);
    if (w < k->uw) {
        kset(out, k->res_log2, 0, k->v, w, k->vw, k->ku + k->uw - w, k->kv, 0);
        k->uw -= w;
    } else {
        *out = *k;
        out->u -= kres(k);
        k->u = 0;
        k->uw = 0;
    }
}
static void split_b(pkernel* k, pkernel* out)
{
    int w = -k->v;
    if (w < k->vw) {
        kset(out, k->res_log2, k->u, kres(k) - w, k->uw, w, k->ku, k->kv, 0);
        k->v = 0;
        drop_front(k->kv, k->vw, w);
        k->vw -= w;
    } else {
        *out = *k;
        out->v += kres(k);
        k->v = 0;
        k->vw = 0;
    }
}
static void split_t(pkernel* k, pkernel* out)
{
    int w = k->v + k->vw - kres(k);
    if (w < k->vw) {
        kset(out, k->res_log2, k->u, 0, k->uw, w, k->ku, k->kv + k->vw - w, 0);
        k->vw -= w;
    } else {
        *out = *k;
        out->v -= kres(k);
        k->v = 0;
        k->vw = 0;
    }
}

static void reverse(float* a, int n)
{
    for (int i = 0, j = n - 1; i < j; ++i, --j) {
        float t = a[i];
        a[i] = a[j];
        a[j] = t;
    }
}
static void flip_u(pkernel* k)
{
    k->u = kres(k) - k->u - k->uw;
    reverse(k->ku, k->uw);
}
static void flip_v(pkernel* k)
{
    k->v = kres(k) - k->v - k->vw;
    reverse(k->kv, k->vw);
}
static void swap_uv(pkernel* k)
{
    float t[KMAX];
    int tu = k->u, tuw = k->uw;
    k->u = k->v;
    k->v = tu;
    memcpy(t, k->ku, sizeof(t));
    memcpy(k->ku, k->kv, sizeof(t));
    memcpy(k->kv, t, sizeof(t));
    k->uw = k->vw;
    k->vw = tuw;
}
/* rotate `r` steps counter-clockwise */
static void krotate(pkernel* k, int r)
{
    switch (r & 3) {
    default: return;
    case 1: flip_u(k); swap_uv(k); break;
    case 2: flip_u(k); flip_v(k); break;
    case 3: flip_v(k); swap_uv(k); break;
    }
    k->rot = (k->rot + r) & 3;
}

static float make_symmetric(pkernel* k, float initial_weight)
{
    int n = k->uw < k->vw ? k->uw : k->vw;
    k->uw = k->vw = n;
    float new_weight = 0;
    for (int i = 0; i < n; ++i) {
        float sum = k->ku[i] + k->kv[i];
        k->ku[i] = k->kv[i] = sum;
        new_weight += sum;
    }
    new_weight *= new_weight;
    float scale = new_weight == 0 ? 1.f : initial_weight / new_weight;
    if (!(scale >= 1)) {
        if (scale < -1) {
            for (int i = 0; i < n; ++i) k->ku[i] *= -1;
            new_weight = -new_weight;
        } else {
            for (int i = 0; i < n; ++i) k->ku[i] *= scale;
            new_weight = initial_weight;
        }
    }
    return new_weight;
}

static int adjface(const orc_ptex* tx, int f, int e) { return tx->adjfaces[4 * f + e]; }
static int adjedge(const orc_ptex* tx, int f, int e) { return tx->adjedges[4 * f + e]; }

/* Apply<uint8_t, 3>: row sums over u first, then weighted by v */
static void kapply(pfilter* F, const pkernel* k, int face)
{
    if (k->uw <= 0 || k->vw <= 0) return;
    const int res = kres(k);
    const uint8_t* data = F->tx->texels + (size_t)face * res * res * 3;
    for (int j = 0; j < k->vw; ++j) {
        const uint8_t* p = data + ((size_t)(k->v + j) * res + k->u) * 3;
        float row[3];
        for (int c = 0; c < 3; ++c) row[c] = (float)p[c] * k->ku[0];
        for (int i = 1; i < k->uw; ++i)
            for (int c = 0; c < 3; ++c) row[c] += (float)p[3 * i + c] * k->ku[i];
        for (int c = 0; c < 3; ++c) F->result[c] += row[c] * k->kv[j];
    }
}

static void apply_across_edge(pfilter* F, pkernel* k, int face, int eid)
{
    int af = adjface(F->tx, face, eid), ae = adjedge(F->tx, face, eid);
    krotate(k, eid - ae + 2);
    kapply(F, k, af);
}

static void apply_to_corner_face(pfilter* F, pkernel* k, int eid, int cf, int ce)
{
    krotate(k, eid - ce + 2);
    kapply(F, k, cf);
}

static void apply_to_corner(pfilter* F, pkernel* k, int face, int eid)
{
    enum { kMaxValence = 10 };
    int cface[kMaxValence], cedge[kMaxValence];
    int af = face, ae = eid, ncorners = 0;
    for (int i = 0; i < kMaxValence; ++i) {
        int next = adjface(F->tx, af, ae);
        int nexte = (adjedge(F->tx, af, ae) + 1) % 4;
        af = next;
        ae = nexte;
        if (af < 0 || (af == face && ae == eid)) {
            ncorners = i - 2;
            break;
        }
        cface[i] = af;
        cedge[i] = ae;
    }
    if (ncorners == 1) {
        apply_to_corner_face(F, k, eid, cface[1], cedge[1]);
    } else if (ncorners > 1) {
        krotate(k, eid + 2);
        float initial = kweight(k);
        float neww = make_symmetric(k, initial);
        for (int i = 1; i <= ncorners; ++i) {
            pkernel kc = *k;
            apply_to_corner_face(F, &kc, 2, cface[i], cedge[i]);
        }
        F->weight += neww * (float)ncorners - initial;
    } else {
        F->weight -= kweight(k);
    }
}

static void split_and_apply(pfilter* F, pkernel* k, int face)
{
    const orc_ptex* tx = F->tx;
    int res = kres(k);
    int sr = k->u + k->uw > res, sl = k->u < 0, st = k->v + k->vw > res, sb = k->v < 0;
    if (sr || sl || st || sb) {
        pkernel ka, kc;
        if (sr) {
            if (adjface(tx, face, 1) >= 0) {
                split_r(k, &ka);
                if (st) {
                    if (adjface(tx, face, 2) >= 0) {
                        split_t(&ka, &kc);
                        apply_to_corner(F, &kc, face, 2);
                    } else
                        merge_t(&ka);
                }
                if (sb) {
                    if (adjface(tx, face, 0) >= 0) {
                        split_b(&ka, &kc);
                        apply_to_corner(F, &kc, face, 1);
                    } else
                        merge_b(&ka);
                }
                apply_across_edge(F, &ka, face, 1);
            } else
                merge_r(k);
        }
        if (sl) {
            if (adjface(tx, face, 3) >= 0) {
                split_l(k, &ka);
                if (st) {
                    if (adjface(tx, face, 2) >= 0) {
                        split_t(&ka, &kc);
                        apply_to_corner(F, &kc, face, 3);
                    } else
                        merge_t(&ka);
                }
                if (sb) {
                    if (adjface(tx, face, 0) >= 0) {
                        split_b(&ka, &kc);
                        apply_to_corner(F, &kc, face, 0);
                    } else
                        merge_b(&ka);
                }
                apply_across_edge(F, &ka, face, 3);
            } else
                merge_l(k);
        }
        if (st) {
            if (adjface(tx, face, 2) >= 0) {
                split_t(k, &ka);
                apply_across_edge(F, &ka, face, 2);
            } else
                merge_t(k);
        }
        if (sb) {
            if (adjface(tx, face, 0) >= 0) {
                split_b(k, &ka);
                apply_across_edge(F, &ka, face, 0);
            } else
                merge_b(k);
        }
    }
    kapply(F, k, face);
}

/* PtexBicubicFilter::kernelFn with sharpness 0 (B = 1, C = 0: the cubic B-spline) */
static float bspline(float x)
{
    const float B = 1.0f;
    const float c0 = 1.5f - B, c1 = 1.5f * B - 2.5f, c2 = 1.0f - (float)(1.0 / 3.0) * B;
    const float c3 = (float)(1.0 / 3.0) * B - 0.5f, c4 = 2.5f - 1.5f * B, c5 = 2.0f * B - 4.0f;
    const float c6 = 2.0f - (float)(2.0 / 3.0) * B;
    x = fabsf(x);
    if (x < 1.0f) return (c0 * x + c1) * x * x + c2;
    if (x < 2.0f) return ((c3 * x + c4) * x + c5) * x + c6;
    return 0.0f;
}

/* 2-unit cubic Hermite used for wide footprints */
static float blur(float x)
{
    x = fabsf(x);
    return x < 1.0f ? (2.0f * x - 3.0f) * x * x + 1.0f : 0.0f;
}

static int res_from_width(float w)
{
    union {
        float f;
        int32_t i;
    } b;
    b.f = w;
    return 127 - ((b.i >> 23) & 0xff);
}

static float recip_pow2(int power)
{
    union {
        float f;
        int32_t i;
    } b;
    b.i = (127 - power) << 23;
    return b.f;
}

/* PtexWidth4Filter::buildKernelAxis with lerp off, for a zero footprint */
static void build_axis(int* k_reslog2, int* k_u, int* k_uw, float* ku, float u, float uw, int f_reslog2)
{
    uw = uw > recip_pow2(f_reslog2) ? uw : recip_pow2(f_reslog2);
    *k_reslog2 = res_from_width(uw);
    int resu = 1 << *k_reslog2;
    const float lerp2 = 0, lerp1 = 1.0f - lerp2;
    if (uw >= .25f) {
        if (uw < .5f) {
            *k_reslog2 = 2;
            float upix = u * 4.0f - 0.5f;
            int u1 = (int)ceilf(upix - 2), u2 = (int)ceilf(upix + 2);
            u1 = u1 & ~1;
            u2 = (u2 + 1) & ~1;
            *k_u = u1;
            *k_uw = u2 - u1;
            float x1 = (float)u1 - upix;
            for (int i = 0; i < *k_uw; i += 2) {
                float xa = x1 + (float)i, xb = xa + 1.0f, xc = (xa + xb) * 0.25f;
                float s = 1.0f / (uw + .75f);
                float ka = bspline(xa), kb = bspline(xb), kc = blur(xc * s);
                ku[i] = ka * lerp1 + kc * lerp2;
                ku[i + 1] = kb * lerp1 + kc * lerp2;
            }
            return;
        } else if (uw < 1) {
            *k_reslog2 = 1;
            float upix = u * 2.0f - 0.5f;
            *k_u = (int)floorf(u - .5f) * 2;
            *k_uw = 4;
            float x1 = (float)*k_u - upix;
            for (int i = 0; i < *k_uw; i += 2) {
                float xa = x1 + (float)i, xb = xa + 1.0f, xc = (xa + xb) * 0.5f;
                float s = 1.0f / (uw * 1.5f + .5f);
                float ka = blur(xa * s), kb = blur(xb * s), kc = blur(xc * s);
                ku[i] = ka * lerp1 + kc * lerp2;
                ku[i + 1] = kb * lerp1 + kc * lerp2;
            }
            return;
        } else {
            *k_reslog2 = 0;
            float upix = u - .5f;
            *k_uw = 2;
            float ui = floorf(upix);
            *k_u = (int)ui;
            ku[0] = blur(upix - ui);
            ku[1] = 1 - ku[0];
            return;
        }
    }
    float upix = u * (float)resu - 0.5f;
    float uwpix = uw * (float)resu;
    float dupix = 2.0f * uwpix;
    int u1 = (int)ceilf(upix - dupix), u2 = (int)ceilf(upix + dupix);
    *k_u = u1;
    *k_uw = u2 - u1;
    float x1 = ((float)u1 - upix) / uwpix, step = 1.0f / uwpix;
    for (int i = 0; i < *k_uw; ++i) ku[i] = bspline(x1 + (float)i * step);
}

void orc_ptex_eval(const orc_ptex* tx, int face, float u, float v, float* out3)
{
    out3[0] = out3[1] = out3[2] = 0.0f;
    if (face < 0 || (uint32_t)face >= tx->nfaces) return;
    u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    pkernel k;
    memset(&k, 0, sizeof(k));
    int ru, rv;
    build_axis(&ru, &k.u, &k.uw, k.ku, u, 0.0f, tx->res_log2);
    build_axis(&rv, &k.v, &k.vw, k.kv, v, 0.0f, tx->res_log2);
    k.res_log2 = ru; /* square faces: both axes choose the same resolution */
    /* stripZeros */
    while (k.ku[0] == 0) {
        drop_front(k.ku, k.uw, 1);
        k.u++;
        k.uw--;
    }
    while (k.ku[k.uw - 1] == 0) k.uw--;
    while (k.kv[0] == 0) {
        drop_front(k.kv, k.vw, 1);
        k.v++;
        k.vw--;
    }
    while (k.kv[k.vw - 1] == 0) k.vw--;
    pfilter F;
    F.tx = tx;
    F.result[0] = F.result[1] = F.result[2] = 0.0f;
    F.weight = kweight(&k);
    split_and_apply(&F, &k, face);
    float scale = 1.0f / (F.weight * 255.0f);
    for (int c = 0; c < 3; ++c) out3[c] = (float)(F.result[c] * scale);
}

/* ---- triangle meshes ------------------------------------------------------------------------
 * PtexFilter::getFilter serves triangle-mesh textures with PtexTriangleFilter whatever the filter
 * type asked for (PtexFilters.cpp:436-440): an elliptical Gaussian over the triangular texel grid.
 * Restated from PtexTriangleFilter.cpp:45-265 (eval, buildKernel, splitAndApply, applyAcrossEdge,
 * apply, applyIter), PtexTriangleKernel.h:48-231 (split, rotate, reorient, clampRes, clampExtent,
 * getIterators) and PtexTriangleKernel.cpp:42-75 (gaussian, Apply for 3 channels of 8-bit data),
 * for the zero footprint Selas asks for (uw1 = vw1 = uw2 = vw2 = 0, width 1, blur 0), one
 * resolution for every face, no neighbourhood-constant faces. Triangle edge e runs from corner e
 * to corner e + 1: edge 0 is v = 0, edge 1 is w = 0, edge 2 is u = 0. The Gaussian is exp of a
 * float argument: the reference's call resolves to the float overload (expf), orc_set_libm mode 1;
 * mode 0 rounds the double exp once, as the device does (orc_fn.h). */
#define TRI_KERNEL_WIDTH 3.5f

typedef struct {
    int res_log2;
    float u, v, u1, v1, w1, u2, v2, w2, A, B, C;
} tkernel;

typedef struct {
    int rowlen;
    float u, v;
    int u1, v1, w1, u2, v2, w2;
    float A, B, C;
    int valid;
    float wscale, weight;
} titer;

static float sqf(float x) { return x * x; }
static float fminf_(float a, float b) { return a < b ? a : b; }
static float fmaxf_(float a, float b) { return a > b ? a : b; }

static void tset8(tkernel* k, float u, float v, float u1, float v1, float w1, float u2, float v2, float w2)
{
    k->u = u;
    k->v = v;
    k->u1 = u1;
    k->v1 = v1;
    k->w1 = w1;
    k->u2 = u2;
    k->v2 = v2;
    k->w2 = w2;
}

static void tset_abc(tkernel* k, float A, float B, float C)
{
    k->A = A;
    k->B = B;
    k->C = C;
}

/* rotate1: u' = w, v' = u, w' = v; rotate2: u' = v, v' = w, w' = u */
static void trotate1(tkernel* k) { tset_abc(k, k->C, 2.0f * k->C - k->B, k->A + k->C - k->B); }
static void trotate2(tkernel* k) { tset_abc(k, k->A + k->C - k->B, 2.0f * k->A - k->B, k->A); }

static void treorient(tkernel* k, int eid, int aeid)
{
    const float u = k->u, v = k->v, w = 1.0f - k->u - k->v;
    const float u1 = k->u1, v1 = k->v1, w1 = k->w1, u2 = k->u2, v2 = k->v2, w2 = k->w2;
    switch (eid * 3 + aeid) {
    case 0: tset8(k, 1.0f - u, -v, 1.0f - u2, -v2, 1.0f - w2, 1.0f - u1, -v1, 1.0f - w1); break;
    case 1: tset8(k, 1.0f - w, 1.0f - u, 1.0f - w2, 1.0f - u2, -v2, 1.0f - w1, 1.0f - u1, -v1); trotate1(k); break;
    case 2: tset8(k, -v, 1.0f - w, -v2, 1.0f - w2, 1.0f - u2, -v1, 1.0f - w1, 1.0f - u1); trotate2(k); break;
    case 3: tset8(k, 1.0f - v, -w, 1.0f - v2, -w2, 1.0f - u2, 1.0f - v1, -w1, 1.0f - u1); trotate2(k); break;
    case 4: tset8(k, 1.0f - u, 1.0f - v, 1.0f - u2, 1.0f - v2, -w2, 1.0f - u1, 1.0f - v1, -w1); break;
    case 5: tset8(k, -w, 1.0f - u, -w2, 1.0f - u2, 1.0f - v2, -w1, 1.0f - u1, 1.0f - v1); trotate1(k); break;
    case 6: tset8(k, 1.0f - w, -u, 1.0f - w2, -u2, 1.0f - v2, 1.0f - w1, -u1, 1.0f - v1); trotate1(k); break;
    case 7: tset8(k, 1.0f - v, 1.0f - w, 1.0f - v2, 1.0f - w2, -u2, 1.0f - v1, 1.0f - w1, -u1); trotate2(k); break;
    case 8: tset8(k, -u, 1.0f - v, -u2, 1.0f - v2, 1.0f - w2, -u1, 1.0f - v1, 1.0f - w1); break;
    }
}

static float tri_gaussian(float q)
{
    const float scale = -0.5f * (TRI_KERNEL_WIDTH * TRI_KERNEL_WIDTH);
    return exp_f(scale * q);
}

static void tri_get_iterators(const tkernel* k, titer* ke, titer* ko)
{
    const int resu = 1 << k->res_log2;
    const float third = (float)(1 / 3.0);
    float Finv = 1.0f / ((float)resu * (float)resu * (k->A * k->C - 0.25f * k->B * k->B));
    float Ak = k->A * Finv, Bk = k->B * Finv, Ck = k->C * Finv;
    ke->rowlen = resu;
    ke->wscale = 1.0f / ((float)resu * (float)resu);
    float scale = (float)ke->rowlen;
    ke->u = k->u * scale - third;
    ke->v = k->v * scale - third;
    ke->u1 = (int)ceilf(k->u1 * scale - third);
    ke->v1 = (int)ceilf(k->v1 * scale - third);
    ke->w1 = (int)ceilf(k->w1 * scale - third);
    ke->u2 = (int)ceilf(k->u2 * scale - third);
    ke->v2 = (int)ceilf(k->v2 * scale - third);
    ke->w2 = (int)ceilf(k->w2 * scale - third);
    ke->A = Ak;
    ke->B = Bk;
    ke->C = Ck;
    ke->valid = ke->u2 > ke->u1 && ke->v2 > ke->v1 && ke->w2 > ke->w1;
    ke->weight = 0;
    ko->rowlen = ke->rowlen;
    ko->wscale = ke->wscale;
    ko->u = (1.0f - k->v) * scale - third;
    ko->v = (1.0f - k->u) * scale - third;
    ko->u1 = (int)ceilf((1.0f - k->v2) * scale - third);
    ko->v1 = (int)ceilf((1.0f - k->u2) * scale - third);
    ko->w1 = (int)ceilf((-k->w2) * scale - third);
    ko->u2 = (int)ceilf((1.0f - k->v1) * scale - third);
    ko->v2 = (int)ceilf((1.0f - k->u1) * scale - third);
    ko->w2 = (int)ceilf((-k->w1) * scale - third);
    ko->A = Ck;
    ko->B = Bk;
    ko->C = Ak;
    ko->valid = ko->u2 > ko->u1 && ko->v2 > ko->v1 && ko->w2 > ko->w1;
    ko->weight = 0;
}

/* Apply<uint8_t, 3>: the ellipse's Q by finite differences along each row (Heckbert) */
static void tri_iter_apply(titer* k, float* result, const uint8_t* data)
{
    float DDQ = 2.0f * k->A;
    for (int vi = k->v1; vi != k->v2; vi++) {
        int xw = k->rowlen - vi;
        int x1 = k->u1 > xw - k->w2 ? k->u1 : xw - k->w2;
        int x2 = k->u2 < xw - k->w1 ? k->u2 : xw - k->w1;
        float U = (float)x1 - k->u;
        float V = (float)vi - k->v;
        float DQ = k->A * (2.0f * U + 1.0f) + k->B * V;
        float Q = k->A * U * U + (k->B * U + k->C * V) * V;
        const uint8_t* p = data + ((size_t)vi * k->rowlen + x1) * 3;
        const uint8_t* pEnd = p + (ptrdiff_t)(x2 - x1) * 3;
        for (; p < pEnd; p += 3) {
            if (Q < 1.0f) {
                float weight = tri_gaussian(Q) * k->wscale;
                k->weight += weight;
                result[0] += (float)p[0] * weight;
                result[1] += (float)p[1] * weight;
                result[2] += (float)p[2] * weight;
            }
            Q += DQ;
            DQ += DDQ;
        }
    }
}

static void tri_apply(pfilter* F, tkernel* k, int face)
{
    /* clampRes (one resolution: the face's), clampExtent */
    if (k->res_log2 > F->tx->res_log2) k->res_log2 = F->tx->res_log2;
    k->u1 = fmaxf_(k->u1, 0.0f);
    k->v1 = fmaxf_(k->v1, 0.0f);
    k->w1 = fmaxf_(k->w1, 0.0f);
    k->u2 = fminf_(k->u2, 1.0f - (k->v1 + k->w1));
    k->v2 = fminf_(k->v2, 1.0f - (k->w1 + k->u1));
    k->w2 = fminf_(k->w2, 1.0f - (k->u1 + k->v1));
    titer ke, ko;
    tri_get_iterators(k, &ke, &ko);
    if (!ke.valid && !ko.valid) return;
    const int res = 1 << F->tx->res_log2;
    const uint8_t* data = F->tx->texels + (size_t)face * res * res * 3;
    if (ke.valid) {
        tri_iter_apply(&ke, F->result, data);
        F->weight += ke.weight;
    }
    if (ko.valid) {
        tri_iter_apply(&ko, F->result, data);
        F->weight += ko.weight;
    }
}

static void tri_split_and_apply(pfilter* F, tkernel* k, int face);

static void tri_apply_across_edge(pfilter* F, tkernel* k, int face, int eid)
{
    int af = adjface(F->tx, face, eid), ae = adjedge(F->tx, face, eid);
    treorient(k, eid, ae);
    tri_split_and_apply(F, k, af);
}

static void tri_split_and_apply(pfilter* F, tkernel* k, int face)
{
    if (k->u1 < 0 && adjface(F->tx, face, 2) >= 0) {
        tkernel ka = *k;
        k->u1 = 0;
        ka.u2 = 0;
        tri_apply_across_edge(F, &ka, face, 2);
    }
    if (k->v1 < 0 && adjface(F->tx, face, 0) >= 0) {
        tkernel ka = *k;
        k->v1 = 0;
        ka.v2 = 0;
        tri_apply_across_edge(F, &ka, face, 0);
    }
    if (k->w1 < 0 && adjface(F->tx, face, 1) >= 0) {
        tkernel ka = *k;
        k->w1 = 0;
        ka.w2 = 0;
        tri_apply_across_edge(F, &ka, face, 1);
    }
    tri_apply(F, k, face);
}

static int calc_res_from_width(float w)
{
    union {
        float f;
        int32_t i;
    } b;
    b.f = w;
    return 127 - ((b.i >> 23) & 0xff);
}

/* buildKernel with uw1 = vw1 = uw2 = vw2 = 0, width 1, blur 0, then splitAndApply and the
 * normalisation of eval */
void orc_ptex_tri_eval(const orc_ptex* tx, int face, float u, float v, float* out3)
{
    out3[0] = out3[1] = out3[2] = 0.0f;
    if (face < 0 || (uint32_t)face >= tx->nfaces) return;
    u = fminf_(fmaxf_(u, 0.0f), 1.0f);
    v = fminf_(fmaxf_(v, 0.0f), 1.0f);
    const float sqrt3 = 1.7320508075688772f;
    const float width = 1.0f, blur = 0.0f, uw1 = 0.0f, vw1 = 0.0f, uw2 = 0.0f, vw2 = 0.0f;
    float scaleAC = 0.25f * width * width;
    float scaleB = -2.0f * scaleAC;
    float A = (vw1 * vw1 + vw2 * vw2) * scaleAC;
    float B = (uw1 * vw1 + uw2 * vw2) * scaleB;
    float C = (uw1 * uw1 + uw2 * uw2) * scaleAC;
    float Ac = 0.75f * A;
    float Bc = (sqrt3 / 2) * (B - A);
    float Cc = 0.25f * A - 0.5f * B + C;
    const float maxEcc = 15.0f;
    const float eccRatio = (maxEcc * maxEcc + 1.0f) / (maxEcc * maxEcc - 1.0f);
    float X = sqrtf(sqf(Ac - Cc) + sqf(Bc));
    float b_e = 0.5f * (eccRatio * X - (Ac + Cc));
    float b_t = sqf(0.5f / (float)(1 << tx->res_log2));
    float b_b = 0.25f * blur * blur;
    float b = fmaxf_(b_b, fmaxf_(b_e, b_t));
    Ac += b;
    Cc += b;
    float m = sqrtf(2.0f * (Ac * Cc - 0.25f * Bc * Bc) / (Ac + Cc + X));
    int reslog2 = calc_res_from_width(2.0f * m);
    if (reslog2 < 0) reslog2 = 0;
    A = (float)(4 / 3.0) * Ac;
    B = (2 / sqrt3) * Bc + A;
    C = -0.25f * A + 0.5f * B + Cc;
    float scale = TRI_KERNEL_WIDTH * TRI_KERNEL_WIDTH;
    A *= scale;
    B *= scale;
    C *= scale;
    float uw = fminf_(sqrtf(C), 1.0f);
    float vw = fminf_(sqrtf(A), 1.0f);
    float ww = fminf_(sqrtf(A - B + C), 1.0f);
    float w = 1.0f - u - v;
    tkernel k;
    k.res_log2 = reslog2;
    tset8(&k, u, v, u - uw, v - vw, w - ww, u + uw, v + vw, w + ww);
    tset_abc(&k, A, B, C);
    pfilter F;
    F.tx = tx;
    F.result[0] = F.result[1] = F.result[2] = 0.0f;
    F.weight = 0.0f;
    tri_split_and_apply(&F, &k, face);
    float s = 1.0f / (F.weight * 255.0f);
    for (int c = 0; c < 3; ++c) out3[c] = (float)(F.result[c] * s);
}

void orc_ptex_eval_batch(const orc_ptex* tx, uint32_t n, const int32_t* faces, const float* u, const float* v,
                         float* out)
{
    for (uint32_t i = 0; i < n; ++i) {
        if (tx->tri)
            orc_ptex_tri_eval(tx, faces[i], u[i], v[i], out + 3 * (size_t)i);
        else
            orc_ptex_eval(tx, faces[i], u[i], v[i], out + 3 * (size_t)i);
    }
}
