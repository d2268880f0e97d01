/*
 * ORACLE -- test infrastructure only (see orc_trace.c): nothing in the product links, loads or
 * calls this code.
 *
 * Round B-spline curves, the geometry Selas builds for hair and grass with
 * rtcNewGeometry(RTC_GEOMETRY_TYPE_ROUND_BSPLINE_CURVE) (Source/Core/SceneLib/ModelResource.cpp:
 * 300-330): per segment an index to the first of four consecutive control points (x, y, z, radius;
 * RTC_FORMAT_FLOAT4 vertices, RTC_FORMAT_UINT indices). The segment is the uniform cubic B-spline
 * P(u), r(u), u in [0, 1], and the surface is that of the sphere swept along it: the union of the
 * balls B(P(u), r(u)), end balls included. A ray o + t d first meets it at the smallest entry
 * t(u) = (-(d.oc) - sqrt(disc(u))) / (d.d) (oc = o - P(u), disc = (d.oc)^2 - (d.d)(oc.oc - r^2))
 * over the u whose ball it crosses; the hit reports u and the normal X - P(u) at X = o + t d (the
 * swept surface touches ball u there; u is refined as the root of (X - P(u)).P'(u) + r(u) r'(u),
 * which is steep in u where the entry is flat). Embree's library is not in the reference (headers only), so
 * its arithmetic and its conventions beyond this definition (v, back-face hits of rays starting
 * inside a curve) are parity unpinned; v is reported as 0 and only entries are hits.
 *
 * Two restatements:
 *   orc_curve_hit_dense  the definition itself: t(u) at 8192 + 1 parameters, the least entry
 *                        refined by golden-section search in double precision (pins the other;
 *                        pinned itself by closed forms: a collapsed segment is a sphere, equal
 *                        collinear points with equal radii a capsule, see tests/test_oracle_curves.py);
 *   orc_curve_hit        the device's algorithm in double: the segment in 16 pieces, each bounded
 *                        by a capsule around its chord (radius: the larger end radius plus h^2/8
 *                        times the second-difference bounds of position and radius), and for each
 *                        capsule the ray meets, the entry minimised over the piece's parameters
 *                        (widened by half a piece each side) by golden-section search.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>

/* uniform cubic B-spline basis at u (the segment's 4 control points, xyz + radius) */
static void bs_eval(const float* cp, double u, double* P, double* r)
{
    const double s = 1.0 - u;
    const double b0 = s * s * s / 6.0;
    const double b1 = (3.0 * u * u * u - 6.0 * u * u + 4.0) / 6.0;
    const double b2 = (-3.0 * u * u * u + 3.0 * u * u + 3.0 * u + 1.0) / 6.0;
    const double b3 = u * u * u / 6.0;
    for (int k = 0; k < 3; ++k) P[k] = b0 * cp[k] + b1 * cp[4 + k] + b2 * cp[8 + k] + b3 * cp[12 + k];
    *r = b0 * cp[3] + b1 * cp[7] + b2 * cp[11] + b3 * cp[15];
}

/* Entry of the ray into ball u: returns 1 and *t when the line meets the ball; phi (the entry, or
 * past the ball the ray parameter of closest approach plus the clearance, continuous in u) always */
static int ball_entry(const float* cp, const float* o, const float* d, double u, double* t, double* phi)
{
    double P[3], r;
    bs_eval(cp, u, P, &r);
    const double oc[3] = {o[0] - P[0], o[1] - P[1], o[2] - P[2]};
    const double a = (double)d[0] * d[0] + (double)d[1] * d[1] + (double)d[2] * d[2];
    const double hb = d[0] * oc[0] + d[1] * oc[1] + d[2] * oc[2];
    const double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - r * r;
    const double disc = hb * hb - a * c;
    if (disc >= 0.0) {
        *t = (-hb - sqrt(disc)) / a;
        *phi = *t;
        return 1;
    }
    *phi = (-hb + sqrt(-disc)) / a;
    return 0;
}

/* distance from P(u) to the ray's line minus r(u) (negative: the line crosses ball u) */
static double clearance_at(const float* cp, const float* o, const float* d, double u)
{
    double P[3], r;
    bs_eval(cp, u, P, &r);
    const double oc[3] = {o[0] - P[0], o[1] - P[1], o[2] - P[2]};
    const double a = (double)d[0] * d[0] + (double)d[1] * d[1] + (double)d[2] * d[2];
    const double hb = d[0] * oc[0] + d[1] * oc[1] + d[2] * oc[2];
    const double d2 = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - hb * hb / a;
    return sqrt(d2 > 0.0 ? d2 : 0.0) - r;
}

/* golden-section search for the least of f over [lo, hi] (f: 0 the entry phi, 1 the clearance) */
static double golden_min(const float* cp, const float* o, const float* d, double lo, double hi, int iters, int f)
{
    const double g = 0.6180339887498949;
    double x1 = hi - g * (hi - lo), x2 = lo + g * (hi - lo), t, f1, f2;
    if (f) {
        f1 = clearance_at(cp, o, d, x1);
        f2 = clearance_at(cp, o, d, x2);
    } else {
        ball_entry(cp, o, d, x1, &t, &f1);
        ball_entry(cp, o, d, x2, &t, &f2);
    }
    for (int i = 0; i < iters; ++i) {
        if (f1 <= f2) {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = hi - g * (hi - lo);
            if (f) f1 = clearance_at(cp, o, d, x1);
            else ball_entry(cp, o, d, x1, &t, &f1);
        } else {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + g * (hi - lo);
            if (f) f2 = clearance_at(cp, o, d, x2);
            else ball_entry(cp, o, d, x2, &t, &f2);
        }
    }
    /* a minimum at u = 0 or 1 (an end ball) is approached as the bracket shrinks onto that end */
    return 0.5 * (lo + hi);
}

/* The least entry over the run of parameters around uf (whose ball the line crosses) within
 * [lo, hi]: the run's ends by bisection (each kept on the crossing side), then golden-section search
 * of the entry over the run. Returns the parameter. */
static double least_entry(const float* cp, const float* o, const float* d, double lo, double hi, double uf, int iters)
{
    double t, phi;
    if (!ball_entry(cp, o, d, lo, &t, &phi)) {
        double a = lo, b = uf;
        for (int i = 0; i < iters; ++i) {
            const double m = 0.5 * (a + b);
            if (ball_entry(cp, o, d, m, &t, &phi)) b = m;
            else a = m;
        }
        lo = b;
    }
    if (!ball_entry(cp, o, d, hi, &t, &phi)) {
        double a = uf, b = hi;
        for (int i = 0; i < iters; ++i) {
            const double m = 0.5 * (a + b);
            if (ball_entry(cp, o, d, m, &t, &phi)) a = m;
            else b = m;
        }
        hi = a;
    }
    return golden_min(cp, o, d, lo, hi, iters, 0);
}

/* The contact parameter of the hit point X: the root near u of F(u) = (X - P(u)).P'(u) + r(u) r'(u)
 * (Newton's method kept in [0, 1]), then the normal X - P(u). Returns the parameter. */
static double hit_normal(const float* cp, const float* o, const float* d, double t, double u, double* ng)
{
    double X[3], P[3], r;
    for (int k = 0; k < 3; ++k) X[k] = o[k] + t * d[k];
    for (int i = 0; i < 4; ++i) {
        const double w = 1.0 - u;
        const double db[4] = {-0.5 * w * w, 1.5 * u * u - 2.0 * u, -1.5 * u * u + u + 0.5, 0.5 * u * u};
        const double eb[4] = {w, 3.0 * u - 2.0, 1.0 - 3.0 * u, u};
        double dP[3] = {0, 0, 0}, ddP[3] = {0, 0, 0}, dr = 0.0, ddr = 0.0;
        for (int j = 0; j < 4; ++j) {
            for (int k = 0; k < 3; ++k) {
                dP[k] += db[j] * cp[4 * j + k];
                ddP[k] += eb[j] * cp[4 * j + k];
            }
            dr += db[j] * cp[4 * j + 3];
            ddr += eb[j] * cp[4 * j + 3];
        }
        bs_eval(cp, u, P, &r);
        double F = r * dr, dF = dr * dr + r * ddr;
        for (int k = 0; k < 3; ++k) {
            F += (X[k] - P[k]) * dP[k];
            dF += (X[k] - P[k]) * ddP[k] - dP[k] * dP[k];
        }
        if (dF < 0.0) {
            u -= F / dF;
            u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
        }
    }
    bs_eval(cp, u, P, &r);
    for (int k = 0; k < 3; ++k) ng[k] = X[k] - P[k];
    return u;
}

int orc_curve_hit_dense(const float* cp, const float* o, const float* d, float tnear, float tfar, double* t_out,
                        double* u_out, double* ng)
{
    const int M = 8192;
    double best = INFINITY;
    int bi = -1;
    for (int i = 0; i <= M; ++i) {
        double t, phi;
        if (ball_entry(cp, o, d, (double)i / M, &t, &phi) && t >= tnear && t < best) {
            best = t;
            bi = i;
        }
    }
    if (bi < 0) return 0;
    const double lo = bi > 0 ? (double)(bi - 1) / M : 0.0, hi = bi < M ? (double)(bi + 1) / M : 1.0;
    double u = least_entry(cp, o, d, lo, hi, (double)bi / M, 60), t, phi;
    if (!ball_entry(cp, o, d, u, &t, &phi) || t < tnear || t > best) {
        u = (double)bi / M;
        t = best;
    }
    if (t > tfar) return 0;
    *t_out = t;
    *u_out = hit_normal(cp, o, d, t, u, ng);
    return 1;
}

int orc_curve_hit(const float* cp, const float* o, const float* d, float tnear, float tfar, int any, double* t_out,
                  double* u_out, double* ng)
{
    const int N = 16;
    const double h = 1.0 / N;
    double m2 = 0.0, m2r = 0.0;
    for (int j = 0; j < 2; ++j) {
        const float* q = cp + 4 * j;
        double s = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double e = (double)q[k] - 2.0 * q[4 + k] + q[8 + k];
            s += e * e;
        }
        s = sqrt(s);
        if (s > m2) m2 = s;
        const double er = fabs((double)q[3] - 2.0 * q[7] + q[11]);
        if (er > m2r) m2r = er;
    }
    const double dev = h * h / 8.0 * (m2 + m2r);
    const double a = (double)d[0] * d[0] + (double)d[1] * d[1] + (double)d[2] * d[2];
    double best = (double)tfar, bu = 0.0;
    int found = 0;
    double A[3], ra;
    bs_eval(cp, 0.0, A, &ra);
    for (int k = 0; k < N; ++k) {
        double B[3], rb;
        bs_eval(cp, (k + 1) * h, B, &rb);
        const double R = (ra > rb ? ra : rb) + dev;
        /* closest approach of the ray's line and the chord AB */
        const double e[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        const double w0[3] = {o[0] - A[0], o[1] - A[1], o[2] - A[2]};
        const double b = d[0] * e[0] + d[1] * e[1] + d[2] * e[2];
        const double c = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        const double dd = d[0] * w0[0] + d[1] * w0[1] + d[2] * w0[2];
        const double ee = e[0] * w0[0] + e[1] * w0[1] + e[2] * w0[2];
        const double den = a * c - b * b;
        double s = den > 1e-300 ? (a * ee - b * dd) / den : 0.0;
        s = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
        const double tc = (s * b - dd) / a;
        double dist2 = 0.0;
        for (int q = 0; q < 3; ++q) {
            const double x = o[q] + tc * d[q] - (A[q] + s * e[q]);
            dist2 += x * x;
        }
        const double reach = (R + sqrt(c)) / sqrt(a); /* ray parameter span of the capsule */
        if (dist2 <= R * R && tc + reach >= tnear && tc - reach <= best) {
            const double lo = k * h - 0.5 * h < 0.0 ? 0.0 : k * h - 0.5 * h;
            const double hi = (k + 1) * h + 0.5 * h > 1.0 ? 1.0 : (k + 1) * h + 0.5 * h;
            /* the parameter nearest the line; where its ball is crossed, the least entry around it */
            const double ug = golden_min(cp, o, d, lo, hi, 40, 1);
            double t, phi;
            if (clearance_at(cp, o, d, ug) > 0.0 || !ball_entry(cp, o, d, ug, &t, &phi)) goto next;
            const double u = least_entry(cp, o, d, lo, hi, ug, 40);
            if (ball_entry(cp, o, d, u, &t, &phi) && t >= tnear && t <= best) {
                best = t;
                bu = u;
                found = 1;
                if (any) break;
            }
        }
    next:
        for (int q = 0; q < 3; ++q) A[q] = B[q];
        ra = rb;
    }
    if (!found) return 0;
    *t_out = best;
    *u_out = hit_normal(cp, o, d, best, bu, ng);
    return 1;
}

/* the control points' box grown by the largest radius (contains the segment: the basis is a
 * partition of unity with non-negative weights) */
static void seg_box(const float* cp, float* lo, float* hi)
{
    float rmax = 0.0f;
    for (int k = 0; k < 3; ++k) {
        lo[k] = INFINITY;
        hi[k] = -INFINITY;
    }
    for (int j = 0; j < 4; ++j) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], cp[4 * j + k]);
            hi[k] = fmaxf(hi[k], cp[4 * j + k]);
        }
        rmax = fmaxf(rmax, cp[4 * j + 3]);
    }
    for (int k = 0; k < 3; ++k) {
        lo[k] -= rmax;
        hi[k] += rmax;
    }
}

static int ray_box(const float* lo, const float* hi, const float* o, const float* d, double tnear, double tfar)
{
    double t0 = tnear, t1 = tfar;
    for (int k = 0; k < 3; ++k) {
        if (d[k] == 0.0f) {
            if (o[k] < lo[k] || o[k] > hi[k]) return 0;
            continue;
        }
        double a = ((double)lo[k] - o[k]) / d[k], b = ((double)hi[k] - o[k]) / d[k];
        if (a > b) {
            double x = a;
            a = b;
            b = x;
        }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
        if (t0 > t1) return 0;
    }
    return 1;
}

int orc_curves_closest1(const float* cps, const float* boxes, uint32_t nseg, const float* o, const float* d,
                        float tnear, float tfar, int dense, float* t, float* u, uint32_t* prim, float* ng)
{
    double best = tfar, bu = 0.0, bng[3] = {0, 0, 0};
    uint32_t bp = 0xFFFFFFFFu;
    for (uint32_t s = 0; s < nseg; ++s) {
        if (!ray_box(boxes + 6 * (size_t)s, boxes + 6 * (size_t)s + 3, o, d, tnear, best)) continue;
        double ts, us, ns[3];
        const float* cp = cps + 16 * (size_t)s;
        const int hit = dense ? orc_curve_hit_dense(cp, o, d, tnear, (float)best, &ts, &us, ns)
                              : orc_curve_hit(cp, o, d, tnear, (float)best, 0, &ts, &us, ns);
        if (hit && (bp == 0xFFFFFFFFu || ts < best)) {
            best = ts;
            bu = us;
            bp = s;
            for (int k = 0; k < 3; ++k) bng[k] = ns[k];
        }
    }
    if (bp == 0xFFFFFFFFu) return 0;
    *t = (float)best;
    *u = (float)bu;
    *prim = bp;
    for (int k = 0; k < 3; ++k) ng[k] = (float)bng[k];
    return 1;
}

int orc_curves_occluded1(const float* cps, const float* boxes, uint32_t nseg, const float* o, const float* d,
                         float tnear, float tfar)
{
    for (uint32_t s = 0; s < nseg; ++s) {
        if (!ray_box(boxes + 6 * (size_t)s, boxes + 6 * (size_t)s + 3, o, d, tnear, tfar)) continue;
        double ts, us, ns[3];
        if (orc_curve_hit(cps + 16 * (size_t)s, o, d, tnear, tfar, 1, &ts, &us, ns)) return 1;
    }
    return 0;
}

void orc_curve_boxes(const float* cps, uint32_t nseg, float* boxes)
{
    for (uint32_t s = 0; s < nseg; ++s) seg_box(cps + 16 * (size_t)s, boxes + 6 * (size_t)s, boxes + 6 * (size_t)s + 3);
}

/* ---- batches over rays (brute force over the segments, box culled) ------------------------------ */

typedef struct {
    const float* cps;
    float* boxes;
    uint32_t nseg;
    orc_rays rays;
    float *t, *u, *ng;
    uint32_t* prim;
    uint32_t* occ;
    int dense;
} curve_ctx;

static void closest_range(void* c, uint32_t b, uint32_t e)
{
    curve_ctx* x = (curve_ctx*)c;
    for (uint32_t i = b; i < e; ++i) {
        const float o[3] = {x->rays.ox[i], x->rays.oy[i], x->rays.oz[i]};
        const float d[3] = {x->rays.dx[i], x->rays.dy[i], x->rays.dz[i]};
        float t, u, ng[3];
        uint32_t p;
        if (orc_curves_closest1(x->cps, x->boxes, x->nseg, o, d, x->rays.tnear[i], x->rays.tfar[i], x->dense, &t, &u,
                                &p, ng)) {
            x->t[i] = t;
            x->u[i] = u;
            x->prim[i] = p;
            for (int k = 0; k < 3; ++k) x->ng[3 * i + k] = ng[k];
        } else {
            x->t[i] = x->rays.tfar[i];
            x->u[i] = 0.0f;
            x->prim[i] = 0xFFFFFFFFu;
            for (int k = 0; k < 3; ++k) x->ng[3 * i + k] = 0.0f;
        }
    }
}

static void occluded_range(void* c, uint32_t b, uint32_t e)
{
    curve_ctx* x = (curve_ctx*)c;
    for (uint32_t i = b; i < e; ++i) {
        const float o[3] = {x->rays.ox[i], x->rays.oy[i], x->rays.oz[i]};
        const float d[3] = {x->rays.dx[i], x->rays.dy[i], x->rays.dz[i]};
        x->occ[i] = (uint32_t)orc_curves_occluded1(x->cps, x->boxes, x->nseg, o, d, x->rays.tnear[i], x->rays.tfar[i]);
    }
}

void orc_curves_closest(const float* cps, uint32_t nseg, const orc_rays* rays, uint32_t n, float* t, float* u,
                        uint32_t* prim, float* ng, int dense, int threads)
{
    curve_ctx c = {cps, malloc(sizeof(float) * 6 * (size_t)(nseg ? nseg : 1)), nseg, *rays, t, u, ng, prim, NULL, dense};
    orc_curve_boxes(cps, nseg, c.boxes);
    orc_parallel_for(closest_range, &c, n, threads);
    free(c.boxes);
}

void orc_curves_occluded(const float* cps, uint32_t nseg, const orc_rays* rays, uint32_t n, uint32_t* occ, int threads)
{
    curve_ctx c = {cps, malloc(sizeof(float) * 6 * (size_t)(nseg ? nseg : 1)), nseg, *rays, NULL, NULL, NULL, NULL, occ, 0};
    orc_curve_boxes(cps, nseg, c.boxes);
    orc_parallel_for(occluded_range, &c, n, threads);
    free(c.boxes);
}

/* Clearance of a ray from a segment: the least over u in [0, 1] (8192 + 1 samples) of the
 * distance from P(u) to the ray's line minus r(u), in units of the largest control radius
 * (negative: the line passes through the swept surface). Tests use it to leave out grazing rays,
 * whose hit or miss depends on rounding. */
double orc_curve_clearance(const float* cp, const float* o, const float* d)
{
    const int M = 8192;
    double rmax = 0.0, best = INFINITY;
    for (int j = 0; j < 4; ++j)
        if (cp[4 * j + 3] > rmax) rmax = cp[4 * j + 3];
    const double a = (double)d[0] * d[0] + (double)d[1] * d[1] + (double)d[2] * d[2];
    for (int i = 0; i <= M; ++i) {
        double P[3], r;
        bs_eval(cp, (double)i / M, P, &r);
        const double oc[3] = {o[0] - P[0], o[1] - P[1], o[2] - P[2]};
        const double hb = d[0] * oc[0] + d[1] * oc[1] + d[2] * oc[2];
        double d2 = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - hb * hb / a;
        const double cl = sqrt(d2 > 0.0 ? d2 : 0.0) - r;
        if (cl < best) best = cl;
    }
    return rmax > 0.0 ? best / rmax : best;
}
