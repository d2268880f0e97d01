// Round B-spline curve segments: the ray test of Selas's hair and grass geometry
// (rtcNewGeometry(RTC_GEOMETRY_TYPE_ROUND_BSPLINE_CURVE), ModelResource.cpp:300-330), usable on the
// host and the device.
//
// A segment is four consecutive control points (x, y, z, radius) of a uniform cubic B-spline P(u),
// r(u), u in [0, 1]; its surface is that of the sphere swept along it (the union of the balls
// B(P(u), r(u)), end balls included). A ray o + t d first meets the segment at the least entry
// t(u) into a ball it crosses; the hit reports that u (refined as the contact parameter of the hit
// point X, curve_contact), v = 0 and the normal X - P(u) (the swept surface touches ball u there).
// Only entries are hits: a ray starting inside a curve's tube does not hit that tube.
//
// Search: a capsule around the whole segment's chord first (rejects most segments whose box the ray
// crosses); then the segment in kCurvePieces pieces, each bounded by a capsule around its chord whose radius is
// the larger end radius plus h^2 / 8 times the second-difference bounds of position and radius (a
// cubic's distance from its chord over a piece of length h). For each capsule the ray's line meets,
// the parameter nearest the line over the piece (widened by half a piece each side) by golden-section
// search; where that ball is crossed, the run of crossed balls around it by bisection and the least
// entry over the run by golden-section search. Distances to the line are taken from the perpendicular
// vector oc - (d.oc / d.d) d, which keeps float rounding at the scale of |oc|'s ulp for thin curves
// seen from afar.
#pragma once

#include "sgpu_math.h"

namespace sgpu {

constexpr int kCurvePieces = 16;
constexpr int kCurveIters = 20;

struct CurveSeg {
    float4 p[4];  // control points: xyz, radius
};

SGPU_HD void curve_eval(const CurveSeg& s, float u, f3& P, float& r)
{
    const float w = 1.0f - u;
    const float b0 = w * w * w * (1.0f / 6.0f);
    const float b1 = (3.0f * u * u * u - 6.0f * u * u + 4.0f) * (1.0f / 6.0f);
    const float b2 = (-3.0f * u * u * u + 3.0f * u * u + 3.0f * u + 1.0f) * (1.0f / 6.0f);
    const float b3 = u * u * u * (1.0f / 6.0f);
    P = mk3(b0 * s.p[0].x + b1 * s.p[1].x + b2 * s.p[2].x + b3 * s.p[3].x,
            b0 * s.p[0].y + b1 * s.p[1].y + b2 * s.p[2].y + b3 * s.p[3].y,
            b0 * s.p[0].z + b1 * s.p[1].z + b2 * s.p[2].z + b3 * s.p[3].z);
    r = b0 * s.p[0].w + b1 * s.p[1].w + b2 * s.p[2].w + b3 * s.p[3].w;
}

// P'(u), P''(u), r'(u), r''(u) of the segment
SGPU_HD void curve_eval_d(const CurveSeg& s, float u, f3& dP, float& dr, f3& ddP, float& ddr)
{
    const float w = 1.0f - u;
    const float d0 = -0.5f * w * w, d1 = 1.5f * u * u - 2.0f * u, d2 = -1.5f * u * u + u + 0.5f, d3 = 0.5f * u * u;
    const float e0 = w, e1 = 3.0f * u - 2.0f, e2 = 1.0f - 3.0f * u, e3 = u;
    dP = mk3(d0 * s.p[0].x + d1 * s.p[1].x + d2 * s.p[2].x + d3 * s.p[3].x,
             d0 * s.p[0].y + d1 * s.p[1].y + d2 * s.p[2].y + d3 * s.p[3].y,
             d0 * s.p[0].z + d1 * s.p[1].z + d2 * s.p[2].z + d3 * s.p[3].z);
    dr = d0 * s.p[0].w + d1 * s.p[1].w + d2 * s.p[2].w + d3 * s.p[3].w;
    ddP = mk3(e0 * s.p[0].x + e1 * s.p[1].x + e2 * s.p[2].x + e3 * s.p[3].x,
              e0 * s.p[0].y + e1 * s.p[1].y + e2 * s.p[2].y + e3 * s.p[3].y,
              e0 * s.p[0].z + e1 * s.p[1].z + e2 * s.p[2].z + e3 * s.p[3].z);
    ddr = e0 * s.p[0].w + e1 * s.p[1].w + e2 * s.p[2].w + e3 * s.p[3].w;
}

// The parameter of the swept surface's contact with the ball through hit point X: the root of
// F(u) = (X - P(u)).P'(u) + r(u) r'(u) (the u-derivative of |X - P|^2 - r^2) near u, by Newton's
// method kept in [0, 1] (an end ball's hit stays at its end). The least entry is flat in u, so the
// search leaves u uncertain to the square root of t's rounding; F is steep in u.
SGPU_HD float curve_contact(const CurveSeg& s, f3 X, float u)
{
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        f3 P, dP, ddP;
        float r, dr, ddr;
        curve_eval(s, u, P, r);
        curve_eval_d(s, u, dP, dr, ddP, ddr);
        const f3 xp = X - P;
        const float F = dot3(xp, dP) + r * dr;
        const float dF = -dot3(dP, dP) + dot3(xp, ddP) + dr * dr + r * ddr;
        if (dF < 0.0f) u = fminf(fmaxf(u - F / dF, 0.0f), 1.0f);
    }
    return u;
}

struct CurveRay {
    f3 o, d;
    float inv_a;  // 1 / (d.d)
    float inv_len;  // 1 / |d|
};

// Ball u against the ray's line: clearance (distance minus radius; <= 0: crossed) and, when
// crossed, the entry parameter t.
SGPU_HD float curve_ball(const CurveSeg& s, const CurveRay& ray, float u, float& t)
{
    f3 P;
    float r;
    curve_eval(s, u, P, r);
    const f3 oc = ray.o - P;
    const float tc = -dot3(oc, ray.d) * ray.inv_a;  // closest approach
    const f3 perp = oc + ray.d * tc;
    const float dist = length3(perp);
    const float cl = dist - r;
    const float h = r * r - dot3(perp, perp);
    t = tc - sqrtf(h > 0.0f ? h : 0.0f) * ray.inv_len;
    return cl;
}

// Lower bound of the entry of the ray into a capsule around chord AB of radius R: the entry into its
// bounding sphere (centre mid-chord, radius |AB| / 2 + R); false when the ray's line misses that sphere.
SGPU_HD bool capsule_entry_bound(const CurveRay& ray, f3 A, f3 B, float R, float& t)
{
    const f3 c = (A + B) * 0.5f;
    const float rs = 0.5f * length3(B - A) + R;
    const f3 oc = ray.o - c;
    const float tc = -dot3(oc, ray.d) * ray.inv_a;
    const f3 perp = oc + ray.d * tc;
    const float h = rs * rs - dot3(perp, perp);
    t = tc - sqrtf(h > 0.0f ? h : 0.0f) * ray.inv_len;
    return h >= 0.0f;
}

// golden-section search over [lo, hi] of the clearance (entry = false) or of the entry t on a run of
// crossed balls (entry = true)
SGPU_HD float curve_golden(const CurveSeg& s, const CurveRay& ray, float lo, float hi, bool entry)
{
    const float g = 0.618033988749895f;
    float x1 = hi - g * (hi - lo), x2 = lo + g * (hi - lo), t1, t2;
    float f1 = curve_ball(s, ray, x1, t1), f2 = curve_ball(s, ray, x2, t2);
    if (entry) {
        f1 = t1;
        f2 = t2;
    }
    for (int i = 0; i < kCurveIters; ++i) {
        if (f1 <= f2) {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = hi - g * (hi - lo);
            f1 = curve_ball(s, ray, x1, t1);
            if (entry) f1 = t1;
        } else {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + g * (hi - lo);
            f2 = curve_ball(s, ray, x2, t2);
            if (entry) f2 = t2;
        }
    }
    return 0.5f * (lo + hi);
}

// First entry of the ray into the segment with t in [tnear, tfar]; ANY: the first found.
template <bool ANY>
SGPU_HD bool curve_intersect(const CurveSeg& s, f3 o, f3 d, float tnear, float tfar, float& t_out, float& u_out,
                             f3& ng)
{
    CurveRay ray;
    ray.o = o;
    ray.d = d;
    const float a = dot3(d, d);
    ray.inv_a = 1.0f / a;
    ray.inv_len = 1.0f / sqrtf(a);
    const float h = 1.0f / kCurvePieces;
    float m2 = 0.0f, m2r = 0.0f;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const f3 e = mk3(s.p[j].x - 2.0f * s.p[j + 1].x + s.p[j + 2].x, s.p[j].y - 2.0f * s.p[j + 1].y + s.p[j + 2].y,
                         s.p[j].z - 2.0f * s.p[j + 1].z + s.p[j + 2].z);
        m2 = fmaxf(m2, length3(e));
        m2r = fmaxf(m2r, fabsf(s.p[j].w - 2.0f * s.p[j + 1].w + s.p[j + 2].w));
    }
    const float dev = h * h * 0.125f * (m2 + m2r);
    float best = tfar, bu = 0.0f;
    bool found = false;
    f3 A;
    float ra;
    curve_eval(s, 0.0f, A, ra);
    {
        // the whole segment first: its balls lie in the capsule around the chord P(0) P(1) whose radius
        // is the larger end radius plus (position and radius second-difference bounds) / 8
        f3 B;
        float rb;
        curve_eval(s, 1.0f, B, rb);
        const float R = fmaxf(ra, rb) + 0.125f * (m2 + m2r);
        const f3 e = B - A, w0 = o - A;
        const float b = dot3(d, e), c = dot3(e, e), dd = dot3(d, w0), ee = dot3(e, w0);
        const float den = a * c - b * b;
        float sc = den > 0.0f ? (a * ee - b * dd) / den : 0.0f;
        sc = fminf(fmaxf(sc, 0.0f), 1.0f);
        const float tc = (sc * b - dd) * ray.inv_a;
        const f3 x = (o + d * tc) - (A + e * sc);
        const float reach = (R + sqrtf(c)) * ray.inv_len;
        if (!(dot3(x, x) <= R * R * 1.0001f && tc + reach >= tnear && tc - reach <= best)) return false;
        float tl;
        if (!capsule_entry_bound(ray, A, B, R * 1.0001f, tl) || tl > best) return false;
    }
    for (int k = 0; k < kCurvePieces; ++k) {
        f3 B;
        float rb;
        curve_eval(s, (float)(k + 1) * h, B, rb);
        const float R = fmaxf(ra, rb) + dev;
        // closest approach of the ray's line and the chord AB
        const f3 e = B - A, w0 = o - A;
        const float b = dot3(d, e), c = dot3(e, e), dd = dot3(d, w0), ee = dot3(e, w0);
        const float den = a * c - b * b;
        float sc = den > 0.0f ? (a * ee - b * dd) / den : 0.0f;
        sc = fminf(fmaxf(sc, 0.0f), 1.0f);
        const float tc = (sc * b - dd) * ray.inv_a;
        const f3 x = (o + d * tc) - (A + e * sc);
        const float reach = (R + sqrtf(c)) * ray.inv_len;
        float tl;
        if (dot3(x, x) <= R * R * 1.0001f && tc + reach >= tnear && tc - reach <= best &&
            capsule_entry_bound(ray, A, B, R * 1.0001f, tl) && tl <= best) {
            const float lo = fmaxf((float)k * h - 0.5f * h, 0.0f);
            const float hi = fminf((float)(k + 1) * h + 0.5f * h, 1.0f);
            const float ug = curve_golden(s, ray, lo, hi, false);
            float tg;
            if (curve_ball(s, ray, ug, tg) <= 0.0f) {
                // the run of crossed balls around ug within [lo, hi]
                float l = lo, r = hi, t;
                if (curve_ball(s, ray, l, t) > 0.0f) {
                    float p = l, q = ug;
                    for (int i = 0; i < kCurveIters; ++i) {
                        const float m = 0.5f * (p + q);
                        if (curve_ball(s, ray, m, t) <= 0.0f) q = m;
                        else p = m;
                    }
                    l = q;
                }
                if (curve_ball(s, ray, r, t) > 0.0f) {
                    float p = ug, q = r;
                    for (int i = 0; i < kCurveIters; ++i) {
                        const float m = 0.5f * (p + q);
                        if (curve_ball(s, ray, m, t) <= 0.0f) p = m;
                        else q = m;
                    }
                    r = p;
                }
                const float u = curve_golden(s, ray, l, r, true);
                if (curve_ball(s, ray, u, t) <= 0.0f && t >= tnear && t <= best) {
                    best = t;
                    bu = u;
                    found = true;
                    if (ANY) break;
                }
            }
        }
        A = B;
        ra = rb;
    }
    if (!found) return false;
    const f3 X = o + d * best;
    bu = curve_contact(s, X, bu);
    t_out = best;
    u_out = bu;
    f3 P;
    float r;
    curve_eval(s, bu, P, r);
    ng = X - P;
    return true;
}

}  // namespace sgpu
