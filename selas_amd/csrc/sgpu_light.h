// Light sampling on the device: quad lights sampled by solid angle (Selas's only direct-lighting
// path, Source/Core/Shading/AreaLighting.cpp:113-153,353-391) and point lights (an addition for
// scenes that ask for one; Selas has none).
#pragma once

#include "sgpu_shade.h"

namespace sgpu {

// SceneLight (Source/Core/SceneLib/SceneResource.h:32-40); same layout as sgpu_light in the C ABI.
struct Light {
    uint32_t type;  // 0 quad, 1 point
    float position[3];
    float direction[3];
    float x[3];
    float z[3];
    float radiance[3];
};
static_assert(sizeof(Light) == 64, "Light must be 64 bytes");

SGPU_HD f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

// Spherical-rectangle sampler (RectangulerLightSampler.cpp:17-98; Urena et al. 2013).
struct RectSampler {
    f3 o, x, y, z;
    float z0, z0sq, x0, y0, y0sq, x1, y1, y1sq, b0, b1, b0sq, k, S;
};

SGPU_HD void rect_init(f3 s, f3 ex, f3 ey, f3 o, RectSampler& r)
{
    r.o = o;
    float exl = length3(ex), eyl = length3(ey);
    r.x = ex * (1.0f / exl);
    r.y = ey * (1.0f / eyl);
    r.z = cross3(r.x, r.y);
    f3 d = s - o;
    r.z0 = dot3(d, r.z);
    if (r.z0 > 0) {
        r.z = -r.z;
        r.z0 *= -1;
    }
    r.z0sq = r.z0 * r.z0;
    r.x0 = dot3(d, r.x);
    r.y0 = dot3(d, r.y);
    r.x1 = r.x0 + exl;
    r.y1 = r.y0 + eyl;
    r.y0sq = r.y0 * r.y0;
    r.y1sq = r.y1 * r.y1;
    f3 v00 = mk3(r.x0, r.y0, r.z0), v01 = mk3(r.x0, r.y1, r.z0), v10 = mk3(r.x1, r.y0, r.z0),
       v11 = mk3(r.x1, r.y1, r.z0);
    f3 n0 = normalize3(cross3(v00, v10)), n1 = normalize3(cross3(v10, v11));
    f3 n2 = normalize3(cross3(v11, v01)), n3 = normalize3(cross3(v01, v00));
    float g0 = acos_f(-dot3(n0, n1)), g1 = acos_f(-dot3(n1, n2)), g2 = acos_f(-dot3(n2, n3)), g3 = acos_f(-dot3(n3, n0));
    r.b0 = n0.z;
    r.b1 = n2.z;
    r.b0sq = r.b0 * r.b0;
    r.k = 2 * kPi - g2 - g3;
    r.S = g0 + g1 - r.k;
}

SGPU_HD f3 rect_sample(const RectSampler& r, float u, float v)
{
    float au = u * r.S + r.k;
    float fu = (cos_f(au) * r.b0 - r.b1) / sin_f(au);
    float cu = 1 / sqrtf(fu * fu + r.b0sq) * (fu > 0 ? +1 : -1);
    cu = clamp_(cu, -1.0f, 1.0f);
    float xu = -(cu * r.z0) / sqrtf(1 - cu * cu);
    xu = clamp_(xu, r.x0, r.x1);
    float d = sqrtf(xu * xu + r.z0sq);
    float h0 = r.y0 / sqrtf(d * d + r.y0sq);
    float h1 = r.y1 / sqrtf(d * d + r.y1sq);
    float hv = h0 + v * (h1 - h0), hv2 = hv * hv;
    float yv = (hv2 < 1 - 1e-10f) ? (hv * d) / sqrtf(1 - hv2) : r.y1;
    return r.o + xu * r.x + yv * r.y + r.z0 * r.z;
}

struct LightSample {
    uint32_t index;
    f3 direction;
    f3 radiance;
    float distance;
    float pdf_w;
};

// SampleRectangleLightSolidAngle (AreaLighting.cpp:113-153). Selas's estimator multiplies the
// radiance by Dot(normal, l) although the BSDF value it is combined with already carries
// |cos theta_l|; `cosine_once` drops that extra factor (the physically consistent form).
template <class U>
SGPU_HD void sample_quad_light(U& u, const Light& L, f3 position, f3 normal, bool cosine_once, LightSample& out)
{
    f3 ex = ld3(L.x), ez = ld3(L.z);
    f3 s = ld3(L.position) - 0.5f * ex - 0.5f * ez;
    RectSampler qs;
    rect_init(s, ex, ez, position, qs);
    float pdf = 1.0f / qs.S;
    float uu = u.next();
    float vv = u.next();
    f3 lp = rect_sample(qs, uu, vv);
    f3 ul = lp - position;
    float dist_sq = dot3(ul, ul);
    float dist = sqrtf(dist_sq);
    f3 l = (1.0f / dist) * ul;
    float area_measure = dot3(ld3(L.direction), -l) / dist_sq;
    if (area_measure > 0) {
        out.direction = l;
        out.distance = dist;
        out.radiance = cosine_once ? ld3(L.radiance) * area_measure : ld3(L.radiance) * area_measure * dot3(normal, l);
        out.pdf_w = pdf * area_measure;
    } else {
        out.direction = mk3(0, 0, 0);
        out.distance = 0.0f;
        out.radiance = mk3(0, 0, 0);
        out.pdf_w = 1.0f;
    }
}

// Point light: a delta in position; radiance = intensity / d^2, pdf 1 (before the light pick).
SGPU_HD void sample_point_light(const Light& L, f3 position, LightSample& out)
{
    f3 ul = ld3(L.position) - position;
    float dist_sq = dot3(ul, ul);
    float dist = sqrtf(dist_sq);
    out.direction = (1.0f / dist) * ul;
    out.distance = dist;
    out.radiance = ld3(L.radiance) * (1.0f / dist_sq);
    out.pdf_w = 1.0f;
}

// NextEventEstimation (AreaLighting.cpp:353-374): uniform choice among the lights, then the
// light's own sampling. Consumes 1 uniform for the pick and 2 for a quad light.
template <class U>
SGPU_HD void next_event_estimation(U& u, const Light* __restrict lights, uint32_t n_lights, f3 position, f3 normal,
                                   bool cosine_once, LightSample& out)
{
    out.radiance = mk3(0, 0, 0);
    out.pdf_w = 1.0f;
    out.distance = 0.0f;
    out.direction = mk3(0, 0, 0);
    out.index = 0;
    if (n_lights == 0) return;
    float per_light = 1.0f / n_lights;
    float p0 = u.next();
    uint32_t idx = (uint32_t)(p0 * n_lights);
    if (idx >= n_lights) idx = n_lights - 1;
    const Light L = lights[idx];
    if (L.type == 1)
        sample_point_light(L, position, out);
    else
        sample_quad_light(u, L, position, normal, cosine_once, out);
    out.pdf_w *= per_light;
    out.index = idx;
}

}  // namespace sgpu
