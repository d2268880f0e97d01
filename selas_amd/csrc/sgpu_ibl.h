// Image-based light (equirectangular RGB FP32), the light Selas's VCM emits from:
//   sampling tables  BuildImageBasedLight.cpp:18-103 (CieLuma, CalculateIntensityMap,
//                    CalculateStrataDistributionFunctions): per texel sin(theta) x CIE luma, a
//                    normalized running sum per row (conditional CDFs) and over the row sums
//                    (marginal CDF); built on the host when the light is set
//   sampling         ImageBasedLightResource.cpp:82-104 (SampleDistributionFunction), :118-151 (Ibl)
//   lookup and pdf   ImageBasedLightResource.cpp:154-195 (SampleIbl(wi)), :221-261 (SampleIBlPdf)
//   emission         AreaLighting.cpp:246-290 (EmitIblLightSample: a direction from the tables,
//                    a position on the disc of the scene's bounding sphere facing it)
//   direct sample    AreaLighting.cpp:293-314 (DirectIblLightSample), :317-331 (IblCalculateRadiance)
//   directions       Projection.cpp:109-145 (NormalizedCartesianToSpherical, SphericalToCartesian; y up)
//   disc             Disc.cpp:13-62 (SampleConcentricDisc, ConcentricDiscPdf = 1/pi)
// Where Selas's code is inconsistent, the consistent form is used, so that emission, direct
// sampling and the lookup of a direction agree on its texel and its pdf (the MIS weights of VCM
// need that):
//   * the texel pdf is converted to solid angle with W H / (2 pi^2): d omega = sin(theta) (pi / H)
//     (2 pi / W) per texel; Selas divides by 2 pi only, so its pdf is pi times too large;
//   * the lookup takes the texel that contains the direction, floor(u W / 2 pi) and
//     floor(theta H / pi); Selas subtracts half a texel first, which maps a sampled texel centre to
//     the texel before it about half the time;
//   * the rotation (normalized to [0, 2 pi) when the light is set) is added to phi by the lookup and
//     subtracted by the sampling, which makes one the inverse of the other; Selas adds it in both;
//   * Selas never assigns the scene's bounding sphere (SceneResource.h:78), so emission starts on a
//     disc of radius 0; here it is the sphere around the triangles' bounding box;
//   * a sample whose pdf is 0 carries no light (the reference would divide by it).
#pragma once

#include "sgpu_internal.h"
#include "sgpu_shade.h"

namespace sgpu {

constexpr float kTwoPi = kPi * 2.0f;      // Math::TwoPi_
constexpr float kTwoPiSq = kTwoPi * kPi;  // 2 pi^2

struct IblView {
    const float* rgb;  // height x width x 3 (row y: theta from 0 at +y to pi at -y)
    const float* mdf;  // height: marginal CDF over the rows
    const float* cdf;  // height x width: conditional CDF of each row
    uint32_t w, h;
    float rotation;    // radians in [0, 2 pi), added to phi by the lookup
    float exposure;    // radiance scale (exposureScale)
    float center[3];   // the scene's bounding sphere
    float radius;
};

// SampleDistributionFunction: the first entry of the CDF at or above r (binary search) and the
// probability of that entry. For r < 1 one is always found (the last entry is 1).
SGPU_HD uint32_t ibl_sample_distribution(const float* d, uint32_t count, float r, float& pdf)
{
    int32_t lo = 0, hi = (int32_t)count - 1;
    uint32_t index = count - 1;
    pdf = count > 1 ? d[count - 1] - d[count - 2] : d[0];
    while (lo <= hi) {
        const int32_t mid = (lo + hi) / 2;
        if (d[mid] >= r) {
            index = (uint32_t)mid;
            pdf = mid > 0 ? d[mid] - d[mid - 1] : d[mid];
            hi = mid - 1;
        } else {
            lo = mid + 1;
        }
    }
    return index;
}

SGPU_HD f3 spherical_to_cartesian(float theta, float phi)
{
    const float st = sin_f(theta), ct = cos_f(theta), sp = sin_f(phi), cp = cos_f(phi);
    return mk3(st * cp, ct, st * sp);
}

// solid-angle pdf of drawing texel (x, y), at polar angle theta
SGPU_HD float ibl_texel_pdf(const IblView& b, uint32_t x, uint32_t y, float theta)
{
    const float mdf = y > 0 ? b.mdf[y] - b.mdf[y - 1] : b.mdf[y];
    const float* row = b.cdf + (size_t)y * b.w;
    const float cdf = x > 0 ? row[x] - row[x - 1] : row[x];
    const float inv_jacobian = ((float)b.w * (float)b.h) / kTwoPiSq;
    const float st = sin_f(theta);
    return st > 0.0f ? mdf * cdf * inv_jacobian / st : 0.0f;
}

// SampleIbl(x, y)
SGPU_HD f3 ibl_texel(const IblView& b, uint32_t x, uint32_t y)
{
    const float* t = b.rgb + 3 * ((size_t)y * b.w + x);
    return mk3(t[0], t[1], t[2]) * b.exposure;
}

// Ibl: a texel drawn from the tables with r0 (row) and r1 (column), its direction (theta, phi) and
// its solid-angle pdf
SGPU_HD void ibl_sample(const IblView& b, float r0, float r1, float& theta, float& phi, uint32_t& x, uint32_t& y,
                        float& pdf)
{
    const float wf = (float)b.w, hf = (float)b.h;
    float mdf, cdf;
    y = ibl_sample_distribution(b.mdf, b.h, r0, mdf);
    x = ibl_sample_distribution(b.cdf + (size_t)y * b.w, b.w, r1, cdf);
    theta = ((float)y + 0.5f) * kPi / hf;
    float u = ((float)x + 0.5f) * kTwoPi / wf - b.rotation;
    if (u < 0.0f) u += kTwoPi;
    phi = u - kPi;
    const float inv_jacobian = (wf * hf) / kTwoPiSq;
    const float st = sin_f(theta);
    pdf = st > 0.0f ? mdf * cdf * inv_jacobian / st : 0.0f;
}

// SampleIbl(wi) / SampleIBlPdf: the radiance arriving from direction wi (unit) and the pdf of
// sampling wi
SGPU_HD f3 ibl_lookup(const IblView& b, f3 wi, float& pdf)
{
    const float wf = (float)b.w, hf = (float)b.h;
    const float cy = wi.y > 1.0f ? 1.0f : (wi.y < -1.0f ? -1.0f : wi.y);
    const float theta = acos_f(cy);
    const float phi = fmodf(atan2_f(wi.z, wi.x) + kPi + b.rotation, kTwoPi);
    int32_t xi = (int32_t)floorf(phi * wf / kTwoPi);
    int32_t yi = (int32_t)floorf(theta * hf / kPi);
    xi = xi < 0 ? 0 : (xi >= (int32_t)b.w ? (int32_t)b.w - 1 : xi);
    yi = yi < 0 ? 0 : (yi >= (int32_t)b.h ? (int32_t)b.h - 1 : yi);
    pdf = ibl_texel_pdf(b, (uint32_t)xi, (uint32_t)yi, theta);
    return ibl_texel(b, (uint32_t)xi, (uint32_t)yi);
}

// SampleConcentricDisc (Shirley and Chiu, "A Low Distortion Map Between Disk and Square")
SGPU_HD void concentric_disc(float r0, float r1, float& dx, float& dy)
{
    float phi, r;
    const float a = 2.0f * r0 - 1.0f, b = 2.0f * r1 - 1.0f;
    if (a > -b) {
        if (a > b) {
            r = a;
            phi = (kPi / 4.0f) * (b / a);
        } else {
            r = b;
            phi = (kPi / 4.0f) * (2.0f - (a / b));
        }
    } else {
        if (a < b) {
            r = -a;
            phi = (kPi / 4.0f) * (4.0f + (b / a));
        } else {
            r = -b;
            phi = absf_(b) > 0.0f ? (kPi / 4.0f) * (6.0f - (a / b)) : 0.0f;
        }
    }
    dx = r * cos_f(phi);
    dy = r * sin_f(phi);
}

// area pdf of the emission disc (ConcentricDiscPdf / R^2)
SGPU_HD float ibl_position_pdf(const IblView& b) { return kInvPi * (1.0f / (b.radius * b.radius)); }

// EmitIblLightSample: origin, direction (into the scene), radiance and pdfs of an emitted ray.
// direction_pdf is the solid-angle pdf of the direction (Selas's directionPdfA).
SGPU_HD void ibl_emit(const IblView& b, float r0, float r1, float r2, float r3, f3& position, f3& direction,
                      f3& radiance, float& emission_pdf_w, float& direction_pdf)
{
    float theta, phi, pdf;
    uint32_t x, y;
    ibl_sample(b, r0, r1, theta, phi, x, y, pdf);
    const f3 to_ibl = spherical_to_cartesian(theta, phi);
    radiance = ibl_texel(b, x, y);
    f3 ex, ez;
    make_frame(to_ibl, ex, ez);
    float ux, uy;
    concentric_disc(r2, r3, ux, uy);
    position = mk3(b.center[0], b.center[1], b.center[2]) + ((to_ibl + ex * ux) + ez * uy) * b.radius;
    direction = -to_ibl;
    direction_pdf = pdf;
    emission_pdf_w = ibl_position_pdf(b) * pdf;
}

// DirectIblLightSample: a direction toward the light, its radiance and solid-angle pdf
SGPU_HD f3 ibl_direct(const IblView& b, float r0, float r1, f3& direction, float& pdf)
{
    float theta, phi;
    uint32_t x, y;
    ibl_sample(b, r0, r1, theta, phi, x, y, pdf);
    direction = spherical_to_cartesian(theta, phi);
    return ibl_texel(b, x, y);
}

// The light of a committed scene as the kernels read it; false when the scene has none.
inline bool ibl_view(const sgpu_scene_s* sc, IblView& v)
{
    v = IblView{};
    if (!sc->d_ibl) return false;
    const size_t n = (size_t)sc->ibl_w * sc->ibl_h;
    v.rgb = sc->d_ibl;
    v.mdf = sc->d_ibl + 3 * n;
    v.cdf = sc->d_ibl + 3 * n + sc->ibl_h;
    v.w = sc->ibl_w;
    v.h = sc->ibl_h;
    v.rotation = sc->ibl_rotation;
    v.exposure = sc->ibl_exposure;
    for (int k = 0; k < 3; ++k) v.center[k] = sc->bsphere[k];
    v.radius = sc->bsphere[3];
    return true;
}

}  // namespace sgpu
