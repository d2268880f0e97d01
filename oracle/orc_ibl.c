/*
 * ORACLE -- test infrastructure only (see orc_trace.c for the rules). Compiled with
 * -ffp-contract=off.
 *
 * CPU restatement of Selas's image-based light (an equirectangular RGB map, y up):
 *   CieLuma, CalculateIntensityMap,
 *   CalculateStrataDistributionFunctions     BuildImageBasedLight.cpp:18-103
 *   SampleDistributionFunction               ImageBasedLightResource.cpp:82-104
 *   Ibl                                      ImageBasedLightResource.cpp:118-151
 *   SampleIbl(wi), SampleIBlPdf              ImageBasedLightResource.cpp:154-195, 221-261
 *   SampleIbl(x, y)                          ImageBasedLightResource.cpp:264-267
 *   EmitIblLightSample                       AreaLighting.cpp:246-290
 *   DirectIblLightSample                     AreaLighting.cpp:293-314
 *   ConcentricDiscPdf                        Disc.cpp:59-62 (SampleConcentricDisc is orc_shade.c's,
 *                                            pinned by the reference's own vectors)
 *   SphericalToCartesian                     Projection.cpp:132-145
 * Selas's lookup and sampling disagree with each other, so the consistent forms are restated
 * (the device code makes the same choices):
 *   - solid-angle pdf = texel probability x W H / (2 pi^2 sin theta) (Selas: W H / (2 pi), pi times
 *     too large; an equirectangular texel spans sin(theta) (pi / H) (2 pi / W) steradians);
 *   - the lookup takes the texel containing the direction, floor(u W / 2 pi), floor(theta H / pi)
 *     (Selas subtracts half a texel, mapping a sampled texel centre to the previous texel about
 *     half the time);
 *   - the lookup adds the rotation and the sampling subtracts it (Selas adds it in both);
 *   - the emission disc belongs to the sphere around the triangles' bounding box (Selas's
 *     boundingSphere is never assigned, SceneResource.h:78);
 *   - the lookup clamps wi.y to [-1, 1] before acos.
 * Pinned by: tables vs closed forms (uniform map: marginal CDF = normalized running sum of
 * sin(theta_y), rows linear); sample -> lookup round trips (same texel, pdf within float rounding);
 * the pdf integrating to 1 over the sphere; sampling frequencies vs the pdf (chi-square); the disc
 * map's area preservation; and, through the integrators, the closed form of a Lambert plane under
 * a uniform sky (tests/test_oracle_ibl.py).
 */
#include "oracle.h"
#include "orc_fn.h"

#include <math.h>
#include <string.h>

static const float kPi = 3.141592653589793f;
#define TWO_PI (kPi * 2.0f)

/* BuildImageBasedLight.cpp:18-103: rows summed left to right, rows' sums top to bottom */
void orc_ibl_tables(uint32_t w, uint32_t h, const float* rgb, float* mdf, float* cdf)
{
    const float hf = (float)h;
    memset(mdf, 0, sizeof(float) * h);
    memset(cdf, 0, sizeof(float) * (size_t)w * h);
    float marginal_sum = 0.0f;
    for (uint32_t y = 0; y < h; ++y) {
        const float theta = ((float)y + 0.5f) * kPi / hf;
        const float st = sin_f(theta);
        float conditional_sum = 0.0f;
        float* c = cdf + (size_t)y * w;
        /* the intensity map is kept in the row of cdf until the row is normalized */
        for (uint32_t x = 0; x < w; ++x) {
            const float* t = rgb + 3 * ((size_t)y * w + x);
            c[x] = st * (t[0] * 0.299f + t[1] * 0.587f + t[2] * 0.114f);
            conditional_sum += c[x];
        }
        mdf[y] = conditional_sum;
        marginal_sum += conditional_sum;
        if (conditional_sum > 0.0f) {
            const float oo = 1.0f / conditional_sum;
            float integration = 0.0f;
            for (uint32_t x = 0; x < w; ++x) {
                integration += oo * c[x];
                c[x] = integration;
            }
        } else {
            memset(c, 0, sizeof(float) * w);
        }
        c[w - 1] = 1.0f;
    }
    if (marginal_sum > 0.0f) {
        const float oo = 1.0f / marginal_sum;
        float integration = 0.0f;
        for (uint32_t y = 0; y < h; ++y) {
            integration += mdf[y] * oo;
            mdf[y] = integration;
        }
    }
    mdf[h - 1] = 1.0f;
}

float orc_ibl_rotation(float radians)
{
    float r = fmodf(radians, TWO_PI);
    if (r < 0.0f) r += TWO_PI;
    if (r >= TWO_PI) r = 0.0f;
    return r;
}

void orc_bounding_sphere(const float* tri9, uint32_t n_tris, float* out4)
{
    float lo[3], hi[3];
    out4[0] = out4[1] = out4[2] = out4[3] = 0.0f;
    if (n_tris == 0) return;
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = tri9[k];
    for (size_t i = 0; i < (size_t)n_tris * 9; i += 3)
        for (int k = 0; k < 3; ++k) {
            if (tri9[i + k] < lo[k]) lo[k] = tri9[i + k];
            if (tri9[i + k] > hi[k]) hi[k] = tri9[i + k];
        }
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    for (int k = 0; k < 3; ++k) out4[k] = (lo[k] + hi[k]) * 0.5f;
    out4[3] = 0.5f * sqrtf(dx * dx + dy * dy + dz * dz);
}

/* SampleDistributionFunction (ImageBasedLightResource.cpp:82-104) */
uint32_t orc_ibl_sample_distribution(const float* d, uint32_t count, float r, float* pdf)
{
    int32_t lo = 0, hi = (int32_t)count - 1;
    uint32_t index = count - 1;
    *pdf = count > 1 ? d[count - 1] - d[count - 2] : d[0];
    while (lo <= hi) {
        const int32_t mid = (lo + hi) / 2;
        if (d[mid] >= r) {
            index = (uint32_t)mid;
            *pdf = mid > 0 ? d[mid] - d[mid - 1] : d[mid];
            hi = mid - 1;
        } else {
            lo = mid + 1;
        }
    }
    return index;
}

static float texel_pdf(const orc_ibl* b, uint32_t x, uint32_t y, float theta)
{
    const float mdf = y > 0 ? b->mdf[y] - b->mdf[y - 1] : b->mdf[y];
    const float* row = b->cdf + (size_t)y * b->w;
    const float cdf = x > 0 ? row[x] - row[x - 1] : row[x];
    const float inv_jacobian = ((float)b->w * (float)b->h) / (TWO_PI * kPi);
    const float st = sin_f(theta);
    return st > 0.0f ? mdf * cdf * inv_jacobian / st : 0.0f;
}

/* Ibl (ImageBasedLightResource.cpp:118-151) */
void orc_ibl_sample(const orc_ibl* b, float r0, float r1, float* out6)
{
    const float wf = (float)b->w, hf = (float)b->h;
    float mdf, cdf;
    const uint32_t y = orc_ibl_sample_distribution(b->mdf, b->h, r0, &mdf);
    const uint32_t x = orc_ibl_sample_distribution(b->cdf + (size_t)y * b->w, b->w, r1, &cdf);
    const float theta = ((float)y + 0.5f) * kPi / hf;
    float u = ((float)x + 0.5f) * TWO_PI / wf - b->rotation;
    if (u < 0.0f) u += TWO_PI;
    const float phi = u - kPi;
    const float inv_jacobian = (wf * hf) / (TWO_PI * kPi);
    const float st = sin_f(theta);
    out6[0] = theta;
    out6[1] = phi;
    out6[2] = (float)x;
    out6[3] = (float)y;
    out6[4] = st > 0.0f ? mdf * cdf * inv_jacobian / st : 0.0f;
    out6[5] = 0.0f;
}

static void texel(const orc_ibl* b, uint32_t x, uint32_t y, float* rgb_out)
{
    const float* t = b->rgb + 3 * ((size_t)y * b->w + x);
    for (int k = 0; k < 3; ++k) rgb_out[k] = t[k] * b->exposure;
}

/* SampleIbl(wi) + SampleIBlPdf (ImageBasedLightResource.cpp:154-195, 221-261) */
float orc_ibl_lookup(const orc_ibl* b, const float* wi, float* rgb_out)
{
    const float wf = (float)b->w, hf = (float)b->h;
    const float cy = wi[1] > 1.0f ? 1.0f : (wi[1] < -1.0f ? -1.0f : wi[1]);
    const float theta = acos_f(cy);
    const float phi = fmodf(atan2_f(wi[2], wi[0]) + kPi + b->rotation, TWO_PI);
    int32_t xi = (int32_t)floorf(phi * wf / TWO_PI);
    int32_t yi = (int32_t)floorf(theta * hf / kPi);
    xi = xi < 0 ? 0 : (xi >= (int32_t)b->w ? (int32_t)b->w - 1 : xi);
    yi = yi < 0 ? 0 : (yi >= (int32_t)b->h ? (int32_t)b->h - 1 : yi);
    texel(b, (uint32_t)xi, (uint32_t)yi, rgb_out);
    return texel_pdf(b, (uint32_t)xi, (uint32_t)yi, theta);
}

/* ConcentricDiscPdf (Disc.cpp:59-62) / R^2 */
float orc_ibl_position_pdf(const orc_ibl* b) { return (1.0f / kPi) * (1.0f / (b->radius * b->radius)); }

static void spherical_to_cartesian(float theta, float phi, float* v)
{
    const float st = sin_f(theta), ct = cos_f(theta), sp = sin_f(phi), cp = cos_f(phi);
    v[0] = st * cp;
    v[1] = ct;
    v[2] = st * sp;
}

/* EmitIblLightSample (AreaLighting.cpp:246-290) */
void orc_ibl_emit(const orc_ibl* b, float r0, float r1, float r2, float r3, float* out14)
{
    float s[6];
    orc_ibl_sample(b, r0, r1, s);
    float to_ibl[3], ex[3], ez[3], disc[2];
    spherical_to_cartesian(s[0], s[1], to_ibl);
    texel(b, (uint32_t)s[2], (uint32_t)s[3], out14 + 6);
    orc_make_frame(to_ibl, ex, ez);
    orc_concentric_disc(r2, r3, disc);
    for (int k = 0; k < 3; ++k) {
        out14[k] = b->center[k] + ((to_ibl[k] + ex[k] * disc[0]) + ez[k] * disc[1]) * b->radius;
        out14[3 + k] = -to_ibl[k];
    }
    out14[9] = orc_ibl_position_pdf(b) * s[4];
    out14[10] = s[4];
    out14[11] = out14[12] = out14[13] = 0.0f;
}

/* DirectIblLightSample (AreaLighting.cpp:293-314) */
void orc_ibl_direct(const orc_ibl* b, float r0, float r1, float* out8)
{
    float s[6];
    orc_ibl_sample(b, r0, r1, s);
    spherical_to_cartesian(s[0], s[1], out8);
    texel(b, (uint32_t)s[2], (uint32_t)s[3], out8 + 3);
    out8[6] = s[4];
    out8[7] = 0.0f;
}
