// Surface shading on the device: Disney BSDF (eval + sample), Lambert, Fresnel and GGX helpers,
// quad-light sampling, concentric disc. Header-only; every function is __host__ __device__.
//
// Semantics follow Selas:
//   Fresnel          Source/Core/Shading/Fresnel.cpp:19-82
//   GGX / Smith      Source/Core/Shading/Ggx.cpp:19-155
//   Disney BSDF      Source/Core/Shading/Disney.cpp:26-659
//   dispatch         Source/Core/Shading/SurfaceScattering.cpp:18-60
//   tangent frame    Source/Core/GeometryLib/CoordinateSystem.h:14-21 (y-up tangent space)
//   quad lights      Source/Core/Shading/AreaLighting.cpp:49-66,113-153,353-391 with the
//                    spherical-rectangle sampler of Source/Core/GeometryLib/RectangulerLightSampler.cpp
//                    (Urena et al. 2013, "An Area-Preserving Parametrization for Spherical Rectangles")
//   RayQuad          Source/Core/MathLib/GeometryIntersection.cpp:183-236 (Lagae & Dutre 2005)
//   disc             Source/Core/GeometryLib/Disc.cpp:13-56 (Shirley & Chiu 1997)
// Values that Selas computes in a particular order are computed in the same order here, so that
// results agree with the reference to the accuracy of the device's libm. Reference quirks that
// change values are kept: the specular pdf of EvaluateDisney is divided by 4|wo.wm| twice;
// the scalar Schlick is lerp(1, w, r0); `float / float3` in CalculateExtinction returns the vector
// divided by the scalar. Lambert is the corrected form (pdf cos/pi, sample weight = albedo): Selas's
// own Lambert (Lambert.cpp) is unreachable in the reference and lacks the 1/pi in its pdf.
#pragma once

#include "sgpu_math.h"

namespace sgpu {

constexpr float kPi = 3.141592653589793f;
constexpr float kInvPi = 1.0f / 3.141592653589793f;

enum Shader : uint32_t { kDisneyThin = 0, kDisneySolid = 1, kLambert = 2 };
enum EventFlags : uint32_t { kScatter = 1, kTransmission = 2, kDirac = 4 };

// SurfaceParameters (Source/Core/Shading/SurfaceParameters.h:18-49): position, frame, material.
struct Surface {
    f3 position;
    f3 n, t, b;  // shading frame; tangent space is y-up (n = local y)
    float error;
    f3 base_color;
    f3 transmittance_color;
    float sheen, sheen_tint, clearcoat, clearcoat_gloss, metallic, spec_trans, diff_trans, flatness;
    float anisotropic, relative_ior, specular_tint, roughness, scatter_distance, ior;
    uint32_t shader;
    uint32_t flags;
};

struct BsdfSample {
    uint32_t flags;
    uint32_t medium;  // 0 vacuum, 1 isotropic
    f3 reflectance;
    f3 wi;
    f3 extinction;
    float fwd_pdf, rev_pdf;
};

// ---- small helpers (Selas's Absf/Max/Min/Clamp/Saturate/Lerp semantics) ------------------------
SGPU_HD float absf_(float x) { return (x < 0) ? -x : x; }
SGPU_HD float max_(float a, float b) { return (a > b) ? a : b; }
SGPU_HD float min_(float a, float b) { return (a < b) ? a : b; }
SGPU_HD float clamp_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
SGPU_HD float saturate_(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }
SGPU_HD float sq(float x) { return x * x; }
SGPU_HD float lerp_(float a, float b, float t) { return (1.0f - t) * a + t * b; }
SGPU_HD f3 lerp3(f3 a, f3 b, float t) { return a * (1.0f - t) + b * t; }
SGPU_HD float sign_(float x) { return x < 0.0f ? -1.0f : (x > 0.0f ? 1.0f : 0.0f); }
SGPU_HD f3 sqrt3(f3 a) { return mk3(sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)); }

SGPU_HD float cos_theta(f3 w) { return w.y; }
SGPU_HD float cos2_theta(f3 w) { return w.y * w.y; }
SGPU_HD float abs_cos_theta(f3 w) { return absf_(w.y); }
SGPU_HD float sin2_theta(f3 w) { return max_(0.0f, 1.0f - cos2_theta(w)); }
SGPU_HD float sin_theta(f3 w) { return sqrtf(sin2_theta(w)); }
SGPU_HD float tan_theta(f3 w) { return sin_theta(w) / cos_theta(w); }
SGPU_HD float cos_phi(f3 w)
{
    float s = sin_theta(w);
    return (s == 0) ? 1.0f : clamp_(w.x / s, -1.0f, 1.0f);
}
SGPU_HD float sin_phi(f3 w)
{
    float s = sin_theta(w);
    return (s == 0) ? 1.0f : clamp_(w.z / s, -1.0f, 1.0f);
}
SGPU_HD float cos2_phi(f3 w)
{
    float c = cos_phi(w);
    return c * c;
}
SGPU_HD float sin2_phi(f3 w)
{
    float s = sin_phi(w);
    return s * s;
}
SGPU_HD f3 reflect_(f3 n, f3 l) { return (2.0f * dot3(n, l)) * n - l; }
SGPU_HD bool transmit_(f3 wm, f3 wi, float n, f3& wo)
{
    float c = dot3(wi, wm);
    if (c < 0.0f) {
        c = -c;
        wm = -wm;
    }
    float root = 1.0f - n * n * (1.0f - c * c);
    if (root <= 0) return false;
    wo = (n * c - sqrtf(root)) * wm - n * wi;
    return true;
}

// MakeOrthogonalCoordinateSystem (CoordinateSystem.h:14-21)
SGPU_HD void make_frame(f3 n, f3& t, f3& b)
{
    if (absf_(n.x) > absf_(n.y))
        t = mk3(-n.z, 0, n.x) * (1.0f / sqrtf(n.x * n.x + n.z * n.z));
    else
        t = mk3(0, n.z, -n.y) * (1.0f / sqrtf(n.y * n.y + n.z * n.z));
    b = cross3(n, t);
}

SGPU_HD f3 to_local(const Surface& s, f3 v) { return mk3(dot3(v, s.t), dot3(v, s.n), dot3(v, s.b)); }
SGPU_HD f3 to_world(const Surface& s, f3 w)
{
    return mk3(w.x * s.t.x + w.y * s.n.x + w.z * s.b.x, w.x * s.t.y + w.y * s.n.y + w.z * s.b.y,
               w.x * s.t.z + w.y * s.n.z + w.z * s.b.z);
}

// ---- Fresnel -----------------------------------------------------------------------------------
SGPU_HD float schlick_weight(float u)
{
    float m = saturate_(1.0f - u);
    float m2 = m * m;
    return m * m2 * m2;
}
SGPU_HD float schlick_scalar(float r0, float c) { return lerp_(1.0f, schlick_weight(c), r0); }
SGPU_HD f3 schlick3(f3 r0, float c)
{
    float e = pow5_f(1.0f - c);
    return r0 + (mk3(1.0f, 1.0f, 1.0f) - r0) * e;
}
SGPU_HD float schlick_r0(float eta) { return sq(eta - 1.0f) / sq(eta + 1.0f); }
SGPU_HD float fresnel_dielectric(float cos_i, float ni, float nt)
{
    cos_i = clamp_(cos_i, -1.0f, 1.0f);
    if (cos_i < 0.0f) {
        float tmp = ni;
        ni = nt;
        nt = tmp;
        cos_i = -cos_i;
    }
    float sin_i = sqrtf(max_(0.0f, 1.0f - cos_i * cos_i));
    float sin_t = ni / nt * sin_i;
    if (sin_t >= 1) return 1;
    float cos_t = sqrtf(max_(0.0f, 1.0f - sin_t * sin_t));
    float rpar = ((nt * cos_i) - (ni * cos_t)) / ((nt * cos_i) + (ni * cos_t));
    float rperp = ((ni * cos_i) - (nt * cos_t)) / ((ni * cos_i) + (nt * cos_t));
    return (rpar * rpar + rperp * rperp) / 2;
}

// ---- GGX ---------------------------------------------------------------------------------------
SGPU_HD float g1_aniso(f3 w, float ax, float ay)
{
    float abs_tan = absf_(tan_theta(w));
    if (isinf(abs_tan)) return 0.0f;
    float a = sqrtf(cos2_phi(w) * ax * ax + sin2_phi(w) * ay * ay);
    float a2tan2 = sq(a * abs_tan);
    float lambda = 0.5f * (-1.0f + sqrtf(1.0f + a2tan2));
    return 1.0f / (1.0f + lambda);
}
SGPU_HD float g1_iso(f3 w, float a)
{
    float a2 = a * a;
    float c = abs_cos_theta(w);
    return 2.0f / (1.0f + sqrtf(a2 + (1 - a2) * c * c));
}
SGPU_HD float d_aniso(f3 wm, float ax, float ay)
{
    float hx2 = sq(wm.x), hy2 = sq(wm.z), c2 = cos2_theta(wm);
    float ax2 = sq(ax), ay2 = sq(ay);
    return 1.0f / (kPi * ax * ay * sq(hx2 / ax2 + hy2 / ay2 + c2));
}
// Heitz 2018, "Sampling the GGX Distribution of Visible Normals"
SGPU_HD f3 sample_vndf(f3 wo, float ax, float ay, float u1, float u2)
{
    f3 v = normalize3(mk3(wo.x * ax, wo.y, wo.z * ay));
    f3 t1 = (v.y < 0.9999f) ? normalize3(cross3(v, mk3(0, 1, 0))) : mk3(1, 0, 0);
    f3 t2 = cross3(t1, v);
    float a = 1.0f / (1.0f + v.y);
    float r = sqrtf(u1);
    float phi = (u2 < a) ? (u2 / a) * kPi : kPi + (u2 - a) / (1.0f - a) * kPi;
    float p1 = r * cos_f(phi);
    float p2 = r * sin_f(phi) * ((u2 < a) ? 1.0f : v.y);
    f3 n = p1 * t1 + p2 * t2 + sqrtf(max_(0.0f, 1.0f - p1 * p1 - p2 * p2)) * v;
    return normalize3(mk3(ax * n.x, n.y, ay * n.z));
}
SGPU_HD void vndf_pdf(f3 wi, f3 wm, f3 wo, float ax, float ay, float& fwd, float& rev)
{
    float d = d_aniso(wm, ax, ay);
    float nl = abs_cos_theta(wi), hl = absf_(dot3(wm, wi));
    float g1v = g1_aniso(wo, ax, ay);
    fwd = g1v * hl * d / nl;
    float nv = abs_cos_theta(wo), hv = absf_(dot3(wm, wo));
    float g1l = g1_aniso(wi, ax, ay);
    rev = g1l * hv * d / nv;
}

// ---- Disney ------------------------------------------------------------------------------------
SGPU_HD void lobe_pdfs(const Surface& s, float& p_spec, float& p_diff, float& p_cc, float& p_trans)
{
    float metallic_brdf = s.metallic;
    float specular_bsdf = (1.0f - s.metallic) * s.spec_trans;
    float dielectric_brdf = (1.0f - s.spec_trans) * (1.0f - s.metallic);
    float spec_w = metallic_brdf + dielectric_brdf;
    float trans_w = specular_bsdf;
    float diff_w = dielectric_brdf;
    float cc_w = 1.0f * saturate_(s.clearcoat);
    float norm = 1.0f / (spec_w + trans_w + diff_w + cc_w);
    p_spec = spec_w * norm;
    p_trans = trans_w * norm;
    p_diff = diff_w * norm;
    p_cc = cc_w * norm;
}
SGPU_HD float thin_trans_roughness(float ior, float roughness) { return saturate_((0.65f * ior - 0.35f) * roughness); }
SGPU_HD void aniso_params(float roughness, float anisotropic, float& ax, float& ay)
{
    float aspect = sqrtf(1.0f - 0.9f * anisotropic);
    ax = max_(0.001f, sq(roughness) / aspect);
    ay = max_(0.001f, sq(roughness) * aspect);
}
SGPU_HD f3 tint_of(f3 c)
{
    float lum = dot3(mk3(0.3f, 0.6f, 1.0f), c);
    return (lum > 0.0f) ? c * (1.0f / lum) : mk3(1, 1, 1);
}
SGPU_HD float gtr1(float abs_hl, float a)
{
    if (a >= 1) return kInvPi;
    float a2 = a * a;
    return (a2 - 1.0f) / (kPi * log2f(a2) * (1.0f + (a2 - 1.0f) * abs_hl * abs_hl));
}
SGPU_HD float eval_clearcoat(float clearcoat, float alpha, f3 wo, f3 wm, f3 wi, float& fpdf, float& rpdf)
{
    if (clearcoat <= 0.0f) return 0.0f;
    float nh = abs_cos_theta(wm);
    float hl = dot3(wm, wi);
    float d = gtr1(nh, lerp_(0.1f, 0.001f, alpha));
    float f = schlick_scalar(0.04f, hl);
    float gl = g1_iso(wi, 0.25f), gv = g1_iso(wo, 0.25f);
    fpdf = d / (4.0f * absf_(dot3(wo, wm)));
    rpdf = d / (4.0f * absf_(dot3(wi, wm)));
    return 0.25f * clearcoat * d * f * gl * gv;
}
SGPU_HD f3 eval_sheen(const Surface& s, f3 wm, f3 wi)
{
    if (s.sheen <= 0.0f) return mk3(0, 0, 0);
    float hl = absf_(dot3(wm, wi));
    f3 tint = tint_of(s.base_color);
    return (lerp3(mk3(1, 1, 1), tint, s.sheen_tint) * s.sheen) * schlick_weight(hl);
}
SGPU_HD f3 disney_fresnel(const Surface& s, f3 wo, f3 wm, f3 wi)
{
    float hv = dot3(wm, wo);
    f3 tint = tint_of(s.base_color);
    f3 r0 = lerp3(mk3(1, 1, 1), tint, s.specular_tint) * schlick_r0(s.relative_ior);
    r0 = lerp3(r0, s.base_color, s.metallic);
    float df = fresnel_dielectric(hv, 1.0f, s.ior);
    f3 mf = schlick3(r0, dot3(wi, wm));
    return lerp3(mk3(df, df, df), mf, s.metallic);
}
SGPU_HD f3 eval_brdf(const Surface& s, f3 wo, f3 wm, f3 wi, float& fpdf, float& rpdf)
{
    fpdf = 0.0f;
    rpdf = 0.0f;
    float nl = cos_theta(wi), nv = cos_theta(wo);
    if (nl <= 0.0f || nv <= 0.0f) return mk3(0, 0, 0);
    float ax, ay;
    aniso_params(s.roughness, s.anisotropic, ax, ay);
    float d = d_aniso(wm, ax, ay);
    float gl = g1_aniso(wi, ax, ay), gv = g1_aniso(wo, ax, ay);
    f3 f = disney_fresnel(s, wo, wm, wi);
    vndf_pdf(wi, wm, wo, ax, ay, fpdf, rpdf);
    fpdf *= (1.0f / (4 * absf_(dot3(wo, wm))));
    rpdf *= (1.0f / (4 * absf_(dot3(wi, wm))));
    float k = d * gl * gv;
    float den = 4.0f * nl * nv;
    return mk3(k * f.x / den, k * f.y / den, k * f.z / den);
}
SGPU_HD f3 eval_spec_trans(const Surface& s, f3 wo, f3 wm, f3 wi, float ax, float ay, bool thin)
{
    float rel = s.relative_ior;
    float n2 = rel * rel;
    float nl = abs_cos_theta(wi), nv = abs_cos_theta(wo);
    float hl = dot3(wm, wi), hv = dot3(wm, wo);
    float ahl = absf_(hl), ahv = absf_(hv);
    float d = d_aniso(wm, ax, ay);
    float gl = g1_aniso(wi, ax, ay), gv = g1_aniso(wo, ax, ay);
    float f = fresnel_dielectric(hv, 1.0f, s.ior);
    f3 color = thin ? sqrt3(s.base_color) : s.base_color;
    float c = (ahl * ahv) / (nl * nv);
    float t = (n2 / sq(hl + rel * hv));
    return color * c * t * (1.0f - f) * gl * gv * d;
}
SGPU_HD float eval_retro(const Surface& s, f3 wo, f3 wi)
{
    float nl = abs_cos_theta(wi), nv = abs_cos_theta(wo);
    float r = s.roughness * s.roughness;
    float rr = 0.5f + 2.0f * nl * nl * r;
    float fl = schlick_weight(nl), fv = schlick_weight(nv);
    return rr * (fl + fv + fl * fv * (rr - 1.0f));
}
SGPU_HD float eval_diffuse(const Surface& s, f3 wo, f3 wm, f3 wi, bool thin)
{
    float nl = abs_cos_theta(wi), nv = abs_cos_theta(wo);
    float fl = schlick_weight(nl), fv = schlick_weight(nv);
    float hk = 0.0f;
    if (thin && s.flatness > 0.0f) {
        float r = s.roughness * s.roughness;
        float hl = dot3(wm, wi);
        float fss90 = hl * hl * r;
        float fss = lerp_(1.0f, fss90, fl) * lerp_(1.0f, fss90, fv);
        hk = 1.25f * (fss * (1.0f / (nl + nv) - 0.5f) + 0.5f);
    }
    float retro = eval_retro(s, wo, wi);
    float sub_approx = lerp_(1.0f, hk, thin ? s.flatness : 0.0f);
    return kInvPi * (retro + sub_approx * (1.0f - 0.5f * fl) * (1.0f - 0.5f * fv));
}

// EvaluateDisney (Disney.cpp:525-614): returns f * |cos(theta_l)|, forward and reverse pdfs.
SGPU_HD f3 disney_eval(const Surface& s, f3 v, f3 l, bool thin, float& fwd, float& rev)
{
    f3 wo = normalize3(to_local(s, v));
    f3 wi = normalize3(to_local(s, l));
    f3 wm = normalize3(wo + wi);
    float nv = cos_theta(wo), nl = cos_theta(wi);
    f3 refl = mk3(0, 0, 0);
    fwd = 0.0f;
    rev = 0.0f;
    float p_brdf, p_diff, p_cc, p_trans;
    lobe_pdfs(s, p_brdf, p_diff, p_cc, p_trans);
    float diff_w = (1.0f - s.metallic) * (1.0f - s.spec_trans);
    float trans_w = (1.0f - s.metallic) * s.spec_trans;
    bool upper = nl > 0.0f && nv > 0.0f;
    if (upper && s.clearcoat > 0.0f) {
        float fcc, rcc;
        float cc = eval_clearcoat(s.clearcoat, s.clearcoat_gloss, wo, wm, wi, fcc, rcc);
        refl += mk3(cc, cc, cc);
        fwd += p_cc * fcc;
        rev += p_cc * rcc;
    }
    if (diff_w > 0.0f) {
        float fd = abs_cos_theta(wi), rd = abs_cos_theta(wo);
        float diffuse = eval_diffuse(s, wo, wm, wi, thin);
        f3 sheen = eval_sheen(s, wm, wi);
        refl += (s.base_color * diffuse + sheen) * diff_w;
        fwd += p_diff * fd;
        rev += p_diff * rd;
    }
    if (trans_w > 0.0f) {
        float rscaled = thin ? thin_trans_roughness(s.ior, s.roughness) : s.roughness;
        float tax, tay;
        aniso_params(rscaled, s.anisotropic, tax, tay);
        f3 tr = eval_spec_trans(s, wo, wm, wi, tax, tay, thin);
        refl += tr * trans_w;
        float ft, rt;
        vndf_pdf(wi, wm, wo, tax, tay, ft, rt);
        float lh = dot3(wm, wi), vh = dot3(wm, wo);
        fwd += p_trans * ft / (sq(lh + s.relative_ior * vh));
        rev += p_trans * rt / (sq(vh + s.relative_ior * lh));
    }
    if (upper) {
        float fm, rm;
        f3 sp = eval_brdf(s, wo, wm, wi, fm, rm);
        refl += sp;
        fwd += p_brdf * fm / (4 * absf_(dot3(wo, wm)));
        rev += p_brdf * rm / (4 * absf_(dot3(wi, wm)));
    }
    return refl * absf_(nl);
}

SGPU_HD f3 extinction_of(f3 a, float sd)
{
    f3 d = a - mk3(0.8f, 0.8f, 0.8f);
    f3 s = (mk3(1.9f, 1.9f, 1.9f) - a) + (d * 3.5f) * d;
    f3 q = s * sd;
    return mk3(q.x / 1.0f, q.y / 1.0f, q.z / 1.0f);
}

SGPU_HD void zero_sample(BsdfSample& o)
{
    o.fwd_pdf = 0.0f;
    o.rev_pdf = 0.0f;
    o.reflectance = mk3(0, 0, 0);
    o.wi = mk3(0, 0, 0);
}

template <class U>
SGPU_HD bool sample_brdf(U& u, const Surface& s, f3 v, BsdfSample& o)
{
    f3 wo = normalize3(to_local(s, v));
    float ax, ay;
    aniso_params(s.roughness, s.anisotropic, ax, ay);
    float r0 = u.next();
    float r1 = u.next();
    f3 wm = sample_vndf(wo, ax, ay, r0, r1);
    f3 wi = normalize3(reflect_(wm, wo));
    if (cos_theta(wi) <= 0.0f) {
        zero_sample(o);
        return false;
    }
    f3 f = disney_fresnel(s, wo, wm, wi);
    float g1v = g1_aniso(wo, ax, ay);
    o.flags = kScatter;
    o.reflectance = f * g1v;
    o.wi = normalize3(to_world(s, wi));
    vndf_pdf(wi, wm, wo, ax, ay, o.fwd_pdf, o.rev_pdf);
    o.fwd_pdf *= (1.0f / (4 * absf_(dot3(wo, wm))));
    o.rev_pdf *= (1.0f / (4 * absf_(dot3(wi, wm))));
    return true;
}

template <class U>
SGPU_HD bool sample_clearcoat(U& u, const Surface& s, f3 v, BsdfSample& o)
{
    f3 wo = normalize3(to_local(s, v));
    const float a = 0.25f, a2 = a * a;
    float r0 = u.next();
    float r1 = u.next();
    float ct = sqrtf(max_(0, (1.0f - pow_f(a2, 1.0f - r0)) / (1.0f - a2)));
    float st = sqrtf(max_(0, 1.0f - ct * ct));
    float phi = (kPi * 2.0f) * r1;
    f3 wm = mk3(st * cos_f(phi), ct, st * sin_f(phi));
    if (dot3(wm, wo) < 0.0f) wm = -wm;
    f3 wi = reflect_(wm, wo);
    if (dot3(wi, wo) < 0.0f) return false;
    float nh = cos_theta(wm), lh = dot3(wm, wi);
    float d = gtr1(absf_(nh), lerp_(0.1f, 0.001f, s.clearcoat_gloss));
    float f = schlick_scalar(0.04f, lh);
    float g = g1_iso(wi, 0.25f) * g1_iso(wo, 0.25f);
    float fpdf = d / (4.0f * dot3(wo, wm));
    float val = 0.25f * s.clearcoat * g * f * d;
    o.reflectance = mk3(val / fpdf, val / fpdf, val / fpdf);
    o.wi = normalize3(to_world(s, wi));
    o.fwd_pdf = fpdf;
    o.rev_pdf = d / (4.0f * dot3(wi, wm));
    return true;
}

template <class U>
SGPU_HD bool sample_spec_trans(U& u, const Surface& s, f3 v, bool thin, BsdfSample& o)
{
    f3 wo = to_local(s, v);
    if (cos_theta(wo) == 0.0f) {
        zero_sample(o);
        return false;
    }
    float rscaled = thin ? thin_trans_roughness(s.ior, s.roughness) : s.roughness;
    float tax, tay;
    aniso_params(rscaled, s.anisotropic, tax, tay);
    float r0 = u.next();
    float r1 = u.next();
    f3 wm = sample_vndf(wo, tax, tay, r0, r1);
    float vh = dot3(wo, wm);
    if (wm.y < 0.0f) vh = -vh;
    float ni = wo.y > 0.0f ? 1.0f : s.ior;
    float nt = wo.y > 0.0f ? s.ior : 1.0f;
    float rel = ni / nt;
    float F = fresnel_dielectric(vh, 1.0f, s.ior);
    float g1v = g1_aniso(wo, tax, tay);
    float pdf;
    f3 wi;
    if (u.next() <= F) {
        wi = normalize3(reflect_(wm, wo));
        o.flags = kScatter;
        o.reflectance = s.base_color * g1v;
        float jac = (4 * absf_(dot3(wo, wm)));
        pdf = F / jac;
    } else {
        if (thin) {
            wi = reflect_(wm, wo);
            wi.y = -wi.y;
            o.reflectance = sqrt3(s.base_color) * g1v;
            o.flags = kScatter;
        } else {
            if (transmit_(wm, wo, rel, wi)) {
                o.flags = kTransmission;
                o.medium = vh > 0.0f ? 1u : 0u;
                o.extinction = extinction_of(s.transmittance_color, s.scatter_distance);
            } else {
                o.flags = kScatter;
                wi = reflect_(wm, wo);
            }
            o.reflectance = s.base_color * g1v;
        }
        wi = normalize3(wi);
        float lh = absf_(dot3(wi, wm));
        float jac = lh / (sq(lh + s.relative_ior * vh));
        pdf = (1.0f - F) / jac;
    }
    if (cos_theta(wi) == 0.0f) {
        zero_sample(o);
        return false;
    }
    if (s.roughness < 0.01f) o.flags |= kDirac;
    vndf_pdf(wi, wm, wo, tax, tay, o.fwd_pdf, o.rev_pdf);
    o.fwd_pdf *= pdf;
    o.rev_pdf *= pdf;
    o.wi = normalize3(to_world(s, wi));
    return true;
}

template <class U>
SGPU_HD bool sample_diffuse(U& u, const Surface& s, f3 v, bool thin, BsdfSample& o)
{
    f3 wo = to_local(s, v);
    float sg = sign_(cos_theta(wo));
    float r0 = u.next();
    float r1 = u.next();
    float r = sqrtf(r0);
    float theta = (kPi * 2.0f) * r1;
    f3 wi = sg * mk3(r * cos_f(theta), sqrtf(max_(0.0f, 1 - r0)), r * sin_f(theta));
    f3 wm = normalize3(wi + wo);
    float nl = cos_theta(wi);
    if (nl == 0.0f) {
        zero_sample(o);
        return false;
    }
    float nv = cos_theta(wo);
    float pdf;
    uint32_t ev = kScatter;
    f3 color = s.base_color;
    float p = u.next();
    if (p <= s.diff_trans) {
        wi = -wi;
        pdf = s.diff_trans;
        if (thin) {
            color = sqrt3(color);
        } else {
            ev = kTransmission;
            o.medium = 1u;
            o.extinction = extinction_of(s.transmittance_color, s.scatter_distance);
        }
    } else {
        pdf = (1.0f - s.diff_trans);
    }
    f3 sheen = eval_sheen(s, wm, wi);
    float diffuse = eval_diffuse(s, wo, wm, wi, thin);
    o.reflectance = sheen + color * (diffuse / pdf);
    o.wi = normalize3(to_world(s, wi));
    o.fwd_pdf = absf_(nl) * pdf;
    o.rev_pdf = absf_(nv) * pdf;
    o.flags = ev;
    return true;
}

// SampleDisney (Disney.cpp:617-659)
template <class U>
SGPU_HD bool disney_sample(U& u, const Surface& s, f3 v, bool thin, BsdfSample& o)
{
    float p_spec, p_diff, p_cc, p_trans;
    lobe_pdfs(s, p_spec, p_diff, p_cc, p_trans);
    bool ok = false;
    float p_lobe = 0.0f;
    float p = u.next();
    if (p <= p_spec) {
        ok = sample_brdf(u, s, v, o);
        p_lobe = p_spec;
    } else if (p > p_spec && p <= (p_spec + p_cc)) {
        ok = sample_clearcoat(u, s, v, o);
        p_lobe = p_cc;
    } else if (p > p_spec + p_cc && p <= (p_spec + p_cc + p_diff)) {
        ok = sample_diffuse(u, s, v, thin, o);
        p_lobe = p_diff;
    } else if (p_trans >= 0.0f) {
        ok = sample_spec_trans(u, s, v, thin, o);
        p_lobe = p_trans;
    } else {
        o.reflectance = mk3(1000000.0f, 0.0f, 0.0f);
        o.fwd_pdf = 0.000000001f;
        o.rev_pdf = 0.000000001f;
    }
    if (p_lobe > 0.0f) {
        o.reflectance = o.reflectance * (1.0f / p_lobe);
        o.fwd_pdf *= p_lobe;
        o.rev_pdf *= p_lobe;
    }
    return ok;
}

// ---- Lambert (corrected form) --------------------------------------------------------------------
SGPU_HD f3 lambert_eval(const Surface& s, f3 v, f3 l, float& fwd, float& rev)
{
    fwd = 0.0f;
    rev = 0.0f;
    float nl = dot3(s.n, l), nv = dot3(s.n, v);
    if (nl <= 0.0f || nv <= 0.0f) return mk3(0, 0, 0);
    fwd = nl * kInvPi;
    rev = nv * kInvPi;
    return s.base_color * (kInvPi * nl);
}

template <class U>
SGPU_HD bool lambert_sample(U& u, const Surface& s, f3 v, BsdfSample& o)
{
    f3 wo = to_local(s, v);
    float r0 = u.next();
    float r1 = u.next();
    float r = sqrtf(r0);
    float theta = (kPi * 2.0f) * r1;
    f3 wi = mk3(r * cos_f(theta), sqrtf(max_(0.0f, 1 - r0)), r * sin_f(theta));
    o.flags = kScatter;
    o.reflectance = s.base_color;
    o.wi = normalize3(to_world(s, wi));
    o.fwd_pdf = cos_theta(wi) * kInvPi;
    o.rev_pdf = max_(cos_theta(wo), 0.0f) * kInvPi;
    return true;
}

// ---- dispatch (SurfaceScattering.cpp:18-60) ---------------------------------------------------------
SGPU_HD f3 bsdf_eval(const Surface& s, f3 v, f3 l, float& fwd, float& rev)
{
    if (s.shader == kLambert) return lambert_eval(s, v, l, fwd, rev);
    return disney_eval(s, v, l, s.shader == kDisneyThin, fwd, rev);
}

template <class U>
SGPU_HD bool bsdf_sample(U& u, const Surface& s, f3 v, BsdfSample& o)
{
    o.flags = 0;
    o.medium = 0;
    o.extinction = mk3(0, 0, 0);
    o.reflectance = mk3(0, 0, 0);
    o.wi = mk3(0, 0, 0);
    o.fwd_pdf = 0.0f;
    o.rev_pdf = 0.0f;
    if (s.shader == kLambert) return lambert_sample(u, s, v, o);
    return disney_sample(u, s, v, s.shader == kDisneyThin, o);
}

// ContinuationProbability (SurfaceParameters.cpp:420-424)
SGPU_HD float continuation_probability(const Surface& s) { return saturate_(maxc3(s.base_color)); }

// OffsetRayOrigin (SurfaceParameters.cpp:403-416): signed bias along the shading normal
SGPU_HD f3 offset_ray_origin(const Surface& s, f3 dir, float bias_scale, float* signed_bias = nullptr)
{
    float dir_off = dot3(dir, s.n) < 0.0f ? -1.0f : 1.0f;
    float sb = dir_off * s.error * bias_scale;
    if (signed_bias) *signed_bias = sb;
    return s.position + sb * s.n;
}

// Uniform stream over an explicit array (the role of CSampler in a test harness).
struct ArrayUniforms {
    const float* v;
    uint32_t count;
    uint32_t pos;
    SGPU_HD float next() { return pos < count ? v[pos++] : (++pos, 0.5f); }
};

// Uniform stream over a PCG32 state.
struct PcgUniforms {
    Pcg32 r;
    SGPU_HD float next() { return pcg32_float(r); }
};

}  // namespace sgpu
