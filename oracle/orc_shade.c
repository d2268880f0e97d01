/*
 * ORACLE -- test infrastructure only (see orc_trace.c for the rules). Compiled with
 * -ffp-contract=off.
 *
 * CPU restatement of Selas's shading math, pinned against golden vectors recorded from the
 * reference's own compiled functions (tests/golden/make_golden.py, tests/test_oracle_shade.py):
 *   Fresnel          Source/Core/Shading/Fresnel.cpp:19-82
 *   GGX / Smith      Source/Core/Shading/Ggx.cpp:19-155
 *   Disney BSDF      Source/Core/Shading/Disney.cpp:26-659 (eval + sample, thin and solid)
 *   tangent frames   Source/Core/GeometryLib/CoordinateSystem.h:14-21, SurfaceParameters.cpp:240-285
 *   CMJ              Source/Core/MathLib/CorrelatedMultiJitter.cpp:12-62
 *   camera           Source/Core/GeometryLib/Camera.cpp:56-112
 *   rect light       Source/Core/GeometryLib/RectangulerLightSampler.cpp:17-98
 *   RayQuad          Source/Core/MathLib/GeometryIntersection.cpp:183-236
 *   concentric disc  Source/Core/GeometryLib/Disc.cpp:13-56
 * Lambert (Source/Core/Shading/Lambert.cpp) is restated in its corrected form (pdf cos/pi, sample
 * weight = albedo); the reference's own Lambert is unreachable in Selas (LambertAllTheThings_ 0) and
 * its pdf lacks 1/pi. Pinned by closed forms, not by golden vectors.
 *
 * Reference quirks kept on purpose (they change values): the specular pdf of EvaluateDisney is
 * divided by 4|wo.wm| twice (Disney.cpp:162,607); Fresnel::Schlick(r0, c) is lerp(1, w(c), r0)
 * (Fresnel.cpp:30-33); `float / float3` returns the vector divided by the scalar
 * (FloatFuncs.h:147-151), so CalculateExtinction returns s * scatterDistance.
 */
#include "oracle.h"
#include "orc_fn.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

int orc_libm_single = 0;

void orc_set_libm(int single) { orc_libm_single = single ? 1 : 0; }

static const float kPi = 3.141592653589793f;

static inline ov3 v3(float x, float y, float z)
{
    ov3 r = {x, y, z};
    return r;
}
static inline ov3 add(ov3 a, ov3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline ov3 sub(ov3 a, ov3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline ov3 mul(ov3 a, ov3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline ov3 scl(ov3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline ov3 neg(ov3 a) { return v3(-a.x, -a.y, -a.z); }
static inline float dot(ov3 a, ov3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
static inline ov3 cross(ov3 a, ov3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float length(ov3 a) { return sqrtf(dot(a, a)); }
static inline ov3 normalize(ov3 a)
{
    float inv = 1.f / length(a);
    return v3(a.x * inv, a.y * inv, a.z * inv);
}
static inline float absf(float x) { return (x < 0) ? -x : x; }
static inline float maxf(float a, float b) { return (a > b) ? a : b; }
static inline float minf(float a, float b) { return (a < b) ? a : b; }
static inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline float saturate(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }
static inline float square(float x) { return x * x; }
static inline float lerp(float a, float b, float t) { return (1.0f - t) * a + t * b; }
static inline ov3 lerp3(ov3 a, ov3 b, float t) { return add(scl(a, 1.0f - t), scl(b, t)); }
static inline float signf_(float x) { return x < 0.0f ? -1.0f : (x > 0.0f ? 1.0f : 0.0f); }

/* y-up tangent space trigonometry (Trigonometric.h:51-110) */
static inline float cos_theta(ov3 w) { return w.y; }
static inline float cos2_theta(ov3 w) { return w.y * w.y; }
static inline float abs_cos_theta(ov3 w) { return absf(w.y); }
static inline float sin2_theta(ov3 w) { return maxf(0.0f, 1.0f - cos2_theta(w)); }
static inline float sin_theta(ov3 w) { return sqrtf(sin2_theta(w)); }
static inline float tan_theta(ov3 w) { return sin_theta(w) / cos_theta(w); }
static inline float cos_phi(ov3 w)
{
    float s = sin_theta(w);
    return (s == 0) ? 1.0f : clampf(w.x / s, -1.0f, 1.0f);
}
static inline float sin_phi(ov3 w)
{
    float s = sin_theta(w);
    return (s == 0) ? 1.0f : clampf(w.z / s, -1.0f, 1.0f);
}
static inline float cos2_phi(ov3 w)
{
    float c = cos_phi(w);
    return c * c;
}
static inline float sin2_phi(ov3 w)
{
    float s = sin_phi(w);
    return s * s;
}
static inline ov3 reflect(ov3 n, ov3 l) { return sub(scl(n, 2.0f * dot(n, l)), l); }
static inline int transmit(ov3 wm, ov3 wi, float n, ov3* wo)
{
    float c = dot(wi, wm);
    if (c < 0.0f) {
        c = -c;
        wm = neg(wm);
    }
    float root = 1.0f - n * n * (1.0f - c * c);
    if (root <= 0) return 0;
    *wo = sub(scl(wm, n * c - sqrtf(root)), scl(wi, n));
    return 1;
}

/* ---- frames ----------------------------------------------------------------------------------- */

void orc_make_frame(const float* n, float* t_out, float* b_out)
{
    ov3 v1 = v3(n[0], n[1], n[2]), t;
    if (absf(v1.x) > absf(v1.y))
        t = scl(v3(-v1.z, 0, v1.x), 1.0f / sqrtf(v1.x * v1.x + v1.z * v1.z));
    else
        t = scl(v3(0, v1.z, -v1.y), 1.0f / sqrtf(v1.y * v1.y + v1.z * v1.z));
    ov3 b = cross(v1, t);
    t_out[0] = t.x;
    t_out[1] = t.y;
    t_out[2] = t.z;
    b_out[0] = b.x;
    b_out[1] = b.y;
    b_out[2] = b.z;
}

/* world -> tangent (MatrixMultiply(v, worldToTangent)) and back (MatrixMultiply(w, tangentToWorld)) */
static inline ov3 to_local(const orc_surface* s, ov3 v) { return v3(dot(v, s->t), dot(v, s->n), dot(v, s->b)); }
static inline ov3 to_world(const orc_surface* s, ov3 w)
{
    return v3(w.x * s->t.x + w.y * s->n.x + w.z * s->b.x, w.x * s->t.y + w.y * s->n.y + w.z * s->b.y,
              w.x * s->t.z + w.y * s->n.z + w.z * s->b.z);
}

void orc_surface_from_block(const float* p, const float* n, orc_surface* s)
{
    memset(s, 0, sizeof(*s));
    s->n = v3(n[0], n[1], n[2]);
    float t[3], b[3];
    orc_make_frame(n, t, b);
    s->t = v3(t[0], t[1], t[2]);
    s->b = v3(b[0], b[1], b[2]);
    s->base_color = v3(p[0], p[1], p[2]);
    s->transmittance_color = v3(p[3], p[4], p[5]);
    s->sheen = p[6];
    s->sheen_tint = p[7];
    s->clearcoat = p[8];
    s->clearcoat_gloss = p[9];
    s->metallic = p[10];
    s->spec_trans = p[11];
    s->diff_trans = p[12];
    s->flatness = p[13];
    s->anisotropic = p[14];
    s->relative_ior = p[15];
    s->specular_tint = p[16];
    s->roughness = p[17];
    s->scatter_distance = p[18];
    s->ior = p[19];
    s->shader = (int)p[20];
    s->flags = (uint32_t)p[21];
}

/* ---- Fresnel ---------------------------------------------------------------------------------- */

float orc_schlick_weight(float u)
{
    float m = saturate(1.0f - u);
    float m2 = m * m;
    return m * m2 * m2;
}
static inline float schlick_scalar(float r0, float c) { return lerp(1.0f, orc_schlick_weight(c), r0); }
static inline ov3 schlick3(ov3 r0, float c)
{
    float e = pow5_f(1.0f - c);
    return add(r0, scl(sub(v3(1.0f, 1.0f, 1.0f), r0), e));
}
float orc_schlick_r0(float eta) { return square(eta - 1.0f) / square(eta + 1.0f); }
float orc_fresnel_dielectric(float cos_i, float ni, float nt)
{
    cos_i = clampf(cos_i, -1.0f, 1.0f);
    if (cos_i < 0.0f) {
        float tmp = ni;
        ni = nt;
        nt = tmp;
        cos_i = -cos_i;
    }
    float sin_i = sqrtf(maxf(0.0f, 1.0f - cos_i * cos_i));
    float sin_t = ni / nt * sin_i;
    if (sin_t >= 1) return 1;
    float cos_t = sqrtf(maxf(0.0f, 1.0f - sin_t * sin_t));
    float rpar = ((nt * cos_i) - (ni * cos_t)) / ((nt * cos_i) + (ni * cos_t));
    float rperp = ((ni * cos_i) - (nt * cos_t)) / ((ni * cos_i) + (nt * cos_t));
    return (rpar * rpar + rperp * rperp) / 2;
}
float orc_schlick(float r0, float c) { return schlick_scalar(r0, c); }

/* ---- GGX -------------------------------------------------------------------------------------- */

static float g1_aniso(ov3 w, ov3 wm, float ax, float ay)
{
    (void)wm;
    float abs_tan = absf(tan_theta(w));
    if (isinf(abs_tan)) return 0.0f;
    float a = sqrtf(cos2_phi(w) * ax * ax + sin2_phi(w) * ay * ay);
    float a2tan2 = square(a * abs_tan);
    float lambda = 0.5f * (-1.0f + sqrtf(1.0f + a2tan2));
    return 1.0f / (1.0f + lambda);
}
static float g1_iso(ov3 w, float a)
{
    float a2 = a * a;
    float c = abs_cos_theta(w);
    return 2.0f / (1.0f + sqrtf(a2 + (1 - a2) * c * c));
}
static float d_aniso(ov3 wm, float ax, float ay)
{
    float hx2 = square(wm.x);
    float hy2 = square(wm.z);
    float c2 = cos2_theta(wm);
    float ax2 = square(ax);
    float ay2 = square(ay);
    return 1.0f / (kPi * ax * ay * square(hx2 / ax2 + hy2 / ay2 + c2));
}
static ov3 sample_vndf(ov3 wo, float ax, float ay, float u1, float u2)
{
    ov3 v = normalize(v3(wo.x * ax, wo.y, wo.z * ay));
    ov3 t1 = (v.y < 0.9999f) ? normalize(cross(v, v3(0, 1, 0))) : v3(1, 0, 0);
    ov3 t2 = cross(t1, v);
    float a = 1.0f / (1.0f + v.y);
    float r = sqrtf(u1);
    float phi = (u2 < a) ? (u2 / a) * kPi : kPi + (u2 - a) / (1.0f - a) * kPi;
    float p1 = r * cos_f(phi);
    float p2 = r * sin_f(phi) * ((u2 < a) ? 1.0f : v.y);
    ov3 n = add(add(scl(t1, p1), scl(t2, p2)), scl(v, sqrtf(maxf(0.0f, 1.0f - p1 * p1 - p2 * p2))));
    return normalize(v3(ax * n.x, n.y, ay * n.z));
}
static void vndf_pdf(ov3 wi, ov3 wm, ov3 wo, float ax, float ay, float* fwd, float* rev)
{
    float d = d_aniso(wm, ax, ay);
    float nl = abs_cos_theta(wi);
    float hl = absf(dot(wm, wi));
    float g1v = g1_aniso(wo, wm, ax, ay);
    *fwd = g1v * hl * d / nl;
    float nv = abs_cos_theta(wo);
    float hv = absf(dot(wm, wo));
    float g1l = g1_aniso(wi, wm, ax, ay);
    *rev = g1l * hv * d / nv;
}

void orc_ggx(const float* wo, const float* wi, const float* wm, float ax, float ay, float u1, float u2, float* out)
{
    ov3 o = v3(wo[0], wo[1], wo[2]), i = v3(wi[0], wi[1], wi[2]), m = v3(wm[0], wm[1], wm[2]);
    out[0] = d_aniso(m, ax, ay);
    out[1] = g1_aniso(o, m, ax, ay);
    ov3 s = sample_vndf(o, ax, ay, u1, u2);
    out[2] = s.x;
    out[3] = s.y;
    out[4] = s.z;
    vndf_pdf(i, m, o, ax, ay, &out[5], &out[6]);
    out[7] = g1_iso(o, 0.25f);
}

/* ---- Disney ------------------------------------------------------------------------------------ */

static void lobe_pdfs(const orc_surface* s, float* p_spec, float* p_diff, float* p_cc, float* p_trans)
{
    float metallic_brdf = s->metallic;
    float specular_bsdf = (1.0f - s->metallic) * s->spec_trans;
    float dielectric_brdf = (1.0f - s->spec_trans) * (1.0f - s->metallic);
    float spec_w = metallic_brdf + dielectric_brdf;
    float trans_w = specular_bsdf;
    float diff_w = dielectric_brdf;
    float cc_w = 1.0f * saturate(s->clearcoat);
    float norm = 1.0f / (spec_w + trans_w + diff_w + cc_w);
    *p_spec = spec_w * norm;
    *p_trans = trans_w * norm;
    *p_diff = diff_w * norm;
    *p_cc = cc_w * norm;
}
static float thin_trans_roughness(float ior, float roughness) { return saturate((0.65f * ior - 0.35f) * roughness); }
static void aniso_params(float roughness, float anisotropic, float* ax, float* ay)
{
    float aspect = sqrtf(1.0f - 0.9f * anisotropic);
    *ax = maxf(0.001f, square(roughness) / aspect);
    *ay = maxf(0.001f, square(roughness) * aspect);
}
static ov3 tint_of(ov3 c)
{
    float lum = dot(v3(0.3f, 0.6f, 1.0f), c);
    return (lum > 0.0f) ? scl(c, 1.0f / lum) : v3(1, 1, 1);
}
static float gtr1(float abs_hl, float a)
{
    if (a >= 1) return 1.0f / kPi;
    float a2 = a * a;
    return (a2 - 1.0f) / (kPi * log2f(a2) * (1.0f + (a2 - 1.0f) * abs_hl * abs_hl));
}
static float eval_clearcoat(float clearcoat, float alpha, ov3 wo, ov3 wm, ov3 wi, float* f_pdf, float* r_pdf)
{
    if (clearcoat <= 0.0f) return 0.0f;
    float nh = abs_cos_theta(wm);
    float hl = dot(wm, wi);
    float d = gtr1(nh, lerp(0.1f, 0.001f, alpha));
    float f = schlick_scalar(0.04f, hl);
    float gl = g1_iso(wi, 0.25f);
    float gv = g1_iso(wo, 0.25f);
    *f_pdf = d / (4.0f * absf(dot(wo, wm)));
    *r_pdf = d / (4.0f * absf(dot(wi, wm)));
    return 0.25f * clearcoat * d * f * gl * gv;
}
static ov3 eval_sheen(const orc_surface* s, ov3 wo, ov3 wm, ov3 wi)
{
    (void)wo;
    if (s->sheen <= 0.0f) return v3(0, 0, 0);
    float hl = absf(dot(wm, wi));
    ov3 tint = tint_of(s->base_color);
    /* sheen * lerp(1, tint, sheenTint) * w: float * float3, then float3 * float */
    return scl(scl(lerp3(v3(1, 1, 1), tint, s->sheen_tint), s->sheen), orc_schlick_weight(hl));
}
static ov3 disney_fresnel(const orc_surface* s, ov3 wo, ov3 wm, ov3 wi)
{
    float hv = dot(wm, wo);
    ov3 tint = tint_of(s->base_color);
    ov3 r0 = scl(lerp3(v3(1, 1, 1), tint, s->specular_tint), orc_schlick_r0(s->relative_ior));
    r0 = lerp3(r0, s->base_color, s->metallic);
    float df = orc_fresnel_dielectric(hv, 1.0f, s->ior);
    ov3 mf = schlick3(r0, dot(wi, wm));
    return lerp3(v3(df, df, df), mf, s->metallic);
}
static ov3 eval_brdf(const orc_surface* s, ov3 wo, ov3 wm, ov3 wi, float* f_pdf, float* r_pdf)
{
    *f_pdf = 0.0f;
    *r_pdf = 0.0f;
    float nl = cos_theta(wi), nv = cos_theta(wo);
    if (nl <= 0.0f || nv <= 0.0f) return v3(0, 0, 0);
    float ax, ay;
    aniso_params(s->roughness, s->anisotropic, &ax, &ay);
    float d = d_aniso(wm, ax, ay);
    float gl = g1_aniso(wi, wm, ax, ay);
    float gv = g1_aniso(wo, wm, ax, ay);
    ov3 f = disney_fresnel(s, wo, wm, wi);
    vndf_pdf(wi, wm, wo, ax, ay, f_pdf, r_pdf);
    *f_pdf *= (1.0f / (4 * absf(dot(wo, wm))));
    *r_pdf *= (1.0f / (4 * absf(dot(wi, wm))));
    float k = d * gl * gv;
    float den = 4.0f * nl * nv;
    return v3(k * f.x / den, k * f.y / den, k * f.z / den);
}
static ov3 eval_spec_trans(const orc_surface* s, ov3 wo, ov3 wm, ov3 wi, float ax, float ay, int thin)
{
    float rel = s->relative_ior;
    float n2 = rel * rel;
    float nl = abs_cos_theta(wi), nv = abs_cos_theta(wo);
    float hl = dot(wm, wi), hv = dot(wm, wo);
    float ahl = absf(hl), ahv = absf(hv);
    float d = d_aniso(wm, ax, ay);
    float gl = g1_aniso(wi, wm, ax, ay);
    float gv = g1_aniso(wo, wm, ax, ay);
    float f = orc_fresnel_dielectric(hv, 1.0f, s->ior);
    ov3 color = thin ? v3(sqrtf(s->base_color.x), sqrtf(s->base_color.y), sqrtf(s->base_color.z)) : s->base_color;
    float c = (ahl * ahv) / (nl * nv);
    float t = (n2 / square(hl + rel * hv));
    /* color * c * t * (1-f) * gl * gv * d, left to right with float3 * float */
    ov3 r = scl(color, c);
    r = scl(r, t);
    r = scl(r, 1.0f - f);
    r = scl(r, gl);
    r = scl(r, gv);
    r = scl(r, d);
    return r;
}
static float eval_retro(const orc_surface* s, ov3 wo, ov3 wm, ov3 wi)
{
    (void)wm;
    float nl = abs_cos_theta(wi), nv = abs_cos_theta(wo);
    float r = s->roughness * s->roughness;
    float rr = 0.5f + 2.0f * nl * nl * r;
    float fl = orc_schlick_weight(nl), fv = orc_schlick_weight(nv);
    return rr * (fl + fv + fl * fv * (rr - 1.0f));
}
static float eval_diffuse(const orc_surface* s, ov3 wo, ov3 wm, ov3 wi, int thin)
{
    float nl = abs_cos_theta(wi), nv = abs_cos_theta(wo);
    float fl = orc_schlick_weight(nl), fv = orc_schlick_weight(nv);
    float hk = 0.0f;
    if (thin && s->flatness > 0.0f) {
        float r = s->roughness * s->roughness;
        float hl = dot(wm, wi);
        float fss90 = hl * hl * r;
        float fss = lerp(1.0f, fss90, fl) * lerp(1.0f, fss90, fv);
        float ss = 1.25f * (fss * (1.0f / (nl + nv) - 0.5f) + 0.5f);
        hk = ss;
    }
    float lambert = 1.0f;
    float retro = eval_retro(s, wo, wm, wi);
    float sub_approx = lerp(lambert, hk, thin ? s->flatness : 0.0f);
    return (1.0f / kPi) * (retro + sub_approx * (1.0f - 0.5f * fl) * (1.0f - 0.5f * fv));
}

void orc_disney_eval(const orc_surface* s, const float* vv, const float* ll, int thin, float* out)
{
    ov3 wo = normalize(to_local(s, v3(vv[0], vv[1], vv[2])));
    ov3 wi = normalize(to_local(s, v3(ll[0], ll[1], ll[2])));
    ov3 wm = normalize(add(wo, wi));
    float nv = cos_theta(wo), nl = cos_theta(wi);
    ov3 refl = v3(0, 0, 0);
    float fwd = 0.0f, rev = 0.0f;
    float p_brdf, p_diff, p_cc, p_trans;
    lobe_pdfs(s, &p_brdf, &p_diff, &p_cc, &p_trans);
    float metallic = s->metallic, spec_trans = s->spec_trans;
    float ax, ay;
    aniso_params(s->roughness, s->anisotropic, &ax, &ay);
    float diff_w = (1.0f - metallic) * (1.0f - spec_trans);
    float trans_w = (1.0f - metallic) * spec_trans;
    int upper = nl > 0.0f && nv > 0.0f;
    if (upper && s->clearcoat > 0.0f) {
        float fcc, rcc;
        float cc = eval_clearcoat(s->clearcoat, s->clearcoat_gloss, wo, wm, wi, &fcc, &rcc);
        refl = add(refl, v3(cc, cc, cc));
        fwd += p_cc * fcc;
        rev += p_cc * rcc;
    }
    if (diff_w > 0.0f) {
        float fd = abs_cos_theta(wi), rd = abs_cos_theta(wo);
        float diffuse = eval_diffuse(s, wo, wm, wi, thin);
        ov3 sheen = eval_sheen(s, wo, wm, wi);
        refl = add(refl, scl(add(scl(s->base_color, diffuse), sheen), diff_w));
        fwd += p_diff * fd;
        rev += p_diff * rd;
    }
    if (trans_w > 0.0f) {
        float rscaled = thin ? thin_trans_roughness(s->ior, s->roughness) : s->roughness;
        float tax, tay;
        aniso_params(rscaled, s->anisotropic, &tax, &tay);
        ov3 tr = eval_spec_trans(s, wo, wm, wi, tax, tay, thin);
        refl = add(refl, scl(tr, trans_w));
        float ft, rt;
        vndf_pdf(wi, wm, wo, tax, tay, &ft, &rt);
        float lh = dot(wm, wi), vh = dot(wm, wo);
        fwd += p_trans * ft / (square(lh + s->relative_ior * vh));
        rev += p_trans * rt / (square(vh + s->relative_ior * lh));
    }
    if (upper) {
        float fm, rm;
        ov3 sp = eval_brdf(s, wo, wm, wi, &fm, &rm);
        refl = add(refl, sp);
        fwd += p_brdf * fm / (4 * absf(dot(wo, wm)));
        rev += p_brdf * rm / (4 * absf(dot(wi, wm)));
    }
    refl = scl(refl, absf(nl));
    out[0] = refl.x;
    out[1] = refl.y;
    out[2] = refl.z;
    out[3] = fwd;
    out[4] = rev;
}

typedef struct {
    int ok;
    uint32_t flags;
    ov3 refl, wi;
    float fwd, rev;
    int medium;
    ov3 ext;
} bsdf_sample;

static ov3 extinction(ov3 a, float sd)
{
    /* float3(1.9) - a + (3.5 * (a - 0.8)) * (a - 0.8) */
    ov3 d = sub(a, v3(0.8f, 0.8f, 0.8f));
    ov3 s = add(sub(v3(1.9f, 1.9f, 1.9f), a), mul(scl(d, 3.5f), d));
    /* Selas's `1.0f / (s * sd)` is `float / float3`, which returns the vector divided by 1 */
    ov3 q = scl(s, sd);
    return v3(q.x / 1.0f, q.y / 1.0f, q.z / 1.0f);
}

static void zero_sample(bsdf_sample* o)
{
    o->fwd = 0.0f;
    o->rev = 0.0f;
    o->refl = v3(0, 0, 0);
    o->wi = v3(0, 0, 0);
}

static int sample_brdf(orc_uniforms* u, const orc_surface* s, ov3 v, bsdf_sample* o)
{
    ov3 wo = normalize(to_local(s, v));
    float ax, ay;
    aniso_params(s->roughness, s->anisotropic, &ax, &ay);
    float r0 = orc_uniform(u), r1 = orc_uniform(u);
    ov3 wm = sample_vndf(wo, ax, ay, r0, r1);
    ov3 wi = normalize(reflect(wm, wo));
    if (cos_theta(wi) <= 0.0f) {
        zero_sample(o);
        return 0;
    }
    ov3 f = disney_fresnel(s, wo, wm, wi);
    float g1v = g1_aniso(wo, wm, ax, ay);
    o->flags = 1;
    o->refl = scl(f, g1v);
    o->wi = normalize(to_world(s, wi));
    vndf_pdf(wi, wm, wo, ax, ay, &o->fwd, &o->rev);
    o->fwd *= (1.0f / (4 * absf(dot(wo, wm))));
    o->rev *= (1.0f / (4 * absf(dot(wi, wm))));
    return 1;
}

static int sample_clearcoat(orc_uniforms* u, const orc_surface* s, ov3 v, bsdf_sample* o)
{
    ov3 wo = normalize(to_local(s, v));
    float a = 0.25f, a2 = a * a;
    float r0 = orc_uniform(u), r1 = orc_uniform(u);
    float ct = sqrtf(maxf(0, (1.0f - pow_f(a2, 1.0f - r0)) / (1.0f - a2)));
    float st = sqrtf(maxf(0, 1.0f - ct * ct));
    float phi = (kPi * 2.0f) * r1;
    ov3 wm = v3(st * cos_f(phi), ct, st * sin_f(phi));
    if (dot(wm, wo) < 0.0f) wm = neg(wm);
    ov3 wi = reflect(wm, wo);
    if (dot(wi, wo) < 0.0f) return 0;
    float nh = cos_theta(wm), lh = dot(wm, wi);
    float d = gtr1(absf(nh), lerp(0.1f, 0.001f, s->clearcoat_gloss));
    float f = schlick_scalar(0.04f, lh);
    float g = g1_iso(wi, 0.25f) * g1_iso(wo, 0.25f);
    float fpdf = d / (4.0f * dot(wo, wm));
    float val = 0.25f * s->clearcoat * g * f * d;
    o->refl = v3(val / fpdf, val / fpdf, val / fpdf);
    o->wi = normalize(to_world(s, wi));
    o->fwd = fpdf;
    o->rev = d / (4.0f * dot(wi, wm));
    return 1;
}

static int sample_spec_trans(orc_uniforms* u, const orc_surface* s, ov3 v, int thin, bsdf_sample* o)
{
    ov3 wo = to_local(s, v);
    if (cos_theta(wo) == 0.0) {
        zero_sample(o);
        return 0;
    }
    float rscaled = thin ? thin_trans_roughness(s->ior, s->roughness) : s->roughness;
    float tax, tay;
    aniso_params(rscaled, s->anisotropic, &tax, &tay);
    float r0 = orc_uniform(u), r1 = orc_uniform(u);
    ov3 wm = sample_vndf(wo, tax, tay, r0, r1);
    float vh = dot(wo, wm);
    if (wm.y < 0.0f) vh = -vh;
    float ni = wo.y > 0.0f ? 1.0f : s->ior;
    float nt = wo.y > 0.0f ? s->ior : 1.0f;
    float rel = ni / nt;
    float F = orc_fresnel_dielectric(vh, 1.0f, s->ior);
    float g1v = g1_aniso(wo, wm, tax, tay);
    float pdf;
    ov3 wi;
    if (orc_uniform(u) <= F) {
        wi = normalize(reflect(wm, wo));
        o->flags = 1;
        o->refl = scl(s->base_color, g1v);
        float jac = (4 * absf(dot(wo, wm)));
        pdf = F / jac;
    } else {
        if (thin) {
            wi = reflect(wm, wo);
            wi.y = -wi.y;
            o->refl = scl(v3(sqrtf(s->base_color.x), sqrtf(s->base_color.y), sqrtf(s->base_color.z)), g1v);
            o->flags = 1;
        } else {
            if (transmit(wm, wo, rel, &wi)) {
                o->flags = 2;
                o->medium = vh > 0.0f ? 1 : 0;
                o->ext = extinction(s->transmittance_color, s->scatter_distance);
            } else {
                o->flags = 1;
                wi = reflect(wm, wo);
            }
            o->refl = scl(s->base_color, g1v);
        }
        wi = normalize(wi);
        float lh = absf(dot(wi, wm));
        float jac = lh / (square(lh + s->relative_ior * vh));
        pdf = (1.0f - F) / jac;
    }
    if (cos_theta(wi) == 0.0f) {
        zero_sample(o);
        return 0;
    }
    if (s->roughness < 0.01f) o->flags |= 4;
    vndf_pdf(wi, wm, wo, tax, tay, &o->fwd, &o->rev);
    o->fwd *= pdf;
    o->rev *= pdf;
    o->wi = normalize(to_world(s, wi));
    return 1;
}

static int sample_diffuse(orc_uniforms* u, const orc_surface* s, ov3 v, int thin, bsdf_sample* o)
{
    ov3 wo = to_local(s, v);
    float sign = signf_(cos_theta(wo));
    float r0 = orc_uniform(u), r1 = orc_uniform(u);
    float r = sqrtf(r0);
    float theta = (kPi * 2.0f) * r1;
    ov3 wi = scl(v3(r * cos_f(theta), sqrtf(maxf(0.0f, 1 - r0)), r * sin_f(theta)), sign);
    ov3 wm = normalize(add(wi, wo));
    float nl = cos_theta(wi);
    if (nl == 0.0f) {
        zero_sample(o);
        return 0;
    }
    float nv = cos_theta(wo);
    float pdf;
    uint32_t ev = 1;
    ov3 color = s->base_color;
    float p = orc_uniform(u);
    if (p <= s->diff_trans) {
        wi = neg(wi);
        pdf = s->diff_trans;
        if (thin) {
            color = v3(sqrtf(color.x), sqrtf(color.y), sqrtf(color.z));
        } else {
            ev = 2;
            o->medium = 1;
            o->ext = extinction(s->transmittance_color, s->scatter_distance);
        }
    } else {
        pdf = (1.0f - s->diff_trans);
    }
    ov3 sheen = eval_sheen(s, wo, wm, wi);
    float diffuse = eval_diffuse(s, wo, wm, wi, thin);
    o->refl = add(sheen, scl(color, diffuse / pdf));
    o->wi = normalize(to_world(s, wi));
    o->fwd = absf(nl) * pdf;
    o->rev = absf(nv) * pdf;
    o->flags = ev;
    return 1;
}

int orc_disney_sample(orc_uniforms* u, const orc_surface* s, const float* vv, int thin, float* out)
{
    ov3 v = v3(vv[0], vv[1], vv[2]);
    float p_spec, p_diff, p_cc, p_trans;
    lobe_pdfs(s, &p_spec, &p_diff, &p_cc, &p_trans);
    bsdf_sample o;
    memset(&o, 0, sizeof(o));
    int ok = 0;
    float p_lobe = 0.0f;
    float p = orc_uniform(u);
    if (p <= p_spec) {
        ok = sample_brdf(u, s, v, &o);
        p_lobe = p_spec;
    } else if (p > p_spec && p <= (p_spec + p_cc)) {
        ok = sample_clearcoat(u, s, v, &o);
        p_lobe = p_cc;
    } else if (p > p_spec + p_cc && p <= (p_spec + p_cc + p_diff)) {
        ok = sample_diffuse(u, s, v, thin, &o);
        p_lobe = p_diff;
    } else if (p_trans >= 0.0f) {
        ok = sample_spec_trans(u, s, v, thin, &o);
        p_lobe = p_trans;
    } else {
        o.refl = v3(1000000.0f, 0.0f, 0.0f);
        o.fwd = 0.000000001f;
        o.rev = 0.000000001f;
    }
    if (p_lobe > 0.0f) {
        o.refl = scl(o.refl, 1.0f / p_lobe);
        o.fwd *= p_lobe;
        o.rev *= p_lobe;
    }
    out[0] = ok ? 1.0f : 0.0f;
    out[1] = (float)o.flags;
    out[2] = o.refl.x;
    out[3] = o.refl.y;
    out[4] = o.refl.z;
    out[5] = o.wi.x;
    out[6] = o.wi.y;
    out[7] = o.wi.z;
    out[8] = o.fwd;
    out[9] = o.rev;
    out[10] = (float)o.medium;
    out[11] = o.ext.x;
    out[12] = o.ext.y;
    out[13] = o.ext.z;
    return ok;
}

/* ---- Lambert (corrected form; see the header comment) ----------------------------------------- */

void orc_lambert_eval(const orc_surface* s, const float* vv, const float* ll, float* out)
{
    ov3 n = s->n;
    float nl = dot(n, v3(ll[0], ll[1], ll[2])), nv = dot(n, v3(vv[0], vv[1], vv[2]));
    out[0] = out[1] = out[2] = out[3] = out[4] = 0.0f;
    if (nl <= 0.0f || nv <= 0.0f) return;
    ov3 f = scl(s->base_color, (1.0f / kPi) * nl);
    out[0] = f.x;
    out[1] = f.y;
    out[2] = f.z;
    out[3] = nl * (1.0f / kPi);
    out[4] = nv * (1.0f / kPi);
}

int orc_lambert_sample(orc_uniforms* u, const orc_surface* s, const float* vv, float* out)
{
    ov3 wo = to_local(s, v3(vv[0], vv[1], vv[2]));
    float r0 = orc_uniform(u), r1 = orc_uniform(u);
    float r = sqrtf(r0);
    float theta = (kPi * 2.0f) * r1;
    ov3 wi = v3(r * cos_f(theta), sqrtf(maxf(0.0f, 1 - r0)), r * sin_f(theta));
    ov3 w = normalize(to_world(s, wi));
    memset(out, 0, 14 * sizeof(float));
    out[0] = 1.0f;
    out[1] = 1.0f;
    out[2] = s->base_color.x;
    out[3] = s->base_color.y;
    out[4] = s->base_color.z;
    out[5] = w.x;
    out[6] = w.y;
    out[7] = w.z;
    out[8] = cos_theta(wi) * (1.0f / kPi);
    out[9] = maxf(cos_theta(wo), 0.0f) * (1.0f / kPi);
    return 1;
}

uint32_t orc_pcg32_next(uint64_t* state, uint64_t inc)
{
    uint64_t old = *state;
    *state = old * 6364136223846793005ULL + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

void orc_pcg32_seed(uint64_t initstate, uint64_t initseq, uint64_t* state, uint64_t* inc)
{
    *state = 0u;
    *inc = (initseq << 1u) | 1u;
    orc_pcg32_next(state, *inc);
    *state += initstate;
    orc_pcg32_next(state, *inc);
}

float orc_uniform(orc_uniforms* u)
{
    if (u->pcg_state) return (float)(orc_pcg32_next(u->pcg_state, u->pcg_inc) >> 8) * (1.0f / 16777216.0f);
    if (u->pos < u->count) return u->values[u->pos++];
    u->pos++;
    return 0.5f;
}

/* ---- CMJ, camera, lights, disc ---------------------------------------------------------------- */

static float cmj_randfloat(uint32_t i, uint32_t p)
{
    i ^= p;
    i ^= i >> 17;
    i ^= i >> 10;
    i *= 0xb36534e5u;
    i ^= i >> 12;
    i ^= i >> 21;
    i *= 0x93fc4795u;
    i ^= 0xdf6e307fu;
    i ^= i >> 17;
    i *= 1u | p >> 18;
    return (float)i * (1.0f / 4294967808.0f);
}
static uint32_t cmj_permute(uint32_t i, uint32_t l, uint32_t p)
{
    uint32_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {
        i ^= p;
        i *= 0xe170893du;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3fu;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1u | p >> 27;
        i *= 0x6935fa69u;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303u;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3u;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfu;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}
void orc_cmj(int32_t s, int32_t m, int32_t n, int32_t p, float* out)
{
    uint32_t up = (uint32_t)p;
    int sx = (int)cmj_permute((uint32_t)(s % m), (uint32_t)m, up * 0xa511e9b3u);
    int sy = (int)cmj_permute((uint32_t)(s / m), (uint32_t)n, up * 0x63d83595u);
    float jx = cmj_randfloat((uint32_t)s, up * 0xa399d265u);
    float jy = cmj_randfloat((uint32_t)s, up * 0x711ad6a5u);
    out[0] = ((float)(s % m) + ((float)sy + jx) / (float)n) / (float)m;
    out[1] = ((float)(s / m) + ((float)sx + jy) / (float)m) / (float)n;
}

void orc_camera_init(const float* cam, uint32_t width, uint32_t height, orc_camera* c)
{
    float widthf = (float)width, heightf = (float)height;
    float aspect = widthf / heightf;
    float hfov = cam[9] * (kPi / 180.f);
    float hlen = tanf(hfov * 0.5f);
    float vfov = 2 * atanf(hlen / aspect);
    float vlen = tanf(vfov * 0.5f);
    ov3 pos = v3(cam[0], cam[1], cam[2]), look = v3(cam[3], cam[4], cam[5]), up = v3(cam[6], cam[7], cam[8]);
    ov3 fwd = normalize(sub(look, pos));
    ov3 right = normalize(cross(up, fwd));
    c->x = scl(right, hlen);
    c->y = scl(cross(fwd, right), vlen);
    c->z = fwd;
    c->pos = pos;
    c->forward = normalize(sub(look, pos));
    c->width = width;
    c->height = height;
    c->vipd = widthf / (2.0f * tanf(hfov));
}

void orc_camera_dir(const orc_camera* c, float x, float y, float* out)
{
    float cx = ((x / (float)c->width) * 2.0f - 1.0f);
    float cy = 1.0f - 2.0f * y / (float)c->height;
    ov3 d = normalize(add(sub(c->z, scl(c->x, cx)), scl(c->y, cy)));
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

void orc_camera_cmj_dir(const orc_camera* c, int32_t x, int32_t y, int32_t s, int32_t m, int32_t n, int32_t p,
                        float* out)
{
    float j[2];
    orc_cmj(s, m, n, p, j);
    orc_camera_dir(c, (float)x + j[0], (float)y + j[1], out);
}

void orc_rect_light(const float* q, float u, float v, float* out)
{
    ov3 s = v3(q[0], q[1], q[2]), ex = v3(q[3], q[4], q[5]), ey = v3(q[6], q[7], q[8]), o = v3(q[9], q[10], q[11]);
    float exl = length(ex), eyl = length(ey);
    ov3 x = scl(ex, 1.0f / exl), y = scl(ey, 1.0f / eyl), z = cross(x, y);
    ov3 d = sub(s, o);
    float z0 = dot(d, z);
    if (z0 > 0) {
        z = neg(z);
        z0 *= -1;
    }
    float z0sq = z0 * z0;
    float x0 = dot(d, x), y0 = dot(d, y);
    float x1 = x0 + exl, y1 = y0 + eyl;
    float y0sq = y0 * y0, y1sq = y1 * y1;
    ov3 v00 = v3(x0, y0, z0), v01 = v3(x0, y1, z0), v10 = v3(x1, y0, z0), v11 = v3(x1, y1, z0);
    ov3 n0 = normalize(cross(v00, v10)), n1 = normalize(cross(v10, v11));
    ov3 n2 = normalize(cross(v11, v01)), n3 = normalize(cross(v01, v00));
    float g0 = acos_f(-dot(n0, n1)), g1 = acos_f(-dot(n1, n2)), g2 = acos_f(-dot(n2, n3)), g3 = acos_f(-dot(n3, n0));
    float b0 = n0.z, b1 = n2.z, b0sq = b0 * b0;
    float k = 2 * kPi - g2 - g3;
    float S = g0 + g1 - k;
    float au = u * S + k;
    float fu = (cos_f(au) * b0 - b1) / sin_f(au);
    float cu = 1 / sqrtf(fu * fu + b0sq) * (fu > 0 ? +1 : -1);
    cu = clampf(cu, -1.0f, 1.0f);
    float xu = -(cu * z0) / sqrtf(1 - cu * cu);
    xu = clampf(xu, x0, x1);
    float dd = sqrtf(xu * xu + z0sq);
    float h0 = y0 / sqrtf(dd * dd + y0sq);
    float h1 = y1 / sqrtf(dd * dd + y1sq);
    float hv = h0 + v * (h1 - h0), hv2 = hv * hv;
    const float small_eps = 1e-10f; /* SmallFloatEpsilon_ (Source/Core/SystemLib/BasicTypes.h:33) */
    float yv = (hv2 < 1 - small_eps) ? (hv * dd) / sqrtf(1 - hv2) : y1;
    ov3 p = add(add(add(o, scl(x, xu)), scl(y, yv)), scl(z, z0));
    out[0] = S;
    out[1] = z0;
    out[2] = p.x;
    out[3] = p.y;
    out[4] = p.z;
}

int orc_ray_quad(const float* q)
{
    const float eps = 0.00001f; /* IntersectionEpsilon_ (GeometryIntersection.cpp:16) */
    ov3 o = v3(q[0], q[1], q[2]), d = v3(q[3], q[4], q[5]);
    ov3 v00 = v3(q[6], q[7], q[8]), v10 = v3(q[9], q[10], q[11]), v01 = v3(q[12], q[13], q[14]),
        v11 = v3(q[15], q[16], q[17]);
    ov3 e01 = sub(v10, v00), e03 = sub(v01, v00);
    ov3 p = cross(d, e03);
    float det = dot(e01, p);
    if (absf(det) < eps) return 0;
    ov3 t = sub(o, v00);
    float alpha = dot(t, p) / det;
    if (alpha < 0 || alpha > 1) return 0;
    ov3 qq = cross(t, e01);
    float beta = dot(d, qq) / det;
    if (beta < 0 || beta > 1) return 0;
    float tau = dot(e03, qq) / det;
    if (tau < 0) return 0;
    if ((alpha + beta) > 1) {
        ov3 e23 = sub(v01, v11), e21 = sub(v10, v11);
        ov3 pp = cross(d, e21);
        float detp = dot(e23, pp);
        if (absf(detp) < eps) return 0;
        ov3 tp = sub(o, v11);
        float alphap = dot(tp, pp) / detp;
        if (alphap < 0) return 0;
        ov3 qp = cross(tp, e23);
        float betap = dot(d, qp) / detp;
        if (betap < 0.0f) return 0;
    }
    return 1;
}

void orc_concentric_disc(float r0, float r1, float* out)
{
    float phi, r;
    float a = 2.0f * r0 - 1.0f, b = 2.0f * r1 - 1.0f;
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
            if (absf(b) > 0.0f)
                phi = (kPi / 4.0f) * (6.0f - (a / b));
            else
                phi = 0;
        }
    }
    out[0] = r * cos_f(phi);
    out[1] = r * sin_f(phi);
}

/* ---- batch forms (one call per array; used by the tests) ------------------------------------- */

void orc_bsdf_eval_batch(const float* blocks, const float* n, const float* v, const float* l, uint32_t count,
                         float* out)
{
    for (uint32_t i = 0; i < count; ++i) {
        orc_surface s;
        orc_surface_from_block(blocks + 22 * (size_t)i, n + 3 * (size_t)i, &s);
        if (s.shader == 2)
            orc_lambert_eval(&s, v + 3 * (size_t)i, l + 3 * (size_t)i, out + 5 * (size_t)i);
        else
            orc_disney_eval(&s, v + 3 * (size_t)i, l + 3 * (size_t)i, s.shader == 0, out + 5 * (size_t)i);
    }
}

void orc_bsdf_sample_batch(const float* blocks, const float* n, const float* v, const float* uniforms,
                           uint32_t n_uniforms, uint32_t count, float* out)
{
    for (uint32_t i = 0; i < count; ++i) {
        orc_surface s;
        orc_surface_from_block(blocks + 22 * (size_t)i, n + 3 * (size_t)i, &s);
        orc_uniforms u = {uniforms + (size_t)n_uniforms * i, n_uniforms, 0, NULL, 0};
        if (s.shader == 2)
            orc_lambert_sample(&u, &s, v + 3 * (size_t)i, out + 14 * (size_t)i);
        else
            orc_disney_sample(&u, &s, v + 3 * (size_t)i, s.shader == 0, out + 14 * (size_t)i);
    }
}
