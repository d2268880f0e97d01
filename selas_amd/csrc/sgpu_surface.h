// Hit record -> shading record on the device: CalculateSurfaceParams
// (Source/Core/Shading/SurfaceParameters.cpp:144-247). The shading normal is the normalised
// interpolated vertex normal of the hit geometry (rtcInterpolate0 of attribute slot 0), else the
// normalised geometric normal; the tangent frame is the interpolated tangent (slot 1) and
// cross(n, t) times its w, else MakeOrthogonalCoordinateSystem(n); the material comes from the hit
// geometry's material record (MaterialResourceData, Source/Core/SceneLib/ModelResource.h:68-88)
// with the same transforms: baseColor^2.2, saturated specTrans and metallic, diffTrans halved, and
// relativeIOR = ior for a transparent material hit from behind, else 1/ior.
#pragma once

#include "sgpu_shade.h"

namespace sgpu {

// Material record: 22 floats with the layout of the surface block in include/selas_gpu.h, holding
// the material's own values (before CalculateSurfaceParams's transforms); [15] is unused (the
// relative IOR is computed per hit) and [21] holds the material flags. The device keeps a shading
// record per material instead (shading_record): the same layout with CalculateSurfaceParams's
// per-material transforms already applied and [15] holding the continuation probability.
constexpr uint32_t kMatTransparent = 1u << 0;  // eTransparent (ModelResource.h:38)
constexpr uint32_t kMatUsesPtex = 1u << 4;     // eUsesPtex (ModelResource.h:42)

// The hit error bound of RayPick (Source/Applications/Selas/Source/PathTracer.cpp:134-135).
SGPU_HD float hit_error(f3 p, float t)
{
    const float kErr = 32.0f * 1.19209e-07f;
    return kErr * max_(max_(absf_(p.x), absf_(p.y)), max_(absf_(p.z), t));
}

// Shading record of a raw material record: baseColor^2.2, saturated metallic and specTrans,
// halved diffTrans (SurfaceParameters.cpp:197-229), and [15] = the Russian-roulette continuation
// probability of the untextured material (saturate(max(baseColor^2.2))).
SGPU_HD void shading_record(const float* __restrict raw, float* __restrict out)
{
    for (int k = 0; k < 22; ++k) out[k] = raw[k];
    out[0] = pow_f(raw[0], 2.2f);
    out[1] = pow_f(raw[1], 2.2f);
    out[2] = pow_f(raw[2], 2.2f);
    out[10] = saturate_(raw[10]);
    out[11] = saturate_(raw[11]);
    out[12] = raw[12] * 0.5f;
    out[15] = saturate_(maxc3(mk3(out[0], out[1], out[2])));
}

// Shading frame (n, t, b) and material of a surface point seen from `view` (the unit direction
// toward the previous path vertex), from the material's shading record. The base colour is the
// record's unless a texel (already filtered, before the 2.2 power) is given (has_texel). The texel
// travels by value: a pointer chosen between a local and NULL would keep the local in scratch memory.
SGPU_HD void surface_material_frame(const float* __restrict mat, f3 n, f3 t, f3 b, f3 view, bool has_texel, f3 texel,
                                    Surface& s)
{
    s.n = n;
    s.t = t;
    s.b = b;
    s.base_color = has_texel ? mk3(pow_f(texel.x, 2.2f), pow_f(texel.y, 2.2f), pow_f(texel.z, 2.2f))
                             : mk3(mat[0], mat[1], mat[2]);
    s.transmittance_color = mk3(mat[3], mat[4], mat[5]);
    s.sheen = mat[6];
    s.sheen_tint = mat[7];
    s.clearcoat = mat[8];
    s.clearcoat_gloss = mat[9];
    s.metallic = mat[10];
    s.spec_trans = mat[11];
    s.diff_trans = mat[12];
    s.flatness = mat[13];
    s.anisotropic = mat[14];
    s.specular_tint = mat[16];
    s.roughness = mat[17];
    s.scatter_distance = mat[18];
    s.ior = mat[19];
    s.shader = (uint32_t)mat[20];
    s.flags = (uint32_t)mat[21];
    s.relative_ior = ((s.flags & kMatTransparent) && dot3(view, s.n) < 0.0f) ? s.ior : 1.0f / s.ior;
}

// The same with the frame of make_frame(n) (MakeOrthogonalCoordinateSystem).
SGPU_HD void surface_material(const float* __restrict mat, f3 n, f3 view, Surface& s)
{
    f3 t, b;
    make_frame(n, t, b);
    surface_material_frame(mat, n, t, b, view, false, mk3(0, 0, 0), s);
}


// Vertex attributes of one geometry (the rtcSetSharedGeometryBuffer(RTC_BUFFER_TYPE_VERTEX_ATTRIBUTE)
// slots of ModelResource.cpp:164-200 that CalculateSurfaceParams interpolates): per primitive its
// corners' values in the primitive's own vertex order (3 corners per triangle, 4 per quad): normals
// 3 floats, tangents 4 (xyz, w), UVs 2 per corner; NULL where the geometry has none.
struct GeomAttr {
    const float* n;
    const float* t;
    const float* uv;
    uint32_t corners;  // 3 or 4
    uint32_t pad;
};
static_assert(sizeof(GeomAttr) == 32, "GeomAttr must be 32 bytes");

// rtcInterpolate0 of a K-float attribute at (prim, u, v), as Embree 3 evaluates it: a triangle
// (p0, p1, p2) gives fma(w, p0, fma(u, p1, v p2)) with w = 1 - u - v; a quad (p0, p1, p2, p3) is
// evaluated in the triangle of its split that holds (u, v): (p0, p1, p3) with (u, v) when
// u + v <= 1, else (p2, p3, p1) with (1 - u, 1 - v).
template <int K>
SGPU_HD void interp_attr(const float* __restrict a, uint32_t corners, uint32_t prim, float u, float v, float* out)
{
    const size_t base = (size_t)prim * corners;
    size_t i0 = base, i1 = base + 1, i2 = base + 2;
    float U = u, V = v;
    if (corners == 4) {
        if (u + v <= 1.0f) {
            i2 = base + 3;
        } else {
            i0 = base + 2;
            i1 = base + 3;
            i2 = base + 1;
            U = 1.0f - u;
            V = 1.0f - v;
        }
    }
    const float W = 1.0f - U - V;
    for (int k = 0; k < K; ++k) out[k] = fmaf(W, a[i0 * K + k], fmaf(U, a[i1 * K + k], V * a[i2 * K + k]));
}

// Shading normal and tangent frame of a hit (SurfaceParameters.cpp:155-180) from the geometry's
// attributes (ga NULL or without normals: the normalised geometric normal ng, already in world
// space). The interpolated normal and tangent are carried to world space by the linear map l2w
// (row-major 3x3; NULL: identity); with tangents, t is the tangent's xyz and b = cross(n, t) * w,
// used as they are (the reference neither normalises nor orthogonalises them); without, the frame
// of make_frame(n). Returns true when the frame came from tangents (then tw holds w).
SGPU_HD bool hit_frame(const GeomAttr* ga, uint32_t prim, float u, float v, f3 ng, const float* l2w, f3& n, f3& t,
                       f3& b, float& tw)
{
    if (ga && ga->n) {
        float a[3];
        interp_attr<3>(ga->n, ga->corners, prim, u, v, a);
        f3 nn = mk3(a[0], a[1], a[2]);
        if (l2w) nn = xform_linear(l2w, nn);
        n = normalize3(nn);
    } else {
        n = normalize3(ng);
    }
    if (ga && ga->t) {
        float a[4];
        interp_attr<4>(ga->t, ga->corners, prim, u, v, a);
        t = mk3(a[0], a[1], a[2]);
        if (l2w) t = xform_linear(l2w, t);
        tw = a[3];
        b = cross3(n, t) * tw;
        return true;
    }
    make_frame(n, t, b);
    return false;
}

// Surface of the hit at distance t along (o, d) on primitive prim at (u, v) of a geometry with
// attributes ga (may be NULL) and geometric normal ng (world space); l2w as for hit_frame. Returns
// true when the frame came from the geometry's tangents; then *tangent (may be NULL) receives
// (t, w), from which the frame is rebuilt with the same n (tangent_frame).
SGPU_HD bool surface_from_hit(f3 o, f3 d, float t, f3 ng, const GeomAttr* ga, uint32_t prim, float u, float v,
                              const float* l2w, const float* __restrict mat, bool has_texel, f3 texel, Surface& s,
                              float4* tangent = nullptr)
{
    s.position = mk3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
    s.error = hit_error(s.position, t);
    f3 n, tt, b;
    float tw = 0.0f;
    const bool tf = hit_frame(ga, prim, u, v, ng, l2w, n, tt, b, tw);
    if (tf && tangent) *tangent = make_float4(tt.x, tt.y, tt.z, tw);
    surface_material_frame(mat, n, tt, b, -d, has_texel, texel, s);
    return tf;
}

// The frame of a stored vertex with shading normal n and stored tangent (t, w), as hit_frame built it.
SGPU_HD void tangent_frame(f3 n, float4 tangent, Surface& s)
{
    s.t = mk3(tangent.x, tangent.y, tangent.z);
    s.b = cross3(n, s.t) * tangent.w;
}


}  // namespace sgpu
