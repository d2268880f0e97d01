// Hit record -> shading record on the device: CalculateSurfaceParams
// (Source/Core/Shading/SurfaceParameters.cpp:144-247) for flattened triangle meshes without vertex
// normals, tangents or UVs: the shading normal is the normalised geometric normal, the frame comes
// from MakeOrthogonalCoordinateSystem, and the material comes from the hit geometry's material
// record (MaterialResourceData, Source/Core/SceneLib/ModelResource.h:68-88) with the same
// transforms: baseColor^2.2, saturated specTrans and metallic, diffTrans halved, and
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

// Shading frame and material of a surface point with shading normal `n` seen from `view` (the
// unit direction toward the previous path vertex), from the material's shading record. The base
// colour is the record's unless `texel` (already filtered, before the 2.2 power) is given.
SGPU_HD void surface_material(const float* __restrict mat, f3 n, f3 view, const f3* texel, Surface& s)
{
    s.n = n;
    make_frame(s.n, s.t, s.b);
    s.base_color = texel ? mk3(pow_f(texel->x, 2.2f), pow_f(texel->y, 2.2f), pow_f(texel->z, 2.2f))
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

// Surface of the hit at distance t along (o, d) with geometric normal ng.
SGPU_HD void surface_from_hit(f3 o, f3 d, float t, f3 ng, const float* __restrict mat, const f3* texel, Surface& s)
{
    s.position = mk3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
    s.error = hit_error(s.position, t);
    surface_material(mat, normalize3(ng), -d, texel, s);
}

}  // namespace sgpu
