// ORACLE -- test infrastructure only. Compiled together with Selas's own source files (listed in
// oracle/ref.mk, taken from /root/reference where they lie) into oracle/_ref/libselas_ref.so.
// This file is this project's code: a flat extern "C" surface over the reference functions so
// that tests/golden/make_golden.py can record their outputs as fixtures. It restates nothing; every
// value it returns is computed by the reference's own compiled code.
#include "GeometryLib/Camera.h"
#include "Ptexture.h"
#include "GeometryLib/CoordinateSystem.h"
#include "GeometryLib/Disc.h"
#include "GeometryLib/RectangulerLightSampler.h"
#include "MathLib/CorrelatedMultiJitter.h"
#include "MathLib/FloatFuncs.h"
#include "MathLib/GeometryIntersection.h"
#include "MathLib/Sampler.h"
#include "Shading/Disney.h"
#include "Shading/Fresnel.h"
#include "Shading/Ggx.h"
#include "Shading/Scattering.h"
#include "Shading/SurfaceParameters.h"
#include "VCMCommon.h"
#include "VCMHashGrid.h"

#include <stdint.h>

// stb_image_write's HDR writer (what StbImageWrite calls for FrameBuffer_Save) and stb_image's
// reader, from the reference's own copies (Middleware/stb)
#define STB_IMAGE_WRITE_IMPLEMENTATION
#include "stb/stb_image_write.h"
#define STB_IMAGE_IMPLEMENTATION
#include "stb/stb_image.h"

using namespace Selas;

namespace {

// Surface block layout shared with tests/golden/make_golden.py (22 floats):
//  0..2 baseColor, 3..5 transmittanceColor, 6 sheen, 7 sheenTint, 8 clearcoat, 9 clearcoatGloss,
//  10 metallic, 11 specTrans, 12 diffTrans, 13 flatness, 14 anisotropic, 15 relativeIOR,
//  16 specularTint, 17 roughness, 18 scatterDistance, 19 ior, 20 shader (0 thin, 1 solid), 21 flags
// The tangent frame is built from the shading normal n exactly as CalculateSurfaceParams does
// without tangent data (SurfaceParameters.cpp:240-250, 283-285).
void fill_surface(const float* p, const float* n, SurfaceParameters& s)
{
    float3 nn = float3(n[0], n[1], n[2]);
    float3 t, b;
    MakeOrthogonalCoordinateSystem(nn, &t, &b);
    float3x3 tangentToWorld = MakeFloat3x3(t, nn, b);
    s.worldToTangent = MatrixTranspose(tangentToWorld);
    s.position = float3(0.0f, 0.0f, 0.0f);
    s.error = 0.0f;
    s.view = float3(0.0f, 0.0f, 0.0f);
    s.baseColor = float3(p[0], p[1], p[2]);
    s.transmittanceColor = float3(p[3], p[4], p[5]);
    s.sheen = p[6];
    s.sheenTint = p[7];
    s.clearcoat = p[8];
    s.clearcoatGloss = p[9];
    s.metallic = p[10];
    s.specTrans = p[11];
    s.diffTrans = p[12];
    s.flatness = p[13];
    s.anisotropic = p[14];
    s.relativeIOR = p[15];
    s.specularTint = p[16];
    s.roughness = p[17];
    s.scatterDistance = p[18];
    s.ior = p[19];
    s.shader = p[20] == 0.0f ? eDisneyThin : eDisneySolid;
    s.materialFlags = (uint32)p[21];
    s.lightSetIndex = 0;
}

}  // namespace

extern "C" {

// EvaluateDisney (Disney.cpp:525): out = reflectance[3], forwardPdf, reversePdf
void ref_disney_eval(const float* params, const float* n, const float* v, const float* l, int thin, float* out)
{
    SurfaceParameters s;
    fill_surface(params, n, s);
    float fwd = 0.0f, rev = 0.0f;
    float3 r = EvaluateDisney(s, float3(v[0], v[1], v[2]), float3(l[0], l[1], l[2]), thin != 0, fwd, rev);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = fwd;
    out[4] = rev;
}

// SampleDisney (Disney.cpp:617) with a CSampler seeded by `seed`. A second CSampler with the same
// seed records the first 8 uniforms the sample can consume (uniforms_out), so a restatement can
// be fed the identical stream. out = success, flags, reflectance[3], wi[3], fwdPdf, revPdf,
// medium phase function, extinction[3]
void ref_disney_sample(const float* params, const float* n, const float* v, int thin, uint32_t seed,
                       float* uniforms_out, float* out)
{
    SurfaceParameters s;
    fill_surface(params, n, s);
    CSampler rec;
    rec.Initialize(seed);
    for (int i = 0; i < 8; ++i) uniforms_out[i] = rec.UniformFloat();
    rec.Shutdown();
    CSampler sampler;
    sampler.Initialize(seed);
    BsdfSample bs;
    bs.flags = 0;
    bool ok = SampleDisney(&sampler, s, float3(v[0], v[1], v[2]), thin != 0, bs);
    sampler.Shutdown();
    out[0] = ok ? 1.0f : 0.0f;
    out[1] = (float)bs.flags;
    out[2] = bs.reflectance.x;
    out[3] = bs.reflectance.y;
    out[4] = bs.reflectance.z;
    out[5] = bs.wi.x;
    out[6] = bs.wi.y;
    out[7] = bs.wi.z;
    out[8] = bs.forwardPdfW;
    out[9] = bs.reversePdfW;
    out[10] = (float)bs.medium.phaseFunction;
    out[11] = bs.medium.extinction.x;
    out[12] = bs.medium.extinction.y;
    out[13] = bs.medium.extinction.z;
}

// GGX helpers (Ggx.cpp): out = D_aniso, G1_aniso(wo), vndf sample m[3], fwd pdf, rev pdf
void ref_ggx(const float* wo, const float* wi, const float* wm, float ax, float ay, float u1, float u2, float* out)
{
    float3 o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]), m(wm[0], wm[1], wm[2]);
    out[0] = Bsdf::GgxAnisotropicD(m, ax, ay);
    out[1] = Bsdf::SeparableSmithGGXG1(o, m, ax, ay);
    float3 s = Bsdf::SampleGgxVndfAnisotropic(o, ax, ay, u1, u2);
    out[2] = s.x;
    out[3] = s.y;
    out[4] = s.z;
    float f = 0.0f, r = 0.0f;
    Bsdf::GgxVndfAnisotropicPdf(i, m, o, ax, ay, f, r);
    out[5] = f;
    out[6] = r;
    out[7] = Bsdf::SeparableSmithGGXG1(o, 0.25f);
}

// Fresnel (Fresnel.cpp): out = Dielectric(cos, ni, nt), SchlickWeight(cos), SchlickR0FromRelativeIOR(eta),
// Schlick(r0 scalar, cos)
void ref_fresnel(float cos_i, float ni, float nt, float eta, float r0, float* out)
{
    out[0] = Fresnel::Dielectric(cos_i, ni, nt);
    out[1] = Fresnel::SchlickWeight(cos_i);
    out[2] = Fresnel::SchlickR0FromRelativeIOR(eta);
    out[3] = Fresnel::Schlick(r0, cos_i);
}

// CorrelatedMultiJitter (CorrelatedMultiJitter.cpp:53)
void ref_cmj(int32_t s, int32_t m, int32_t n, int32_t p, float* out)
{
    float2 r = CorrelatedMultiJitter(s, m, n, p);
    out[0] = r.x;
    out[1] = r.y;
}

// InitializeRayCastCamera (Camera.cpp:89) + JitteredCameraRay with CMJ (Camera.cpp:77).
// cam = position[3], lookAt[3], up[3], fovDegrees; px = (x, y, s, m, n, p) per ray; out = origin[3], dir[3]
// Also returns the camera's virtualImagePlaneDistance and basis in cam_out[10].
void ref_camera_rays(const float* cam, uint32_t width, uint32_t height, const int32_t* px, uint32_t count,
                     float* out, float* cam_out)
{
    CameraSettings cs;
    DefaultCameraSettings(&cs);
    cs.position = float3(cam[0], cam[1], cam[2]);
    cs.lookAt = float3(cam[3], cam[4], cam[5]);
    cs.up = float3(cam[6], cam[7], cam[8]);
    cs.fovDegrees = cam[9];
    RayCastCameraSettings rc;
    InitializeRayCastCamera(cs, width, height, rc);
    cam_out[0] = rc.cameraX.x;
    cam_out[1] = rc.cameraX.y;
    cam_out[2] = rc.cameraX.z;
    cam_out[3] = rc.cameraY.x;
    cam_out[4] = rc.cameraY.y;
    cam_out[5] = rc.cameraY.z;
    cam_out[6] = rc.cameraZ.x;
    cam_out[7] = rc.cameraZ.y;
    cam_out[8] = rc.cameraZ.z;
    cam_out[9] = rc.virtualImagePlaneDistance;
    for (uint32_t i = 0; i < count; ++i) {
        const int32_t* q = px + 6 * i;
        Ray r = JitteredCameraRay(&rc, q[0], q[1], q[2], q[3], q[4], q[5]);
        out[6 * i + 0] = r.origin.x;
        out[6 * i + 1] = r.origin.y;
        out[6 * i + 2] = r.origin.z;
        out[6 * i + 3] = r.direction.x;
        out[6 * i + 4] = r.direction.y;
        out[6 * i + 5] = r.direction.z;
    }
}

// InitializeRectangleLightSampler + SampleRectangleLight (RectangulerLightSampler.cpp:17,75).
// q = s[3], eX[3], eY[3], o[3]; out = S, z0, point[3]
void ref_rect_light(const float* q, float u, float v, float* out)
{
    RectangleLightSampler rs;
    InitializeRectangleLightSampler(float3(q[0], q[1], q[2]), float3(q[3], q[4], q[5]), float3(q[6], q[7], q[8]),
                                    float3(q[9], q[10], q[11]), rs);
    float3 p = SampleRectangleLight(rs, u, v);
    out[0] = rs.S;
    out[1] = rs.z0;
    out[2] = p.x;
    out[3] = p.y;
    out[4] = p.z;
}

// Intersection::RayQuad (GeometryIntersection.cpp:183); q = o[3], d[3], v00, v10, v01, v11
int ref_ray_quad(const float* q)
{
    return Intersection::RayQuad(float3(q[0], q[1], q[2]), float3(q[3], q[4], q[5]), float3(q[6], q[7], q[8]),
                                 float3(q[9], q[10], q[11]), float3(q[12], q[13], q[14]),
                                 float3(q[15], q[16], q[17]))
               ? 1
               : 0;
}

// SampleConcentricDisc (Disc.cpp:13)
void ref_concentric_disc(float r0, float r1, float* out)
{
    float2 d = SampleConcentricDisc(r0, r1);
    out[0] = d.x;
    out[1] = d.y;
}

// VCMCommon's iteration constants are in a file that cannot be linked here (it needs the scene and
// IBL code); the hash grid is self-contained. BuildHashGrid (VCMHashGrid.cpp:48) over `count`
// points, then SearchHashGrid (VCMHashGrid.cpp:107) for each query. For query i, the found point
// indices are written to found[i*max_found ...] (in visiting order, duplicates kept) and the count
// to found_count[i]. Also returns the grid's cell range ends (cell_ends, cellCount entries) and
// the point order (cell_indices, count entries).
struct GatherCtx {
    const VCMVertex* base;
    uint32_t* dst;
    uint32_t cap;
    uint32_t n;
};

static void gather_cb(const VCMVertex& v, void* user)
{
    GatherCtx* g = (GatherCtx*)user;
    uint32_t idx = (uint32_t)(&v - g->base);
    if (g->n < g->cap) g->dst[g->n] = idx;
    ++g->n;
}

void ref_hash_grid(const float* points, uint32_t count, uint32_t cell_count, float radius, const float* queries,
                   uint32_t n_queries, uint32_t max_found, uint32_t* found, uint32_t* found_count,
                   uint32_t* cell_ends, uint32_t* cell_indices, float* box_out)
{
    CArray<VCMVertex> verts;
    verts.Resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        VCMVertex& v = verts[i];
        v.hit.position = float3(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
    }
    VCMHashGrid grid;
    BuildHashGrid(&grid, cell_count, radius, verts);
    for (uint32_t c = 0; c < cell_count; ++c) cell_ends[c] = grid.cellRangeEnds[c];
    for (uint32_t i = 0; i < count; ++i) cell_indices[i] = grid.cellIndices[i];
    box_out[0] = grid.aaBox.min.x;
    box_out[1] = grid.aaBox.min.y;
    box_out[2] = grid.aaBox.min.z;
    box_out[3] = grid.aaBox.max.x;
    box_out[4] = grid.aaBox.max.y;
    box_out[5] = grid.aaBox.max.z;
    for (uint32_t q = 0; q < n_queries; ++q) {
        GatherCtx g = {&verts[0], found + (size_t)q * max_found, max_found, 0};
        SearchHashGrid(&grid, verts, float3(queries[3 * q], queries[3 * q + 1], queries[3 * q + 2]), &g, gather_cb);
        found_count[q] = g.n;
    }
    ShutdownHashGrid(&grid);
    verts.Shutdown();
}

// PtexFilter{f_bspline}.eval as CalculateSurfaceParams calls it (SurfaceParameters.cpp:196-203):
// 3 channels from channel 0, zero filter footprint. Returns -1 when the file cannot be opened.
int ref_ptex_eval(const char* path, uint32_t n, const int32_t* faces, const float* u, const float* v, float* out)
{
    Ptex::String err;
    PtexTexture* tex = PtexTexture::open(path, err);
    if (!tex) return -1;
    PtexFilter::Options opts(PtexFilter::f_bspline);
    PtexFilter* filter = PtexFilter::getFilter(tex, opts);
    for (uint32_t i = 0; i < n; ++i) filter->eval(out + 3 * (size_t)i, 0, 3, faces[i], u[i], v[i], 0, 0, 0, 0);
    filter->release();
    tex->release();
    return 0;
}

// stbi_write_hdr as StbImageWrite calls it (StbImage.cpp:78): 3 channels, rows top first
int ref_write_hdr(const char* path, int w, int h, const float* rgb) { return stbi_write_hdr(path, w, h, 3, rgb); }

// stbi_loadf of an .hdr file: 3 channels into out (at most cap floats); returns w * h, or -1
int ref_read_hdr(const char* path, float* out, int cap)
{
    int w = 0, h = 0, comp = 0;
    float* data = stbi_loadf(path, &w, &h, &comp, 3);
    if (!data) return -1;
    const int n = w * h;
    for (int i = 0; i < 3 * n && i < cap; ++i) out[i] = data[i];
    stbi_image_free(data);
    return n;
}

}  // extern "C"
