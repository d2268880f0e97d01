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
#include "IoLib/BinaryStreamSerializer.h"
#include "SceneLib/ImageBasedLightResource.h"
#include "SceneLib/ModelResource.h"
#include "SceneLib/SceneResource.h"
#include "SceneLib/SubsceneResource.h"
#include "UtilityLib/BinarySearch.h"
#include "UtilityLib/MurmurHash.h"

#include <stdint.h>
#include <string.h>

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

// ---- resource files: reference-typed objects filled field by field from flat arrays, then written
// by the reference's SerializeToBinary (BinaryStreamSerializer.h:14-33) over its own Serialize
// functions (ModelResource.cpp:36-101, SubsceneResource.cpp:37-46, SceneResource.cpp:37-66,
// ImageBasedLightResource.cpp:23-43). Fixed strings are copied whole (every byte of the buffer is
// serialized, the terminator's tail included).
template <int T>
void put_str(FixedString<T>& s, const char* src)
{
    memcpy(s.str, src, T);
}

template <class E>
void sized(CArray<E>& a, uint64_t count, uint64_t capacity)
{
    a.Reserve(capacity);
    a.Resize(count);
}

float3 f3_at(const float* f) { return float3(f[0], f[1], f[2]); }

// camera: name (128 bytes); f = position[3], fovDegrees, lookAt[3], znear, up[3], zfar
void put_camera(CameraSettings& c, const char* name, const float* f)
{
    put_str(c.name, name);
    c.position = f3_at(f);
    c.fovDegrees = f[3];
    c.lookAt = f3_at(f + 4);
    c.znear = f[7];
    c.up = f3_at(f + 8);
    c.zfar = f[11];
}

// material: baseColorTexture (512 bytes); u = shader, flags; f = baseColor[3], transmittanceColor[3],
// the 14 scalar attributes in ScalarMaterialProperties order
void put_material(MaterialResourceData& m, const char* tex, const uint32_t* u, const float* f)
{
    put_str(m.baseColorTexture, tex);
    m.shader = (ShaderType)u[0];
    m.flags = u[1];
    m.baseColor = f3_at(f);
    m.transmittanceColor = f3_at(f + 3);
    for (int k = 0; k < eMaterialPropertyCount; ++k) m.scalarAttributeValues[k] = f[6 + k];
}

void put_instance(Instance& in, uint64_t index, const float* m16)
{
    in.index = index;
    memcpy(&in.localToWorld, m16, sizeof(float4x4));
}

template <class D>
uint64_t write_blob(D& obj, uint8_t* out, uint64_t cap)
{
    uint8* data = nullptr;
    Selas::uint size = 0;
    SerializeToBinary(obj, data, size);
    if (out && cap >= size) memcpy(out, data, size);
    FreeAligned_(data);
    return size;
}

template <class E>
void carray_at(const CArray<E>& a, const uint8_t* blob, uint64_t* out)
{
    out[0] = a.Count();
    out[1] = (uint64_t)((const uint8_t*)a.DataPointer() - blob);
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

// PtexTexture::open + getFaceInfo + getData(faceid, buffer, 0) (PtexReader.cpp:714-765): face
// `face`'s texels at its own resolution, pixels interleaved, rows along v, written to out (at most
// cap bytes); info = meshType, dataType, numChannels, numFaces, alphaChannel, ulog2, vlog2,
// adjedges packed, flags, adjfaces[4]. Returns the face's byte count, or -1 when the file cannot
// be opened.
int ref_ptex_face(const char* path, int32_t face, uint8_t* out, int64_t cap, int32_t* info)
{
    Ptex::String err;
    PtexTexture* tex = PtexTexture::open(path, err);
    if (!tex) return -1;
    const Ptex::FaceInfo& fi = tex->getFaceInfo(face);
    const int psize = Ptex::DataSize(tex->dataType()) * tex->numChannels();
    const int bytes = psize * fi.res.u() * fi.res.v();
    info[0] = tex->meshType();
    info[1] = tex->dataType();
    info[2] = tex->numChannels();
    info[3] = tex->numFaces();
    info[4] = tex->alphaChannel();
    info[5] = fi.res.ulog2;
    info[6] = fi.res.vlog2;
    info[7] = fi.adjedges;
    info[8] = fi.flags;
    for (int k = 0; k < 4; ++k) info[9 + k] = fi.adjfaces[k];
    if (bytes <= cap) tex->getData(face, out, 0);
    tex->release();
    return bytes;
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

// sizeof of the resource roots and of the records their arrays hold raw: ModelResourceData,
// ModelGeometryData, SubsceneResourceData, SceneResourceData, ImageBasedLightResourceData,
// CameraSettings, FilePathString, MaterialResourceData, MeshMetaData, CurveMetaData, Instance,
// SceneLight, SceneLightSetRange, CArray<uint32>
void ref_resource_sizes(uint64_t* out)
{
    const uint64_t s[] = {sizeof(ModelResourceData), sizeof(ModelGeometryData), sizeof(SubsceneResourceData),
                          sizeof(SceneResourceData), sizeof(ImageBasedLightResourceData), sizeof(CameraSettings),
                          sizeof(FilePathString), sizeof(MaterialResourceData), sizeof(MeshMetaData),
                          sizeof(CurveMetaData), sizeof(Instance), sizeof(SceneLight), sizeof(SceneLightSetRange),
                          sizeof(CArray<uint32>)};
    for (size_t i = 0; i < sizeof(s) / sizeof(s[0]); ++i) out[i] = s[i];
}

// A ModelResource file. box = min[3], max[3]; u4 = totalVertexCount, totalCurveVertexCount,
// curveModelName, pad; sizes = the eight geometry sizes; counts/caps = count and capacity of
// cameras, textureResourceNames, materials, materialHashes, meshes, curves; mesh_u = indexCount,
// indexOffset, vertexCount, vertexOffset, materialHash, indicesPerFace, nameHash per mesh, names 64
// bytes each; curve_u = indexOffset, indexCount. Returns the file's size (written to out when cap
// holds it).
uint64_t ref_model_resource_write(const float* box, const uint32_t* u4, const uint64_t* sizes, const uint64_t* counts,
                                  const uint64_t* caps, const char* cam_names, const float* cam_f,
                                  const char* tex_names, const char* mat_tex, const uint32_t* mat_u,
                                  const float* mat_f, const uint32_t* hashes, const uint32_t* mesh_u,
                                  const char* mesh_names, const uint32_t* curve_u, uint8_t* out, uint64_t cap)
{
    ModelResourceData d;
    d.aaBox.min = f3_at(box);
    d.aaBox.max = f3_at(box + 3);
    d.totalVertexCount = u4[0];
    d.totalCurveVertexCount = u4[1];
    d.curveModelName = u4[2];
    d.pad = u4[3];
    d.indexSize = sizes[0];
    d.faceIndexSize = sizes[1];
    d.positionSize = sizes[2];
    d.normalsSize = sizes[3];
    d.tangentsSize = sizes[4];
    d.uvsSize = sizes[5];
    d.curveIndexSize = sizes[6];
    d.curveVertexSize = sizes[7];
    sized(d.cameras, counts[0], caps[0]);
    for (uint64_t i = 0; i < counts[0]; ++i) put_camera(d.cameras[i], cam_names + 128 * i, cam_f + 12 * i);
    sized(d.textureResourceNames, counts[1], caps[1]);
    for (uint64_t i = 0; i < counts[1]; ++i) put_str(d.textureResourceNames[i], tex_names + 512 * i);
    sized(d.materials, counts[2], caps[2]);
    for (uint64_t i = 0; i < counts[2]; ++i)
        put_material(d.materials[i], mat_tex + 512 * i, mat_u + 2 * i, mat_f + 20 * i);
    sized(d.materialHashes, counts[3], caps[3]);
    for (uint64_t i = 0; i < counts[3]; ++i) d.materialHashes[i] = hashes[i];
    sized(d.meshes, counts[4], caps[4]);
    for (uint64_t i = 0; i < counts[4]; ++i) {
        MeshMetaData& m = d.meshes[i];
        const uint32_t* u = mesh_u + 7 * i;
        m.indexCount = u[0];
        m.indexOffset = u[1];
        m.vertexCount = u[2];
        m.vertexOffset = u[3];
        m.materialHash = u[4];
        m.indicesPerFace = u[5];
        m.nameHash = u[6];
        put_str(m.name, mesh_names + 64 * i);
    }
    sized(d.curves, counts[5], caps[5]);
    for (uint64_t i = 0; i < counts[5]; ++i) {
        d.curves[i].indexOffset = curve_u[2 * i];
        d.curves[i].indexCount = curve_u[2 * i + 1];
    }
    return write_blob(d, out, cap);
}

// A ModelGeometryResource file: the eight sizes (bytes) and data of indices, faceIndexCounts,
// positions, normals, tangents, uvs, curveIndices, curveVertices.
uint64_t ref_model_geometry_write(const uint64_t* sizes, const void* const* data, uint8_t* out, uint64_t cap)
{
    ModelGeometryData g;
    uint64* sz[8] = {&g.indexSize, &g.faceIndexSize, &g.positionSize, &g.normalsSize,
                     &g.tangentsSize, &g.uvsSize, &g.curveIndexSize, &g.curveVertexSize};
    void** p[8] = {(void**)&g.indices, (void**)&g.faceIndexCounts, (void**)&g.positions, (void**)&g.normals,
                   (void**)&g.tangents, (void**)&g.uvs, (void**)&g.curveIndices, (void**)&g.curveVertices};
    for (int k = 0; k < 8; ++k) {
        *sz[k] = sizes[k];
        *p[k] = (void*)data[k];
    }
    return write_blob(g, out, cap);
}

// A SubsceneResource file. counts/caps: modelNames, modelInstances, sceneMaterialNames,
// sceneMaterials; instances: index and 16 floats (localToWorld as stored) each.
uint64_t ref_subscene_write(const char* name, uint64_t light_set, const uint64_t* counts, const uint64_t* caps,
                            const char* model_names, const uint64_t* inst_index, const float* inst_m,
                            const uint32_t* mat_names, const char* mat_tex, const uint32_t* mat_u, const float* mat_f,
                            uint8_t* out, uint64_t cap)
{
    SubsceneResourceData d;
    put_str(d.name, name);
    d.lightSetIndex = light_set;
    sized(d.modelNames, counts[0], caps[0]);
    for (uint64_t i = 0; i < counts[0]; ++i) put_str(d.modelNames[i], model_names + 512 * i);
    sized(d.modelInstances, counts[1], caps[1]);
    for (uint64_t i = 0; i < counts[1]; ++i) put_instance(d.modelInstances[i], inst_index[i], inst_m + 16 * i);
    sized(d.sceneMaterialNames, counts[2], caps[2]);
    for (uint64_t i = 0; i < counts[2]; ++i) d.sceneMaterialNames[i] = mat_names[i];
    sized(d.sceneMaterials, counts[3], caps[3]);
    for (uint64_t i = 0; i < counts[3]; ++i)
        put_material(d.sceneMaterials[i], mat_tex + 512 * i, mat_u + 2 * i, mat_f + 20 * i);
    return write_blob(d, out, cap);
}

// A SceneResource file. bg = backgroundIntensity[4]; counts/caps: subsceneNames,
// subsceneInstances, lights, lightSetRanges, cameras; light_f = position, direction, x, z, radiance
// (15 floats) per light; ranges = start, count.
uint64_t ref_scene_write(const char* name, const char* ibl, const float* bg, const uint64_t* counts,
                         const uint64_t* caps, const char* sub_names, const uint64_t* inst_index, const float* inst_m,
                         const uint32_t* light_type, const float* light_f, const uint32_t* ranges,
                         const char* cam_names, const float* cam_f, uint8_t* out, uint64_t cap)
{
    SceneResourceData d;
    put_str(d.name, name);
    put_str(d.iblName, ibl);
    d.backgroundIntensity = float4(bg[0], bg[1], bg[2], bg[3]);
    sized(d.subsceneNames, counts[0], caps[0]);
    for (uint64_t i = 0; i < counts[0]; ++i) put_str(d.subsceneNames[i], sub_names + 512 * i);
    sized(d.subsceneInstances, counts[1], caps[1]);
    for (uint64_t i = 0; i < counts[1]; ++i) put_instance(d.subsceneInstances[i], inst_index[i], inst_m + 16 * i);
    sized(d.lights, counts[2], caps[2]);
    for (uint64_t i = 0; i < counts[2]; ++i) {
        SceneLight& l = d.lights[i];
        const float* f = light_f + 15 * i;
        l.type = light_type[i];
        l.position = f3_at(f);
        l.direction = f3_at(f + 3);
        l.x = f3_at(f + 6);
        l.z = f3_at(f + 9);
        l.radiance = f3_at(f + 12);
    }
    sized(d.lightSetRanges, counts[3], caps[3]);
    for (uint64_t i = 0; i < counts[3]; ++i) {
        d.lightSetRanges[i].start = ranges[2 * i];
        d.lightSetRanges[i].count = ranges[2 * i + 1];
    }
    sized(d.cameras, counts[4], caps[4]);
    for (uint64_t i = 0; i < counts[4]; ++i) put_camera(d.cameras[i], cam_names + 128 * i, cam_f + 12 * i);
    return write_blob(d, out, cap);
}

// An IBL file: density functions (height marginal and width x height conditional floats), miss
// map size, rotation, exposure scale, light and miss texels (3 floats each).
uint64_t ref_ibl_write(uint64_t width, uint64_t height, const float* marginal, const float* conditional,
                       uint64_t miss_width, uint64_t miss_height, float rotation, float exposure, const float* light,
                       const float* miss, uint8_t* out, uint64_t cap)
{
    ImageBasedLightResourceData d;
    d.densityfunctions.width = width;
    d.densityfunctions.height = height;
    d.densityfunctions.marginalDensityFunction = (float*)marginal;
    d.densityfunctions.conditionalDensityFunctions = (float*)conditional;
    d.missWidth = miss_width;
    d.missHeight = miss_height;
    d.rotationRadians = rotation;
    d.exposureScale = exposure;
    d.lightData = (float3*)light;
    d.missData = (float3*)miss;
    return write_blob(d, out, cap);
}

// AttachToBinary (BinaryStreamSerializer.h:35-44) of a resource file in place (kind 0 model,
// 1 geometry, 2 subscene, 3 scene, 4 IBL), then where the reference finds each array: (count,
// offset of its data from the file's start) per CArray, or the offset of each raw pointer (geometry:
// the 8 buffers; IBL: marginal, conditional, light, miss). Returns the number of values written.
int ref_attach_offsets(int kind, uint8_t* blob, uint64_t size, uint64_t* out)
{
    int n = 0;
    if (kind == 0) {
        ModelResourceData* d = nullptr;
        AttachToBinary(d, blob, size);
        carray_at(d->cameras, blob, out + 0);
        carray_at(d->textureResourceNames, blob, out + 2);
        carray_at(d->materials, blob, out + 4);
        carray_at(d->materialHashes, blob, out + 6);
        carray_at(d->meshes, blob, out + 8);
        carray_at(d->curves, blob, out + 10);
        n = 12;
    }
    else if (kind == 1) {
        ModelGeometryData* g = nullptr;
        AttachToBinary(g, blob, size);
        const void* p[8] = {g->indices, g->faceIndexCounts, g->positions, g->normals,
                            g->tangents, g->uvs, g->curveIndices, g->curveVertices};
        for (int k = 0; k < 8; ++k) out[k] = (uint64_t)((const uint8_t*)p[k] - blob);
        n = 8;
    }
    else if (kind == 2) {
        SubsceneResourceData* d = nullptr;
        AttachToBinary(d, blob, size);
        carray_at(d->modelNames, blob, out + 0);
        carray_at(d->modelInstances, blob, out + 2);
        carray_at(d->sceneMaterialNames, blob, out + 4);
        carray_at(d->sceneMaterials, blob, out + 6);
        n = 8;
    }
    else if (kind == 3) {
        SceneResourceData* d = nullptr;
        AttachToBinary(d, blob, size);
        carray_at(d->subsceneNames, blob, out + 0);
        carray_at(d->subsceneInstances, blob, out + 2);
        carray_at(d->lights, blob, out + 4);
        carray_at(d->lightSetRanges, blob, out + 6);
        carray_at(d->cameras, blob, out + 8);
        n = 10;
    }
    else if (kind == 4) {
        ImageBasedLightResourceData* d = nullptr;
        AttachToBinary(d, blob, size);
        const void* p[4] = {d->densityfunctions.marginalDensityFunction, d->densityfunctions.conditionalDensityFunctions,
                            d->lightData, d->missData};
        for (int k = 0; k < 4; ++k) out[k] = (uint64_t)((const uint8_t*)p[k] - blob);
        n = 4;
    }
    return n;
}

// SetupSceneCamera (SceneResource.cpp:336-358) on a SceneResource file attached in place: the
// RayCastCameraSettings basis and image-plane distance as ref_camera_rays reports them, then the
// camera's position.
void ref_scene_camera(uint8_t* blob, uint64_t size, uint64_t index, uint32_t width, uint32_t height, float* out)
{
    SceneResource scene;
    AttachToBinary(scene.data, blob, size);
    RayCastCameraSettings rc;
    SetupSceneCamera(&scene, index, width, height, rc);
    scene.data = nullptr;
    const float v[13] = {rc.cameraX.x, rc.cameraX.y, rc.cameraX.z, rc.cameraY.x, rc.cameraY.y, rc.cameraY.z,
                         rc.cameraZ.x, rc.cameraZ.y, rc.cameraZ.z, rc.virtualImagePlaneDistance,
                         rc.position.x, rc.position.y, rc.position.z};
    for (int k = 0; k < 13; ++k) out[k] = v[k];
}

// MurmurHash3_x86_32 (MurmurHash.cpp), the hash of names and material names in resource files
uint32_t ref_murmur3_32(const char* key, int32_t len, uint32_t seed) { return MurmurHash3_x86_32(key, len, seed); }

// BinarySearch (BinarySearch.h:11-35) over uint32 keys, as FindMeshMaterial calls it
// (ModelResource.cpp:213-229); (uint64)-1 when not found. Callers keep key <= the largest element:
// beyond it the reference reads data[(uint)-1].
uint64_t ref_binary_search(const uint32_t* data, uint64_t count, uint32_t key)
{
    return BinarySearch((uint32*)data, (Selas::uint)count, (uint32)key);
}

}  // extern "C"
