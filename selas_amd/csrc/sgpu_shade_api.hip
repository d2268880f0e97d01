// Batched BSDF entry points: the device replacements of EvaluateBsdf / SampleBsdfFunction
// (Source/Core/Shading/SurfaceScattering.cpp:18-60) over arrays of surface points.
#include "sgpu_internal.h"
#include "sgpu_shade.h"

namespace {

using namespace sgpu;

constexpr int kBlock = 256;

// Surface block: 22 floats, layout documented in include/selas_gpu.h (sgpu_bsdf_eval).
__device__ __forceinline__ Surface surface_from_block(const float* __restrict p, const float* __restrict n)
{
    Surface s;
    s.position = mk3(0, 0, 0);
    s.error = 0.0f;
    s.n = mk3(n[0], n[1], n[2]);
    make_frame(s.n, s.t, s.b);
    s.base_color = mk3(p[0], p[1], p[2]);
    s.transmittance_color = mk3(p[3], p[4], p[5]);
    s.sheen = p[6];
    s.sheen_tint = p[7];
    s.clearcoat = p[8];
    s.clearcoat_gloss = p[9];
    s.metallic = p[10];
    s.spec_trans = p[11];
    s.diff_trans = p[12];
    s.flatness = p[13];
    s.anisotropic = p[14];
    s.relative_ior = p[15];
    s.specular_tint = p[16];
    s.roughness = p[17];
    s.scatter_distance = p[18];
    s.ior = p[19];
    s.shader = (uint32_t)p[20];
    s.flags = (uint32_t)p[21];
    return s;
}

__global__ __launch_bounds__(kBlock) void k_bsdf_eval(const float* __restrict blocks, const float* __restrict nrm,
                                                      const float* __restrict v, const float* __restrict l,
                                                      uint32_t count, float* __restrict out)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    Surface s = surface_from_block(blocks + 22 * (size_t)i, nrm + 3 * (size_t)i);
    float fwd, rev;
    f3 r = bsdf_eval(s, mk3(v[3 * i], v[3 * i + 1], v[3 * i + 2]), mk3(l[3 * i], l[3 * i + 1], l[3 * i + 2]), fwd, rev);
    float* o = out + 5 * (size_t)i;
    o[0] = r.x;
    o[1] = r.y;
    o[2] = r.z;
    o[3] = fwd;
    o[4] = rev;
}

__global__ __launch_bounds__(kBlock) void k_bsdf_sample(const float* __restrict blocks, const float* __restrict nrm,
                                                        const float* __restrict v, const float* __restrict uni,
                                                        uint32_t per, uint32_t count, float* __restrict out)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    Surface s = surface_from_block(blocks + 22 * (size_t)i, nrm + 3 * (size_t)i);
    ArrayUniforms u{uni + (size_t)per * i, per, 0};
    BsdfSample bs;
    bool ok = bsdf_sample(u, s, mk3(v[3 * i], v[3 * i + 1], v[3 * i + 2]), bs);
    float* o = out + 14 * (size_t)i;
    o[0] = ok ? 1.0f : 0.0f;
    o[1] = (float)bs.flags;
    o[2] = bs.reflectance.x;
    o[3] = bs.reflectance.y;
    o[4] = bs.reflectance.z;
    o[5] = bs.wi.x;
    o[6] = bs.wi.y;
    o[7] = bs.wi.z;
    o[8] = bs.fwd_pdf;
    o[9] = bs.rev_pdf;
    o[10] = (float)bs.medium;
    o[11] = bs.extinction.x;
    o[12] = bs.extinction.y;
    o[13] = bs.extinction.z;
}

}  // namespace

extern "C" {

int sgpu_bsdf_eval(const float* surfaces, const float* normals, const float* v, const float* l, uint32_t count,
                   float* out, void* stream)
{
    if (count == 0) return SGPU_OK;
    if (!surfaces || !normals || !v || !l || !out) return sgpu_fail(SGPU_E_INVALID, "sgpu_bsdf_eval: NULL pointer");
    hipLaunchKernelGGL(k_bsdf_eval, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, (hipStream_t)stream,
                       surfaces, normals, v, l, count, out);
    SGPU_HIP_TRY(hipGetLastError());
    return SGPU_OK;
}

int sgpu_bsdf_sample(const float* surfaces, const float* normals, const float* v, const float* uniforms,
                     uint32_t uniforms_per_sample, uint32_t count, float* out, void* stream)
{
    if (count == 0) return SGPU_OK;
    if (!surfaces || !normals || !v || !uniforms || !out)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_bsdf_sample: NULL pointer");
    if (uniforms_per_sample == 0) return sgpu_fail(SGPU_E_INVALID, "sgpu_bsdf_sample: uniforms_per_sample is 0");
    hipLaunchKernelGGL(k_bsdf_sample, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, (hipStream_t)stream,
                       surfaces, normals, v, uniforms, uniforms_per_sample, count, out);
    SGPU_HIP_TRY(hipGetLastError());
    return SGPU_OK;
}

}  // extern "C"
