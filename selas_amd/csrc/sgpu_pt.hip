// Wavefront unidirectional path tracer: the device form of PathTracer::GenerateImage /
// EvaluatePath (Source/Applications/Selas/Source/PathTracer.cpp:142-348).
//
// One path per (pixel, sample). Each bounce is a sequence of launches on one stream:
//   sgpu_intersect over the active queue -> k_pt_shade (surface, NEE, BSDF sample, append the
//   shadow ray and the continuation ray with ballot/prefix compaction) -> sgpu_occluded over the
//   shadow queue -> k_pt_resolve (add unoccluded light contributions).
// Per path the arithmetic and the order of random numbers follow EvaluatePath: camera jitter
// (2), then per bounce NEE (light pick + 2), BSDF sample (1 + 2..3), Russian roulette after 8
// bounces (1). Random numbers come from a PCG32 stream per path (stream = path index), so a
// path's result does not depend on scheduling. Samples of a pixel are summed in sample order
// and the frame is scaled by 1/spp, as FrameBuffer_Scale does (PathTracer.cpp:346).
// Layers: 0 = first-bounce reflectance, 1 = radiance (PathTracer.cpp:191,224,241).
//
// The same wavefront also runs the deferred path tracer (DeferredPathTracer.cpp:61-430): CMJ
// primary rays, at each hit NEE toward the lights (weight 1) and a sky sample weighted by the
// balance heuristic against the BSDF, the continuation's sky weighted the other way on a miss,
// Russian roulette after 7 tracked bounces, one output layer; and, as PathTracingBatcher orders its
// hit batches (PathTracingBatcher.cpp:221-258, 799-830), hits shaded in (instance, geometry,
// primitive) order by a device radix sort of the hit queue. A path's random numbers and the order
// of its contributions do not depend on the shading order, so sorted and unsorted renders are
// bit-identical.
#include "sgpu_camera.h"
#include "sgpu_ibl.h"
#include "sgpu_internal.h"
#include "sgpu_light.h"
#include "sgpu_surface.h"
#include "sgpu_ptex.h"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstring>

namespace {

using namespace sgpu;

constexpr int kBlock = 256;
constexpr float kFloatMax = 3.402823466e+38F;

// Queue appends with one atomic per block (sgpu_internal.h).
__device__ __forceinline__ uint32_t block_append(bool pred, uint32_t* counter)
{
    return sgpu_block_append<kBlock>(pred, counter);
}

struct PathState {
    float* thr;    // 3 per path
    float* ld0;    // 3 per path, layer 0
    float* ld1;    // 3 per path, layer 1
    uint64_t* rng; // PCG32 state per path (increment = 2 * path_id + 1)
    uint32_t* flags;  // bit 0: every scatter so far was a Dirac event; bit 1: the last one was
    float* pdf;       // deferred: solid-angle pdf of the last BSDF sample
};

struct RayBuf {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tnear, *tfar;
};

struct HitBuf {
    float *t, *u, *v, *ngx, *ngy, *ngz;
    uint32_t *prim, *geom;
    uint32_t* inst;  // composed instance of the hit (instanced scenes)
};

struct ShadowBuf {
    RayBuf ray;
    uint32_t* path;
    float* c;  // 3 per shadow ray
    uint32_t* occluded;
};

__device__ __forceinline__ void put_ray(const RayBuf& r, uint32_t i, f3 o, f3 d, float tnear, float tfar)
{
    r.ox[i] = o.x;
    r.oy[i] = o.y;
    r.oz[i] = o.z;
    r.dx[i] = d.x;
    r.dy[i] = d.y;
    r.dz[i] = d.z;
    r.tnear[i] = tnear;
    r.tfar[i] = tfar;
}

__device__ __forceinline__ Pcg32 load_rng(const PathState& ps, uint64_t path_global, uint32_t p)
{
    Pcg32 r;
    r.state = ps.rng[p];
    r.inc = (path_global << 1u) | 1u;
    return r;
}

__device__ __forceinline__ void add3(float* a, uint32_t p, f3 v)
{
    a[3 * p] += v.x;
    a[3 * p + 1] += v.y;
    a[3 * p + 2] += v.z;
}

__device__ __forceinline__ f3 get3(const float* a, uint32_t p) { return mk3(a[3 * p], a[3 * p + 1], a[3 * p + 2]); }
__device__ __forceinline__ void set3(float* a, uint32_t p, f3 v)
{
    a[3 * p] = v.x;
    a[3 * p + 1] = v.y;
    a[3 * p + 2] = v.z;
}

__global__ __launch_bounds__(kBlock) void k_pt_generate(CameraParams cam, uint32_t spp, uint64_t seed,
                                                        uint64_t path_base, uint32_t n_paths, PathState ps,
                                                        RayBuf rays, uint32_t* __restrict queue)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_paths) return;
    uint64_t path = path_base + i;
    uint32_t pixel = (uint32_t)(path / spp);
    uint32_t x = pixel % cam.width, y = pixel / cam.width;
    PcgUniforms u{pcg32_seed(seed ^ (path * 0x9E3779B97F4A7C15ull), path)};
    // JitteredCameraRay (Camera.cpp:63-75)
    float vx = (float)x + u.next();
    float vy = (float)y + u.next();
    f3 d = camera_image_to_world_dir(cam, vx, vy);
    put_ray(rays, i, cam.position, d, 0.0f, kFloatMax);
    queue[i] = i;
    set3(ps.thr, i, mk3(1.0f, 1.0f, 1.0f));
    set3(ps.ld0, i, mk3(0, 0, 0));
    set3(ps.ld1, i, mk3(0, 0, 0));
    ps.rng[i] = u.r.state;
    ps.flags[i] = 1u;
}

// GeneratePrimaryRays (DeferredPathTracer.cpp:338-370): sample s of the pixel's spp_x x spp_y
// CMJ pattern (JitteredCameraRay, Camera.cpp:77-87, pattern = pixel index); no random numbers.
__global__ __launch_bounds__(kBlock) void k_dpt_generate(CameraParams cam, uint32_t spp_x, uint32_t spp_y,
                                                         uint64_t seed, uint64_t path_base, uint32_t n_paths,
                                                         PathState ps, RayBuf rays, uint32_t* __restrict queue)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_paths) return;
    const uint32_t spp = spp_x * spp_y;
    uint64_t path = path_base + i;
    uint32_t pixel = (uint32_t)(path / spp);
    uint32_t s = (uint32_t)(path - (uint64_t)pixel * spp);
    uint32_t x = pixel % cam.width, y = pixel / cam.width;
    f3 d = camera_cmj_dir(cam, (int32_t)x, (int32_t)y, (int32_t)s, (int32_t)spp_x, (int32_t)spp_y, (int32_t)pixel);
    put_ray(rays, i, cam.position, d, 0.0f, kFloatMax);
    queue[i] = i;
    set3(ps.thr, i, mk3(1.0f, 1.0f, 1.0f));
    set3(ps.ld1, i, mk3(0, 0, 0));
    ps.rng[i] = pcg32_seed(seed ^ (path * 0x9E3779B97F4A7C15ull), path).state;
    ps.flags[i] = 3u;
    ps.pdf[i] = 0.0f;
}

// Sort keys of the hit queue: (geometry in the scene's list, primitive), misses last. For an
// instanced scene the list numbers geometries in composition order, which is (instID[0],
// instID[1]) order, so the key order is PathTracingBatcher's (instance, geometry, primitive).
__global__ __launch_bounds__(kBlock) void k_dpt_keys(uint32_t count, const uint32_t* __restrict geom,
                                                     const uint32_t* __restrict prim, const uint32_t* __restrict inst,
                                                     const uint32_t* __restrict inst_geom_base, uint32_t instanced,
                                                     uint32_t n_geoms, uint64_t* __restrict keys,
                                                     uint32_t* __restrict vals)
{
    uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= count) return;
    uint32_t g = geom[j];
    uint64_t key = (uint64_t)n_geoms << 32;
    if (g != SGPU_INVALID_ID) {
        if (instanced) g += inst_geom_base[inst[j]];
        key = ((uint64_t)min(g, n_geoms) << 32) | prim[j];
    }
    keys[j] = key;
    vals[j] = j;
}

// PathTracingBatcher::GetSortedBatch (PathTracingBatcher.cpp:529-560, 705-750): rays binned by the
// axis and sign of their direction's largest component as DetermineRayCategory (:170-177) bins them
// (strict comparisons, x before y before z on ties; a zero component goes to the negative bin),
// sorted by origin along that axis (quantised to 29 bits over the scene box). The reference's
// quicksort orders partitions of 64 rays or fewer by the direction component instead (Sample,
// :183-186); here the whole bin is ordered by origin. The image does not depend on the order.
__global__ __launch_bounds__(kBlock) void k_dpt_ray_keys(uint32_t count, RayBuf r, f3 lo, f3 inv_ext,
                                                         uint32_t* __restrict keys, uint32_t* __restrict vals)
{
    uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= count) return;
    const float ax = fabsf(r.dx[j]), ay = fabsf(r.dy[j]), az = fabsf(r.dz[j]);
    uint32_t axis = 2;
    float dv = r.dz[j], o = r.oz[j], l = lo.z, s = inv_ext.z;
    if (ax > ay && ax > az) {
        axis = 0;
        dv = r.dx[j];
        o = r.ox[j];
        l = lo.x;
        s = inv_ext.x;
    } else if (ay > az) {
        axis = 1;
        dv = r.dy[j];
        o = r.oy[j];
        l = lo.y;
        s = inv_ext.y;
    }
    const float t = fminf(fmaxf((o - l) * s, 0.0f), 1.0f);
    const uint32_t q = min((uint32_t)(t * 536870911.0f), 536870911u);
    keys[j] = ((2u * axis + (dv > 0.0f ? 0u : 1u)) << 29) | q;
    vals[j] = j;
}

// Rays and their path ids copied into sorted order.
__global__ __launch_bounds__(kBlock) void k_dpt_ray_gather(uint32_t count, const uint32_t* __restrict order, RayBuf src,
                                                           const uint32_t* __restrict qsrc, RayBuf dst,
                                                           uint32_t* __restrict qdst)
{
    uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t j = order[k];
    put_ray(dst, k, mk3(src.ox[j], src.oy[j], src.oz[j]), mk3(src.dx[j], src.dy[j], src.dz[j]), src.tnear[j],
            src.tfar[j]);
    qdst[k] = qsrc[j];
}

struct ShadeParams {
    const float* materials;
    const uint32_t* geom_material;
    uint32_t n_geoms;
    const PtexTex* geom_tex;  // per-face texture per geometry (texels == nullptr: none)
    uint32_t n_geom_tex;
    const GeomAttr* geom_attr;  // vertex attributes per geometry (NULL: no geometry has any)
    const Light* lights;
    uint32_t n_lights;
    f3 background;
    uint32_t has_ibl;  // 1: misses look up the image-based light instead of the background
    IblView ibl;
    // instanced scenes: geometry numbering and normal transform per composed instance
    uint32_t instanced;
    const uint32_t* inst_geom_base;
    const float* inst_l2w;
    uint32_t bounce;  // bounces completed before this one
    uint32_t max_len;
    uint32_t flags;
    uint64_t path_base;
};

constexpr float kInv4Pi = 0.0795774715459476678844f;

// DPT: the deferred path tracer's estimator (ShadeHitPosition, DeferredPathTracer.cpp:61-148; miss
// in TraceRayBatch, :190-200), with its sky sample appended to `sky` and hits taken in the order
// `order` gives (nullptr: queue order).
template <bool DPT>
__global__ __launch_bounds__(kBlock) void k_pt_shade(ShadeParams sp, uint32_t count, const uint32_t* __restrict queue,
                                                     const uint32_t* __restrict order, RayBuf rays, HitBuf hits,
                                                     PathState ps, RayBuf next_rays, uint32_t* __restrict next_queue,
                                                     ShadowBuf shadow, ShadowBuf sky, uint32_t* __restrict counters)
{
    const uint32_t jj = blockIdx.x * kBlock + threadIdx.x;
    bool active = jj < count;
    const uint32_t j = active && order ? order[jj] : jj;
    bool cont = false, emit = false, emit_sky = false, hit = false;
    uint32_t p = 0;
    f3 next_o = mk3(0, 0, 0), next_d = mk3(0, 0, 0), sh_o = mk3(0, 0, 0), sh_d = mk3(0, 0, 0), c = mk3(0, 0, 0);
    f3 sk_o = mk3(0, 0, 0), sk_d = mk3(0, 0, 0), sk_c = mk3(0, 0, 0);
    float sh_tnear = 0.0f, sh_tfar = 0.0f;
    if (active) {
        p = queue[j];
        f3 o = mk3(rays.ox[j], rays.oy[j], rays.oz[j]);
        f3 d = mk3(rays.dx[j], rays.dy[j], rays.dz[j]);
        f3 thr = get3(ps.thr, p);
        uint32_t geom = hits.geom[j];
        f3 ng = mk3(hits.ngx[j], hits.ngy[j], hits.ngz[j]);
        const float* l2w = nullptr;
        if (sp.instanced && geom != SGPU_INVALID_ID) {
            // the placed scene's geometry in the scene's list, its normal in world space
            const uint32_t k = hits.inst[j];
            geom += sp.inst_geom_base[k];
            l2w = sp.inst_l2w + 9 * (size_t)k;
            ng = xform_linear(l2w, ng);
        }
        if (geom == SGPU_INVALID_ID || geom >= sp.n_geoms) {
            // miss: EvaluateBackground (AreaLighting.cpp:415-424): the image-based light's radiance
            // from the ray's direction, else the constant background (EvaluateBackgroundMiss reads
            // a second image of the same light, which here is the same image)
            float pdf = kInv4Pi;
            const f3 bg = sp.has_ibl ? ibl_lookup(sp.ibl, d, pdf) : sp.background;
            if (DPT) {
                // BackgroundLightingPdf against the pdf the direction was sampled with, unless
                // the last scatter (or every one, for camera rays) was a Dirac event
                float w = 1.0f;
                if (!(ps.flags[p] & 2u)) {
                    const float lp = ps.pdf[p];
                    w = lp + pdf > 0.0f ? lp / (lp + pdf) : 0.0f;
                }
                add3(ps.ld1, p, (bg * w) * thr);
            } else {
                add3(ps.ld1, p, bg * thr);
            }
        } else {
            hit = true;
            PcgUniforms u{load_rng(ps, sp.path_base + p, p)};
            const float* mat = sp.materials + 22 * (size_t)sp.geom_material[geom];
            Surface s;
            // CalculateSurfaceParams' eUsesPtex branch (SurfaceParameters.cpp:196-203)
            f3 texel = mk3(0, 0, 0);
            bool tex = false;
            if (((uint32_t)mat[21] & kMatUsesPtex) && geom < sp.n_geom_tex && sp.geom_tex[geom].texels) {
                texel = ptex_eval(sp.geom_tex[geom], (int)hits.prim[j], hits.u[j], hits.v[j]);
                tex = true;
            }
            surface_from_hit(o, d, hits.t[j], ng, sp.geom_attr ? sp.geom_attr + geom : nullptr, hits.prim[j], hits.u[j],
                             hits.v[j], l2w, mat, tex, texel, s);
            f3 v = -d;
            LightSample ls;
            next_event_estimation(u, sp.lights, sp.n_lights, s.position, s.n,
                                  DPT || (sp.flags & SGPU_PT_NEE_COSINE_ONCE) != 0, ls);
            if (DPT) {
                if (dot3(ls.radiance, mk3(1, 1, 1)) > 0) {
                    float fp, rp;
                    const f3 f = bsdf_eval(s, v, ls.direction, fp, rp);
                    const f3 smp = (f * ls.radiance) * (1.0f / ls.pdf_w);
                    if (dot3(smp, mk3(1, 1, 1)) > 0) {
                        // TraceOcclusionBatch: tnear 0 from the origin biased by 0.1 error
                        emit = true;
                        sh_o = offset_ray_origin(s, ls.direction, 0.1f);
                        sh_d = ls.direction;
                        sh_tfar = ls.distance;
                        c = smp * thr;
                    }
                }
                // SampleBackground: the image-based light, else the constant background sampled
                // uniformly over the sphere (pdf 1 / (4 pi))
                const float r0 = u.next(), r1 = u.next();
                f3 sd, srad;
                float spdf;
                if (sp.has_ibl) {
                    srad = ibl_direct(sp.ibl, r0, r1, sd, spdf);
                } else {
                    const float uu = 2.0f * r1 - 1.0f;
                    const float norm = sqrtf(max_(0.0f, 1.0f - uu * uu));
                    const float theta = (kPi * 2.0f) * r0;
                    sd = mk3(norm * cos_f(theta), uu, norm * sin_f(theta));
                    srad = sp.background;
                    spdf = kInv4Pi;
                }
                if (dot3(srad, mk3(1, 1, 1)) > 0 && spdf > 0.0f) {
                    float fp, rp;
                    const f3 f = bsdf_eval(s, v, sd, fp, rp);
                    const float w = spdf / (spdf + fp);
                    const f3 smp = ((f * w) * srad) * (1.0f / spdf);
                    if (dot3(smp, mk3(1, 1, 1)) > 0) {
                        emit_sky = true;
                        sk_o = offset_ray_origin(s, sd, 0.1f);
                        sk_d = sd;
                        sk_c = smp * thr;
                    }
                }
            } else if (dot3(ls.radiance, mk3(1, 1, 1)) > 0) {
                float fp, rp;
                f3 f = bsdf_eval(s, v, ls.direction, fp, rp);
                if (dot3(f, mk3(1, 1, 1)) > 0) {
                    // OcclusionRay (PathTracer.cpp:71-92): origin biased by 0.1 error, tnear = error
                    emit = true;
                    sh_o = offset_ray_origin(s, ls.direction, 0.1f);
                    sh_d = ls.direction;
                    sh_tnear = s.error;
                    sh_tfar = ls.distance;
                    c = (f * ls.radiance * (1.0f / ls.pdf_w)) * thr;
                }
            }
            BsdfSample bs;
            if (DPT) {
                if (bsdf_sample(u, s, v, bs) && bs.fwd_pdf > 0.0f) {
                    thr = thr * bs.reflectance;
                    if (!((thr.x * thr.x + thr.y * thr.y) + thr.z * thr.z == 0.0f)) {
                        const uint32_t dirac = (bs.flags & kDirac) ? 1u : 0u;
                        ps.flags[p] = (ps.flags[p] & dirac) | (dirac << 1);
                        ps.pdf[p] = bs.fwd_pdf;
                        cont = sp.bounce + 1 < sp.max_len;
                        if (sp.bounce >= 7) {
                            // Russian roulette once the tracked bounces saturate (MaxTrackedBounces_ = 7)
                            const float cp = max_(max_(thr.x, thr.y), thr.z);
                            if (u.next() >= cp)
                                cont = false;
                            else
                                thr = thr * (1.0f / cp);
                        }
                        next_o = offset_ray_origin(s, bs.wi, 1.0f);
                        next_d = bs.wi;
                        set3(ps.thr, p, thr);
                    }
                }
            } else if (bsdf_sample(u, s, v, bs) && bs.fwd_pdf > 0.0f) {
                uint32_t fl = ps.flags[p];
                if (!(bs.flags & kDirac)) fl &= ~1u;
                ps.flags[p] = fl;
                thr = (1.0f * thr) * bs.reflectance;
                next_o = offset_ray_origin(s, bs.wi, 1.0f);
                next_d = bs.wi;
                if (sp.bounce == 0) set3(ps.ld0, p, bs.reflectance);
                uint32_t bounce = sp.bounce + 1;
                cont = bounce < sp.max_len;
                if (cont && bounce > 8) {
                    float cp = max_(max_(thr.x, thr.y), thr.z);
                    if (u.next() > cp)
                        cont = false;
                    else
                        thr = thr * (1.0f / cp);
                }
                set3(ps.thr, p, thr);
            }
            ps.rng[p] = u.r.state;
        }
    }
    uint32_t ks = block_append(emit, &counters[1]);
    if (emit) {
        put_ray(shadow.ray, ks, sh_o, sh_d, sh_tnear, sh_tfar);
        shadow.path[ks] = p;
        set3(shadow.c, ks, c);
    }
    if (DPT) {
        uint32_t kk = block_append(emit_sky, &counters[2]);
        if (emit_sky) {
            put_ray(sky.ray, kk, sk_o, sk_d, 0.0f, 1e36f);
            sky.path[kk] = p;
            set3(sky.c, kk, sk_c);
        }
    }
    uint32_t kn = block_append(cont, &counters[0]);
    (void)block_append(hit, &counters[3]);
    if (cont) {
        put_ray(next_rays, kn, next_o, next_d, 0.0f, kFloatMax);
        next_queue[kn] = p;
    }
}

__global__ __launch_bounds__(kBlock) void k_pt_resolve(uint32_t count, ShadowBuf shadow, PathState ps)
{
    uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    if (shadow.occluded[k]) return;
    add3(ps.ld1, shadow.path[k], get3(shadow.c, k));
}

// Sum each pixel's samples in sample order into the frame: both layers, or (one_layer) the
// radiance alone.
__global__ __launch_bounds__(kBlock) void k_pt_accumulate(uint32_t spp, uint32_t pixel_base, uint32_t n_pixels,
                                                          uint32_t n_frame_pixels, uint32_t one_layer, PathState ps,
                                                          float* __restrict fb)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_pixels) return;
    f3 a0 = mk3(0, 0, 0), a1 = mk3(0, 0, 0);
    for (uint32_t s = 0; s < spp; ++s) {
        uint32_t p = i * spp + s;
        if (!one_layer) a0 += get3(ps.ld0, p);
        a1 += get3(ps.ld1, p);
    }
    uint32_t pix = pixel_base + i;
    if (one_layer) {
        add3(fb, pix, a1);
        return;
    }
    add3(fb, pix, a0);
    add3(fb + 3 * (size_t)n_frame_pixels, pix, a1);
}

__global__ __launch_bounds__(kBlock) void k_scale(float* __restrict fb, size_t n, float s)
{
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) fb[i] = fb[i] * s;
}

struct Arena {
    std::vector<void*> ptrs;
    hipStream_t stream;
    hipError_t err = hipSuccess;
    template <typename T>
    T* get(size_t count)
    {
        void* p = nullptr;
        if (err == hipSuccess) err = hipMallocAsync(&p, count * sizeof(T) + 64, stream);
        if (err == hipSuccess) ptrs.push_back(p);
        return (T*)p;
    }
    RayBuf rays(size_t n)
    {
        RayBuf r;
        r.ox = get<float>(n);
        r.oy = get<float>(n);
        r.oz = get<float>(n);
        r.dx = get<float>(n);
        r.dy = get<float>(n);
        r.dz = get<float>(n);
        r.tnear = get<float>(n);
        r.tfar = get<float>(n);
        return r;
    }
    ~Arena()
    {
        for (void* p : ptrs) (void)hipFreeAsync(p, stream);
    }
};

sgpu_rays_soa as_soa(const RayBuf& r) { return sgpu_rays_soa{r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, r.tnear, r.tfar}; }

}  // namespace

namespace {

// One render of either tracer: the path tracer (two layers) or, with dpt, the deferred path
// tracer (one layer, CMJ primary rays, optional hit sort).
struct PathJob {
    sgpu_camera camera;
    uint32_t spp_x, spp_y;  // the path tracer: spp_x = paths per pixel, spp_y = 1
    uint32_t max_path_length;
    uint64_t seed;
    uint32_t flags;
    bool dpt;
    bool sort_hits;
    bool sort_rays;
};

enum { kPtClosest, kPtShadow, kPtShade, kPtOther, kPtPhases };

// stats (may be NULL): ray counts and per-phase device times of the render (events around every
// launch), traversal counters with SGPU_PT_COUNT_TRAVERSAL.
int render_paths(sgpu_scene scene, const PathJob& job, float* fb, hipStream_t stream, sgpu_pt_stats* stats = nullptr)
{
    const uint32_t W = job.camera.width, H = job.camera.height, spp = job.spp_x * job.spp_y;
    const uint64_t n_pixels = (uint64_t)W * H;
    const uint32_t layers = job.dpt ? 1u : 2u;
    const uint32_t max_chunk_paths = 1u << 22;
    uint32_t chunk_pixels = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n_pixels, max_chunk_paths / spp));
    const uint32_t cap = chunk_pixels * spp;

    CameraParams cam = camera_init(mk3(job.camera.position[0], job.camera.position[1], job.camera.position[2]),
                                   mk3(job.camera.look_at[0], job.camera.look_at[1], job.camera.look_at[2]),
                                   mk3(job.camera.up[0], job.camera.up[1], job.camera.up[2]),
                                   job.camera.fov_degrees, W, H);

    Arena A;
    A.stream = stream;
    PathState ps;
    ps.thr = A.get<float>(3 * (size_t)cap);
    ps.ld0 = job.dpt ? nullptr : A.get<float>(3 * (size_t)cap);
    ps.ld1 = A.get<float>(3 * (size_t)cap);
    ps.rng = A.get<uint64_t>(cap);
    ps.flags = A.get<uint32_t>(cap);
    ps.pdf = job.dpt ? A.get<float>(cap) : nullptr;
    RayBuf rays[3] = {A.rays(cap), A.rays(cap), RayBuf{}};
    uint32_t* queue[3] = {A.get<uint32_t>(cap), A.get<uint32_t>(cap), nullptr};
    if (job.sort_rays) {
        rays[2] = A.rays(cap);
        queue[2] = A.get<uint32_t>(cap);
    }
    HitBuf hb;
    hb.t = A.get<float>(cap);
    hb.u = A.get<float>(cap);
    hb.v = A.get<float>(cap);
    hb.ngx = A.get<float>(cap);
    hb.ngy = A.get<float>(cap);
    hb.ngz = A.get<float>(cap);
    hb.prim = A.get<uint32_t>(cap);
    hb.geom = A.get<uint32_t>(cap);
    hb.inst = A.get<uint32_t>(cap);
    auto shadow_buf = [&]() {
        ShadowBuf b;
        b.ray = A.rays(cap);
        b.path = A.get<uint32_t>(cap);
        b.c = A.get<float>(3 * (size_t)cap);
        b.occluded = A.get<uint32_t>(cap);
        return b;
    };
    ShadowBuf sb = shadow_buf();
    ShadowBuf sky = job.dpt ? shadow_buf() : ShadowBuf{};
    uint32_t* spill = A.get<uint32_t>(sgpu_trace_spill_words(scene));
    uint32_t* counters = A.get<uint32_t>(4);
    const bool count_trav = stats && (job.flags & SGPU_PT_COUNT_TRAVERSAL);
    uint64_t* d_trav = count_trav ? A.get<uint64_t>(2) : nullptr;
    sgpu_phase_timer T(stats != nullptr, stream);
    uint64_t n_closest = 0, n_shadow = 0, n_hits = 0, n_bounces = 0;

    ShadeParams sp;
    sp.materials = scene->d_materials;
    sp.geom_material = scene->d_geom_material;
    sp.n_geoms = scene->n_geom_material;
    sp.instanced = scene->instances.empty() ? 0u : 1u;
    sp.inst_geom_base = scene->d_inst_geom_base;
    sp.inst_l2w = scene->d_inst_l2w;
    sp.lights = (const Light*)scene->d_lights;
    sp.n_lights = scene->n_lights;
    sp.background = mk3(scene->background[0], scene->background[1], scene->background[2]);
    sp.has_ibl = ibl_view(scene, sp.ibl) ? 1u : 0u;
    sp.geom_tex = scene->d_geom_tex;
    sp.n_geom_tex = scene->n_geom_tex;
    sp.geom_attr = scene->d_geom_attr;
    if (sp.instanced) {
        sp.materials = scene->d_g_materials;
        sp.geom_material = scene->d_g_geom_material;
        sp.n_geoms = scene->n_g_geoms;
        sp.geom_tex = scene->d_g_geom_tex;
        sp.n_geom_tex = scene->n_g_geoms;
        sp.geom_attr = scene->d_g_geom_attr;
    }
    sp.max_len = job.max_path_length;
    sp.flags = job.flags;

    // hit sort: 64-bit keys (geometry << 32 | primitive), the geometry part sized to the scene
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    void* sort_tmp = nullptr;
    size_t sort_bytes = 0;
    uint32_t end_bit = 32;
    while (end_bit < 64 && ((uint64_t)sp.n_geoms >> (end_bit - 32)) != 0) ++end_bit;
    uint32_t* rkeys[2] = {nullptr, nullptr};
    uint32_t* rvals[2] = {nullptr, nullptr};
    void* rsort_tmp = nullptr;
    size_t rsort_bytes = 0;
    if (job.sort_rays) {
        for (int k = 0; k < 2; ++k) {
            rkeys[k] = A.get<uint32_t>(cap);
            rvals[k] = A.get<uint32_t>(cap);
        }
        SGPU_HIP_TRY(rocprim::radix_sort_pairs(nullptr, rsort_bytes, rkeys[0], rkeys[1], rvals[0], rvals[1], (size_t)cap,
                                               0, 32, stream));
        rsort_tmp = A.get<uint8_t>(rsort_bytes);
    }
    const f3 box_lo = mk3(scene->box[0], scene->box[1], scene->box[2]);
    auto inv_extent = [](float lo, float hi) { return hi > lo ? 1.0f / (hi - lo) : 0.0f; };
    const f3 box_inv = mk3(inv_extent(scene->box[0], scene->box[3]), inv_extent(scene->box[1], scene->box[4]),
                           inv_extent(scene->box[2], scene->box[5]));
    if (job.sort_hits) {
        keys[0] = A.get<uint64_t>(cap);
        keys[1] = A.get<uint64_t>(cap);
        vals[0] = A.get<uint32_t>(cap);
        vals[1] = A.get<uint32_t>(cap);
        SGPU_HIP_TRY(rocprim::radix_sort_pairs(nullptr, sort_bytes, keys[0], keys[1], vals[0], vals[1], (size_t)cap, 0,
                                               end_bit, stream));
        sort_tmp = A.get<uint8_t>(sort_bytes);
    }
    uint32_t* h_counters = nullptr;
    if (A.err != hipSuccess) return sgpu_fail(SGPU_E_NOMEM, "sgpu_render: device allocation failed");
    SGPU_HIP_TRY(hipHostMalloc((void**)&h_counters, 4 * sizeof(uint32_t), hipHostMallocDefault));
    struct HostFree {
        uint32_t* p;
        ~HostFree() { (void)hipHostFree(p); }
    } hf{h_counters};

    SGPU_HIP_TRY(hipMemsetAsync(fb, 0, layers * 3 * n_pixels * sizeof(float), stream));
    if (d_trav) SGPU_HIP_TRY(hipMemsetAsync(d_trav, 0, 2 * sizeof(uint64_t), stream));
    sgpu_hits_soa hsoa{hb.t, hb.u, hb.v, hb.prim, hb.geom, hb.ngx, hb.ngy, hb.ngz};

    auto resolve = [&](ShadowBuf& b, uint32_t n) -> int {
        if (n == 0) return SGPU_OK;
        n_shadow += n;
        sgpu_rays_soa ss = as_soa(b.ray);
        SGPU_HIP_TRY(T.begin(kPtShadow));
        int rc = sgpu_trace_launch(scene, &ss, nullptr, b.occluded, n, true, nullptr, stream, spill);
        if (rc) return rc;
        SGPU_HIP_TRY(T.end());
        SGPU_HIP_TRY(T.begin(kPtOther));
        hipLaunchKernelGGL(k_pt_resolve, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, n, b, ps);
        SGPU_HIP_TRY(hipGetLastError());
        SGPU_HIP_TRY(T.end());
        return SGPU_OK;
    };

    for (uint64_t pix0 = 0; pix0 < n_pixels; pix0 += chunk_pixels) {
        const uint32_t npix = (uint32_t)std::min<uint64_t>(chunk_pixels, n_pixels - pix0);
        const uint32_t n_paths = npix * spp;
        const uint64_t path_base = pix0 * spp;
        SGPU_HIP_TRY(T.begin(kPtOther));
        if (job.dpt)
            hipLaunchKernelGGL(k_dpt_generate, dim3((n_paths + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, cam,
                               job.spp_x, job.spp_y, job.seed, path_base, n_paths, ps, rays[0], queue[0]);
        else
            hipLaunchKernelGGL(k_pt_generate, dim3((n_paths + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, cam, spp,
                               job.seed, path_base, n_paths, ps, rays[0], queue[0]);
        SGPU_HIP_TRY(hipGetLastError());
        SGPU_HIP_TRY(T.end());
        uint32_t count = n_paths;
        int cur = 0;
        sp.path_base = path_base;
        for (uint32_t bounce = 0; bounce < job.max_path_length && count > 0; ++bounce) {
            sgpu_rays_soa rs = as_soa(rays[cur]);
            n_closest += count;
            ++n_bounces;
            SGPU_HIP_TRY(T.begin(kPtClosest));
            int rc = sgpu_trace_launch(scene, &rs, &hsoa, nullptr, count, false, d_trav, stream, spill, nullptr, nullptr,
                                       hb.inst);
            if (rc) return rc;
            SGPU_HIP_TRY(T.end());
            const uint32_t* order = nullptr;
            if (job.sort_hits) {
                hipLaunchKernelGGL(k_dpt_keys, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, count,
                                   hb.geom, hb.prim, hb.inst, sp.inst_geom_base, sp.instanced, sp.n_geoms, keys[0],
                                   vals[0]);
                SGPU_HIP_TRY(hipGetLastError());
                size_t bytes = sort_bytes;
                SGPU_HIP_TRY(rocprim::radix_sort_pairs(sort_tmp, bytes, keys[0], keys[1], vals[0], vals[1], (size_t)count,
                                                       0, end_bit, stream));
                order = vals[1];
            }
            SGPU_HIP_TRY(hipMemsetAsync(counters, 0, 4 * sizeof(uint32_t), stream));
            sp.bounce = bounce;
            SGPU_HIP_TRY(T.begin(kPtShade));
            if (job.dpt)
                hipLaunchKernelGGL(k_pt_shade<true>, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, sp,
                                   count, queue[cur], order, rays[cur], hb, ps, rays[1 - cur], queue[1 - cur], sb, sky,
                                   counters);
            else
                hipLaunchKernelGGL(k_pt_shade<false>, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, sp,
                                   count, queue[cur], order, rays[cur], hb, ps, rays[1 - cur], queue[1 - cur], sb, sky,
                                   counters);
            SGPU_HIP_TRY(hipGetLastError());
            SGPU_HIP_TRY(T.end());
            SGPU_HIP_TRY(hipMemcpyAsync(h_counters, counters, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
            SGPU_HIP_TRY(hipStreamSynchronize(stream));
            const uint32_t n_next = h_counters[0];
            n_hits += h_counters[3];
            // per path: the light sample's contribution, then the sky sample's
            if ((rc = resolve(sb, h_counters[1]))) return rc;
            if (job.dpt && (rc = resolve(sky, h_counters[2]))) return rc;
            if (job.sort_rays && n_next > 1) {
                const int nx = 1 - cur;
                const dim3 g((n_next + kBlock - 1) / kBlock);
                hipLaunchKernelGGL(k_dpt_ray_keys, g, dim3(kBlock), 0, stream, n_next, rays[nx], box_lo, box_inv,
                                   rkeys[0], rvals[0]);
                SGPU_HIP_TRY(hipGetLastError());
                size_t bytes = rsort_bytes;
                SGPU_HIP_TRY(rocprim::radix_sort_pairs(rsort_tmp, bytes, rkeys[0], rkeys[1], rvals[0], rvals[1],
                                                       (size_t)n_next, 0, 32, stream));
                hipLaunchKernelGGL(k_dpt_ray_gather, g, dim3(kBlock), 0, stream, n_next, rvals[1], rays[nx], queue[nx],
                                   rays[2], queue[2]);
                SGPU_HIP_TRY(hipGetLastError());
                std::swap(rays[nx], rays[2]);
                std::swap(queue[nx], queue[2]);
            }
            count = n_next;
            cur = 1 - cur;
        }
        SGPU_HIP_TRY(T.begin(kPtOther));
        hipLaunchKernelGGL(k_pt_accumulate, dim3((npix + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, spp,
                           (uint32_t)pix0, npix, (uint32_t)n_pixels, job.dpt ? 1u : 0u, ps, fb);
        SGPU_HIP_TRY(hipGetLastError());
        SGPU_HIP_TRY(T.end());
    }
    const size_t nf = layers * 3 * n_pixels;
    hipLaunchKernelGGL(k_scale, dim3((uint32_t)((nf + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, fb, nf,
                       1.0f / (float)spp);
    SGPU_HIP_TRY(hipGetLastError());
    SGPU_HIP_TRY(hipStreamSynchronize(stream));
    if (stats) {
        SGPU_HIP_TRY(T.collect());
        std::memset(stats, 0, sizeof(*stats));
        stats->closest_rays = n_closest;
        stats->shadow_rays = n_shadow;
        stats->hits = n_hits;
        stats->bounces = n_bounces;
        if (d_trav) {
            uint64_t tr[2] = {0, 0};
            SGPU_HIP_TRY(hipMemcpy(tr, d_trav, sizeof(tr), hipMemcpyDeviceToHost));
            stats->nodes_visited = tr[0];
            stats->tris_tested = tr[1];
        }
        stats->closest_ms = (float)T.ms[kPtClosest];
        stats->shadow_ms = (float)T.ms[kPtShadow];
        stats->shade_ms = (float)T.ms[kPtShade];
        stats->other_ms = (float)T.ms[kPtOther];
    }
    return SGPU_OK;
}

}  // namespace

int sgpu_pt_render_impl(sgpu_scene scene, const sgpu_pt_job* job, float* fb, hipStream_t stream, sgpu_pt_stats* stats)
{
    PathJob pj{job->camera, job->spp, 1u, job->max_path_length, job->seed, job->flags, false, false, false};
    return render_paths(scene, pj, fb, stream, stats);
}

int sgpu_dpt_render_impl(sgpu_scene scene, const sgpu_dpt_job* job, float* fb, hipStream_t stream)
{
    PathJob pj{job->camera, job->spp_x, job->spp_y, job->max_path_length, job->seed, job->flags, true,
               (job->flags & SGPU_DPT_SORT_HITS) != 0, (job->flags & SGPU_DPT_SORT_RAYS) != 0};
    return render_paths(scene, pj, fb, stream);
}
