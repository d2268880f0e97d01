// Wavefront vertex connection and merging: the device form of Selas's VCM integrator
// (Source/Applications/Selas/Source/VCM.cpp: VertexConnectionAndMerging :669, EvaluateLightPaths
// :502, EvaluateCameraPaths :576, ConnectLightPathToCamera :190, ConnectCameraPathToLight :245,
// ConnectPathVertices :293, MergeVertices :362, SampleBsdfScattering :412; VCMCommon.cpp:37-113;
// VCMHashGrid.cpp:15-157), following Georgiev et al., "Implementing Vertex Connection and
// Merging" (2012) where the commented reference cannot run as written. The integrator semantics
// (and every deviation from the commented code) are those of the CPU restatement the tests
// compare against; see the repository's design notes for the list.
//
// Work is organised in slices: a slice is one VCM iteration of one render job (W x H light paths,
// then W x H camera paths, a photon grid of W x H cells). A wavefront runs the paths of many
// slices together -- several iterations of one job, or iterations of many jobs -- so that every
// launch is large enough to fill the chip; each path carries its slice index, and every
// slice-dependent quantity (camera, merge radius, MIS weights, random streams, grid, image) is
// looked up through it, so a slice's result does not depend on what it is batched with.
//
// One wavefront is a fixed sequence of launches:
//   light pass   k_light_gen, then per bounce: closest-hit traversal -> k_light_shade (store the
//                vertex in its path's slot, queue the camera connection and the continuation ray)
//                -> occlusion traversal -> k_splat (atomic adds into the slice's frame)
//   vertices     exclusive scan of per-path counts -> k_compact: the vertex array in path order,
//                the order in which Selas appends them (so each slice's photon grid is built over
//                the same list), slice-contiguous
//   photon grid  k_grid_bounds (box per slice), k_grid_keys (slice's cell base + hashed cell),
//                rocPRIM radix sort of (cell, vertex) pairs (stable: equal cells keep vertex order,
//                as the counting sort of BuildHashGrid does), inclusive scan of cell counts ->
//                cell range ends
//   camera pass  k_cam_gen, then per bounce: closest-hit traversal -> k_cam_shade (NEE and one
//                vertex connection per light vertex of the same path index, queued as shadow rays
//                into fixed per-path slots; hands the vertex to the merge kernel) -> k_cam_merge
//                (range search in the slice's grid) -> occlusion traversal -> k_cam_resolve ->
//                k_cam_gather (adds the unoccluded slots of each path in slot order)
//   k_frame_add  adds each slice's frame to its job's running sum, slices in iteration order.
#include "sgpu_camera.h"
#include "sgpu_internal.h"
#include "sgpu_light.h"
#include "sgpu_ptex.h"
#include "sgpu_surface.h"
#include "sgpu_vcm_common.h"

#include <cstring>  // rocprim/iterator/texture_cache_iterator.hpp uses memset without including it
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

using namespace sgpu;
using namespace sgpu_vcm;

// ---- light pass ----------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_light_gen(Wave w, PathState ps, RayBuf rays, uint32_t* __restrict queue,
                                                      uint32_t* __restrict counters)
{
    const uint32_t total = w.dims[1];
    // blocks past the end leave at once (captured launches are sized for capacity)
    if (blockIdx.x * kBlock >= total) return;
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    bool live = false;
    uint32_t scn = 0;
    f3 o = mk3(0, 0, 0), d = mk3(0, 0, 0);
    if (g < total) {
        const uint32_t si = find_slice(w, g);
        const Slice& sl = w.slices[si];
        scn = sl.scene;
        ps.slice[g] = si;
        PcgUniforms u{rng_init(sl.seed, stream_id(sl.iteration, g - sl.path_base, 0))};
        ps.plen[g] = 1;
        // GenerateLightSample (VCMCommon.cpp:37-63): one light picked uniformly, the image-based
        // light being light n_lights
        const uint32_t n_pick = sl.n_lights + sl.has_ibl;
        if (n_pick > 0) {
            float p_pick = 1.0f / n_pick;
            uint32_t idx = (uint32_t)(u.next() * n_pick);
            if (idx >= n_pick) idx = n_pick - 1;
            const bool infinite = idx == sl.n_lights;
            float emission_pdf_w, direct_pdf_a;
            f3 radiance;
            if (infinite) {
                // EmitIblLightSample (AreaLighting.cpp:246-290)
                float r0 = u.next(), r1 = u.next(), r2 = u.next(), r3 = u.next();
                float dir_pdf;
                ibl_emit(sl.ibl, r0, r1, r2, r3, o, d, radiance, emission_pdf_w, dir_pdf);
                emission_pdf_w *= p_pick;
                direct_pdf_a = dir_pdf * p_pick;
            } else {
                const Light L = sl.lights[idx];
                if (L.type == 1) {
                    float r0 = u.next(), r1 = u.next();
                    float uu = 2.0f * r1 - 1.0f;
                    float norm = sqrtf(max_(0.0f, 1.0f - uu * uu));
                    float theta = (kPi * 2.0f) * r0;
                    d = mk3(norm * cos_f(theta), uu, norm * sin_f(theta));
                    o = ld3(L.position);
                    emission_pdf_w = (1.0f / (4.0f * kPi)) * p_pick;
                    direct_pdf_a = 1.0f * p_pick;
                    radiance = ld3(L.radiance);
                } else {
                    f3 ex = ld3(L.x), ez = ld3(L.z);
                    float a = u.next(), b = u.next();
                    o = ld3(L.position) - ex * 0.5f - ez * 0.5f + ex * a + ez * b;
                    float r0 = u.next(), r1 = u.next();
                    float r = sqrtf(r0);
                    float th = (kPi * 2.0f) * r1;
                    f3 local = mk3(r * cos_f(th), sqrtf(max_(0.0f, 1 - r0)), r * sin_f(th));
                    f3 n = ld3(L.direction), t, bb;
                    make_frame(n, t, bb);
                    d = normalize3(t * local.x + n * local.y + bb * local.z);
                    float cos_l = local.y;
                    float area = length3(ex) * length3(ez);
                    float pdf_a = 1.0f / area;
                    emission_pdf_w = pdf_a * (cos_l / kPi) * p_pick;
                    direct_pdf_a = pdf_a * p_pick;
                    radiance = ld3(L.radiance) * cos_l;
                }
            }
            if (emission_pdf_w > 0.0f) {
                s3(ps.pos, g, o);
                s3(ps.dir, g, d);
                s3(ps.thr, g, radiance * (1.0f / emission_pdf_w));
                ps.dvcm[g] = direct_pdf_a / emission_pdf_w;
                if (infinite) {
                    // the image-based light can be hit by camera paths (cosThetaLight = 1)
                    const float dvc = 1.0f / emission_pdf_w;
                    ps.dvc[g] = dvc;
                    ps.dvm[g] = dvc * sl.vc_weight;
                    ps.plen[g] = 1u | kInfiniteOrigin;
                } else {
                    // quad and point lights cannot be hit: no "camera path hits the light" technique
                    ps.dvc[g] = 0.0f;
                    ps.dvm[g] = 0.0f;
                }
                live = 1 + 2 < sl.L;
            }
        }
        ps.rng[g] = u.r.state;
    }
    uint32_t k = block_append(live, &counters[0]);
    if (live) {
        put_ray(rays, k, o, d, kRayEps, kFloatMax, scn);
        queue[k] = g;
    }
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_light_shade(Wave w, const uint32_t* __restrict count_ptr, const uint32_t* __restrict queue,
                                                        RayBuf rays, HitBuf hits, PathState ps, LightVertex* __restrict verts,
                                                        uint32_t* __restrict vcount, RayBuf next_rays,
                                                        uint32_t* __restrict next_queue, RayBuf shadow,
                                                        float* __restrict splat_c, uint32_t* __restrict splat_pix,
                                                        uint32_t* __restrict next_count, uint32_t* __restrict shadow_count,
                                                        const uint32_t* __restrict order)
{
    const uint32_t count = *count_ptr;
    // blocks past the end leave at once (captured launches are sized for capacity)
    if (blockIdx.x * kBlock >= count) return;
    // queue entry of this lane: in queue order, or (order != NULL) in the order of a bin sort
    const uint32_t jj = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t j = order && jj < count ? order[jj] : jj;
    bool cont = false, emit = false;
    uint32_t p = 0, pix = 0, scn = 0;
    f3 next_o = mk3(0, 0, 0), next_d = mk3(0, 0, 0), sh_o = mk3(0, 0, 0), sh_d = mk3(0, 0, 0), c = mk3(0, 0, 0);
    float sh_tnear = 0.0f, sh_tfar = 0.0f;
    if (jj < count) {
        p = queue[j];
        const Slice& sl = w.slices[ps.slice[p]];
        scn = sl.scene;
        uint32_t geom;
        f3 ng;
        const float* l2w;
        hit_geometry(sl, hits, j, geom, ng, l2w);
        if (geom != SGPU_INVALID_ID && geom < sl.n_geoms) {
            PcgUniforms u{rng_load(ps, p, sl, 0)};
            f3 o = mk3(rays.ox[j], rays.oy[j], rays.oz[j]);
            f3 d = mk3(rays.dx[j], rays.dy[j], rays.dz[j]);
            const float* mat = sl.materials + 22 * (size_t)sl.geom_material[geom];
            Surface s;
            float4 tan;
            const bool tf = hit_surface(sl, hits, j, geom, ng, l2w, o, d, mat, s, &tan);
            f3 x = s.position;
            f3 view = -d;
            f3 pos = g3(ps.pos, p);
            f3 thr = g3(ps.thr, p);
            float dvcm = ps.dvcm[p], dvc = ps.dvc[p], dvm = ps.dvm[p];
            uint32_t plen = ps.plen[p];
            const bool from_infinite = (plen & kInfiniteOrigin) != 0u;
            plen &= ~kInfiniteOrigin;
            f3 dd = pos - x;
            float d2 = dot3(dd, dd);
            float cos_v = absf_(dot3(s.n, view));
            // the first segment is converted to area measure, except from a light at infinity
            if (!from_infinite) dvcm *= d2;
            dvcm *= (1.0f / cos_v);
            dvc *= (1.0f / cos_v);
            dvm *= (1.0f / cos_v);
            // store the vertex in the path's slot
            uint32_t k = vcount[p];
            if (k < sl.mlv) {
                LightVertex lv;
                lv.pos[0] = x.x;
                lv.pos[1] = x.y;
                lv.pos[2] = x.z;
                lv.thr[0] = thr.x;
                lv.thr[1] = thr.y;
                lv.thr[2] = thr.z;
                lv.n[0] = s.n.x;
                lv.n[1] = s.n.y;
                lv.n[2] = s.n.z;
                lv.wfix[0] = view.x;
                lv.wfix[1] = view.y;
                lv.wfix[2] = view.z;
                lv.dvcm = dvcm;
                lv.dvc = dvc;
                lv.dvm = dvm;
                lv.geom_plen = (geom << 5) | (plen & 31u);
                lv.base[0] = s.base_color.x;
                lv.base[1] = s.base_color.y;
                lv.base[2] = s.base_color.z;
                lv.cp = cont_prob(s);
                verts[(size_t)p * w.mlv + k] = lv;
                if (tf) w.vtan[(size_t)p * w.mlv + k] = tan;
                vcount[p] = k + 1;
            }
            // ConnectLightPathToCamera (VCM.cpp:190-242)
            const CameraParams& cam = sl.cam;
            f3 to_pos = x - cam.position;
            if (dot3(cam.forward, to_pos) > 0.0f) {
                float zc = dot3(to_pos, cam.cam_z);
                if (zc > 0.0f) {
                    float cx = -dot3(to_pos, cam.cam_x) / (dot3(cam.cam_x, cam.cam_x) * zc);
                    float cy = dot3(to_pos, cam.cam_y) / (dot3(cam.cam_y, cam.cam_y) * zc);
                    float px = (cx + 1.0f) * 0.5f * (float)sl.W;
                    float py = (1.0f - cy) * 0.5f * (float)sl.H;
                    if (px >= 0.0f && py >= 0.0f && px < (float)sl.W && py < (float)sl.H) {
                        float dist = length3(to_pos);
                        f3 to_cam = to_pos * (-1.0f / dist);
                        float cos_cam = dot3(cam.forward, -to_cam);
                        float fwd, rev;
                        // a connection in the surface's plane carries nothing (its BSDF value holds
                        // the cosine) and the BSDF is not evaluated there (0/0 in Disney's terms)
                        f3 f = mk3(0, 0, 0);
                        if (dot3(s.n, to_cam) != 0.0f) f = bsdf_eval(s, view, to_cam, fwd, rev);
                        if (!is_zero3(f)) {
                            float cp = cont_prob(s);
                            rev *= cp;
                            float img_dist = sl.vipd / cos_cam;
                            float img_to_solid = img_dist * img_dist / cos_cam;
                            float cos_to_cam = absf_(dot3(s.n, to_cam));
                            float img_to_surface = img_to_solid * cos_to_cam / (dist * dist);
                            float w_light = (img_to_surface / (float)sl.N) * (sl.vm_weight + dvcm + dvc * rev);
                            float mis = 1.0f / (w_light + 1.0f);
                            c = (thr * f) * (mis * img_to_solid / ((float)sl.N * dist * dist));
                            if (!is_zero3(c)) {
                                int ix = (int)floorf(px), iy = (int)floorf(py);
                                if (ix >= 0 && iy >= 0 && ix < (int)sl.W && iy < (int)sl.H) {
                                    emit = true;
                                    pix = sl.path_base + (uint32_t)iy * sl.W + (uint32_t)ix;
                                    sh_o = offset_ray_origin(s, to_cam, 0.1f);
                                    sh_d = to_cam;
                                    sh_tnear = s.error;
                                    sh_tfar = dist;
                                }
                            }
                        }
                    }
                }
            }
            f3 dir = d;
            if (scatter(u, sl, s, dir, thr, dvcm, dvc, dvm, plen)) {
                cont = plen + 2 < sl.L;
                next_o = x;
                next_d = dir;
                s3(ps.pos, p, x);
                s3(ps.dir, p, dir);
                s3(ps.thr, p, thr);
                ps.dvcm[p] = dvcm;
                ps.dvc[p] = dvc;
                ps.dvm[p] = dvm;
                ps.plen[p] = plen;
            }
            ps.rng[p] = u.r.state;
        }
    }
    uint32_t ks = block_append(emit, shadow_count);
    if (emit) {
        put_ray(shadow, ks, sh_o, sh_d, sh_tnear, sh_tfar, scn);
        s3(splat_c, ks, c);
        splat_pix[ks] = pix;
    }
    uint32_t kn = block_append(cont, next_count);
    if (cont) {
        put_ray(next_rays, kn, next_o, next_d, kRayEps, kFloatMax, scn);
        next_queue[kn] = p;
    }
}

__global__ __launch_bounds__(kBlock) void k_splat(const uint32_t* __restrict count_ptr, const uint32_t* __restrict occluded,
                                                  const float* __restrict splat_c, const uint32_t* __restrict splat_pix,
                                                  float* __restrict frame)
{
    const uint32_t count = *count_ptr;
    for (uint32_t k0 = blockIdx.x * kBlock; k0 < count; k0 += gridDim.x * kBlock) {
        const uint32_t k = k0 + threadIdx.x;
        if (k >= count || occluded[k]) continue;
        uint32_t pix = splat_pix[k];
        atomicAdd(&frame[3 * pix], splat_c[3 * k]);
        atomicAdd(&frame[3 * pix + 1], splat_c[3 * k + 1]);
        atomicAdd(&frame[3 * pix + 2], splat_c[3 * k + 2]);
    }
}

// path-ordered compaction of the light vertices (the order of Selas's lightVertices array)
__global__ __launch_bounds__(kBlock) void k_compact(Wave w, const LightVertex* __restrict verts,
                                                    const uint32_t* __restrict vcount, const uint32_t* __restrict offset,
                                                    const uint32_t* __restrict path_slice, LightVertex* __restrict cverts,
                                                    float* __restrict cpos, uint32_t* __restrict cslice)
{
    const uint32_t total = w.dims[1];
    for (uint32_t g0 = blockIdx.x * kBlock; g0 < total; g0 += gridDim.x * kBlock) {
        const uint32_t g = g0 + threadIdx.x;
        if (g >= total) continue;
        uint32_t n = vcount[g], base = offset[g], si = path_slice[g];
        for (uint32_t k = 0; k < n; ++k) {
            LightVertex v = verts[(size_t)g * w.mlv + k];
            cverts[base + k] = v;
            if (w.vtan) w.cvtan[base + k] = w.vtan[(size_t)g * w.mlv + k];
            cpos[3 * (base + k)] = v.pos[0];
            cpos[3 * (base + k) + 1] = v.pos[1];
            cpos[3 * (base + k) + 2] = v.pos[2];
            cslice[base + k] = si;
        }
    }
}

// ---- photon grid (BuildHashGrid, VCMHashGrid.cpp:48-97) ---------------------------------------------

// Per slice s: box[8 s + 0..2] = ord(min), box[8 s + 3..5] = ~ord(max), both reduced with atomicMin.
// Vertices are slice-contiguous, so a block almost always holds one slice: it then reduces through
// its waves and LDS and issues one atomic per value; a block that straddles slices falls back to
// one atomic per value per wave (or per lane, in a wave that straddles them).
__global__ __launch_bounds__(kBlock) void k_grid_bounds(const float* __restrict cpos, const uint32_t* __restrict cslice,
                                                        const uint32_t* __restrict nv_ptr, uint32_t* __restrict box)
{
    __shared__ uint32_t red[6][kBlock / 64];
    __shared__ uint32_t wslice[kBlock / 64];
    const uint32_t nv = *nv_ptr;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t wave = threadIdx.x >> 6;
    const bool valid = i < nv;
    uint32_t si = valid ? cslice[i] : 0xFFFFFFFFu;
    uint32_t v[6];
    for (int k = 0; k < 6; ++k) v[k] = 0xFFFFFFFFu;
    if (valid)
        for (int k = 0; k < 3; ++k) {
            float x = cpos[3 * i + k];
            v[k] = f2ord(x);
            v[3 + k] = ~f2ord(x);
        }
    // slice of the wave's first valid lane; the wave is uniform when every valid lane has it
    const uint64_t vmask = __ballot(valid);
    const uint32_t leader = vmask ? (uint32_t)(__ffsll((unsigned long long)vmask) - 1) : 0u;
    const uint32_t s0 = __shfl(si, leader, 64);
    const bool uniform = __ballot(valid && si != s0) == 0;
    if (uniform) {
        for (int k = 0; k < 6; ++k) {
            uint32_t x = v[k];
            for (int off = 32; off > 0; off >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, off, 64));
            v[k] = x;
        }
    } else if (valid) {
        for (int k = 0; k < 6; ++k) atomicMin(&box[8 * si + k], v[k]);
    }
    if (__lane_id() == 0) {
        // 0xFFFFFFFF: wave without valid lanes; 0xFFFFFFFE: wave already flushed per lane
        wslice[wave] = vmask == 0 ? 0xFFFFFFFFu : (uniform ? s0 : 0xFFFFFFFEu);
        for (int k = 0; k < 6; ++k) red[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t bs = 0xFFFFFFFFu;
        bool one = true;
        for (int q = 0; q < kBlock / 64; ++q) {
            const uint32_t ws = wslice[q];
            if (ws == 0xFFFFFFFFu || ws == 0xFFFFFFFEu) continue;
            if (bs == 0xFFFFFFFFu) bs = ws;
            else if (ws != bs) one = false;
        }
        if (one) {
            if (bs != 0xFFFFFFFFu) {
                uint32_t x = 0xFFFFFFFFu;
                for (int q = 0; q < kBlock / 64; ++q)
                    if (wslice[q] == bs) x = min(x, red[k][q]);
                atomicMin(&box[8 * bs + k], x);
            }
        } else {
            for (int q = 0; q < kBlock / 64; ++q) {
                const uint32_t ws = wslice[q];
                if (ws != 0xFFFFFFFFu && ws != 0xFFFFFFFEu) atomicMin(&box[8 * ws + k], red[k][q]);
            }
        }
    }
}

// key = the slice's first cell + hashed cell (a slice of N paths owns cells [path_base, path_base + N))
__global__ __launch_bounds__(kBlock) void k_grid_keys(Wave w, const float* __restrict cpos,
                                                      const uint32_t* __restrict cslice, const uint32_t* __restrict nv_ptr,
                                                      const uint32_t* __restrict box, uint32_t* __restrict keys,
                                                      uint32_t* __restrict vals, uint32_t* __restrict cell_counts)
{
    uint32_t nv = *nv_ptr;
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nv) return;
    const uint32_t si = cslice[i];
    const Slice& sl = w.slices[si];
    float lo[3] = {ord2f(box[8 * si]), ord2f(box[8 * si + 1]), ord2f(box[8 * si + 2])};
    int32_t c[3];
    for (int k = 0; k < 3; ++k) c[k] = (int32_t)floorf(sl.inv_cell * (cpos[3 * i + k] - lo[k]));
    uint32_t cell = sl.path_base + cell_hash(c[0], c[1], c[2], sl.N);
    keys[i] = cell;
    vals[i] = i;
    atomicAdd(&cell_counts[cell], 1u);
}

// Photons in grid order: slot q of the sorted (cell, vertex) list gets its vertex and position
// copied next to each other, so a cell's range is read as contiguous memory by the range search.
__global__ __launch_bounds__(kBlock) void k_grid_gather(const uint32_t* __restrict nv_ptr,
                                                        const uint32_t* __restrict sorted_idx,
                                                        const LightVertex* __restrict cverts, LightVertex* __restrict sverts,
                                                        float4* __restrict spos)
{
    uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= *nv_ptr) return;
    const LightVertex v = cverts[sorted_idx[q]];
    sverts[q] = v;
    spos[q] = make_float4(v.pos[0], v.pos[1], v.pos[2], 0.0f);
}

// ---- camera pass ----------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_cam_gen(Wave w, PathState ps, RayBuf rays, uint32_t* __restrict queue,
                                                    uint32_t* __restrict counters)
{
    const uint32_t total = w.dims[1];
    // blocks past the end leave at once (captured launches are sized for capacity)
    if (blockIdx.x * kBlock >= total) return;
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    bool live = false;
    uint32_t scn = 0;
    f3 o = mk3(0, 0, 0), d = mk3(0, 0, 0);
    if (g < total) {
        const uint32_t si = find_slice(w, g);
        const Slice& sl = w.slices[si];
        scn = sl.scene;
        const uint32_t i = g - sl.path_base;
        PcgUniforms u{rng_init(sl.seed, stream_id(sl.iteration, i, 1))};
        uint32_t x = i % sl.W, y = i / sl.W;
        float jx = (float)x + u.next(), jy = (float)y + u.next();
        d = camera_image_to_world_dir(sl.cam, jx, jy);
        o = sl.cam.position;
        float cos_cam = dot3(sl.cam.forward, d);
        float img_dist = sl.vipd / cos_cam;
        float cam_pdf_w = img_dist * img_dist / cos_cam;
        ps.slice[g] = si;
        s3(ps.pos, g, o);
        s3(ps.dir, g, d);
        s3(ps.thr, g, mk3(1, 1, 1));
        s3(ps.color, g, mk3(0, 0, 0));
        ps.dvcm[g] = (float)sl.N * (1.0f / cam_pdf_w);
        ps.dvc[g] = 0.0f;
        ps.dvm[g] = 0.0f;
        ps.plen[g] = 1;
        ps.rng[g] = u.r.state;
        live = sl.L > 1;  // a camera path of length 1 ends at the lens (VCM.cpp:585)
    }
    uint32_t k = block_append(live, &counters[0]);
    if (live) {
        put_ray(rays, k, o, d, kRayEps, kFloatMax, scn);
        queue[k] = g;
    }
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_cam_shade(Wave w, const uint32_t* __restrict count_ptr, const uint32_t* __restrict queue,
                                                      RayBuf rays, HitBuf hits, PathState ps, GridView G, CamSlots S,
                                                      CamVertex* __restrict cam_vm, RayBuf next_rays,
                                                      uint32_t* __restrict next_queue, RayBuf shadow,
                                                      uint32_t* __restrict next_count,
                                                      uint32_t* __restrict shadow_count,
                                                      const uint32_t* __restrict order, uint32_t* __restrict merge_key,
                                                      float2* __restrict cam_aux, uint32_t* __restrict pairs,
                                                      uint32_t* __restrict pair_count)
{
    const uint32_t nslots = w.nslots;
    const uint32_t count = *count_ptr;
    // blocks past the end leave at once (captured launches are sized for capacity)
    if (blockIdx.x * kBlock >= count) return;
    // queue entry of this lane: in queue order, or (order != NULL) in the order of a bin sort
    const uint32_t jj = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t j = order && jj < count ? order[jj] : jj;
    bool cont = false;
    uint32_t p = 0, si = 0;
    f3 next_o = mk3(0, 0, 0), next_d = mk3(0, 0, 0);
    // shadow rays of this lane, emitted below with wave-level appends (one per slot index)
    bool active = jj < count;
    Surface s;
    f3 x = mk3(0, 0, 0), view = mk3(0, 0, 0);
    bool hit = false;
    uint32_t geom_hit = 0;
    if (active) {
        p = queue[j];
        si = ps.slice[p];
        for (uint32_t k = 0; k < nslots; ++k) S.valid[(size_t)p * nslots + k] = 0;
        const Slice& sl = w.slices[si];
        uint32_t geom;
        f3 ng;
        const float* l2w;
        hit_geometry(sl, hits, j, geom, ng, l2w);
        f3 d = mk3(rays.dx[j], rays.dy[j], rays.dz[j]);
        if (geom == SGPU_INVALID_ID || geom >= sl.n_geoms) {
            const uint32_t plen_miss = ps.plen[p];
            if (sl.has_ibl) {
                // ConnectToSkyLight (VCM.cpp:171-186): unweighted when seen from the camera
                float pdf;
                const f3 rad = ibl_lookup(sl.ibl, d, pdf);
                float mis = 1.0f;
                if (plen_miss != 1) {
                    const float p_pick = 1.0f / (sl.n_lights + 1);
                    const float direct_pdf_a = pdf * p_pick;
                    const float emission_pdf_w = (ibl_position_pdf(sl.ibl) * pdf) * p_pick;
                    mis = 1.0f / (1.0f + (direct_pdf_a * ps.dvcm[p] + emission_pdf_w * ps.dvc[p]));
                }
                s3(ps.color, p, g3(ps.color, p) + g3(ps.thr, p) * (rad * mis));
            } else if (plen_miss == 1) {
                f3 col = g3(ps.color, p) + g3(ps.thr, p) * sl.background;
                s3(ps.color, p, col);
            }
        } else {
            hit = true;
            geom_hit = geom;
            f3 o = mk3(rays.ox[j], rays.oy[j], rays.oz[j]);
            const float* mat = sl.materials + 22 * (size_t)sl.geom_material[geom];
            float4 tan;
            if (hit_surface(sl, hits, j, geom, ng, l2w, o, d, mat, s, &tan)) w.cam_tan[j] = tan;
            x = s.position;
            view = -d;
        }
    }
    const Slice& sl = w.slices[si];
    PcgUniforms u{Pcg32{0, 1}};
    f3 thr = mk3(0, 0, 0);
    float dvcm = 0, dvc = 0, dvm = 0;
    uint32_t plen = 0;
    if (hit) {
        u.r = rng_load(ps, p, sl, 1);
        f3 pos = g3(ps.pos, p);
        thr = g3(ps.thr, p);
        dvcm = ps.dvcm[p];
        dvc = ps.dvc[p];
        dvm = ps.dvm[p];
        plen = ps.plen[p];
        f3 dd = pos - x;
        float d2 = dot3(dd, dd);
        float cos_v = absf_(dot3(s.n, view));
        dvcm *= d2;
        dvcm /= cos_v;
        dvc /= cos_v;
        dvm /= cos_v;
    }
    // ---- NEE (slot 0) ----
    {
        bool emit = false;
        f3 so = mk3(0, 0, 0), ld = mk3(0, 0, 0);
        float tn = 0, tf = 0;
        const uint32_t n_pick = sl.n_lights + sl.has_ibl;
        if (hit && plen + 1 < sl.L && n_pick > 0) {
            float p_pick = 1.0f / n_pick;
            uint32_t idx = (uint32_t)(u.next() * n_pick);
            if (idx >= n_pick) idx = n_pick - 1;
            const bool infinite = idx == sl.n_lights;
            float a = u.next(), b = u.next();
            bool ok = true;
            f3 rad;
            float dist, dpdf, epdf, cl;
            if (infinite) {
                // DirectIblLightSample (AreaLighting.cpp:293-314)
                float pdf;
                rad = ibl_direct(sl.ibl, a, b, ld, pdf);
                ok = pdf > 0.0f;
                dist = 1e36f;
                dpdf = pdf * p_pick;
                epdf = (ibl_position_pdf(sl.ibl) * pdf) * p_pick;
                cl = 1.0f;
            } else {
                const Light L = sl.lights[idx];
                if (L.type == 1) {
                    f3 dv = ld3(L.position) - x;
                    float dd2 = dot3(dv, dv);
                    dist = sqrtf(dd2);
                    ld = dv * (1.0f / dist);
                    rad = ld3(L.radiance);
                    dpdf = dd2 * p_pick;
                    epdf = (1.0f / (4.0f * kPi)) * p_pick;
                    cl = 1.0f;
                } else {
                    f3 ex = ld3(L.x), ez = ld3(L.z);
                    f3 y = ld3(L.position) - ex * 0.5f - ez * 0.5f + ex * a + ez * b;
                    f3 dv = y - x;
                    float dd2 = dot3(dv, dv);
                    dist = sqrtf(dd2);
                    ld = dv * (1.0f / dist);
                    cl = dot3(ld3(L.direction), -ld);
                    ok = cl > 0.0f;
                    float pdf_a = 1.0f / (length3(ex) * length3(ez));
                    rad = ld3(L.radiance);
                    dpdf = pdf_a * dd2 / cl * p_pick;
                    epdf = pdf_a * (cl / kPi) * p_pick;
                }
            }
            if (ok) {
                float fwd, rev;
                f3 f = mk3(0, 0, 0);
                if (dot3(s.n, ld) != 0.0f) f = bsdf_eval(s, view, ld, fwd, rev);
                if (!is_zero3(f)) {
                    float cp = cont_prob(s);
                    rev *= cp;
                    float cos_s = absf_(dot3(s.n, ld));
                    // only the image-based light can also be reached by the camera path itself
                    float w_light = infinite ? (fwd * cp) / dpdf : 0.0f;
                    float w_cam = (epdf * cos_s / (dpdf * cl)) * (sl.vm_weight + dvcm + dvc * rev);
                    float mis = 1.0f / (w_light + 1.0f + w_cam);
                    f3 c = (f * rad) * (mis / dpdf);
                    if (!is_zero3(c)) {
                        emit = true;
                        s3(S.c, (size_t)p * nslots + 0, thr * c);
                        so = offset_ray_origin(s, ld, 0.1f);
                        tn = s.error;
                        tf = dist;
                    }
                }
            }
        }
        uint32_t ks = block_append(emit, shadow_count);
        if (emit) {
            put_ray(shadow, ks, so, ld, tn, tf, sl.scene);
            S.ray_slot[ks] = p * nslots + 0;
            S.valid[(size_t)p * nslots + 0] = 2;
        }
    }
    // ---- vertex connections to light path p (slots 1..mlv): one pair per light vertex, queued
    // for k_cam_connect (entry j in the low 26 bits, the vertex's index on its path above) ----
    {
        uint32_t npair = 0;
        if (hit) {
            const uint32_t vb = p == 0 ? 0 : G.path_end[p - 1];
            npair = min(G.path_end[p] - vb, w.mlv);
        }
        // the block's pairs: one atomic, then each lane's run at its exclusive prefix
        const uint32_t base = block_reserve(npair, pair_count);
        for (uint32_t k = 0; k < npair; ++k) pairs[base + k] = j | (k << kPairShift);
        if (hit) cam_aux[j] = make_float2(dvc, s.error);
    }
    // ---- vertex merging (slot nslots - 1): handed to k_cam_merge ----
    if (active) {
        CamVertex v;
        v.geom_plen = kNoQuery;
        if (hit) {
            v.pos[0] = x.x;
            v.pos[1] = x.y;
            v.pos[2] = x.z;
            v.n[0] = s.n.x;
            v.n[1] = s.n.y;
            v.n[2] = s.n.z;
            v.view[0] = view.x;
            v.view[1] = view.y;
            v.view[2] = view.z;
            v.thr[0] = thr.x;
            v.thr[1] = thr.y;
            v.thr[2] = thr.z;
            v.dvcm = dvcm;
            v.dvm = dvm;
            v.geom_plen = (geom_hit << 5) | (plen & 31u);
            v.path = p;
            v.base[0] = s.base_color.x;
            v.base[1] = s.base_color.y;
            v.base[2] = s.base_color.z;
            v.pad = 0.0f;
        }
        cam_vm[j] = v;
        // the query's bin for the spatial order of the merging (none for a vertex without a query)
        if (merge_key) merge_key[j] = hit && *G.nv_ptr != 0 ? merge_bin(x, G.box + 8 * si) : kNoBin;
    }
    // ---- continuation ----
    if (hit) {
        f3 dir = -view;
        if (scatter(u, sl, s, dir, thr, dvcm, dvc, dvm, plen)) {
            cont = plen < sl.L;
            next_o = x;
            next_d = dir;
            s3(ps.pos, p, x);
            s3(ps.dir, p, dir);
            s3(ps.thr, p, thr);
            ps.dvcm[p] = dvcm;
            ps.dvc[p] = dvc;
            ps.dvm[p] = dvm;
            ps.plen[p] = plen;
        }
        ps.rng[p] = u.r.state;
    }
    uint32_t kn = block_append(cont, next_count);
    if (cont) {
        put_ray(next_rays, kn, next_o, next_d, kRayEps, kFloatMax, sl.scene);
        next_queue[kn] = p;
    }
}

// Vertex connections (ConnectVertices, VCM.cpp:631-666), one queued (camera vertex, light vertex)
// pair per lane, grid-stride: the camera vertex's surface is rebuilt from its merge record (the
// same inputs CalculateSurfaceParams had, so the same values), the light vertex's from its stored
// record; a connection that carries light gets its contribution slot and a shadow ray.
__global__ __launch_bounds__(kBlock) void k_cam_connect(Wave w, const uint32_t* __restrict count_ptr,
                                                        const uint32_t* __restrict pairs,
                                                        const CamVertex* __restrict cam_vm,
                                                        const float2* __restrict cam_aux, PathState ps, GridView G,
                                                        CamSlots S, RayBuf shadow, uint32_t* __restrict shadow_count)
{
    const uint32_t count = *count_ptr;
    const uint32_t nslots = w.nslots;
    for (uint32_t i0 = blockIdx.x * kBlock; i0 < count; i0 += gridDim.x * kBlock) {
        const uint32_t i = i0 + threadIdx.x;
        bool emit = false;
        f3 so = mk3(0, 0, 0), cd = mk3(0, 0, 0);
        float tn = 0, tf = 0;
        uint32_t slot = 0, scn = 0;
        if (i < count) {
            const uint32_t pr = pairs[i];
            const uint32_t j = pr & ((1u << kPairShift) - 1u), k = pr >> kPairShift;
            const CamVertex v = cam_vm[j];
            const uint32_t p = v.path;
            const Slice& sl = w.slices[ps.slice[p]];
            scn = sl.scene;
            const uint32_t vi = (p == 0 ? 0 : G.path_end[p - 1]) + k;
            const LightVertex lv = G.cverts[vi];
            const uint32_t plen = v.geom_plen & 31u;
            const uint32_t lplen = lv.geom_plen & 31u;
            // the reference stops at the first light vertex that makes the path too long; the
            // vertices of a light path are stored in order of length, so this test is the same
            if (lplen + 1 + plen <= sl.L) {
                const float2 aux = cam_aux[j];
                const f3 x = mk3(v.pos[0], v.pos[1], v.pos[2]);
                const f3 view = mk3(v.view[0], v.view[1], v.view[2]);
                const uint32_t geom = v.geom_plen >> 5;
                Surface s;
                s.position = x;
                s.error = aux.y;
                surface_material(sl.materials + 22 * (size_t)sl.geom_material[geom], mk3(v.n[0], v.n[1], v.n[2]), view,
                                 s);
                if (geom_has_tangents(sl, geom)) tangent_frame(s.n, w.cam_tan[j], s);
                s.base_color = mk3(v.base[0], v.base[1], v.base[2]);
                const float dvcm = v.dvcm, dvc = aux.x;
                f3 lpos = mk3(lv.pos[0], lv.pos[1], lv.pos[2]);
                f3 dv = lpos - x;
                float dist2 = dot3(dv, dv);
                float dist = sqrtf(dist2);
                cd = dv * (1.0f / dist);
                float cfwd, crev;
                // a connection in either surface's plane carries nothing (see the splats)
                f3 fc = mk3(0, 0, 0);
                if (dot3(cd, s.n) != 0.0f && dot3(-cd, mk3(lv.n[0], lv.n[1], lv.n[2])) != 0.0f)
                    fc = bsdf_eval(s, view, cd, cfwd, crev);
                if (!is_zero3(fc)) {
                    Surface ls;
                    vertex_surface(sl, lv, w.cvtan ? &w.cvtan[vi] : nullptr, ls);
                    float lfwd, lrev;
                    f3 lw = mk3(lv.wfix[0], lv.wfix[1], lv.wfix[2]);
                    f3 fl = bsdf_eval(ls, lw, -cd, lfwd, lrev);
                    if (!is_zero3(fl)) {
                        float cos_c = absf_(dot3(cd, s.n));
                        float cos_l = absf_(dot3(-cd, ls.n));
                        float ccp = cont_prob(s), lcp = cont_prob(ls);
                        cfwd *= ccp;
                        crev *= ccp;
                        lfwd *= lcp;
                        lrev *= lcp;
                        float c_pdf_a = cfwd * cos_l / dist2;
                        float l_pdf_a = lfwd * cos_c / dist2;
                        float w_light = c_pdf_a * (sl.vm_weight + lv.dvcm + lv.dvc * lrev);
                        float w_cam = l_pdf_a * (sl.vm_weight + dvcm + dvc * crev);
                        float mis = 1.0f / (w_light + 1.0f + w_cam);
                        f3 c = (fc * fl) * (mis / dist2);
                        if (!is_zero3(c)) {
                            f3 lthr = mk3(lv.thr[0], lv.thr[1], lv.thr[2]);
                            float sb;
                            so = offset_ray_origin(s, cd, 0.1f, &sb);
                            tn = s.error;
                            tf = dist - 16.0f * absf_(sb);
                            emit = true;
                            slot = p * nslots + 1 + k;
                            s3(S.c, (size_t)slot, (mk3(v.thr[0], v.thr[1], v.thr[2]) * lthr) * c);
                        }
                    }
                }
            }
        }
        const uint32_t ks = block_append(emit, shadow_count);
        if (emit) {
            put_ray(shadow, ks, so, cd, tn, tf, scn);
            S.ray_slot[ks] = slot;
            S.valid[(size_t)slot] = 2;
        }
    }
}

// One merge query per camera vertex, eight lanes per query: lane k of the group scans grid cell k of
// the 2 x 2 x 2 block SearchHashGrid visits (x fastest, then y, then z, VCMHashGrid.cpp:131-152)
// and the group adds its eight partial sums with a fixed butterfly ((c0+c1)+(c2+c3))+((c4+c5)+
// (c6+c7)). A cell's vertices are contiguous (grid order); positions are tested against the radius
// before the 64-byte vertex is read. MergeVertices (VCM.cpp:362-409) per vertex found.
__global__ __launch_bounds__(kBlock) void k_cam_resolve(const uint32_t* __restrict count_ptr, const uint32_t* __restrict occluded, CamSlots S)
{
    const uint32_t count = *count_ptr;
    for (uint32_t k0 = blockIdx.x * kBlock; k0 < count; k0 += gridDim.x * kBlock) {
        const uint32_t k = k0 + threadIdx.x;
        if (k >= count) continue;
        S.valid[S.ray_slot[k]] = occluded[k] ? 3u : 1u;
    }
}

// color += the path's valid slots, in slot order (NEE, connections, merging), as the reference adds them
__global__ __launch_bounds__(kBlock) void k_cam_gather(const uint32_t* __restrict count_ptr, const uint32_t* __restrict queue, uint32_t nslots,
                                                       CamSlots S, PathState ps)
{
    const uint32_t count = *count_ptr;
    for (uint32_t j0 = blockIdx.x * kBlock; j0 < count; j0 += gridDim.x * kBlock) {
        const uint32_t j = j0 + threadIdx.x;
        if (j >= count) continue;
        uint32_t p = queue[j];
        f3 col = g3(ps.color, p);
        for (uint32_t k = 0; k < nslots; ++k)
            if (S.valid[(size_t)p * nslots + k] == 1u) col += g3(S.c, (size_t)p * nslots + k);
        s3(ps.color, p, col);
    }
}

// acc (a job's running sum, n pixels) += frame + color of each of the job's slices first..first+count-1
__global__ __launch_bounds__(kBlock) void k_frame_add(Wave w, uint32_t first, uint32_t count, uint32_t n,
                                                      const float* __restrict color, const float* __restrict frame,
                                                      float* __restrict acc)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float a0 = acc[3 * i], a1 = acc[3 * i + 1], a2 = acc[3 * i + 2];
    for (uint32_t s = first; s < first + count; ++s) {
        const size_t g = (size_t)w.slices[s].path_base + i;
        a0 += frame[3 * g] + color[3 * g];
        a1 += frame[3 * g + 1] + color[3 * g + 1];
        a2 += frame[3 * g + 2] + color[3 * g + 2];
    }
    acc[3 * i] = a0;
    acc[3 * i + 1] = a1;
    acc[3 * i + 2] = a2;
}

__global__ __launch_bounds__(kBlock) void k_scale(float* __restrict fb, size_t n, float s)
{
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) fb[i] = fb[i] * s;
}

// ---- work order: counting sort of a bounce's queue entries by key (BinBufs) ---------------------------

// Shading key of queue entry j: the material of its hit's geometry, spread over the bins per scene of
// the wavefront; misses share the last bin.
__device__ __forceinline__ uint32_t material_bin(const Wave& w, const uint32_t* queue, const HitBuf& hits,
                                                 const PathState& ps, uint32_t j)
{
    const Slice& sl = w.slices[ps.slice[queue[j]]];
    uint32_t geom;
    f3 ng;
    const float* l2w;
    hit_geometry(sl, hits, j, geom, ng, l2w);
    if (geom == SGPU_INVALID_ID || geom >= sl.n_geoms) return kBins - 1;
    return (sl.geom_material[geom] + 97u * sl.scene) % (kBins - 1);
}

// Bin totals of the entries 0 .. *count_ptr - 1: one LDS histogram per block of kBlock x kBinItems
// entries, then one atomic per non-empty bin. MATERIAL: the keys are computed here (material_bin)
// and stored; otherwise they are read.
template <bool MATERIAL>
__global__ __launch_bounds__(kBlock) void k_bin_count(Wave w, const uint32_t* __restrict count_ptr,
                                                      const uint32_t* __restrict queue, HitBuf hits, PathState ps,
                                                      uint32_t* __restrict key, uint32_t* __restrict total)
{
    __shared__ uint32_t hist[kBins];
    const uint32_t count = *count_ptr;
    const uint32_t first = blockIdx.x * kBlock * kBinItems;
    if (first >= count) return;
    for (uint32_t b = threadIdx.x; b < kBins; b += kBlock) hist[b] = 0;
    __syncthreads();
    for (uint32_t k = 0; k < kBinItems; ++k) {
        const uint32_t j = first + k * kBlock + threadIdx.x;
        if (j >= count) break;
        uint32_t kb;
        if (MATERIAL) {
            kb = material_bin(w, queue, hits, ps, j);
            key[j] = kb;
        } else {
            kb = key[j];
        }
        if (kb != kNoBin) atomicAdd(&hist[kb], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kBins; b += kBlock)
        if (hist[b]) atomicAdd(&total[b], hist[b]);
}

// Exclusive scan of the bin totals into the cursors (one block of 1024 lanes, 4 bins each); the
// totals are cleared for the next sort and the entry count is written to *n.
constexpr uint32_t kScanLanes = 1024;
static_assert(kBins == 4 * kScanLanes, "k_bin_scan covers kBins with 4 bins per lane");
__global__ __launch_bounds__(kScanLanes) void k_bin_scan(uint32_t* __restrict total, uint32_t* __restrict cursor,
                                                         uint32_t* __restrict n)
{
    __shared__ uint32_t wave_sum[kScanLanes / 64];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t v[4], sum = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = total[4 * t + k];
        total[4 * t + k] = 0;
        sum += v[k];
    }
    uint32_t incl = sum;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 64);
        if (lane >= d) incl += y;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t q = 0; q < wave; ++q) base += wave_sum[q];
    uint32_t run = base + incl - sum;
    for (int k = 0; k < 4; ++k) {
        cursor[4 * t + k] = run;
        run += v[k];
    }
    if (t == kScanLanes - 1) *n = run;
}

// order[cursor[key] ...] = the entries, each block reserving its run of every bin with one atomic
// and placing its entries by their rank within the block (LDS atomics).
__global__ __launch_bounds__(kBlock) void k_bin_scatter(const uint32_t* __restrict count_ptr,
                                                        const uint32_t* __restrict key, uint32_t* __restrict cursor,
                                                        uint32_t* __restrict order)
{
    __shared__ uint32_t hist[kBins];
    const uint32_t count = *count_ptr;
    const uint32_t first = blockIdx.x * kBlock * kBinItems;
    if (first >= count) return;
    for (uint32_t b = threadIdx.x; b < kBins; b += kBlock) hist[b] = 0;
    __syncthreads();
    uint32_t kb[kBinItems], rank[kBinItems];
    for (uint32_t k = 0; k < kBinItems; ++k) {
        const uint32_t j = first + k * kBlock + threadIdx.x;
        kb[k] = j < count ? key[j] : kNoBin;
        rank[k] = kb[k] != kNoBin ? atomicAdd(&hist[kb[k]], 1u) : 0u;
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kBins; b += kBlock)
        if (hist[b]) hist[b] = atomicAdd(&cursor[b], hist[b]);
    __syncthreads();
    for (uint32_t k = 0; k < kBinItems; ++k)
        if (kb[k] != kNoBin) order[hist[kb[k]] + rank[k]] = first + k * kBlock + threadIdx.x;
}

// Stream-ordered device allocations freed together; a dry-run arena only adds up their sizes.
struct Arena {
    std::vector<void*> ptrs;
    hipStream_t stream = nullptr;
    hipError_t err = hipSuccess;
    bool dry = false;
    uint64_t bytes = 0;
    template <typename T>
    T* get(size_t count)
    {
        const size_t n = count * sizeof(T) + 64;
        bytes += n;
        if (dry) return reinterpret_cast<T*>((uintptr_t)256);
        void* p = nullptr;
        if (err == hipSuccess) err = hipMallocAsync(&p, n, stream);
        if (err == hipSuccess) ptrs.push_back(p);
        return (T*)p;
    }
    RayBuf rays(size_t n, bool with_scene)
    {
        RayBuf r;
        r.scene = with_scene ? get<uint32_t>(n) : nullptr;
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

// Device time per phase, from event pairs recorded around each phase's launches on the render's
// stream (only when asked for: SGPU_VCM_KERNEL_TIMES). collect() runs after a stream sync.
enum { kPhaseClosest, kPhaseShadow, kPhaseLight, kPhaseCamera, kPhaseGrid, kPhaseMerge, kPhases };
using PhaseTimer = sgpu_phase_timer;
static_assert(kPhases <= sgpu_phase_timer::kMaxPhases, "too many phases for the timer");

sgpu_rays_soa as_soa(const RayBuf& r) { return sgpu_rays_soa{r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, r.tnear, r.tfar}; }

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// One admitted job of a server.
struct ServerJob {
    uint64_t id;
    sgpu_scene scene;
    sgpu_vcm_job job;
    float* fb;
    Slice base;      // everything but the per-iteration values
    uint32_t next;   // next iteration offset to schedule
};

}  // namespace

// A persistent VCM executor: wavefront buffers sized once, jobs admitted at any time and advanced
// wavefront by wavefront (see the file comment for what a slice and a wavefront are).
// What a captured pass depends on besides the server's own buffers.
struct GraphKey {
    const void* table = nullptr;
    const void* spill = nullptr;
    const void* tan = nullptr;
    uint32_t max_stack = 0, tree = 0, blocks = 0, lds = 0, inst = 0, order = 0;
    bool operator==(const GraphKey& o) const
    {
        return table == o.table && spill == o.spill && tan == o.tan && max_stack == o.max_stack && tree == o.tree &&
               blocks == o.blocks && lds == o.lds && inst == o.inst && order == o.order;
    }
};

struct sgpu_vcm_server_s {
    sgpu_context ctx = nullptr;
    sgpu_vcm_server_config cfg;
    hipStream_t alloc_stream = nullptr;
    uint32_t TM = 0, mlv = 0, nslots = 0, max_slices = 0;
    uint32_t cap = 0;  // paths per wavefront in use (<= TM: a cached server may serve smaller requests)
    unsigned key_bits = 1;
    Arena A;
    // wavefront buffers
    PathState ps;
    RayBuf rays[2], shadow;
    uint32_t* queue[2] = {nullptr, nullptr};
    HitBuf hb;
    uint32_t *occl = nullptr, *splat_pix = nullptr, *vcount = nullptr, *voffset = nullptr, *path_end = nullptr;
    float *splat_c = nullptr, *cpos = nullptr, *frame = nullptr;
    LightVertex *verts = nullptr, *cverts = nullptr, *sverts = nullptr;
    float4* spos = nullptr;
    uint32_t *cslice = nullptr, *keys = nullptr, *keys2 = nullptr, *vals = nullptr, *vals2 = nullptr;
    uint32_t *cell_counts = nullptr, *cell_ends = nullptr, *box = nullptr, *nv_dev = nullptr;
    // queue counts: per pass (light, camera) and bounce b, 4 words: [0] rays of bounce b, [1] its
    // shadow rays; the last word of the array is the traversal stack-overflow flag
    uint32_t Lb = 0;  // bounce bound of a pass (the largest path length)
    uint32_t* cnt = nullptr;
    uint32_t* ovf = nullptr;
    size_t cnt_words = 0;
    uint32_t* dims = nullptr;  // device: slices and paths of the current wavefront
    uint32_t* h_dims = nullptr;
    uint32_t* h_cnt = nullptr;
    // the two passes captured as graphs (valid while the key's pointers and tuning are unchanged)
    hipGraphExec_t g_pass[2] = {nullptr, nullptr};
    hipStream_t cap_stream = nullptr;
    GraphKey gkey;
    bool graph_failed = false;
    Slice* d_slices = nullptr;
    CamVertex* cam_vm = nullptr;
    float2* cam_aux = nullptr;  // per camera entry: dvc and the hit's error bound (vertex connections)
    uint32_t* pairs = nullptr;  // queued vertex connections of a bounce (TM x mlv)
    BinBufs bins;  // the work-order sort (tuning vcm_order)
    uint32_t* spill = nullptr;
    uint32_t spill_cap = 0;           // stack entries the spill buffer is sized for
    uint32_t* spill_grown = nullptr;  // a larger spill buffer allocated at submit (freed at destroy)
    // tangents of vertices shaded with a tangent frame (Wave::vtan, cvtan, cam_tan): one allocation
    // made when the first job on a scene with vertex tangents is admitted (freed at destroy)
    float4* vtan = nullptr;
    float4* cvtan = nullptr;
    float4* cam_tan = nullptr;
    CamSlots S;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    uint64_t* d_trav = nullptr;
    uint8_t* hbuf = nullptr;  // pinned: counters, slice table, scene table
    uint32_t* h = nullptr;
    Slice* h_slices = nullptr;
    TraceScene* h_table = nullptr;
    bool any_instanced = false;  // an instanced scene is in the table (traversal by k_trace8i)
    bool any_curves = false;     // a scene of the table has curves (the curve pass follows traversal)
    // the scenes of the admitted jobs (traversal table, rebuilt at every step) and their unique
    // ids, jobs in admission order
    std::vector<sgpu_scene> scenes;
    std::vector<uint64_t> scene_uids;
    TraceScene* d_table = nullptr;
    uint32_t table_cap = 0;
    bool table_dirty = false;
    uint32_t max_stack = 0;  // largest traversal stack of the scenes
    std::vector<ServerJob> active;
    uint64_t next_id = 1;
    // totals
    uint64_t n_closest = 0, n_shadow = 0, n_light_verts = 0, n_launches = 0, n_samples = 0, n_waves = 0;
    uint64_t trav_nodes = 0, trav_tris = 0;
    uint64_t n_merge_launches = 0, merge_cnt[4] = {0, 0, 0, 0};
    double ms[kPhases] = {};
};

namespace {

constexpr uint32_t kMaxSlicesPerWave = 4096;

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict p, uint32_t v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) p[i] = v;
}

// n words of p = v: a kernel, not a memset node, so that it may sit in a captured graph
void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t stream)
{
    if (n == 0) return;
    const uint32_t grid = (uint32_t)std::min<size_t>((n + kBlock - 1) / kBlock, 2048);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(kBlock), 0, stream, p, v, n);
}

// Captures the light and camera passes (pass_fn(pass, capture_stream)) as graphs unless the ones
// held are still valid. A capture that fails leaves the server on eager launches.
template <class F>
int server_graphs(sgpu_vcm_server sv, F&& pass_fn)
{
    const sgpu_context ctx = sv->ctx;
    GraphKey key;
    key.table = sv->d_table;
    key.spill = sv->spill;
    key.tan = sv->vtan;
    key.max_stack = sv->max_stack;
    key.tree = ctx->trace_tree;
    key.blocks = ctx->trace_blocks_per_cu;
    key.lds = ctx->trace_lds_nodes;
    key.inst = (sv->any_instanced ? 1u : 0u) | (sv->any_curves ? 2u : 0u);
    key.order = ctx->vcm_order | (ctx->vcm_merge_lanes << 8);
    if (sv->g_pass[0] && sv->g_pass[1] && key == sv->gkey) return SGPU_OK;
    for (int k = 0; k < 2; ++k) {
        if (sv->g_pass[k]) (void)hipGraphExecDestroy(sv->g_pass[k]);
        sv->g_pass[k] = nullptr;
    }
    if (!sv->cap_stream) SGPU_HIP_TRY(hipStreamCreateWithFlags(&sv->cap_stream, hipStreamNonBlocking));
    for (int pass = 0; pass < 2; ++pass) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(sv->cap_stream, hipStreamCaptureModeThreadLocal);
        int rc = e == hipSuccess ? pass_fn(pass, sv->cap_stream) : SGPU_E_HIP;
        if (e == hipSuccess) e = hipStreamEndCapture(sv->cap_stream, &g);
        if (rc == SGPU_OK && e == hipSuccess && g) e = hipGraphInstantiate(&sv->g_pass[pass], g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (rc != SGPU_OK || e != hipSuccess || !sv->g_pass[pass]) {
            (void)hipGetLastError();
            for (int k = 0; k < 2; ++k) {
                if (sv->g_pass[k]) (void)hipGraphExecDestroy(sv->g_pass[k]);
                sv->g_pass[k] = nullptr;
            }
            sv->graph_failed = true;
            return SGPU_OK;
        }
    }
    sv->gkey = key;
    return SGPU_OK;
}
constexpr uint32_t kDefaultStackCap = 96;  // stack spill sized at creation (grown for deeper trees)

int server_alloc(sgpu_vcm_server sv)
{
    Arena& A = sv->A;
    const uint32_t TM = sv->TM, mlv = sv->mlv, nslots = sv->nslots;
    PathState& ps = sv->ps;
    ps.pos = A.get<float>(3 * (size_t)TM);
    ps.dir = A.get<float>(3 * (size_t)TM);
    ps.thr = A.get<float>(3 * (size_t)TM);
    ps.dvcm = A.get<float>(TM);
    ps.dvc = A.get<float>(TM);
    ps.dvm = A.get<float>(TM);
    ps.plen = A.get<uint32_t>(TM);
    ps.slice = A.get<uint32_t>(TM);
    ps.rng = A.get<uint64_t>(TM);
    ps.color = A.get<float>(3 * (size_t)TM);
    sv->rays[0] = A.rays(TM, true);
    sv->rays[1] = A.rays(TM, true);
    sv->queue[0] = A.get<uint32_t>(TM);
    sv->queue[1] = A.get<uint32_t>(TM);
    HitBuf& hb = sv->hb;
    hb.t = A.get<float>(TM);
    hb.u = A.get<float>(TM);
    hb.v = A.get<float>(TM);
    hb.ngx = A.get<float>(TM);
    hb.ngy = A.get<float>(TM);
    hb.ngz = A.get<float>(TM);
    hb.prim = A.get<uint32_t>(TM);
    hb.geom = A.get<uint32_t>(TM);
    hb.inst = A.get<uint32_t>(TM);
    const size_t max_shadow = (size_t)TM * (mlv + 1);
    const size_t max_verts = (size_t)TM * mlv;
    sv->shadow = A.rays(max_shadow, true);
    sv->occl = A.get<uint32_t>(max_shadow);
    sv->splat_c = A.get<float>(3 * (size_t)TM);
    sv->splat_pix = A.get<uint32_t>(TM);
    sv->verts = A.get<LightVertex>(max_verts);
    sv->cverts = A.get<LightVertex>(max_verts);
    sv->sverts = A.get<LightVertex>(max_verts);
    sv->spos = A.get<float4>(max_verts);
    sv->cpos = A.get<float>(3 * max_verts);
    sv->cslice = A.get<uint32_t>(max_verts);
    sv->vcount = A.get<uint32_t>((size_t)TM + 1);
    sv->voffset = A.get<uint32_t>((size_t)TM + 1);
    sv->path_end = A.get<uint32_t>(TM);
    sv->keys = A.get<uint32_t>(max_verts);
    sv->keys2 = A.get<uint32_t>(max_verts);
    sv->vals = A.get<uint32_t>(max_verts);
    sv->vals2 = A.get<uint32_t>(max_verts);
    sv->cell_counts = A.get<uint32_t>(TM);
    sv->cell_ends = A.get<uint32_t>(TM);
    sv->box = A.get<uint32_t>(8 * (size_t)sv->max_slices);
    sv->Lb = sv->cfg.max_path_length;
    sv->cnt_words = 8 * ((size_t)sv->Lb + 2) + 4;
    sv->cnt = A.get<uint32_t>(sv->cnt_words);
    sv->ovf = sv->cnt + sv->cnt_words - 1;
    sv->dims = A.get<uint32_t>(4);
    sv->nv_dev = A.get<uint32_t>(1);
    sv->frame = A.get<float>(3 * (size_t)TM);
    sv->d_slices = A.get<Slice>(sv->max_slices);
    sv->cam_vm = A.get<CamVertex>(TM);
    sv->cam_aux = A.get<float2>(TM);
    sv->pairs = A.get<uint32_t>((size_t)TM * mlv);
    sv->bins.key = A.get<uint32_t>(TM);
    sv->bins.order = A.get<uint32_t>(TM);
    sv->bins.total = A.get<uint32_t>(kBins);
    sv->bins.cursor = A.get<uint32_t>(kBins);
    sv->bins.n = A.get<uint32_t>(4);
    sv->d_trav = A.get<uint64_t>(6);  // [0..1] traversal counters, [2..5] merging counters
    sv->S.c = A.get<float>(3 * (size_t)TM * nslots);
    sv->S.valid = A.get<uint32_t>((size_t)TM * nslots);
    sv->S.ray_slot = A.get<uint32_t>(max_shadow);
    if (A.err != hipSuccess) return sgpu_fail(SGPU_E_NOMEM, "sgpu_vcm_server: device allocation failed");
    // cell keys are below the wavefront's path count: sort only the bits that can be set
    sv->key_bits = 1;
    while (sv->key_bits < 32 && (TM - 1) >> sv->key_bits) ++sv->key_bits;
    size_t sort_bytes = 0, scan_bytes = 0, scan2_bytes = 0;
    SGPU_HIP_TRY(rocprim::radix_sort_pairs(nullptr, sort_bytes, sv->keys, sv->keys2, sv->vals, sv->vals2, max_verts, 0,
                                           sv->key_bits, A.stream));
    SGPU_HIP_TRY(rocprim::exclusive_scan(nullptr, scan_bytes, sv->vcount, sv->voffset, 0u, (size_t)TM + 1,
                                         rocprim::plus<uint32_t>(), A.stream));
    SGPU_HIP_TRY(rocprim::inclusive_scan(nullptr, scan2_bytes, sv->cell_counts, sv->cell_ends, (size_t)TM,
                                         rocprim::plus<uint32_t>(), A.stream));
    sv->tmp_bytes = std::max(sort_bytes, std::max(scan_bytes, scan2_bytes));
    sv->tmp = A.get<uint8_t>(sv->tmp_bytes);
    sv->spill_cap = kDefaultStackCap;
    sv->spill = A.get<uint32_t>(sgpu_trace_spill_words(sv->ctx, kDefaultStackCap));
    if (A.err != hipSuccess) return sgpu_fail(SGPU_E_NOMEM, "sgpu_vcm_server: device allocation failed");
    if (A.dry) return SGPU_OK;
    const size_t slice_bytes = sizeof(Slice) * (size_t)sv->max_slices;
    const size_t hbytes = 64 + 64 + slice_bytes + sv->cnt_words * sizeof(uint32_t);
    SGPU_HIP_TRY(hipHostMalloc((void**)&sv->hbuf, hbytes, hipHostMallocDefault));
    sv->h = (uint32_t*)sv->hbuf;
    sv->h_dims = (uint32_t*)(sv->hbuf + 64);
    sv->h_slices = (Slice*)(sv->hbuf + 128);
    sv->h_cnt = (uint32_t*)(sv->hbuf + 128 + slice_bytes);
    return SGPU_OK;
}

// The scene table of the next step: the scenes of the admitted jobs in order of first admission,
// each job's table index in its base slice. Built anew at every step, so a retired job's scene
// (which may since have been released, and a new scene created at its address) is never reached
// through a stale entry; the table is re-uploaded when the list of scene ids changed.
void server_sync_scenes(sgpu_vcm_server sv)
{
    std::vector<sgpu_scene> list;
    std::vector<uint64_t> uids;
    bool inst = false, curves = false;
    uint32_t stack = 0;
    for (ServerJob& j : sv->active) {
        uint32_t k = 0;
        while (k < list.size() && list[k] != j.scene) ++k;  // admitted scenes are alive: addresses differ
        if (k == list.size()) {
            list.push_back(j.scene);
            uids.push_back(j.scene->uid);
            inst |= !j.scene->instances.empty();
            curves |= j.scene->n_curves > 0;
            stack = std::max(stack, j.scene->stack_cap);
        }
        j.base.scene = k;
    }
    if (uids != sv->scene_uids) {
        sv->scenes.swap(list);
        sv->scene_uids.swap(uids);
        sv->table_dirty = true;
    }
    // kept from earlier steps (reset with the server): a larger stack bound and the instanced
    // traversal serve every scene, and the captured graphs, which bake both in, stay valid
    sv->any_instanced |= inst;
    sv->any_curves |= curves;
    sv->max_stack = std::max(sv->max_stack, stack);
}

// A job leaves the server (retired or dropped): its scene may be changed or released again.
void server_forget_job(ServerJob& j)
{
    if (j.scene) j.scene->server_jobs.fetch_sub(1);
    j.scene = nullptr;
}

int server_upload_table(sgpu_vcm_server sv, hipStream_t stream)
{
    if (!sv->table_dirty) return SGPU_OK;
    SGPU_HIP_TRY(hipStreamSynchronize(stream));
    const uint32_t n = (uint32_t)sv->scenes.size();
    if (n > sv->table_cap) {
        if (sv->d_table) (void)hipFree(sv->d_table);
        if (sv->h_table) (void)hipHostFree(sv->h_table);
        sv->d_table = nullptr;
        sv->h_table = nullptr;
        sv->table_cap = std::max<uint32_t>(16, 2 * n);
        SGPU_HIP_TRY(hipMalloc((void**)&sv->d_table, sizeof(TraceScene) * sv->table_cap));
        SGPU_HIP_TRY(hipHostMalloc((void**)&sv->h_table, sizeof(TraceScene) * sv->table_cap, hipHostMallocDefault));
    }
    for (uint32_t k = 0; k < n; ++k)
        sv->h_table[k] = TraceScene{sv->scenes[k]->d_nodes,  sv->scenes[k]->d_wide,   sv->scenes[k]->d_tris,
                                    sv->scenes[k]->n_tris,   sv->scenes[k]->n_inst,   sv->scenes[k]->d_inst,
                                    sv->scenes[k]->d_cnodes, sv->scenes[k]->d_curves, sv->scenes[k]->n_curves};
    SGPU_HIP_TRY(hipMemcpyAsync(sv->d_table, sv->h_table, sizeof(TraceScene) * n, hipMemcpyHostToDevice, stream));
    SGPU_HIP_TRY(hipStreamSynchronize(stream));
    sv->table_dirty = false;
    return SGPU_OK;
}

}  // namespace

// A server made ready for a new set of jobs: no jobs, no scenes, zero totals, `cfg`'s wavefront
// size (at most the allocated one) and flags.
static void server_reset(sgpu_vcm_server sv, const sgpu_vcm_server_config& cfg)
{
    for (ServerJob& j : sv->active) server_forget_job(j);
    sv->active.clear();
    sv->scenes.clear();
    sv->scene_uids.clear();
    sv->table_dirty = true;
    sv->max_stack = 0;
    sv->any_instanced = false;
    sv->any_curves = false;
    sv->cfg.flags = cfg.flags;
    sv->cap = (uint32_t)std::min<uint64_t>(cfg.wavefront_paths, sv->TM);
    sv->n_closest = sv->n_shadow = sv->n_light_verts = sv->n_launches = sv->n_samples = sv->n_waves = 0;
    sv->trav_nodes = sv->trav_tris = 0;
    sv->n_merge_launches = 0;
    for (int k = 0; k < 4; ++k) sv->merge_cnt[k] = 0;
    for (int k = 0; k < kPhases; ++k) sv->ms[k] = 0.0;
}

int sgpu_vcm_server_create_impl(sgpu_context ctx, const sgpu_vcm_server_config* cfg, hipStream_t stream,
                                sgpu_vcm_server* out)
{
    sgpu_vcm_server sv = new (std::nothrow) sgpu_vcm_server_s;
    if (!sv) return sgpu_fail(SGPU_E_NOMEM, "sgpu_vcm_server_create: out of host memory");
    sv->ctx = ctx;
    sv->cfg = *cfg;
    sv->A.stream = stream;
    sv->TM = (uint32_t)cfg->wavefront_paths;
    sv->cap = sv->TM;
    sv->mlv = cfg->max_path_length > 2 ? cfg->max_path_length - 2 : 1;
    sv->nslots = sv->mlv + 2;
    sv->max_slices = kMaxSlicesPerWave;
    int rc = server_alloc(sv);
    if (rc) {
        sgpu_vcm_server_destroy_impl(sv, stream);
        return rc;
    }
    *out = sv;
    return SGPU_OK;
}

uint64_t sgpu_vcm_server_bytes(sgpu_context ctx, const sgpu_vcm_server_config* cfg)
{
    sgpu_vcm_server_s sv;
    sv.ctx = ctx;
    sv.cfg = *cfg;
    sv.A.dry = true;
    sv.TM = (uint32_t)cfg->wavefront_paths;
    sv.mlv = cfg->max_path_length > 2 ? cfg->max_path_length - 2 : 1;
    sv.nslots = sv.mlv + 2;
    sv.max_slices = kMaxSlicesPerWave;
    server_alloc(&sv);
    return sv.A.bytes;
}

int sgpu_vcm_server_destroy_impl(sgpu_vcm_server sv, hipStream_t stream)
{
    if (!sv) return SGPU_OK;
    sv->A.stream = stream;
    for (ServerJob& j : sv->active) server_forget_job(j);
    if (sv->hbuf) (void)hipHostFree(sv->hbuf);
    if (sv->h_table) (void)hipHostFree(sv->h_table);
    if (sv->d_table) (void)hipFree(sv->d_table);
    if (sv->spill_grown) (void)hipFreeAsync(sv->spill_grown, stream);
    if (sv->vtan) (void)hipFreeAsync(sv->vtan, stream);
    for (int k = 0; k < 2; ++k)
        if (sv->g_pass[k]) (void)hipGraphExecDestroy(sv->g_pass[k]);
    if (sv->cap_stream) (void)hipStreamDestroy(sv->cap_stream);
    delete sv;  // the arena frees its buffers, stream-ordered
    return SGPU_OK;
}

int sgpu_vcm_server_submit_impl(sgpu_vcm_server sv, sgpu_scene sc, const sgpu_vcm_job* jb, float* fb,
                                uint64_t* id_out, hipStream_t stream)
{
    const float pi = 3.141592653589793f;
    const uint32_t W = jb->camera.width, H = jb->camera.height, L = jb->max_path_length;
    if ((uint64_t)W * H > sv->cap) return sgpu_fail(SGPU_E_RANGE, "sgpu_vcm_server_submit: image larger than a wavefront");
    if (L > sv->cfg.max_path_length)
        return sgpu_fail(SGPU_E_RANGE, "sgpu_vcm_server_submit: path length above the server's maximum");
    if (sc->stack_cap > sv->spill_cap) {
        // a deeper tree than the spill buffer holds stacks for: a larger one, stream-ordered
        uint32_t* grown = nullptr;
        SGPU_HIP_TRY(hipMallocAsync((void**)&grown, sgpu_trace_spill_words(sv->ctx, sc->stack_cap) * sizeof(uint32_t),
                                    stream));
        if (sv->spill_grown) SGPU_HIP_TRY(hipFreeAsync(sv->spill_grown, stream));
        sv->spill_grown = grown;
        sv->spill = grown;
        sv->spill_cap = sc->stack_cap;
    }
    if (sc->has_tangents && !sv->vtan) {
        // the tangent arrays: 2 x (paths x vertex slots) for light vertices, paths for camera vertices
        const size_t nv = (size_t)sv->TM * sv->mlv;
        float4* t = nullptr;
        SGPU_HIP_TRY(hipMallocAsync((void**)&t, (2 * nv + sv->TM) * sizeof(float4), stream));
        sv->vtan = t;
        sv->cvtan = t + nv;
        sv->cam_tan = t + 2 * nv;
    }
    ServerJob j;
    j.id = sv->next_id++;
    j.scene = sc;
    j.job = *jb;
    j.fb = fb;
    j.next = 0;
    Slice& b = j.base;
    std::memset(&b, 0, sizeof(b));
    b.cam = camera_init(mk3(jb->camera.position[0], jb->camera.position[1], jb->camera.position[2]),
                        mk3(jb->camera.look_at[0], jb->camera.look_at[1], jb->camera.look_at[2]),
                        mk3(jb->camera.up[0], jb->camera.up[1], jb->camera.up[2]), jb->camera.fov_degrees, W, H);
    float hfov = jb->camera.fov_degrees * (pi / 180.f);
    b.vipd = (float)W / (2.0f * tanf(hfov * 0.5f));
    b.W = W;
    b.H = H;
    b.N = W * H;
    b.L = L;
    b.mlv = L > 2 ? L - 2 : 1;
    b.seed = jb->seed;
    b.scene = 0;  // the job's scene table index, set at each step (server_sync_scenes)
    b.materials = sc->d_materials;
    b.geom_material = sc->d_geom_material;
    b.n_geoms = sc->n_geom_material;
    b.lights = (const Light*)sc->d_lights;
    b.n_lights = sc->n_lights;
    b.background = mk3(sc->background[0], sc->background[1], sc->background[2]);
    b.geom_tex = sc->d_geom_tex;
    b.n_geom_tex = sc->n_geom_tex;
    b.geom_attr = sc->d_geom_attr;
    b.has_ibl = ibl_view(sc, b.ibl) ? 1u : 0u;
    if (!sc->instances.empty()) {
        b.materials = sc->d_g_materials;
        b.geom_material = sc->d_g_geom_material;
        b.n_geoms = sc->n_g_geoms;
        b.geom_tex = sc->d_g_geom_tex;
        b.n_geom_tex = sc->n_g_geoms;
        b.geom_attr = sc->d_g_geom_attr;
        b.instanced = 1;
        b.inst_geom_base = sc->d_inst_geom_base;
        b.inst_l2w = sc->d_inst_l2w;
    }
    SGPU_HIP_TRY(hipMemsetAsync(fb, 0, 3 * (size_t)b.N * sizeof(float), stream));
    sc->server_jobs.fetch_add(1);
    sv->active.push_back(j);
    if (id_out) *id_out = j.id;
    return SGPU_OK;
}

int sgpu_vcm_server_step_impl(sgpu_vcm_server sv, hipStream_t stream, uint64_t* done_ids, uint32_t max_done,
                              uint32_t* n_done, uint32_t* n_active)
{
    const float pi = 3.141592653589793f;
    if (n_done) *n_done = 0;
    if (sv->active.empty()) {
        if (n_active) *n_active = 0;
        return SGPU_OK;
    }
    const uint32_t flags = sv->cfg.flags;
    const bool one_slice = flags & SGPU_VCM_ONE_SLICE_PER_WAVEFRONT;
    server_sync_scenes(sv);
    int rc = server_upload_table(sv, stream);
    if (rc) return rc;

    // ---- the wavefront: consecutive iterations of the jobs in admission order ----
    struct Run {
        uint32_t job, first, count;  // job index in `active`, first slice, slices
    };
    std::vector<Run> runs;
    std::vector<uint32_t> order(sv->active.size());
    for (uint32_t k = 0; k < order.size(); ++k) order[k] = k;
    if (flags & SGPU_VCM_SHORTEST_FIRST) {
        auto left = [&](uint32_t k) {
            const ServerJob& j = sv->active[k];
            return (uint64_t)j.base.N * (j.job.iterations - j.next);
        };
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return left(a) < left(b); });
    }
    uint32_t ns = 0, total = 0;
    for (uint32_t oi = 0; oi < order.size() && ns < sv->max_slices; ++oi) {
        const uint32_t ji = order[oi];
        ServerJob& j = sv->active[ji];
        Run run{ji, ns, 0};
        while (j.next + run.count < j.job.iterations && ns < sv->max_slices) {
            if (ns > 0 && (one_slice || (uint64_t)total + j.base.N > sv->cap)) break;
            Slice sl = j.base;
            const uint32_t gi = j.job.iteration_base + j.next + run.count;
            // CalculateIterationConstants (VCMCommon.cpp:97-113), vmCount = W*H, vcCount = 1
            float iteration_index = (float)gi + 1.0f;
            float r = j.job.base_radius / pow_f(iteration_index, 0.5f * (1.0f - j.job.radius_alpha));
            float r2 = r * r;
            sl.radius_sq = r2;
            sl.vm_norm = 1.0f / (pi * r2 * sl.N);
            sl.vm_weight = pi * r2 * sl.N / 1;
            sl.vc_weight = 1 / (pi * r2 * sl.N);
            sl.inv_cell = 1.0f / (2.0f * r);
            sl.iteration = gi;
            sl.path_base = total;
            sv->h_slices[ns] = sl;
            total += sl.N;
            ++ns;
            ++run.count;
        }
        if (run.count) runs.push_back(run);
        if (one_slice && ns > 0) break;
        if (run.count == 0 || j.next + run.count < j.job.iterations) break;  // the wavefront is full
    }
    bool single = true;
    for (uint32_t k = 1; k < ns; ++k) single &= sv->h_slices[k].scene == sv->h_slices[0].scene;
    const uint32_t scene0 = sv->h_slices[0].scene;
    sgpu_scene first_scene = sv->scenes[scene0];

    const bool kernel_times = flags & SGPU_VCM_KERNEL_TIMES;
    uint64_t* d_trav = (flags & SGPU_VCM_COUNT_TRAVERSAL) ? sv->d_trav : nullptr;
    unsigned long long* d_mstat = (flags & SGPU_VCM_COUNT_MERGE) ? (unsigned long long*)(sv->d_trav + 2) : nullptr;
    // the two passes replay captured graphs unless launches are timed or counted one by one
    const bool graphs = sv->ctx->vcm_graphs && !kernel_times && !d_trav && !d_mstat && !sv->graph_failed;
    if (graphs) single = false;  // graph launches traverse through the scene table
    PhaseTimer T(kernel_times, stream);
    // per-ray scene indices are only written and read in multi-scene wavefronts
    RayBuf rays[2] = {sv->rays[0], sv->rays[1]}, shadow = sv->shadow;
    if (single) rays[0].scene = rays[1].scene = shadow.scene = nullptr;
    PathState& ps = sv->ps;
    HitBuf& hb = sv->hb;
    CamSlots& S = sv->S;
    uint32_t* h = sv->h;
    uint32_t** queue = sv->queue;
    const uint32_t TM = sv->TM;
    const uint32_t Lb = sv->Lb;
    const uint64_t max_shadow = (uint64_t)TM * (sv->mlv + 1);
    auto cnt_at = [&](int pass, uint32_t b) { return sv->cnt + 4 * ((size_t)pass * (Lb + 2) + b); };
    const uint32_t cus = (uint32_t)sv->ctx->props.multiProcessorCount;
    // eager launches: one thread per item; captured ones (sized for capacity): a grid-stride cap
    auto grid_for = [&](uint64_t n) {
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(blocks_for(n), 64ull * cus));
    };

    sv->h_dims[0] = ns;
    sv->h_dims[1] = total;
    SGPU_HIP_TRY(hipMemcpyAsync(sv->d_slices, sv->h_slices, sizeof(Slice) * ns, hipMemcpyHostToDevice, stream));
    SGPU_HIP_TRY(hipMemcpyAsync(sv->dims, sv->h_dims, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    if (d_trav || d_mstat) SGPU_HIP_TRY(hipMemsetAsync(sv->d_trav, 0, 6 * sizeof(uint64_t), stream));
    Wave wv;
    wv.slices = sv->d_slices;
    wv.dims = sv->dims;
    wv.mlv = sv->mlv;
    wv.nslots = sv->nslots;
    wv.vtan = sv->vtan;
    wv.cvtan = sv->cvtan;
    wv.cam_tan = sv->cam_tan;
    sv->n_samples += total;
    ++sv->n_waves;
    sgpu_hits_soa hsoa{hb.t, hb.u, hb.v, hb.prim, hb.geom, hb.ngx, hb.ngy, hb.ngz};
    GridView G;
    G.box = sv->box;
    G.cell_ends = sv->cell_ends;
    G.sverts = sv->sverts;
    G.spos = sv->spos;
    G.cverts = sv->cverts;
    G.path_end = sv->path_end;
    G.nv_ptr = sv->nv_dev;

    // counts, the frame and the per-path vertex counts start at zero
    fill_u32(sv->cnt, 0u, sv->cnt_words, stream);
    fill_u32((uint32_t*)sv->frame, 0u, 3 * (size_t)total, stream);
    fill_u32(sv->vcount, 0u, (size_t)total + 1, stream);
    fill_u32(sv->bins.total, 0u, kBins, stream);
    SGPU_HIP_TRY(hipGetLastError());
    const uint32_t vorder = sv->ctx->vcm_order;

    // One pass (0: light subpaths, 1: camera subpaths) on stream `st`. Eager: the host reads each
    // bounce's counts to size its launches and stops at the first empty bounce. Captured: every
    // bounce up to the path-length bound is launched at capacity and the kernels read the counts.
    auto run_pass = [&](int pass, hipStream_t st, bool eager) -> int {
        // one item per thread (the heavy kernels, which exit per block past the count) or grid-stride
        auto grid1 = [&](uint64_t n) { return std::max<uint32_t>(1, blocks_for(n)); };
        auto grid = [&](uint64_t n) { return eager ? grid1(n) : grid_for(n); };
        auto read_counts = [&](const uint32_t* c) -> int {
            SGPU_HIP_TRY(hipMemcpyAsync(h, c, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            SGPU_HIP_TRY(hipStreamSynchronize(st));
            return SGPU_OK;
        };
        auto trace = [&](const RayBuf& r, const sgpu_hits_soa* hits, uint32_t* occ, uint32_t n, const uint32_t* n_ptr,
                         bool any, uint64_t* d_st) -> int {
            sgpu_rays_soa rs = as_soa(r);
            uint32_t* comp = hits ? hb.inst : nullptr;
            if (!single)
                return sgpu_trace_launch_multi(sv->ctx, sv->d_table, r.scene, sv->max_stack, sv->ovf, &rs, hits, occ, n,
                                               any, d_st, st, sv->spill, n_ptr, sv->any_instanced, comp,
                                               sv->any_curves);
            return sgpu_trace_launch(first_scene, &rs, hits, occ, n, any, d_st, st, sv->spill, n_ptr, sv->ovf, comp);
        };
        // the entries 0 .. *c - 1 of this bounce into bins.order by key (material keys computed from
        // the hits, else read from bins.key); n bounds the count (the capacity for captured launches)
        auto bin_sort = [&](const uint32_t* c, uint32_t n, bool material, const uint32_t* q) {
            const uint32_t g = std::max<uint32_t>(1, (uint32_t)(((uint64_t)n + kBlock * kBinItems - 1) /
                                                                (kBlock * kBinItems)));
            if (material)
                hipLaunchKernelGGL(k_bin_count<true>, dim3(g), dim3(kBlock), 0, st, wv, c, q, hb, ps, sv->bins.key,
                                   sv->bins.total);
            else
                hipLaunchKernelGGL(k_bin_count<false>, dim3(g), dim3(kBlock), 0, st, wv, c, q, hb, ps, sv->bins.key,
                                   sv->bins.total);
            hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kScanLanes), 0, st, sv->bins.total, sv->bins.cursor, sv->bins.n);
            hipLaunchKernelGGL(k_bin_scatter, dim3(g), dim3(kBlock), 0, st, c, sv->bins.key, sv->bins.cursor,
                               sv->bins.order);
        };
        const bool shade_sorted = (vorder & SGPU_ORDER_SHADE_MATERIAL) != 0;
        const bool merge_sorted = (vorder & SGPU_ORDER_MERGE_CELLS) != 0;
        const int ph_pass = pass == 0 ? kPhaseLight : kPhaseCamera;
        SGPU_HIP_TRY(T.begin(ph_pass));
        if (pass == 0)
            hipLaunchKernelGGL(k_light_gen, dim3(grid1(eager ? total : TM)), dim3(kBlock), 0, st, wv, ps, rays[0],
                               queue[0], cnt_at(0, 0));
        else
            hipLaunchKernelGGL(k_cam_gen, dim3(grid1(eager ? total : TM)), dim3(kBlock), 0, st, wv, ps, rays[0],
                               queue[0], cnt_at(1, 0));
        SGPU_HIP_TRY(hipGetLastError());
        SGPU_HIP_TRY(T.end());
        for (uint32_t b = 0; b < Lb; ++b) {
            const int cur = b & 1;
            uint32_t* c_in = cnt_at(pass, b);
            uint32_t* c_next = cnt_at(pass, b + 1);
            uint32_t* c_shadow = c_in + 1;
            uint32_t n = TM;
            if (eager) {
                int rc2 = read_counts(c_in);
                if (rc2) return rc2;
                n = h[0];
                if (n == 0) break;
            }
            SGPU_HIP_TRY(T.begin(kPhaseClosest));
            int rc2 = trace(rays[cur], &hsoa, nullptr, n, c_in, false, d_trav);
            if (rc2) return rc2;
            SGPU_HIP_TRY(T.end());
            SGPU_HIP_TRY(T.begin(ph_pass));
            if (shade_sorted) bin_sort(c_in, n, true, queue[cur]);
            const uint32_t* shade_order = shade_sorted ? sv->bins.order : nullptr;
            if (pass == 0)
                hipLaunchKernelGGL(k_light_shade, dim3(grid1(n)), dim3(kBlock), 0, st, wv, c_in, queue[cur], rays[cur],
                                   hb, ps, sv->verts, sv->vcount, rays[1 - cur], queue[1 - cur], shadow, sv->splat_c,
                                   sv->splat_pix, c_next, c_shadow, shade_order);
            else {
                uint32_t* c_pairs = c_in + 2;
                hipLaunchKernelGGL(k_cam_shade, dim3(grid1(n)), dim3(kBlock), 0, st, wv, c_in, queue[cur], rays[cur],
                                   hb, ps, G, S, sv->cam_vm, rays[1 - cur], queue[1 - cur], shadow, c_next, c_shadow,
                                   shade_order, merge_sorted ? sv->bins.key : nullptr, sv->cam_aux, sv->pairs,
                                   c_pairs);
                SGPU_HIP_TRY(hipGetLastError());
                // the bounce's vertex connections, one pair per lane (grid-stride over the device count)
                hipLaunchKernelGGL(k_cam_connect, dim3(grid_for((uint64_t)n * sv->mlv)), dim3(kBlock), 0, st, wv,
                                   c_pairs, sv->pairs, sv->cam_vm, sv->cam_aux, ps, G, S, shadow, c_shadow);
            }
            SGPU_HIP_TRY(hipGetLastError());
            SGPU_HIP_TRY(T.end());
            if (pass == 1) {
                SGPU_HIP_TRY(T.begin(kPhaseMerge));
                if (merge_sorted) bin_sort(c_in, n, false, nullptr);
                // sorted: a multiple of the XCD count, so that the blocks can be numbered xcd-major
                const uint32_t mlanes = sv->ctx->vcm_merge_lanes;
                const uint32_t per_block = (64 / mlanes) * kMergeWaves;  // queries per block and stride
                uint32_t merge_blocks = std::min<uint32_t>((n + per_block - 1) / per_block, 16u * cus);
                if (merge_sorted) merge_blocks = (merge_blocks + kXcds - 1) / kXcds * kXcds;
                launch_cam_merge(mlanes, merge_blocks, st, wv, merge_sorted ? sv->bins.n : c_in,
                                 merge_sorted ? sv->bins.order : nullptr, sv->cam_vm, ps.slice, G, S, d_mstat);
                SGPU_HIP_TRY(hipGetLastError());
                SGPU_HIP_TRY(T.end());
            }
            uint64_t n_sh = max_shadow;
            if (eager) {
                rc2 = read_counts(c_in);
                if (rc2) return rc2;
                n_sh = h[1];
            }
            if (n_sh > 0) {
                SGPU_HIP_TRY(T.begin(kPhaseShadow));
                rc2 = trace(shadow, nullptr, sv->occl, (uint32_t)n_sh, c_shadow, true, nullptr);
                if (rc2) return rc2;
                SGPU_HIP_TRY(T.end());
                SGPU_HIP_TRY(T.begin(ph_pass));
                if (pass == 0)
                    hipLaunchKernelGGL(k_splat, dim3(grid(n_sh)), dim3(kBlock), 0, st, c_shadow, sv->occl,
                                       sv->splat_c, sv->splat_pix, sv->frame);
                else
                    hipLaunchKernelGGL(k_cam_resolve, dim3(grid(n_sh)), dim3(kBlock), 0, st, c_shadow, sv->occl, S);
                SGPU_HIP_TRY(hipGetLastError());
                SGPU_HIP_TRY(T.end());
            }
            if (pass == 1) {
                SGPU_HIP_TRY(T.begin(kPhaseCamera));
                hipLaunchKernelGGL(k_cam_gather, dim3(grid(n)), dim3(kBlock), 0, st, c_in, queue[cur], sv->nslots, S,
                                   ps);
                SGPU_HIP_TRY(hipGetLastError());
                SGPU_HIP_TRY(T.end());
            }
        }
        return SGPU_OK;
    };
    if (graphs) {
        rc = server_graphs(sv, [&](int pass, hipStream_t cs) { return run_pass(pass, cs, false); });
        if (rc) return rc;
    }

    // ---- light pass ----
    if (graphs && sv->g_pass[0]) {
        SGPU_HIP_TRY(hipGraphLaunch(sv->g_pass[0], stream));
    } else {
        rc = run_pass(0, stream, true);
        if (rc) return rc;
    }

    // ---- light vertices, in path order ----
    SGPU_HIP_TRY(T.begin(kPhaseGrid));
    SGPU_HIP_TRY(rocprim::exclusive_scan(sv->tmp, sv->tmp_bytes, sv->vcount, sv->voffset, 0u, (size_t)total + 1,
                                         rocprim::plus<uint32_t>(), stream));
    hipLaunchKernelGGL(k_compact, dim3(std::max<uint32_t>(1, blocks_for(total))), dim3(kBlock), 0, stream, wv, sv->verts, sv->vcount,
                       sv->voffset, ps.slice, sv->cverts, sv->cpos, sv->cslice);
    SGPU_HIP_TRY(hipGetLastError());
    // path_end[g] = offset[g + 1] (inclusive end); the vertex count is offset[total]
    SGPU_HIP_TRY(hipMemcpyAsync(sv->path_end, sv->voffset + 1, (size_t)total * sizeof(uint32_t),
                                hipMemcpyDeviceToDevice, stream));
    SGPU_HIP_TRY(hipMemcpyAsync(sv->nv_dev, sv->voffset + total, sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
    SGPU_HIP_TRY(T.end());
    SGPU_HIP_TRY(hipMemcpyAsync(h, sv->voffset + total, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    SGPU_HIP_TRY(hipStreamSynchronize(stream));
    const uint32_t nv = h[0];
    sv->n_light_verts += nv;

    // ---- photon grids ----
    SGPU_HIP_TRY(T.begin(kPhaseGrid));
    SGPU_HIP_TRY(hipMemsetAsync(sv->box, 0xFF, 8 * sizeof(uint32_t) * ns, stream));
    SGPU_HIP_TRY(hipMemsetAsync(sv->cell_counts, 0, (size_t)total * sizeof(uint32_t), stream));
    if (nv > 0) {
        hipLaunchKernelGGL(k_grid_bounds, dim3(blocks_for(nv)), dim3(kBlock), 0, stream, sv->cpos, sv->cslice,
                           sv->nv_dev, sv->box);
        SGPU_HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_grid_keys, dim3(blocks_for(nv)), dim3(kBlock), 0, stream, wv, sv->cpos, sv->cslice,
                           sv->nv_dev, sv->box, sv->keys, sv->vals, sv->cell_counts);
        SGPU_HIP_TRY(hipGetLastError());
        size_t need = 0;
        SGPU_HIP_TRY(rocprim::radix_sort_pairs(nullptr, need, sv->keys, sv->keys2, sv->vals, sv->vals2, (size_t)nv, 0,
                                               sv->key_bits, stream));
        if (need > sv->tmp_bytes) return sgpu_fail(SGPU_E_NOMEM, "sgpu_vcm_server: sort scratch too small");
        SGPU_HIP_TRY(rocprim::radix_sort_pairs(sv->tmp, sv->tmp_bytes, sv->keys, sv->keys2, sv->vals, sv->vals2,
                                               (size_t)nv, 0, sv->key_bits, stream));
        hipLaunchKernelGGL(k_grid_gather, dim3(blocks_for(nv)), dim3(kBlock), 0, stream, sv->nv_dev, sv->vals2,
                           sv->cverts, sv->sverts, sv->spos);
        SGPU_HIP_TRY(hipGetLastError());
    }
    SGPU_HIP_TRY(rocprim::inclusive_scan(sv->tmp, sv->tmp_bytes, sv->cell_counts, sv->cell_ends, (size_t)total,
                                         rocprim::plus<uint32_t>(), stream));
    SGPU_HIP_TRY(T.end());

    // ---- camera pass ----
    if (graphs && sv->g_pass[1]) {
        SGPU_HIP_TRY(hipGraphLaunch(sv->g_pass[1], stream));
    } else {
        rc = run_pass(1, stream, true);
        if (rc) return rc;
    }
    // each job's slices of this wavefront (consecutive, in iteration order) into its sum; a job
    // whose last iteration ran is scaled by 1 / iterations and retired
    SGPU_HIP_TRY(T.begin(kPhaseCamera));
    std::vector<uint32_t> finished;
    for (const Run& run : runs) {
        ServerJob& j = sv->active[run.job];
        hipLaunchKernelGGL(k_frame_add, dim3(blocks_for(j.base.N)), dim3(kBlock), 0, stream, wv, run.first, run.count,
                           j.base.N, ps.color, sv->frame, j.fb);
        SGPU_HIP_TRY(hipGetLastError());
        j.next += run.count;
        if (j.next == j.job.iterations) {
            const size_t nf = 3 * (size_t)j.base.N;
            hipLaunchKernelGGL(k_scale, dim3(blocks_for(nf)), dim3(kBlock), 0, stream, j.fb, nf,
                               1.0f / (float)j.job.iterations);
            SGPU_HIP_TRY(hipGetLastError());
            finished.push_back(run.job);
        }
    }
    SGPU_HIP_TRY(T.end());
    SGPU_HIP_TRY(T.collect());
    // per-bounce ray counts (and the stack-overflow word after them) for the totals
    SGPU_HIP_TRY(hipMemcpyAsync(sv->h_cnt, sv->cnt, sv->cnt_words * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    SGPU_HIP_TRY(hipStreamSynchronize(stream));
    for (int pass = 0; pass < 2; ++pass)
        for (uint32_t b = 0; b < Lb; ++b) {
            const uint32_t* c = sv->h_cnt + 4 * ((size_t)pass * (Lb + 2) + b);
            if (c[0] == 0) continue;
            sv->n_closest += c[0];
            sv->n_shadow += c[1];
            ++sv->n_launches;
            if (pass == 1) ++sv->n_merge_launches;
        }
    if (sv->h_cnt[sv->cnt_words - 1] != 0)
        return sgpu_fail(SGPU_E_RANGE, "sgpu_vcm_server_step: traversal stack overflow (results incomplete)");
    for (int k = kPhaseClosest; k < kPhases; ++k) sv->ms[k] += T.ms[k];
    if (d_trav || d_mstat) {
        uint64_t tr[6] = {0, 0, 0, 0, 0, 0};
        SGPU_HIP_TRY(hipMemcpy(tr, sv->d_trav, sizeof(tr), hipMemcpyDeviceToHost));
        sv->trav_nodes += tr[0];
        sv->trav_tris += tr[1];
        for (int k = 0; k < 4; ++k) sv->merge_cnt[k] += tr[2 + k];
    }
    std::sort(finished.begin(), finished.end());  // erase from the back
    uint32_t nd = 0;
    for (size_t k = finished.size(); k-- > 0;) {
        const uint32_t ji = finished[k];
        if (done_ids && nd < max_done) done_ids[nd] = sv->active[ji].id;
        ++nd;
        server_forget_job(sv->active[ji]);
        sv->active.erase(sv->active.begin() + ji);
    }
    if (n_done) *n_done = nd;
    if (n_active) *n_active = (uint32_t)sv->active.size();
    return SGPU_OK;
}

int sgpu_vcm_server_stats_impl(sgpu_vcm_server sv, sgpu_vcm_stats* st)
{
    std::memset(st, 0, sizeof(*st));
    st->rays = sv->n_closest + sv->n_shadow;
    st->closest_rays = sv->n_closest;
    st->shadow_rays = sv->n_shadow;
    st->closest_launches = sv->n_launches;
    st->light_vertices = sv->n_light_verts;
    st->samples = sv->n_samples;
    st->wavefronts = sv->n_waves;
    st->nodes_visited = sv->trav_nodes;
    st->tris_tested = sv->trav_tris;
    st->closest_ms = (float)sv->ms[kPhaseClosest];
    st->shadow_ms = (float)sv->ms[kPhaseShadow];
    st->light_ms = (float)sv->ms[kPhaseLight];
    st->camera_ms = (float)sv->ms[kPhaseCamera];
    st->grid_ms = (float)sv->ms[kPhaseGrid];
    st->merge_ms = (float)sv->ms[kPhaseMerge];
    st->merge_launches = sv->n_merge_launches;
    st->merge_entries = sv->merge_cnt[0];
    st->merge_queries = sv->merge_cnt[1];
    st->merge_visited = sv->merge_cnt[2];
    st->merge_near = sv->merge_cnt[3];
    return SGPU_OK;
}

// Paths per wavefront the one-shot renders aim for: enough lanes to fill the chip in every
// bounce's launch.

int sgpu_vcm_render_batch_impl(const sgpu_scene* scenes, const sgpu_vcm_job* jobs, float* const* framebuffers,
                               uint32_t n_jobs, uint32_t flags, sgpu_vcm_stats* stats, hipStream_t stream)
{
    sgpu_vcm_server_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    uint64_t largest = 1;
    uint32_t L = 1;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        largest = std::max<uint64_t>(largest, (uint64_t)jobs[j].camera.width * jobs[j].camera.height);
        L = std::max(L, jobs[j].max_path_length);
    }
    // wavefront capacity: the target, unless a slice is larger; never more than the batch needs
    uint64_t needed = 0;
    for (uint32_t j = 0; j < n_jobs; ++j)
        needed += (uint64_t)jobs[j].camera.width * jobs[j].camera.height * jobs[j].iterations;
    cfg.wavefront_paths = std::max<uint64_t>(largest, std::min<uint64_t>(scenes[0]->ctx->vcm_wavefront_paths, needed));
    cfg.max_path_length = L;
    cfg.flags = flags;
    // the context keeps the last server it built and reuses it while it is large enough (its
    // buffers are then already mapped); a call that finds it in use builds a private one
    sgpu_context ctx = scenes[0]->ctx;
    std::unique_lock<std::mutex> lk(ctx->vcm_mu, std::try_to_lock);
    sgpu_vcm_server sv = nullptr;
    int rc = SGPU_OK;
    if (lk.owns_lock() && ctx->vcm_cache && ctx->vcm_cache->TM >= cfg.wavefront_paths &&
        ctx->vcm_cache->cfg.max_path_length >= cfg.max_path_length) {
        sv = ctx->vcm_cache;
        server_reset(sv, cfg);
    } else {
        if (lk.owns_lock() && ctx->vcm_cache) {
            sgpu_vcm_server_destroy_impl(ctx->vcm_cache, stream);
            ctx->vcm_cache = nullptr;
        }
        rc = sgpu_vcm_server_create_impl(ctx, &cfg, stream, &sv);
        if (rc) return rc;
        if (lk.owns_lock()) ctx->vcm_cache = sv;
    }
    for (uint32_t j = 0; j < n_jobs && rc == SGPU_OK; ++j)
        rc = sgpu_vcm_server_submit_impl(sv, scenes[j], &jobs[j], framebuffers[j], nullptr, stream);
    uint32_t n_active = n_jobs;
    while (rc == SGPU_OK && n_active > 0) rc = sgpu_vcm_server_step_impl(sv, stream, nullptr, 0, nullptr, &n_active);
    if (rc == SGPU_OK && stats) sgpu_vcm_server_stats_impl(sv, stats);
    if (sv != ctx->vcm_cache || !lk.owns_lock()) {
        sgpu_vcm_server_destroy_impl(sv, stream);
    } else {
        // nothing of this call stays referenced: jobs that failed mid-way are dropped, scenes forgotten
        sgpu_vcm_server_config keep = sv->cfg;
        keep.flags = 0;
        server_reset(sv, keep);
    }
    return rc;
}

void sgpu_vcm_release_cached(sgpu_context ctx)
{
    std::lock_guard<std::mutex> lk(ctx->vcm_mu);
    if (ctx->vcm_cache) sgpu_vcm_server_destroy_impl(ctx->vcm_cache, nullptr);
    ctx->vcm_cache = nullptr;
}

int sgpu_vcm_render_impl(sgpu_scene scene, const sgpu_vcm_job* job, float* fb, sgpu_vcm_stats* stats,
                         hipStream_t stream)
{
    float* fbs[1] = {fb};
    sgpu_vcm_job j = *job;
    j.flags = 0;
    return sgpu_vcm_render_batch_impl(&scene, &j, fbs, 1, job->flags, stats, stream);
}
