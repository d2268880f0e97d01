// Types and device helpers shared by the VCM translation units (sgpu_vcm.hip: the passes, the
// photon grid and the server; sgpu_vcm_merge.hip: vertex merging, built with its own division
// and square-root settings). Everything here is the device-side layout of one wavefront.
#pragma once

#include "sgpu_camera.h"
#include "sgpu_ibl.h"
#include "sgpu_internal.h"
#include "sgpu_light.h"
#include "sgpu_ptex.h"
#include "sgpu_surface.h"

namespace sgpu_vcm {

using namespace sgpu;


constexpr int kBlock = 256;
constexpr float kFloatMax = 3.402823466e+38F;
// PathState::plen bit of a light path that starts at the image-based light, cleared at its first hit
constexpr uint32_t kInfiniteOrigin = 1u << 31;
constexpr float kRayEps = 0.00001f;  // RayPick's tnear in the VCM code (VCM.cpp:137)
constexpr uint32_t kNoQuery = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0xFFFFFFFFu;
    const uint32_t lane = __lane_id();
    const uint32_t below = __popcll(mask & ((1ull << lane) - 1ull));
    const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    return pred ? base + below : 0xFFFFFFFFu;
}

// Queue appends with one atomic per block (sgpu_internal.h), for this file's block size.
__device__ __forceinline__ uint32_t block_append(bool pred, uint32_t* counter)
{
    return sgpu_block_append<kBlock>(pred, counter);
}
__device__ __forceinline__ uint32_t block_reserve(uint32_t n, uint32_t* counter)
{
    return sgpu_block_reserve<kBlock>(n, counter);
}

struct RayBuf {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tnear, *tfar;
    uint32_t* scene;  // per ray, the scene it traverses (multi-scene wavefronts only, else nullptr)
};
struct HitBuf {
    float *t, *u, *v, *ngx, *ngy, *ngz;
    uint32_t *prim, *geom;
    uint32_t* inst;  // composed instance of the hit (instanced scenes)
};

__device__ __forceinline__ void put_ray(const RayBuf& r, uint32_t i, f3 o, f3 d, float tnear, float tfar,
                                        uint32_t scene)
{
    if (r.scene) r.scene[i] = scene;
    r.ox[i] = o.x;
    r.oy[i] = o.y;
    r.oz[i] = o.z;
    r.dx[i] = d.x;
    r.dy[i] = d.y;
    r.dz[i] = d.z;
    r.tnear[i] = tnear;
    r.tfar[i] = tfar;
}

// Light vertex, 80 B (VCMVertex, VCMCommon.h:36-46, with the surface data a connection needs in
// place of the hit record: position, normal, incoming direction, material, and the base colour
// CalculateSurfaceParams derived -- the filtered texel for textured materials).
struct LightVertex {
    float pos[3];
    float thr[3];
    float n[3];
    float wfix[3];  // direction toward the previous vertex of the light subpath
    float dvcm, dvc, dvm;
    uint32_t geom_plen;  // geometry << 5 | path length (path length < 32, as in PathState)
    float base[3];       // base colour (^2.2 applied)
    float cp;            // Russian-roulette continuation probability of the surface
};
static_assert(sizeof(LightVertex) == 80, "LightVertex must be 80 bytes");

// A queued vertex connection: the camera entry j | (the light vertex's index on its path << kPairShift)
// (wavefronts hold at most 2^26 paths and paths at most 31 vertices).
constexpr uint32_t kPairShift = 26;

// Camera vertex handed from k_cam_shade to k_cam_merge and k_cam_connect (80 B); geom_plen =
// kNoQuery: no hit.
struct CamVertex {
    float pos[3];
    float n[3];
    float view[3];
    float thr[3];
    float dvcm, dvm;
    uint32_t geom_plen;  // geometry << 5 | path length
    uint32_t path;
    float base[3];       // base colour (^2.2 applied)
    float pad;
};
static_assert(sizeof(CamVertex) == 80, "CamVertex must be 80 bytes");

// Subpath state (PathState, VCMCommon.h:48-60), SoA over the paths of a wavefront.
struct PathState {
    float* pos;  // 3
    float* dir;  // 3
    float* thr;  // 3
    float* dvcm;
    float* dvc;
    float* dvm;
    uint32_t* plen;
    uint32_t* slice;
    uint64_t* rng;
    float* color;  // 3, camera paths only
};

// One VCM iteration of one job (see the file comment).
struct Slice {
    CameraParams cam;
    float vipd;  // W / (2 tan(hfov / 2))
    uint32_t W, H, N, L, mlv;
    uint32_t path_base;  // first wavefront path of the slice; also its first grid cell
    uint32_t iteration;  // global iteration index (radius schedule and random streams)
    uint32_t job;        // index of the job whose image the slice adds to
    uint32_t scene;      // index of the job's scene in the wavefront's traversal table
    uint64_t seed;
    float vm_weight, vc_weight, vm_norm, radius_sq, inv_cell;
    const float* materials;  // shading records (sgpu::shading_record)
    const uint32_t* geom_material;
    uint32_t n_geoms;
    const Light* lights;
    uint32_t n_lights;
    f3 background;
    const PtexTex* geom_tex;  // per-face texture per geometry (texels == nullptr: none)
    uint32_t n_geom_tex;
    const GeomAttr* geom_attr;  // vertex attributes per geometry (NULL: no geometry has any)
    uint32_t has_ibl;         // 1: the image-based light is light n_lights of the uniform pick
    IblView ibl;
    // instanced scene: a hit's geometry is numbered inst_geom_base[k] + geom in the scene's list
    // (materials, geom_material and geom_tex above are the scene's gathered ones) and its normal is
    // carried to world space by inst_l2w[k]
    uint32_t instanced;
    const uint32_t* inst_geom_base;
    const float* inst_l2w;
};

// The geometry (in the slice's list) and world-space geometric normal of hit j.
__device__ __forceinline__ void hit_geometry(const Slice& sl, const HitBuf& hits, uint32_t j, uint32_t& geom, f3& ng,
                                             const float*& l2w)
{
    geom = hits.geom[j];
    ng = mk3(hits.ngx[j], hits.ngy[j], hits.ngz[j]);
    l2w = nullptr;
    if (sl.instanced && geom != SGPU_INVALID_ID) {
        const uint32_t k = hits.inst[j];
        geom += sl.inst_geom_base[k];
        l2w = sl.inst_l2w + 9 * (size_t)k;
        ng = xform_linear(l2w, ng);
    }
}

// CalculateSurfaceParams' eUsesPtex branch (SurfaceParameters.cpp:196-203): the filtered texel of
// the hit's face at its (u, v) when the material asks for it and its geometry has a texture.
__device__ __forceinline__ bool hit_texel(const Slice& sl, uint32_t geom, const float* mat, uint32_t prim, float u,
                                          float v, f3& texel)
{
    if (!((uint32_t)mat[21] & kMatUsesPtex) || geom >= sl.n_geom_tex) return false;
    const PtexTex tx = sl.geom_tex[geom];
    if (!tx.texels) return false;
    texel = ptex_eval(tx, (int)prim, u, v);
    return true;
}

// Surface of hit j of geometry geom (hit_geometry's outputs): CalculateSurfaceParams with the
// geometry's vertex attributes and per-face texture. Returns true when the frame came from the
// geometry's tangents (then *tangent holds (t, w)).
__device__ __forceinline__ bool hit_surface(const Slice& sl, const HitBuf& hits, uint32_t j, uint32_t geom, f3 ng,
                                            const float* l2w, f3 o, f3 d, const float* mat, Surface& s,
                                            float4* tangent)
{
    f3 texel = mk3(0, 0, 0);
    const bool tex = hit_texel(sl, geom, mat, hits.prim[j], hits.u[j], hits.v[j], texel);
    return surface_from_hit(o, d, hits.t[j], ng, sl.geom_attr ? sl.geom_attr + geom : nullptr, hits.prim[j], hits.u[j],
                            hits.v[j], l2w, mat, tex, texel, s, tangent);
}

// Whether geometry geom of a slice shades with a tangent frame (vertex tangents).
__device__ __forceinline__ bool geom_has_tangents(const Slice& sl, uint32_t geom)
{
    return sl.geom_attr && sl.geom_attr[geom].t != nullptr;
}

// Wavefront-wide arguments.
struct Wave {
    const Slice* slices;
    const uint32_t* dims;  // device: [0] slices, [1] paths in the wavefront (sum of the slices' N)
    uint32_t mlv;          // vertex slots per path (largest slice mlv)
    uint32_t nslots;       // camera contribution slots per path (mlv + 2)
    // tangents (t, w) of the vertices shaded with a tangent frame, beside the light vertices (per
    // path slot, then compacted in path order) and the camera merging vertices; NULL until a job on
    // a scene with vertex tangents is admitted
    float4* vtan;
    float4* cvtan;
    float4* cam_tan;
};

// Element i of an array of float triples; the index is 64-bit because 3 i exceeds 2^32 for the
// camera contribution slots of the largest wavefronts (2^26 paths x 31 slots).
__device__ __forceinline__ f3 g3(const float* a, size_t i) { return mk3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
__device__ __forceinline__ void s3(float* a, size_t i, f3 v)
{
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
}

__device__ __forceinline__ uint64_t stream_id(uint32_t iteration, uint32_t path, uint32_t pass)
{
    return ((uint64_t)iteration << 33) | ((uint64_t)path << 1) | pass;
}

__device__ __forceinline__ Pcg32 rng_init(uint64_t seed, uint64_t sid)
{
    return pcg32_seed(seed ^ (sid * 0x9E3779B97F4A7C15ull), sid);
}

__device__ __forceinline__ Pcg32 rng_load(const PathState& ps, uint32_t g, const Slice& sl, uint32_t pass)
{
    Pcg32 r;
    r.state = ps.rng[g];
    r.inc = (stream_id(sl.iteration, g - sl.path_base, pass) << 1u) | 1u;
    return r;
}

// slice of wavefront path g (slices are sorted by path_base)
__device__ __forceinline__ uint32_t find_slice(const Wave& w, uint32_t g)
{
    uint32_t lo = 0, hi = w.dims[0] - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (w.slices[mid].path_base <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ float cont_prob(const Surface& s) { return saturate_(max_(max_(s.base_color.x, s.base_color.y), s.base_color.z)); }

// Surface of a stored light vertex (the material transforms of CalculateSurfaceParams); the frame
// of its geometry's tangents when it has them (tan: the vertex's stored tangent).
__device__ __forceinline__ void vertex_surface(const Slice& sl, const LightVertex& v, const float4* tan, Surface& s)
{
    const uint32_t geom = v.geom_plen >> 5;
    const float* mat = sl.materials + 22 * (size_t)sl.geom_material[geom];
    s.position = mk3(v.pos[0], v.pos[1], v.pos[2]);
    s.error = 0.0f;
    surface_material(mat, mk3(v.n[0], v.n[1], v.n[2]), mk3(v.wfix[0], v.wfix[1], v.wfix[2]), s);
    if (tan && geom_has_tangents(sl, geom)) tangent_frame(s.n, *tan, s);
    s.base_color = mk3(v.base[0], v.base[1], v.base[2]);
}


// SampleBsdfScattering (VCM.cpp:412-446) with the Russian-roulette-compensated throughput.
template <class U>
__device__ __forceinline__ bool scatter(U& u, const Slice& sl, const Surface& s, f3& dir, f3& thr, float& dvcm,
                                        float& dvc, float& dvm, uint32_t& plen)
{
    BsdfSample bs;
    if (!bsdf_sample(u, s, -dir, bs)) return false;
    // a sample of pdf 0 (SampleDisney's diffuse transmission taken at diffTrans = 0 when the
    // uniform is exactly 0, with an infinite reflectance) ends the path
    if (!(bs.fwd_pdf > 0.0f)) return false;
    if (is_zero3(bs.reflectance)) return false;
    float cp = cont_prob(s);
    if (u.next() > cp) return false;
    float fwd = bs.fwd_pdf * cp, rev = bs.rev_pdf * cp;
    float cos_b = absf_(dot3(bs.wi, s.n));
    thr = (thr * bs.reflectance) * (1.0f / cp);
    float k = cos_b / fwd;
    float nvc = k * (dvc * rev + dvcm + sl.vm_weight);
    float nvm = k * (dvm * rev + dvcm * sl.vc_weight + 1.0f);
    dvc = nvc;
    dvm = nvm;
    dvcm = 1.0f / fwd;
    dir = bs.wi;
    ++plen;
    return true;
}

__device__ __forceinline__ uint32_t f2ord(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

__device__ __forceinline__ uint32_t cell_hash(int32_t x, int32_t y, int32_t z, uint32_t cell_count)
{
    uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
    return ((ux * 73856093u) ^ (uy * 19349663u) ^ (uz * 83492791u)) % cell_count;
}

struct GridView {
    const uint32_t* box;        // 8 per slice
    const uint32_t* cell_ends;  // inclusive ends over all cells of the wavefront
    const LightVertex* sverts;  // vertices in grid order (cell ranges are contiguous)
    const float4* spos;         // their positions
    const LightVertex* cverts;  // vertices in path order
    const uint32_t* path_end;   // inclusive end of each light path's vertices
    const uint32_t* nv_ptr;
};

struct CamSlots {
    float* c;            // nslots x 3 per path: NEE, VC 0..mlv-1, VM
    uint32_t* valid;     // nslots per path: 0 none, 1 unconditional, 2 pending shadow ray, 3 occluded
    uint32_t* ray_slot;  // per queued shadow ray: path * nslots + slot
};

// ---- work order: a device counting sort of a bounce's queue entries by a small key ----------------
// key[j] in [0, kBins) or kNoBin (entry j left out); the sort writes order[0 .. n) = the entries by
// key (entries of one key in no particular order) and n to its count word. Per bin: a block
// histogram in LDS, one global atomic per non-empty bin and block, a one-block scan of the bin
// totals, and the scatter reserving each block's run of every bin with one atomic.
constexpr uint32_t kBins = 4096;
constexpr uint32_t kNoBin = 0xFFFFFFFFu;
constexpr uint32_t kBinItems = 8;  // entries per thread of the binning kernels

struct BinBufs {
    uint32_t* key;     // per queue entry
    uint32_t* order;   // queue entries in key order
    uint32_t* total;   // per bin; zero between sorts (the scan clears it)
    uint32_t* cursor;  // per bin: next free position of the bin in `order`
    uint32_t* n;       // entries in `order`
};

// Bin of a merge query at x: the Morton code of x's cell in a 16 x 16 x 16 division of the photon
// grid's box (box: the slice's 8 words, ordered-integer min then complemented max), kNoBin for a
// point outside the box (the range search skips those: VCMHashGrid.cpp:113-117).
__device__ __forceinline__ uint32_t merge_bin(f3 x, const uint32_t* box)
{
    const float p[3] = {x.x, x.y, x.z};
    uint32_t code = 0;
    for (int k = 0; k < 3; ++k) {
        const float lo = ord2f(box[k]), hi = ord2f(~box[3 + k]);
        if (p[k] - lo < 0.0f || hi - p[k] < 0.0f) return kNoBin;
        const float ext = hi - lo;
        float q = ext > 0.0f ? (p[k] - lo) * (16.0f / ext) : 0.0f;
        q = fminf(fmaxf(q, 0.0f), 15.0f);
        const uint32_t c = (uint32_t)q;
        for (int b = 0; b < 4; ++b) code |= ((c >> b) & 1u) << (3 * b + k);
    }
    return code;
}

// XCDs of the chip: consecutive workgroups go to different XCDs in turn, so a launch that wants
// neighbouring work on one XCD's L2 numbers its blocks xcd-major.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_major_block(uint32_t b, uint32_t grid)
{
    if (grid % kXcds) return b;
    return (b % kXcds) * (grid / kXcds) + b / kXcds;
}

// Vertex merging of the camera vertices of one bounce (sgpu_vcm_merge.hip).
constexpr uint32_t kMergeWaves = kBlock / 64;
// lanes: lanes per range search, 8, 16 or 32 (tuning vcm_merge_lanes). count_ptr: entries of cam_vm
// to merge, or (order != NULL) entries of order, which names them. blocks: a grid-stride cap.
// mstat (NULL: not counted): [0] camera vertices read, [1] range searches, [2] photons visited,
// [3] photons within the radius, summed with atomics.
void launch_cam_merge(uint32_t lanes, uint32_t blocks, hipStream_t stream, const Wave& w, const uint32_t* count_ptr,
                      const uint32_t* order, const CamVertex* cam_vm, const uint32_t* path_slice, const GridView& G,
                      const CamSlots& S, unsigned long long* mstat);

}  // namespace sgpu_vcm
