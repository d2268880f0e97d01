// Internal declarations shared by the translation units of libselas_gpu.
#pragma once

#include "../../include/selas_gpu.h"
#include "sgpu_math.h"
#include "sgpu_ptex.h"
#include "sgpu_surface.h"

#include <hip/hip_runtime.h>
#include <atomic>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

// ---- device data layout of a committed scene ---------------------------------------------------
//
// BVH node: a binary node that stores the boxes of both children, laid out per axis so that
// the two slab tests of one visit use three float4 loads plus one int4 load (64 B):
//   x = (c0.lo.x, c0.hi.x, c1.lo.x, c1.hi.x), y = (...), z = (...), child = (c0, c1, 0, 0).
// A child index >= 0 names an internal node; a negative index ~j names the triangle record j
// (leaves hold one triangle).
struct BvhNode {
    float4 x;
    float4 y;
    float4 z;
    int4 child;
};
static_assert(sizeof(BvhNode) == 64, "BvhNode must be 64 bytes");

// Triangle record, 48 B, in BVH leaf order:
//   a = (v0.x, v0.y, v0.z, bits(global triangle index)),
//   b = (e1.x, e1.y, e1.z, bits(geomID)),    e1 = v1 - v0
//   c = (e2.x, e2.y, e2.z, bits(primID)),    e2 = v2 - v0
// The global triangle index orders triangles by (geomID, primID); it is the tie-break key.
struct TriRecord {
    float4 a;
    float4 b;
    float4 c;
};
static_assert(sizeof(TriRecord) == 48, "TriRecord must be 48 bytes");

// 8-wide node, 128 B (one cache line), built by collapsing the binary tree. Child boxes are
// quantized to 8 bits per plane relative to the node's box: plane = fma(q, 2^(e - 127), origin)
// in FP32 (q * 2^(e - 127) is exact, so the decode rounds once), with q chosen in the builder so
// that the decoded box contains the binary tree's (inflated) child box. child[c] >= 0: wide node
// index; < 0: ~(triangle record index); slots c >= nchild are unused.
struct WideNode {
    float ox, oy, oz;
    uint32_t meta;       // ex | ey << 8 | ez << 16 | nchild << 24 (biased exponents of the plane scales)
    uint32_t qlo[3][2];  // axis a, child c: byte (c & 3) of qlo[a][c >> 2]
    uint32_t qhi[3][2];
    int32_t child[8];
    uint32_t pad[8];
};
static_assert(sizeof(WideNode) == 128, "WideNode must be 128 bytes");
constexpr uint32_t kWideMaxPush = 7;  // stack entries a wide node can push

// One composed instance of an instanced scene as the traversal reads it (64 bytes): the rows of
// its world-to-local affine transform (x' = w.x x + w.y y + w.z z + w.w) and the 8-wide tree and
// triangle records of the scene of geometries it places.
struct InstRecord {
    float4 w2l[3];
    const WideNode* wide;
    const TriRecord* tris;
};
static_assert(sizeof(InstRecord) == 64, "InstRecord must be 64 bytes");
// Round B-spline curve segment, 80 B, in curve-tree leaf order: the four control points (xyz,
// radius; in world space, also for the curves of an instanced scene's placed scenes, which are
// carried into world space at its commit), the geometry id (in the placed scene for an instanced
// scene), the segment's index within its geometry and the composed instance (SGPU_INVALID_ID in a
// scene of geometries).
struct CurveRecord {
    float4 p[4];
    uint32_t geom, prim, inst, pad;
};
static_assert(sizeof(CurveRecord) == 80, "CurveRecord must be 80 bytes");
// Stack entries of the curve tree's traversal (a binary tree: its depth must stay below this).
constexpr uint32_t kCurveStack = 96;

// The traversal stack entry that returns from an instance's tree to the top tree.
constexpr int kInstExit = (int)0x80000000;

// Slots of the LDS part of the per-lane traversal stack; deeper entries spill to global memory.
#define SGPU_STACK_LDS 16
#define SGPU_TRACE_BLOCK 256

// The traversal grid is at most kMaxTraceBlocksPerCu blocks per CU (sgpu_tuning bound); spill
// buffers are sized for it, so any tuning within the bound fits any buffer.
constexpr uint32_t kMaxTraceBlocksPerCu = 32;

struct sgpu_context_s {
    int device;
    hipDeviceProp_t props;
    uint64_t vcm_wavefront_paths = 1u << 20;
    uint32_t trace_blocks_per_cu = 8;
    uint32_t trace_tree = 1;  // 0: binary nodes, 1: 8-wide compressed nodes
    uint32_t trace_lds_nodes = 0;  // top wide nodes read from LDS (single-scene launches)
    uint32_t vcm_graphs = 1;       // VCM passes replayed as captured graphs
    uint32_t vcm_order = 0;        // SGPU_ORDER_* bits: work order inside a VCM bounce
    uint32_t vcm_merge_lanes = 16;  // lanes per vertex-merging range search (8, 16 or 32)
    uint32_t bvh_builder = 0;       // binary tree under the 8-wide one: 0 Karras (LBVH), 1 PLOC
    uint32_t* noalloc_scratch = nullptr;  // device address of this device's sgpu_intersect_noalloc scratch
    // the VCM server of the last sgpu_render_vcm(_batch) call, reused while large enough
    std::mutex vcm_mu;
    sgpu_vcm_server vcm_cache = nullptr;
};

struct sgpu_scene_s {
    sgpu_context ctx;
    std::mutex mu;
    uint64_t uid = 0;  // unique in the process (addresses of released scenes are reused)
    // jobs on this scene admitted to VCM servers and not yet retired: while non-zero the scene
    // cannot be released and its shading data cannot be replaced (the jobs hold its device pointers)
    std::atomic<uint32_t> server_jobs{0};
    // host copy of the world-space triangles (9 floats each) until commit, in global index order
    std::vector<float> tri_pos;
    std::vector<uint32_t> tri_geom;   // geomID per triangle
    std::vector<uint32_t> geom_first; // first global index of each geometry
    std::vector<uint8_t> geom_quad;   // 1: quad geometry (triangles 2q and 2q+1 form quad q), 2: curves
    // round B-spline curves (sgpu_scene_add_curves): 16 floats per segment (control points, placed),
    // its geometry and index within it; kept after commit for the scenes that place this one
    std::vector<float> curve_cp;
    std::vector<uint32_t> curve_geom, curve_prim;
    // the curve tree (binary, over the segments, or over the placed scenes' curves carried into world
    // space for an instanced scene) and its records in leaf order
    BvhNode* d_cnodes = nullptr;
    CurveRecord* d_curves = nullptr;
    uint32_t n_curves = 0;
    bool committed = false;
    uint32_t n_tris = 0;
    uint32_t n_nodes = 0;
    uint32_t max_depth = 0;
    BvhNode* d_nodes = nullptr;
    TriRecord* d_tris = nullptr;
    WideNode* d_wide = nullptr;  // the same tree collapsed to 8-wide nodes (breadth-first order)
    uint32_t n_wide = 0;
    uint32_t wide_depth = 0;
    uint32_t stack_cap = 0;      // traversal stack entries either tree can need
    // set to non-zero by a traversal whose stack would exceed the tree depth (never expected)
    uint32_t* d_overflow = nullptr;
    // shading data (sgpu_scene_set_materials / sgpu_scene_set_lights)
    float* d_materials = nullptr;      // shading record per material (22 floats, sgpu::shading_record)
    uint32_t n_materials = 0;
    uint32_t* d_geom_material = nullptr;  // material index per geometry
    uint32_t n_geom_material = 0;
    void* d_lights = nullptr;          // sgpu_light records
    uint32_t n_lights = 0;
    float background[3] = {0.0f, 0.0f, 0.0f};
    // vertex attributes (sgpu_scene_set_vertex_attributes). Until commit, per geometry: its index
    // buffer in primitive vertex order (3 or 4 per primitive), vertex count and placement transform
    // (column-major 4x4, empty for none), to expand per-vertex values to corners. attr_host[g] points
    // into attr_allocs (device corner arrays); d_geom_attr is the device table of attr_host (NULL
    // when no geometry has attributes)
    std::vector<std::vector<uint32_t>> geom_idx;
    std::vector<uint32_t> geom_nverts;
    std::vector<std::vector<float>> geom_xform;
    std::vector<sgpu::GeomAttr> attr_host;
    std::vector<std::pair<void*, size_t>> attr_allocs;  // device allocation, bytes
    sgpu::GeomAttr* d_geom_attr = nullptr;
    bool has_tangents = false;  // some geometry (placed geometry, for an instanced scene) has tangents
    // per-face textures (sgpu_scene_set_ptex): one record per geometry, texels == nullptr for none
    std::vector<sgpu::PtexTex> tex_host;
    std::vector<void*> tex_allocs;
    sgpu::PtexTex* d_geom_tex = nullptr;
    uint32_t n_geom_tex = 0;
    // image-based light (sgpu_scene_set_ibl): texels, marginal CDF and conditional CDFs in one
    // allocation (d_ibl: w h 3 + h + w h floats); none when d_ibl == nullptr
    float* d_ibl = nullptr;
    uint32_t ibl_w = 0, ibl_h = 0;
    float ibl_rotation = 0.0f, ibl_exposure = 1.0f;
    // sphere around the triangles' bounding box (centre, radius), set at commit
    float bsphere[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // bounding box of the committed content (lo xyz, hi xyz; empty: lo > hi)
    float box[6] = {1.0f, 1.0f, 1.0f, -1.0f, -1.0f, -1.0f};
    // instances (sgpu_scene_add_instance): children and local-to-world transforms (row-major 3x4)
    // until commit; then the composed instances on the device, in top-tree leaf order
    struct Instance {
        sgpu_scene child;
        double l2w[12];
    };
    std::vector<Instance> instances;
    uint32_t refs = 0;                 // instances of other scenes that place this one
    InstRecord* d_inst = nullptr;      // composed instances (traversal records)
    uint32_t* d_inst_ids = nullptr;    // inst_id0, inst_id1 per composed instance
    uint32_t n_inst = 0;
    uint32_t inst_levels = 0;          // 1 or 2 for an instanced scene
    uint32_t blas_stack_cap = 0;       // the largest stack of the placed scenes' trees
    // shading data of an instanced scene, gathered at commit from the placed scenes: their
    // geometries numbered as one list (geometry g of placed scene m is d_inst_geom_base[k] + g for
    // every composed instance k of m), with the materials and textures those scenes held then
    float* d_g_materials = nullptr;       // shading records of all placed scenes
    uint32_t* d_g_geom_material = nullptr;
    sgpu::PtexTex* d_g_geom_tex = nullptr;
    sgpu::GeomAttr* d_g_geom_attr = nullptr;  // NULL when no placed geometry has attributes
    uint32_t n_g_geoms = 0;
    uint32_t* d_inst_geom_base = nullptr;  // per composed instance
    float* d_inst_l2w = nullptr;           // per composed instance: local-to-world linear part, 9 floats
    std::string render_error;              // why the scene cannot be rendered ("": it can)
    // host: the placed scene and local-to-world linear part (row-major 3x3) of each composed
    // instance, in composition order (kept for the shading data)
    std::vector<sgpu_scene> comp_blas;
    std::vector<float> comp_l2w;
};

// ---- per-phase device times ----------------------------------------------------------------------
// Events recorded on `stream` around each launch of a phase (begin/end), summed per phase by
// collect(); a disabled timer records nothing.
struct sgpu_phase_timer {
    static constexpr int kMaxPhases = 8;
    bool on;
    hipStream_t stream;
    std::vector<hipEvent_t> ev;
    std::vector<int> phase;
    size_t used = 0;
    double ms[kMaxPhases] = {};
    sgpu_phase_timer(bool enabled, hipStream_t s) : on(enabled), stream(s) {}
    sgpu_phase_timer(const sgpu_phase_timer&) = delete;
    sgpu_phase_timer& operator=(const sgpu_phase_timer&) = delete;
    hipError_t begin(int ph)
    {
        if (!on) return hipSuccess;
        while (ev.size() < 2 * (used + 1)) {
            hipEvent_t e;
            hipError_t err = hipEventCreate(&e);
            if (err != hipSuccess) return err;
            ev.push_back(e);
        }
        if (phase.size() < used + 1) phase.resize(used + 1);
        phase[used] = ph;
        return hipEventRecord(ev[2 * used], stream);
    }
    hipError_t end()
    {
        if (!on) return hipSuccess;
        hipError_t err = hipEventRecord(ev[2 * used + 1], stream);
        ++used;
        return err;
    }
    hipError_t collect()
    {
        if (!on) return hipSuccess;
        hipError_t err = hipStreamSynchronize(stream);
        for (size_t i = 0; i < used && err == hipSuccess; ++i) {
            float t = 0.0f;
            err = hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]);
            ms[phase[i]] += t;
        }
        used = 0;
        return err;
    }
    ~sgpu_phase_timer()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

// ---- error plumbing ------------------------------------------------------------------------------
void sgpu_set_error(const std::string& msg);
int sgpu_fail(int code, const std::string& msg);
#define SGPU_HIP_TRY(expr)                                                                       \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return sgpu_fail(SGPU_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));   \
    } while (0)

// Appends to a device queue with one atomic per block of BLOCK threads: returns the slot of a
// thread with pred (0xFFFFFFFF for the others), slots of a block consecutive in thread order. Every
// thread of the block calls it, the block's appends in the same order. Same-address atomics
// serialize in the L2: one per wave costs milliseconds on a launch of millions of threads, one per
// block a quarter of that.
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t sgpu_block_append(bool pred, uint32_t* counter)
{
    __shared__ uint32_t s_wave[BLOCK / 64];
    __shared__ uint32_t s_base;
    const uint64_t mask = __ballot(pred);
    const uint32_t lane = __lane_id(), wid = threadIdx.x >> 6;
    if (lane == 0) s_wave[wid] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t k = 0; k < BLOCK / 64; ++k) {
            const uint32_t c = s_wave[k];
            s_wave[k] = run;
            run += c;
        }
        s_base = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    const uint32_t r = s_base + s_wave[wid] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    __syncthreads();  // the next append reuses s_wave and s_base
    return pred ? r : 0xFFFFFFFFu;
}

// The first of n consecutive slots for each thread (n may differ per thread), one atomic per
// block; every thread of the block calls it.
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t sgpu_block_reserve(uint32_t n, uint32_t* counter)
{
    __shared__ uint32_t s_wave[BLOCK / 64];
    __shared__ uint32_t s_base;
    const uint32_t lane = __lane_id(), wid = threadIdx.x >> 6;
    uint32_t incl = n;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 64);
        if (lane >= d) incl += y;
    }
    if (lane == 63) s_wave[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t k = 0; k < BLOCK / 64; ++k) {
            const uint32_t c = s_wave[k];
            s_wave[k] = run;
            run += c;
        }
        s_base = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    const uint32_t r = s_base + s_wave[wid] + incl - n;
    __syncthreads();  // the next reservation reuses s_wave and s_base
    return r;
}

// ---- internal entry points between translation units --------------------------------------------
// Builds d_nodes/d_tris (and the 8-wide d_wide) for the scene's triangles (sgpu_bvh.hip).
int sgpu_bvh_build(sgpu_scene scene, hipStream_t stream);
// Composes an instanced scene's instances and builds its top tree (d_nodes, d_wide over the
// composed instances, d_inst, d_inst_ids) (sgpu_bvh.hip).
int sgpu_bvh_build_instances(sgpu_scene scene, hipStream_t stream);
// Launches the traversal kernels (sgpu_trace.hip), on the context's tree kind (trace_tree).
// `spill` (may be NULL: then the call allocates its own, stream-ordered) holds the launch's ray
// counters (zeroed by the launch) and the spilled part of the per-lane stacks: at least
// sgpu_trace_spill_words(scene) 32-bit words, not shared with a
// concurrent launch. With n_ptr, the kernels read the ray count from that device word (n then only
// sizes the grid and must bound it); `overflow` replaces the scene's stack-overflow flag.
// For an instanced scene, comp_out (may be NULL) receives each closest hit's composed instance.
// max_blocks (0: none) caps the grid below the tuned one; a lent spill buffer then needs only
// (stack_cap - SGPU_STACK_LDS) x max_blocks x SGPU_TRACE_BLOCK words.
int sgpu_trace_launch(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits,
                      uint32_t* occluded, uint32_t n, bool any_hit, uint64_t* d_stats, hipStream_t stream,
                      uint32_t* spill = nullptr, const uint32_t* n_ptr = nullptr, uint32_t* overflow = nullptr,
                      uint32_t* comp_out = nullptr, uint32_t max_blocks = 0, uint64_t* curve_stats = nullptr);
// Argument checks shared by the query entry points (sgpu_capi.hip).
int sgpu_check_query(sgpu_scene scene, const sgpu_rays_soa* rays, uint32_t n, const char* who);
// Looks up this device's no-allocation scratch (sgpu_noalloc.hip) with the context's device current.
int sgpu_noalloc_init(sgpu_context ctx);
size_t sgpu_trace_spill_words(sgpu_scene scene);
size_t sgpu_trace_spill_words(sgpu_context ctx, uint32_t stack_cap);
// A spill buffer starts with kTraceWorkWords words of ray counters; a traversal launch clears them
// on its stream (a small kernel, so that it may be captured) before it starts.
constexpr uint32_t kTraceWorkWords = 64;
void sgpu_trace_work_zero(uint32_t* spill, hipStream_t stream);
// A scene's trees as a multi-scene traversal launch reads them (one table entry per scene).
struct TraceScene {
    const BvhNode* nodes;
    const WideNode* wide;
    const TriRecord* tris;
    uint32_t n_tris;
    uint32_t n_inst;           // > 0: an instanced scene (wide = its top tree over `insts`)
    const InstRecord* insts;
    const BvhNode* cnodes;     // curve tree (n_curves == 0: none)
    const CurveRecord* curves;
    uint32_t n_curves;
};
// One launch over rays of several scenes: ray i traverses table[ray_scene[i]] (device arrays);
// stack_cap is the largest stack_cap of the table's scenes.
// instanced: some entries are instanced scenes (then comp_out, if not NULL, receives each closest
// hit's composed instance, SGPU_INVALID_ID for hits in scenes of geometries).
// curves: some entries have curves (the curve pass runs after the triangle pass).
int sgpu_trace_launch_multi(sgpu_context ctx, const TraceScene* table, const uint32_t* ray_scene, uint32_t stack_cap,
                            uint32_t* overflow, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits,
                            uint32_t* occluded, uint32_t n, bool any_hit, uint64_t* d_stats, hipStream_t stream,
                            uint32_t* spill = nullptr, const uint32_t* n_ptr = nullptr, bool instanced = false,
                            uint32_t* comp_out = nullptr, bool curves = false);
// The curve pass (sgpu_curve.hip): after a triangle pass over the same rays, the curve tree's hits
// closer than the ray's current hit (hits->t, which the triangle pass wrote: its hit or tfar)
// replace it; for any-hit queries, rays not yet occluded are tested. One scene (nodes, curves,
// n_curves; insts / inst_ids of an instanced scene, or NULL) or a table (then ray_scene).
int sgpu_curve_launch(const BvhNode* nodes, const CurveRecord* curves, uint32_t n_curves, const InstRecord* insts,
                      const uint32_t* inst_ids, const TraceScene* table, const uint32_t* ray_scene,
                      const sgpu_rays_soa* rays, const sgpu_hits_soa* hits, uint32_t* occluded, uint32_t n,
                      bool any_hit, const uint32_t* n_ptr, uint32_t* comp_out, uint32_t* overflow,
                      hipStream_t stream, uint64_t* stats = nullptr);
// Builds the curve tree of a scene of geometries from its segments, or of an instanced scene from
// the curves of its composed instances (sgpu_bvh.hip; called by the builds above).
int sgpu_curve_build(sgpu_scene scene, const std::vector<float>& cp, const std::vector<uint32_t>& geom,
                     const std::vector<uint32_t>& prim, const std::vector<uint32_t>& inst, hipStream_t stream);
// Wavefront path tracer (sgpu_pt.hip); arguments validated by sgpu_render_pt.
int sgpu_pt_render_impl(sgpu_scene scene, const sgpu_pt_job* job, float* framebuffer, hipStream_t stream,
                        sgpu_pt_stats* stats = nullptr);
int sgpu_dpt_render_impl(sgpu_scene scene, const sgpu_dpt_job* job, float* framebuffer, hipStream_t stream);
int sgpu_vcm_render_impl(sgpu_scene scene, const sgpu_vcm_job* job, float* framebuffer, sgpu_vcm_stats* stats,
                         hipStream_t stream);
int sgpu_vcm_render_batch_impl(const sgpu_scene* scenes, const sgpu_vcm_job* jobs, float* const* framebuffers,
                               uint32_t n_jobs, uint32_t flags, sgpu_vcm_stats* stats, hipStream_t stream);
int sgpu_vcm_server_create_impl(sgpu_context ctx, const sgpu_vcm_server_config* cfg, hipStream_t stream,
                                sgpu_vcm_server* out);
int sgpu_vcm_server_destroy_impl(sgpu_vcm_server sv, hipStream_t stream);
int sgpu_vcm_server_submit_impl(sgpu_vcm_server sv, sgpu_scene scene, const sgpu_vcm_job* job, float* fb,
                                uint64_t* id_out, hipStream_t stream);
int sgpu_vcm_server_step_impl(sgpu_vcm_server sv, hipStream_t stream, uint64_t* done_ids, uint32_t max_done,
                              uint32_t* n_done, uint32_t* n_active);
int sgpu_vcm_server_stats_impl(sgpu_vcm_server sv, sgpu_vcm_stats* st);
uint64_t sgpu_vcm_server_bytes(sgpu_context ctx, const sgpu_vcm_server_config* cfg);
// Frees the server a context keeps between sgpu_render_vcm(_batch) calls (at context destruction).
void sgpu_vcm_release_cached(sgpu_context ctx);
