// Closest-hit and any-hit BVH traversal: the replacement of Embree's rtcIntersect1/8 and
// rtcOccluded1/8 (called from RayPick / OcclusionRay, Source/Applications/Selas/Source/PathTracer.cpp:71-139,
// and the packet tracers, Source/Applications/Selas/Source/DeferredPathTracer.cpp:151-280).
//
// One ray per lane, 64-lane waves, rays read from SoA arrays with coalesced loads. The traversal
// stack is per lane: its top SGPU_STACK_LDS entries live in LDS laid out [slot][lane] (one bank
// per lane), deeper entries spill to a global buffer laid out [depth][thread] (coalesced).
//
// Triangle test (the arithmetic restated in oracle/sgpu_oracle.c, orc_tri_test):
//   pvec = cross_fma(d, e2); det = dot_fma(e1, pvec); tvec = o - v0;
//   U = dot_fma(tvec, pvec); qvec = cross_fma(tvec, e1); V = dot_fma(d, qvec); T = dot_fma(e2, qvec)
//   det == 0 -> miss; if det < 0 negate U, V, T; a = |det|
//   barycentric accept: U >= 0 && V >= 0 && U + V <= a
//   t = T / a, u = U / a, v = V / a (IEEE division); accept tnear < t <= tfar.
// Closest hit: smallest t, ties to the smaller global triangle index (= smaller (geomID, primID)).
#include "sgpu_internal.h"

#include <algorithm>

namespace {

__device__ __forceinline__ bool tri_test(const TriRecord& r, const f3& o, const f3& d, float tnear, float& t,
                                         float& u, float& v)
{
    f3 v0 = mk3(r.a.x, r.a.y, r.a.z);
    f3 e1 = mk3(r.b.x, r.b.y, r.b.z);
    f3 e2 = mk3(r.c.x, r.c.y, r.c.z);
    f3 pvec = cross_fma(d, e2);
    float det = dot_fma(e1, pvec);
    f3 tvec = o - v0;
    float U = dot_fma(tvec, pvec);
    f3 qvec = cross_fma(tvec, e1);
    float V = dot_fma(d, qvec);
    float T = dot_fma(e2, qvec);
    if (det == 0.0f) return false;
    if (det < 0.0f) {
        U = -U;
        V = -V;
        T = -T;
    }
    float a = fabsf(det);
    if (!(U >= 0.0f && V >= 0.0f && U + V <= a)) return false;
    t = T / a;
    if (!(t > tnear)) return false;
    u = U / a;
    v = V / a;
    return true;
}

__device__ __forceinline__ float safe_rcp(float x)
{
    float ax = fabsf(x);
    float y = ax < 1e-20f ? copysignf(1e-20f, x) : x;
    return __builtin_amdgcn_rcpf(y);
}

template <bool ANY, bool STATS, bool MULTI>
__global__ __launch_bounds__(SGPU_TRACE_BLOCK) void k_trace(const BvhNode* __restrict nodes,
                                                            const TriRecord* __restrict tris, uint32_t n_tris,
                                                            const TraceScene* __restrict table,
                                                            const uint32_t* __restrict ray_scene,
                                                            sgpu_rays_soa rays, sgpu_hits_soa hits,
                                                            uint32_t* __restrict occluded, uint32_t n,
                                                            uint32_t* __restrict spill, uint32_t stack_cap,
                                                            unsigned long long* __restrict stats,
                                                            uint32_t* __restrict overflow)
{
    __shared__ uint32_t lds_stack[SGPU_STACK_LDS][SGPU_TRACE_BLOCK];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * SGPU_TRACE_BLOCK + threadIdx.x;
    const uint32_t stride = gridDim.x * SGPU_TRACE_BLOCK;
    unsigned long long cnt_nodes = 0, cnt_tris = 0;

    for (uint32_t i = gtid; i < n; i += stride) {
        f3 o = mk3(rays.org_x[i], rays.org_y[i], rays.org_z[i]);
        f3 d = mk3(rays.dir_x[i], rays.dir_y[i], rays.dir_z[i]);
        float tnear = rays.tnear[i];
        float tfar = rays.tfar[i];
        f3 inv = mk3(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
        f3 oinv = mk3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
        // the ray's scene: the launch's one, or its own entry of the table in a multi-scene launch
        const BvhNode* __restrict rn = nodes;
        const TriRecord* __restrict rt = tris;
        uint32_t rnt = n_tris;
        if (MULTI) {
            const TraceScene ts = table[ray_scene[i]];
            rn = ts.nodes;
            rt = ts.tris;
            rnt = ts.n_tris;
        }

        float best_t = tfar;
        uint32_t best_key = 0xFFFFFFFFu;
        uint32_t best_tri = 0xFFFFFFFFu;
        float best_u = 0.0f, best_v = 0.0f;
        bool hit_any = false;

        uint32_t sp = 0;
        int node = rnt > 0 ? 0 : 0x7FFFFFFF;
        bool done = rnt == 0;
        while (!done) {
            if (node >= 0) {
                const BvhNode nd = rn[node];
                if (STATS) ++cnt_nodes;
                float l0x = fmaf(nd.x.x, inv.x, -oinv.x), h0x = fmaf(nd.x.y, inv.x, -oinv.x);
                float l1x = fmaf(nd.x.z, inv.x, -oinv.x), h1x = fmaf(nd.x.w, inv.x, -oinv.x);
                float l0y = fmaf(nd.y.x, inv.y, -oinv.y), h0y = fmaf(nd.y.y, inv.y, -oinv.y);
                float l1y = fmaf(nd.y.z, inv.y, -oinv.y), h1y = fmaf(nd.y.w, inv.y, -oinv.y);
                float l0z = fmaf(nd.z.x, inv.z, -oinv.z), h0z = fmaf(nd.z.y, inv.z, -oinv.z);
                float l1z = fmaf(nd.z.z, inv.z, -oinv.z), h1z = fmaf(nd.z.w, inv.z, -oinv.z);
                float tmin0 = fmaxf(fmaxf(fminf(l0x, h0x), fminf(l0y, h0y)), fmaxf(fminf(l0z, h0z), tnear));
                float tmax0 = fminf(fminf(fmaxf(l0x, h0x), fmaxf(l0y, h0y)), fminf(fmaxf(l0z, h0z), best_t));
                float tmin1 = fmaxf(fmaxf(fminf(l1x, h1x), fminf(l1y, h1y)), fmaxf(fminf(l1z, h1z), tnear));
                float tmax1 = fminf(fminf(fmaxf(l1x, h1x), fmaxf(l1y, h1y)), fminf(fmaxf(l1z, h1z), best_t));
                bool hit0 = tmin0 <= tmax0;
                bool hit1 = tmin1 <= tmax1;
                if (hit0 && hit1) {
                    int near_c = tmin0 <= tmin1 ? nd.child.x : nd.child.y;
                    int far_c = tmin0 <= tmin1 ? nd.child.y : nd.child.x;
                    if (sp >= stack_cap) {
                        atomicOr(overflow, 1u);
                        done = true;
                        break;
                    }
                    if (sp < SGPU_STACK_LDS) {
                        lds_stack[sp][lane] = (uint32_t)far_c;
                    } else {
                        // ring over the LDS slots: evict position sp-K to global before reuse
                        uint32_t slot = sp % SGPU_STACK_LDS;
                        spill[(size_t)(sp - SGPU_STACK_LDS) * stride + gtid] = lds_stack[slot][lane];
                        lds_stack[slot][lane] = (uint32_t)far_c;
                    }
                    ++sp;
                    node = near_c;
                    continue;
                } else if (hit0) {
                    node = nd.child.x;
                    continue;
                } else if (hit1) {
                    node = nd.child.y;
                    continue;
                }
            } else {
                const uint32_t j = (uint32_t)(~node);
                const TriRecord r = rt[j];
                if (STATS) ++cnt_tris;
                float t, u, v;
                if (tri_test(r, o, d, tnear, t, u, v)) {
                    if (ANY) {
                        if (t <= tfar) {
                            hit_any = true;
                            break;
                        }
                    } else {
                        uint32_t key = __float_as_uint(r.a.w);
                        if (t < best_t || (t == best_t && key < best_key)) {
                            best_t = t;
                            best_key = key;
                            best_tri = j;
                            best_u = u;
                            best_v = v;
                        }
                    }
                }
            }
            // pop
            if (sp == 0) break;
            --sp;
            uint32_t slot = sp % SGPU_STACK_LDS;
            node = (int)lds_stack[slot][lane];
            if (sp >= SGPU_STACK_LDS) lds_stack[slot][lane] = spill[(size_t)(sp - SGPU_STACK_LDS) * stride + gtid];
        }

        if (ANY) {
            occluded[i] = hit_any ? 1u : 0u;
        } else {
            bool found = best_tri != 0xFFFFFFFFu;
            uint32_t prim = SGPU_INVALID_ID, geom = SGPU_INVALID_ID;
            f3 ng = mk3(0.0f, 0.0f, 0.0f);
            if (found) {
                const TriRecord r = rt[best_tri];
                geom = __float_as_uint(r.b.w);
                prim = __float_as_uint(r.c.w);
                if (hits.ng_x) ng = cross3(mk3(r.b.x, r.b.y, r.b.z), mk3(r.c.x, r.c.y, r.c.z));
                if (prim & 0x80000000u) {
                    // triangle 2q + h of quad q: report the quad and its own (u, v)
                    if (prim & 1u) {
                        best_u = 1.0f - best_u;
                        best_v = 1.0f - best_v;
                    }
                    prim = (prim & 0x7FFFFFFFu) >> 1;
                }
            }
            if (hits.t) hits.t[i] = best_t;
            if (hits.u) hits.u[i] = found ? best_u : 0.0f;
            if (hits.v) hits.v[i] = found ? best_v : 0.0f;
            if (hits.prim_id) hits.prim_id[i] = prim;
            if (hits.geom_id) hits.geom_id[i] = geom;
            if (hits.ng_x) {
                hits.ng_x[i] = ng.x;
                hits.ng_y[i] = ng.y;
                hits.ng_z[i] = ng.z;
            }
        }
    }
    if (STATS) {
        atomicAdd(&stats[0], cnt_nodes);
        atomicAdd(&stats[1], cnt_tris);
    }
}

}  // namespace


namespace {

template <bool ANY, bool STATS>
void launch_trace(bool multi, uint32_t blocks, hipStream_t stream, const BvhNode* nodes, const TriRecord* tris,
                  uint32_t n_tris, const TraceScene* table, const uint32_t* ray_scene, const sgpu_rays_soa& rays,
                  const sgpu_hits_soa& h, uint32_t* occluded, uint32_t n, uint32_t* spill, uint32_t cap,
                  unsigned long long* st, uint32_t* overflow)
{
    if (multi)
        hipLaunchKernelGGL((k_trace<ANY, STATS, true>), dim3(blocks), dim3(SGPU_TRACE_BLOCK), 0, stream, nodes, tris,
                           n_tris, table, ray_scene, rays, h, occluded, n, spill, cap, st, overflow);
    else
        hipLaunchKernelGGL((k_trace<ANY, STATS, false>), dim3(blocks), dim3(SGPU_TRACE_BLOCK), 0, stream, nodes,
                           tris, n_tris, table, ray_scene, rays, h, occluded, n, spill, cap, st, overflow);
}

// One traversal launch over n rays: of `scene`, or (table != nullptr) of the scenes of a batch,
// ray i in table[ray_scene[i]]; max_depth bounds every scene's tree depth.
int trace_launch(sgpu_context ctx, const BvhNode* nodes, const TriRecord* tris, uint32_t n_tris, uint32_t max_depth,
                 uint32_t* overflow, const TraceScene* table, const uint32_t* ray_scene, const sgpu_rays_soa* rays,
                 const sgpu_hits_soa* hits, uint32_t* occluded, uint32_t n, bool any_hit, uint64_t* d_stats,
                 hipStream_t stream, uint32_t* spill_buf)
{
    if (n == 0) return SGPU_OK;
    const uint32_t cus = (uint32_t)ctx->props.multiProcessorCount;
    uint32_t blocks = (n + SGPU_TRACE_BLOCK - 1) / SGPU_TRACE_BLOCK;
    const uint32_t max_blocks = cus * std::min(ctx->trace_blocks_per_cu, kMaxTraceBlocksPerCu);
    if (blocks > max_blocks) blocks = max_blocks;
    const uint32_t cap = max_depth + 1;

    // stream-ordered scratch for the spilled part of the stacks, private to this call, unless the
    // caller lends a buffer of sgpu_trace_spill_words (sized for the largest grid)
    uint32_t* spill = spill_buf;
    if (!spill) SGPU_HIP_TRY(hipMallocAsync((void**)&spill, sgpu_trace_spill_words(ctx, max_depth) * sizeof(uint32_t),
                                            stream));
    sgpu_hits_soa h = {};
    if (hits) h = *hits;
    unsigned long long* st = (unsigned long long*)d_stats;
    const bool multi = table != nullptr;
    if (any_hit) {
        if (st)
            launch_trace<true, true>(multi, blocks, stream, nodes, tris, n_tris, table, ray_scene, *rays, h, occluded,
                                     n, spill, cap, st, overflow);
        else
            launch_trace<true, false>(multi, blocks, stream, nodes, tris, n_tris, table, ray_scene, *rays, h, occluded,
                                      n, spill, cap, st, overflow);
    } else {
        if (st)
            launch_trace<false, true>(multi, blocks, stream, nodes, tris, n_tris, table, ray_scene, *rays, h, occluded,
                                      n, spill, cap, st, overflow);
        else
            launch_trace<false, false>(multi, blocks, stream, nodes, tris, n_tris, table, ray_scene, *rays, h,
                                       occluded, n, spill, cap, st, overflow);
    }
    SGPU_HIP_TRY(hipGetLastError());
    if (!spill_buf) SGPU_HIP_TRY(hipFreeAsync(spill, stream));
    return SGPU_OK;
}

}  // namespace

size_t sgpu_trace_spill_words(sgpu_context ctx, uint32_t max_depth)
{
    const uint32_t cus = (uint32_t)ctx->props.multiProcessorCount;
    const uint32_t threads = cus * kMaxTraceBlocksPerCu * SGPU_TRACE_BLOCK;
    const uint32_t cap = max_depth + 1;
    const uint32_t spill_depth = cap > SGPU_STACK_LDS ? cap - SGPU_STACK_LDS : 1;
    return (size_t)spill_depth * threads;
}

size_t sgpu_trace_spill_words(sgpu_scene scene) { return sgpu_trace_spill_words(scene->ctx, scene->max_depth); }

int sgpu_trace_launch(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits, uint32_t* occluded,
                      uint32_t n, bool any_hit, uint64_t* d_stats, hipStream_t stream, uint32_t* spill_buf)
{
    return trace_launch(scene->ctx, scene->d_nodes, scene->d_tris, scene->n_tris, scene->max_depth, scene->d_overflow,
                        nullptr, nullptr, rays, hits, occluded, n, any_hit, d_stats, stream, spill_buf);
}

int sgpu_trace_launch_multi(sgpu_context ctx, const TraceScene* table, const uint32_t* ray_scene, uint32_t max_depth,
                            uint32_t* overflow, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits,
                            uint32_t* occluded, uint32_t n, bool any_hit, uint64_t* d_stats, hipStream_t stream,
                            uint32_t* spill_buf)
{
    return trace_launch(ctx, nullptr, nullptr, 0, max_depth, overflow, table, ray_scene, rays, hits, occluded, n,
                        any_hit, d_stats, stream, spill_buf);
}
