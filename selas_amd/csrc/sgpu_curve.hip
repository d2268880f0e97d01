// The curve pass: rays against a scene's round B-spline curves (rtcIntersect1 / rtcOccluded1 on
// RTC_GEOMETRY_TYPE_ROUND_BSPLINE_CURVE geometry, ModelResource.cpp:300-330), after the triangle
// traversal over the same rays. Curves sit in their own binary tree over segment boxes (built by
// sgpu_curve_build), so that the triangle kernels keep their registers; this pass runs only for
// scenes that hold curves. One lane per ray, grid-stride; a short per-lane stack (the tree's depth
// is checked against it at build).
#include "sgpu_curve.h"
#include "sgpu_internal.h"

namespace {

constexpr int kCurveBlock = 256;
constexpr uint32_t kCurveMaxBlocks = 4096;

__device__ __forceinline__ bool child_box(const BvhNode& nd, int c, f3 o, f3 inv, float tmin, float tmax, float& tent)
{
    const float lx = c ? nd.x.z : nd.x.x, hx = c ? nd.x.w : nd.x.y;
    const float ly = c ? nd.y.z : nd.y.x, hy = c ? nd.y.w : nd.y.y;
    const float lz = c ? nd.z.z : nd.z.x, hz = c ? nd.z.w : nd.z.y;
    const float ax = (lx - o.x) * inv.x, bx = (hx - o.x) * inv.x;
    const float ay = (ly - o.y) * inv.y, by = (hy - o.y) * inv.y;
    const float az = (lz - o.z) * inv.z, bz = (hz - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(tmin, fminf(ax, bx)), fmaxf(fminf(ay, by), fminf(az, bz)));
    const float tf = fminf(fminf(tmax, fmaxf(ax, bx)), fminf(fmaxf(ay, by), fmaxf(az, bz)));
    tent = tn;
    return tn <= tf;
}

__device__ __forceinline__ float safe_inv(float x)
{
    return 1.0f / (fabsf(x) > 1e-30f ? x : copysignf(1e-30f, x));
}

template <bool ANY, bool MULTI, bool STATS>
__global__ __launch_bounds__(kCurveBlock) void k_curves(const BvhNode* __restrict nodes,
                                                        const CurveRecord* __restrict curves, uint32_t n_curves,
                                                        const InstRecord* __restrict insts,
                                                        const uint32_t* __restrict inst_ids,
                                                        const TraceScene* __restrict table,
                                                        const uint32_t* __restrict ray_scene, sgpu_rays_soa rays,
                                                        sgpu_hits_soa hits, uint32_t* __restrict occluded, uint32_t n,
                                                        const uint32_t* __restrict n_ptr,
                                                        uint32_t* __restrict comp_out, uint32_t* __restrict overflow,
                                                        unsigned long long* __restrict stats)
{
    const uint32_t count = n_ptr ? min(*n_ptr, n) : n;
    unsigned long long cnt_nodes = 0, cnt_segs = 0;
    for (uint32_t i = blockIdx.x * kCurveBlock + threadIdx.x; i < count; i += gridDim.x * kCurveBlock) {
        const BvhNode* nd_base = nodes;
        const CurveRecord* recs = curves;
        uint32_t nc = n_curves;
        const InstRecord* ins = insts;
        if (MULTI) {
            const TraceScene ts = table[ray_scene[i]];
            nd_base = ts.cnodes;
            recs = ts.curves;
            nc = ts.n_curves;
            ins = ts.insts;
        }
        if (nc == 0) continue;
        if (ANY && occluded[i]) continue;
        const f3 o = mk3(rays.org_x[i], rays.org_y[i], rays.org_z[i]);
        const f3 d = mk3(rays.dir_x[i], rays.dir_y[i], rays.dir_z[i]);
        const f3 inv = mk3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
        const float tnear = rays.tnear[i];
        // the triangle pass left its hit distance (or tfar on a miss) in hits.t
        float best = (!ANY && hits.t) ? hits.t[i] : rays.tfar[i];
        int best_rec = -1;
        float best_u = 0.0f;
        f3 best_ng = mk3(0.0f, 0.0f, 0.0f);
        bool found = false, ovf = false;
        int stack[kCurveStack];
        uint32_t sp = 0;
        int node = 0;
        auto test_leaf = [&](int leaf) {
            if (STATS) ++cnt_segs;
            const CurveRecord r = recs[leaf];
            sgpu::CurveSeg s;
            s.p[0] = r.p[0];
            s.p[1] = r.p[1];
            s.p[2] = r.p[2];
            s.p[3] = r.p[3];
            float t, u;
            f3 ng;
            if (sgpu::curve_intersect<ANY>(s, o, d, tnear, best, t, u, ng) && (ANY || t < best)) {
                found = true;
                best = t;
                best_rec = leaf;
                best_u = u;
                best_ng = ng;
            }
        };
        while (true) {
            if (STATS) ++cnt_nodes;
            const BvhNode nd = nd_base[node];
            float t0, t1;
            bool h0 = child_box(nd, 0, o, inv, tnear, best, t0);
            bool h1 = child_box(nd, 1, o, inv, tnear, best, t1);
            if (h0 && nd.child.x < 0) {
                test_leaf(~nd.child.x);
                h0 = false;
            }
            if (ANY && found) break;
            if (h1 && nd.child.y < 0) {
                test_leaf(~nd.child.y);
                h1 = false;
            }
            if (ANY && found) break;
            if (h0 && h1) {
                const int nearc = t0 <= t1 ? nd.child.x : nd.child.y;
                const int farc = t0 <= t1 ? nd.child.y : nd.child.x;
                if (sp < kCurveStack) stack[sp++] = farc;
                else ovf = true;
                node = nearc;
                continue;
            }
            if (h0) {
                node = nd.child.x;
                continue;
            }
            if (h1) {
                node = nd.child.y;
                continue;
            }
            if (sp == 0) break;
            node = stack[--sp];
        }
        if (ovf) atomicOr(overflow, 1u);
        if (ANY) {
            if (found) occluded[i] = 1u;
            continue;
        }
        if (!found) continue;
        const CurveRecord r = recs[best_rec];
        f3 ng = best_ng;
        uint32_t id0 = SGPU_INVALID_ID, id1 = SGPU_INVALID_ID;
        if (r.inst != SGPU_INVALID_ID && ins) {
            // Embree reports an instanced geometry's normal in its own space: through the linear part
            // of the world-to-local transform (a similarity, so the normal's direction is kept)
            const InstRecord ir = ins[r.inst];
            ng = mk3(ir.w2l[0].x * best_ng.x + ir.w2l[0].y * best_ng.y + ir.w2l[0].z * best_ng.z,
                     ir.w2l[1].x * best_ng.x + ir.w2l[1].y * best_ng.y + ir.w2l[1].z * best_ng.z,
                     ir.w2l[2].x * best_ng.x + ir.w2l[2].y * best_ng.y + ir.w2l[2].z * best_ng.z);
            if (inst_ids) {
                id0 = inst_ids[2 * r.inst];
                id1 = inst_ids[2 * r.inst + 1];
            }
        }
        if (hits.t) hits.t[i] = best;
        if (hits.u) hits.u[i] = best_u;
        if (hits.v) hits.v[i] = 0.0f;
        if (hits.prim_id) hits.prim_id[i] = r.prim;
        if (hits.geom_id) hits.geom_id[i] = r.geom;
        if (hits.ng_x) {
            hits.ng_x[i] = ng.x;
            hits.ng_y[i] = ng.y;
            hits.ng_z[i] = ng.z;
        }
        if (hits.inst_id0) hits.inst_id0[i] = id0;
        if (hits.inst_id1) hits.inst_id1[i] = id1;
        if (comp_out) comp_out[i] = r.inst;
    }
    if (STATS) {
        atomicAdd(&stats[0], cnt_nodes);
        atomicAdd(&stats[1], cnt_segs);
    }
}

}  // namespace

int sgpu_curve_launch(const BvhNode* nodes, const CurveRecord* curves, uint32_t n_curves, const InstRecord* insts,
                      const uint32_t* inst_ids, const TraceScene* table, const uint32_t* ray_scene,
                      const sgpu_rays_soa* rays, const sgpu_hits_soa* hits, uint32_t* occluded, uint32_t n,
                      bool any_hit, const uint32_t* n_ptr, uint32_t* comp_out, uint32_t* overflow, hipStream_t stream,
                      uint64_t* stats)
{
    if (n == 0) return SGPU_OK;
    sgpu_hits_soa h = {};
    if (hits) h = *hits;
    const uint32_t blocks = std::min<uint32_t>((n + kCurveBlock - 1) / kCurveBlock, kCurveMaxBlocks);
    const dim3 g(blocks), b(kCurveBlock);
    const bool multi = table != nullptr;
    unsigned long long* st = (unsigned long long*)stats;
#define SGPU_LAUNCH_C(A, M, S)                                                                                    \
    hipLaunchKernelGGL((k_curves<A, M, S>), g, b, 0, stream, nodes, curves, n_curves, insts, inst_ids, table,   \
                       ray_scene, *rays, h, occluded, n, n_ptr, comp_out, overflow, st)
    if (st) {  // counting: single-scene queries (sgpu_intersect_stats)
        if (any_hit) SGPU_LAUNCH_C(true, false, true);
        else SGPU_LAUNCH_C(false, false, true);
    } else if (any_hit) {
        if (multi) SGPU_LAUNCH_C(true, true, false);
        else SGPU_LAUNCH_C(true, false, false);
    } else {
        if (multi) SGPU_LAUNCH_C(false, true, false);
        else SGPU_LAUNCH_C(false, false, false);
    }
#undef SGPU_LAUNCH_C
    SGPU_HIP_TRY(hipGetLastError());
    return SGPU_OK;
}
