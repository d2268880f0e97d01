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


// The query's result: occlusion flag, or Embree's hit fields (quad hits reported as the quad and its
// own (u, v): triangle 2q + h of quad q, second triangle (1 - u, 1 - v)).
template <bool ANY>
__device__ __forceinline__ void write_result(uint32_t i, bool hit_any, uint32_t best_tri, float best_t, float best_u,
                                             float best_v, const TriRecord* __restrict rt, sgpu_hits_soa& hits,
                                             uint32_t* __restrict occluded, uint32_t id0 = SGPU_INVALID_ID,
                                             uint32_t id1 = SGPU_INVALID_ID)
{
    if (ANY) {
        occluded[i] = hit_any ? 1u : 0u;
        return;
    }
    bool found = best_tri != 0xFFFFFFFFu;
    uint32_t prim = SGPU_INVALID_ID, geom = SGPU_INVALID_ID;
    f3 ng = mk3(0.0f, 0.0f, 0.0f);
    if (found) {
        const TriRecord r = rt[best_tri];
        geom = __float_as_uint(r.b.w);
        prim = __float_as_uint(r.c.w);
        if (hits.ng_x) ng = cross3(mk3(r.b.x, r.b.y, r.b.z), mk3(r.c.x, r.c.y, r.c.z));
        if (prim & 0x80000000u) {
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
    if (hits.inst_id0) hits.inst_id0[i] = found ? id0 : SGPU_INVALID_ID;
    if (hits.inst_id1) hits.inst_id1[i] = found ? id1 : SGPU_INVALID_ID;
}

template <bool ANY, bool STATS, bool MULTI>
__global__ __launch_bounds__(SGPU_TRACE_BLOCK) void k_trace(const BvhNode* __restrict nodes,
                                                            const TriRecord* __restrict tris, uint32_t n_tris,
                                                            const TraceScene* __restrict table,
                                                            const uint32_t* __restrict ray_scene,
                                                            sgpu_rays_soa rays, sgpu_hits_soa hits,
                                                            uint32_t* __restrict occluded, uint32_t n,
                                                            const uint32_t* __restrict n_ptr,
                                                            uint32_t* __restrict spill, uint32_t stack_cap,
                                                            unsigned long long* __restrict stats,
                                                            uint32_t* __restrict overflow)
{
    __shared__ uint32_t lds_stack[SGPU_STACK_LDS][SGPU_TRACE_BLOCK];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * SGPU_TRACE_BLOCK + threadIdx.x;
    const uint32_t stride = gridDim.x * SGPU_TRACE_BLOCK;
    unsigned long long cnt_nodes = 0, cnt_tris = 0;
    if (n_ptr) n = *n_ptr;  // the count written on the device by the kernel that queued the rays

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

        write_result<ANY>(i, hit_any, best_tri, best_t, best_u, best_v, rt, hits, occluded);
    }
    if (STATS) {
        atomicAdd(&stats[0], cnt_nodes);
        atomicAdd(&stats[1], cnt_tris);
    }
}

// Compare-exchange of (key, ref) pairs: ascending keys.
__device__ __forceinline__ void cswap(float& ka, int& ra, float& kb, int& rb)
{
    const bool sw = kb < ka;
    const float k = sw ? kb : ka;
    const int r = sw ? rb : ra;
    kb = sw ? ka : kb;
    rb = sw ? ra : rb;
    ka = k;
    ra = r;
}

// 8-wide traversal of the collapsed tree (WideNode). Per node the 8 quantized child boxes are
// decoded and slab-tested in registers. Triangle children whose box is hit are tested at once;
// internal children whose box is hit are ordered by entry distance with a fixed sorting network
// (closest hit; the any-hit query keeps slot order), the nearest is entered and the rest are pushed
// far to near. Hits are the triangle test's, independent of the tree and of the visiting order.
// Waves take rays in chunks of 64 and a lane whose ray is done takes the next one of its wave's
// chunk at once (one node visit per loop trip), so lanes do not wait for the wave's longest ray;
// a wave whose chunk is used up takes the next from a counter, so waves do not wait for the
// launch's longest chunk. Measured on MI355X against the one-ray-per-lane grid-stride loop: the
// traversal workload (16M rays, 2M triangles) 1027 -> 1824 Mrays/s, the VCM job's closest-hit and
// shadow traversals 13.2 + 12.2 -> 13.4 + 11.2 ms, both at 8 blocks per CU; held to 7 waves per
// SIMD (72 VGPRs): at 8 the refill state spills and costs more than the eighth wave gives.
// Wide nodes of a single-scene launch that may be kept in LDS (sgpu_tuning.trace_lds_nodes): the root
// and the next breadth-first nodes (the root's children when it has 8 internal ones), visited by
// nearly every ray. Measured on MI355X they hit in L1/L2 anyway, and the LDS branch costs 1-3 %.
constexpr uint32_t kTopNodes = 9;
constexpr uint32_t kTraceChunk = 64;  // rays a wave takes from the launch's counter at a time

template <bool ANY, bool STATS, bool MULTI>
__global__ __launch_bounds__(SGPU_TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(7))) void k_trace8(const WideNode* __restrict wnodes, uint32_t n_top,
                                                             const TriRecord* __restrict tris, uint32_t n_tris,
                                                             const TraceScene* __restrict table,
                                                             const uint32_t* __restrict ray_scene,
                                                             sgpu_rays_soa rays, sgpu_hits_soa hits,
                                                             uint32_t* __restrict occluded, uint32_t n,
                                                            const uint32_t* __restrict n_ptr,
                                                             uint32_t* __restrict spill, uint32_t stack_cap,
                                                             unsigned long long* __restrict stats,
                                                             uint32_t* __restrict overflow, uint32_t* __restrict work)
{
    __shared__ uint32_t lds_stack[SGPU_STACK_LDS][SGPU_TRACE_BLOCK];
    __shared__ uint4 lds_top[kTopNodes * 6];  // the 96 bytes of each top node the traversal reads
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * SGPU_TRACE_BLOCK + threadIdx.x;
    const uint32_t stride = gridDim.x * SGPU_TRACE_BLOCK;
    unsigned long long cnt_nodes = 0, cnt_tris = 0;
    if (n_ptr) n = *n_ptr;  // the count written on the device by the kernel that queued the rays
    if (MULTI) n_top = 0;
    for (uint32_t k = threadIdx.x; k < 6 * n_top; k += SGPU_TRACE_BLOCK)
        lds_top[k] = reinterpret_cast<const uint4*>(wnodes + k / 6)[k % 6];
    __syncthreads();

    // Rays come in chunks of 64: wave w's first chunk is chunk w, the next ones come from the
    // launch's counter (work[0], zeroed before the launch) past the grid's first chunks. A lane
    // whose ray is done takes the next ray of its wave's chunk at once, and a wave whose chunk is
    // used up takes the next chunk, so that lanes do not idle while a wave's longest ray is
    // traversed and waves do not idle while others hold rays; a launch of at most one chunk per
    // wave takes no atomic. Positions in a chunk come from a ballot, so `next` and `end` stay
    // wave-uniform.
    const uint32_t grid_rays = (gridDim.x * SGPU_TRACE_BLOCK / 64) * kTraceChunk;
    uint32_t next = min(n, (gtid >> 6) * kTraceChunk);
    uint32_t end = min(n, next + kTraceChunk);
    bool drained = false;
    const unsigned long long below = (1ull << (threadIdx.x & 63u)) - 1ull;

    bool active = false;
    uint32_t i = 0;
    f3 o = mk3(0, 0, 0), d = mk3(0, 0, 0), inv = mk3(0, 0, 0), oinv = mk3(0, 0, 0);
    float tnear = 0.0f, tfar = 0.0f;
    const WideNode* __restrict rw = wnodes;
    const TriRecord* __restrict rt = tris;
    uint32_t rnt = n_tris;
    float best_t = 0.0f;
    uint32_t best_key = 0xFFFFFFFFu, best_tri = 0xFFFFFFFFu;
    float best_u = 0.0f, best_v = 0.0f;
    bool hit_any = false;
    // the per-lane stack: top SGPU_STACK_LDS entries in LDS (a ring), deeper ones spilled
    uint32_t sp = 0;
    bool overflowed = false;
    int node = -1;

    auto push = [&](int x) {
        if (sp >= stack_cap) {
            overflowed = true;
            return;
        }
        if (sp < SGPU_STACK_LDS) {
            lds_stack[sp][lane] = (uint32_t)x;
        } else {
            const uint32_t slot = sp % SGPU_STACK_LDS;
            spill[(size_t)(sp - SGPU_STACK_LDS) * stride + gtid] = lds_stack[slot][lane];
            lds_stack[slot][lane] = (uint32_t)x;
        }
        ++sp;
    };
    auto test_tri = [&](int ref) -> bool {  // true: any-hit found
        const uint32_t j = (uint32_t)(~ref);
        const TriRecord r = rt[j];
        if (STATS) ++cnt_tris;
        float t, u, v;
        if (tri_test(r, o, d, tnear, t, u, v)) {
            if (ANY) {
                if (t <= tfar) return true;
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
        return false;
    };

    for (;;) {
        // lanes without a ray take the next rays of the chunk (a new chunk when it is used up)
        const unsigned long long idle = __ballot(!active);
        if (idle && next >= end && !drained) {
            if (grid_rays >= n) {
                drained = true;  // the first chunks held every ray
            } else {
                uint32_t base = 0;
                if ((threadIdx.x & 63u) == 0) base = grid_rays + atomicAdd(&work[0], kTraceChunk);
                base = (uint32_t)__shfl((int)base, 0, 64);
                if (base >= n) {
                    drained = true;
                } else {
                    next = base;
                    end = min(n, base + kTraceChunk);
                }
            }
        }
        if (idle && next < end) {
            const uint32_t r = next + (uint32_t)__popcll(idle & below);
            if (!active && r < end) {
                i = r;
                active = true;
                o = mk3(rays.org_x[i], rays.org_y[i], rays.org_z[i]);
                d = mk3(rays.dir_x[i], rays.dir_y[i], rays.dir_z[i]);
                tnear = rays.tnear[i];
                tfar = rays.tfar[i];
                inv = mk3(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
                oinv = mk3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
                if (MULTI) {
                    const TraceScene ts = table[ray_scene[i]];
                    rw = ts.wide;
                    rt = ts.tris;
                    rnt = ts.n_tris;
                }
                best_t = tfar;
                best_key = 0xFFFFFFFFu;
                best_tri = 0xFFFFFFFFu;
                best_u = 0.0f;
                best_v = 0.0f;
                hit_any = false;
                sp = 0;
                overflowed = false;
                node = rnt > 0 ? 0 : -1;
            }
            next = min(end, next + (uint32_t)__popcll(idle));
        }
        if (__ballot(active) == 0) break;
        if (!active) continue;
        if (node >= 0) {
            // ---- one node visit ----
            uint4 w0, w1, w2, w3, w4, w5;
            if ((uint32_t)node < n_top) {
                const uint4* t = lds_top + 6 * node;
                w0 = t[0], w1 = t[1], w2 = t[2], w3 = t[3], w4 = t[4], w5 = t[5];
            } else {
                const uint4* __restrict np = reinterpret_cast<const uint4*>(rw + node);
                w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4], w5 = np[5];
            }
            if (STATS) ++cnt_nodes;
            const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
            const uint32_t meta = w0.w;
            const float sx = __uint_as_float((meta & 0xFFu) << 23);
            const float sy = __uint_as_float(((meta >> 8) & 0xFFu) << 23);
            const float sz = __uint_as_float(((meta >> 16) & 0xFFu) << 23);
            const uint32_t nch = meta >> 24;
            const uint32_t lox[2] = {w1.x, w1.y}, loy[2] = {w1.z, w1.w}, loz[2] = {w2.x, w2.y};
            const uint32_t hix[2] = {w2.z, w2.w}, hiy[2] = {w3.x, w3.y}, hiz[2] = {w3.z, w3.w};
            int ref[8] = {(int)w4.x, (int)w4.y, (int)w4.z, (int)w4.w, (int)w5.x, (int)w5.y, (int)w5.z, (int)w5.w};
            float key[8];
            uint32_t leaves = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int sh = 8 * (c & 3), wi = c >> 2;
                const float bx0 = fmaf((float)((lox[wi] >> sh) & 0xFFu), sx, ox);
                const float bx1 = fmaf((float)((hix[wi] >> sh) & 0xFFu), sx, ox);
                const float by0 = fmaf((float)((loy[wi] >> sh) & 0xFFu), sy, oy);
                const float by1 = fmaf((float)((hiy[wi] >> sh) & 0xFFu), sy, oy);
                const float bz0 = fmaf((float)((loz[wi] >> sh) & 0xFFu), sz, oz);
                const float bz1 = fmaf((float)((hiz[wi] >> sh) & 0xFFu), sz, oz);
                const float l_x = fmaf(bx0, inv.x, -oinv.x), h_x = fmaf(bx1, inv.x, -oinv.x);
                const float l_y = fmaf(by0, inv.y, -oinv.y), h_y = fmaf(by1, inv.y, -oinv.y);
                const float l_z = fmaf(bz0, inv.z, -oinv.z), h_z = fmaf(bz1, inv.z, -oinv.z);
                const float tmin = fmaxf(fmaxf(fminf(l_x, h_x), fminf(l_y, h_y)), fmaxf(fminf(l_z, h_z), tnear));
                const float tmax = fminf(fminf(fmaxf(l_x, h_x), fmaxf(l_y, h_y)), fminf(fmaxf(l_z, h_z), best_t));
                const bool hit = (uint32_t)c < nch && tmin <= tmax;
                // internal children: entry distance as the sort key (+inf: not to visit)
                key[c] = hit && ref[c] >= 0 ? tmin : __int_as_float(0x7f800000);
                if (hit && ref[c] < 0) leaves |= 1u << c;
            }
            // triangle children, in slot order
            bool found = false;
            while (leaves) {
                const int c = __builtin_ctz(leaves);
                leaves &= leaves - 1;
                int r = ref[0];
#pragma unroll
                for (int k = 1; k < 8; ++k) r = c == k ? ref[k] : r;
                if (test_tri(r)) {
                    found = true;
                    break;
                }
            }
            if (ANY && found) {
                hit_any = true;
                node = -1;
            } else {
                const float inf = __int_as_float(0x7f800000);
                if (!ANY) {
                    // entry distances beyond the closest hit so far cannot hold a closer one
#pragma unroll
                    for (int c = 0; c < 8; ++c) key[c] = key[c] <= best_t ? key[c] : inf;
                    // ascending (key, ref): Batcher's odd-even merge sort network for 8 inputs
                    cswap(key[0], ref[0], key[1], ref[1]); cswap(key[2], ref[2], key[3], ref[3]);
                    cswap(key[4], ref[4], key[5], ref[5]); cswap(key[6], ref[6], key[7], ref[7]);
                    cswap(key[0], ref[0], key[2], ref[2]); cswap(key[1], ref[1], key[3], ref[3]);
                    cswap(key[4], ref[4], key[6], ref[6]); cswap(key[5], ref[5], key[7], ref[7]);
                    cswap(key[1], ref[1], key[2], ref[2]); cswap(key[5], ref[5], key[6], ref[6]);
                    cswap(key[0], ref[0], key[4], ref[4]); cswap(key[1], ref[1], key[5], ref[5]);
                    cswap(key[2], ref[2], key[6], ref[6]); cswap(key[3], ref[3], key[7], ref[7]);
                    cswap(key[2], ref[2], key[4], ref[4]); cswap(key[3], ref[3], key[5], ref[5]);
                    cswap(key[1], ref[1], key[2], ref[2]); cswap(key[3], ref[3], key[4], ref[4]);
                    cswap(key[5], ref[5], key[6], ref[6]);
                }
                // the nearest (first valid) is entered, the others pushed far to near
                int nxt = -1;
#pragma unroll
                for (int c = 7; c >= 0; --c) {
                    if (key[c] != inf) {
                        if (nxt >= 0) push(nxt);
                        nxt = ref[c];
                    }
                }
                if (nxt >= 0) {
                    node = nxt;
                } else if (sp == 0) {
                    node = -1;  // the ray is done
                } else {
                    // pop
                    --sp;
                    const uint32_t slot = sp % SGPU_STACK_LDS;
                    node = (int)lds_stack[slot][lane];
                    if (sp >= SGPU_STACK_LDS)
                        lds_stack[slot][lane] = spill[(size_t)(sp - SGPU_STACK_LDS) * stride + gtid];
                }
            }
            if (overflowed) node = -1;
        }
        if (node < 0) {
            if (overflowed) atomicOr(overflow, 1u);
            write_result<ANY>(i, hit_any, best_tri, best_t, best_u, best_v, rt, hits, occluded);
            active = false;
        }
    }
    if (STATS) {
        atomicAdd(&stats[0], cnt_nodes);
        atomicAdd(&stats[1], cnt_tris);
    }
}

// A ray into an instance's space: x' = w.x x + w.y y + w.z z (+ w.w for points), summed by fused
// multiply-adds from the last term (the arithmetic the oracle restates).
__device__ __forceinline__ f3 inst_point(const float4* w, f3 p)
{
    return mk3(fmaf(w[0].x, p.x, fmaf(w[0].y, p.y, fmaf(w[0].z, p.z, w[0].w))),
               fmaf(w[1].x, p.x, fmaf(w[1].y, p.y, fmaf(w[1].z, p.z, w[1].w))),
               fmaf(w[2].x, p.x, fmaf(w[2].y, p.y, fmaf(w[2].z, p.z, w[2].w))));
}
__device__ __forceinline__ f3 inst_vector(const float4* w, f3 v)
{
    return mk3(fmaf(w[0].x, v.x, fmaf(w[0].y, v.y, w[0].z * v.z)), fmaf(w[1].x, v.x, fmaf(w[1].y, v.y, w[1].z * v.z)),
               fmaf(w[2].x, v.x, fmaf(w[2].y, v.y, w[2].z * v.z)));
}

// Traversal of an instanced scene: the top 8-wide tree over the composed instances, whose leaves
// (~k: instance record k) are entered in entry-distance order like internal children; entering
// one pushes kInstExit, moves the ray into the instance's space (t keeps its meaning: the
// direction is not renormalized) and walks the instance's own 8-wide tree with triangles tested
// inline; popping kInstExit restores the world ray. Composed instances are numbered in
// composition order (instance i of the scene, then its instance j), so the closest hit is the
// smallest t with ties to the smaller (composed instance, triangle) = (inst_id0, inst_id1, geomID,
// primID).
template <bool ANY, bool STATS, bool MULTI>
__global__ __launch_bounds__(SGPU_TRACE_BLOCK) void k_trace8i(const WideNode* __restrict top, const TriRecord* __restrict top_tris,
                                                              uint32_t top_n_tris, uint32_t n_inst,
                                                              const InstRecord* __restrict insts,
                                                              const uint32_t* __restrict inst_ids,
                                                              const TraceScene* __restrict table,
                                                              const uint32_t* __restrict ray_scene, sgpu_rays_soa rays,
                                                              sgpu_hits_soa hits, uint32_t* __restrict occluded,
                                                              uint32_t n, const uint32_t* __restrict n_ptr,
                                                              uint32_t* __restrict spill, uint32_t stack_cap,
                                                              unsigned long long* __restrict stats,
                                                              uint32_t* __restrict overflow, uint32_t* __restrict comp_out)
{
    __shared__ uint32_t lds_stack[SGPU_STACK_LDS][SGPU_TRACE_BLOCK];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * SGPU_TRACE_BLOCK + threadIdx.x;
    const uint32_t stride = gridDim.x * SGPU_TRACE_BLOCK;
    unsigned long long cnt_nodes = 0, cnt_tris = 0;
    const float inf = __int_as_float(0x7f800000);
    if (n_ptr) n = *n_ptr;  // the count written on the device by the kernel that queued the rays

    for (uint32_t i = gtid; i < n; i += stride) {
        const f3 wo = mk3(rays.org_x[i], rays.org_y[i], rays.org_z[i]);
        const f3 wd = mk3(rays.dir_x[i], rays.dir_y[i], rays.dir_z[i]);
        const float tnear = rays.tnear[i];
        const float tfar = rays.tfar[i];
        f3 o = wo, d = wd;
        f3 inv = mk3(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
        f3 oinv = mk3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
        // the ray's scene: an instanced one (top tree over instances) or, in a multi-scene launch,
        // possibly a scene of geometries (top tree over triangles)
        const WideNode* __restrict tw = top;
        const TriRecord* __restrict tt = top_tris;
        uint32_t tn = top_n_tris, ni = n_inst;
        const InstRecord* __restrict ri = insts;
        if (MULTI) {
            const TraceScene ts = table[ray_scene[i]];
            tw = ts.wide;
            tt = ts.tris;
            tn = ts.n_tris;
            ni = ts.n_inst;
            ri = ts.insts;
        }
        const bool has_inst = ni > 0;
        const WideNode* __restrict rw = tw;
        const TriRecord* __restrict rt = tt;
        uint32_t cur = SGPU_INVALID_ID;  // the instance being walked; SGPU_INVALID_ID: the top tree

        float best_t = tfar;
        uint32_t best_key = 0xFFFFFFFFu, best_tri = 0xFFFFFFFFu, best_inst = 0xFFFFFFFFu;
        float best_u = 0.0f, best_v = 0.0f;
        bool hit_any = false;

        uint32_t sp = 0;
        bool overflowed = false;
        auto push = [&](int x) {
            if (sp >= stack_cap) {
                overflowed = true;
                return;
            }
            if (sp < SGPU_STACK_LDS) {
                lds_stack[sp][lane] = (uint32_t)x;
            } else {
                const uint32_t slot = sp % SGPU_STACK_LDS;
                spill[(size_t)(sp - SGPU_STACK_LDS) * stride + gtid] = lds_stack[slot][lane];
                lds_stack[slot][lane] = (uint32_t)x;
            }
            ++sp;
        };
        auto pop = [&](int& x) -> bool {
            if (sp == 0) return false;
            --sp;
            const uint32_t slot = sp % SGPU_STACK_LDS;
            x = (int)lds_stack[slot][lane];
            if (sp >= SGPU_STACK_LDS) lds_stack[slot][lane] = spill[(size_t)(sp - SGPU_STACK_LDS) * stride + gtid];
            return true;
        };
        auto test_tri = [&](int ref) -> bool {  // true: any-hit found
            const uint32_t j = (uint32_t)(~ref);
            const TriRecord r = rt[j];
            if (STATS) ++cnt_tris;
            float t, u, v;
            if (tri_test(r, o, d, tnear, t, u, v)) {
                if (ANY) {
                    if (t <= tfar) return true;
                } else {
                    const uint32_t key = __float_as_uint(r.a.w);
                    if (t < best_t || (t == best_t && (cur < best_inst || (cur == best_inst && key < best_key)))) {
                        best_t = t;
                        best_key = key;
                        best_tri = j;
                        best_inst = cur;
                        best_u = u;
                        best_v = v;
                    }
                }
            }
            return false;
        };

        int node = 0;
        const bool walking = has_inst || tn > 0;
        while (walking && !overflowed) {
            // Leaving or entering an instance falls through to the node visit of the same trip: the
            // entry popped after an exit is a top-tree entry (exits are only pushed on entering an
            // instance from the top tree), and an entered instance starts at its root.
            if (node == kInstExit) {
                // back in the top tree
                o = wo;
                d = wd;
                inv = mk3(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
                oinv = mk3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
                rw = tw;
                cur = SGPU_INVALID_ID;
                if (!pop(node)) break;
            }
            if (node < 0) {
                // a leaf of the top tree: enter composed instance ~node
                cur = (uint32_t)(~node);
                const InstRecord ir = ri[cur];
                o = inst_point(ir.w2l, wo);
                d = inst_vector(ir.w2l, wd);
                inv = mk3(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
                oinv = mk3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
                rw = ir.wide;
                rt = ir.tris;
                push(kInstExit);
                if (overflowed) break;
                node = 0;
            }
            const uint4* __restrict np = reinterpret_cast<const uint4*>(rw + node);
            const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4], w5 = np[5];
            if (STATS) ++cnt_nodes;
            const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
            const uint32_t meta = w0.w;
            const float sx = __uint_as_float((meta & 0xFFu) << 23);
            const float sy = __uint_as_float(((meta >> 8) & 0xFFu) << 23);
            const float sz = __uint_as_float(((meta >> 16) & 0xFFu) << 23);
            const uint32_t nch = meta >> 24;
            const uint32_t lox[2] = {w1.x, w1.y}, loy[2] = {w1.z, w1.w}, loz[2] = {w2.x, w2.y};
            const uint32_t hix[2] = {w2.z, w2.w}, hiy[2] = {w3.x, w3.y}, hiz[2] = {w3.z, w3.w};
            int ref[8] = {(int)w4.x, (int)w4.y, (int)w4.z, (int)w4.w, (int)w5.x, (int)w5.y, (int)w5.z, (int)w5.w};
            const bool in_top = cur == SGPU_INVALID_ID && has_inst;
            float key[8];
            uint32_t leaves = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int sh = 8 * (c & 3), wi = c >> 2;
                const float bx0 = fmaf((float)((lox[wi] >> sh) & 0xFFu), sx, ox);
                const float bx1 = fmaf((float)((hix[wi] >> sh) & 0xFFu), sx, ox);
                const float by0 = fmaf((float)((loy[wi] >> sh) & 0xFFu), sy, oy);
                const float by1 = fmaf((float)((hiy[wi] >> sh) & 0xFFu), sy, oy);
                const float bz0 = fmaf((float)((loz[wi] >> sh) & 0xFFu), sz, oz);
                const float bz1 = fmaf((float)((hiz[wi] >> sh) & 0xFFu), sz, oz);
                const float l_x = fmaf(bx0, inv.x, -oinv.x), h_x = fmaf(bx1, inv.x, -oinv.x);
                const float l_y = fmaf(by0, inv.y, -oinv.y), h_y = fmaf(by1, inv.y, -oinv.y);
                const float l_z = fmaf(bz0, inv.z, -oinv.z), h_z = fmaf(bz1, inv.z, -oinv.z);
                const float tmin = fmaxf(fmaxf(fminf(l_x, h_x), fminf(l_y, h_y)), fmaxf(fminf(l_z, h_z), tnear));
                const float tmax = fminf(fminf(fmaxf(l_x, h_x), fmaxf(l_y, h_y)), fminf(fmaxf(l_z, h_z), best_t));
                const bool hit = (uint32_t)c < nch && tmin <= tmax;
                // the top tree's instance leaves are visited like internal children
                const bool visit = hit && (ref[c] >= 0 || in_top);
                key[c] = visit ? tmin : inf;
                if (hit && ref[c] < 0 && !in_top) leaves |= 1u << c;
            }
            bool found = false;
            while (leaves) {
                const int c = __builtin_ctz(leaves);
                leaves &= leaves - 1;
                int r = ref[0];
#pragma unroll
                for (int k = 1; k < 8; ++k) r = c == k ? ref[k] : r;
                if (test_tri(r)) {
                    found = true;
                    break;
                }
            }
            if (ANY && found) {
                hit_any = true;
                break;
            }
            if (!ANY) {
#pragma unroll
                for (int c = 0; c < 8; ++c) key[c] = key[c] <= best_t ? key[c] : inf;
                cswap(key[0], ref[0], key[1], ref[1]); cswap(key[2], ref[2], key[3], ref[3]);
                cswap(key[4], ref[4], key[5], ref[5]); cswap(key[6], ref[6], key[7], ref[7]);
                cswap(key[0], ref[0], key[2], ref[2]); cswap(key[1], ref[1], key[3], ref[3]);
                cswap(key[4], ref[4], key[6], ref[6]); cswap(key[5], ref[5], key[7], ref[7]);
                cswap(key[1], ref[1], key[2], ref[2]); cswap(key[5], ref[5], key[6], ref[6]);
                cswap(key[0], ref[0], key[4], ref[4]); cswap(key[1], ref[1], key[5], ref[5]);
                cswap(key[2], ref[2], key[6], ref[6]); cswap(key[3], ref[3], key[7], ref[7]);
                cswap(key[2], ref[2], key[4], ref[4]); cswap(key[3], ref[3], key[5], ref[5]);
                cswap(key[1], ref[1], key[2], ref[2]); cswap(key[3], ref[3], key[4], ref[4]);
                cswap(key[5], ref[5], key[6], ref[6]);
            }
            // the nearest is entered, the others pushed far to near
            bool have = false;
            int next = 0;
#pragma unroll
            for (int c = 7; c >= 0; --c) {
                if (key[c] != inf) {
                    if (have) push(next);
                    next = ref[c];
                    have = true;
                }
            }
            if (have) {
                node = next;
                continue;
            }
            if (!pop(node)) break;
        }
        if (overflowed) atomicOr(overflow, 1u);

        uint32_t id0 = SGPU_INVALID_ID, id1 = SGPU_INVALID_ID;
        const TriRecord* __restrict bt = tt;
        if (!ANY && best_tri != 0xFFFFFFFFu && best_inst != SGPU_INVALID_ID) {
            bt = ri[best_inst].tris;
            if (inst_ids) {
                id0 = inst_ids[2 * best_inst];
                id1 = inst_ids[2 * best_inst + 1];
            }
        }
        write_result<ANY>(i, hit_any, best_tri, best_t, best_u, best_v, bt, hits, occluded, id0, id1);
        if (!ANY && comp_out) comp_out[i] = best_tri != 0xFFFFFFFFu ? best_inst : SGPU_INVALID_ID;
    }
    if (STATS) {
        atomicAdd(&stats[0], cnt_nodes);
        atomicAdd(&stats[1], cnt_tris);
    }
}

}  // namespace


namespace {

template <bool ANY, bool STATS>
void launch_trace(bool wide, bool multi, uint32_t blocks, hipStream_t stream, const BvhNode* nodes,
                  const WideNode* wnodes, uint32_t n_top, const TriRecord* tris, uint32_t n_tris, const TraceScene* table,
                  const uint32_t* ray_scene, const sgpu_rays_soa& rays, const sgpu_hits_soa& h, uint32_t* occluded,
                  uint32_t n, const uint32_t* n_ptr, uint32_t* spill, uint32_t cap, unsigned long long* st,
                  uint32_t* overflow, uint32_t* work)
{
    if (wide) {
        sgpu_trace_work_zero(work, stream);
        if (multi)
            hipLaunchKernelGGL((k_trace8<ANY, STATS, true>), dim3(blocks), dim3(SGPU_TRACE_BLOCK), 0, stream, wnodes,
                               n_top, tris, n_tris, table, ray_scene, rays, h, occluded, n, n_ptr, spill, cap, st, overflow,
                               work);
        else
            hipLaunchKernelGGL((k_trace8<ANY, STATS, false>), dim3(blocks), dim3(SGPU_TRACE_BLOCK), 0, stream, wnodes,
                               n_top, tris, n_tris, table, ray_scene, rays, h, occluded, n, n_ptr, spill, cap, st, overflow,
                               work);
    } else {
        if (multi)
            hipLaunchKernelGGL((k_trace<ANY, STATS, true>), dim3(blocks), dim3(SGPU_TRACE_BLOCK), 0, stream, nodes,
                               tris, n_tris, table, ray_scene, rays, h, occluded, n, n_ptr, spill, cap, st, overflow);
        else
            hipLaunchKernelGGL((k_trace<ANY, STATS, false>), dim3(blocks), dim3(SGPU_TRACE_BLOCK), 0, stream, nodes,
                               tris, n_tris, table, ray_scene, rays, h, occluded, n, n_ptr, spill, cap, st, overflow);
    }
}

// One traversal launch over n rays: of one scene (its trees `nodes` / `wnodes`), or (table !=
// nullptr) of the scenes of a batch, ray i in table[ray_scene[i]]; stack_cap bounds every scene's
// stack. The context's trace_tree picks the tree kind.
int trace_launch(sgpu_context ctx, const BvhNode* nodes, const WideNode* wnodes, uint32_t n_wide, const TriRecord* tris,
                 uint32_t n_tris, uint32_t stack_cap, uint32_t* overflow, const TraceScene* table,
                 const uint32_t* ray_scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits, uint32_t* occluded,
                 uint32_t n, bool any_hit, uint64_t* d_stats, hipStream_t stream, uint32_t* spill_buf,
                 const uint32_t* n_ptr, bool instanced = false, uint32_t* comp_out = nullptr, uint32_t cap_blocks = 0)
{
    if (n == 0) return SGPU_OK;
    const uint32_t cus = (uint32_t)ctx->props.multiProcessorCount;
    uint32_t blocks = (n + SGPU_TRACE_BLOCK - 1) / SGPU_TRACE_BLOCK;
    uint32_t max_blocks = cus * std::min(ctx->trace_blocks_per_cu, kMaxTraceBlocksPerCu);
    if (cap_blocks && max_blocks > cap_blocks) max_blocks = cap_blocks;
    if (blocks > max_blocks) blocks = max_blocks;

    // stream-ordered scratch for the spilled part of the stacks, private to this call, unless the
    // caller lends a buffer of sgpu_trace_spill_words (sized for the largest grid); its first
    // kTraceWorkWords words are the launch's ray counters (zero between launches)
    uint32_t* spill = spill_buf;
    if (!spill)
        SGPU_HIP_TRY(hipMallocAsync((void**)&spill, sgpu_trace_spill_words(ctx, stack_cap) * sizeof(uint32_t), stream));
    uint32_t* const buf = spill;
    uint32_t* const work = spill;
    spill += kTraceWorkWords;
    sgpu_hits_soa h = {};
    if (hits) h = *hits;
    unsigned long long* st = (unsigned long long*)d_stats;
    const bool multi = table != nullptr;
    const bool wide = ctx->trace_tree == 1;
    const uint32_t n_top = std::min(n_wide, std::min(ctx->trace_lds_nodes, kTopNodes));
    if (multi && instanced) {
        // a table with instanced scenes: k_trace8i walks both kinds of entry
        const dim3 g(blocks), b(SGPU_TRACE_BLOCK);
#define SGPU_LAUNCH_IM(A, S)                                                                                       \
    hipLaunchKernelGGL((k_trace8i<A, S, true>), g, b, 0, stream, nullptr, nullptr, 0u, 0u, nullptr, nullptr, table, \
                       ray_scene, *rays, h, occluded, n, n_ptr, spill, stack_cap, st, overflow, comp_out)
        if (any_hit) {
            if (st) SGPU_LAUNCH_IM(true, true);
            else SGPU_LAUNCH_IM(true, false);
        } else {
            if (st) SGPU_LAUNCH_IM(false, true);
            else SGPU_LAUNCH_IM(false, false);
        }
#undef SGPU_LAUNCH_IM
        SGPU_HIP_TRY(hipGetLastError());
        if (!spill_buf) SGPU_HIP_TRY(hipFreeAsync(buf, stream));
        return SGPU_OK;
    }
    if (any_hit) {
        if (st)
            launch_trace<true, true>(wide, multi, blocks, stream, nodes, wnodes, n_top, tris, n_tris, table, ray_scene, *rays,
                                     h, occluded, n, n_ptr, spill, stack_cap, st, overflow, work);
        else
            launch_trace<true, false>(wide, multi, blocks, stream, nodes, wnodes, n_top, tris, n_tris, table, ray_scene, *rays,
                                      h, occluded, n, n_ptr, spill, stack_cap, st, overflow, work);
    } else {
        if (st)
            launch_trace<false, true>(wide, multi, blocks, stream, nodes, wnodes, n_top, tris, n_tris, table, ray_scene, *rays,
                                      h, occluded, n, n_ptr, spill, stack_cap, st, overflow, work);
        else
            launch_trace<false, false>(wide, multi, blocks, stream, nodes, wnodes, n_top, tris, n_tris, table, ray_scene,
                                       *rays, h, occluded, n, n_ptr, spill, stack_cap, st, overflow, work);
    }
    SGPU_HIP_TRY(hipGetLastError());
    if (!spill_buf) SGPU_HIP_TRY(hipFreeAsync(buf, stream));
    return SGPU_OK;
}

}  // namespace

size_t sgpu_trace_spill_words(sgpu_context ctx, uint32_t stack_cap)
{
    const uint32_t cus = (uint32_t)ctx->props.multiProcessorCount;
    const uint32_t threads = cus * kMaxTraceBlocksPerCu * SGPU_TRACE_BLOCK;
    const uint32_t spill_depth = stack_cap > SGPU_STACK_LDS ? stack_cap - SGPU_STACK_LDS : 1;
    return kTraceWorkWords + (size_t)spill_depth * threads;
}

namespace {
__global__ void k_zero_work(uint32_t* __restrict w)
{
    if (threadIdx.x < kTraceWorkWords) w[threadIdx.x] = 0u;
}
}  // namespace

void sgpu_trace_work_zero(uint32_t* spill, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_work, dim3(1), dim3(kTraceWorkWords), 0, stream, spill);
}

size_t sgpu_trace_spill_words(sgpu_scene scene) { return sgpu_trace_spill_words(scene->ctx, scene->stack_cap); }

namespace {

// One traversal launch over the rays of an instanced scene (k_trace8i).
int trace_launch_instanced(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits, uint32_t* occluded,
                           uint32_t n, bool any_hit, uint64_t* d_stats, hipStream_t stream, uint32_t* overflow,
                           uint32_t* spill_buf, const uint32_t* n_ptr, uint32_t* comp_out, uint32_t cap_blocks)
{
    if (n == 0) return SGPU_OK;
    sgpu_context ctx = scene->ctx;
    const uint32_t cus = (uint32_t)ctx->props.multiProcessorCount;
    uint32_t blocks = (n + SGPU_TRACE_BLOCK - 1) / SGPU_TRACE_BLOCK;
    uint32_t max_blocks = cus * std::min(ctx->trace_blocks_per_cu, kMaxTraceBlocksPerCu);
    if (cap_blocks && max_blocks > cap_blocks) max_blocks = cap_blocks;
    if (blocks > max_blocks) blocks = max_blocks;
    uint32_t* spill = spill_buf;
    if (!spill)
        SGPU_HIP_TRY(hipMallocAsync((void**)&spill, sgpu_trace_spill_words(ctx, scene->stack_cap) * sizeof(uint32_t),
                                    stream));
    uint32_t* const buf = spill;
    spill += kTraceWorkWords;  // the counter words (unused by the instanced traversal) precede the spill
    sgpu_hits_soa h = {};
    if (hits) h = *hits;
    unsigned long long* st = (unsigned long long*)d_stats;
    const dim3 g(blocks), b(SGPU_TRACE_BLOCK);
#define SGPU_LAUNCH_I(A, S)                                                                                        \
    hipLaunchKernelGGL((k_trace8i<A, S, false>), g, b, 0, stream, scene->d_wide, nullptr, 0u, scene->n_inst,      \
                       scene->d_inst, scene->d_inst_ids, nullptr, nullptr, *rays, h, occluded, n, n_ptr, spill,    \
                       scene->stack_cap, st, overflow, comp_out)
    if (any_hit) {
        if (st) SGPU_LAUNCH_I(true, true);
        else SGPU_LAUNCH_I(true, false);
    } else {
        if (st) SGPU_LAUNCH_I(false, true);
        else SGPU_LAUNCH_I(false, false);
    }
#undef SGPU_LAUNCH_I
    SGPU_HIP_TRY(hipGetLastError());
    if (!spill_buf) SGPU_HIP_TRY(hipFreeAsync(buf, stream));
    return SGPU_OK;
}

}  // namespace

int sgpu_trace_launch(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits, uint32_t* occluded,
                      uint32_t n, bool any_hit, uint64_t* d_stats, hipStream_t stream, uint32_t* spill_buf,
                      const uint32_t* n_ptr, uint32_t* overflow, uint32_t* comp_out, uint32_t max_blocks,
                      uint64_t* curve_stats)
{
    int rc;
    if (!scene->instances.empty())
        rc = trace_launch_instanced(scene, rays, hits, occluded, n, any_hit, d_stats, stream,
                                    overflow ? overflow : scene->d_overflow, spill_buf, n_ptr, comp_out, max_blocks);
    else
        rc = trace_launch(scene->ctx, scene->d_nodes, scene->d_wide, scene->n_wide, scene->d_tris, scene->n_tris,
                          scene->stack_cap, overflow ? overflow : scene->d_overflow, nullptr, nullptr, rays, hits,
                          occluded, n, any_hit, d_stats, stream, spill_buf, n_ptr, false, nullptr, max_blocks);
    if (rc != SGPU_OK || scene->n_curves == 0) return rc;
    return sgpu_curve_launch(scene->d_cnodes, scene->d_curves, scene->n_curves, scene->d_inst, scene->d_inst_ids,
                             nullptr, nullptr, rays, hits, occluded, n, any_hit, n_ptr, comp_out,
                             overflow ? overflow : scene->d_overflow, stream, curve_stats);
}

int sgpu_trace_launch_multi(sgpu_context ctx, const TraceScene* table, const uint32_t* ray_scene, uint32_t stack_cap,
                            uint32_t* overflow, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits,
                            uint32_t* occluded, uint32_t n, bool any_hit, uint64_t* d_stats, hipStream_t stream,
                            uint32_t* spill_buf, const uint32_t* n_ptr, bool instanced, uint32_t* comp_out,
                            bool curves)
{
    int rc = trace_launch(ctx, nullptr, nullptr, 0, nullptr, 0, stack_cap, overflow, table, ray_scene, rays, hits,
                          occluded, n, any_hit, d_stats, stream, spill_buf, n_ptr, instanced, comp_out);
    if (rc != SGPU_OK || !curves) return rc;
    return sgpu_curve_launch(nullptr, nullptr, 0, nullptr, nullptr, table, ray_scene, rays, hits, occluded, n, any_hit,
                             n_ptr, comp_out, overflow, stream);
}
