// Device BVH construction for sgpu_scene_commit (the replacement of Embree's rtcCommitScene,
// Source/Core/SceneLib/ModelResource.cpp:447-472).
//
// Pipeline (all on the device, one stream):
//   1. k_prep     triangle records (v0, e1, e2, ids), per-triangle boxes and centroids, and a
//                 block-reduced scene box and centroid box.
//   2. k_morton   63-bit Morton codes (21 bits per axis) of centroids in the centroid box.
//   3. rocPRIM    device radix sort of (code, triangle) pairs.
//   4. k_karras   binary radix tree over the sorted codes (Karras 2012, "Maximizing Parallelism
//                 in the Construction of BVHs, Octrees, and k-d Trees", Alg. 1 + 2); equal codes
//                 are split by position, so every key is distinct.
//   5. k_fit      bottom-up box fitting with one arrival counter per node: the second thread to
//                 arrive at a node writes that node's record (both child boxes + child links).
//   6. k_reorder  triangle records in leaf order; k_depth the deepest leaf (traversal stack size).
//
// Boxes are inflated by 2^-18 of the scene's largest absolute coordinate. The boxes only cull;
// the inflation keeps culling conservative with respect to the triangle test's rounding, so
// closest hits do not depend on the tree's shape.
#include "sgpu_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>  // rocprim/iterator/texture_cache_iterator.hpp uses memset without including it
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t f2ord(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// bounds layout (12 ordered-uint slots): [0..2] scene lo, [3..5] scene hi (stored as ~ord so that
// atomicMin works for both), [6..8] centroid lo, [9..11] centroid hi (also as ~ord).
__global__ __launch_bounds__(kBlock) void k_prep(const float* __restrict pos, uint32_t n,
                                                 const uint32_t* __restrict tri_geom,
                                                 const uint32_t* __restrict geom_first,
                                                 TriRecord* __restrict rec, float4* __restrict blo,
                                                 float4* __restrict bhi, float4* __restrict cent,
                                                 uint32_t* __restrict bounds)
{
    __shared__ uint32_t red[12][kBlock / 64];
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t v[12];
    for (int k = 0; k < 12; ++k) v[k] = 0xFFFFFFFFu;
    if (i < n) {
        const float* p = pos + (size_t)i * 9;
        f3 v0 = mk3(p[0], p[1], p[2]);
        f3 v1 = mk3(p[3], p[4], p[5]);
        f3 v2 = mk3(p[6], p[7], p[8]);
        f3 e1 = v1 - v0;
        f3 e2 = v2 - v0;
        uint32_t g = tri_geom[i];
        // quad geometries (bit 31 of their first index) keep prim = 2 * quad + half, flagged in bit 31
        const uint32_t fg = geom_first[g];
        uint32_t prim = (i - (fg & 0x7FFFFFFFu)) | (fg & 0x80000000u);
        TriRecord r;
        r.a = make_float4(v0.x, v0.y, v0.z, __uint_as_float(i));
        r.b = make_float4(e1.x, e1.y, e1.z, __uint_as_float(g));
        r.c = make_float4(e2.x, e2.y, e2.z, __uint_as_float(prim));
        rec[i] = r;
        f3 lo = mk3(fminf(v0.x, fminf(v1.x, v2.x)), fminf(v0.y, fminf(v1.y, v2.y)), fminf(v0.z, fminf(v1.z, v2.z)));
        f3 hi = mk3(fmaxf(v0.x, fmaxf(v1.x, v2.x)), fmaxf(v0.y, fmaxf(v1.y, v2.y)), fmaxf(v0.z, fmaxf(v1.z, v2.z)));
        f3 c = mk3(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z));
        blo[i] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        bhi[i] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        cent[i] = make_float4(c.x, c.y, c.z, 0.0f);
        v[0] = f2ord(lo.x); v[1] = f2ord(lo.y); v[2] = f2ord(lo.z);
        v[3] = ~f2ord(hi.x); v[4] = ~f2ord(hi.y); v[5] = ~f2ord(hi.z);
        v[6] = f2ord(c.x); v[7] = f2ord(c.y); v[8] = f2ord(c.z);
        v[9] = ~f2ord(c.x); v[10] = ~f2ord(c.y); v[11] = ~f2ord(c.z);
    }
    // wave64 min-reduction, then across the block's waves, then one atomic per slot per block
    for (int k = 0; k < 12; ++k) {
        uint32_t x = v[k];
        for (int off = 32; off > 0; off >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, off, 64));
        v[k] = x;
    }
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 12; ++k) red[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x < 12) {
        uint32_t x = red[threadIdx.x][0];
        for (int w = 1; w < kBlock / 64; ++w) x = min(x, red[threadIdx.x][w]);
        atomicMin(&bounds[threadIdx.x], x);
    }
}

__device__ __forceinline__ uint64_t spread21(uint64_t x)
{
    x &= 0x1fffffULL;
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

__global__ __launch_bounds__(kBlock) void k_morton(const float4* __restrict cent, uint32_t n,
                                                   const uint32_t* __restrict bounds,
                                                   uint64_t* __restrict keys, uint32_t* __restrict vals)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float lo[3], ext[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = ord2f(bounds[6 + k]);
        float hi = ord2f(~bounds[9 + k]);
        ext[k] = hi - lo[k];
    }
    float4 c = cent[i];
    float cc[3] = {c.x, c.y, c.z};
    uint64_t q[3];
    for (int k = 0; k < 3; ++k) {
        float f = ext[k] > 0.0f ? (cc[k] - lo[k]) / ext[k] : 0.0f;
        f = fminf(fmaxf(f, 0.0f), 1.0f);
        q[k] = (uint64_t)fminf(f * 2097152.0f, 2097151.0f);
    }
    keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[i] = i;
}

__device__ __forceinline__ int delta(const uint64_t* __restrict k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clzll((long long)(a ^ b));
}

// Internal node i of n-1. Children are stored as internal index, or ~leaf for leaves.
__global__ __launch_bounds__(kBlock) void k_karras(const uint64_t* __restrict keys, int n,
                                                   int2* __restrict child, int* __restrict parent_int,
                                                   int* __restrict parent_leaf)
{
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0;
    int div = 2;
    while (true) {
        int t = (l + div - 1) / div;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
        div <<= 1;
    }
    int gamma = i + s * d + min(d, 0);
    int left, right;
    if (min(i, j) == gamma) {
        left = ~gamma;
        parent_leaf[gamma] = i;
    } else {
        left = gamma;
        parent_int[gamma] = i;
    }
    if (max(i, j) == gamma + 1) {
        right = ~(gamma + 1);
        parent_leaf[gamma + 1] = i;
    } else {
        right = gamma + 1;
        parent_int[gamma + 1] = i;
    }
    child[i] = make_int2(left, right);
    if (i == 0) parent_int[0] = -1;
}

__device__ __forceinline__ float inflate_eps(const uint32_t* __restrict bounds)
{
    float m = 0.0f;
    for (int k = 0; k < 3; ++k) {
        m = fmaxf(m, fabsf(ord2f(bounds[k])));
        m = fmaxf(m, fabsf(ord2f(~bounds[3 + k])));
    }
    return m * (1.0f / 262144.0f);
}

__device__ __forceinline__ void leaf_box(const float4* __restrict blo, const float4* __restrict bhi,
                                         const uint32_t* __restrict order, int j, float eps,
                                         float4& lo, float4& hi)
{
    uint32_t t = order[j];
    lo = blo[t];
    hi = bhi[t];
    lo.x -= eps; lo.y -= eps; lo.z -= eps;
    hi.x += eps; hi.y += eps; hi.z += eps;
}

__global__ __launch_bounds__(kBlock) void k_fit(int n, const int2* __restrict child,
                                                const int* __restrict parent_int,
                                                const int* __restrict parent_leaf,
                                                const float4* __restrict blo, const float4* __restrict bhi,
                                                const uint32_t* __restrict order,
                                                const uint32_t* __restrict bounds, float4* ilo, float4* ihi,
                                                uint32_t* flags, BvhNode* __restrict nodes)
{
    int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    float eps = inflate_eps(bounds);
    int p = parent_leaf[j];
    // every climb step moves one level up; the bound and the range check only guard against a
    // malformed tree (they never trigger for a radix tree)
    for (int steps = 0; p >= 0 && p < n - 1 && steps < n; ++steps) {
        __threadfence();
        uint32_t old = atomicAdd(&flags[p], 1u);
        if (old == 0) return;
        __threadfence();
        int2 c = child[p];
        float4 lo0, hi0, lo1, hi1;
        if (c.x >= 0) {
            lo0 = ilo[c.x];
            hi0 = ihi[c.x];
        } else {
            leaf_box(blo, bhi, order, ~c.x, eps, lo0, hi0);
        }
        if (c.y >= 0) {
            lo1 = ilo[c.y];
            hi1 = ihi[c.y];
        } else {
            leaf_box(blo, bhi, order, ~c.y, eps, lo1, hi1);
        }
        BvhNode nd;
        nd.x = make_float4(lo0.x, hi0.x, lo1.x, hi1.x);
        nd.y = make_float4(lo0.y, hi0.y, lo1.y, hi1.y);
        nd.z = make_float4(lo0.z, hi0.z, lo1.z, hi1.z);
        nd.child = make_int4(c.x, c.y, 0, 0);
        nodes[p] = nd;
        ilo[p] = make_float4(fminf(lo0.x, lo1.x), fminf(lo0.y, lo1.y), fminf(lo0.z, lo1.z), 0.0f);
        ihi[p] = make_float4(fmaxf(hi0.x, hi1.x), fmaxf(hi0.y, hi1.y), fmaxf(hi0.z, hi1.z), 0.0f);
        p = parent_int[p];
    }
}

// Single-triangle scene: one root whose two children are the same leaf.
__global__ void k_single(const float4* __restrict blo, const float4* __restrict bhi,
                         const uint32_t* __restrict bounds, BvhNode* __restrict nodes)
{
    float eps = inflate_eps(bounds);
    float4 lo = blo[0], hi = bhi[0];
    lo.x -= eps; lo.y -= eps; lo.z -= eps;
    hi.x += eps; hi.y += eps; hi.z += eps;
    BvhNode nd;
    nd.x = make_float4(lo.x, hi.x, lo.x, hi.x);
    nd.y = make_float4(lo.y, hi.y, lo.y, hi.y);
    nd.z = make_float4(lo.z, hi.z, lo.z, hi.z);
    nd.child = make_int4(~0, ~0, 0, 0);
    nodes[0] = nd;
}

__global__ __launch_bounds__(kBlock) void k_reorder(const TriRecord* __restrict in,
                                                    const uint32_t* __restrict order, uint32_t n,
                                                    TriRecord* __restrict out)
{
    uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j < n) out[j] = in[order[j]];
}

// ---- PLOC: the binary tree by parallel locally-ordered clustering ----------------------------------
//
// Meister and Bittner, "Parallel Locally-Ordered Clustering for Bounding Volume Hierarchy
// Construction" (IEEE TVCG 2018). The clusters start as the leaves in Morton order; every round,
// each cluster finds the neighbour within kPlocRadius positions whose union box has the smallest
// surface area, mutual nearest neighbours merge into a new internal node (at the lower position),
// and the surviving clusters are compacted in order. Edges are ordered by (area, lower position,
// higher position), so the nearest-neighbour relation has no cycles and every round merges at
// least one pair. Node indices are handed out from n - 2 downwards in position order, so the last
// merge is the root, node 0, and the tree is deterministic.

constexpr int kPlocRadius = 16;

__device__ __forceinline__ float union_area(float4 alo, float4 ahi, float4 blo, float4 bhi)
{
    const float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
    const float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
    const float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
    return dx * dy + dy * dz + dz * dx;
}

// The leaves as the first clusters: leaf j (Morton order) with its inflated box, height 0.
__global__ __launch_bounds__(kBlock) void k_ploc_init(uint32_t n, const float4* __restrict blo,
                                                      const float4* __restrict bhi, const uint32_t* __restrict order,
                                                      const uint32_t* __restrict bounds, int* __restrict cid,
                                                      float4* __restrict clo, float4* __restrict chi,
                                                      uint32_t* __restrict height)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const float eps = inflate_eps(bounds);
    float4 lo, hi;
    leaf_box(blo, bhi, order, (int)j, eps, lo, hi);
    cid[j] = ~(int)j;
    clo[j] = lo;
    chi[j] = hi;
    height[j] = 0;
}

// nn[i] = the neighbour of cluster i (within kPlocRadius positions) of least union area, edges
// ordered by (area, lower position, higher position). The block's clusters and the radius on both
// sides are staged in LDS.
__global__ __launch_bounds__(kBlock) void k_ploc_nn(uint32_t m, const float4* __restrict clo,
                                                    const float4* __restrict chi, uint32_t* __restrict nn)
{
    __shared__ float4 slo[kBlock + 2 * kPlocRadius], shi[kBlock + 2 * kPlocRadius];
    const int first = (int)(blockIdx.x * kBlock) - kPlocRadius;
    for (int k = threadIdx.x; k < kBlock + 2 * kPlocRadius; k += kBlock) {
        const int g = first + k;
        if (g >= 0 && g < (int)m) {
            slo[k] = clo[g];
            shi[k] = chi[g];
        }
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= (int)m) return;
    const int li = i - first;
    const float4 lo = slo[li], hi = shi[li];
    float best = 0.0f;
    int bj = -1;
    const int j0 = i - kPlocRadius < 0 ? 0 : i - kPlocRadius;
    const int j1 = i + kPlocRadius >= (int)m ? (int)m - 1 : i + kPlocRadius;
    for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const float a = union_area(lo, hi, slo[j - first], shi[j - first]);
        // (a, min, max) against the best edge so far; with i fixed, a smaller other end wins ties
        if (bj < 0 || a < best || (a == best && j < bj)) {
            best = a;
            bj = j;
        }
    }
    nn[i] = (uint32_t)bj;
}

// Per position: bit 0 of keep = the cluster survives this round (it is not the higher end of a
// merging pair), bit 32 = it is the lower end of a merging pair (a new node).
__global__ __launch_bounds__(kBlock) void k_ploc_flags(uint32_t m, const uint32_t* __restrict nn,
                                                       uint64_t* __restrict flags)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const uint32_t j = nn[i];
    const bool mutual = nn[j] == i;
    const uint64_t keep = (mutual && j < i) ? 0ull : 1ull;
    const uint64_t lead = (mutual && i < j) ? 1ull : 0ull;
    flags[i] = keep | (lead << 32);
}

// The round's merges (node base + rank among this round's merges, written as BvhNodes) and the
// surviving clusters, compacted in position order into the other buffers.
__global__ __launch_bounds__(kBlock) void k_ploc_merge(uint32_t m, uint32_t node_base, const uint32_t* __restrict nn,
                                                       const uint64_t* __restrict flags,
                                                       const uint64_t* __restrict prefix, const int* __restrict cid,
                                                       const float4* __restrict clo, const float4* __restrict chi,
                                                       const uint32_t* __restrict height, int* __restrict ncid,
                                                       float4* __restrict nlo, float4* __restrict nhi,
                                                       uint32_t* __restrict nheight, BvhNode* __restrict nodes)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const uint64_t f = flags[i];
    if (!(f & 1ull)) return;
    const uint32_t pos = (uint32_t)(prefix[i] & 0xFFFFFFFFull);
    if (f >> 32) {
        const uint32_t j = nn[i];
        const uint32_t id = node_base + (uint32_t)(prefix[i] >> 32);
        const float4 lo0 = clo[i], hi0 = chi[i], lo1 = clo[j], hi1 = chi[j];
        BvhNode nd;
        nd.x = make_float4(lo0.x, hi0.x, lo1.x, hi1.x);
        nd.y = make_float4(lo0.y, hi0.y, lo1.y, hi1.y);
        nd.z = make_float4(lo0.z, hi0.z, lo1.z, hi1.z);
        nd.child = make_int4(cid[i], cid[j], 0, 0);
        nodes[id] = nd;
        ncid[pos] = (int)id;
        nlo[pos] = make_float4(fminf(lo0.x, lo1.x), fminf(lo0.y, lo1.y), fminf(lo0.z, lo1.z), 0.0f);
        nhi[pos] = make_float4(fmaxf(hi0.x, hi1.x), fmaxf(hi0.y, hi1.y), fmaxf(hi0.z, hi1.z), 0.0f);
        nheight[pos] = 1 + max(height[i], height[j]);
    } else {
        ncid[pos] = cid[i];
        nlo[pos] = clo[i];
        nhi[pos] = chi[i];
        nheight[pos] = height[i];
    }
}

__global__ __launch_bounds__(kBlock) void k_depth(int n, const int* __restrict parent_int,
                                                  const int* __restrict parent_leaf, uint32_t* maxdepth)
{
    int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    uint32_t d = 1;
    int p = parent_leaf[j];
    while (p > 0 && p < n - 1 && d <= (uint32_t)n) {
        p = parent_int[p];
        ++d;
    }
    atomicMax(maxdepth, d);
}

// ---- 8-wide collapse (breadth-first, one level per pair of launches) ----
//
// A wide node starts from its binary node's two children and repeatedly opens the internal child
// of largest surface area (the first such child on ties) until it has 8 children or only leaves.
// k_wide_expand writes each task's children (binary refs and boxes) and its count of internal
// children; an exclusive scan of those counts numbers the next level's nodes in task order, so
// the numbering (and the tree) is deterministic; k_wide_emit quantizes the boxes and writes the
// node and the next level's tasks.

__device__ __forceinline__ float half_area(const float* b)
{
    float dx = fmaxf(b[3] - b[0], 0.0f), dy = fmaxf(b[4] - b[1], 0.0f), dz = fmaxf(b[5] - b[2], 0.0f);
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ void child_box(const BvhNode& nd, int c, float* b)
{
    b[0] = c ? nd.x.z : nd.x.x;
    b[1] = c ? nd.y.z : nd.y.x;
    b[2] = c ? nd.z.z : nd.z.x;
    b[3] = c ? nd.x.w : nd.x.y;
    b[4] = c ? nd.y.w : nd.y.y;
    b[5] = c ? nd.z.w : nd.z.y;
}

__global__ __launch_bounds__(kBlock) void k_wide_expand(const int* __restrict tasks, uint32_t n_tasks,
                                                        const BvhNode* __restrict bn, int* __restrict cref,
                                                        float* __restrict cbox, uint32_t* __restrict nchild,
                                                        uint32_t* __restrict ninternal)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_tasks) return;
    int ref[8];
    float box[8][6];
    const BvhNode root = bn[tasks[t]];
    ref[0] = root.child.x;
    ref[1] = root.child.y;
    child_box(root, 0, box[0]);
    child_box(root, 1, box[1]);
    uint32_t m = 2;
    while (m < 8) {
        int best = -1;
        float best_a = -1.0f;
        for (uint32_t k = 0; k < m; ++k)
            if (ref[k] >= 0) {
                float a = half_area(box[k]);
                if (a > best_a) {
                    best_a = a;
                    best = (int)k;
                }
            }
        if (best < 0) break;
        const BvhNode nd = bn[ref[best]];
        ref[best] = nd.child.x;
        child_box(nd, 0, box[best]);
        ref[m] = nd.child.y;
        child_box(nd, 1, box[m]);
        ++m;
    }
    uint32_t internal = 0;
    for (uint32_t k = 0; k < m; ++k) {
        cref[8 * (size_t)t + k] = ref[k];
        for (int a = 0; a < 6; ++a) cbox[48 * (size_t)t + 6 * k + a] = box[k][a];
        internal += ref[k] >= 0 ? 1u : 0u;
    }
    nchild[t] = m;
    ninternal[t] = internal;
}

// Smallest e in [-126, 127] with ext <= 254 * 2^e (ext >= 0).
__device__ __forceinline__ int plane_exponent(double ext)
{
    if (!(ext > 0.0)) return -126;
    int k;
    double f = frexp(ext / 254.0, &k);  // ext / 254 = f * 2^k, f in [0.5, 1)
    int e = f == 0.5 ? k - 1 : k;
    return e < -126 ? -126 : (e > 127 ? 127 : e);
}

__global__ __launch_bounds__(kBlock) void k_wide_emit(uint32_t n_tasks, uint32_t level_base, uint32_t next_base,
                                                      const uint32_t* __restrict offsets, const int* __restrict cref,
                                                      const float* __restrict cbox, const uint32_t* __restrict nchild,
                                                      WideNode* __restrict wide, int* __restrict next_tasks,
                                                      uint32_t* __restrict bad)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_tasks) return;
    const uint32_t m = nchild[t];
    const float* B = cbox + 48 * (size_t)t;
    float lo[3] = {B[0], B[1], B[2]}, hi[3] = {B[3], B[4], B[5]};
    for (uint32_t k = 1; k < m; ++k)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], B[6 * k + a]);
            hi[a] = fmaxf(hi[a], B[6 * k + 3 + a]);
        }
    WideNode w;
    w.ox = lo[0];
    w.oy = lo[1];
    w.oz = lo[2];
    uint32_t meta = m << 24;
    for (int a = 0; a < 3; ++a) {
        w.qlo[a][0] = w.qlo[a][1] = 0;
        w.qhi[a][0] = w.qhi[a][1] = 0;
        const int e = plane_exponent((double)hi[a] - (double)lo[a]);
        const float sc = __uint_as_float((uint32_t)(e + 127) << 23);
        const double sd = (double)sc;
        meta |= (uint32_t)(e + 127) << (8 * a);
        for (uint32_t k = 0; k < m; ++k) {
            const float clo = B[6 * k + a], chi = B[6 * k + 3 + a];
            double ql = floor(((double)clo - (double)lo[a]) / sd);
            double qh = ceil(((double)chi - (double)lo[a]) / sd);
            int qlo = ql < 0.0 ? 0 : (ql > 255.0 ? 255 : (int)ql);
            int qhi = qh < 0.0 ? 0 : (qh > 255.0 ? 255 : (int)qh);
            // the decoded planes (one FP32 rounding, as the traversal decodes them) must contain the box
            while (qlo > 0 && fmaf((float)qlo, sc, lo[a]) > clo) --qlo;
            while (qhi < 255 && fmaf((float)qhi, sc, lo[a]) < chi) ++qhi;
            if (fmaf((float)qlo, sc, lo[a]) > clo || fmaf((float)qhi, sc, lo[a]) < chi) atomicOr(bad, 1u);
            w.qlo[a][k >> 2] |= (uint32_t)qlo << (8 * (k & 3));
            w.qhi[a][k >> 2] |= (uint32_t)qhi << (8 * (k & 3));
        }
    }
    w.meta = meta;
    uint32_t rank = 0;
    for (uint32_t k = 0; k < 8; ++k) {
        int r = k < m ? cref[8 * (size_t)t + k] : -1;
        if (k < m && r >= 0) {
            next_tasks[offsets[t] + rank] = r;
            r = (int)(next_base + offsets[t] + rank);
            ++rank;
        }
        w.child[k] = r;
    }
    for (int k = 0; k < 8; ++k) w.pad[k] = 0;
    wide[level_base + t] = w;
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    hipError_t alloc(size_t count) { return hipMalloc((void**)&p, count * sizeof(T) + 16); }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// Collapses the binary tree (nodes, n_int internal nodes, root 0) into 8-wide nodes.
int build_wide(sgpu_scene scene, const BvhNode* nodes, uint32_t n_int, hipStream_t stream)
{
    DevBuf<int> d_tasks, d_next, d_cref;
    DevBuf<float> d_cbox;
    DevBuf<uint32_t> d_nch, d_nint, d_off, d_bad;
    DevBuf<WideNode> d_all;
    DevBuf<uint8_t> d_tmp;
    SGPU_HIP_TRY(d_tasks.alloc(n_int));
    SGPU_HIP_TRY(d_next.alloc(n_int));
    SGPU_HIP_TRY(d_cref.alloc(8 * (size_t)n_int));
    SGPU_HIP_TRY(d_cbox.alloc(48 * (size_t)n_int));
    SGPU_HIP_TRY(d_nch.alloc(n_int));
    SGPU_HIP_TRY(d_nint.alloc(n_int));
    SGPU_HIP_TRY(d_off.alloc(n_int));
    SGPU_HIP_TRY(d_bad.alloc(1));
    SGPU_HIP_TRY(d_all.alloc(n_int));
    size_t tmp_bytes = 0;
    SGPU_HIP_TRY(rocprim::exclusive_scan(nullptr, tmp_bytes, d_nint.p, d_off.p, 0u, (size_t)n_int,
                                         rocprim::plus<uint32_t>(), stream));
    SGPU_HIP_TRY(d_tmp.alloc(tmp_bytes));
    const int root = 0;
    SGPU_HIP_TRY(hipMemcpyAsync(d_tasks.p, &root, sizeof(int), hipMemcpyHostToDevice, stream));
    SGPU_HIP_TRY(hipMemsetAsync(d_bad.p, 0, sizeof(uint32_t), stream));
    uint32_t level_base = 0, n_tasks = 1, depth = 0;
    while (n_tasks > 0) {
        const uint32_t grid = (n_tasks + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_wide_expand, dim3(grid), dim3(kBlock), 0, stream, d_tasks.p, n_tasks, nodes, d_cref.p,
                           d_cbox.p, d_nch.p, d_nint.p);
        SGPU_HIP_TRY(hipGetLastError());
        SGPU_HIP_TRY(rocprim::exclusive_scan(d_tmp.p, tmp_bytes, d_nint.p, d_off.p, 0u, (size_t)n_tasks,
                                             rocprim::plus<uint32_t>(), stream));
        uint32_t last[2] = {0, 0};
        SGPU_HIP_TRY(hipMemcpyAsync(&last[0], d_off.p + n_tasks - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        SGPU_HIP_TRY(hipMemcpyAsync(&last[1], d_nint.p + n_tasks - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        const uint32_t next_base = level_base + n_tasks;
        hipLaunchKernelGGL(k_wide_emit, dim3(grid), dim3(kBlock), 0, stream, n_tasks, level_base, next_base, d_off.p,
                           d_cref.p, d_cbox.p, d_nch.p, d_all.p, d_next.p, d_bad.p);
        SGPU_HIP_TRY(hipGetLastError());
        SGPU_HIP_TRY(hipStreamSynchronize(stream));
        const uint32_t n_next = last[0] + last[1];
        if ((uint64_t)next_base + n_next > n_int) return sgpu_fail(SGPU_E_RANGE, "wide tree larger than the binary tree");
        std::swap(d_tasks.p, d_next.p);
        level_base = next_base;
        n_tasks = n_next;
        ++depth;
    }
    uint32_t bad = 0;
    SGPU_HIP_TRY(hipMemcpy(&bad, d_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (bad) return sgpu_fail(SGPU_E_RANGE, "wide node quantization could not contain a child box");
    WideNode* wide = nullptr;
    SGPU_HIP_TRY(hipMalloc((void**)&wide, (size_t)level_base * sizeof(WideNode)));
    hipError_t e = hipMemcpyAsync(wide, d_all.p, (size_t)level_base * sizeof(WideNode), hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        (void)hipFree(wide);
        return sgpu_fail(SGPU_E_HIP, std::string("wide tree copy: ") + hipGetErrorString(e));
    }
    scene->d_wide = wide;
    scene->n_wide = level_base;
    scene->wide_depth = depth;
    return SGPU_OK;
}

}  // namespace

namespace {

float host_ord2f(uint32_t u)
{
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &b, sizeof(f));
    return f;
}

uint32_t host_f2ord(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, sizeof(u));
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// The binary tree over n primitive boxes (d_blo, d_bhi, centroids d_cent, bounds as k_prep
// leaves them): Morton codes, radix sort, Karras hierarchy, bottom-up fit. On success *nodes_out
// holds max(n - 1, 1) nodes, d_order the primitive of each leaf in leaf order and *depth the
// deepest leaf.
int build_binary(uint32_t n, const float4* d_blo, const float4* d_bhi, const float4* d_cent, const uint32_t* d_bounds,
                 hipStream_t stream, BvhNode** nodes_out, DevBuf<uint32_t>& d_order, uint32_t* depth_out,
                 uint32_t builder)
{
    const uint32_t n_int = n > 1 ? n - 1 : 1;
    BvhNode* nodes = nullptr;
    SGPU_HIP_TRY(hipMalloc((void**)&nodes, (size_t)n_int * sizeof(BvhNode)));
    auto fail_free = [&](int code, const std::string& msg) {
        (void)hipFree(nodes);
        return sgpu_fail(code, msg);
    };
#define SGPU_BUILD_TRY(expr)                                                                  \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return fail_free(SGPU_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)
    SGPU_BUILD_TRY(d_order.alloc(n));
    if (n == 1) {
        hipLaunchKernelGGL(k_single, dim3(1), dim3(1), 0, stream, d_blo, d_bhi, d_bounds, nodes);
        SGPU_BUILD_TRY(hipGetLastError());
        SGPU_BUILD_TRY(hipMemsetAsync(d_order.p, 0, sizeof(uint32_t), stream));
        SGPU_BUILD_TRY(hipStreamSynchronize(stream));
        *nodes_out = nodes;
        *depth_out = 1;
        return SGPU_OK;
    }
    DevBuf<uint32_t> d_vals, d_flags, d_depth;
    DevBuf<float4> d_ilo, d_ihi;
    DevBuf<uint64_t> d_keys, d_keys2;
    DevBuf<int2> d_child;
    DevBuf<int> d_pint, d_pleaf;
    DevBuf<uint8_t> d_tmp;
    SGPU_BUILD_TRY(d_keys.alloc(n));
    SGPU_BUILD_TRY(d_keys2.alloc(n));
    SGPU_BUILD_TRY(d_vals.alloc(n));
    SGPU_BUILD_TRY(d_child.alloc(n_int));
    SGPU_BUILD_TRY(d_pint.alloc(n_int));
    SGPU_BUILD_TRY(d_pleaf.alloc(n));
    SGPU_BUILD_TRY(d_flags.alloc(n_int));
    SGPU_BUILD_TRY(d_ilo.alloc(n_int));
    SGPU_BUILD_TRY(d_ihi.alloc(n_int));
    SGPU_BUILD_TRY(d_depth.alloc(1));
    SGPU_BUILD_TRY(hipMemsetAsync(d_flags.p, 0, (size_t)n_int * sizeof(uint32_t), stream));
    SGPU_BUILD_TRY(hipMemsetAsync(d_pint.p, 0xFF, (size_t)n_int * sizeof(int), stream));
    SGPU_BUILD_TRY(hipMemsetAsync(d_pleaf.p, 0xFF, (size_t)n * sizeof(int), stream));
    SGPU_BUILD_TRY(hipMemsetAsync(d_depth.p, 0, sizeof(uint32_t), stream));

    const uint32_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_morton, dim3(grid), dim3(kBlock), 0, stream, d_cent, n, d_bounds, d_keys.p, d_vals.p);
    SGPU_BUILD_TRY(hipGetLastError());
    size_t tmp_bytes = 0;
    SGPU_BUILD_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, d_keys.p, d_keys2.p, d_vals.p, d_order.p,
                                             (size_t)n, 0, 63, stream));
    SGPU_BUILD_TRY(d_tmp.alloc(tmp_bytes));
    SGPU_BUILD_TRY(rocprim::radix_sort_pairs(d_tmp.p, tmp_bytes, d_keys.p, d_keys2.p, d_vals.p, d_order.p,
                                             (size_t)n, 0, 63, stream));
    if (builder == 1) {
        // PLOC over the Morton-ordered leaves
        DevBuf<int> cid[2];
        DevBuf<float4> clo[2], chi[2];
        DevBuf<uint32_t> hgt[2], nn;
        DevBuf<uint64_t> flags, prefix, last;
        for (int b = 0; b < 2; ++b) {
            SGPU_BUILD_TRY(cid[b].alloc(n));
            SGPU_BUILD_TRY(clo[b].alloc(n));
            SGPU_BUILD_TRY(chi[b].alloc(n));
            SGPU_BUILD_TRY(hgt[b].alloc(n));
        }
        SGPU_BUILD_TRY(nn.alloc(n));
        SGPU_BUILD_TRY(flags.alloc(n));
        SGPU_BUILD_TRY(prefix.alloc(n));
        size_t scan_bytes = 0;
        SGPU_BUILD_TRY(rocprim::exclusive_scan(nullptr, scan_bytes, flags.p, prefix.p, (uint64_t)0, (size_t)n,
                                               rocprim::plus<uint64_t>(), stream));
        DevBuf<uint8_t> d_scan;
        SGPU_BUILD_TRY(d_scan.alloc(scan_bytes));
        hipLaunchKernelGGL(k_ploc_init, dim3(grid), dim3(kBlock), 0, stream, n, d_blo, d_bhi, d_order.p, d_bounds,
                           cid[0].p, clo[0].p, chi[0].p, hgt[0].p);
        SGPU_BUILD_TRY(hipGetLastError());
        uint32_t m = n, next_id = n - 1, cur = 0;
        uint64_t tail[2] = {0, 0};
        while (m > 1) {
            const uint32_t g = (m + kBlock - 1) / kBlock;
            hipLaunchKernelGGL(k_ploc_nn, dim3(g), dim3(kBlock), 0, stream, m, clo[cur].p, chi[cur].p, nn.p);
            hipLaunchKernelGGL(k_ploc_flags, dim3(g), dim3(kBlock), 0, stream, m, nn.p, flags.p);
            SGPU_BUILD_TRY(hipGetLastError());
            SGPU_BUILD_TRY(rocprim::exclusive_scan(d_scan.p, scan_bytes, flags.p, prefix.p, (uint64_t)0, (size_t)m,
                                                   rocprim::plus<uint64_t>(), stream));
            SGPU_BUILD_TRY(hipMemcpyAsync(&tail[0], prefix.p + m - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
            SGPU_BUILD_TRY(hipMemcpyAsync(&tail[1], flags.p + m - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
            SGPU_BUILD_TRY(hipStreamSynchronize(stream));
            const uint64_t tot = tail[0] + tail[1];
            const uint32_t kept = (uint32_t)(tot & 0xFFFFFFFFull), merges = (uint32_t)(tot >> 32);
            if (merges == 0 || merges > next_id || kept != m - merges)
                return fail_free(SGPU_E_STATE, "PLOC: a clustering round made no progress");
            next_id -= merges;
            hipLaunchKernelGGL(k_ploc_merge, dim3(g), dim3(kBlock), 0, stream, m, next_id, nn.p, flags.p, prefix.p,
                               cid[cur].p, clo[cur].p, chi[cur].p, hgt[cur].p, cid[1 - cur].p, clo[1 - cur].p,
                               chi[1 - cur].p, hgt[1 - cur].p, nodes);
            SGPU_BUILD_TRY(hipGetLastError());
            cur = 1 - cur;
            m = kept;
        }
        uint32_t h = 0;
        SGPU_BUILD_TRY(hipMemcpyAsync(&h, hgt[cur].p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        SGPU_BUILD_TRY(hipStreamSynchronize(stream));
        if (next_id != 0) return fail_free(SGPU_E_STATE, "PLOC: the tree does not end at node 0");
        *nodes_out = nodes;
        *depth_out = h;
        return SGPU_OK;
    }
    const uint32_t grid_int = (n - 1 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_karras, dim3(grid_int), dim3(kBlock), 0, stream, d_keys2.p, (int)n, d_child.p, d_pint.p,
                       d_pleaf.p);
    SGPU_BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_fit, dim3(grid), dim3(kBlock), 0, stream, (int)n, d_child.p, d_pint.p, d_pleaf.p, d_blo, d_bhi,
                       d_order.p, d_bounds, d_ilo.p, d_ihi.p, d_flags.p, nodes);
    SGPU_BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_depth, dim3(grid), dim3(kBlock), 0, stream, (int)n, d_pint.p, d_pleaf.p, d_depth.p);
    SGPU_BUILD_TRY(hipGetLastError());
    uint32_t depth = 0;
    SGPU_BUILD_TRY(hipMemcpyAsync(&depth, d_depth.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    SGPU_BUILD_TRY(hipStreamSynchronize(stream));
#undef SGPU_BUILD_TRY
    *nodes_out = nodes;
    *depth_out = depth;
    return SGPU_OK;
}

}  // namespace

namespace {

// Box of each curve segment (its control points' box grown by the largest radius: the B-spline
// basis weights are non-negative and sum to one, so the swept balls stay inside), centroid, and the
// ordered-uint bounds k_prep computes.
__global__ __launch_bounds__(kBlock) void k_curve_prep(const CurveRecord* __restrict rec, uint32_t n,
                                                       float4* __restrict blo, float4* __restrict bhi,
                                                       float4* __restrict cent, uint32_t* __restrict bounds)
{
    __shared__ uint32_t red[12][kBlock / 64];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t v[12];
    for (int k = 0; k < 12; ++k) v[k] = 0xFFFFFFFFu;
    if (i < n) {
        const CurveRecord r = rec[i];
        f3 lo = mk3(r.p[0].x, r.p[0].y, r.p[0].z), hi = lo;
        float rmax = r.p[0].w;
        for (int j = 1; j < 4; ++j) {
            lo = mk3(fminf(lo.x, r.p[j].x), fminf(lo.y, r.p[j].y), fminf(lo.z, r.p[j].z));
            hi = mk3(fmaxf(hi.x, r.p[j].x), fmaxf(hi.y, r.p[j].y), fmaxf(hi.z, r.p[j].z));
            rmax = fmaxf(rmax, r.p[j].w);
        }
        lo = mk3(lo.x - rmax, lo.y - rmax, lo.z - rmax);
        hi = mk3(hi.x + rmax, hi.y + rmax, hi.z + rmax);
        const f3 c = mk3(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z));
        blo[i] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        bhi[i] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        cent[i] = make_float4(c.x, c.y, c.z, 0.0f);
        v[0] = f2ord(lo.x); v[1] = f2ord(lo.y); v[2] = f2ord(lo.z);
        v[3] = ~f2ord(hi.x); v[4] = ~f2ord(hi.y); v[5] = ~f2ord(hi.z);
        v[6] = f2ord(c.x); v[7] = f2ord(c.y); v[8] = f2ord(c.z);
        v[9] = ~f2ord(c.x); v[10] = ~f2ord(c.y); v[11] = ~f2ord(c.z);
    }
    for (int k = 0; k < 12; ++k) {
        uint32_t x = v[k];
        for (int off = 32; off > 0; off >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, off, 64));
        v[k] = x;
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 12; ++k) red[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x < 12) {
        uint32_t x = red[threadIdx.x][0];
        for (int w = 1; w < kBlock / 64; ++w) x = min(x, red[threadIdx.x][w]);
        atomicMin(&bounds[threadIdx.x], x);
    }
}

__global__ __launch_bounds__(kBlock) void k_curve_reorder(const CurveRecord* __restrict in,
                                                          const uint32_t* __restrict order, uint32_t n,
                                                          CurveRecord* __restrict out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = in[order[i]];
}

}  // namespace

int sgpu_curve_build(sgpu_scene scene, const std::vector<float>& cp, const std::vector<uint32_t>& geom,
                     const std::vector<uint32_t>& prim, const std::vector<uint32_t>& inst, hipStream_t stream)
{
    const uint32_t n = (uint32_t)geom.size();
    scene->n_curves = 0;
    if (n == 0) return SGPU_OK;
    if (n > (1u << 28)) return sgpu_fail(SGPU_E_RANGE, "scene has more than 2^28 curve segments");
    std::vector<CurveRecord> rec(n);
    for (uint32_t s = 0; s < n; ++s) {
        const float* c = cp.data() + 16 * (size_t)s;
        for (int j = 0; j < 4; ++j) rec[s].p[j] = make_float4(c[4 * j], c[4 * j + 1], c[4 * j + 2], c[4 * j + 3]);
        rec[s].geom = geom[s];
        rec[s].prim = prim[s];
        rec[s].inst = inst[s];
        rec[s].pad = 0;
    }
    DevBuf<CurveRecord> d_rec;
    DevBuf<float4> d_blo, d_bhi, d_cent;
    DevBuf<uint32_t> d_bounds, d_order;
    SGPU_HIP_TRY(d_rec.alloc(n));
    SGPU_HIP_TRY(d_blo.alloc(n));
    SGPU_HIP_TRY(d_bhi.alloc(n));
    SGPU_HIP_TRY(d_cent.alloc(n));
    SGPU_HIP_TRY(d_bounds.alloc(12));
    SGPU_HIP_TRY(hipMemcpyAsync(d_rec.p, rec.data(), n * sizeof(CurveRecord), hipMemcpyHostToDevice, stream));
    SGPU_HIP_TRY(hipMemsetAsync(d_bounds.p, 0xFF, 12 * sizeof(uint32_t), stream));
    const uint32_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_curve_prep, dim3(grid), dim3(kBlock), 0, stream, d_rec.p, n, d_blo.p, d_bhi.p, d_cent.p,
                       d_bounds.p);
    SGPU_HIP_TRY(hipGetLastError());
    uint32_t hb[12];
    SGPU_HIP_TRY(hipMemcpyAsync(hb, d_bounds.p, sizeof(hb), hipMemcpyDeviceToHost, stream));
    BvhNode* nodes = nullptr;
    uint32_t depth = 0;
    int rc = build_binary(n, d_blo.p, d_bhi.p, d_cent.p, d_bounds.p, stream, &nodes, d_order, &depth, 0);
    if (rc != SGPU_OK) return rc;
    if (depth + 1 > kCurveStack) {
        (void)hipFree(nodes);
        return sgpu_fail(SGPU_E_RANGE, "curve tree deeper than the curve traversal's stack");
    }
    CurveRecord* out = nullptr;
    hipError_t e = hipMalloc((void**)&out, (size_t)n * sizeof(CurveRecord));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_curve_reorder, dim3(grid), dim3(kBlock), 0, stream, d_rec.p, d_order.p, n, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        (void)hipFree(nodes);
        if (out) (void)hipFree(out);
        return sgpu_fail(SGPU_E_HIP, std::string("curve records: ") + hipGetErrorString(e));
    }
    scene->d_cnodes = nodes;
    scene->d_curves = out;
    scene->n_curves = n;
    // the scene's box takes in the curves
    for (int k = 0; k < 3; ++k) {
        const float lo = host_ord2f(hb[k]), hi = host_ord2f(~hb[3 + k]);
        if (scene->box[k] > scene->box[3 + k]) {
            scene->box[k] = lo;
            scene->box[3 + k] = hi;
        } else {
            scene->box[k] = std::min(scene->box[k], lo);
            scene->box[3 + k] = std::max(scene->box[3 + k], hi);
        }
    }
    return SGPU_OK;
}

int sgpu_bvh_build(sgpu_scene scene, hipStream_t stream)
{
    const uint32_t n = (uint32_t)scene->tri_geom.size();
    scene->n_tris = n;
    scene->n_nodes = 0;
    scene->max_depth = 0;
    if (n == 0) return SGPU_OK;
    if (n > (1u << 30)) return sgpu_fail(SGPU_E_RANGE, "scene has more than 2^30 triangles");

    DevBuf<float> d_pos;
    DevBuf<uint32_t> d_geom, d_first, d_bounds, d_order;
    DevBuf<TriRecord> d_rec;
    DevBuf<float4> d_blo, d_bhi, d_cent;
    SGPU_HIP_TRY(d_pos.alloc((size_t)n * 9));
    SGPU_HIP_TRY(d_geom.alloc(n));
    SGPU_HIP_TRY(d_first.alloc(scene->geom_first.size()));
    SGPU_HIP_TRY(d_bounds.alloc(12));
    SGPU_HIP_TRY(d_rec.alloc(n));
    SGPU_HIP_TRY(d_blo.alloc(n));
    SGPU_HIP_TRY(d_bhi.alloc(n));
    SGPU_HIP_TRY(d_cent.alloc(n));

    SGPU_HIP_TRY(hipMemcpyAsync(d_pos.p, scene->tri_pos.data(), (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice,
                                stream));
    SGPU_HIP_TRY(hipMemcpyAsync(d_geom.p, scene->tri_geom.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice,
                                stream));
    // device copy of the geometries' first indices with bit 31 set for quad geometries
    std::vector<uint32_t> first_flagged(scene->geom_first.size());
    for (size_t g = 0; g < first_flagged.size(); ++g)
        first_flagged[g] = scene->geom_first[g] | (scene->geom_quad[g] == 1 ? 0x80000000u : 0u);
    SGPU_HIP_TRY(hipMemcpy(d_first.p, first_flagged.data(), first_flagged.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice));
    SGPU_HIP_TRY(hipMemsetAsync(d_bounds.p, 0xFF, 12 * sizeof(uint32_t), stream));
    const uint32_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_prep, dim3(grid), dim3(kBlock), 0, stream, d_pos.p, n, d_geom.p, d_first.p, d_rec.p, d_blo.p,
                       d_bhi.p, d_cent.p, d_bounds.p);
    SGPU_HIP_TRY(hipGetLastError());
    uint32_t hb[12];
    SGPU_HIP_TRY(hipMemcpyAsync(hb, d_bounds.p, sizeof(hb), hipMemcpyDeviceToHost, stream));

    BvhNode* nodes = nullptr;
    uint32_t depth = 0;
    int rc = build_binary(n, d_blo.p, d_bhi.p, d_cent.p, d_bounds.p, stream, &nodes, d_order, &depth,
                          scene->ctx->bvh_builder);
    if (rc != SGPU_OK) return rc;
    for (int k = 0; k < 3; ++k) {
        scene->box[k] = host_ord2f(hb[k]);
        scene->box[3 + k] = host_ord2f(~hb[3 + k]);
    }
    TriRecord* tris = nullptr;
    hipError_t e = hipMalloc((void**)&tris, (size_t)n * sizeof(TriRecord));
    if (e != hipSuccess) {
        (void)hipFree(nodes);
        return sgpu_fail(SGPU_E_NOMEM, "hipMalloc of triangle records failed");
    }
    hipLaunchKernelGGL(k_reorder, dim3(grid), dim3(kBlock), 0, stream, d_rec.p, d_order.p, n, tris);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        (void)hipFree(nodes);
        (void)hipFree(tris);
        return sgpu_fail(SGPU_E_HIP, std::string("triangle reorder: ") + hipGetErrorString(e));
    }
    scene->d_nodes = nodes;
    scene->d_tris = tris;
    scene->n_nodes = n > 1 ? n - 1 : 1;
    scene->max_depth = depth;
    rc = build_wide(scene, nodes, scene->n_nodes, stream);
    scene->stack_cap = std::max(scene->max_depth + 1, kWideMaxPush * scene->wide_depth + 1);
    return rc;
}

namespace {

// r = a * b for row-major 3x4 affine transforms (the implicit last row 0 0 0 1)
void affine_mul(const double* a, const double* b, double* r)
{
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 4; ++j) {
            double v = a[4 * i + 0] * b[0 + j] + a[4 * i + 1] * b[4 + j] + a[4 * i + 2] * b[8 + j];
            if (j == 3) v += a[4 * i + 3];
            r[4 * i + j] = v;
        }
    }
}

// inverse of a row-major 3x4 affine transform; false when singular
bool affine_inverse(const double* m, double* r)
{
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], k = m[10];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    const double inv[9] = {A * id, -(b * k - c * h) * id, (b * f - c * e) * id,
                           B * id, (a * k - c * g) * id,  -(a * f - c * d) * id,
                           C * id, -(a * h - b * g) * id, (a * e - b * d) * id};
    for (int i = 0; i < 3; ++i) {
        r[4 * i + 0] = inv[3 * i + 0];
        r[4 * i + 1] = inv[3 * i + 1];
        r[4 * i + 2] = inv[3 * i + 2];
        r[4 * i + 3] = -(inv[3 * i + 0] * m[3] + inv[3 * i + 1] * m[7] + inv[3 * i + 2] * m[11]);
    }
    return true;
}

}  // namespace

namespace {

// The curves of the composed instances (placed scene, local-to-world row-major 3 x 4 in double) in
// world space, each segment tagged with its composed instance, and the instanced scene's curve
// tree over them. A placement that scales unevenly or shears would turn the swept spheres into
// ellipsoids, which the world-space curve test does not represent: such a placement is refused.
int bake_instance_curves(sgpu_scene scene, const std::vector<std::pair<sgpu_scene, const double*>>& comp,
                         hipStream_t stream)
{
    std::vector<float> cp;
    std::vector<uint32_t> geom, prim, inst;
    for (uint32_t k = 0; k < comp.size(); ++k) {
        const sgpu_scene b = comp[k].first;
        const double* m = comp[k].second;
        if (b->curve_geom.empty()) continue;
        double n2[3], s2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            n2[c] = m[c] * m[c] + m[4 + c] * m[4 + c] + m[8 + c] * m[8 + c];
            s2 += n2[c] / 3.0;
        }
        bool similar = s2 > 0.0;
        for (int c = 0; c < 3 && similar; ++c) {
            similar &= std::fabs(n2[c] - s2) <= 1e-5 * s2;
            for (int e = c + 1; e < 3; ++e)
                similar &= std::fabs(m[c] * m[e] + m[4 + c] * m[4 + e] + m[8 + c] * m[8 + e]) <= 1e-5 * s2;
        }
        if (!similar)
            return sgpu_fail(SGPU_E_INVALID, "an instance placing curves is not a similarity (rotation, uniform scale, "
                                             "translation)");
        const double sc = std::sqrt(s2);
        const size_t nseg = b->curve_geom.size(), base = cp.size();
        cp.resize(base + 16 * nseg);
        for (size_t i = 0; i < nseg; ++i) {
            const float* src = b->curve_cp.data() + 16 * i;
            float* dst = cp.data() + base + 16 * i;
            for (int j = 0; j < 4; ++j) {
                const double p[3] = {src[4 * j], src[4 * j + 1], src[4 * j + 2]};
                for (int r = 0; r < 3; ++r)
                    dst[4 * j + r] = (float)(m[4 * r] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3]);
                dst[4 * j + 3] = (float)(src[4 * j + 3] * sc);
            }
            geom.push_back(b->curve_geom[i]);
            prim.push_back(b->curve_prim[i]);
            inst.push_back(k);
        }
    }
    return sgpu_curve_build(scene, cp, geom, prim, inst, stream);
}

}  // namespace

int sgpu_bvh_build_instances(sgpu_scene scene, hipStream_t stream)
{
    // compose the levels: (instance i of a scene of geometries) or (instance i of an instanced
    // scene, its instance j), in that order
    struct Flat {
        sgpu_scene blas;
        double l2w[12];
        uint32_t id0, id1;
    };
    std::vector<Flat> flat;
    uint32_t levels = 1;
    for (size_t i = 0; i < scene->instances.size(); ++i) {
        const auto& in = scene->instances[i];
        if (in.child->instances.empty()) {
            Flat f{in.child, {}, (uint32_t)i, SGPU_INVALID_ID};
            std::memcpy(f.l2w, in.l2w, sizeof(f.l2w));
            flat.push_back(f);
        } else {
            levels = 2;
            for (size_t j = 0; j < in.child->instances.size(); ++j) {
                const auto& in2 = in.child->instances[j];
                Flat f{in2.child, {}, (uint32_t)i, (uint32_t)j};
                affine_mul(in.l2w, in2.l2w, f.l2w);
                flat.push_back(f);
            }
        }
    }
    // scenes without triangles are not in the top tree; those with curves are composed after the
    // tree's instances (the curve pass reaches them through their curves), the others place nothing
    std::vector<Flat> extra;
    for (const Flat& f : flat)
        if (f.blas->n_tris == 0 && !f.blas->curve_geom.empty()) extra.push_back(f);
    flat.erase(std::remove_if(flat.begin(), flat.end(), [](const Flat& f) { return f.blas->n_tris == 0; }), flat.end());
    if (flat.size() + extra.size() > (1u << 30)) return sgpu_fail(SGPU_E_RANGE, "scene has more than 2^30 composed instances");
    const uint32_t n = (uint32_t)flat.size();
    scene->n_tris = 0;
    scene->n_nodes = 0;
    scene->max_depth = 0;
    scene->inst_levels = levels;
    scene->comp_blas.clear();
    scene->comp_l2w.clear();
    // composed instances past the tree's: records (world-to-local), ids, placed scene and transform
    std::vector<InstRecord> xrec(extra.size());
    std::vector<uint32_t> xids(2 * extra.size());
    for (size_t e = 0; e < extra.size(); ++e) {
        double w2l[12];
        if (!affine_inverse(extra[e].l2w, w2l)) return sgpu_fail(SGPU_E_INVALID, "instance transform is singular");
        for (int r = 0; r < 3; ++r)
            xrec[e].w2l[r] = make_float4((float)w2l[4 * r], (float)w2l[4 * r + 1], (float)w2l[4 * r + 2], (float)w2l[4 * r + 3]);
        xrec[e].wide = nullptr;
        xrec[e].tris = nullptr;
        xids[2 * e] = extra[e].id0;
        xids[2 * e + 1] = extra[e].id1;
    }
    auto composed = [&]() {
        std::vector<std::pair<sgpu_scene, const double*>> c;
        for (const Flat& f : flat) c.emplace_back(f.blas, f.l2w);
        for (const Flat& f : extra) c.emplace_back(f.blas, f.l2w);
        return c;
    };
    auto push_extra_shading = [&]() {
        for (const Flat& f : extra) {
            scene->comp_blas.push_back(f.blas);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) scene->comp_l2w.push_back((float)f.l2w[4 * r + c]);
        }
    };
    if (n == 0) {
        if (extra.empty()) return SGPU_OK;
        // only curves: no top tree; the composed instances' records for the curve pass and shading
        push_extra_shading();
        hipError_t e = hipMalloc((void**)&scene->d_inst, xrec.size() * sizeof(InstRecord));
        if (e == hipSuccess) e = hipMalloc((void**)&scene->d_inst_ids, xids.size() * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemcpy(scene->d_inst, xrec.data(), xrec.size() * sizeof(InstRecord), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(scene->d_inst_ids, xids.data(), xids.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) return sgpu_fail(SGPU_E_HIP, std::string("instance records: ") + hipGetErrorString(e));
        scene->n_inst = 0;
        scene->stack_cap = 1;
        return bake_instance_curves(scene, composed(), stream);
    }

    // The top tree's box of an instance: the boxes of the placed scene's root children (decoded
    // from its 8-wide root as the traversal decodes them) through the transform, their 8 corners in
    // double, united and rounded outward: much tighter than the transformed box of the whole scene
    // when the transform rotates. The scene's own box (the IBL's emission sphere) takes the 8
    // corners of each placed scene's box.
    std::vector<std::pair<sgpu_scene, std::vector<float>>> root_boxes;
    auto child_boxes = [&](sgpu_scene b, std::vector<float>& out) -> int {
        for (const auto& rb : root_boxes)
            if (rb.first == b) {
                out = rb.second;
                return SGPU_OK;
            }
        WideNode w;
        SGPU_HIP_TRY(hipMemcpy(&w, b->d_wide, sizeof(WideNode), hipMemcpyDeviceToHost));
        const uint32_t nch = w.meta >> 24;
        const float o[3] = {w.ox, w.oy, w.oz};
        float sc[3];
        for (int a = 0; a < 3; ++a) {
            const uint32_t bits = ((w.meta >> (8 * a)) & 0xFFu) << 23;
            std::memcpy(&sc[a], &bits, sizeof(float));
        }
        out.clear();
        for (uint32_t c = 0; c < nch && c < 8; ++c)
            for (int side = 0; side < 2; ++side)
                for (int a = 0; a < 3; ++a) {
                    const uint32_t word = side ? w.qhi[a][c >> 2] : w.qlo[a][c >> 2];
                    out.push_back(std::fma((float)((word >> (8 * (c & 3))) & 0xFFu), sc[a], o[a]));
                }
        root_boxes.emplace_back(b, out);
        return SGPU_OK;
    };
    std::vector<float4> blo(n), bhi(n), cent(n);
    std::vector<InstRecord> rec(n);
    std::vector<uint32_t> ids(2 * (size_t)n);
    float slo[3] = {INFINITY, INFINITY, INFINITY}, shi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t blas_cap = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const Flat& f = flat[k];
        double w2l[12];
        if (!affine_inverse(f.l2w, w2l)) return sgpu_fail(SGPU_E_INVALID, "instance transform is singular");
        // the 8 corners of a local box through the transform, in double
        auto world_box = [&](const float* b6, double* lo, double* hi) {
            for (int c = 0; c < 8; ++c) {
                const double p[3] = {b6[(c & 1) ? 3 : 0], b6[(c & 2) ? 4 : 1], b6[(c & 4) ? 5 : 2]};
                for (int r = 0; r < 3; ++r) {
                    const double w =
                        f.l2w[4 * r] * p[0] + f.l2w[4 * r + 1] * p[1] + f.l2w[4 * r + 2] * p[2] + f.l2w[4 * r + 3];
                    lo[r] = std::min(lo[r], w);
                    hi[r] = std::max(hi[r], w);
                }
            }
        };
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        world_box(f.blas->box, lo, hi);
        std::vector<float> cb;
        int rcb = child_boxes(f.blas, cb);
        if (rcb != SGPU_OK) return rcb;
        double tlo[3] = {INFINITY, INFINITY, INFINITY}, thi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t c = 0; c + 6 <= cb.size(); c += 6) world_box(cb.data() + c, tlo, thi);
        if (cb.empty())
            for (int r = 0; r < 3; ++r) {
                tlo[r] = lo[r];
                thi[r] = hi[r];
            }
        float l[3], h[3];
        for (int r = 0; r < 3; ++r) {
            const float sl = std::nextafter((float)lo[r], -INFINITY), sh = std::nextafter((float)hi[r], INFINITY);
            l[r] = std::nextafter((float)std::max(tlo[r], lo[r]), -INFINITY);
            h[r] = std::nextafter((float)std::min(thi[r], hi[r]), INFINITY);
            if (!std::isfinite(sl) || !std::isfinite(sh))
                return sgpu_fail(SGPU_E_RANGE, "instance box is not finite in single precision");
            slo[r] = std::min(slo[r], sl);
            shi[r] = std::max(shi[r], sh);
            const float cc = 0.5f * (l[r] + h[r]);
            clo[r] = std::min(clo[r], cc);
            chi[r] = std::max(chi[r], cc);
        }
        blo[k] = make_float4(l[0], l[1], l[2], 0.0f);
        bhi[k] = make_float4(h[0], h[1], h[2], 0.0f);
        cent[k] = make_float4(0.5f * (l[0] + h[0]), 0.5f * (l[1] + h[1]), 0.5f * (l[2] + h[2]), 0.0f);
        for (int r = 0; r < 3; ++r)
            rec[k].w2l[r] = make_float4((float)w2l[4 * r], (float)w2l[4 * r + 1], (float)w2l[4 * r + 2], (float)w2l[4 * r + 3]);
        rec[k].wide = f.blas->d_wide;
        rec[k].tris = f.blas->d_tris;
        scene->comp_blas.push_back(f.blas);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) scene->comp_l2w.push_back((float)f.l2w[4 * r + c]);
        ids[2 * k] = f.id0;
        ids[2 * k + 1] = f.id1;
        blas_cap = std::max(blas_cap, f.blas->stack_cap);
    }
    push_extra_shading();
    rec.insert(rec.end(), xrec.begin(), xrec.end());
    ids.insert(ids.end(), xids.begin(), xids.end());
    uint32_t hb[12];
    for (int r = 0; r < 3; ++r) {
        hb[r] = host_f2ord(slo[r]);
        hb[3 + r] = ~host_f2ord(shi[r]);
        hb[6 + r] = host_f2ord(clo[r]);
        hb[9 + r] = ~host_f2ord(chi[r]);
        scene->box[r] = slo[r];
        scene->box[3 + r] = shi[r];
    }
    DevBuf<float4> d_blo, d_bhi, d_cent;
    DevBuf<uint32_t> d_bounds, d_order;
    SGPU_HIP_TRY(d_blo.alloc(n));
    SGPU_HIP_TRY(d_bhi.alloc(n));
    SGPU_HIP_TRY(d_cent.alloc(n));
    SGPU_HIP_TRY(d_bounds.alloc(12));
    SGPU_HIP_TRY(hipMemcpyAsync(d_blo.p, blo.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
    SGPU_HIP_TRY(hipMemcpyAsync(d_bhi.p, bhi.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
    SGPU_HIP_TRY(hipMemcpyAsync(d_cent.p, cent.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
    SGPU_HIP_TRY(hipMemcpyAsync(d_bounds.p, hb, sizeof(hb), hipMemcpyHostToDevice, stream));
    BvhNode* nodes = nullptr;
    uint32_t depth = 0;
    int rc = build_binary(n, d_blo.p, d_bhi.p, d_cent.p, d_bounds.p, stream, &nodes, d_order, &depth,
                          scene->ctx->bvh_builder);
    if (rc != SGPU_OK) return rc;
    // the top tree's leaves name composed instances by their composition index (instance i, then
    // its instance j), so the closest-hit tie rule follows (inst_id0, inst_id1)
    std::vector<uint32_t> order(n);
    const uint32_t n_int = n > 1 ? n - 1 : 1;
    std::vector<BvhNode> hn(n_int);
    hipError_t e = hipMemcpy(order.data(), d_order.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hn.data(), nodes, n_int * sizeof(BvhNode), hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
        for (BvhNode& nd : hn) {
            if (nd.child.x < 0) nd.child.x = ~(int)order[~nd.child.x];
            if (nd.child.y < 0) nd.child.y = ~(int)order[~nd.child.y];
        }
        e = hipMemcpy(nodes, hn.data(), n_int * sizeof(BvhNode), hipMemcpyHostToDevice);
    }
    InstRecord* d_inst = nullptr;
    uint32_t* d_ids = nullptr;
    if (e == hipSuccess) e = hipMalloc((void**)&d_inst, rec.size() * sizeof(InstRecord));
    if (e == hipSuccess) e = hipMalloc((void**)&d_ids, ids.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpy(d_inst, rec.data(), rec.size() * sizeof(InstRecord), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_ids, ids.data(), ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(nodes);
        if (d_inst) (void)hipFree(d_inst);
        if (d_ids) (void)hipFree(d_ids);
        return sgpu_fail(SGPU_E_HIP, std::string("instance records: ") + hipGetErrorString(e));
    }
    scene->d_nodes = nodes;
    scene->d_inst = d_inst;
    scene->d_inst_ids = d_ids;
    scene->n_inst = n;
    scene->n_nodes = n > 1 ? n - 1 : 1;
    scene->max_depth = depth;
    scene->blas_stack_cap = blas_cap;
    rc = build_wide(scene, nodes, scene->n_nodes, stream);
    // the top tree's stack, the exit entry, and the deepest placed tree's stack
    scene->stack_cap = std::max(scene->max_depth + 1, kWideMaxPush * scene->wide_depth + 1) + 1 + blas_cap;
    if (rc != SGPU_OK) return rc;
    return bake_instance_curves(scene, composed(), stream);
}
