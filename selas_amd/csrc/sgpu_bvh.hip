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

int sgpu_bvh_build(sgpu_scene scene, hipStream_t stream)
{
    const uint32_t n = (uint32_t)scene->tri_geom.size();
    scene->n_tris = n;
    scene->n_nodes = 0;
    scene->max_depth = 0;
    if (n == 0) return SGPU_OK;
    if (n > (1u << 30)) return sgpu_fail(SGPU_E_RANGE, "scene has more than 2^30 triangles");

    DevBuf<float> d_pos;
    DevBuf<uint32_t> d_geom, d_first, d_bounds, d_vals, d_vals2, d_flags, d_depth;
    DevBuf<TriRecord> d_rec;
    DevBuf<float4> d_blo, d_bhi, d_cent, d_ilo, d_ihi;
    DevBuf<uint64_t> d_keys, d_keys2;
    DevBuf<int2> d_child;
    DevBuf<int> d_pint, d_pleaf;
    DevBuf<uint8_t> d_tmp;

    const uint32_t n_int = n > 1 ? n - 1 : 1;
    SGPU_HIP_TRY(d_pos.alloc((size_t)n * 9));
    SGPU_HIP_TRY(d_geom.alloc(n));
    SGPU_HIP_TRY(d_first.alloc(scene->geom_first.size()));
    SGPU_HIP_TRY(d_bounds.alloc(12));
    SGPU_HIP_TRY(d_rec.alloc(n));
    SGPU_HIP_TRY(d_blo.alloc(n));
    SGPU_HIP_TRY(d_bhi.alloc(n));
    SGPU_HIP_TRY(d_cent.alloc(n));
    SGPU_HIP_TRY(d_keys.alloc(n));
    SGPU_HIP_TRY(d_keys2.alloc(n));
    SGPU_HIP_TRY(d_vals.alloc(n));
    SGPU_HIP_TRY(d_vals2.alloc(n));
    SGPU_HIP_TRY(d_child.alloc(n_int));
    SGPU_HIP_TRY(d_pint.alloc(n_int));
    SGPU_HIP_TRY(d_pleaf.alloc(n));
    SGPU_HIP_TRY(d_flags.alloc(n_int));
    SGPU_HIP_TRY(d_ilo.alloc(n_int));
    SGPU_HIP_TRY(d_ihi.alloc(n_int));
    SGPU_HIP_TRY(d_depth.alloc(1));

    BvhNode* nodes = nullptr;
    TriRecord* tris = nullptr;
    SGPU_HIP_TRY(hipMalloc((void**)&nodes, (size_t)n_int * sizeof(BvhNode)));
    hipError_t e = hipMalloc((void**)&tris, (size_t)n * sizeof(TriRecord));
    if (e != hipSuccess) {
        (void)hipFree(nodes);
        return sgpu_fail(SGPU_E_NOMEM, "hipMalloc of triangle records failed");
    }
    auto fail_free = [&](int code, const std::string& msg) {
        (void)hipFree(nodes);
        (void)hipFree(tris);
        return sgpu_fail(code, msg);
    };
#define SGPU_BUILD_TRY(expr)                                                                  \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return fail_free(SGPU_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

    SGPU_BUILD_TRY(hipMemcpyAsync(d_pos.p, scene->tri_pos.data(), (size_t)n * 9 * sizeof(float),
                                  hipMemcpyHostToDevice, stream));
    SGPU_BUILD_TRY(hipMemcpyAsync(d_geom.p, scene->tri_geom.data(), (size_t)n * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, stream));
    // device copy of the geometries' first indices with bit 31 set for quad geometries
    std::vector<uint32_t> first_flagged(scene->geom_first.size());
    for (size_t g = 0; g < first_flagged.size(); ++g)
        first_flagged[g] = scene->geom_first[g] | (scene->geom_quad[g] ? 0x80000000u : 0u);
    SGPU_BUILD_TRY(hipMemcpy(d_first.p, first_flagged.data(), first_flagged.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice));
    SGPU_BUILD_TRY(hipMemsetAsync(d_bounds.p, 0xFF, 12 * sizeof(uint32_t), stream));
    SGPU_BUILD_TRY(hipMemsetAsync(d_flags.p, 0, (size_t)n_int * sizeof(uint32_t), stream));
    SGPU_BUILD_TRY(hipMemsetAsync(d_pint.p, 0xFF, (size_t)n_int * sizeof(int), stream));
    SGPU_BUILD_TRY(hipMemsetAsync(d_pleaf.p, 0xFF, (size_t)n * sizeof(int), stream));
    SGPU_BUILD_TRY(hipMemsetAsync(d_depth.p, 0, sizeof(uint32_t), stream));

    const uint32_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_prep, dim3(grid), dim3(kBlock), 0, stream, d_pos.p, n, d_geom.p, d_first.p, d_rec.p,
                       d_blo.p, d_bhi.p, d_cent.p, d_bounds.p);
    SGPU_BUILD_TRY(hipGetLastError());

    if (n == 1) {
        hipLaunchKernelGGL(k_single, dim3(1), dim3(1), 0, stream, d_blo.p, d_bhi.p, d_bounds.p, nodes);
        SGPU_BUILD_TRY(hipGetLastError());
        SGPU_BUILD_TRY(hipMemcpyAsync(tris, d_rec.p, sizeof(TriRecord), hipMemcpyDeviceToDevice, stream));
        SGPU_BUILD_TRY(hipStreamSynchronize(stream));
        scene->d_nodes = nodes;
        scene->d_tris = tris;
        scene->n_nodes = 1;
        scene->max_depth = 1;
        int rc = build_wide(scene, nodes, 1, stream);
        scene->stack_cap = std::max(scene->max_depth + 1, kWideMaxPush * scene->wide_depth + 1);
        return rc;
    }

    hipLaunchKernelGGL(k_morton, dim3(grid), dim3(kBlock), 0, stream, d_cent.p, n, d_bounds.p, d_keys.p, d_vals.p);
    SGPU_BUILD_TRY(hipGetLastError());

    size_t tmp_bytes = 0;
    SGPU_BUILD_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, d_keys.p, d_keys2.p, d_vals.p, d_vals2.p,
                                             (size_t)n, 0, 63, stream));
    SGPU_BUILD_TRY(d_tmp.alloc(tmp_bytes));
    SGPU_BUILD_TRY(rocprim::radix_sort_pairs(d_tmp.p, tmp_bytes, d_keys.p, d_keys2.p, d_vals.p, d_vals2.p,
                                             (size_t)n, 0, 63, stream));

    const uint32_t grid_int = (n - 1 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_karras, dim3(grid_int), dim3(kBlock), 0, stream, d_keys2.p, (int)n, d_child.p, d_pint.p,
                       d_pleaf.p);
    SGPU_BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_fit, dim3(grid), dim3(kBlock), 0, stream, (int)n, d_child.p, d_pint.p, d_pleaf.p, d_blo.p,
                       d_bhi.p, d_vals2.p, d_bounds.p, d_ilo.p, d_ihi.p, d_flags.p, nodes);
    SGPU_BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_reorder, dim3(grid), dim3(kBlock), 0, stream, d_rec.p, d_vals2.p, n, tris);
    SGPU_BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_depth, dim3(grid), dim3(kBlock), 0, stream, (int)n, d_pint.p, d_pleaf.p, d_depth.p);
    SGPU_BUILD_TRY(hipGetLastError());

    uint32_t depth = 0;
    SGPU_BUILD_TRY(hipMemcpyAsync(&depth, d_depth.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    SGPU_BUILD_TRY(hipStreamSynchronize(stream));
#undef SGPU_BUILD_TRY

    scene->d_nodes = nodes;
    scene->d_tris = tris;
    scene->n_nodes = n - 1;
    scene->max_depth = depth;
    int rc = build_wide(scene, nodes, n - 1, stream);
    scene->stack_cap = std::max(scene->max_depth + 1, kWideMaxPush * scene->wide_depth + 1);
    return rc;
}
