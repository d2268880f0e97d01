// Per-face texture lookup on the device: Ptex 2.x's separable cubic B-spline filter with a zero
// footprint, as CalculateSurfaceParams calls it for eUsesPtex materials
// (SurfaceParameters.cpp:196-203: PtexFilter{f_bspline}.eval(out, 0, 3, primId, u, v, 0, 0, 0, 0)).
// Restated from Middleware/ptex/Include/PtexSeparableFilter.cpp:45-330 (eval, splitAndApply,
// applyAcrossEdge, applyToCorner, applyToCornerFace, apply), PtexFilters.cpp:116-282
// (PtexWidth4Filter::buildKernelAxis, PtexBicubicFilter with sharpness 0) and
// PtexSeparableKernel.h:48-480 / .cpp:40-75, for the textures the renderer uses: quad meshes,
// 8-bit RGB, one resolution for every face (no subfaces), clamp border mode, no edge filter mode.
// The operations and their order are the reference's, so results agree with its compiled code bit
// for bit. A kernel that stays inside its face (most lookups) takes a direct 4 x 4 path that
// performs the same arithmetic as the general one.
#pragma once

#include "sgpu_math.h"

#include <cstring>

namespace sgpu {

// One texture: texels (nfaces x res x res x 3, row = v texel) and per face 8 words: the adjacent
// face across edges bottom, right, top, left (-1 = border), then the neighbour's edge id for each.
// A triangle-mesh texture (tri = 1) uses the first 3 of each: edges v = 0, w = 0, u = 0.
struct PtexTex {
    const uint8_t* texels;
    const int32_t* adj;
    uint32_t nfaces;
    int32_t res_log2;
    int32_t tri;
};

namespace ptex_detail {

// Widest kernel of a zero-footprint lookup: 6 taps on 4 x 4 faces (PtexWidth4Filter's even-aligned
// half-resolution branch), 4 at every other resolution. Every array below is indexed only by
// compile-time indices after unrolling (dynamic positions become selects), so that the filter
// lives in registers: a dynamically indexed array would go to per-lane scratch memory.
constexpr int kMax = 6;

struct Kernel {
    int res_log2;
    int u, v, uw, vw;
    float ku[kMax], kv[kMax];
    int rot;
};

struct Filter {
    PtexTex tx;
    float result[3];
    float weight;
};

// The value itself, opaque to the optimizer: a select among pinned values cannot be folded back
// into a load at a run-time address (which would put the array in scratch memory).
SGPU_HD float pin(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(x));
#endif
    return x;
}

// a[i] for a run-time i in [0, kMax)
SGPU_HD float get(const float* a, int i)
{
    float r = pin(a[0]);
#pragma unroll
    for (int j = 1; j < kMax; ++j) r = j == i ? pin(a[j]) : r;
    return r;
}
SGPU_HD void add_at(float* a, int i, float x)
{
#pragma unroll
    for (int j = 0; j < kMax; ++j) {
        const float y = pin(a[j]);
        a[j] = j == i ? y + x : y;
    }
}
// a[lo] + ... + a[lo + n - 1], summed in index order
SGPU_HD float accumulate(const float* a, int lo, int n)
{
    float r = 0;
#pragma unroll
    for (int j = 0; j < kMax; ++j)
        if (j >= lo && j < lo + n) r += a[j];
    return r;
}
SGPU_HD float kweight(const Kernel& k) { return accumulate(k.ku, 0, k.uw) * accumulate(k.kv, 0, k.vw); }
SGPU_HD int kres(const Kernel& k) { return 1 << k.res_log2; }
// a[i] = a[i + w] for i + w < n (the rest is left as it was)
SGPU_HD void drop_front(float* a, int n, int w)
{
    float b[kMax];
#pragma unroll
    for (int i = 0; i < kMax; ++i) b[i] = pin(a[i]);
#pragma unroll
    for (int i = 0; i < kMax; ++i) {
        float x = b[i];
#pragma unroll
        for (int s = 1; i + s < kMax; ++s) x = s == w ? b[i + s] : x;
        a[i] = i + w < n ? x : b[i];
    }
}
// dst[i] = src[i + off] for i + off < kMax
SGPU_HD void copy_from(float* dst, const float* src, int off)
{
    float b[kMax];
#pragma unroll
    for (int i = 0; i < kMax; ++i) b[i] = pin(src[i]);
#pragma unroll
    for (int i = 0; i < kMax; ++i) {
        float x = b[i];
#pragma unroll
        for (int s = 1; i + s < kMax; ++s) x = s == off ? b[i + s] : x;
        dst[i] = x;
    }
}

SGPU_HD void merge_l(Kernel& k)
{
    int w = -k.u;
    add_at(k.ku, w, accumulate(k.ku, 0, w));
    drop_front(k.ku, k.uw, w);
    k.uw -= w;
    k.u = 0;
}
SGPU_HD void merge_r(Kernel& k)
{
    int w = k.uw + k.u - kres(k);
    int at = k.uw - w;
    add_at(k.ku, at - 1, accumulate(k.ku, at, w));
    k.uw -= w;
}
SGPU_HD void merge_b(Kernel& k)
{
    int w = -k.v;
    add_at(k.kv, w, accumulate(k.kv, 0, w));
    drop_front(k.kv, k.vw, w);
    k.vw -= w;
    k.v = 0;
}
SGPU_HD void merge_t(Kernel& k)
{
    int w = k.vw + k.v - kres(k);
    int at = k.vw - w;
    add_at(k.kv, at - 1, accumulate(k.kv, at, w));
    k.vw -= w;
}

// d = a kernel of the given placement whose taps are ku[ku_off ..] and kv[kv_off ..]
SGPU_HD void kset(Kernel& d, int res_log2, int u, int v, int uw, int vw, const float* ku, int ku_off, const float* kv,
                  int kv_off)
{
    float tu[kMax], tv[kMax];
    copy_from(tu, ku, ku_off);
    copy_from(tv, kv, kv_off);
    d.res_log2 = res_log2;
    d.u = u;
    d.v = v;
    d.uw = uw;
    d.vw = vw;
#pragma unroll
    for (int i = 0; i < kMax; ++i) {
        d.ku[i] = tu[i];
        d.kv[i] = tv[i];
    }
    d.rot = 0;
}

SGPU_HD void split_l(Kernel& k, Kernel& out)
{
    int w = -k.u;
    if (w < k.uw) {
        kset(out, k.res_log2, kres(k) - w, k.v, w, k.vw, k.ku, 0, k.kv, 0);
        k.u = 0;
        drop_front(k.ku, k.uw, w);
        k.uw -= w;
    } else {
        out = k;
        out.u += kres(k);
        k.u = 0;
        k.uw = 0;
    }
}
SGPU_HD void split_r(Kernel& k, Kernel& out)
{
    int w = k.u + k.uw - kres(k);
    if (w < k.uw) {
        kset(out, k.res_log2, 0, k.v, w, k.vw, k.ku, k.uw - w, k.kv, 0);
        k.uw -= w;
    } else {
        out = k;
        out.u -= kres(k);
        k.u = 0;
        k.uw = 0;
    }
}
SGPU_HD void split_b(Kernel& k, Kernel& out)
{
    int w = -k.v;
    if (w < k.vw) {
        kset(out, k.res_log2, k.u, kres(k) - w, k.uw, w, k.ku, 0, k.kv, 0);
        k.v = 0;
        drop_front(k.kv, k.vw, w);
        k.vw -= w;
    } else {
        out = k;
        out.v += kres(k);
        k.v = 0;
        k.vw = 0;
    }
}
SGPU_HD void split_t(Kernel& k, Kernel& out)
{
    int w = k.v + k.vw - kres(k);
    if (w < k.vw) {
        kset(out, k.res_log2, k.u, 0, k.uw, w, k.ku, 0, k.kv, k.vw - w);
        k.vw -= w;
    } else {
        out = k;
        out.v -= kres(k);
        k.v = 0;
        k.vw = 0;
    }
}

// the first n taps in reverse order
SGPU_HD void reverse(float* a, int n)
{
    float b[kMax];
#pragma unroll
    for (int i = 0; i < kMax; ++i) b[i] = pin(a[i]);
#pragma unroll
    for (int i = 0; i < kMax; ++i) {
        float x = b[i];
#pragma unroll
        for (int j = 0; j < kMax; ++j) x = (i < n && j == n - 1 - i) ? b[j] : x;
        a[i] = x;
    }
}
SGPU_HD void flip_u(Kernel& k)
{
    k.u = kres(k) - k.u - k.uw;
    reverse(k.ku, k.uw);
}
SGPU_HD void flip_v(Kernel& k)
{
    k.v = kres(k) - k.v - k.vw;
    reverse(k.kv, k.vw);
}
SGPU_HD void swap_uv(Kernel& k)
{
    int t = k.u;
    k.u = k.v;
    k.v = t;
    t = k.uw;
    k.uw = k.vw;
    k.vw = t;
#pragma unroll
    for (int i = 0; i < kMax; ++i) {
        float f = k.ku[i];
        k.ku[i] = k.kv[i];
        k.kv[i] = f;
    }
}
// rotate `r` steps counter-clockwise
SGPU_HD void krotate(Kernel& k, int r)
{
    switch (r & 3) {
    default: return;
    case 1:
        flip_u(k);
        swap_uv(k);
        break;
    case 2:
        flip_u(k);
        flip_v(k);
        break;
    case 3:
        flip_v(k);
        swap_uv(k);
        break;
    }
    k.rot = (k.rot + r) & 3;
}

SGPU_HD float make_symmetric(Kernel& k, float initial_weight)
{
    int n = k.uw < k.vw ? k.uw : k.vw;
    k.uw = k.vw = n;
    float new_weight = 0;
#pragma unroll
    for (int i = 0; i < kMax; ++i) {
        if (i < n) {
            float sum = k.ku[i] + k.kv[i];
            k.ku[i] = k.kv[i] = sum;
            new_weight += sum;
        }
    }
    new_weight *= new_weight;
    float scale = new_weight == 0 ? 1.f : initial_weight / new_weight;
    if (!(scale >= 1)) {
        if (scale < -1) {
#pragma unroll
            for (int i = 0; i < kMax; ++i)
                if (i < n) k.ku[i] *= -1;
            new_weight = -new_weight;
        } else {
#pragma unroll
            for (int i = 0; i < kMax; ++i)
                if (i < n) k.ku[i] *= scale;
            new_weight = initial_weight;
        }
    }
    return new_weight;
}

SGPU_HD int adjface(const PtexTex& tx, int f, int e) { return tx.adj[8 * f + e]; }
SGPU_HD int adjedge(const PtexTex& tx, int f, int e) { return tx.adj[8 * f + 4 + e]; }

// Apply<uint8_t, 3>: row sums over u first, then weighted by v
SGPU_HD void kapply(Filter& F, const Kernel& k, int face)
{
    if (k.uw <= 0 || k.vw <= 0) return;
    const int res = kres(k);
    const uint8_t* data = F.tx.texels + (size_t)face * res * res * 3;
#pragma unroll
    for (int j = 0; j < kMax; ++j) {
        if (j >= k.vw) break;
        const uint8_t* p = data + ((size_t)(k.v + j) * res + k.u) * 3;
        float r0 = (float)p[0] * k.ku[0], r1 = (float)p[1] * k.ku[0], r2 = (float)p[2] * k.ku[0];
#pragma unroll
        for (int i = 1; i < kMax; ++i) {
            if (i < k.uw) {
                r0 += (float)p[3 * i] * k.ku[i];
                r1 += (float)p[3 * i + 1] * k.ku[i];
                r2 += (float)p[3 * i + 2] * k.ku[i];
            }
        }
        F.result[0] += r0 * k.kv[j];
        F.result[1] += r1 * k.kv[j];
        F.result[2] += r2 * k.kv[j];
    }
}

SGPU_HD void apply_across_edge(Filter& F, Kernel& k, int face, int eid)
{
    int af = adjface(F.tx, face, eid), ae = adjedge(F.tx, face, eid);
    krotate(k, eid - ae + 2);
    kapply(F, k, af);
}

// One step around a corner: the face across edge e of face f and the edge that follows.
SGPU_HD void corner_step(const PtexTex& tx, int& f, int& e)
{
    const int next = adjface(tx, f, e);
    const int nexte = (adjedge(tx, f, e) + 1) % 4;
    f = next;
    e = nexte;
}

SGPU_HD void apply_to_corner(Filter& F, Kernel& k, int face, int eid)
{
    // the faces around the corner (the reference records them; here the ring is walked again
    // where they are needed, so that no array is indexed at run time)
    constexpr int kMaxValence = 10;
    int af = face, ae = eid, ncorners = 0, face1 = 0, edge1 = 0;
    for (int i = 0; i < kMaxValence; ++i) {
        corner_step(F.tx, af, ae);
        if (af < 0 || (af == face && ae == eid)) {
            ncorners = i - 2;
            break;
        }
        if (i == 1) {
            face1 = af;
            edge1 = ae;
        }
    }
    if (ncorners == 1) {
        krotate(k, eid - edge1 + 2);
        kapply(F, k, face1);
    } else if (ncorners > 1) {
        krotate(k, eid + 2);
        float initial = kweight(k);
        float neww = make_symmetric(k, initial);
        af = face;
        ae = eid;
        corner_step(F.tx, af, ae);  // ring position 0
        for (int i = 1; i <= ncorners; ++i) {
            corner_step(F.tx, af, ae);
            Kernel kc = k;
            krotate(kc, 2 - ae + 2);
            kapply(F, kc, af);
        }
        F.weight += neww * (float)ncorners - initial;
    } else {
        F.weight -= kweight(k);
    }
}

SGPU_HD void split_and_apply(Filter& F, Kernel& k, int face)
{
    const PtexTex& tx = F.tx;
    int res = kres(k);
    bool sr = k.u + k.uw > res, sl = k.u < 0, st = k.v + k.vw > res, sb = k.v < 0;
    if (sr || sl || st || sb) {
        Kernel ka, kc;
        if (sr) {
            if (adjface(tx, face, 1) >= 0) {
                split_r(k, ka);
                if (st) {
                    if (adjface(tx, face, 2) >= 0) {
                        split_t(ka, kc);
                        apply_to_corner(F, kc, face, 2);
                    } else
                        merge_t(ka);
                }
                if (sb) {
                    if (adjface(tx, face, 0) >= 0) {
                        split_b(ka, kc);
                        apply_to_corner(F, kc, face, 1);
                    } else
                        merge_b(ka);
                }
                apply_across_edge(F, ka, face, 1);
            } else
                merge_r(k);
        }
        if (sl) {
            if (adjface(tx, face, 3) >= 0) {
                split_l(k, ka);
                if (st) {
                    if (adjface(tx, face, 2) >= 0) {
                        split_t(ka, kc);
                        apply_to_corner(F, kc, face, 3);
                    } else
                        merge_t(ka);
                }
                if (sb) {
                    if (adjface(tx, face, 0) >= 0) {
                        split_b(ka, kc);
                        apply_to_corner(F, kc, face, 0);
                    } else
                        merge_b(ka);
                }
                apply_across_edge(F, ka, face, 3);
            } else
                merge_l(k);
        }
        if (st) {
            if (adjface(tx, face, 2) >= 0) {
                split_t(k, ka);
                apply_across_edge(F, ka, face, 2);
            } else
                merge_t(k);
        }
        if (sb) {
            if (adjface(tx, face, 0) >= 0) {
                split_b(k, ka);
                apply_across_edge(F, ka, face, 0);
            } else
                merge_b(k);
        }
    }
    kapply(F, k, face);
}

// PtexBicubicFilter::kernelFn with sharpness 0 (B = 1, C = 0: the cubic B-spline)
SGPU_HD float bspline(float x)
{
    const float B = 1.0f;
    const float c0 = 1.5f - B, c1 = 1.5f * B - 2.5f, c2 = 1.0f - (float)(1.0 / 3.0) * B;
    const float c3 = (float)(1.0 / 3.0) * B - 0.5f, c4 = 2.5f - 1.5f * B, c5 = 2.0f * B - 4.0f;
    const float c6 = 2.0f - (float)(2.0 / 3.0) * B;
    x = fabsf(x);
    if (x < 1.0f) return (c0 * x + c1) * x * x + c2;
    if (x < 2.0f) return ((c3 * x + c4) * x + c5) * x + c6;
    return 0.0f;
}

// 2-unit cubic Hermite used for wide footprints
SGPU_HD float blur(float x)
{
    x = fabsf(x);
    return x < 1.0f ? (2.0f * x - 3.0f) * x * x + 1.0f : 0.0f;
}

SGPU_HD float recip_pow2(int power)
{
    uint32_t b = (uint32_t)(127 - power) << 23;
    float f;
    memcpy(&f, &b, 4);
    return f;
}
SGPU_HD int res_from_width(float w)
{
    int32_t b;
    memcpy(&b, &w, 4);
    return 127 - ((b >> 23) & 0xff);
}

// PtexWidth4Filter::buildKernelAxis with lerp off, for a zero footprint (uw is at least one texel:
// the generic branch then has exactly 4 taps, the half-resolution branch at most 6)
SGPU_HD void build_axis(int& k_reslog2, int& k_u, int& k_uw, float* ku, float u, float uw, int f_reslog2)
{
    uw = uw > recip_pow2(f_reslog2) ? uw : recip_pow2(f_reslog2);
    k_reslog2 = res_from_width(uw);
    const int resu = 1 << k_reslog2;
    const float lerp2 = 0, lerp1 = 1.0f - lerp2;
    if (uw >= .25f) {
        if (uw < .5f) {
            k_reslog2 = 2;
            float upix = u * 4.0f - 0.5f;
            int u1 = (int)ceilf(upix - 2), u2 = (int)ceilf(upix + 2);
            u1 = u1 & ~1;
            u2 = (u2 + 1) & ~1;
            k_u = u1;
            k_uw = u2 - u1;
            float x1 = (float)u1 - upix;
#pragma unroll
            for (int i = 0; i < kMax; i += 2) {
                if (i >= k_uw) break;
                float xa = x1 + (float)i, xb = xa + 1.0f, xc = (xa + xb) * 0.25f;
                float s = 1.0f / (uw + .75f);
                float ka = bspline(xa), kb = bspline(xb), kc = blur(xc * s);
                ku[i] = ka * lerp1 + kc * lerp2;
                ku[i + 1] = kb * lerp1 + kc * lerp2;
            }
            return;
        } else if (uw < 1) {
            k_reslog2 = 1;
            float upix = u * 2.0f - 0.5f;
            k_u = (int)floorf(u - .5f) * 2;
            k_uw = 4;
            float x1 = (float)k_u - upix;
#pragma unroll
            for (int i = 0; i < 4; i += 2) {
                float xa = x1 + (float)i, xb = xa + 1.0f, xc = (xa + xb) * 0.5f;
                float s = 1.0f / (uw * 1.5f + .5f);
                float ka = blur(xa * s), kb = blur(xb * s), kc = blur(xc * s);
                ku[i] = ka * lerp1 + kc * lerp2;
                ku[i + 1] = kb * lerp1 + kc * lerp2;
            }
            return;
        } else {
            k_reslog2 = 0;
            float upix = u - .5f;
            k_uw = 2;
            float ui = floorf(upix);
            k_u = (int)ui;
            ku[0] = blur(upix - ui);
            ku[1] = 1 - ku[0];
            return;
        }
    }
    float upix = u * (float)resu - 0.5f;
    float uwpix = uw * (float)resu;
    float dupix = 2.0f * uwpix;
    int u1 = (int)ceilf(upix - dupix), u2 = (int)ceilf(upix + dupix);
    k_u = u1;
    k_uw = u2 - u1;
    float x1 = ((float)u1 - upix) / uwpix, step = 1.0f / uwpix;
#pragma unroll
    for (int i = 0; i < kMax; ++i)
        if (i < k_uw) ku[i] = bspline(x1 + (float)i * step);
}

}  // namespace ptex_detail

// ---- triangle meshes ------------------------------------------------------------------------------
// PtexFilter::getFilter serves triangle-mesh textures with PtexTriangleFilter whatever the filter type
// (PtexFilters.cpp:436-440): an elliptical Gaussian over the triangular texel grid. Restated from
// PtexTriangleFilter.cpp:45-265 (eval, buildKernel, splitAndApply, applyAcrossEdge, apply,
// applyIter), PtexTriangleKernel.h:48-231 and PtexTriangleKernel.cpp:42-75 (Apply for 3 channels of
// 8-bit data) for Selas's zero footprint and one resolution for every face. The reference recurses:
// splitAndApply(k, f) clips k against each edge it crosses (u, then v, then w), applies the piece
// beyond the edge to the neighbour (splitAndApply again, in the neighbour's coordinates) and then
// applies what is left to f, all summed into one running total. Here the same pieces are visited in
// the same order without a stack: a piece is named by its path of edge slots from the first face
// (2 bits a level), and a piece's kernel is rebuilt from the first one by replaying its path, so the
// walk keeps only the current piece in registers. Splits nest at most 3 deep on the meshes tested
// (closed and open, vertex valences 3 to 16, 1 x 1 to 32 x 32 texels); the walk allows 8.
// The Gaussian is exp in double rounded once to float: the reference's expf agrees except where it
// is not correctly rounded itself (about 1 call in 1,500; the filtered value then moves by an ulp or
// two).
namespace ptex_tri_detail {

constexpr float kWidth = 3.5f;  // PtexTriangleKernelWidth
constexpr int kMaxDepth = 8;

struct TriKernel {
    int res_log2;
    float u, v, u1, v1, w1, u2, v2, w2, A, B, C;
};

// edge slot s of splitAndApply: u1 < 0 crosses edge 2 (u = 0), v1 < 0 edge 0, w1 < 0 edge 1. The
// field a slot names is chosen by selects over pinned values and every field is written: a field
// picked by a runtime index would otherwise become an address offset, and the kernel would live in
// per-lane scratch memory instead of registers.
SGPU_HD int slot_edge(int s) { return s == 0 ? 2 : (s == 1 ? 0 : 1); }
SGPU_HD float lower(const TriKernel& k, int s)
{
    const float a = ptex_detail::pin(k.u1), b = ptex_detail::pin(k.v1), c = ptex_detail::pin(k.w1);
    return s == 0 ? a : (s == 1 ? b : c);
}
SGPU_HD void clear_lower(TriKernel& k, int s)
{
    k.u1 = s == 0 ? 0.0f : ptex_detail::pin(k.u1);
    k.v1 = s == 1 ? 0.0f : ptex_detail::pin(k.v1);
    k.w1 = s == 2 ? 0.0f : ptex_detail::pin(k.w1);
}
SGPU_HD void clear_upper(TriKernel& k, int s)
{
    k.u2 = s == 0 ? 0.0f : ptex_detail::pin(k.u2);
    k.v2 = s == 1 ? 0.0f : ptex_detail::pin(k.v2);
    k.w2 = s == 2 ? 0.0f : ptex_detail::pin(k.w2);
}
SGPU_HD bool taken(const PtexTex& tx, const TriKernel& k, int face, int s)
{
    return lower(k, s) < 0.0f && tx.adj[8 * face + slot_edge(s)] >= 0;
}

SGPU_HD void set8(TriKernel& k, float u, float v, float u1, float v1, float w1, float u2, float v2, float w2)
{
    k.u = u;
    k.v = v;
    k.u1 = u1;
    k.v1 = v1;
    k.w1 = w1;
    k.u2 = u2;
    k.v2 = v2;
    k.w2 = w2;
}

// PtexTriangleKernel::reorient: the kernel in the coordinates of the face across edge eid, entered
// through its edge aeid. The reference's nine cases share one form: the neighbour's (u, v, w) are a
// rotation of (u, v, w) by r = (aeid - eid) mod 3 (r = 1: (w, u, v), r = 2: (v, w, u)), each
// coordinate x becoming 1 - x except the one at position (aeid + 1) mod 3, which becomes -x; a
// bound's lower and upper ends swap. Straight-line selects here (the same float operations as each
// case of the switch), with A, B, C rotated as rotate1 / rotate2 do.
SGPU_HD float flip(float x, bool neg) { return neg ? -x : 1.0f - x; }
SGPU_HD float rot3(float a, float b, float c, int r) { return r == 0 ? a : (r == 1 ? c : b); }
SGPU_HD void reorient(TriKernel& k, int eid, int aeid)
{
    const float u = k.u, v = k.v, w = 1.0f - k.u - k.v;
    const float u1 = k.u1, v1 = k.v1, w1 = k.w1, u2 = k.u2, v2 = k.v2, w2 = k.w2;
    const float A = k.A, B = k.B, C = k.C;
    const int r = (aeid - eid + 3) % 3;
    const int z = (aeid + 1) % 3;
    // position p of the rotated triple takes element (p - r) mod 3 of the original
    const float x0 = rot3(u, v, w, r), x1 = rot3(v, w, u, r);
    const float l0 = rot3(u1, v1, w1, r), l1 = rot3(v1, w1, u1, r), l2 = rot3(w1, u1, v1, r);
    const float h0 = rot3(u2, v2, w2, r), h1 = rot3(v2, w2, u2, r), h2 = rot3(w2, u2, v2, r);
    set8(k, flip(x0, z == 0), flip(x1, z == 1), flip(h0, z == 0), flip(h1, z == 1), flip(h2, z == 2),
         flip(l0, z == 0), flip(l1, z == 1), flip(l2, z == 2));
    const float AC_B = A + C - B;
    k.A = r == 0 ? A : (r == 1 ? C : AC_B);
    k.B = r == 0 ? B : (r == 1 ? 2.0f * C - B : 2.0f * A - B);
    k.C = r == 0 ? C : (r == 1 ? AC_B : A);
}

// the piece of (k, face) beyond slot s's edge, in the neighbour's coordinates: k with the lower
// bounds of the slots taken before s cleared, its own upper bound cleared, then reoriented
SGPU_HD void child(const PtexTex& tx, const TriKernel& k, int face, int s, TriKernel& c, int& cface)
{
    c = k;
#pragma unroll
    for (int t = 0; t < 2; ++t)
        if (t < s && taken(tx, k, face, t)) clear_lower(c, t);
    clear_upper(c, s);
    const int e = slot_edge(s);
    cface = tx.adj[8 * face + e];
    reorient(c, e, tx.adj[8 * face + 4 + e]);
}

SGPU_HD float gaussian(float q)
{
    const float scale = -0.5f * (kWidth * kWidth);
    return (float)exp((double)(scale * q));
}

struct Iter {
    int rowlen;
    float u, v;
    int u1, v1, w1, u2, v2, w2;
    float A, B, C;
    bool valid;
    float wscale;
};

// PtexTriangleKernel::getIterators: the even iterator (odd = false) or the odd one, flipped across
// the diagonal (odd = true). Built one at a time, so that only one is live while it is applied.
SGPU_HD void iterator(const TriKernel& k, bool odd, Iter& it)
{
    const int resu = 1 << k.res_log2;
    const float third = (float)(1 / 3.0);
    const float Finv = 1.0f / ((float)resu * (float)resu * (k.A * k.C - 0.25f * k.B * k.B));
    const float Ak = k.A * Finv, Bk = k.B * Finv, Ck = k.C * Finv;
    const float scale = (float)resu;
    it.rowlen = resu;
    it.wscale = 1.0f / ((float)resu * (float)resu);
    if (!odd) {
        it.u = k.u * scale - third;
        it.v = k.v * scale - third;
        it.u1 = (int)ceilf(k.u1 * scale - third);
        it.v1 = (int)ceilf(k.v1 * scale - third);
        it.w1 = (int)ceilf(k.w1 * scale - third);
        it.u2 = (int)ceilf(k.u2 * scale - third);
        it.v2 = (int)ceilf(k.v2 * scale - third);
        it.w2 = (int)ceilf(k.w2 * scale - third);
        it.A = Ak;
        it.C = Ck;
    } else {
        it.u = (1.0f - k.v) * scale - third;
        it.v = (1.0f - k.u) * scale - third;
        it.u1 = (int)ceilf((1.0f - k.v2) * scale - third);
        it.v1 = (int)ceilf((1.0f - k.u2) * scale - third);
        it.w1 = (int)ceilf((-k.w2) * scale - third);
        it.u2 = (int)ceilf((1.0f - k.v1) * scale - third);
        it.v2 = (int)ceilf((1.0f - k.u1) * scale - third);
        it.w2 = (int)ceilf((-k.w1) * scale - third);
        it.A = Ck;
        it.C = Ak;
    }
    it.B = Bk;
    it.valid = it.u2 > it.u1 && it.v2 > it.v1 && it.w2 > it.w1;
}

// Apply<uint8_t, 3>: Q of the ellipse by finite differences along each row (Heckbert); returns the
// iterator's summed weight
SGPU_HD float apply_iter(const Iter& k, const uint8_t* data, f3& result)
{
    float weight_sum = 0.0f;
    const float DDQ = 2.0f * k.A;
    for (int vi = k.v1; vi != k.v2; vi++) {
        const int xw = k.rowlen - vi;
        const int x1 = k.u1 > xw - k.w2 ? k.u1 : xw - k.w2;
        const int x2 = k.u2 < xw - k.w1 ? k.u2 : xw - k.w1;
        const float U = (float)x1 - k.u;
        const float V = (float)vi - k.v;
        float DQ = k.A * (2.0f * U + 1.0f) + k.B * V;
        float Q = k.A * U * U + (k.B * U + k.C * V) * V;
        const uint8_t* p = data + ((size_t)vi * k.rowlen + x1) * 3;
        for (int x = x1; x < x2; ++x, p += 3) {
            if (Q < 1.0f) {
                const float weight = gaussian(Q) * k.wscale;
                weight_sum += weight;
                result.x += (float)p[0] * weight;
                result.y += (float)p[1] * weight;
                result.z += (float)p[2] * weight;
            }
            Q += DQ;
            DQ += DDQ;
        }
    }
    return weight_sum;
}

// PtexTriangleFilter::apply: clampRes (one resolution: the face's), clampExtent, both iterators
SGPU_HD void apply(const PtexTex& tx, TriKernel k, int face, f3& result, float& weight)
{
    if (k.res_log2 > tx.res_log2) k.res_log2 = tx.res_log2;
    k.u1 = k.u1 > 0.0f ? k.u1 : 0.0f;
    k.v1 = k.v1 > 0.0f ? k.v1 : 0.0f;
    k.w1 = k.w1 > 0.0f ? k.w1 : 0.0f;
    const float cu = 1.0f - (k.v1 + k.w1), cv = 1.0f - (k.w1 + k.u1), cw = 1.0f - (k.u1 + k.v1);
    k.u2 = k.u2 < cu ? k.u2 : cu;
    k.v2 = k.v2 < cv ? k.v2 : cv;
    k.w2 = k.w2 < cw ? k.w2 : cw;
    const size_t res = (size_t)1 << tx.res_log2;
    const uint8_t* data = tx.texels + (size_t)face * res * res * 3;
    Iter it;
    iterator(k, false, it);
    if (it.valid) weight += apply_iter(it, data, result);
    iterator(k, true, it);
    if (it.valid) weight += apply_iter(it, data, result);
}

// the piece named by the first `depth` slots of `path` (2 bits a level), replayed from the root
SGPU_HD void replay(const PtexTex& tx, const TriKernel& root, int root_face, uint32_t path, int depth, TriKernel& k,
                    int& face)
{
    k = root;
    face = root_face;
    for (int l = 0; l < depth; ++l) {
        TriKernel c;
        int cf;
        child(tx, k, face, (int)((path >> (2 * l)) & 3u), c, cf);
        k = c;
        face = cf;
    }
}

// the first slot from s on that (k, face) splits off, or 3
SGPU_HD int next_slot(const PtexTex& tx, const TriKernel& k, int face, int s)
{
    for (; s < 3; ++s)
        if (taken(tx, k, face, s)) return s;
    return 3;
}

// from piece (k, f) at `depth`, down its first split-off pieces to one that splits nothing off
SGPU_HD void descend(const PtexTex& tx, TriKernel& k, int& f, uint32_t& path, int& depth)
{
    for (int s = next_slot(tx, k, f, 0); s < 3 && depth < kMaxDepth; s = next_slot(tx, k, f, 0)) {
        TriKernel c;
        int cf;
        child(tx, k, f, s, c, cf);
        path = (path & ~(3u << (2 * depth))) | ((uint32_t)s << (2 * depth));
        ++depth;
        k = c;
        f = cf;
    }
}

}  // namespace ptex_tri_detail

// PtexTriangleFilter::eval with buildKernel for uw1 = vw1 = uw2 = vw2 = 0, width 1, blur 0.
SGPU_HD f3 ptex_tri_eval(const PtexTex& tx, int face, float u, float v)
{
    using namespace ptex_tri_detail;
    if (face < 0 || (uint32_t)face >= tx.nfaces) return mk3(0, 0, 0);
    u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    const float sqrt3 = 1.7320508075688772f;
    const float width = 1.0f, blur = 0.0f, uw1 = 0.0f, vw1 = 0.0f, uw2 = 0.0f, vw2 = 0.0f;
    const float scaleAC = 0.25f * width * width;
    const float scaleB = -2.0f * scaleAC;
    float A = (vw1 * vw1 + vw2 * vw2) * scaleAC;
    float B = (uw1 * vw1 + uw2 * vw2) * scaleB;
    float C = (uw1 * uw1 + uw2 * uw2) * scaleAC;
    float Ac = 0.75f * A;
    const float Bc = (sqrt3 / 2) * (B - A);
    float Cc = 0.25f * A - 0.5f * B + C;
    const float maxEcc = 15.0f;
    const float eccRatio = (maxEcc * maxEcc + 1.0f) / (maxEcc * maxEcc - 1.0f);
    const float dAC = Ac - Cc;
    const float X = sqrtf(dAC * dAC + Bc * Bc);
    const float b_e = 0.5f * (eccRatio * X - (Ac + Cc));
    const float hr = 0.5f / (float)(1 << tx.res_log2);
    const float b_t = hr * hr;
    const float b_b = 0.25f * blur * blur;
    const float b_et = b_e > b_t ? b_e : b_t;
    const float b = b_b > b_et ? b_b : b_et;
    Ac += b;
    Cc += b;
    const float m = sqrtf(2.0f * (Ac * Cc - 0.25f * Bc * Bc) / (Ac + Cc + X));
    const float width2 = 2.0f * m;
    uint32_t wbits;
    memcpy(&wbits, &width2, sizeof(wbits));
    int reslog2 = 127 - (int)((wbits >> 23) & 0xffu);  // calcResFromWidth: the exponent of 2m
    if (reslog2 < 0) reslog2 = 0;
    A = (float)(4 / 3.0) * Ac;
    B = (2 / sqrt3) * Bc + A;
    C = -0.25f * A + 0.5f * B + Cc;
    const float scale = kWidth * kWidth;
    A *= scale;
    B *= scale;
    C *= scale;
    const float sC = sqrtf(C), sA = sqrtf(A), sW = sqrtf(A - B + C);
    const float uw = sC < 1.0f ? sC : 1.0f;
    const float vw = sA < 1.0f ? sA : 1.0f;
    const float ww = sW < 1.0f ? sW : 1.0f;
    const float w = 1.0f - u - v;
    TriKernel root;
    root.res_log2 = reslog2;
    set8(root, u, v, u - uw, v - vw, w - ww, u + uw, v + vw, w + ww);
    root.A = A;
    root.B = B;
    root.C = C;

    // splitAndApply's depth-first order: each piece's pieces beyond its edges (slots in order), then
    // the piece itself with those parts clipped off
    f3 result = mk3(0.0f, 0.0f, 0.0f);
    float weight = 0.0f;
    uint32_t path = 0;
    int depth = 0;
    TriKernel k = root;
    int f = face;
    descend(tx, k, f, path, depth);
    while (true) {
        TriKernel a = k;  // the piece with every slot it split off cleared
#pragma unroll
        for (int t = 0; t < 3; ++t)
            if (taken(tx, k, f, t)) clear_lower(a, t);
        apply(tx, a, f, result, weight);
        if (depth == 0) break;
        --depth;
        const int s = (int)((path >> (2 * depth)) & 3u);
        replay(tx, root, face, path, depth, k, f);  // the parent
        const int s2 = next_slot(tx, k, f, s + 1);
        if (s2 < 3) {
            TriKernel c;
            int cf;
            child(tx, k, f, s2, c, cf);
            path = (path & ~(3u << (2 * depth))) | ((uint32_t)s2 << (2 * depth));
            ++depth;
            k = c;
            f = cf;
            descend(tx, k, f, path, depth);
        }
    }
    const float sc = 1.0f / (weight * 255.0f);
    return mk3(result.x * sc, result.y * sc, result.z * sc);
}

// The filtered texel (channels 0..2 scaled to [0, 1]) of face `face` at (u, v).
SGPU_HD f3 ptex_eval(const PtexTex& tx, int face, float u, float v)
{
    using namespace ptex_detail;
    if (tx.tri) return ptex_tri_eval(tx, face, u, v);
    if (face < 0 || (uint32_t)face >= tx.nfaces) return mk3(0, 0, 0);
    u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    Kernel k;
    int ru, rv;
    build_axis(ru, k.u, k.uw, k.ku, u, 0.0f, tx.res_log2);
    build_axis(rv, k.v, k.vw, k.kv, v, 0.0f, tx.res_log2);
    k.res_log2 = ru;  // square faces: both axes choose the same resolution
    k.rot = 0;
    // stripZeros
    while (get(k.ku, 0) == 0) {
        drop_front(k.ku, k.uw, 1);
        k.u++;
        k.uw--;
    }
    while (get(k.ku, k.uw - 1) == 0) k.uw--;
    while (get(k.kv, 0) == 0) {
        drop_front(k.kv, k.vw, 1);
        k.v++;
        k.vw--;
    }
    while (get(k.kv, k.vw - 1) == 0) k.vw--;
    Filter F;
    F.tx = tx;
    F.result[0] = F.result[1] = F.result[2] = 0.0f;
    F.weight = kweight(k);
    split_and_apply(F, k, face);
    float scale = 1.0f / (F.weight * 255.0f);
    return mk3(F.result[0] * scale, F.result[1] * scale, F.result[2] * scale);
}

}  // namespace sgpu
