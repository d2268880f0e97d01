// Vertex merging (the camera side of VCM's photon range search), in a translation unit of its own.
// Measured on MI355X: building it with fast FP32 division and square root cuts its static
// instructions by a quarter but needs more registers (179: two waves per SIMD instead of three),
// and held to three waves it runs no faster (19.8 vs 19.9 ms on the 512 x 512 x 32 job); it is
// built with correctly rounded arithmetic like the rest of the path.
#include "sgpu_vcm_common.h"

namespace sgpu_vcm {

namespace {

// Vertex merging (SearchHashGrid + MergeVertices, VCMHashGrid.cpp:107-157, VCM.cpp:362-409), one
// query per group of GW lanes (64 / GW queries per wave), the queries of a bounce taken grid-stride.
// The lanes of a group scan the photons of their query's 2x2x2 cells as one list (the cells' ranges
// concatenated in the reference's visiting order, strided over the lanes), append those within the
// radius to the group's LDS list (ballot and prefix count, list order = scan order) and evaluate
// the camera BSDF for up to GW listed photons at a time, one per lane, every group in the same
// pass: an evaluation pass serves 64 / GW queries, and the acceptance test does not make whole
// waves wait on every BSDF evaluation. With about 14 photons within the radius per query on the
// 1M-triangle workload, 16-lane groups keep most lanes of an evaluation pass busy where 32-lane
// groups leave more than half of them idle. Each lane sums its photons in list order; a fixed
// butterfly over the group's lanes gives the query's sum. With `order` (the queries of the bounce
// in spatial order) blocks are numbered xcd-major, so that the queries running at once on one XCD
// are neighbours in space and read the same photon cells through that XCD's L2.
template <uint32_t GW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_cam_merge(Wave w, const uint32_t* __restrict count_ptr,
                                                      const uint32_t* __restrict order,
                                                      const CamVertex* __restrict cam_vm,
                                                      const uint32_t* __restrict path_slice, GridView G, CamSlots S,
                                                      unsigned long long* __restrict mstat)
{
    static_assert(GW == 8 || GW == 16 || GW == 32, "a group covers the 8 cells and divides the wave");
    constexpr uint32_t kQ = 64 / GW;  // queries per wave
    const uint32_t count = *count_ptr;
    __shared__ uint32_t list_lds[kMergeWaves][kQ][2 * GW];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t grp = lane / GW, gl = lane % GW;  // group and lane within it
    uint32_t* list = list_lds[wid][grp];
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long group_bits = ((GW == 64 ? ~0ull : ((1ull << GW) - 1ull))) << (grp * GW);
    if (*G.nv_ptr == 0) return;
    const uint32_t stride = kQ * gridDim.x * kMergeWaves;
    const uint32_t block = order ? xcd_major_block(blockIdx.x, gridDim.x) : blockIdx.x;
    // work counters of this wave (lane 0 of each group adds its query's), flushed at the end
    __shared__ uint32_t cnt_lds[kMergeWaves][4];
    if (mstat && lane < 4) cnt_lds[wid][lane] = 0;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t j0 = kQ * (block * kMergeWaves + wid); j0 < count; j0 += stride) {
        // ---- this group's query (none past the end, or without a query, or outside the photon box) ----
        const uint32_t jj = j0 + grp;
        const uint32_t j = order && jj < count ? order[jj] : jj;
        CamVertex v;
        v.geom_plen = kNoQuery;
        if (jj < count) v = cam_vm[j];
        bool active = v.geom_plen != kNoQuery;
        uint32_t si = 0;
        float dmin[3] = {0.0f, 0.0f, 0.0f};
        if (active) {
            si = path_slice[v.path];
            const uint32_t* box = G.box + 8 * si;
            for (int k = 0; k < 3; ++k) {
                dmin[k] = v.pos[k] - ord2f(box[k]);
                const float dmax = ord2f(~box[3 + k]) - v.pos[k];
                if (dmin[k] < 0.0f || dmax < 0.0f) active = false;
            }
        }
        if (mstat && gl == 0 && jj < count) {
            atomicAdd(&cnt_lds[wid][0], 1u);
            if (active) atomicAdd(&cnt_lds[wid][1], 1u);
        }
        if (__ballot(active) == 0) continue;
        const Slice& sl = w.slices[si];
        // lanes 0..7 of the group: cell k = (x, y, z) bits of k, its photon range, offset in the scan list
        uint32_t cb = 0, cn = 0;
        if (active && gl < 8) {
            int32_t mn[3];
            for (int k = 0; k < 3; ++k) {
                float cp = sl.inv_cell * dmin[k];
                float cf = floorf(cp);
                float fr = cp - cf;
                mn[k] = (int32_t)cf;
                if (fr < 0.5f) --mn[k];
            }
            const uint32_t cell = sl.path_base + cell_hash(mn[0] + (int32_t)(gl & 1u), mn[1] + (int32_t)((gl >> 1) & 1u),
                                                           mn[2] + (int32_t)(gl >> 2), sl.N);
            cb = cell == 0 ? 0 : G.cell_ends[cell - 1];
            cn = G.cell_ends[cell] - cb;
        }
        uint32_t incl = cn;
        for (uint32_t d = 1; d < 8; d <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, d, GW);
            if (gl >= d) incl += y;
        }
        const uint32_t excl = incl - cn;
        uint32_t ends[8];
        for (int k = 0; k < 8; ++k) ends[k] = (uint32_t)__shfl((int)incl, k, GW);
        const uint32_t total = ends[7];  // 0 for a group without a query

        const f3 x = mk3(v.pos[0], v.pos[1], v.pos[2]);
        const f3 view = mk3(v.view[0], v.view[1], v.view[2]);
        const uint32_t plen = v.geom_plen & 31u;
        Surface s;
        s.position = x;
        s.error = 0.0f;
        float ccp = 0.0f;
        if (active) {
            const uint32_t geom = v.geom_plen >> 5;
            surface_material(sl.materials + 22 * (size_t)sl.geom_material[geom], mk3(v.n[0], v.n[1], v.n[2]), view, s);
            if (geom_has_tangents(sl, geom)) tangent_frame(s.n, w.cam_tan[j], s);
            s.base_color = mk3(v.base[0], v.base[1], v.base[2]);
            ccp = cont_prob(s);
        }
        f3 acc = mk3(0, 0, 0);
        auto merge = [&](uint32_t q) {
            const LightVertex& lv = G.sverts[q];
            const uint32_t lgp = lv.geom_plen;
            if (plen + (lgp & 31u) > sl.L) return;
            const float* lmat = sl.materials + 22 * (size_t)sl.geom_material[lgp >> 5];
            // the reference's transparency rule (VCM.cpp:381-383)
            if ((s.flags & kMatTransparent) != ((uint32_t)lmat[21] & kMatTransparent)) return;
            f3 lw = mk3(lv.wfix[0], lv.wfix[1], lv.wfix[2]);
            // a photon arriving in the surface's plane is not evaluated (measure zero; 0/0 below)
            if (dot3(s.n, lw) == 0.0f) return;
            float fwd, rev;
            f3 f = bsdf_eval(s, view, lw, fwd, rev);
            if (is_zero3(f)) return;
            float cos_l = absf_(dot3(s.n, lw));
            fwd *= ccp;
            rev *= lv.cp;
            float w_light = lv.dvcm * sl.vc_weight + lv.dvm * fwd;
            float w_cam = v.dvcm * sl.vc_weight + v.dvm * rev;
            float mis = 1.0f / (w_light + 1.0f + w_cam);
            acc += (f * mk3(lv.thr[0], lv.thr[1], lv.thr[2])) * (mis / cos_l);
        };
        uint32_t listed = 0;  // same in every lane of the group
        uint32_t n_near = 0;  // photons within the radius (counted only for mstat)
        for (uint32_t c0 = 0; __ballot(c0 < total) != 0; c0 += GW) {
            const uint32_t i = c0 + gl;
            // every lane takes part in the shuffles (lanes past the end look at their last photon)
            const uint32_t ic = i < total ? i : (total > 0 ? total - 1 : 0);
            uint32_t k = 0;
            for (int m = 0; m < 7; ++m) k += ic >= ends[m] ? 1u : 0u;
            const uint32_t q = (uint32_t)__shfl((int)cb, (int)k, GW) + (ic - (uint32_t)__shfl((int)excl, (int)k, GW));
            bool near = false;
            if (i < total) {
                const float4 lp = G.spos[q];
                const f3 dv = x - mk3(lp.x, lp.y, lp.z);
                near = dot3(dv, dv) <= sl.radius_sq;
            }
            const unsigned long long m = __ballot(near) & group_bits;
            if (near) list[listed + (uint32_t)__popcll(m & below)] = q;
            listed += (uint32_t)__popcll(m);
            n_near += (uint32_t)__popcll(m);
            __builtin_amdgcn_wave_barrier();
            if (__ballot(listed >= GW) != 0) {
                // every group evaluates up to GW listed photons and keeps the rest
                const uint32_t mine = gl < listed ? list[gl] : 0u;
                const uint32_t rest = gl + GW < listed ? list[gl + GW] : 0u;
                __builtin_amdgcn_wave_barrier();
                if (gl + GW < listed) list[gl] = rest;
                __builtin_amdgcn_wave_barrier();
                if (gl < listed) merge(mine);
                listed = listed > GW ? listed - GW : 0;
            }
        }
        if (gl < listed) merge(list[gl]);
        __builtin_amdgcn_wave_barrier();
        for (uint32_t off = 1; off < GW; off <<= 1) {
            acc.x += __shfl_xor(acc.x, off, GW);
            acc.y += __shfl_xor(acc.y, off, GW);
            acc.z += __shfl_xor(acc.z, off, GW);
        }
        if (mstat && gl == 0 && active) {
            atomicAdd(&cnt_lds[wid][2], total);
            atomicAdd(&cnt_lds[wid][3], n_near);
        }
        if (active && gl == 0) {
            const uint32_t p = v.path;
            s3(S.c, (size_t)p * w.nslots + w.nslots - 1, (mk3(v.thr[0], v.thr[1], v.thr[2]) * acc) * sl.vm_norm);
            S.valid[(size_t)p * w.nslots + w.nslots - 1] = 1;
        }
    }
    if (mstat) {
        __builtin_amdgcn_wave_barrier();
        if (lane < 4 && cnt_lds[wid][lane]) atomicAdd(mstat + lane, (unsigned long long)cnt_lds[wid][lane]);
    }
}

}  // namespace

void launch_cam_merge(uint32_t lanes, uint32_t blocks, hipStream_t stream, const Wave& w, const uint32_t* count_ptr,
                      const uint32_t* order, const CamVertex* cam_vm, const uint32_t* path_slice, const GridView& G,
                      const CamSlots& S, unsigned long long* mstat)
{
    if (lanes == 32)
        hipLaunchKernelGGL(k_cam_merge<32>, dim3(blocks), dim3(kBlock), 0, stream, w, count_ptr, order, cam_vm,
                           path_slice, G, S, mstat);
    else if (lanes == 8)
        hipLaunchKernelGGL(k_cam_merge<8>, dim3(blocks), dim3(kBlock), 0, stream, w, count_ptr, order, cam_vm,
                           path_slice, G, S, mstat);
    else
        hipLaunchKernelGGL(k_cam_merge<16>, dim3(blocks), dim3(kBlock), 0, stream, w, count_ptr, order, cam_vm,
                           path_slice, G, S, mstat);
}

}  // namespace sgpu_vcm
