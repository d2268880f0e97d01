// sgpu_intersect_noalloc: the closest-hit query of sgpu_intersect without any device allocation,
// for callers that cannot allocate (code under stream capture, tight serving loops, hosts that
// forbid hipMalloc). The only device memory sgpu_intersect allocates per call is the spill area of
// the per-lane traversal stacks (entries deeper than the SGPU_STACK_LDS kept in LDS, laid out
// [depth][thread]); here it is a fixed array in device memory, one per GPU, held until the process
// ends. The traversal grid is capped so that its spill fits the array (rays beyond the grid are
// taken in turn by the grid's waves, as in every traversal launch); the results are those of
// sgpu_intersect. The array's first words are the traversal's ray counters, cleared by each launch.
#include <mutex>

#include "sgpu_internal.h"

// 64 MiB: a spill depth of 62 entries (the 8-wide tree of a million triangles) still fits 1,056
// blocks of 256 lanes (4 per CU of MI355X).
static constexpr size_t kScratchWords = (size_t)16 << 20;
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device, looked up once

}  // namespace

int sgpu_noalloc_init(sgpu_context ctx)
{
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return sgpu_fail(SGPU_E_RANGE, "sgpu_context_create: device index beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[ctx->device]) {
        // the calling thread has made ctx->device current (sgpu_context_create)
        uint32_t* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return sgpu_fail(SGPU_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[ctx->device] = ptr;
    }
    ctx->noalloc_scratch = g_scratch[ctx->device];
    return SGPU_OK;
}

extern "C" int sgpu_intersect_noalloc(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits,
                                      uint32_t n, void* stream)
{
    int rc = sgpu_check_query(scene, rays, n, "sgpu_intersect_noalloc");
    if (rc) return rc;
    if (!hits) return sgpu_fail(SGPU_E_INVALID, "sgpu_intersect_noalloc: NULL hits");
    if ((hits->ng_x || hits->ng_y || hits->ng_z) && !(hits->ng_x && hits->ng_y && hits->ng_z))
        return sgpu_fail(SGPU_E_INVALID, "sgpu_intersect_noalloc: ng_x/ng_y/ng_z must be all set or all NULL");
    if (scene->n_curves > 0 && !hits->t)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_intersect_noalloc: a scene with curves needs hits.t");
    uint32_t* ptr = scene->ctx->noalloc_scratch;
    if (!ptr) return sgpu_fail(SGPU_E_STATE, "sgpu_intersect_noalloc: scratch not set up for this device");
    // blocks whose spilled stacks fit the scratch
    const size_t depth = scene->stack_cap > SGPU_STACK_LDS ? scene->stack_cap - SGPU_STACK_LDS : 1;
    const size_t fit = (kScratchWords - kTraceWorkWords) / (depth * SGPU_TRACE_BLOCK);
    if (fit == 0)
        return sgpu_fail(SGPU_E_RANGE, "sgpu_intersect_noalloc: the scene's traversal stack (" +
                                           std::to_string(scene->stack_cap) + " entries) needs more than the " +
                                           std::to_string(kScratchWords * 4 >> 20) + " MiB scratch per block");
    const uint32_t cap = (uint32_t)std::min<size_t>(fit, UINT32_MAX);
    return sgpu_trace_launch(scene, rays, hits, nullptr, n, false, nullptr, (hipStream_t)stream, ptr, nullptr, nullptr,
                             nullptr, cap);
}
