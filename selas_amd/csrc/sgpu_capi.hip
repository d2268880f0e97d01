// C ABI of libselas_gpu: contexts, device memory helpers, scenes and ray queries.
// Declarations and the Selas/Embree interfaces each entry point replaces are in include/selas_gpu.h.
#include "sgpu_ibl.h"
#include "sgpu_internal.h"
#include "sgpu_surface.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace {
thread_local std::string g_last_error;

hipStream_t as_stream(void* s) { return (hipStream_t)s; }

__global__ __launch_bounds__(256) void k_copy16(uint4* __restrict dst, const uint4* __restrict src, size_t n16)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

__global__ void k_copy1(uint8_t* __restrict dst, const uint8_t* __restrict src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Column-major 4x4 (Embree RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR): element (row r, col c) at m[c*4+r].
// A point p maps to M * (p, 1).
void xform_point(const float* m, const float* p, float* out)
{
    for (int r = 0; r < 3; ++r) out[r] = m[0 * 4 + r] * p[0] + m[1 * 4 + r] * p[1] + m[2 * 4 + r] * p[2] + m[3 * 4 + r];
}
}  // namespace

void sgpu_set_error(const std::string& msg) { g_last_error = msg; }

int sgpu_fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

extern "C" {

const char* sgpu_last_error(void) { return g_last_error.c_str(); }

int sgpu_context_create(int device, sgpu_context* out)
{
    if (!out) return sgpu_fail(SGPU_E_INVALID, "sgpu_context_create: out is NULL");
    *out = nullptr;
    int count = 0;
    SGPU_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return sgpu_fail(SGPU_E_INVALID, "sgpu_context_create: no such device");
    SGPU_HIP_TRY(hipSetDevice(device));
    sgpu_context ctx = new (std::nothrow) sgpu_context_s();
    if (!ctx) return sgpu_fail(SGPU_E_NOMEM, "sgpu_context_create: out of host memory");
    ctx->device = device;
    hipError_t e = hipGetDeviceProperties(&ctx->props, device);
    if (e != hipSuccess) {
        delete ctx;
        return sgpu_fail(SGPU_E_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    }
    // stream-ordered allocations freed by one call stay mapped for the next (wavefront buffers are
    // tens of GB at the tuned sizes; mapping them anew on every call costs more than the render)
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    const int rc = sgpu_noalloc_init(ctx);
    if (rc) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return SGPU_OK;
}

int sgpu_context_destroy(sgpu_context ctx)
{
    if (!ctx) return SGPU_OK;
    sgpu_vcm_release_cached(ctx);
    delete ctx;
    return SGPU_OK;
}

int sgpu_set_tuning(sgpu_context ctx, const sgpu_tuning* t)
{
    if (!ctx || !t) return sgpu_fail(SGPU_E_INVALID, "sgpu_set_tuning: NULL argument");
    if (t->vcm_wavefront_paths < (1u << 12) || t->vcm_wavefront_paths > (1u << 24))
        return sgpu_fail(SGPU_E_RANGE, "sgpu_set_tuning: vcm_wavefront_paths outside 2^12..2^24");
    if (t->trace_blocks_per_cu < 1 || t->trace_blocks_per_cu > kMaxTraceBlocksPerCu)
        return sgpu_fail(SGPU_E_RANGE, "sgpu_set_tuning: trace_blocks_per_cu outside 1..32");
    if (t->trace_tree > 1) return sgpu_fail(SGPU_E_RANGE, "sgpu_set_tuning: trace_tree must be 0 or 1");
    if (t->trace_lds_nodes > 9) return sgpu_fail(SGPU_E_RANGE, "sgpu_set_tuning: trace_lds_nodes outside 0..9");
    if (t->vcm_graphs > 1) return sgpu_fail(SGPU_E_RANGE, "sgpu_set_tuning: vcm_graphs must be 0 or 1");
    if (t->vcm_order & ~(SGPU_ORDER_MERGE_CELLS | SGPU_ORDER_SHADE_MATERIAL))
        return sgpu_fail(SGPU_E_RANGE, "sgpu_set_tuning: unknown vcm_order bits");
    if (t->vcm_merge_lanes != 8 && t->vcm_merge_lanes != 16 && t->vcm_merge_lanes != 32)
        return sgpu_fail(SGPU_E_RANGE, "sgpu_set_tuning: vcm_merge_lanes must be 8, 16 or 32");
    if (t->bvh_builder > 1) return sgpu_fail(SGPU_E_RANGE, "sgpu_set_tuning: bvh_builder must be 0 or 1");
    ctx->vcm_wavefront_paths = t->vcm_wavefront_paths;
    ctx->trace_blocks_per_cu = t->trace_blocks_per_cu;
    ctx->trace_tree = t->trace_tree;
    ctx->trace_lds_nodes = t->trace_lds_nodes;
    ctx->vcm_graphs = t->vcm_graphs;
    ctx->vcm_order = t->vcm_order;
    ctx->vcm_merge_lanes = t->vcm_merge_lanes;
    ctx->bvh_builder = t->bvh_builder;
    return SGPU_OK;
}

int sgpu_get_tuning(sgpu_context ctx, sgpu_tuning* t)
{
    if (!ctx || !t) return sgpu_fail(SGPU_E_INVALID, "sgpu_get_tuning: NULL argument");
    t->vcm_wavefront_paths = ctx->vcm_wavefront_paths;
    t->trace_blocks_per_cu = ctx->trace_blocks_per_cu;
    t->trace_tree = ctx->trace_tree;
    t->trace_lds_nodes = ctx->trace_lds_nodes;
    t->vcm_graphs = ctx->vcm_graphs;
    t->vcm_order = ctx->vcm_order;
    t->vcm_merge_lanes = ctx->vcm_merge_lanes;
    t->bvh_builder = ctx->bvh_builder;
    return SGPU_OK;
}

int sgpu_get_device_info(sgpu_context ctx, sgpu_device_info* info)
{
    if (!ctx || !info) return sgpu_fail(SGPU_E_INVALID, "sgpu_get_device_info: NULL argument");
    std::memset(info, 0, sizeof(*info));
    // some runtimes leave the marketing name empty: then the architecture name (gfx950:...)
    std::strncpy(info->name, ctx->props.name[0] ? ctx->props.name : ctx->props.gcnArchName, sizeof(info->name) - 1);
    info->compute_units = (uint32_t)ctx->props.multiProcessorCount;
    SGPU_HIP_TRY(hipSetDevice(ctx->device));
    size_t free_b = 0, total_b = 0;
    SGPU_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    info->total_bytes = total_b;
    info->free_bytes = free_b;
    return SGPU_OK;
}

int sgpu_malloc(sgpu_context ctx, size_t bytes, void** dptr)
{
    if (!ctx || !dptr) return sgpu_fail(SGPU_E_INVALID, "sgpu_malloc: NULL argument");
    *dptr = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) return sgpu_fail(SGPU_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return SGPU_OK;
}

int sgpu_free(sgpu_context ctx, void* dptr)
{
    (void)ctx;
    if (dptr) SGPU_HIP_TRY(hipFree(dptr));
    return SGPU_OK;
}

int sgpu_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream)
{
    if (bytes == 0) return SGPU_OK;
    if (!dst || !src) return sgpu_fail(SGPU_E_INVALID, "sgpu_memcpy_htod: NULL pointer");
    SGPU_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    SGPU_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return SGPU_OK;
}

int sgpu_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream)
{
    if (bytes == 0) return SGPU_OK;
    if (!dst || !src) return sgpu_fail(SGPU_E_INVALID, "sgpu_memcpy_dtoh: NULL pointer");
    SGPU_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    SGPU_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return SGPU_OK;
}

int sgpu_memset(void* dst, int value, size_t bytes, void* stream)
{
    if (bytes == 0) return SGPU_OK;
    if (!dst) return sgpu_fail(SGPU_E_INVALID, "sgpu_memset: NULL pointer");
    SGPU_HIP_TRY(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
    return SGPU_OK;
}

int sgpu_device_copy(void* dst, const void* src, size_t bytes, void* stream)
{
    if (bytes == 0) return SGPU_OK;
    if (!dst || !src) return sgpu_fail(SGPU_E_INVALID, "sgpu_device_copy: NULL pointer");
    if ((((uintptr_t)dst | (uintptr_t)src | bytes) & 15) == 0) {
        size_t n16 = bytes / 16;
        size_t blocks = (n16 + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(k_copy16, dim3((uint32_t)blocks), dim3(256), 0, as_stream(stream), (uint4*)dst,
                           (const uint4*)src, n16);
    } else {
        size_t blocks = (bytes + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(k_copy1, dim3((uint32_t)blocks), dim3(256), 0, as_stream(stream), (uint8_t*)dst,
                           (const uint8_t*)src, bytes);
    }
    SGPU_HIP_TRY(hipGetLastError());
    return SGPU_OK;
}

int sgpu_stream_create(void** stream)
{
    if (!stream) return sgpu_fail(SGPU_E_INVALID, "sgpu_stream_create: NULL argument");
    hipStream_t s;
    SGPU_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return SGPU_OK;
}

int sgpu_stream_destroy(void* stream)
{
    if (stream) SGPU_HIP_TRY(hipStreamDestroy(as_stream(stream)));
    return SGPU_OK;
}

int sgpu_stream_synchronize(void* stream)
{
    SGPU_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return SGPU_OK;
}

int sgpu_device_synchronize(void)
{
    SGPU_HIP_TRY(hipDeviceSynchronize());
    return SGPU_OK;
}

int sgpu_event_create(void** ev)
{
    if (!ev) return sgpu_fail(SGPU_E_INVALID, "sgpu_event_create: NULL argument");
    hipEvent_t e;
    SGPU_HIP_TRY(hipEventCreate(&e));
    *ev = (void*)e;
    return SGPU_OK;
}

int sgpu_event_destroy(void* ev)
{
    if (ev) SGPU_HIP_TRY(hipEventDestroy((hipEvent_t)ev));
    return SGPU_OK;
}

int sgpu_event_record(void* ev, void* stream)
{
    if (!ev) return sgpu_fail(SGPU_E_INVALID, "sgpu_event_record: NULL event");
    SGPU_HIP_TRY(hipEventRecord((hipEvent_t)ev, as_stream(stream)));
    return SGPU_OK;
}

int sgpu_event_elapsed_ms(void* start, void* stop, float* ms)
{
    if (!start || !stop || !ms) return sgpu_fail(SGPU_E_INVALID, "sgpu_event_elapsed_ms: NULL argument");
    SGPU_HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
    SGPU_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return SGPU_OK;
}

// ---- scenes --------------------------------------------------------------------------------------

int sgpu_scene_create(sgpu_context ctx, sgpu_scene* out)
{
    if (!ctx || !out) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_create: NULL argument");
    static std::atomic<uint64_t> next_uid{1};
    sgpu_scene s = new (std::nothrow) sgpu_scene_s();
    if (!s) return sgpu_fail(SGPU_E_NOMEM, "sgpu_scene_create: out of host memory");
    s->ctx = ctx;
    s->uid = next_uid.fetch_add(1);
    *out = s;
    return SGPU_OK;
}

int sgpu_scene_add_triangles(sgpu_scene scene, const float* positions, uint32_t n_verts, const uint32_t* indices,
                             uint32_t n_tris, const float* colmajor4x4, uint32_t* geom_id_out)
{
    if (!scene) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_triangles: NULL scene");
    if (n_tris > 0 && (!positions || !indices))
        return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_triangles: NULL positions or indices");
    std::lock_guard<std::mutex> lock(scene->mu);
    if (scene->committed) return sgpu_fail(SGPU_E_STATE, "sgpu_scene_add_triangles: scene already committed");
    if (!scene->instances.empty())
        return sgpu_fail(SGPU_E_STATE, "sgpu_scene_add_triangles: the scene holds instances, not geometries");
    const uint64_t total = (uint64_t)scene->tri_geom.size() + n_tris;
    if (total > (1ull << 30)) return sgpu_fail(SGPU_E_RANGE, "sgpu_scene_add_triangles: more than 2^30 triangles");
    for (uint64_t t = 0; t < (uint64_t)n_tris * 3; ++t)
        if (indices[t] >= n_verts)
            return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_triangles: index out of range");
    const uint32_t geom = (uint32_t)scene->geom_first.size();
    const size_t base = scene->tri_pos.size();
    try {
        scene->tri_pos.resize(base + (size_t)n_tris * 9);
        scene->tri_geom.resize(scene->tri_geom.size() + n_tris, geom);
        scene->geom_idx.emplace_back(indices, indices + (size_t)n_tris * 3);
        scene->geom_nverts.push_back(n_verts);
        scene->geom_xform.emplace_back();
        if (colmajor4x4) scene->geom_xform.back().assign(colmajor4x4, colmajor4x4 + 16);
    } catch (...) {
        scene->tri_pos.resize(base);
        scene->tri_geom.resize(scene->geom_first.empty() ? 0 : scene->tri_geom.size() - n_tris);
        scene->geom_idx.resize(geom);
        scene->geom_nverts.resize(geom);
        scene->geom_xform.resize(geom);
        return sgpu_fail(SGPU_E_NOMEM, "sgpu_scene_add_triangles: out of host memory");
    }
    scene->geom_first.push_back((uint32_t)scene->tri_geom.size() - n_tris);
    scene->geom_quad.push_back(0);
    float* dst = scene->tri_pos.data() + base;
    for (uint32_t t = 0; t < n_tris; ++t) {
        for (int k = 0; k < 3; ++k) {
            const float* p = positions + (size_t)indices[(size_t)t * 3 + k] * 3;
            float* o = dst + (size_t)t * 9 + k * 3;
            if (colmajor4x4) {
                xform_point(colmajor4x4, p, o);
            } else {
                o[0] = p[0];
                o[1] = p[1];
                o[2] = p[2];
            }
        }
    }
    if (geom_id_out) *geom_id_out = geom;
    return SGPU_OK;
}

int sgpu_scene_add_quads(sgpu_scene scene, const float* positions, uint32_t n_verts, const uint32_t* indices,
                         uint32_t n_quads, const float* colmajor4x4, uint32_t* geom_id_out)
{
    if (!scene) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_quads: NULL scene");
    if (n_quads > 0 && (!positions || !indices))
        return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_quads: NULL positions or indices");
    if ((uint64_t)n_quads * 2 > (1ull << 30)) return sgpu_fail(SGPU_E_RANGE, "sgpu_scene_add_quads: too many quads");
    // Embree's split of quad (v0, v1, v2, v3): triangles (v0, v1, v3) and (v2, v3, v1)
    std::vector<uint32_t> tri;
    try {
        tri.resize((size_t)n_quads * 6);
    } catch (...) {
        return sgpu_fail(SGPU_E_NOMEM, "sgpu_scene_add_quads: out of host memory");
    }
    for (uint32_t q = 0; q < n_quads; ++q) {
        const uint32_t* v = indices + (size_t)q * 4;
        uint32_t* t = tri.data() + (size_t)q * 6;
        t[0] = v[0];
        t[1] = v[1];
        t[2] = v[3];
        t[3] = v[2];
        t[4] = v[3];
        t[5] = v[1];
    }
    uint32_t geom = 0;
    int rc = sgpu_scene_add_triangles(scene, positions, n_verts, tri.data(), n_quads * 2, colmajor4x4, &geom);
    if (rc != SGPU_OK) return rc;
    std::lock_guard<std::mutex> lock(scene->mu);
    scene->geom_quad[geom] = 1;
    // attribute corners follow the quad's own vertex order
    scene->geom_idx[geom].assign(indices, indices + (size_t)n_quads * 4);
    if (geom_id_out) *geom_id_out = geom;
    return SGPU_OK;
}

int sgpu_scene_add_curves(sgpu_scene scene, const float* vertices, uint32_t n_verts, const uint32_t* indices,
                          uint32_t n_segments, const float* colmajor4x4, uint32_t* geom_id_out)
{
    if (!scene) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_curves: NULL scene");
    if (n_segments > 0 && (!vertices || !indices))
        return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_curves: NULL vertices or indices");
    std::lock_guard<std::mutex> lock(scene->mu);
    if (scene->committed) return sgpu_fail(SGPU_E_STATE, "sgpu_scene_add_curves: scene already committed");
    if (!scene->instances.empty())
        return sgpu_fail(SGPU_E_STATE, "sgpu_scene_add_curves: the scene holds instances, not geometries");
    if ((uint64_t)scene->curve_geom.size() + n_segments > (1ull << 28))
        return sgpu_fail(SGPU_E_RANGE, "sgpu_scene_add_curves: more than 2^28 curve segments");
    for (uint32_t s = 0; s < n_segments; ++s)
        if ((uint64_t)indices[s] + 3 >= n_verts)
            return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_curves: a segment's control points run past the vertices");
    for (uint64_t v = 0; v < (uint64_t)n_verts; ++v) {
        const float r = vertices[4 * v + 3];
        if (!(r >= 0.0f) || !std::isfinite(r))
            return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_curves: a radius is negative or not finite");
    }
    // the placement scales radii by its scale factor: a similarity keeps the swept spheres round
    float scale = 1.0f;
    if (colmajor4x4) {
        double c[3][3], n2[3];
        for (int j = 0; j < 3; ++j) {
            n2[j] = 0.0;
            for (int i = 0; i < 3; ++i) {
                c[j][i] = colmajor4x4[4 * j + i];
                n2[j] += c[j][i] * c[j][i];
            }
        }
        const double s2 = (n2[0] + n2[1] + n2[2]) / 3.0;
        bool similar = s2 > 0.0;
        for (int j = 0; j < 3 && similar; ++j) {
            similar &= std::fabs(n2[j] - s2) <= 1e-5 * s2;
            for (int k = j + 1; k < 3; ++k)
                similar &= std::fabs(c[j][0] * c[k][0] + c[j][1] * c[k][1] + c[j][2] * c[k][2]) <= 1e-5 * s2;
        }
        if (!similar)
            return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_curves: the placement is not a similarity (rotation, "
                                             "uniform scale, translation)");
        scale = (float)std::sqrt(s2);
    }
    const uint32_t geom = (uint32_t)scene->geom_first.size();
    const size_t base = scene->curve_cp.size();
    try {
        scene->curve_cp.resize(base + (size_t)n_segments * 16);
        scene->curve_geom.resize(scene->curve_geom.size() + n_segments, geom);
        scene->curve_prim.reserve(scene->curve_prim.size() + n_segments);
        scene->geom_idx.emplace_back();
        scene->geom_nverts.push_back(0);
        scene->geom_xform.emplace_back();
    } catch (...) {
        scene->curve_cp.resize(base);
        scene->curve_geom.resize(scene->curve_prim.size());
        scene->geom_idx.resize(geom);
        scene->geom_nverts.resize(geom);
        scene->geom_xform.resize(geom);
        return sgpu_fail(SGPU_E_NOMEM, "sgpu_scene_add_curves: out of host memory");
    }
    float* dst = scene->curve_cp.data() + base;
    for (uint32_t s = 0; s < n_segments; ++s) {
        scene->curve_prim.push_back(s);
        for (int j = 0; j < 4; ++j) {
            const float* p = vertices + 4 * ((size_t)indices[s] + j);
            float* o = dst + 16 * (size_t)s + 4 * j;
            if (colmajor4x4) xform_point(colmajor4x4, p, o);
            else {
                o[0] = p[0];
                o[1] = p[1];
                o[2] = p[2];
            }
            o[3] = p[3] * scale;
        }
    }
    // a geometry of no triangles: its triangle range is empty
    scene->geom_first.push_back((uint32_t)scene->tri_geom.size());
    scene->geom_quad.push_back(2);
    if (geom_id_out) *geom_id_out = geom;
    return SGPU_OK;
}

int sgpu_scene_set_vertex_attributes(sgpu_scene scene, uint32_t geom_id, const float* normals, const float* tangents,
                                     const float* uvs, uint32_t n_verts)
{
    const char* who = "sgpu_scene_set_vertex_attributes";
    if (!scene) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": NULL scene");
    if ((tangents && !normals) || (uvs && !(normals && tangents)))
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": tangents need normals, UVs need normals and tangents");
    std::lock_guard<std::mutex> lock(scene->mu);
    if (scene->committed) return sgpu_fail(SGPU_E_STATE, std::string(who) + ": scene already committed");
    if (geom_id >= scene->geom_first.size()) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": no such geometry");
    if (scene->geom_quad[geom_id] == 2)
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": a curve geometry (curves take no vertex attributes)");
    if (n_verts != scene->geom_nverts[geom_id])
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": vertex count differs from the geometry's");
    const size_t nv = n_verts;
    for (size_t i = 0; i < 3 * nv && normals; ++i)
        if (!std::isfinite(normals[i])) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": non-finite normal");
    for (size_t i = 0; i < 4 * nv && tangents; ++i)
        if (!std::isfinite(tangents[i])) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": non-finite tangent");
    for (size_t i = 0; i < 2 * nv && uvs; ++i)
        if (!std::isfinite(uvs[i])) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": non-finite UV");
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    if (scene->attr_host.size() < scene->geom_first.size())
        scene->attr_host.resize(scene->geom_first.size(), sgpu::GeomAttr{nullptr, nullptr, nullptr, 0, 0});
    sgpu::GeomAttr& ga = scene->attr_host[geom_id];
    // replaced attributes: their allocation is released
    const void* olds[3] = {ga.n, ga.t, ga.uv};
    for (const void* o : olds) {
        if (!o) continue;
        auto it = std::find_if(scene->attr_allocs.begin(), scene->attr_allocs.end(),
                               [&](const std::pair<void*, size_t>& a) { return a.first == o; });
        if (it != scene->attr_allocs.end()) {
            (void)hipFree(it->first);
            scene->attr_allocs.erase(it);
        }
    }
    const std::vector<uint32_t>& idx = scene->geom_idx[geom_id];
    const uint32_t corners = scene->geom_quad[geom_id] == 1 ? 4u : 3u;
    const std::vector<float>& m = scene->geom_xform[geom_id];
    ga = sgpu::GeomAttr{nullptr, nullptr, nullptr, corners, 0};
    // per corner: the vertex's value, normals and tangents' xyz through the placement's linear part
    auto upload = [&](const float* src, int K, bool vec, const float** out) -> int {
        std::vector<float> c(idx.size() * (size_t)K);
        for (size_t i = 0; i < idx.size(); ++i) {
            const float* a = src + (size_t)idx[i] * K;
            float* o = c.data() + i * K;
            for (int k = 0; k < K; ++k) o[k] = a[k];
            if (vec && !m.empty())
                for (int r = 0; r < 3; ++r) o[r] = m[0 * 4 + r] * a[0] + m[1 * 4 + r] * a[1] + m[2 * 4 + r] * a[2];
        }
        void* d = nullptr;
        SGPU_HIP_TRY(hipMalloc(&d, std::max<size_t>(c.size(), 1) * sizeof(float)));
        scene->attr_allocs.emplace_back(d, c.size() * sizeof(float));
        SGPU_HIP_TRY(hipMemcpy(d, c.data(), c.size() * sizeof(float), hipMemcpyHostToDevice));
        *out = (const float*)d;
        return SGPU_OK;
    };
    int rc = SGPU_OK;
    if (normals) rc = upload(normals, 3, true, &ga.n);
    if (rc == SGPU_OK && tangents) rc = upload(tangents, 4, true, &ga.t);
    if (rc == SGPU_OK && uvs) rc = upload(uvs, 2, false, &ga.uv);
    if (rc != SGPU_OK) ga = sgpu::GeomAttr{nullptr, nullptr, nullptr, corners, 0};
    return rc;
}

int sgpu_scene_add_instance(sgpu_scene scene, sgpu_scene child, const float* colmajor4x4, uint32_t* inst_id_out)
{
    if (!scene || !child || !colmajor4x4) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_instance: NULL argument");
    if (child == scene) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_instance: a scene cannot place itself");
    if (child->ctx != scene->ctx)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_instance: the child belongs to another context");
    for (int k = 0; k < 16; ++k)
        if (!std::isfinite(colmajor4x4[k]))
            return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_instance: transform is not finite");
    std::lock(scene->mu, child->mu);
    std::lock_guard<std::mutex> la(scene->mu, std::adopt_lock), lb(child->mu, std::adopt_lock);
    if (scene->committed) return sgpu_fail(SGPU_E_STATE, "sgpu_scene_add_instance: scene already committed");
    if (!child->committed) return sgpu_fail(SGPU_E_STATE, "sgpu_scene_add_instance: child scene not committed");
    if (!scene->tri_geom.empty())
        return sgpu_fail(SGPU_E_STATE, "sgpu_scene_add_instance: the scene holds geometries, not instances");
    if (child->inst_levels > 1)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_add_instance: more than two instance levels");
    if (scene->instances.size() >= (1u << 30))
        return sgpu_fail(SGPU_E_RANGE, "sgpu_scene_add_instance: more than 2^30 instances");
    sgpu_scene_s::Instance in;
    in.child = child;
    // column-major 4x4 -> row-major 3x4 (Embree's RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR)
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) in.l2w[4 * r + c] = (double)colmajor4x4[4 * c + r];
    try {
        scene->instances.push_back(in);
    } catch (...) {
        return sgpu_fail(SGPU_E_NOMEM, "sgpu_scene_add_instance: out of host memory");
    }
    ++child->refs;
    if (inst_id_out) *inst_id_out = (uint32_t)(scene->instances.size() - 1);
    return SGPU_OK;
}

// The shading data of an instanced scene (see sgpu_scene_s): the placed scenes' geometries as one
// list, their shading records and textures as they are now. A placed scene without materials for
// every geometry leaves the instanced scene queryable but not renderable (render_error).
static int gather_instance_shading(sgpu_scene scene)
{
    std::vector<sgpu_scene> models;
    std::vector<uint32_t> model_of(scene->comp_blas.size());
    for (size_t k = 0; k < scene->comp_blas.size(); ++k) {
        sgpu_scene m = scene->comp_blas[k];
        size_t j = std::find(models.begin(), models.end(), m) - models.begin();
        if (j == models.size()) models.push_back(m);
        model_of[k] = (uint32_t)j;
    }
    std::vector<uint32_t> g0(models.size()), m0(models.size());
    uint64_t n_geoms = 0, n_mats = 0;
    for (size_t j = 0; j < models.size(); ++j) {
        sgpu_scene m = models[j];
        if (!m->d_materials || m->n_geom_material < m->geom_first.size()) {
            scene->render_error = "a placed scene has no materials for every geometry (set them before committing "
                                  "the scene that places it)";
            return SGPU_OK;
        }
        g0[j] = (uint32_t)n_geoms;
        m0[j] = (uint32_t)n_mats;
        n_geoms += m->geom_first.size();
        n_mats += m->n_materials;
    }
    if (n_geoms >= (1u << 27)) {
        scene->render_error = "the placed scenes hold more than 2^27 geometries";
        return SGPU_OK;
    }
    std::vector<uint32_t> gmat(n_geoms);
    std::vector<sgpu::PtexTex> gtex(n_geoms, sgpu::PtexTex{nullptr, nullptr, 0, 0, 0});
    std::vector<sgpu::GeomAttr> gattr(n_geoms, sgpu::GeomAttr{nullptr, nullptr, nullptr, 0, 0});
    bool any_attr = false;
    for (size_t j = 0; j < models.size(); ++j) {
        sgpu_scene m = models[j];
        const size_t ng = m->geom_first.size();
        SGPU_HIP_TRY(hipMemcpy(gmat.data() + g0[j], m->d_geom_material, ng * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (size_t g = 0; g < ng; ++g) {
            gmat[g0[j] + g] += m0[j];
            if (g < m->tex_host.size()) gtex[g0[j] + g] = m->tex_host[g];
            if (m->d_geom_attr && g < m->attr_host.size()) {
                gattr[g0[j] + g] = m->attr_host[g];
                any_attr |= m->attr_host[g].n != nullptr;
                scene->has_tangents |= m->attr_host[g].t != nullptr;
            }
        }
    }
    if (any_attr) {
        SGPU_HIP_TRY(hipMalloc((void**)&scene->d_g_geom_attr, sizeof(sgpu::GeomAttr) * (size_t)n_geoms));
        SGPU_HIP_TRY(hipMemcpy(scene->d_g_geom_attr, gattr.data(), sizeof(sgpu::GeomAttr) * n_geoms,
                               hipMemcpyHostToDevice));
    }
    std::vector<uint32_t> base(model_of.size());
    for (size_t k = 0; k < base.size(); ++k) base[k] = g0[model_of[k]];
    SGPU_HIP_TRY(hipMalloc((void**)&scene->d_g_materials, 22 * sizeof(float) * (size_t)std::max<uint64_t>(n_mats, 1)));
    for (size_t j = 0; j < models.size(); ++j)
        SGPU_HIP_TRY(hipMemcpy(scene->d_g_materials + 22 * (size_t)m0[j], models[j]->d_materials,
                               22 * sizeof(float) * (size_t)models[j]->n_materials, hipMemcpyDeviceToDevice));
    SGPU_HIP_TRY(hipMalloc((void**)&scene->d_g_geom_material, sizeof(uint32_t) * (size_t)std::max<uint64_t>(n_geoms, 1)));
    SGPU_HIP_TRY(hipMemcpy(scene->d_g_geom_material, gmat.data(), sizeof(uint32_t) * n_geoms, hipMemcpyHostToDevice));
    SGPU_HIP_TRY(hipMalloc((void**)&scene->d_g_geom_tex, sizeof(sgpu::PtexTex) * (size_t)std::max<uint64_t>(n_geoms, 1)));
    SGPU_HIP_TRY(hipMemcpy(scene->d_g_geom_tex, gtex.data(), sizeof(sgpu::PtexTex) * n_geoms, hipMemcpyHostToDevice));
    SGPU_HIP_TRY(hipMalloc((void**)&scene->d_inst_geom_base, sizeof(uint32_t) * std::max<size_t>(base.size(), 1)));
    SGPU_HIP_TRY(hipMemcpy(scene->d_inst_geom_base, base.data(), sizeof(uint32_t) * base.size(), hipMemcpyHostToDevice));
    SGPU_HIP_TRY(hipMalloc((void**)&scene->d_inst_l2w, sizeof(float) * std::max<size_t>(scene->comp_l2w.size(), 1)));
    SGPU_HIP_TRY(hipMemcpy(scene->d_inst_l2w, scene->comp_l2w.data(), sizeof(float) * scene->comp_l2w.size(),
                           hipMemcpyHostToDevice));
    scene->n_g_geoms = (uint32_t)n_geoms;
    return SGPU_OK;
}

int sgpu_scene_commit(sgpu_scene scene, void* stream)
{
    if (!scene) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_commit: NULL scene");
    std::lock_guard<std::mutex> lock(scene->mu);
    if (scene->committed) return sgpu_fail(SGPU_E_STATE, "sgpu_scene_commit: scene already committed");
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    SGPU_HIP_TRY(hipMalloc((void**)&scene->d_overflow, sizeof(uint32_t)));
    SGPU_HIP_TRY(hipMemset(scene->d_overflow, 0, sizeof(uint32_t)));
    if (!scene->instances.empty()) {
        int rc = sgpu_bvh_build_instances(scene, as_stream(stream));
        if (rc == SGPU_OK) rc = gather_instance_shading(scene);
        if (rc != SGPU_OK) return rc;
        if (scene->n_inst > 0 || scene->n_curves > 0) {
            const float* b = scene->box;
            const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
            for (int k = 0; k < 3; ++k) scene->bsphere[k] = (b[k] + b[3 + k]) * 0.5f;
            scene->bsphere[3] = 0.5f * sqrtf(dx * dx + dy * dy + dz * dz);
        }
        scene->committed = true;
        return SGPU_OK;
    }
    int rc = sgpu_bvh_build(scene, as_stream(stream));
    if (rc == SGPU_OK && !scene->curve_geom.empty())
        rc = sgpu_curve_build(scene, scene->curve_cp, scene->curve_geom, scene->curve_prim,
                              std::vector<uint32_t>(scene->curve_geom.size(), SGPU_INVALID_ID), as_stream(stream));
    if (rc != SGPU_OK) return rc;
    // the attribute table (one record per geometry) when any geometry has attributes
    bool any_attr = false;
    for (const sgpu::GeomAttr& ga : scene->attr_host) {
        any_attr |= ga.n != nullptr;
        scene->has_tangents |= ga.t != nullptr;
    }
    if (any_attr) {
        scene->attr_host.resize(scene->geom_first.size(), sgpu::GeomAttr{nullptr, nullptr, nullptr, 0, 0});
        SGPU_HIP_TRY(hipMalloc((void**)&scene->d_geom_attr, sizeof(sgpu::GeomAttr) * scene->attr_host.size()));
        SGPU_HIP_TRY(hipMemcpy(scene->d_geom_attr, scene->attr_host.data(),
                               sizeof(sgpu::GeomAttr) * scene->attr_host.size(), hipMemcpyHostToDevice));
    }
    std::vector<std::vector<uint32_t>>().swap(scene->geom_idx);
    std::vector<std::vector<float>>().swap(scene->geom_xform);
    // the sphere around the triangles' (and curves') bounding box (the IBL emits from its disc)
    if (!scene->tri_pos.empty() || scene->n_curves > 0) {
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = scene->tri_pos.empty() ? scene->box[k] : scene->tri_pos[k];
            hi[k] = scene->tri_pos.empty() ? scene->box[3 + k] : scene->tri_pos[k];
        }
        for (size_t i = 0; i < scene->tri_pos.size(); i += 3)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], scene->tri_pos[i + k]);
                hi[k] = std::max(hi[k], scene->tri_pos[i + k]);
            }
        if (scene->n_curves > 0)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], scene->box[k]);
                hi[k] = std::max(hi[k], scene->box[3 + k]);
            }
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        for (int k = 0; k < 3; ++k) scene->bsphere[k] = (lo[k] + hi[k]) * 0.5f;
        scene->bsphere[3] = 0.5f * sqrtf(dx * dx + dy * dy + dz * dz);
    }
    scene->committed = true;
    std::vector<float>().swap(scene->tri_pos);
    return SGPU_OK;
}

int sgpu_scene_release(sgpu_scene scene)
{
    if (!scene) return SGPU_OK;
    {
        std::lock_guard<std::mutex> lock(scene->mu);
        if (scene->refs > 0)
            return sgpu_fail(SGPU_E_STATE, "sgpu_scene_release: the scene is placed by an instance of another scene");
        if (scene->server_jobs.load() > 0)
            return sgpu_fail(SGPU_E_STATE, "sgpu_scene_release: jobs on the scene are admitted to a VCM server");
    }
    for (const auto& in : scene->instances) {
        std::lock_guard<std::mutex> lock(in.child->mu);
        --in.child->refs;
    }
    if (scene->d_inst) (void)hipFree(scene->d_inst);
    if (scene->d_inst_ids) (void)hipFree(scene->d_inst_ids);
    if (scene->d_g_materials) (void)hipFree(scene->d_g_materials);
    if (scene->d_g_geom_material) (void)hipFree(scene->d_g_geom_material);
    if (scene->d_g_geom_tex) (void)hipFree(scene->d_g_geom_tex);
    if (scene->d_g_geom_attr) (void)hipFree(scene->d_g_geom_attr);
    if (scene->d_geom_attr) (void)hipFree(scene->d_geom_attr);
    for (const auto& a : scene->attr_allocs) (void)hipFree(a.first);
    if (scene->d_inst_geom_base) (void)hipFree(scene->d_inst_geom_base);
    if (scene->d_inst_l2w) (void)hipFree(scene->d_inst_l2w);
    if (scene->d_nodes) (void)hipFree(scene->d_nodes);
    if (scene->d_tris) (void)hipFree(scene->d_tris);
    if (scene->d_wide) (void)hipFree(scene->d_wide);
    if (scene->d_cnodes) (void)hipFree(scene->d_cnodes);
    if (scene->d_curves) (void)hipFree(scene->d_curves);
    if (scene->d_overflow) (void)hipFree(scene->d_overflow);
    if (scene->d_materials) (void)hipFree(scene->d_materials);
    if (scene->d_geom_material) (void)hipFree(scene->d_geom_material);
    if (scene->d_lights) (void)hipFree(scene->d_lights);
    for (void* a : scene->tex_allocs) (void)hipFree(a);
    if (scene->d_geom_tex) (void)hipFree(scene->d_geom_tex);
    if (scene->d_ibl) (void)hipFree(scene->d_ibl);
    delete scene;
    return SGPU_OK;
}

int sgpu_scene_set_materials(sgpu_scene scene, const float* materials, uint32_t n_materials,
                             const uint32_t* geom_material, uint32_t n_geoms)
{
    if (!scene || !materials || !geom_material || n_materials == 0 || n_geoms == 0)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_materials: NULL or empty argument");
    for (uint32_t g = 0; g < n_geoms; ++g)
        if (geom_material[g] >= n_materials)
            return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_materials: material index out of range");
    for (uint32_t m = 0; m < n_materials; ++m) {
        float sh = materials[22 * (size_t)m + 20];
        if (!(sh == 0.0f || sh == 1.0f || sh == 2.0f))
            return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_materials: shader must be 0, 1 or 2");
    }
    std::lock_guard<std::mutex> lock(scene->mu);
    if (scene->server_jobs.load() > 0)
        return sgpu_fail(SGPU_E_STATE, "sgpu_scene_set_materials: jobs on the scene are admitted to a VCM server");
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    float* dm = nullptr;
    uint32_t* dg = nullptr;
    SGPU_HIP_TRY(hipMalloc((void**)&dm, 22 * sizeof(float) * (size_t)n_materials));
    hipError_t e = hipMalloc((void**)&dg, sizeof(uint32_t) * (size_t)n_geoms);
    if (e != hipSuccess) {
        (void)hipFree(dm);
        return sgpu_fail(SGPU_E_NOMEM, "sgpu_scene_set_materials: hipMalloc failed");
    }
    std::vector<float> shading(22 * (size_t)n_materials);
    for (uint32_t m = 0; m < n_materials; ++m) sgpu::shading_record(materials + 22 * (size_t)m, &shading[22 * (size_t)m]);
    e = hipMemcpy(dm, shading.data(), 22 * sizeof(float) * (size_t)n_materials, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dg, geom_material, sizeof(uint32_t) * (size_t)n_geoms, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(dm);
        (void)hipFree(dg);
        return sgpu_fail(SGPU_E_HIP, std::string("sgpu_scene_set_materials: ") + hipGetErrorString(e));
    }
    if (scene->d_materials) (void)hipFree(scene->d_materials);
    if (scene->d_geom_material) (void)hipFree(scene->d_geom_material);
    scene->d_materials = dm;
    scene->n_materials = n_materials;
    scene->d_geom_material = dg;
    scene->n_geom_material = n_geoms;
    return SGPU_OK;
}

int sgpu_scene_set_lights(sgpu_scene scene, const sgpu_light* lights, uint32_t n_lights, const float* background_rgb)
{
    if (!scene || (n_lights > 0 && !lights)) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_lights: NULL argument");
    for (uint32_t i = 0; i < n_lights; ++i)
        if (lights[i].type > 1) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_lights: light type must be 0 or 1");
    std::lock_guard<std::mutex> lock(scene->mu);
    if (scene->server_jobs.load() > 0)
        return sgpu_fail(SGPU_E_STATE, "sgpu_scene_set_lights: jobs on the scene are admitted to a VCM server");
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    void* dl = nullptr;
    if (n_lights > 0) {
        SGPU_HIP_TRY(hipMalloc(&dl, sizeof(sgpu_light) * (size_t)n_lights));
        hipError_t e = hipMemcpy(dl, lights, sizeof(sgpu_light) * (size_t)n_lights, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(dl);
            return sgpu_fail(SGPU_E_HIP, std::string("sgpu_scene_set_lights: ") + hipGetErrorString(e));
        }
    }
    if (scene->d_lights) (void)hipFree(scene->d_lights);
    scene->d_lights = dl;
    scene->n_lights = n_lights;
    for (int k = 0; k < 3; ++k) scene->background[k] = background_rgb ? background_rgb[k] : 0.0f;
    return SGPU_OK;
}

// CalculateIntensityMap + CalculateStrataDistributionFunctions (BuildImageBasedLight.cpp:18-103):
// mdf (h entries) and cdf (w h entries), summed in the same order, the last entry of each forced
// to 1.
static void ibl_build_tables(uint32_t w, uint32_t h, const float* rgb, float* mdf, float* cdf)
{
    const float hf = (float)h;
    std::fill(mdf, mdf + h, 0.0f);
    std::fill(cdf, cdf + (size_t)w * h, 0.0f);
    std::vector<float> row(w);
    float marginal_sum = 0.0f;
    for (uint32_t y = 0; y < h; ++y) {
        const float theta = ((float)y + 0.5f) * sgpu::kPi / hf;
        const float st = sin_f(theta);
        float conditional_sum = 0.0f;
        for (uint32_t x = 0; x < w; ++x) {
            const float* t = rgb + 3 * ((size_t)y * w + x);
            row[x] = st * (t[0] * 0.299f + t[1] * 0.587f + t[2] * 0.114f);
            conditional_sum += row[x];
        }
        mdf[y] = conditional_sum;
        marginal_sum += conditional_sum;
        float* c = cdf + (size_t)y * w;
        if (conditional_sum > 0.0f) {
            const float oo = 1.0f / conditional_sum;
            float integration = 0.0f;
            for (uint32_t x = 0; x < w; ++x) {
                integration += oo * row[x];
                c[x] = integration;
            }
        }
        c[w - 1] = 1.0f;
    }
    if (marginal_sum > 0.0f) {
        const float oo = 1.0f / marginal_sum;
        float integration = 0.0f;
        for (uint32_t y = 0; y < h; ++y) {
            integration += mdf[y] * oo;
            mdf[y] = integration;
        }
    }
    mdf[h - 1] = 1.0f;
}

int sgpu_scene_set_ibl(sgpu_scene scene, uint32_t width, uint32_t height, const float* rgb, float rotation_radians,
                       float exposure)
{
    if (!scene) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_ibl: NULL scene");
    std::lock_guard<std::mutex> lock(scene->mu);
    if (scene->server_jobs.load() > 0)
        return sgpu_fail(SGPU_E_STATE, "sgpu_scene_set_ibl: jobs on the scene are admitted to a VCM server");
    if (!scene->committed) return sgpu_fail(SGPU_E_STATE, "sgpu_scene_set_ibl: scene not committed");
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    if (!rgb) {
        if (scene->d_ibl) SGPU_HIP_TRY(hipFree(scene->d_ibl));
        scene->d_ibl = nullptr;
        scene->ibl_w = scene->ibl_h = 0;
        return SGPU_OK;
    }
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_ibl: width and height must be in [1, 16384]");
    if (!std::isfinite(rotation_radians) || !std::isfinite(exposure) || exposure < 0.0f)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_ibl: rotation must be finite and exposure finite and >= 0");
    const size_t n = (size_t)width * height;
    for (size_t i = 0; i < 3 * n; ++i)
        if (!(rgb[i] >= 0.0f) || !std::isfinite(rgb[i]))
            return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_ibl: texels must be finite and >= 0");
    std::vector<float> host(3 * n + height + n);
    std::memcpy(host.data(), rgb, 3 * n * sizeof(float));
    ibl_build_tables(width, height, rgb, host.data() + 3 * n, host.data() + 3 * n + height);
    float* d = nullptr;
    SGPU_HIP_TRY(hipMalloc((void**)&d, host.size() * sizeof(float)));
    hipError_t e = hipMemcpy(d, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return sgpu_fail(SGPU_E_HIP, std::string("sgpu_scene_set_ibl: ") + hipGetErrorString(e));
    }
    if (scene->d_ibl) (void)hipFree(scene->d_ibl);
    scene->d_ibl = d;
    scene->ibl_w = width;
    scene->ibl_h = height;
    float r = fmodf(rotation_radians, sgpu::kTwoPi);
    if (r < 0.0f) r += sgpu::kTwoPi;
    if (r >= sgpu::kTwoPi) r = 0.0f;
    scene->ibl_rotation = r;
    scene->ibl_exposure = exposure;
    return SGPU_OK;
}

int sgpu_scene_set_ptex(sgpu_scene scene, uint32_t geom_id, uint32_t res_log2, uint32_t n_faces,
                        const uint8_t* texels, const int32_t* adjfaces, const int32_t* adjedges)
{
    if (!scene || !texels || !adjfaces || !adjedges || n_faces == 0)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_ptex: NULL or empty argument");
    std::lock_guard<std::mutex> lock(scene->mu);
    if (scene->server_jobs.load() > 0)
        return sgpu_fail(SGPU_E_STATE, "sgpu_scene_set_ptex: jobs on the scene are admitted to a VCM server");
    const size_t n_geoms = scene->geom_first.size();
    if (geom_id >= n_geoms) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_ptex: no such geometry");
    if (scene->geom_quad[geom_id] == 2) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_ptex: a curve geometry");
    const bool tri = !scene->geom_quad[geom_id];
    const uint32_t first = scene->geom_first[geom_id];
    const uint32_t end = geom_id + 1 < n_geoms ? scene->geom_first[geom_id + 1] : (uint32_t)scene->tri_geom.size();
    if ((tri ? end - first : (end - first) / 2) != n_faces)
        return sgpu_fail(SGPU_E_INVALID, tri ? "sgpu_scene_set_ptex: face count differs from the geometry's triangle count"
                                             : "sgpu_scene_set_ptex: face count differs from the geometry's quad count");
    if (res_log2 > 8) return sgpu_fail(SGPU_E_RANGE, "sgpu_scene_set_ptex: res_log2 above 8");
    const int edges = tri ? 3 : 4;
    for (size_t f = 0; f < n_faces; ++f)
        for (int e = 0; e < edges; ++e) {
            const size_t i = 4 * f + e;
            if (adjfaces[i] < -1 || adjfaces[i] >= (int32_t)n_faces)
                return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_ptex: adjacent face out of range");
            if (adjedges[i] < 0 || adjedges[i] >= edges)
                return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_set_ptex: adjacent edge out of range");
        }
    const size_t res = (size_t)1 << res_log2;
    const size_t texel_bytes = (size_t)n_faces * res * res * 3;
    std::vector<int32_t> adj((size_t)n_faces * 8);
    for (size_t f = 0; f < n_faces; ++f)
        for (int e = 0; e < 4; ++e) {
            adj[8 * f + e] = e < edges ? adjfaces[4 * f + e] : -1;
            adj[8 * f + 4 + e] = e < edges ? adjedges[4 * f + e] : 0;
        }
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    uint8_t* dt = nullptr;
    int32_t* da = nullptr;
    SGPU_HIP_TRY(hipMalloc((void**)&dt, texel_bytes));
    scene->tex_allocs.push_back(dt);
    SGPU_HIP_TRY(hipMalloc((void**)&da, adj.size() * sizeof(int32_t)));
    scene->tex_allocs.push_back(da);
    SGPU_HIP_TRY(hipMemcpy(dt, texels, texel_bytes, hipMemcpyHostToDevice));
    SGPU_HIP_TRY(hipMemcpy(da, adj.data(), adj.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    scene->tex_host.resize(n_geoms, sgpu::PtexTex{nullptr, nullptr, 0, 0, 0});
    scene->tex_host[geom_id] = sgpu::PtexTex{dt, da, n_faces, (int32_t)res_log2, tri ? 1 : 0};
    // the per-geometry table the shading kernels read
    if (scene->d_geom_tex) (void)hipFree(scene->d_geom_tex);
    scene->d_geom_tex = nullptr;
    scene->n_geom_tex = 0;
    SGPU_HIP_TRY(hipMalloc((void**)&scene->d_geom_tex, n_geoms * sizeof(sgpu::PtexTex)));
    SGPU_HIP_TRY(hipMemcpy(scene->d_geom_tex, scene->tex_host.data(), n_geoms * sizeof(sgpu::PtexTex),
                           hipMemcpyHostToDevice));
    scene->n_geom_tex = (uint32_t)n_geoms;
    return SGPU_OK;
}

namespace {
__global__ void k_ptex_eval(sgpu::PtexTex tx, const int32_t* __restrict faces, const float* __restrict u,
                            const float* __restrict v, uint32_t n, float* __restrict out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 c = sgpu::ptex_eval(tx, faces[i], u[i], v[i]);
    out[3 * i] = c.x;
    out[3 * i + 1] = c.y;
    out[3 * i + 2] = c.z;
}
}  // namespace

int sgpu_ptex_eval(sgpu_scene scene, uint32_t geom_id, const int32_t* faces, const float* u, const float* v,
                   uint32_t n, float* out, void* stream)
{
    if (!scene) return sgpu_fail(SGPU_E_INVALID, "sgpu_ptex_eval: NULL scene");
    if (n > 0 && (!faces || !u || !v || !out)) return sgpu_fail(SGPU_E_INVALID, "sgpu_ptex_eval: NULL array");
    sgpu::PtexTex tx;
    {
        std::lock_guard<std::mutex> lock(scene->mu);
        if (geom_id >= scene->tex_host.size() || !scene->tex_host[geom_id].texels)
            return sgpu_fail(SGPU_E_STATE, "sgpu_ptex_eval: geometry has no texture");
        tx = scene->tex_host[geom_id];
    }
    if (n == 0) return SGPU_OK;
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    hipLaunchKernelGGL(k_ptex_eval, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), tx, faces, u, v, n, out);
    SGPU_HIP_TRY(hipGetLastError());
    return SGPU_OK;
}

static int check_pt_job(sgpu_scene scene, const sgpu_pt_job* job, const float* framebuffer, const std::string& who)
{
    if (!scene || !job || !framebuffer) return sgpu_fail(SGPU_E_INVALID, who + ": NULL argument");
    if (!scene->committed) return sgpu_fail(SGPU_E_STATE, who + ": scene not committed");
    if (!scene->instances.empty()) {
        if (!scene->render_error.empty()) return sgpu_fail(SGPU_E_STATE, who + ": " + scene->render_error);
    } else if (!scene->d_materials || scene->n_geom_material < scene->geom_first.size()) {
        return sgpu_fail(SGPU_E_STATE, who + ": materials not set for every geometry");
    }
    if (job->camera.width == 0 || job->camera.height == 0 || job->spp == 0 || job->max_path_length == 0)
        return sgpu_fail(SGPU_E_INVALID, who + ": zero width, height, spp or path length");
    if ((uint64_t)job->camera.width * job->camera.height > (1ull << 26))
        return sgpu_fail(SGPU_E_RANGE, who + ": more than 2^26 pixels");
    if (job->spp > (1u << 20)) return sgpu_fail(SGPU_E_RANGE, who + ": spp above 2^20");
    if (job->max_path_length > 1024) return sgpu_fail(SGPU_E_RANGE, who + ": path length above 1024");
    if (job->flags & ~(SGPU_PT_NEE_COSINE_ONCE | SGPU_PT_COUNT_TRAVERSAL))
        return sgpu_fail(SGPU_E_INVALID, who + ": unknown flags");
    return SGPU_OK;
}

int sgpu_render_pt(sgpu_scene scene, const sgpu_pt_job* job, float* framebuffer, void* stream)
{
    int rc = check_pt_job(scene, job, framebuffer, "sgpu_render_pt");
    if (rc) return rc;
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    return sgpu_pt_render_impl(scene, job, framebuffer, as_stream(stream));
}

int sgpu_render_pt_stats(sgpu_scene scene, const sgpu_pt_job* job, float* framebuffer, sgpu_pt_stats* stats,
                         void* stream)
{
    int rc = check_pt_job(scene, job, framebuffer, "sgpu_render_pt_stats");
    if (rc) return rc;
    if (!stats) return sgpu_fail(SGPU_E_INVALID, "sgpu_render_pt_stats: NULL stats");
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    return sgpu_pt_render_impl(scene, job, framebuffer, as_stream(stream), stats);
}

int sgpu_render_deferred_pt(sgpu_scene scene, const sgpu_dpt_job* job, float* framebuffer, void* stream)
{
    const char* who = "sgpu_render_deferred_pt";
    if (!scene || !job || !framebuffer) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": NULL argument");
    if (!scene->committed) return sgpu_fail(SGPU_E_STATE, std::string(who) + ": scene not committed");
    if (!scene->instances.empty()) {
        if (!scene->render_error.empty()) return sgpu_fail(SGPU_E_STATE, std::string(who) + ": " + scene->render_error);
    } else if (!scene->d_materials || scene->n_geom_material < scene->geom_first.size()) {
        return sgpu_fail(SGPU_E_STATE, std::string(who) + ": materials not set for every geometry");
    }
    if (job->camera.width == 0 || job->camera.height == 0 || job->spp_x == 0 || job->spp_y == 0 ||
        job->max_path_length == 0)
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": zero width, height, strata or path length");
    if ((uint64_t)job->camera.width * job->camera.height > (1ull << 26))
        return sgpu_fail(SGPU_E_RANGE, std::string(who) + ": more than 2^26 pixels");
    if ((uint64_t)job->spp_x * job->spp_y > (1u << 20))
        return sgpu_fail(SGPU_E_RANGE, std::string(who) + ": more than 2^20 samples per pixel");
    if (job->max_path_length > 1024) return sgpu_fail(SGPU_E_RANGE, std::string(who) + ": path length above 1024");
    if (job->flags & ~(SGPU_DPT_SORT_HITS | SGPU_DPT_SORT_RAYS)) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": unknown flags");
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    return sgpu_dpt_render_impl(scene, job, framebuffer, as_stream(stream));
}

static int check_vcm_job(sgpu_scene scene, const sgpu_vcm_job* job, const void* framebuffer, const char* who)
{
    if (!scene || !job || !framebuffer) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": NULL argument");
    if (!scene->committed) return sgpu_fail(SGPU_E_STATE, std::string(who) + ": scene not committed");
    if (!scene->instances.empty()) {
        if (!scene->render_error.empty()) return sgpu_fail(SGPU_E_STATE, std::string(who) + ": " + scene->render_error);
    } else if (!scene->d_materials || scene->n_geom_material < scene->geom_first.size()) {
        return sgpu_fail(SGPU_E_STATE, std::string(who) + ": materials not set for every geometry");
    }
    if (scene->n_geom_material >= (1u << 27))
        return sgpu_fail(SGPU_E_RANGE, std::string(who) + ": more than 2^27 geometries");
    if (job->camera.width == 0 || job->camera.height == 0 || job->iterations == 0 || job->max_path_length == 0)
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": zero width, height, iterations or path length");
    if (job->flags & ~(SGPU_VCM_KERNEL_TIMES | SGPU_VCM_COUNT_TRAVERSAL | SGPU_VCM_ONE_SLICE_PER_WAVEFRONT |
                       SGPU_VCM_COUNT_MERGE))
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": unknown flags");
    if (!(job->base_radius > 0.0f) || !(job->base_radius < 3.0e38f))
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": base_radius must be positive and finite");
    if (!(job->radius_alpha >= 0.0f && job->radius_alpha <= 1.0f))
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": radius_alpha must lie in [0, 1]");
    if ((uint64_t)job->camera.width * job->camera.height > (1ull << 24))
        return sgpu_fail(SGPU_E_RANGE, std::string(who) + ": more than 2^24 pixels");
    if (job->max_path_length > 31) return sgpu_fail(SGPU_E_RANGE, std::string(who) + ": path length above 31");
    if ((uint64_t)job->iteration_base + job->iterations > (1ull << 30))
        return sgpu_fail(SGPU_E_RANGE, std::string(who) + ": iteration index above 2^30");
    return SGPU_OK;
}

int sgpu_render_vcm(sgpu_scene scene, const sgpu_vcm_job* job, float* framebuffer, sgpu_vcm_stats* stats, void* stream)
{
    int rc = check_vcm_job(scene, job, framebuffer, "sgpu_render_vcm");
    if (rc) return rc;
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    return sgpu_vcm_render_impl(scene, job, framebuffer, stats, as_stream(stream));
}

int sgpu_render_vcm_batch(const sgpu_scene* scenes, const sgpu_vcm_job* jobs, float* const* framebuffers,
                          uint32_t n_jobs, uint32_t flags, sgpu_vcm_stats* stats, void* stream)
{
    if (!scenes || !jobs || !framebuffers || n_jobs == 0)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_render_vcm_batch: NULL or empty argument");
    if (flags & ~(SGPU_VCM_KERNEL_TIMES | SGPU_VCM_COUNT_TRAVERSAL | SGPU_VCM_ONE_SLICE_PER_WAVEFRONT |
                  SGPU_VCM_COUNT_MERGE))
        return sgpu_fail(SGPU_E_INVALID, "sgpu_render_vcm_batch: unknown flags");
    for (uint32_t j = 0; j < n_jobs; ++j) {
        int rc = check_vcm_job(scenes[j], &jobs[j], framebuffers[j], "sgpu_render_vcm_batch");
        if (rc) return rc;
        if (jobs[j].flags != 0) return sgpu_fail(SGPU_E_INVALID, "sgpu_render_vcm_batch: job flags must be 0");
        if (scenes[j]->ctx != scenes[0]->ctx)
            return sgpu_fail(SGPU_E_INVALID, "sgpu_render_vcm_batch: scenes of different contexts");
        for (uint32_t k = 0; k < j; ++k)
            if (framebuffers[k] == framebuffers[j])
                return sgpu_fail(SGPU_E_INVALID, "sgpu_render_vcm_batch: two jobs share a framebuffer");
    }
    SGPU_HIP_TRY(hipSetDevice(scenes[0]->ctx->device));
    return sgpu_vcm_render_batch_impl(scenes, jobs, framebuffers, n_jobs, flags, stats, as_stream(stream));
}

static int check_server_config(const sgpu_vcm_server_config* cfg, const char* who)
{
    if (!cfg) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": NULL config");
    if (cfg->wavefront_paths == 0 || cfg->wavefront_paths > (1ull << 26))
        return sgpu_fail(SGPU_E_RANGE, std::string(who) + ": wavefront_paths must lie in 1..2^26");
    if (cfg->max_path_length == 0 || cfg->max_path_length > 31)
        return sgpu_fail(SGPU_E_RANGE, std::string(who) + ": max_path_length must lie in 1..31");
    // camera contribution slots (paths x (max_path_length - 2 + 2)) are numbered in 32 bits
    static_assert((1ull << 26) * 31 < (1ull << 32), "slot numbers of the largest server exceed 32 bits");
    if (cfg->flags & ~(SGPU_VCM_KERNEL_TIMES | SGPU_VCM_COUNT_TRAVERSAL | SGPU_VCM_ONE_SLICE_PER_WAVEFRONT |
                       SGPU_VCM_SHORTEST_FIRST | SGPU_VCM_COUNT_MERGE))
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": unknown flags");
    return SGPU_OK;
}

int sgpu_vcm_server_create(sgpu_context ctx, const sgpu_vcm_server_config* config, void* stream,
                           sgpu_vcm_server* out)
{
    if (!ctx || !out) return sgpu_fail(SGPU_E_INVALID, "sgpu_vcm_server_create: NULL argument");
    int rc = check_server_config(config, "sgpu_vcm_server_create");
    if (rc) return rc;
    SGPU_HIP_TRY(hipSetDevice(ctx->device));
    return sgpu_vcm_server_create_impl(ctx, config, as_stream(stream), out);
}

int sgpu_vcm_server_submit(sgpu_vcm_server server, sgpu_scene scene, const sgpu_vcm_job* job, float* framebuffer,
                           uint64_t* job_id_out, void* stream)
{
    if (!server) return sgpu_fail(SGPU_E_INVALID, "sgpu_vcm_server_submit: NULL server");
    int rc = check_vcm_job(scene, job, framebuffer, "sgpu_vcm_server_submit");
    if (rc) return rc;
    if (job->flags != 0) return sgpu_fail(SGPU_E_INVALID, "sgpu_vcm_server_submit: job flags must be 0");
    SGPU_HIP_TRY(hipSetDevice(scene->ctx->device));
    return sgpu_vcm_server_submit_impl(server, scene, job, framebuffer, job_id_out, as_stream(stream));
}

int sgpu_vcm_server_step(sgpu_vcm_server server, void* stream, uint64_t* done_ids, uint32_t max_done,
                         uint32_t* n_done, uint32_t* n_active)
{
    if (!server) return sgpu_fail(SGPU_E_INVALID, "sgpu_vcm_server_step: NULL server");
    return sgpu_vcm_server_step_impl(server, as_stream(stream), done_ids, max_done, n_done, n_active);
}

int sgpu_vcm_server_stats(sgpu_vcm_server server, sgpu_vcm_stats* stats)
{
    if (!server || !stats) return sgpu_fail(SGPU_E_INVALID, "sgpu_vcm_server_stats: NULL argument");
    return sgpu_vcm_server_stats_impl(server, stats);
}

int sgpu_vcm_server_footprint(sgpu_context ctx, const sgpu_vcm_server_config* config, uint64_t* bytes)
{
    if (!ctx || !bytes) return sgpu_fail(SGPU_E_INVALID, "sgpu_vcm_server_footprint: NULL argument");
    int rc = check_server_config(config, "sgpu_vcm_server_footprint");
    if (rc) return rc;
    *bytes = sgpu_vcm_server_bytes(ctx, config);
    return SGPU_OK;
}

int sgpu_vcm_server_destroy(sgpu_vcm_server server, void* stream)
{
    return sgpu_vcm_server_destroy_impl(server, as_stream(stream));
}

int sgpu_scene_info(sgpu_scene scene, uint32_t* n_tris, uint32_t* n_nodes, uint32_t* max_depth)
{
    if (!scene) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_info: NULL scene");
    if (!scene->committed) return sgpu_fail(SGPU_E_STATE, "sgpu_scene_info: scene not committed");
    if (n_tris) *n_tris = scene->n_tris;
    if (n_nodes) *n_nodes = scene->n_nodes;
    if (max_depth) *max_depth = scene->max_depth;
    return SGPU_OK;
}

int sgpu_scene_device_bytes(sgpu_scene scene, uint64_t* bytes)
{
    if (!scene || !bytes) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_device_bytes: NULL argument");
    if (!scene->committed) return sgpu_fail(SGPU_E_STATE, "sgpu_scene_device_bytes: scene not committed");
    uint64_t b = (uint64_t)scene->n_nodes * sizeof(BvhNode) + (uint64_t)scene->n_wide * sizeof(WideNode) +
                 (uint64_t)scene->n_tris * sizeof(TriRecord) +
                 sizeof(uint32_t) + (uint64_t)scene->n_materials * 22 * sizeof(float) +
                 (uint64_t)scene->n_geom_material * sizeof(uint32_t) + (uint64_t)scene->n_lights * sizeof(sgpu_light) +
                 (uint64_t)scene->n_geom_tex * sizeof(sgpu::PtexTex) +
                 (uint64_t)scene->ibl_w * scene->ibl_h * 4 * sizeof(float) + (uint64_t)scene->ibl_h * sizeof(float) +
                 (uint64_t)scene->n_inst * (sizeof(InstRecord) + 2 * sizeof(uint32_t)) +
                 (uint64_t)scene->n_curves * (sizeof(CurveRecord) + sizeof(BvhNode));
    for (const sgpu::PtexTex& t : scene->tex_host)
        if (t.texels) b += (uint64_t)t.nfaces * (3ull << (2 * t.res_log2)) + (uint64_t)t.nfaces * 8 * sizeof(int32_t);
    for (const auto& a : scene->attr_allocs) b += a.second;
    if (scene->d_geom_attr) b += scene->attr_host.size() * sizeof(sgpu::GeomAttr);
    *bytes = b;
    return SGPU_OK;
}

int sgpu_scene_check(sgpu_scene scene)
{
    if (!scene) return sgpu_fail(SGPU_E_INVALID, "sgpu_scene_check: NULL scene");
    if (!scene->committed) return sgpu_fail(SGPU_E_STATE, "sgpu_scene_check: scene not committed");
    uint32_t flag = 0;
    SGPU_HIP_TRY(hipMemcpy(&flag, scene->d_overflow, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (flag) return sgpu_fail(SGPU_E_RANGE, "traversal stack overflow: results of some query are incomplete");
    return SGPU_OK;
}

// ---- queries -------------------------------------------------------------------------------------

extern "C++" int sgpu_check_query(sgpu_scene scene, const sgpu_rays_soa* rays, uint32_t n, const char* who)
{
    if (!scene || !rays) return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": NULL argument");
    if (!scene->committed) return sgpu_fail(SGPU_E_STATE, std::string(who) + ": scene not committed");
    if (n > 0 && (!rays->org_x || !rays->org_y || !rays->org_z || !rays->dir_x || !rays->dir_y || !rays->dir_z ||
                  !rays->tnear || !rays->tfar))
        return sgpu_fail(SGPU_E_INVALID, std::string(who) + ": NULL ray array");
    return SGPU_OK;
}

int sgpu_intersect(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits, uint32_t n, void* stream)
{
    int rc = sgpu_check_query(scene, rays, n, "sgpu_intersect");
    if (rc) return rc;
    if (!hits) return sgpu_fail(SGPU_E_INVALID, "sgpu_intersect: NULL hits");
    if ((hits->ng_x || hits->ng_y || hits->ng_z) && !(hits->ng_x && hits->ng_y && hits->ng_z))
        return sgpu_fail(SGPU_E_INVALID, "sgpu_intersect: ng_x/ng_y/ng_z must be all set or all NULL");
    if (scene->n_curves > 0 && !hits->t)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_intersect: a scene with curves needs hits.t (the curve pass compares "
                                         "against the triangle hit)");
    return sgpu_trace_launch(scene, rays, hits, nullptr, n, false, nullptr, as_stream(stream));
}

int sgpu_occluded(sgpu_scene scene, const sgpu_rays_soa* rays, uint32_t* occluded, uint32_t n, void* stream)
{
    int rc = sgpu_check_query(scene, rays, n, "sgpu_occluded");
    if (rc) return rc;
    if (n > 0 && !occluded) return sgpu_fail(SGPU_E_INVALID, "sgpu_occluded: NULL output");
    return sgpu_trace_launch(scene, rays, nullptr, occluded, n, true, nullptr, as_stream(stream));
}

int sgpu_intersect_stats(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits, uint32_t n,
                         int any_hit, sgpu_trace_stats* stats)
{
    int rc = sgpu_check_query(scene, rays, n, "sgpu_intersect_stats");
    if (rc) return rc;
    if (!stats) return sgpu_fail(SGPU_E_INVALID, "sgpu_intersect_stats: NULL stats");
    uint64_t* d_stats = nullptr;
    uint32_t* d_occ = nullptr;
    SGPU_HIP_TRY(hipMalloc((void**)&d_stats, 4 * sizeof(uint64_t)));
    SGPU_HIP_TRY(hipMemset(d_stats, 0, 4 * sizeof(uint64_t)));
    if (any_hit) {
        hipError_t e = hipMalloc((void**)&d_occ, (size_t)(n ? n : 1) * sizeof(uint32_t));
        if (e != hipSuccess) {
            (void)hipFree(d_stats);
            return sgpu_fail(SGPU_E_NOMEM, "sgpu_intersect_stats: hipMalloc failed");
        }
    }
    sgpu_hits_soa none = {};
    rc = sgpu_trace_launch(scene, rays, any_hit ? &none : (hits ? hits : &none), d_occ, n, any_hit != 0, d_stats,
                           nullptr, nullptr, nullptr, nullptr, nullptr, 0, d_stats + 2);
    uint64_t h[4] = {0, 0, 0, 0};
    if (rc == SGPU_OK) {
        hipError_t e = hipMemcpy(h, d_stats, sizeof(h), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = sgpu_fail(SGPU_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    (void)hipFree(d_stats);
    if (d_occ) (void)hipFree(d_occ);
    if (rc) return rc;
    stats->rays = n;
    stats->nodes_visited = h[0];
    stats->tris_tested = h[1];
    stats->bytes_node = scene->ctx->trace_tree == 1 ? sizeof(WideNode) : sizeof(BvhNode);
    stats->bytes_tri = sizeof(TriRecord);
    stats->curve_nodes_visited = h[2];
    stats->curve_segments_tested = h[3];
    return SGPU_OK;
}

}  // extern "C"
