/*
 * selas_gpu.h -- C ABI of libselas_gpu, the MI355X (gfx950) ray-cast and render path for Selas.
 *
 * Every entry point returns an int status: 0 on success, a negative SGPU_E* code on failure.
 * Nothing throws across this boundary. sgpu_last_error() gives the message of the last failure
 * on the calling host thread.
 *
 * Device buffers passed to the query functions are device pointers (allocated with
 * sgpu_malloc, or by any HIP allocator in the same process). "stream" arguments are hipStream_t
 * values passed as void*; NULL selects the null stream.
 *
 * The entry points replace the interfaces Selas programs against (file:line into the Selas
 * source tree):
 *   - the Embree 3.2 scene side: rtcNewDevice / rtcNewScene / rtcNewGeometry(TRIANGLE) /
 *     rtcSetSharedGeometryBuffer / rtcCommitScene, as used in
 *     Source/Core/SceneLib/ModelResource.cpp:236-298,447-472 and
 *     Source/Core/SceneLib/SceneResource.cpp:218-249,286-312;
 *   - the Embree 3.2 query side: rtcIntersect1/rtcIntersect8 and rtcOccluded1/rtcOccluded8,
 *     as called by RayPick / OcclusionRay (Source/Applications/Selas/Source/PathTracer.cpp:71-139)
 *     and the batch tracers (Source/Applications/Selas/Source/DeferredPathTracer.cpp:151-280).
 */
#ifndef SELAS_GPU_H
#define SELAS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------- */
#define SGPU_OK 0
#define SGPU_E_INVALID (-22)   /* bad argument (EINVAL) */
#define SGPU_E_NOMEM (-12)     /* device or host allocation failed (ENOMEM) */
#define SGPU_E_STATE (-16)     /* object in the wrong state, e.g. query before commit (EBUSY) */
#define SGPU_E_HIP (-5)        /* a HIP runtime call failed (EIO) */
#define SGPU_E_RANGE (-34)     /* a size exceeds a documented limit (ERANGE) */

/* Embree's RTC_INVALID_GEOMETRY_ID: primID/geomID/instID of a miss. */
#define SGPU_INVALID_ID 0xFFFFFFFFu

typedef struct sgpu_context_s* sgpu_context;
typedef struct sgpu_scene_s* sgpu_scene;

/* Rays in structure-of-arrays form (device pointers, n entries each). Replaces RTCRay /
 * RTCRay8 (embree3/rtcore_ray.h:26-40): org, dir, tnear, tfar. A hit is accepted when
 * tnear < t <= tfar. */
typedef struct sgpu_rays_soa {
    const float* org_x;
    const float* org_y;
    const float* org_z;
    const float* dir_x;
    const float* dir_y;
    const float* dir_z;
    const float* tnear;
    const float* tfar;
} sgpu_rays_soa;

/* Closest-hit results in structure-of-arrays form (device pointers). Replaces RTCHit
 * (embree3/rtcore_ray.h:43-55). Any pointer may be NULL to skip that output.
 *   t       hit distance (Embree writes it to ray.tfar); tfar is written on a miss
 *   u, v    barycentric weights of vertices 1 and 2
 *   prim_id primitive index within its geometry (the quad index for sgpu_scene_add_quads
 *           geometries), SGPU_INVALID_ID on a miss
 *   geom_id geometry id returned by sgpu_scene_add_triangles, SGPU_INVALID_ID on a miss
 *   ng_*    unnormalised geometric normal (v1-v0)x(v2-v0), in the space of the hit geometry
 *           (an instanced geometry's own space, as Embree reports it)
 *   inst_id0, inst_id1  instID[0], instID[1]: the instance of the hit in the queried scene and
 *           the instance inside that one (sgpu_scene_add_instance); SGPU_INVALID_ID where there is
 *           no such level and on a miss
 * Passing rays.tfar's storage as hits.t reproduces Embree's in-place tfar update. */
typedef struct sgpu_hits_soa {
    float* t;
    float* u;
    float* v;
    uint32_t* prim_id;
    uint32_t* geom_id;
    float* ng_x;
    float* ng_y;
    float* ng_z;
    uint32_t* inst_id0;
    uint32_t* inst_id1;
} sgpu_hits_soa;

/* Per-query traversal counters (filled by sgpu_intersect_stats). */
typedef struct sgpu_trace_stats {
    uint64_t rays;
    uint64_t nodes_visited;  /* internal nodes whose child boxes were tested */
    uint64_t tris_tested;    /* triangle tests */
    uint64_t bytes_node;     /* size of one node record */
    uint64_t bytes_tri;      /* size of one triangle record */
    uint64_t curve_nodes_visited;    /* curve-tree nodes whose child boxes were tested (curve pass) */
    uint64_t curve_segments_tested;  /* curve segment tests (curve pass) */
} sgpu_trace_stats;

/* ---- errors ----------------------------------------------------------------------------- */
const char* sgpu_last_error(void);

/* ---- context (replaces rtcNewDevice / rtcReleaseDevice, Source/Applications/Selas/Source/main.cpp:99)
 * A context keeps device memory between calls: the stream-ordered pool of its device holds freed
 * memory for reuse, and the VCM wavefront buffers of the last sgpu_render_vcm(_batch) call stay
 * allocated until a call needs larger ones or the context is destroyed. */
int sgpu_context_create(int device, sgpu_context* out);
int sgpu_context_destroy(sgpu_context ctx);

/* Launch-shape parameters of a context (no counterpart in Selas; an autotuner sets them). Not to be
 * changed while renders on the context are running. */
typedef struct sgpu_tuning {
    uint64_t vcm_wavefront_paths; /* paths per wavefront of sgpu_render_vcm(_batch): 2^12..2^24
                                     (default 2^20); a larger image than this gets one slice each */
    uint32_t trace_blocks_per_cu; /* traversal grid: at most this many 256-lane blocks per CU; waves
                                     take the rays in chunks of 64 (the next chunk from a counter when
                                     one is used up), a lane the next ray of its wave's chunk when its
                                     ray is done: 1..32 (default 8) */
    uint32_t trace_tree;          /* traversal tree: 0 binary nodes (64 B), 1 8-wide compressed
                                     nodes (128 B, default); both are built at commit */
    uint32_t trace_lds_nodes;     /* 8-wide single-scene launches: this many top nodes (breadth-first)
                                     read from LDS instead of memory: 0..9 (default 0) */
    uint32_t vcm_graphs;          /* 1 (default): each VCM wavefront replays its light and camera
                                     passes as captured hipGraphs (queue counts read on the device);
                                     0: launches one by one, sized from counts read back per bounce */
    uint32_t vcm_order;           /* work order inside a VCM bounce, SGPU_ORDER_* bits (default 0:
                                     the order the queues were filled in); images do not depend on it */
    uint32_t vcm_merge_lanes;     /* lanes per vertex-merging range search: 8, 16 (default) or 32 */
    uint32_t bvh_builder;         /* binary tree the 8-wide tree is collapsed from, built at
                                     sgpu_scene_commit: 0 (default) linear BVH (Morton order,
                                     Karras 2012), 1 locally-ordered clustering (PLOC, Meister and
                                     Bittner 2018, radius 16) over the same Morton order; queries
                                     give the same results on either */
} sgpu_tuning;
/* vcm_order bits. SGPU_ORDER_MERGE_CELLS: the vertex-merging range searches of a bounce run in
 * spatial order (binned by a 12-bit Morton code of the query point in its photon grid's box, by a
 * device counting sort), each XCD taking a contiguous run of the bins. SGPU_ORDER_SHADE_MATERIAL:
 * the hits of a bounce are shaded grouped by material (the same counting sort keyed by the hit
 * geometry's material), the sort-by-material of PathTracingBatcher::GetSortedHits
 * (Source/Core/Shading/PathTracingBatcher.cpp:221-258). */
#define SGPU_ORDER_MERGE_CELLS 1u
#define SGPU_ORDER_SHADE_MATERIAL 2u
int sgpu_set_tuning(sgpu_context ctx, const sgpu_tuning* tuning);
int sgpu_get_tuning(sgpu_context ctx, sgpu_tuning* tuning);

/* The context's device: name, compute units, device memory (total and free now). */
typedef struct sgpu_device_info {
    char name[128];
    uint32_t compute_units;
    uint32_t reserved;
    uint64_t total_bytes;
    uint64_t free_bytes;
} sgpu_device_info;
int sgpu_get_device_info(sgpu_context ctx, sgpu_device_info* info);

/* ---- device memory helpers (plain hipMalloc/hipMemcpy wrappers for hosts without HIP headers) */
int sgpu_malloc(sgpu_context ctx, size_t bytes, void** dptr);
int sgpu_free(sgpu_context ctx, void* dptr);
int sgpu_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int sgpu_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int sgpu_memset(void* dst, int value, size_t bytes, void* stream);
int sgpu_stream_create(void** stream);
int sgpu_stream_destroy(void* stream);
int sgpu_stream_synchronize(void* stream);
int sgpu_device_synchronize(void);

/* Device-to-device copy by a plain streaming kernel (16 B per lane, grid-stride). Besides being a
 * copy, it is the calibration kernel for HBM byte counters: its traffic is exactly 2 x bytes. */
int sgpu_device_copy(void* dst, const void* src, size_t bytes, void* stream);

/* HIP events for timing on a given stream. */
int sgpu_event_create(void** ev);
int sgpu_event_destroy(void* ev);
int sgpu_event_record(void* ev, void* stream);
int sgpu_event_elapsed_ms(void* start, void* stop, float* ms);

/* ---- scenes --------------------------------------------------------------------------------
 * sgpu_scene_create replaces rtcNewScene (SceneResource.cpp:228 / ModelResource.cpp:447).
 * sgpu_scene_add_triangles replaces rtcNewGeometry(RTC_GEOMETRY_TYPE_TRIANGLE) +
 * rtcSetSharedGeometryBuffer(VERTEX float3, INDEX uint3) + rtcAttachGeometry
 * (ModelResource.cpp:260-298). Unlike Embree's shared buffers, the host arrays are copied; the
 * caller may free them on return. An optional column-major 4x4 transform (Embree's
 * RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR, SubsceneResource.cpp:72) places the mesh in world space;
 * the transformed vertices are what the BVH is built over (a flattened instance).
 * sgpu_scene_commit replaces rtcCommitScene (ModelResource.cpp:468): the device builds the BVH. */
int sgpu_scene_create(sgpu_context ctx, sgpu_scene* out);
int sgpu_scene_add_triangles(sgpu_scene scene, const float* positions, uint32_t n_verts,
                             const uint32_t* indices, uint32_t n_tris,
                             const float* colmajor4x4, uint32_t* geom_id_out);
/* Quad mesh (4 vertex indices per quad, v0 v1 v2 v3 counter-clockwise): replaces
 * rtcNewGeometry(RTC_GEOMETRY_TYPE_QUAD) (ModelResource.cpp:273). As Embree does, each quad is
 * intersected as the triangles (v0, v1, v3) and (v2, v3, v1); a hit reports the quad index as
 * prim_id and (u, v) as the quad's own coordinates (u along v0->v1, v along v0->v3: the second
 * triangle's barycentrics (u', v') become (1 - u', 1 - v')), the coordinates Ptex faces use. */
int sgpu_scene_add_quads(sgpu_scene scene, const float* positions, uint32_t n_verts, const uint32_t* indices,
                         uint32_t n_quads, const float* colmajor4x4, uint32_t* geom_id_out);
/* Round B-spline curves: replaces rtcNewGeometry(RTC_GEOMETRY_TYPE_ROUND_BSPLINE_CURVE) +
 * rtcSetSharedGeometryBuffer(INDEX uint, VERTEX float4) + rtcAttachGeometry (InitializeCurves,
 * ModelResource.cpp:300-330). vertices: n_verts control points (x, y, z, radius); indices: per segment
 * the first of its four consecutive control points. A segment is the uniform cubic B-spline through
 * them and its surface the sphere swept along it (end spheres included); a hit reports prim_id = the
 * segment, u = the curve parameter in [0, 1], v = 0 and ng = the surface normal (the hit point minus
 * the curve point at u); only entries are hits (a ray starting inside a curve's tube does not hit that
 * tube). Curves take no vertex attributes and no per-face texture. The optional placement
 * (colmajor4x4) must be a similarity (rotation, uniform scale, translation): radii scale with it;
 * an instance placing a scene with curves must be one too. Queries on a scene with curves need
 * hits.t (the curve pass after the triangle traversal compares against it); sgpu_intersect_stats
 * counts the curve pass apart (curve_nodes_visited, curve_segments_tested; bounded by the rays' tfar
 * when it is given no hits). Curves count as one geometry id each call, like meshes. */
int sgpu_scene_add_curves(sgpu_scene scene, const float* vertices, uint32_t n_verts, const uint32_t* indices,
                          uint32_t n_segments, const float* colmajor4x4, uint32_t* geom_id_out);
/* Per-vertex shading attributes of geometry geom_id (host arrays, copied; before commit): replaces
 * rtcSetGeometryVertexAttributeCount + rtcSetSharedGeometryBuffer(RTC_BUFFER_TYPE_VERTEX_ATTRIBUTE)
 * with slot 0 = normals (FLOAT3), slot 1 = tangents (FLOAT4: xyz and the bitangent sign w), slot 2 =
 * UVs (FLOAT2) (ModelResource.cpp:164-200), which CalculateSurfaceParams interpolates with
 * rtcInterpolate0 at the hit (SurfaceParameters.cpp:155-190): the renderers then shade with the
 * normalised interpolated normal and the frame (t, n, cross(n, t) * w) of the interpolated tangent.
 * n_verts must equal the vertex count the geometry was added with; the geometry's transform
 * (colmajor4x4) is applied to the normals and the tangents' xyz. Tangents need normals, UVs need
 * normals and tangents (as ModelResource.cpp asserts); normals == NULL removes the attributes. UVs
 * are kept for Selas's 2D texture lookups, which this library does not implement (its base colour
 * textures are per-face Ptex, sgpu_scene_set_ptex): they do not change the image. */
int sgpu_scene_set_vertex_attributes(sgpu_scene scene, uint32_t geom_id, const float* normals, const float* tangents,
                                     const float* uvs, uint32_t n_verts);
/* Instance of a committed scene `child`, placed by a column-major 4x4 local-to-world transform
 * (an affine one: the last row is ignored): replaces rtcNewGeometry(RTC_GEOMETRY_TYPE_INSTANCE) +
 * rtcSetGeometryInstancedScene + rtcSetGeometryTransform(RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR)
 * (SubsceneResource.cpp:61-78, a model instanced in a subscene) and Selas's user-geometry
 * subscene instances (SceneResource.cpp:87-169, 220-249, a subscene instanced in the scene). A
 * scene holds either geometries or instances; a child holds geometries or instances of scenes
 * with geometries, so a query crosses at most two instance levels (MaxInstanceLevelCount_ = 2),
 * reported as inst_id0 and inst_id1 (SceneResource.cpp:136-137). Instance ids count from 0 in
 * call order. The two levels are composed at commit (in double precision, rounded once) into one
 * level of instances with world-to-local transforms, over which the device builds the top tree; a
 * ray is transformed into the geometry's space, where the triangle test runs, and t is the world
 * ray's parameter. The child must stay alive while an instance refers to it (its release fails
 * with SGPU_E_STATE). Instanced scenes answer queries (sgpu_intersect, sgpu_occluded,
 * sgpu_intersect_stats) and are rendered by every renderer: at commit the placed scenes' materials,
 * textures and vertex attributes are gathered (set them before committing the scene that places
 * them). */
int sgpu_scene_add_instance(sgpu_scene scene, sgpu_scene child, const float* colmajor4x4, uint32_t* inst_id_out);
int sgpu_scene_commit(sgpu_scene scene, void* stream);
int sgpu_scene_release(sgpu_scene scene);
/* Number of triangles, BVH nodes and maximum leaf depth of a committed scene (of an instanced
 * scene: 0 triangles, and the top tree's nodes and depth over its composed instances). */
int sgpu_scene_info(sgpu_scene scene, uint32_t* n_tris, uint32_t* n_nodes, uint32_t* max_depth);
/* Device memory the scene holds (tree, triangles, shading records, lights, textures), in bytes:
 * what keeping it resident costs. No counterpart in Selas (its scenes live in host memory). */
int sgpu_scene_device_bytes(sgpu_scene scene, uint64_t* bytes);
/* Synchronous health check after queries: SGPU_E_RANGE if any traversal on this scene ran out of
 * stack (the stack is sized from the tree depth, so this is never expected), else 0. */
int sgpu_scene_check(sgpu_scene scene);

/* ---- queries --------------------------------------------------------------------------------
 * sgpu_intersect replaces rtcIntersect1/8 (PathTracer.cpp:95-139, DeferredPathTracer.cpp:151-231):
 * closest hit of each ray in (tnear, tfar]. Ties in t go to the smaller (geomID, primID).
 * sgpu_occluded replaces rtcOccluded1/8 (PathTracer.cpp:71-92, DeferredPathTracer.cpp:234-280):
 * occluded[i] = 1 when any triangle is hit in (tnear, tfar], else 0 (Embree signals the same
 * by tfar = -inf). */
int sgpu_intersect(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits,
                   uint32_t n, void* stream);
int sgpu_occluded(sgpu_scene scene, const sgpu_rays_soa* rays, uint32_t* occluded, uint32_t n,
                  void* stream);
/* sgpu_intersect without any device allocation. No counterpart in Selas or Embree: for callers that
 * cannot allocate device memory (code under HIP stream capture, tight serving loops, hosts that
 * forbid hipMalloc). The traversal stacks' spill area, which sgpu_intersect allocates per call, is
 * a fixed 64 MiB array in device memory instead, one per GPU (looked up at sgpu_context_create and
 * held until the process ends); the grid is capped so that its spill fits (rays beyond it are taken
 * grid-stride). Results equal sgpu_intersect's. Fails with SGPU_E_RANGE when even one block's
 * stacks would not fit (a scene whose stack bound, sgpu_scene_info's depth x 7, exceeds 65,552
 * entries). Calls that overlap on one GPU (other streams, other threads) share the array and race:
 * order them. */
int sgpu_intersect_noalloc(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits, uint32_t n,
                           void* stream);
/* Same query as sgpu_intersect with per-ray node/triangle counters, summed into *stats.
 * Synchronous. Used to state the algorithmic bytes per ray. */
int sgpu_intersect_stats(sgpu_scene scene, const sgpu_rays_soa* rays, const sgpu_hits_soa* hits,
                         uint32_t n, int any_hit, sgpu_trace_stats* stats);

/* ---- shading ----------------------------------------------------------------------------------
 * Batched replacements of EvaluateBsdf / SampleBsdfFunction (Source/Core/Shading/SurfaceScattering.cpp:18-60,
 * dispatching to EvaluateDisney / SampleDisney, Disney.cpp:525,617). All pointers are device pointers.
 * A surface is 22 floats (the material part of Selas's SurfaceParameters, SurfaceParameters.h:18-49):
 *   [0..2] baseColor (linear), [3..5] transmittanceColor, [6] sheen, [7] sheenTint, [8] clearcoat,
 *   [9] clearcoatGloss, [10] metallic, [11] specTrans, [12] diffTrans (already halved, as
 *   CalculateSurfaceParams stores it), [13] flatness, [14] anisotropic, [15] relativeIOR,
 *   [16] specularTint, [17] roughness, [18] scatterDistance, [19] ior,
 *   [20] shader (0 = DisneyThin, 1 = DisneySolid, 2 = Lambert), [21] material flags.
 * normals: 3 floats per point, the unit shading normal; the tangent frame is built from it as
 * CalculateSurfaceParams does (MakeOrthogonalCoordinateSystem, y-up tangent space).
 * v, l: unit world-space directions (v toward the viewer, l toward the light), 3 floats each.
 * sgpu_bsdf_eval out (5 per point): f*|cos theta_l| rgb, forward pdf, reverse pdf.
 * sgpu_bsdf_sample: uniforms_per_sample uniforms in [0,1) per point, consumed in the order Selas's
 * CSampler would deliver them; out (14 per point): success, event flags (1 scatter, 2 transmission,
 * 4 Dirac), reflectance rgb, wi xyz, forward pdf, reverse pdf, medium (0 vacuum, 1 isotropic),
 * extinction rgb. */
int sgpu_bsdf_eval(const float* surfaces, const float* normals, const float* v, const float* l,
                   uint32_t count, float* out, void* stream);
int sgpu_bsdf_sample(const float* surfaces, const float* normals, const float* v, const float* uniforms,
                     uint32_t uniforms_per_sample, uint32_t count, float* out, void* stream);

/* ---- scene shading data ------------------------------------------------------------------------
 * A light, with the fields of Selas's SceneLight (Source/Core/SceneLib/SceneResource.h:32-40).
 * type 0: quad light centred at `position`, edges `x` and `z` (full lengths), emitting `radiance`
 * into the half-space `direction` points to; it is not intersectable geometry, as in Selas.
 * type 1: point light at `position` with intensity `radiance` (an addition; Selas has none). */
typedef struct sgpu_light {
    uint32_t type;
    float position[3];
    float direction[3];
    float x[3];
    float z[3];
    float radiance[3];
} sgpu_light;

/* Materials of a scene (host arrays, copied): 22 floats per material with the surface-block layout
 * above, holding the material's own values (MaterialResourceData, Source/Core/SceneLib/ModelResource.h:68-88):
 * the renderer applies CalculateSurfaceParams's transforms (baseColor^2.2, saturated metallic and
 * specTrans, diffTrans x 0.5, relative IOR per hit); slot [15] is unused and slot [21] holds the
 * material flags (1 = transparent, 16 = per-face texture). geom_material[g] is the material index
 * of geometry g (one entry per geometry of the scene). */
int sgpu_scene_set_materials(sgpu_scene scene, const float* materials, uint32_t n_materials,
                             const uint32_t* geom_material, uint32_t n_geoms);
/* Lights of a scene (host array, copied; one light set holding all of them) and the constant
 * background radiance (SceneResourceData::backgroundIntensity, SceneResource.h:53). */
int sgpu_scene_set_lights(sgpu_scene scene, const sgpu_light* lights, uint32_t n_lights, const float* background_rgb);

/* Image-based light of a scene: replaces ImageBasedLightResource (ImageBasedLightResource.h:20-40)
 * with the sampling tables BuildImageBasedLight.cpp:18-103 computes (CDFs over sin(theta) x CIE
 * luma), built here from the texels. rgb: height x width x 3 floats (host array, copied), an
 * equirectangular map whose row 0 is at +y; every value finite and >= 0; width, height in
 * [1, 16384]. rotation_radians is added to phi (normalized to [0, 2 pi)); exposure scales the
 * radiance (exposureScale). rgb == NULL removes the light. Emission starts on the disc of the
 * sphere around the triangles' bounding box, so the scene must be committed first. VCM picks it
 * like one more light (emission, NEE, and camera rays that leave the scene); the path tracer adds
 * it where camera paths leave the scene (EvaluateBackground, AreaLighting.cpp:415). It replaces
 * the constant background. */
int sgpu_scene_set_ibl(sgpu_scene scene, uint32_t width, uint32_t height, const float* rgb, float rotation_radians,
                       float exposure);

/* Per-face texture of a quad or triangle geometry: replaces TextureCache::FetchPtex +
 * PtexTexture::open (TextureCache.cpp:168, SurfaceParameters.cpp:197) for 8-bit RGB Ptex data with
 * one resolution for every face. texels: n_faces x (1 << res_log2)^2 x 3 bytes (face-major, rows
 * along v); adjfaces / adjedges: 4 per face, the adjacent face (-1 = border) and that face's edge id
 * of the shared edge (Ptex FaceInfo, Ptexture.h:240-300). Quad geometry: edges in Ptex's quad order
 * (bottom, right, top, left), n_faces = the quad count, filtered with Ptex's cubic B-spline filter
 * (PtexSeparableFilter). Triangle geometry: edges 0..2 in Ptex's triangle order (edge e from corner
 * e to corner e + 1: v = 0, w = 0, u = 0; the fourth entry is ignored), n_faces = the triangle
 * count, filtered with Ptex's triangle filter (PtexTriangleFilter, the filter getFilter returns for
 * triangle meshes, PtexFilters.cpp:436-440) at the hit's barycentrics. res_log2 <= 8. Host arrays,
 * copied. A material with flag 16 (eUsesPtex) takes its base colour from its geometry's texture. */
int sgpu_scene_set_ptex(sgpu_scene scene, uint32_t geom_id, uint32_t res_log2, uint32_t n_faces,
                        const uint8_t* texels, const int32_t* adjfaces, const int32_t* adjedges);

/* PtexFilter{f_bspline}.eval(out, 0, 3, face, u, v, 0, 0, 0, 0) (Ptexture.h:989-1010, as
 * SurfaceParameters.cpp:199-203 calls it) for n lookups on the texture of geometry geom_id:
 * faces, u, v and out (3 floats per lookup) are device arrays. */
int sgpu_ptex_eval(sgpu_scene scene, uint32_t geom_id, const int32_t* faces, const float* u, const float* v,
                   uint32_t n, float* out, void* stream);

/* ---- integrators -------------------------------------------------------------------------------- */
typedef struct sgpu_camera {
    float position[3];
    float look_at[3];
    float up[3];
    float fov_degrees; /* horizontal field of view (Selas CameraSettings::fovDegrees) */
    uint32_t width;
    uint32_t height;
} sgpu_camera;

#define SGPU_PT_NEE_COSINE_ONCE 1u /* drop the extra Dot(normal, l) of Selas's quad-light estimator */

typedef struct sgpu_pt_job {
    sgpu_camera camera;
    uint32_t spp;             /* paths per pixel (PathsPerPixel_) */
    uint32_t max_path_length; /* bounces (MaxBounceCount_); Russian roulette after 8 */
    uint64_t seed;            /* PCG32 seed; path i uses stream i */
    uint32_t flags;           /* SGPU_PT_* */
} sgpu_pt_job;

/* Unidirectional path tracing: replaces PathTracer::GenerateImage / EvaluatePath
 * (Source/Applications/Selas/Source/PathTracer.cpp:142-348). framebuffer is a device array of
 * 2 x height x width x 3 floats: layer 0 = first-bounce reflectance, layer 1 = radiance, both
 * averaged over spp. Synchronous on `stream`. */
int sgpu_render_pt(sgpu_scene scene, const sgpu_pt_job* job, float* framebuffer, void* stream);

#define SGPU_PT_COUNT_TRAVERSAL 2u /* sgpu_render_pt_stats: count nodes and triangles of closest-hit rays */

/* Work and device time of one sgpu_render_pt_stats call (no counterpart in Selas). */
typedef struct sgpu_pt_stats {
    uint64_t closest_rays;   /* closest-hit rays: camera rays and continuations */
    uint64_t shadow_rays;    /* occlusion rays (NEE) */
    uint64_t hits;           /* closest-hit rays that hit a surface (the rest are misses) */
    uint64_t bounces;        /* closest-hit launches (= shading launches) */
    uint64_t nodes_visited;  /* SGPU_PT_COUNT_TRAVERSAL: over all closest-hit rays */
    uint64_t tris_tested;    /* SGPU_PT_COUNT_TRAVERSAL: over all closest-hit rays */
    float closest_ms;        /* device time of the closest-hit traversals (events around each launch) */
    float shadow_ms;         /* ... of the occlusion traversals */
    float shade_ms;          /* ... of the shading kernels (k_pt_shade) */
    float other_ms;          /* ... of path generation, shadow resolution and accumulation */
} sgpu_pt_stats;

/* sgpu_render_pt with per-phase device times (HIP events recorded around every launch on `stream`)
 * and ray counts; job->flags may add SGPU_PT_COUNT_TRAVERSAL. Same image as sgpu_render_pt. */
int sgpu_render_pt_stats(sgpu_scene scene, const sgpu_pt_job* job, float* framebuffer, sgpu_pt_stats* stats,
                         void* stream);

#define SGPU_DPT_SORT_HITS 1u /* shade each bounce's hits in (instID[0], instID[1], geomID, primID) order */
#define SGPU_DPT_SORT_RAYS 2u /* trace each bounce's rays binned by the axis and sign of their direction's largest
                                 component (DetermineRayCategory's bins), each bin ordered by origin along its axis
                                 (GetSortedBatch; the reference orders partitions of 64 rays or fewer by direction) */

typedef struct sgpu_dpt_job {
    sgpu_camera camera;
    uint32_t spp_x, spp_y;    /* CMJ strata per pixel (SamplesPerPixelX_ x SamplesPerPixelY_) */
    uint32_t max_path_length; /* bound on segments per path (Selas ends paths by Russian roulette
                                 alone, after MaxTrackedBounces_ = 7) */
    uint64_t seed;            /* PCG32 seed; path i uses stream i */
    uint32_t flags;           /* SGPU_DPT_* */
} sgpu_dpt_job;

/* The deferred path tracer: replaces DeferredPathTracer::GenerateImage
 * (Source/Applications/Selas/Source/DeferredPathTracer.cpp:388-430) with its ShadeHitPosition
 * estimator (:61-148): CMJ primary rays, NEE toward the quad and point lights, a sky sample (the
 * image-based light, else the constant background) and the BSDF's continuation weighted against
 * each other by the balance heuristic, Russian roulette after 7 bounces. SGPU_DPT_SORT_HITS orders
 * the shading as PathTracingBatcher::GetSortedHits (PathTracingBatcher.cpp:221-258, 799-830) does;
 * SGPU_DPT_SORT_RAYS orders each bounce's continuation rays as GetSortedBatch (:705-750) does; the
 * image depends on neither. framebuffer: a device array of height x width x 3 floats (one
 * layer, OutputLayers_ = 1), averaged over spp_x x spp_y. Synchronous on `stream`. */
int sgpu_render_deferred_pt(sgpu_scene scene, const sgpu_dpt_job* job, float* framebuffer, void* stream);

#define SGPU_VCM_KERNEL_TIMES 1u    /* time each phase with events (sgpu_vcm_stats.*_ms) */
#define SGPU_VCM_COUNT_TRAVERSAL 2u /* count nodes and triangles visited by closest-hit rays (slower) */
#define SGPU_VCM_ONE_SLICE_PER_WAVEFRONT 4u /* run each iteration of each job as its own wavefront (the
                                               un-batched schedule, for comparisons) */
#define SGPU_VCM_SHORTEST_FIRST 8u /* servers only: fill each wavefront with the admitted jobs of least
                                      remaining work (paths x iterations left) first, ties in admission
                                      order, instead of in admission order */
#define SGPU_VCM_COUNT_MERGE 16u /* count the vertex-merging work: camera vertices read, range searches,
                                    photons visited and photons within the radius (slower) */

typedef struct sgpu_vcm_job {
    sgpu_camera camera;
    uint32_t iterations;      /* VCM iterations: one light path and one camera path per pixel each */
    uint32_t iteration_base;  /* global index of the first iteration (radius schedule and streams) */
    uint32_t max_path_length; /* maxPathLength, 1..31 */
    float base_radius;        /* merge radius of iteration 0 (vmSearchRadius) */
    float radius_alpha;       /* radius shrink exponent (vmRadiusAlpha) */
    uint32_t flags;           /* SGPU_VCM_* */
    uint64_t seed;            /* PCG32 seed; path p of iteration i, pass s uses stream (i << 33 | p << 1 | s) */
} sgpu_vcm_job;

typedef struct sgpu_vcm_stats {
    uint64_t rays;             /* closest-hit and shadow rays traced */
    uint64_t closest_rays;     /* closest-hit rays (light and camera subpath segments) */
    uint64_t shadow_rays;      /* occlusion rays (NEE, vertex connections, light-tracing splats) */
    uint64_t closest_launches; /* closest-hit traversal launches */
    uint64_t light_vertices;   /* light vertices stored, summed over iterations */
    uint64_t samples;          /* camera paths = width x height x iterations */
    uint64_t wavefronts;       /* wavefronts the iterations were packed into */
    uint64_t nodes_visited;    /* SGPU_VCM_COUNT_TRAVERSAL: over all closest-hit rays */
    uint64_t tris_tested;      /* SGPU_VCM_COUNT_TRAVERSAL: over all closest-hit rays */
    float closest_ms;          /* SGPU_VCM_KERNEL_TIMES: device time of the closest-hit traversals */
    float shadow_ms;           /* ... of the occlusion traversals */
    float light_ms;            /* ... of the light-subpath kernels (emission, shading, splats) */
    float camera_ms;           /* ... of the camera-subpath kernels (NEE, connections, continuation) */
    float grid_ms;             /* ... of the vertex compaction and photon grid build */
    float merge_ms;            /* ... of the vertex merging (photon range searches) */
    uint64_t merge_launches;   /* vertex-merging launches (camera bounces with rays) */
    uint64_t merge_entries;    /* SGPU_VCM_COUNT_MERGE: camera vertices read by the merging kernel */
    uint64_t merge_queries;    /* ... of them range searches (a merging vertex inside its slice's photon box) */
    uint64_t merge_visited;    /* ... photons whose positions were tested (the 2x2x2 cells' ranges) */
    uint64_t merge_near;       /* ... photons within the radius (their vertex records read) */
} sgpu_vcm_stats;

/* Writes each layer of a device framebuffer (layers x height x width x 3 floats, the layout the
 * renderers write) to "<path_prefix>_<layer>.hdr": replaces FrameBuffer_Save
 * (Source/Core/TextureLib/Framebuffer.cpp:43-59, stb_image_write's HDR writer through
 * StbImageWrite, StbImage.cpp:64-78). Radiance RGBE, run-length coded scanlines (flat for widths
 * below 8 or above 32767), top row first. The RGBE conversion runs on the device; synchronous on
 * `stream`. */
int sgpu_framebuffer_save_hdr(const char* path_prefix, const float* framebuffer, uint32_t width, uint32_t height,
                              uint32_t layers, void* stream);

/* Vertex connection and merging: replaces VCM::GenerateImage and its helpers
 * (Source/Applications/Selas/Source/VCM.h:18, VCM.cpp:669-760, VCMCommon.cpp:37-113,
 * VCMHashGrid.cpp:15-157) for quad and point lights (sgpu_scene_set_lights) and the image-based
 * light (sgpu_scene_set_ibl; EmitIblLightSample, DirectIblLightSample, ConnectToSkyLight,
 * AreaLighting.cpp:246-331, VCM.cpp:171-186), the lights picked uniformly. framebuffer is a
 * device array of height x width x 3 floats, the radiance averaged over the iterations. The merge
 * radius of global iteration k is base_radius / (k + 1)^((1 - radius_alpha) / 2). stats may be
 * NULL. Synchronous on `stream`. */
int sgpu_render_vcm(sgpu_scene scene, const sgpu_vcm_job* job, float* framebuffer, sgpu_vcm_stats* stats,
                    void* stream);

/* Batched form: n_jobs render jobs of any size, each on its own scene (scenes may repeat), run as
 * one schedule. A job's iterations are slices; slices of all jobs are packed in job order into
 * wavefronts of about 2^20 paths, so small jobs share launches and large ones fill the chip; rays
 * of different scenes share traversal launches. Each framebuffers[j] receives exactly what
 * sgpu_render_vcm(scenes[j], &jobs[j], ...) writes. flags: SGPU_VCM_* for the whole batch (the jobs'
 * own flags must be 0); stats (may be NULL) sums the batch. Synchronous on `stream`. */
int sgpu_render_vcm_batch(const sgpu_scene* scenes, const sgpu_vcm_job* jobs, float* const* framebuffers,
                          uint32_t n_jobs, uint32_t flags, sgpu_vcm_stats* stats, void* stream);

/* A persistent VCM executor for continuous batching (request loops): wavefront buffers are sized
 * once; jobs are admitted at any time (sgpu_vcm_server_submit) and every sgpu_vcm_server_step runs
 * one wavefront over the next iterations of the admitted jobs in admission order (the oldest job's
 * iterations first, then the next job's, up to wavefront_paths paths; with SGPU_VCM_SHORTEST_FIRST
 * the jobs of least remaining work first), retiring the jobs whose last iteration ran. A job's framebuffer (device, height x width x 3 floats, owned by the caller and
 * untouched by anything else until the job is retired) then holds exactly what sgpu_render_vcm
 * writes for it. Scenes must stay alive while jobs on them are admitted. Calls on one server must
 * not overlap; submit and step on the same stream, or order the streams. This has no counterpart
 * in Selas (its renders are one job at a time); sgpu_render_vcm_batch is built on it. */
typedef struct sgpu_vcm_server_s* sgpu_vcm_server;

typedef struct sgpu_vcm_server_config {
    uint64_t wavefront_paths;  /* paths per wavefront; also the largest admissible image (W x H) */
    uint32_t max_path_length;  /* largest path length of the admitted jobs, 1..31 */
    uint32_t flags;            /* SGPU_VCM_* */
} sgpu_vcm_server_config;

int sgpu_vcm_server_create(sgpu_context ctx, const sgpu_vcm_server_config* config, void* stream,
                           sgpu_vcm_server* out);
/* Admits a job (its flags must be 0); *job_id_out (may be NULL) receives its id. */
int sgpu_vcm_server_submit(sgpu_vcm_server server, sgpu_scene scene, const sgpu_vcm_job* job, float* framebuffer,
                           uint64_t* job_id_out, void* stream);
/* Runs one wavefront (nothing when no job is admitted); writes up to max_done ids of the jobs it
 * retired to done_ids, their count to *n_done and the number still admitted to *n_active (any of
 * the three may be NULL). Synchronous on `stream`. */
int sgpu_vcm_server_step(sgpu_vcm_server server, void* stream, uint64_t* done_ids, uint32_t max_done,
                         uint32_t* n_done, uint32_t* n_active);
/* Totals since creation. */
int sgpu_vcm_server_stats(sgpu_vcm_server server, sgpu_vcm_stats* stats);
/* Device bytes a server with this configuration allocates (admission planning). */
int sgpu_vcm_server_footprint(sgpu_context ctx, const sgpu_vcm_server_config* config, uint64_t* bytes);
int sgpu_vcm_server_destroy(sgpu_vcm_server server, void* stream);

/* ---- procedural inputs (host side, deterministic from a seed) ------------------------------
 * Gaussian-cluster triangle soup: n_clusters x tris_per_cluster triangles. Cluster centres are
 * uniform in [0,1]^3, triangle centroids normal around them with sigma, two edge lengths
 * log-uniform in [edge_min, edge_max], random orientation. Writes 9 floats per triangle
 * (v0,v1,v2) to positions_out (host). */
int sgpu_gen_cluster_soup(uint64_t seed, uint32_t n_clusters, uint32_t tris_per_cluster, float sigma,
                          float edge_min, float edge_max, float* positions_out);
/* Primary camera rays of a width x height pinhole camera (Selas InitializeRayCastCamera +
 * JitteredCameraRay with a 4x4 correlated multi-jitter pattern, Camera.cpp:63-87,89-112),
 * samples 0..spp-1 of each pixel, pixel-major. Writes 8 floats per ray into the SoA arrays
 * (host pointers, n = width*height*spp). */
int sgpu_gen_camera_rays(const float* cam_pos, const float* cam_look_at, const float* cam_up,
                         float fov_degrees, uint32_t width, uint32_t height, uint32_t spp,
                         float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                         float* tnear, float* tfar);
/* Cosine-distributed rays leaving uniformly chosen points of uniformly chosen triangles of a
 * soup (9 floats per triangle), on a random side, offset like Selas's OffsetRayOrigin. */
int sgpu_gen_bounce_rays(uint64_t seed, const float* tri_positions, uint32_t n_tris, uint32_t n_rays,
                         float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                         float* tnear, float* tfar);

#ifdef __cplusplus
}
#endif

#endif /* SELAS_GPU_H */
