// Procedural inputs for the ray-cast workloads (host side, deterministic from a seed).
// Selas loads assets from disk; these generators stand in for the asset pipeline so that every
// workload can be rebuilt from a seed on any machine.
#include "sgpu_camera.h"
#include "sgpu_internal.h"

#include <cmath>

namespace {

const float kPi = 3.141592653589793f;

struct Rng {
    Pcg32 s;
    explicit Rng(uint64_t seed, uint64_t stream) : s(pcg32_seed(seed, stream)) {}
    float uni() { return pcg32_float(s); }
    uint32_t below(uint32_t n) { return (uint32_t)(((uint64_t)pcg32_next(s) * n) >> 32); }
    float normal()
    {
        // Box-Muller; u1 in (0, 1]
        float u1 = 1.0f - uni();
        float u2 = uni();
        return sqrtf(-2.0f * logf(u1)) * cosf(2.0f * kPi * u2);
    }
    f3 unit_vector()
    {
        float z = 2.0f * uni() - 1.0f;
        float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float phi = 2.0f * kPi * uni();
        return mk3(r * cosf(phi), r * sinf(phi), z);
    }
};

// Orthonormal basis around n (Selas MakeOrthogonalCoordinateSystem, CoordinateSystem.h:14-21).
void ortho_basis(f3 n, f3& t, f3& b)
{
    if (fabsf(n.x) > fabsf(n.y))
        t = mk3(-n.z, 0, n.x) * (1.0f / sqrtf(n.x * n.x + n.z * n.z));
    else
        t = mk3(0, n.z, -n.y) * (1.0f / sqrtf(n.y * n.y + n.z * n.z));
    b = cross3(n, t);
}

}  // namespace

extern "C" {

int sgpu_gen_cluster_soup(uint64_t seed, uint32_t n_clusters, uint32_t tris_per_cluster, float sigma, float edge_min,
                          float edge_max, float* out)
{
    if (!out || n_clusters == 0 || tris_per_cluster == 0 || !(edge_min > 0.0f) || !(edge_max >= edge_min))
        return sgpu_fail(SGPU_E_INVALID, "sgpu_gen_cluster_soup: bad argument");
    const float lmin = logf(edge_min), lmax = logf(edge_max);
    for (uint32_t c = 0; c < n_clusters; ++c) {
        Rng rc(seed, 2ull * c + 1);
        f3 centre = mk3(rc.uni(), rc.uni(), rc.uni());
        for (uint32_t k = 0; k < tris_per_cluster; ++k) {
            Rng r(seed ^ 0x9E3779B97F4A7C15ull, (uint64_t)c * tris_per_cluster + k + 0x100000000ull);
            f3 cen = centre + mk3(sigma * r.normal(), sigma * r.normal(), sigma * r.normal());
            f3 nrm = r.unit_vector();
            f3 a, b;
            ortho_basis(nrm, a, b);
            float phi = 2.0f * kPi * r.uni();
            f3 a2 = cosf(phi) * a + sinf(phi) * b;
            f3 b2 = cosf(phi) * b - sinf(phi) * a;
            float l1 = expf(lmin + (lmax - lmin) * r.uni());
            float l2 = expf(lmin + (lmax - lmin) * r.uni());
            float th = kPi / 6.0f + (2.0f * kPi / 3.0f) * r.uni();
            f3 p0 = mk3(0, 0, 0);
            f3 p1 = l1 * a2;
            f3 p2 = l2 * (cosf(th) * a2 + sinf(th) * b2);
            f3 shift = cen - (1.0f / 3.0f) * (p0 + p1 + p2);
            f3 v[3] = {p0 + shift, p1 + shift, p2 + shift};
            float* o = out + ((size_t)c * tris_per_cluster + k) * 9;
            for (int q = 0; q < 3; ++q) {
                o[q * 3 + 0] = v[q].x;
                o[q * 3 + 1] = v[q].y;
                o[q * 3 + 2] = v[q].z;
            }
        }
    }
    return SGPU_OK;
}

int sgpu_gen_camera_rays(const float* cam_pos, const float* cam_look_at, const float* cam_up, float fov_degrees,
                         uint32_t width, uint32_t height, uint32_t spp, float* ox, float* oy, float* oz, float* dx,
                         float* dy, float* dz, float* tnear, float* tfar)
{
    if (!cam_pos || !cam_look_at || !cam_up || !ox || !oy || !oz || !dx || !dy || !dz || !tnear || !tfar ||
        width == 0 || height == 0 || spp == 0 || spp > 16)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_gen_camera_rays: bad argument (spp must be 1..16)");
    CameraParams c = camera_init(mk3(cam_pos[0], cam_pos[1], cam_pos[2]),
                                 mk3(cam_look_at[0], cam_look_at[1], cam_look_at[2]),
                                 mk3(cam_up[0], cam_up[1], cam_up[2]), fov_degrees, width, height);
    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            uint32_t pix = y * width + x;
            for (uint32_t s = 0; s < spp; ++s) {
                size_t i = (size_t)pix * spp + s;
                f3 d = camera_cmj_dir(c, (int32_t)x, (int32_t)y, (int32_t)s, 4, 4, (int32_t)pix);
                ox[i] = c.position.x;
                oy[i] = c.position.y;
                oz[i] = c.position.z;
                dx[i] = d.x;
                dy[i] = d.y;
                dz[i] = d.z;
                tnear[i] = 0.0f;
                tfar[i] = INFINITY;
            }
        }
    }
    return SGPU_OK;
}

int sgpu_gen_bounce_rays(uint64_t seed, const float* tri, uint32_t n_tris, uint32_t n_rays, float* ox, float* oy,
                         float* oz, float* dx, float* dy, float* dz, float* tnear, float* tfar)
{
    if (!tri || n_tris == 0 || !ox || !oy || !oz || !dx || !dy || !dz || !tnear || !tfar)
        return sgpu_fail(SGPU_E_INVALID, "sgpu_gen_bounce_rays: bad argument");
    const float kErr = 32.0f * 1.19209e-07f;  // Selas RayPick hit error scale (PathTracer.cpp:134)
    for (uint32_t i = 0; i < n_rays; ++i) {
        Rng r(seed, 0x200000000ull + i);
        uint32_t t = r.below(n_tris);
        const float* p = tri + (size_t)t * 9;
        f3 v0 = mk3(p[0], p[1], p[2]), v1 = mk3(p[3], p[4], p[5]), v2 = mk3(p[6], p[7], p[8]);
        float b1 = r.uni(), b2 = r.uni();
        if (b1 + b2 > 1.0f) {
            b1 = 1.0f - b1;
            b2 = 1.0f - b2;
        }
        f3 pos = v0 + b1 * (v1 - v0) + b2 * (v2 - v0);
        f3 n = normalize3(cross3(v1 - v0, v2 - v0));
        if (r.uni() < 0.5f) n = -n;
        // cosine-weighted hemisphere around n (y-up local frame, as Selas's SampleCosineWeightedHemisphere)
        float u0 = r.uni(), u1 = r.uni();
        float rr = sqrtf(u0), th = 2.0f * kPi * u1;
        f3 l = mk3(rr * cosf(th), sqrtf(fmaxf(0.0f, 1.0f - u0)), rr * sinf(th));
        f3 tx, bz;
        ortho_basis(n, tx, bz);
        f3 dir = normalize3(l.x * tx + l.y * n + l.z * bz);
        float err = kErr * fmaxf(fmaxf(fabsf(pos.x), fabsf(pos.y)), fabsf(pos.z));
        f3 org = pos + err * n;  // OffsetRayOrigin with biasScale 1 (SurfaceParameters.cpp:403-408)
        ox[i] = org.x;
        oy[i] = org.y;
        oz[i] = org.z;
        dx[i] = dir.x;
        dy[i] = dir.y;
        dz[i] = dir.z;
        tnear[i] = 0.0f;
        tfar[i] = INFINITY;
    }
    return SGPU_OK;
}

}  // extern "C"
