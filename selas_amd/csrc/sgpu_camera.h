// Pinhole camera and correlated multi-jitter sampling, device and host.
//
// Semantics follow Selas:
//   * InitializeRayCastCamera (Source/Core/GeometryLib/Camera.cpp:89-112): horizontal fov,
//     cameraX = tan(hfov/2) * right, cameraY = tan(vfov/2) * (forward x right), right = up x forward,
//     virtualImagePlaneDistance = width / (2 tan(hfov)).
//   * ImageToWorldDirection (Camera.cpp:56-61): clip = (2x/W - 1, 1 - 2y/H),
//     dir = normalize(Z - clip.x X + clip.y Y)  (note the minus on x).
//   * CorrelatedMultiJitter (Source/Core/MathLib/CorrelatedMultiJitter.cpp:12-62), Kensler 2013,
//     "Correlated Multi-Jittered Sampling", Pixar Technical Memo 13-01: integer permutation and
//     hash jitter, bit-exact in unsigned 32-bit arithmetic.
#pragma once

#include "sgpu_math.h"

struct CameraParams {
    f3 cam_x, cam_y, cam_z;  // scaled image-plane axes and forward
    f3 position;
    f3 forward;
    float width_f, height_f;
    uint32_t width, height;
    float virtual_image_plane_distance;
};

// Host-side set-up (uses libm tanf/atanf, as Selas does; computed once per camera).
inline CameraParams camera_init(f3 position, f3 look_at, f3 up, float fov_degrees, uint32_t width, uint32_t height)
{
    const float pi = 3.141592653589793f;
    float widthf = (float)width;
    float heightf = (float)height;
    float aspect = widthf / heightf;
    float hfov = fov_degrees * (pi / 180.f);
    float h_len = tanf(hfov * 0.5f);
    float vfov = 2 * atanf(h_len / aspect);
    float v_len = tanf(vfov * 0.5f);
    f3 forward = normalize3(look_at - position);
    f3 right = normalize3(cross3(up, forward));
    CameraParams c;
    c.cam_x = h_len * right;
    c.cam_y = v_len * cross3(forward, right);
    c.cam_z = forward;
    c.position = position;
    c.forward = normalize3(look_at - position);
    c.width_f = widthf;
    c.height_f = heightf;
    c.width = width;
    c.height = height;
    c.virtual_image_plane_distance = widthf / (2.0f * tanf(hfov));
    return c;
}

SGPU_HD f3 camera_image_to_world_dir(const CameraParams& c, float x, float y)
{
    float clip_x = (x / (float)c.width) * 2.0f - 1.0f;
    float clip_y = 1.0f - 2.0f * y / (float)c.height;
    return normalize3(c.cam_z - clip_x * c.cam_x + clip_y * c.cam_y);
}

SGPU_HD float cmj_randfloat(uint32_t i, uint32_t p)
{
    i ^= p;
    i ^= i >> 17;
    i ^= i >> 10;
    i *= 0xb36534e5u;
    i ^= i >> 12;
    i ^= i >> 21;
    i *= 0x93fc4795u;
    i ^= 0xdf6e307fu;
    i ^= i >> 17;
    i *= 1u | p >> 18;
    return (float)i * (1.0f / 4294967808.0f);
}

SGPU_HD uint32_t cmj_permute(uint32_t i, uint32_t l, uint32_t p)
{
    uint32_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {
        i ^= p;
        i *= 0xe170893du;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3fu;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1u | p >> 27;
        i *= 0x6935fa69u;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303u;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3u;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfu;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}

// Sample s of an m x n pattern with permutation seed p; returns (x, y) in [0,1)^2.
SGPU_HD void cmj_sample(int32_t s, int32_t m, int32_t n, int32_t p, float& rx, float& ry)
{
    uint32_t up = (uint32_t)p;
    int sx = (int)cmj_permute((uint32_t)(s % m), (uint32_t)m, up * 0xa511e9b3u);
    int sy = (int)cmj_permute((uint32_t)(s / m), (uint32_t)n, up * 0x63d83595u);
    float jx = cmj_randfloat((uint32_t)s, up * 0xa399d265u);
    float jy = cmj_randfloat((uint32_t)s, up * 0x711ad6a5u);
    rx = ((float)(s % m) + ((float)sy + jx) / (float)n) / (float)m;
    ry = ((float)(s / m) + ((float)sx + jy) / (float)m) / (float)n;
}

// JitteredCameraRay with a CMJ offset (Camera.cpp:77-87): pixel (x, y), sample s of m x n, seed p.
SGPU_HD f3 camera_cmj_dir(const CameraParams& c, int32_t x, int32_t y, int32_t s, int32_t m, int32_t n, int32_t p)
{
    float jx, jy;
    cmj_sample(s, m, n, p, jx, jy);
    float vx = (float)x + jx;
    float vy = (float)y + jy;
    return camera_image_to_world_dir(c, vx, vy);
}
