// Small float3 toolkit shared by the HIP kernels and the host-side generators of libselas_gpu.
//
// Conventions follow Selas's MathLib:
//   * tangent space is y-up: CosTheta(w) = w.y (Source/Core/MathLib/Trigonometric.h:51-110);
//   * Reflect(n, l) = 2 dot(n,l) n - l (Source/Core/MathLib/FloatFuncs.h:348);
//   * Normalize multiplies by 1/Length (Source/Core/MathLib/FloatFuncs.h:401).
//
// Every file that includes this header is compiled with -ffp-contract=off, so the only fused
// multiply-adds are the explicit fmaf() calls below. The triangle and box tests depend on that:
// their arithmetic is specified operation by operation and restated on the CPU in oracle/.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define SGPU_HD __host__ __device__ __forceinline__

struct f3 {
    float x, y, z;
};

SGPU_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
SGPU_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
SGPU_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
SGPU_HD f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
SGPU_HD f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
SGPU_HD f3 operator*(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
SGPU_HD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
SGPU_HD f3& operator+=(f3& a, f3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }

// Plain (unfused) dot and cross, evaluated left to right like Selas's FloatFuncs.
SGPU_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SGPU_HD f3 cross3(f3 a, f3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
SGPU_HD float length3(f3 a) { return sqrtf(dot3(a, a)); }
SGPU_HD f3 normalize3(f3 a)
{
    float inv = 1.0f / length3(a);
    return mk3(a.x * inv, a.y * inv, a.z * inv);
}
SGPU_HD float maxc3(f3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
SGPU_HD bool is_zero3(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// Fused forms used by the intersection arithmetic. Each is a fixed sequence of IEEE operations:
//   cross_fma(a,b).x = fma(a.y, b.z, -(a.z*b.y)), cyclically for y and z;
//   dot_fma(a,b)     = fma(a.x, b.x, fma(a.y, b.y, a.z*b.z)).
SGPU_HD f3 cross_fma(f3 a, f3 b)
{
    return mk3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
SGPU_HD float dot_fma(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

// PCG32 (O'Neill 2014, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
// Algorithms for Random Number Generation", pcg32_random_r / pcg32_srandom_r).
struct Pcg32 {
    uint64_t state;
    uint64_t inc;
};

SGPU_HD uint32_t pcg32_next(Pcg32& r)
{
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

SGPU_HD Pcg32 pcg32_seed(uint64_t initstate, uint64_t initseq)
{
    Pcg32 r;
    r.state = 0u;
    r.inc = (initseq << 1u) | 1u;
    pcg32_next(r);
    r.state += initstate;
    pcg32_next(r);
    return r;
}

// Single-precision transcendentals of the sampling and shading code, evaluated in double and
// rounded once to float. The double results of the device library and of the host's libm are both
// within one double ulp of the exact value, so their float roundings agree except when the exact
// value lies within ~2^-29 (relative) of a float rounding boundary: host and device then make the
// same sampling decisions and hit the same triangles, path for path. The float values are within
// one float ulp of the single-precision libm functions Selas calls (std::cos, std::pow, ...).
SGPU_HD float cos_f(float x) { return (float)cos((double)x); }
SGPU_HD float sin_f(float x) { return (float)sin((double)x); }
SGPU_HD float acos_f(float x) { return (float)acos((double)x); }
SGPU_HD float atan2_f(float y, float x) { return (float)atan2((double)y, (double)x); }
SGPU_HD float pow_f(float x, float y) { return (float)pow((double)x, (double)y); }
// A vector through a row-major 3x3 matrix m (the linear part of an instance's local-to-world
// transform: how Selas carries hit normals to world space, MatrixMultiplyVector(normal,
// localToWorld), SurfaceParameters.cpp:167), by fused multiply-adds from the last term.
SGPU_HD f3 xform_linear(const float* m, f3 v)
{
    return mk3(fmaf(m[0], v.x, fmaf(m[1], v.y, m[2] * v.z)), fmaf(m[3], v.x, fmaf(m[4], v.y, m[5] * v.z)),
               fmaf(m[6], v.x, fmaf(m[7], v.y, m[8] * v.z)));
}
// x^5 (Schlick's weight), by double multiplications: the same bits on host and device.
SGPU_HD float pow5_f(float x)
{
    double d = (double)x;
    double d2 = d * d;
    return (float)(d2 * d2 * d);
}

// Uniform float in [0, 1): the top 24 bits scaled by 2^-24 (exact, never returns 1.0).
SGPU_HD float pcg32_float(Pcg32& r) { return (float)(pcg32_next(r) >> 8) * (1.0f / 16777216.0f); }
