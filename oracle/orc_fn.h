/*
 * ORACLE -- test infrastructure only. Single-precision transcendentals of the sampling and shading
 * restatements, in one of two modes (orc_set_libm):
 *   0 (default) evaluated in double and rounded once to float, as the device code evaluates them.
 *     glibc's double functions are within one double ulp of the exact value, so the float result
 *     is the correctly rounded one except when the exact value lies within ~2^-29 (relative) of a
 *     float rounding boundary: the restatement then makes the same sampling decisions as the
 *     device, path for path.
 *   1 the single-precision libm functions Selas calls (std::cos, std::pow, ...), for comparisons
 *     with vectors recorded from the reference's own code. The two modes differ by at most one
 *     float ulp per call.
 */
#ifndef SGPU_ORC_FN_H
#define SGPU_ORC_FN_H

#include <math.h>

extern int orc_libm_single;

static inline float cos_f(float x) { return orc_libm_single ? cosf(x) : (float)cos((double)x); }
static inline float sin_f(float x) { return orc_libm_single ? sinf(x) : (float)sin((double)x); }
static inline float acos_f(float x) { return orc_libm_single ? acosf(x) : (float)acos((double)x); }
static inline float atan2_f(float y, float x)
{
    return orc_libm_single ? atan2f(y, x) : (float)atan2((double)y, (double)x);
}
static inline float exp_f(float x) { return orc_libm_single ? expf(x) : (float)exp((double)x); }
static inline float pow_f(float x, float y) { return orc_libm_single ? powf(x, y) : (float)pow((double)x, (double)y); }
/* x^5 (Schlick's weight): powf in mode 1, else double multiplications (the device's way) */
static inline float pow5_f(float x)
{
    if (orc_libm_single) return powf(x, 5.0f);
    double d = (double)x;
    double d2 = d * d;
    return (float)(d2 * d2 * d);
}

#endif
