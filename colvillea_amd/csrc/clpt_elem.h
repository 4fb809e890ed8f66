/*
 * clpt_elem.h -- float32 elementary functions written from +, -, *, /, sqrt
 * and exact rounding primitives only, so that a CPU build (gcc,
 * -ffp-contract=off) and a gfx950 build (hipcc, -ffp-contract=off,
 * correctly rounded divide/sqrt) return the same bits for the same input.
 *
 * The renderer needs sin/cos (GGX azimuth, concentric disk, HDRI direction
 * from uv), atan2/acos (HDRI uv from direction) and, for the learned
 * proposal, exp/log.  The reference calls CUDA fast-math versions of these
 * (include/colvillea/libkernel/base/warp.h:46,
 * include/colvillea/libkernel/emitters/hdridome.h:210-211,452,474-475);
 * those carry no reproducible bit pattern, so the renderer defines its own.
 *
 * Method: Cody-Waite argument reduction (three-part pi/2) and polynomial
 * kernels on the reduced range.  Coefficients come from
 * gen_elem_coeffs.py (Lawson iteratively reweighted least squares, i.e. an
 * approximate minimax fit, rounded to float32).  Every polynomial is
 * evaluated in a fixed Horner order with no fused multiply-add.
 *
 * Accuracy (checked against float64 libm in the CPU test suite): sin/cos
 * <= 2 ulp for |x| <= 8 (the renderer's arguments lie in [-2pi, 2pi];
 * larger arguments are reduced with a float tail and lose accuracy),
 * atan2/acos/asin <= 3 ulp, exp/log <= 2 ulp.
 *
 * The same header serves C (gcc) and HIP (hipcc); it depends on nothing but
 * <math.h> for floorf/rintf/sqrtf/fabsf, all of which are exact or
 * correctly rounded on both sides.
 */
#ifndef CLPT_ELEM_H
#define CLPT_ELEM_H

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define CLPT_FN __host__ __device__ static inline
#else
#include <math.h>
#include <stdint.h>
#include <string.h>
#define CLPT_FN static inline
#endif

/* ---- polynomial coefficients (gen_elem_coeffs.py) ---- */
/* sin(r) = r + r*z*S(z), z = r^2, |r| <= pi/4 */
#define CLPT_SIN_S0 -0.1666666716337204f
#define CLPT_SIN_S1 0.008333329111337662f
#define CLPT_SIN_S2 -0.0001983933470910415f
#define CLPT_SIN_S3 2.718312771321507e-06f
/* cos(r) = (1 - z/2) + z*z*C(z) */
#define CLPT_COS_C0 0.0416666679084301f
#define CLPT_COS_C1 -0.0013888883404433727f
#define CLPT_COS_C2 2.4799519451335073e-05f
#define CLPT_COS_C3 -2.721024543461681e-07f
/* atan(t) = t + t*z*A(z), |t| <= tan(pi/8) */
#define CLPT_ATAN_A0 -0.3333333432674408f
#define CLPT_ATAN_A1 0.19999994337558746f
#define CLPT_ATAN_A2 -0.14285427331924438f
#define CLPT_ATAN_A3 0.11104058474302292f
#define CLPT_ATAN_A4 -0.08997320383787155f
#define CLPT_ATAN_A5 0.06993293017148972f
#define CLPT_ATAN_A6 -0.03795938938856125f
/* asin(x) = x + x*z*P(z), |x| <= 1/2 */
#define CLPT_ASIN_P0 0.16666659712791443f
#define CLPT_ASIN_P1 0.07500510662794113f
#define CLPT_ASIN_P2 0.04451572522521019f
#define CLPT_ASIN_P3 0.03181758150458336f
#define CLPT_ASIN_P4 0.014402728527784348f
#define CLPT_ASIN_P5 0.037565022706985474f
/* exp: 2^k * e^r, |r| <= ln2/2, e^r = 1 + r + r*r*E(r) */
#define CLPT_EXP_E0 0.5f
#define CLPT_EXP_E1 0.1666651964187622f
#define CLPT_EXP_E2 0.04166620597243309f
#define CLPT_EXP_E3 0.008368880487978458f
#define CLPT_EXP_E4 0.0013950478751212358f
/* log(1+f) = 2s + 2s*w*L(w), s = f/(2+f), w = s^2 */
#define CLPT_LOG_L0 0.3333333432674408f
#define CLPT_LOG_L1 0.19999997317790985f
#define CLPT_LOG_L2 0.142860546708107f
#define CLPT_LOG_L3 0.11088097840547562f
#define CLPT_LOG_L4 0.09792017191648483f

/* Three-part pi/2: k*C1 exact for |k| < 2^16, k*C2 exact for |k| < 2^12. */
#define CLPT_PIO2_C1 1.5703125f
#define CLPT_PIO2_C2 0.0004837512969970703f
#define CLPT_PIO2_C3 7.549790126404332e-08f
#define CLPT_2_OVER_PI 0.636619772367581343f
/* Two-part constants: value = HI + LO to ~2^-48. */
#define CLPT_PIO2_HI 1.5707963705062866f
#define CLPT_PIO2_LO -4.371139000186243e-08f
#define CLPT_PIO4_HI 0.7853981852531433f
#define CLPT_PIO4_LO -2.1855695000931214e-08f
#define CLPT_PI_HI 3.1415927410125732f
#define CLPT_PI_LO -8.742278000372485e-08f
#define CLPT_LN2_HI 0.693145751953125f
#define CLPT_LN2_LO 1.428606765330187e-06f
#define CLPT_LOG2E 1.4426950408889634f

CLPT_FN uint32_t clpt_f2u(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}

CLPT_FN float clpt_u2f(uint32_t u)
{
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

CLPT_FN int clpt_isnan(float x) { return x != x; }

/* sin and cos of x in one reduction. */
CLPT_FN void clpt_sincosf(float x, float* s_out, float* c_out)
{
    const int bad = clpt_isnan(x) || fabsf(x) == __builtin_huge_valf();
    if (bad)
        x = 0.0f;
    const float j = rintf(x * CLPT_2_OVER_PI);
    const int   q = (int)j;
    float       r = x - j * CLPT_PIO2_C1;
    r             = r - j * CLPT_PIO2_C2;
    r             = r - j * CLPT_PIO2_C3;
    const float z = r * r;
    const float sp = CLPT_SIN_S0 + z * (CLPT_SIN_S1 + z * (CLPT_SIN_S2 + z * CLPT_SIN_S3));
    const float s  = r + (r * z) * sp;
    const float cp = CLPT_COS_C0 + z * (CLPT_COS_C1 + z * (CLPT_COS_C2 + z * CLPT_COS_C3));
    const float c  = (1.0f - 0.5f * z) + (z * z) * cp;
    /* quadrant q & 3: (s, c), (c, -s), (-s, -c), (-c, s), as selects */
    const int   swap = (q & 1) != 0;
    const float s1   = swap ? c : s;
    const float c1   = swap ? s : c;
    const float nan  = clpt_u2f(0x7fc00000u);
    *s_out           = bad ? nan : ((q & 2) != 0 ? -s1 : s1);
    *c_out           = bad ? nan : (((q + 1) & 2) != 0 ? -c1 : c1);
}

CLPT_FN float clpt_sinf(float x)
{
    float s, c;
    clpt_sincosf(x, &s, &c);
    return s;
}

CLPT_FN float clpt_cosf(float x)
{
    float s, c;
    clpt_sincosf(x, &s, &c);
    return c;
}

/* atan(t) for |t| <= tan(pi/8). */
CLPT_FN float clpt_atan_kernel(float t)
{
    const float z = t * t;
    const float p = CLPT_ATAN_A0 +
        z * (CLPT_ATAN_A1 + z * (CLPT_ATAN_A2 + z * (CLPT_ATAN_A3 + z * (CLPT_ATAN_A4 + z * (CLPT_ATAN_A5 + z * CLPT_ATAN_A6)))));
    return t + (t * z) * p;
}

CLPT_FN float clpt_atan2f(float y, float x)
{
    if (clpt_isnan(x) || clpt_isnan(y))
        return x + y;
    const int   ysign = (clpt_f2u(y) >> 31) != 0;
    const int   xneg  = (clpt_f2u(x) >> 31) != 0;
    const float ax    = fabsf(x);
    const float ay    = fabsf(y);
    const float inf   = __builtin_huge_valf();
    float       r;
    if (ay == 0.0f)
    {
        r = xneg ? CLPT_PI_HI : 0.0f;
    }
    else if (ax == 0.0f)
    {
        r = CLPT_PIO2_HI;
    }
    else if (ax == inf || ay == inf)
    {
        if (ax == inf && ay == inf)
            r = xneg ? 3.0f * CLPT_PIO4_HI : CLPT_PIO4_HI;
        else if (ay == inf)
            r = CLPT_PIO2_HI;
        else
            r = xneg ? CLPT_PI_HI : 0.0f;
    }
    else
    {
        const int   swap = ay > ax;
        const float mx   = swap ? ay : ax;
        const float mn   = swap ? ax : ay;
        const float a    = mn / mx;
        float       at;
        if (a > 0.41421356237309503f)
        {
            const float t = (a - 1.0f) / (a + 1.0f);
            at            = CLPT_PIO4_HI + (clpt_atan_kernel(t) + CLPT_PIO4_LO);
        }
        else
        {
            at = clpt_atan_kernel(a);
        }
        if (swap)
            at = CLPT_PIO2_HI - (at - CLPT_PIO2_LO);
        if (xneg)
            at = CLPT_PI_HI - (at - CLPT_PI_LO);
        r = at;
    }
    return ysign ? -r : r;
}

/* asin(x) for |x| <= 1/2 (z = x*x supplied). */
CLPT_FN float clpt_asin_kernel(float x, float z)
{
    const float p = CLPT_ASIN_P0 + z * (CLPT_ASIN_P1 + z * (CLPT_ASIN_P2 + z * (CLPT_ASIN_P3 + z * (CLPT_ASIN_P4 + z * CLPT_ASIN_P5))));
    return x + (x * z) * p;
}

CLPT_FN float clpt_acosf(float x)
{
    if (clpt_isnan(x) || fabsf(x) > 1.0f)
        return clpt_u2f(0x7fc00000u);
    const float ax = fabsf(x);
    if (ax <= 0.5f)
    {
        const float as = clpt_asin_kernel(x, x * x);
        return CLPT_PIO2_HI - (as - CLPT_PIO2_LO);
    }
    const float z  = (1.0f - ax) * 0.5f;
    const float s  = sqrtf(z);
    const float as = clpt_asin_kernel(s, z);
    if (x > 0.0f)
        return 2.0f * as;
    return CLPT_PI_HI - (2.0f * as - CLPT_PI_LO);
}

CLPT_FN float clpt_asinf(float x)
{
    if (clpt_isnan(x) || fabsf(x) > 1.0f)
        return clpt_u2f(0x7fc00000u);
    const float ax = fabsf(x);
    float       r;
    if (ax <= 0.5f)
        return clpt_asin_kernel(x, x * x);
    const float z  = (1.0f - ax) * 0.5f;
    const float s  = sqrtf(z);
    const float as = clpt_asin_kernel(s, z);
    r              = CLPT_PIO2_HI - (2.0f * as - CLPT_PIO2_LO);
    return x < 0.0f ? -r : r;
}

CLPT_FN float clpt_expf(float x)
{
    if (clpt_isnan(x))
        return x;
    if (x > 88.72283935546875f)
        return __builtin_huge_valf();
    if (x < -103.97208404541016f)
        return 0.0f;
    const float k = rintf(x * (float)CLPT_LOG2E);
    const float r = (x - k * CLPT_LN2_HI) - k * CLPT_LN2_LO;
    const float p = CLPT_EXP_E0 + r * (CLPT_EXP_E1 + r * (CLPT_EXP_E2 + r * (CLPT_EXP_E3 + r * CLPT_EXP_E4)));
    const float e = 1.0f + (r + (r * r) * p);
    /* scale by 2^k in two steps so that subnormal results stay exact-ish */
    int       ki = (int)k;
    const int k1 = ki / 2;
    const int k2 = ki - k1;
    return (e * clpt_u2f((uint32_t)(127 + k1) << 23)) * clpt_u2f((uint32_t)(127 + k2) << 23);
}

CLPT_FN float clpt_logf(float x)
{
    if (clpt_isnan(x) || x < 0.0f)
        return clpt_u2f(0x7fc00000u);
    if (x == 0.0f)
        return -__builtin_huge_valf();
    if (x == __builtin_huge_valf())
        return x;
    uint32_t u = clpt_f2u(x);
    int      e = 0;
    if (u < 0x00800000u) /* subnormal: scale up by 2^25 */
    {
        u = clpt_f2u(x * 33554432.0f);
        e = -25;
    }
    e += (int)(u >> 23) - 127;
    /* mantissa m in [sqrt(1/2), sqrt(2)) */
    uint32_t mu = (u & 0x007fffffu) | 0x3f800000u;
    float    m  = clpt_u2f(mu);
    if (m > 1.41421356f)
    {
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float w = s * s;
    const float p = CLPT_LOG_L0 + w * (CLPT_LOG_L1 + w * (CLPT_LOG_L2 + w * (CLPT_LOG_L3 + w * CLPT_LOG_L4)));
    const float l = 2.0f * s + (2.0f * s * w) * p;
    const float k = (float)e;
    return (k * CLPT_LN2_HI) + (l + k * CLPT_LN2_LO);
}

CLPT_FN float clpt_tanf(float x)
{
    float s, c;
    clpt_sincosf(x, &s, &c);
    return s / c;
}

#endif /* CLPT_ELEM_H */
