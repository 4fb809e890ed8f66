/* orc_f64.c -- float64 instantiation of the oracle core (libm elementary
 * functions): the image-level reference.  TEST INFRASTRUCTURE ONLY (see oracle.h). */
#include <math.h>

static inline void orc_sincos64(double x, double* s, double* c)
{
    *s = sin(x);
    *c = cos(x);
}

#define R double
#define SFX(x) x##_f64
#define RC(lit) ((double)(lit))
#define RSINCOS(x, s, c) orc_sincos64((x), (s), (c))
#define RATAN2(y, x) atan2((y), (x))
#define RACOS(x) acos(x)
#define RSQRT(x) sqrt(x)
#define RFABS(x) fabs(x)
#define RFMIN(a, b) fmin((a), (b))
#define RFMAX(a, b) fmax((a), (b))
#define RFLOOR(x) floor(x)
#define RFMA(a, b, c) fma((a), (b), (c))
#define RPOW(x, y) pow((x), (y))
#define ORC_PI 3.14159265358979323846
#define ORC_INV_PI 0.318309886183790671538
#define ORC_IS_F32 0
#include "orc_real.h"
