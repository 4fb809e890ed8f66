/* orc_f32.c -- float32 instantiation of the oracle core (bit-comparable to
 * the gfx950 kernels).  TEST INFRASTRUCTURE ONLY (see oracle.h). */
#include "../colvillea_amd/csrc/clpt_elem.h"

#define R float
#define SFX(x) x##_f32
#define RC(lit) ((float)(lit))
#define RSINCOS(x, s, c) clpt_sincosf((x), (s), (c))
#define RATAN2(y, x) clpt_atan2f((y), (x))
#define RACOS(x) clpt_acosf(x)
#define RSQRT(x) sqrtf(x)
#define RFABS(x) fabsf(x)
#define RFMIN(a, b) fminf((a), (b))
#define RFMAX(a, b) fmaxf((a), (b))
#define RFLOOR(x) floorf(x)
#define RFMA(a, b, c) fmaf((a), (b), (c))
#define RPOW(x, y) powf((x), (y)) /* libm: the display encoding only */
#define ORC_PI 3.14159265358979323846f
#define ORC_INV_PI 0.318309886183790671538f
#define ORC_IS_F32 1
#include "orc_real.h"
