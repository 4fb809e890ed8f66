/*
 * orc_real.h -- the oracle's path-tracing core, written once for a real type
 * R and instantiated twice (orc_f32.c: R = float, the renderer's elementary
 * functions; orc_f64.c: R = double, libm).  Expressions keep the reference's
 * evaluation order (left-associative products and sums as written), which
 * together with -ffp-contract=off makes the float32 variant bit-comparable
 * to the gfx950 kernels.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The instantiating file defines:
 *   R, SFX(x), RC(lit)      real type, name suffixer, literal of type R
 *   RSINCOS(x,s,c), RATAN2(y,x), RACOS(x)
 *   RSQRT, RFABS, RFMIN, RFMAX, RFLOOR, RFMA (fused multiply-add), RPOW
 *   ORC_PI, ORC_INV_PI      the reference's M_PIf / M_1_PIf in type R
 *   ORC_IS_F32              1 for the float32 variant
 */
#include "orc_internal.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define V3 SFX(v3)
typedef struct
{
    R x, y, z;
} V3;

static inline V3 SFX(mk)(R x, R y, R z)
{
    V3 v = {x, y, z};
    return v;
}
static inline V3 SFX(ldf)(const float* p) { return SFX(mk)((R)p[0], (R)p[1], (R)p[2]); }
static inline V3 SFX(add)(V3 a, V3 b) { return SFX(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 SFX(sub)(V3 a, V3 b) { return SFX(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 SFX(mul)(V3 a, V3 b) { return SFX(mk)(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 SFX(divv)(V3 a, V3 b) { return SFX(mk)(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline V3 SFX(scl)(V3 a, R s) { return SFX(mk)(a.x * s, a.y * s, a.z * s); }
static inline V3 SFX(sclL)(R s, V3 a) { return SFX(mk)(s * a.x, s * a.y, s * a.z); }
static inline V3 SFX(neg)(V3 a) { return SFX(mk)(-a.x, -a.y, -a.z); }
static inline R  SFX(dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 SFX(cross)(V3 a, V3 b)
{
    return SFX(mk)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* OWL normalize: v * (1 / sqrt(dot(v, v))) -- rounding form chosen here (OWL
 * is absent from the image; the form is shared with the renderer). */
static inline V3 SFX(normalize)(V3 v) { return SFX(scl)(v, RC(1.0) / RSQRT(SFX(dot)(v, v))); }
static inline int SFX(nonzero)(V3 v) { return v.x > RC(0.0) || v.y > RC(0.0) || v.z > RC(0.0); }

/* ------------------------------------------------------------ warps */

/* warp::squareToUniformConcentricDisk: warp.h:25-50.  The reference
 * evaluates phi in double (M_PI / 4.0f is a double). */
static void SFX(concentric_disk)(R ux, R uy, R* ox, R* oy)
{
    const R r1 = RC(2.0) * ux - RC(1.0);
    const R r2 = RC(2.0) * uy - RC(1.0);
    R       phi, r;
    if (r1 == RC(0.0) && r2 == RC(0.0))
    {
        r = phi = RC(0.0);
    }
    else if (r1 * r1 > r2 * r2)
    {
        r   = r1;
        phi = (R)((3.14159265358979323846 / 4.0) * (double)(r2 / r1));
    }
    else
    {
        r   = r2;
        phi = (R)((3.14159265358979323846 / 2.0) - (double)(r1 / r2) * (3.14159265358979323846 / 4.0));
    }
    R s, c;
    RSINCOS(phi, &s, &c);
    *ox = r * c;
    *oy = r * s;
}

/* warp::squareToCosineHemisphere: warp.h:58-64.  The radicand is clamped
 * at 0 (the reference can take sqrt of a rounding-negative value at |p| = 1). */
static V3 SFX(cosine_hemisphere)(R ux, R uy)
{
    R px, py;
    SFX(concentric_disk)(ux, uy, &px, &py);
    const R z = RSQRT(RFMAX(RC(0.0), RC(1.0) - px * px - py * py));
    return SFX(mk)(px, py, z);
}

/* makeFrame: math.h:33-48 -- b = cross(c, a). */
static void SFX(make_frame)(V3 a, V3* b, V3* c)
{
    if (RFABS(a.x) > RFABS(a.y))
    {
        const R inv = RC(1.0) / RSQRT(a.x * a.x + a.z * a.z);
        *c          = SFX(mk)(a.z * inv, RC(0.0), -a.x * inv);
    }
    else
    {
        const R inv = RC(1.0) / RSQRT(a.y * a.y + a.z * a.z);
        *c          = SFX(mk)(RC(0.0), a.z * inv, -a.y * inv);
    }
    *b = SFX(cross)(*c, a);
}

/* Frame (frame.h:24-130): s = secondary tangent, t = tangent, n = normal;
 * toLocal = (v.t, v.s, v.n), toWorld = x*t + y*s + z*n. */
typedef struct
{
    V3 s, t, n;
} SFX(frame);
static inline V3 SFX(to_local)(const SFX(frame) * f, V3 v)
{
    return SFX(mk)(SFX(dot)(v, f->t), SFX(dot)(v, f->s), SFX(dot)(v, f->n));
}
static inline V3 SFX(to_world)(const SFX(frame) * f, V3 v)
{
    return SFX(add)(SFX(add)(SFX(sclL)(v.x, f->t), SFX(sclL)(v.y, f->s)), SFX(sclL)(v.z, f->n));
}

/* ------------------------------------------------------------ GGX */

/* hypot2: math.h:55-73 */
static R SFX(hypot2)(R a, R b)
{
    R r;
    if (RFABS(a) > RFABS(b))
    {
        r = b / a;
        r = RFABS(a) * RSQRT(RC(1.0) + r * r);
    }
    else if (b != RC(0.0))
    {
        r = a / b;
        r = RFABS(b) * RSQRT(RC(1.0) + r * r);
    }
    else
        r = RC(0.0);
    return r;
}

/* MicrofacetDistribution: microfacet.h:24-134 (alpha clamped at 1e-4). */
static inline R SFX(ggx_alpha)(R alpha) { return RFMAX(alpha, RC(1e-4)); }

static R SFX(ggx_D)(R alpha, V3 m)
{
    if (m.z <= RC(0.0))
        return RC(0.0);
    const R cos2 = m.z * m.z;
    const R be   = (m.x * m.x + m.y * m.y) / (alpha * alpha * cos2);
    const R root = (RC(1.0) + be) * cos2;
    return RC(1.0) / (ORC_PI * alpha * alpha * root * root);
}

/* Frame::sinTheta/tanTheta: frame.h:80-89 */
static inline R SFX(sin_theta)(V3 v)
{
    const R t = RC(1.0) - v.z * v.z;
    return t <= RC(0.0) ? RC(0.0) : RSQRT(t);
}

static R SFX(ggx_G1)(R alpha, V3 v, V3 m)
{
    if (SFX(dot)(v, m) * v.z <= RC(0.0))
        return RC(0.0);
    const R tan_theta = RFABS(SFX(sin_theta)(v) / v.z);
    if (tan_theta == RC(0.0))
        return RC(1.0);
    const R root = alpha * tan_theta;
    return RC(2.0) / (RC(1.0) + SFX(hypot2)(RC(1.0), root));
}

static inline R SFX(ggx_G)(R alpha, V3 wi, V3 wo, V3 m) { return SFX(ggx_G1)(alpha, wi, m) * SFX(ggx_G1)(alpha, wo, m); }

/* sampleAll: microfacet.h:92-121 */
static V3 SFX(ggx_sample_all)(R alpha, R ux, R uy, R* pdf)
{
    const R alpha_sqr = alpha * alpha;
    R       sin_phi, cos_phi;
    RSINCOS((RC(2.0) * ORC_PI) * uy, &sin_phi, &cos_phi);
    const R tan2 = alpha_sqr * ux / (RC(1.0) - ux);
    const R cosm = RC(1.0) / RSQRT(RC(1.0) + tan2);
    const R temp = RC(1.0) + tan2 / alpha_sqr;
    *pdf         = ORC_INV_PI / (alpha_sqr * cosm * cosm * cosm * temp * temp);
    if (*pdf < RC(1e-20))
        *pdf = RC(0.0);
    const R sinm = RSQRT(RFMAX(RC(0.0), RC(1.0) - cosm * cosm));
    return SFX(mk)(sinm * cos_phi, sinm * sin_phi, cosm);
}

/* ------------------------------------------------------------ Fresnel */

/* fresnelConductor: math.h:172-193 (Mitsuba 0.6), per channel. */
static R SFX(fresnel_conductor1)(R c, R eta, R k)
{
    const R cos2  = c * c;
    const R sin2  = RC(1.0) - cos2;
    const R sin4  = sin2 * sin2;
    const R temp1 = eta * eta - k * k - sin2;
    const R a2pb2 = RSQRT(temp1 * temp1 + k * k * eta * eta * RC(4.0));
    const R a     = RSQRT((a2pb2 + temp1) * RC(0.5));
    const R term1 = a2pb2 + cos2;
    const R term2 = a * (RC(2.0) * c);
    const R rs2   = (term1 - term2) / (term1 + term2);
    const R term3 = a2pb2 * cos2 + sin4;
    const R term4 = term2 * sin2;
    const R rp2   = rs2 * (term3 - term4) / (term3 + term4);
    return RC(0.5) * (rp2 + rs2);
}

/* fresnelDielectricExt: math.h:255-287 (fresnelDielectric, :208-235, is the
 * same value without cos_t). */
static R SFX(fresnel_dielectric_ext)(R cos_i_, R* cos_t_, R eta)
{
    if (eta == RC(1.0))
    {
        *cos_t_ = -cos_i_;
        return RC(0.0);
    }
    const R scale = (cos_i_ > RC(0.0)) ? RC(1.0) / eta : eta;
    const R ct2   = RC(1.0) - (RC(1.0) - cos_i_ * cos_i_) * (scale * scale);
    if (ct2 <= RC(0.0))
    {
        *cos_t_ = RC(0.0);
        return RC(1.0);
    }
    const R cos_i = RFABS(cos_i_);
    const R cos_t = RSQRT(ct2);
    const R rs    = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    const R rp    = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    *cos_t_       = (cos_i_ > RC(0.0)) ? -cos_t : cos_t;
    return RC(0.5) * (rs * rs + rp * rp);
}

static inline R SFX(fresnel_dielectric)(R cos_i, R eta)
{
    R ct;
    return SFX(fresnel_dielectric_ext)(cos_i, &ct, eta);
}

/* reflect / refract: math.h:299-329 */
static inline V3 SFX(reflect)(V3 wo, V3 n) { return SFX(add)(SFX(neg)(wo), SFX(sclL)(RC(2.0) * SFX(dot)(wo, n), n)); }
static inline V3 SFX(refract)(V3 wi, V3 n, R eta, R cos_t)
{
    if (cos_t < RC(0.0))
        eta = RC(1.0) / eta;
    return SFX(add)(SFX(sclL)(-eta, wi), SFX(sclL)(eta * SFX(dot)(wi, n) + cos_t, n));
}

/* ------------------------------------------------------------ BSDFs */

/* Material -> BSDF parameters: DiffuseMtl/MetalMtl/GlassMtl::getBSDF
 * (materials/diffusemtl.h:20-29, metal.h:37-55 with roughness^2 at :52,
 * glass.h:19-29 with roughness^2 at :26). */
static inline R SFX(mtl_alpha)(const orc_material* m)
{
    const R r = (R)m->roughness;
    return SFX(ggx_alpha)(r * r);
}

/* SmoothDiffuse::eval (smoothdiffuse.h:31-44); RoughConductor::eval
 * (roughconductor.h:39-70); RoughDielectric::eval (roughdielectric.h:36-103,
 * whose local names swap wi/wo).  Arguments: wi = sampled/light direction,
 * wo = direction towards the previous vertex, both local. */
static V3 SFX(bsdf_eval)(const orc_material* m, V3 wi, V3 wo)
{
    const V3 zero = SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
    if (m->type == ORC_MTL_DIFFUSE)
    {
        if (wo.z <= RC(0.0) || wi.z <= RC(0.0))
            return zero;
        return SFX(scl)(SFX(ldf)(m->reflectance), ORC_INV_PI);
    }
    if (m->type == ORC_MTL_METAL)
    {
        if (wo.z <= RC(0.0) || wi.z <= RC(0.0))
            return zero;
        const R  alpha = SFX(mtl_alpha)(m);
        const V3 h     = SFX(normalize)(SFX(add)(wo, wi));
        const R  D     = SFX(ggx_D)(alpha, h);
        if (D == RC(0.0))
            return zero;
        const R  c = SFX(dot)(wo, h);
        const V3 F = SFX(mul)(SFX(mk)(SFX(fresnel_conductor1)(c, (R)m->eta[0], (R)m->k[0]),
                                      SFX(fresnel_conductor1)(c, (R)m->eta[1], (R)m->k[1]),
                                      SFX(fresnel_conductor1)(c, (R)m->eta[2], (R)m->k[2])),
                              SFX(ldf)(m->reflectance));
        const R  G     = SFX(ggx_G)(alpha, wo, wi, h);
        const R  denom = RC(4.0) * wo.z * wi.z;
        return SFX(mk)(F.x * D * G / denom, F.y * D * G / denom, F.z * D * G / denom);
    }
    /* dielectric: reference names wi := bRec.woLocal, wo := bRec.wiLocal */
    const V3 dwi = wo, dwo = wi;
    if (dwi.z == RC(0.0))
        return zero;
    const R ior     = (R)m->ior;
    const R alpha   = SFX(mtl_alpha)(m);
    const int refl  = dwi.z * dwo.z > RC(0.0);
    V3      H;
    if (refl)
        H = SFX(normalize)(SFX(add)(dwo, dwi));
    else
    {
        const R eta = dwi.z > RC(0.0) ? ior / RC(1.0) : RC(1.0) / ior;
        H           = SFX(normalize)(SFX(add)(dwi, SFX(scl)(dwo, eta)));
    }
    H       = SFX(scl)(H, H.z > RC(0.0) ? RC(1.0) : RC(-1.0));
    const R D = SFX(ggx_D)(alpha, H);
    if (D == RC(0.0))
        return zero;
    const R F = SFX(fresnel_dielectric)(SFX(dot)(dwi, H), ior / RC(1.0));
    const R G = SFX(ggx_G)(alpha, dwi, dwo, H);
    R       value;
    if (refl)
        value = F * D * G / (RC(4.0) * RFABS(dwi.z) * RFABS(dwo.z));
    else
    {
        const R eta = dwi.z > RC(0.0) ? ior / RC(1.0) : RC(1.0) / ior;
        const R sd  = SFX(dot)(dwi, H) + eta * SFX(dot)(dwo, H);
        value = RFABS((RC(1.0) - F) * D * G * SFX(dot)(dwi, H) * SFX(dot)(dwo, H) /
                      (sd * sd * RFABS(dwi.z) * RFABS(dwo.z)));
    }
    return SFX(mk)(value, value, value);
}

/* pdf of bsdf_sample for direction wi.  Corrected forms:
 *  D4 smoothdiffuse.h:56 uses woLocal -> cosine pdf of wi;
 *  D5 roughconductor.h:89 uses pdfAll(wo) -> D(h) cos(h) / (4 |wo.h|);
 *  D6 roughdielectric.h:149 multiplies pdfAll(H) by cos(H) again -> dropped.
 * Each equals the density sample() reports for the same direction. */
static R SFX(bsdf_pdf)(const orc_material* m, V3 wi, V3 wo)
{
    if (m->type == ORC_MTL_DIFFUSE)
    {
        if (wo.z <= RC(0.0) || wi.z <= RC(0.0))
            return RC(0.0);
        return wi.z * ORC_INV_PI;
    }
    if (m->type == ORC_MTL_METAL)
    {
        if (wo.z <= RC(0.0) || wi.z <= RC(0.0))
            return RC(0.0);
        const R  alpha = SFX(mtl_alpha)(m);
        const V3 h     = SFX(normalize)(SFX(add)(wo, wi));
        return SFX(ggx_D)(alpha, h) * h.z / (RC(4.0) * RFABS(SFX(dot)(wo, h)));
    }
    const V3 dwi = wo, dwo = wi;
    const R  ior   = (R)m->ior;
    const R  alpha = SFX(mtl_alpha)(m);
    const int refl = dwi.z * dwo.z > RC(0.0);
    V3       H;
    R        dwh;
    if (refl)
    {
        H   = SFX(normalize)(SFX(add)(dwo, dwi));
        dwh = RC(1.0) / (RC(4.0) * SFX(dot)(dwo, H));
    }
    else
    {
        const R eta = dwi.z > RC(0.0) ? ior / RC(1.0) : RC(1.0) / ior;
        H           = SFX(normalize)(SFX(add)(dwi, SFX(scl)(dwo, eta)));
        const R sd  = SFX(dot)(dwi, H) + eta * SFX(dot)(dwo, H);
        dwh         = (eta * eta * SFX(dot)(dwo, H)) / (sd * sd);
    }
    H      = SFX(scl)(H, H.z < RC(0.0) ? RC(-1.0) : RC(1.0));
    R prob = SFX(ggx_D)(alpha, H) * H.z;
    const R F = SFX(fresnel_dielectric)(SFX(dot)(dwi, H), ior / RC(1.0));
    prob *= refl ? F : (RC(1.0) - F);
    return RFABS(prob * dwh);
}

/* BSDF::sample.  Returns f (zero when no direction is produced); *wi and
 * *pdf are set when f is nonzero.  u_lobe is the dielectric's
 * reflect/refract draw (D7: roughdielectric.h:191 reuses sample.x, which also
 * set theta_m; the renderer draws it independently). */
static V3 SFX(bsdf_sample)(const orc_material* m, V3 wo, R ux, R uy, R u_lobe, V3* wi, R* pdf)
{
    const V3 zero = SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
    if (m->type == ORC_MTL_DIFFUSE)
    {
        /* SmoothDiffuse::sample: smoothdiffuse.h:59-79 */
        if (wo.z <= RC(0.0))
            return zero;
        *wi  = SFX(cosine_hemisphere)(ux, uy);
        *pdf = wi->z * ORC_INV_PI;
        return SFX(scl)(SFX(ldf)(m->reflectance), ORC_INV_PI);
    }
    if (m->type == ORC_MTL_METAL)
    {
        /* RoughConductor::sample: roughconductor.h:97-137 */
        if (wo.z <= RC(0.0))
            return zero;
        const R alpha = SFX(mtl_alpha)(m);
        R       p;
        const V3 mf = SFX(ggx_sample_all)(alpha, ux, uy, &p);
        if (p == RC(0.0))
            return zero;
        const V3 w = SFX(reflect)(wo, mf);
        if (w.z <= RC(0.0))
            return zero;
        p /= (RC(4.0) * SFX(dot)(wo, mf));
        const R D = SFX(ggx_D)(alpha, mf);
        if (D == RC(0.0))
            return zero;
        const R  c = SFX(dot)(wo, mf);
        const V3 F = SFX(mul)(SFX(mk)(SFX(fresnel_conductor1)(c, (R)m->eta[0], (R)m->k[0]),
                                      SFX(fresnel_conductor1)(c, (R)m->eta[1], (R)m->k[1]),
                                      SFX(fresnel_conductor1)(c, (R)m->eta[2], (R)m->k[2])),
                              SFX(ldf)(m->reflectance));
        const R  G     = SFX(ggx_G)(alpha, wo, w, mf);
        const R  denom = RC(4.0) * wo.z * w.z;
        *wi            = w;
        *pdf           = p;
        return SFX(mk)(F.x * D * G / denom, F.y * D * G / denom, F.z * D * G / denom);
    }
    /* RoughDielectric::sample: roughdielectric.h:158-251 */
    const V3 dwi   = wo;
    const R  ior   = (R)m->ior;
    const R  alpha = SFX(mtl_alpha)(m);
    R        mpdf;
    const V3 mf = SFX(ggx_sample_all)(alpha, ux, uy, &mpdf);
    if (mpdf == RC(0.0))
        return zero;
    R       p = mpdf;
    R       cos_t;
    const R F = SFX(fresnel_dielectric_ext)(SFX(dot)(dwi, mf), &cos_t, ior / RC(1.0));
    int     refl = 1;
    if (u_lobe > F)
    {
        refl = 0;
        p *= RC(1.0) - F;
    }
    else
        p *= F;
    const R D = SFX(ggx_D)(alpha, mf);
    if (refl)
    {
        const V3 dwo = SFX(reflect)(dwi, mf);
        if (dwi.z * dwo.z <= RC(0.0))
            return zero;
        const R G   = SFX(ggx_G)(alpha, dwo, dwi, mf);
        const R dwh = RFABS(RC(1.0) / (RC(4.0) * SFX(dot)(dwo, mf)));
        p *= dwh;
        const R v = RFABS(F * D * G / (RC(4.0) * dwo.z * dwi.z));
        *wi       = dwo;
        *pdf      = p;
        return SFX(mk)(v, v, v);
    }
    if (cos_t == RC(0.0))
        return zero;
    const V3 dwo = SFX(refract)(dwi, mf, ior / RC(1.0), cos_t);
    if (dwo.z * dwi.z >= RC(0.0))
        return zero;
    const R G   = SFX(ggx_G)(alpha, dwo, dwi, mf);
    const R eta = dwi.z > RC(0.0) ? ior / RC(1.0) : RC(1.0) / ior;
    const R sd  = SFX(dot)(dwi, mf) + eta * SFX(dot)(dwo, mf);
    const R dwh = RFABS(SFX(dot)(dwo, mf) * eta * eta / (sd * sd));
    p *= dwh;
    const R v = RFABS((RC(1.0) - F) * D * G * SFX(dot)(dwi, mf) * SFX(dot)(dwo, mf) /
                      (sd * sd * RFABS(dwi.z) * RFABS(dwo.z)));
    *wi  = dwo;
    *pdf = p;
    return SFX(mk)(v, v, v);
}

/* ------------------------------------------------------------ textures */

/* Exact bilinear fetch with wrap addressing and normalized coordinates --
 * the renderer's definition of ImageTexture2D::eval2D (imagetex2d.h:34-37;
 * filter/address modes src/core/nodes/textures/imagetex2d.cpp:50-66).  The
 * CUDA texture unit's 9-bit weights are not reproduced. */
static void SFX(tex_eval)(const float* rgba, uint32_t w, uint32_t h, R u, R v, R out[4])
{
    const R  uu = u - RFLOOR(u);
    const R  vv = v - RFLOOR(v);
    const R  x  = uu * (R)w - RC(0.5);
    const R  y  = vv * (R)h - RC(0.5);
    const R  fx = RFLOOR(x), fy = RFLOOR(y);
    const R  ax = x - fx, ay = y - fy;
    int64_t  ix0 = (int64_t)fx, iy0 = (int64_t)fy;
    int64_t  ix1 = ix0 + 1, iy1 = iy0 + 1;
    const int64_t W = (int64_t)w, H = (int64_t)h;
    ix0 = ((ix0 % W) + W) % W;
    ix1 = ((ix1 % W) + W) % W;
    iy0 = ((iy0 % H) + H) % H;
    iy1 = ((iy1 % H) + H) % H;
    const float* t00 = rgba + 4 * (iy0 * W + ix0);
    const float* t10 = rgba + 4 * (iy0 * W + ix1);
    const float* t01 = rgba + 4 * (iy1 * W + ix0);
    const float* t11 = rgba + 4 * (iy1 * W + ix1);
    const R w00 = (RC(1.0) - ax) * (RC(1.0) - ay);
    const R w10 = ax * (RC(1.0) - ay);
    const R w01 = (RC(1.0) - ax) * ay;
    const R w11 = ax * ay;
    for (int c = 0; c < 4; ++c)
        out[c] = w00 * (R)t00[c] + w10 * (R)t10[c] + w01 * (R)t01[c] + w11 * (R)t11[c];
}

/* ------------------------------------------------------------ emitters */

/* upper_bound: math.h:84-96 */
static uint32_t SFX(upper_bound)(const float* a, uint32_t begin, uint32_t end, R value)
{
    while (begin < end)
    {
        const uint32_t mid = begin + (end - begin) / 2;
        if ((R)a[mid] <= value)
            begin = mid + 1;
        else
            end = mid;
    }
    return begin;
}

/* Distribution1D::sample1D: hdridome.h:31-60.  Returns the continuous
 * sample; *off is the chosen cell relative to cdf_off. */
static R SFX(sample1d)(const float* f, const float* cdf, uint32_t N, R c, R u, uint32_t cdf_off, R* pdf, uint32_t* off)
{
    if (c == RC(0.0))
        return RC(0.0);
    uint32_t o = SFX(upper_bound)(cdf, cdf_off, cdf_off + N + 1, u);
    /* clamp(o - 1, cdf_off, cdf_off + N - 1) on uint32 (o - 1 wraps to
     * UINT32_MAX when o == 0, which clamps to the upper bound) */
    o = o - 1u;
    if (o < cdf_off)
        o = cdf_off;
    if (o > cdf_off + N - 1)
        o = cdf_off + N - 1;
    *pdf = (R)f[o] / c;
    R du = u - (R)cdf[o];
    if ((R)cdf[o + 1] - (R)cdf[o] != RC(0.0))
        du /= (R)cdf[o + 1] - (R)cdf[o];
    o -= cdf_off;
    *off = o;
    return ((R)o + du) / (R)N;
}

/* Distribution2D::sample2D: hdridome.h:96-123 */
static void SFX(sample2d)(const orc_dome* d, R ux, R uy, R* su, R* sv, R* pdf)
{
    const uint32_t nu = d->w, nv = d->h;
    R              pu = RC(0.0), pvv = RC(0.0);
    uint32_t       off = 0;
    *sv = SFX(sample1d)(d->pv, d->cdfv, nv, (R)d->pv[nv], uy, 0, &pvv, &off);
    const R cv = (R)d->pcond[(size_t)off * (nu + 1) + nu];
    *su = SFX(sample1d)(d->pcond, d->cdfcond, nu, cv, ux, off * (nu + 1), &pu, &off);
    *pdf = pu * pvv;
}

/* Distribution2D::pdf2D: hdridome.h:125-135 */
static R SFX(pdf2d)(const orc_dome* d, R u, R v)
{
    if (u < RC(0.0) || v < RC(0.0))
        return RC(0.0);
    uint32_t iu = (uint32_t)((R)d->w * u);
    uint32_t iv = (uint32_t)((R)d->h * v);
    if (iu > d->w - 1)
        iu = d->w - 1;
    if (iv > d->h - 1)
        iv = d->h - 1;
    return (R)d->pcond[(size_t)iv * (d->w + 1) + iu] / (R)d->pv[d->h];
}

/* cos(theta) for acos, clamped to [-1, 1]: a unit vector rotated through a
 * float32 frame can have |y| just above 1, where the reference's unclamped
 * acos(dir.y) (hdridome.h:452, 475) returns NaN.  Pinned by the
 * pole-direction test of the dome lookups. */
static inline R SFX(cos_theta_clamped)(V3 dir) { return RFMIN(RC(1.0), RFMAX(RC(-1.0), dir.y)); }

/* HDRIDome::directionToUVCoords: hdridome.h:472-476 */
static void SFX(dir_to_uv)(V3 dir, R* u, R* v)
{
    *u = RATAN2(dir.x, dir.z) * -ORC_INV_PI * RC(0.5) + RC(0.5);
    *v = RACOS(SFX(cos_theta_clamped)(dir)) * ORC_INV_PI;
}

/* HDRIDome::evalEnvironment: hdridome.h:291-297 */
static V3 SFX(dome_le)(const orc_dome* d, V3 dir)
{
    R u, v, c[4];
    SFX(dir_to_uv)(dir, &u, &v);
    SFX(tex_eval)(d->rgba, d->w, d->h, u, v, c);
    return SFX(mk)(c[0], c[1], c[2]);
}

/* HDRIDome::pdfDirect: hdridome.h:248-279 */
static R SFX(dome_pdf)(const orc_dome* d, V3 dir)
{
    const R phi   = RATAN2(dir.x, dir.z);
    const R theta = RACOS(SFX(cos_theta_clamped)(dir));
    R       st, ct;
    RSINCOS(theta, &st, &ct);
    if (st == RC(0.0))
        return RC(0.0);
    const R u = phi * -ORC_INV_PI * RC(0.5) + RC(0.5);
    const R v = theta * ORC_INV_PI;
    return SFX(pdf2d)(d, u, v) / (RC(2.0) * ORC_PI * ORC_PI * st);
}

/* HDRIDome::sampleDirect: hdridome.h:186-228 */
static V3 SFX(dome_sample)(const orc_dome* d, R ux, R uy, V3* dir, R* pdf)
{
    R su, sv, p;
    SFX(sample2d)(d, ux, uy, &su, &sv, &p);
    const R theta = ORC_PI * sv;
    const R phi   = RC(-2.0) * ORC_PI * (su - RC(0.5));
    R       st, ct, sp, cp;
    RSINCOS(theta, &st, &ct);
    RSINCOS(phi, &sp, &cp);
    if (st == RC(0.0))
    {
        *pdf = RC(0.0);
        return SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
    }
    *dir = SFX(mk)(st * sp, ct, st * cp);
    *pdf = p / (RC(2.0) * ORC_PI * ORC_PI * st);
    R c[4];
    SFX(tex_eval)(d->rgba, d->w, d->h, su, sv, c);
    return SFX(mk)(c[0], c[1], c[2]);
}

/* DirectionalEmitter::sampleDirect: directional.h:30-56 (tan of the angular
 * radius is precomputed once per emitter, the same value on both sides). */
static V3 SFX(sun_sample)(const orc_emitter* e, R tan_r, R ux, R uy, V3* dir, R* pdf)
{
    const V3 rev = SFX(neg)(SFX(ldf)(e->direction));
    R        dx, dy;
    SFX(concentric_disk)(ux, uy, &dx, &dy);
    dx *= tan_r;
    dy *= tan_r;
    V3 ax, ay;
    SFX(make_frame)(rev, &ax, &ay);
    *dir            = SFX(normalize)(SFX(add)(SFX(add)(rev, SFX(scl)(ax, dx)), SFX(scl)(ay, dy)));
    const R tan_t   = RSQRT(dx * dx + dy * dy);
    const R cos_t   = RC(1.0) / RSQRT(RC(1.0) + tan_t * tan_t);
    *pdf            = ORC_INV_PI / (tan_r * tan_r * cos_t * cos_t * cos_t);
    return SFX(ldf)(e->intensity);
}

/* LightSampler::sampleEmitterDirect: emitter.h:168-202.  Returns the
 * selection-scaled value (value * n); *which = chosen emitter. */
static V3 SFX(sample_emitters)(const orc_scene* s, R ux, R uy, V3* dir, R* pdf, uint32_t* which)
{
    const uint32_t n = s->d.nemitters;
    *pdf             = RC(0.0);
    if (n == 0)
        return SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
    uint32_t idx = (uint32_t)RFLOOR(ux * (R)n);
    if (idx > n - 1) /* ux < 1 keeps idx < n; the guard only protects memory */
        idx = n - 1;
    const R ux2 = ux * (R)n - (R)idx;
    *which      = idx;
    const orc_emitter* e = &s->emitters[idx];
    V3                 value;
    if (e->type == ORC_EMITTER_DOME)
        value = SFX(dome_sample)(&s->d.dome, ux2, uy, dir, pdf);
    else
        value = SFX(sun_sample)(e, (R)s->sun_tan[idx], ux2, uy, dir, pdf);
    value = SFX(scl)(value, (R)n);
    if (*pdf == RC(0.0))
        return SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
    return value;
}

/* ------------------------------------------------------------ geometry */

typedef struct
{
    V3  o, d;
    R   tmin, tmax;
    int kx, ky, kz;
    R   Sx, Sy, Sz;
    V3  inv; /* box-test reciprocals; zero components replaced by 1e-20 */
} SFX(ray);

static inline R SFX(comp)(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

/* Per-ray set-up of the watertight test (Woop, Benthin, Wald, "Watertight
 * Ray/Triangle Intersection", JCGT 2(1), 2013). */
static void SFX(ray_setup)(SFX(ray) * r)
{
    const R ax = RFABS(r->d.x), ay = RFABS(r->d.y), az = RFABS(r->d.z);
    int     kz = 0;
    if (ay > ax)
        kz = 1;
    if (az > (kz == 0 ? ax : ay))
        kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (SFX(comp)(r->d, kz) < RC(0.0))
    {
        const int t = kx;
        kx          = ky;
        ky          = t;
    }
    r->kx = kx;
    r->ky = ky;
    r->kz = kz;
    const R dz = SFX(comp)(r->d, kz);
    r->Sz      = RC(1.0) / dz;
    r->Sx      = SFX(comp)(r->d, kx) * r->Sz;
    r->Sy      = SFX(comp)(r->d, ky) * r->Sz;
    const R tiny = RC(1e-20);
    r->inv = SFX(mk)(RC(1.0) / (RFABS(r->d.x) > tiny ? r->d.x : (r->d.x < RC(0.0) ? -tiny : tiny)),
                     RC(1.0) / (RFABS(r->d.y) > tiny ? r->d.y : (r->d.y < RC(0.0) ? -tiny : tiny)),
                     RC(1.0) / (RFABS(r->d.z) > tiny ? r->d.z : (r->d.z < RC(0.0) ? -tiny : tiny)));
}

/* Watertight two-sided test.  Returns 1 with t, b1 (weight of v1), b2
 * (weight of v2) when tmin <= t <= tmax. */
static int SFX(tri_test)(const SFX(ray) * r, V3 v0, V3 v1, V3 v2, R* t, R* b1, R* b2)
{
    const V3 A = SFX(sub)(v0, r->o), B = SFX(sub)(v1, r->o), C = SFX(sub)(v2, r->o);
    const R  Az0 = SFX(comp)(A, r->kz), Bz0 = SFX(comp)(B, r->kz), Cz0 = SFX(comp)(C, r->kz);
    const R  Ax = SFX(comp)(A, r->kx) - r->Sx * Az0;
    const R  Ay = SFX(comp)(A, r->ky) - r->Sy * Az0;
    const R  Bx = SFX(comp)(B, r->kx) - r->Sx * Bz0;
    const R  By = SFX(comp)(B, r->ky) - r->Sy * Bz0;
    const R  Cx = SFX(comp)(C, r->kx) - r->Sx * Cz0;
    const R  Cy = SFX(comp)(C, r->ky) - r->Sy * Cz0;
    R        U  = Cx * By - Cy * Bx;
    R        V  = Ax * Cy - Ay * Cx;
    R        W  = Bx * Ay - By * Ax;
#if ORC_IS_F32
    if (U == 0.0f || V == 0.0f || W == 0.0f)
    {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
#endif
    if ((U < RC(0.0) || V < RC(0.0) || W < RC(0.0)) && (U > RC(0.0) || V > RC(0.0) || W > RC(0.0)))
        return 0;
    const R det = U + V + W;
    if (det == RC(0.0))
        return 0;
    const R Az = r->Sz * Az0, Bz = r->Sz * Bz0, Cz = r->Sz * Cz0;
    const R T  = U * Az + V * Bz + W * Cz;
    const R tt = T / det;
    if (!(tt >= r->tmin && tt <= r->tmax))
        return 0;
    *t  = tt;
    *b1 = V / det;
    *b2 = W / det;
    return 1;
}

/* Conservative slab test: the interval is widened by 1e-5 relative so that
 * a triangle hit computed inside a leaf box is never culled by rounding. */
static int SFX(box_test)(const SFX(ray) * r, const float lo[3], const float hi[3], R tmax, R* tnear)
{
    const R t0x = ((R)lo[0] - r->o.x) * r->inv.x, t1x = ((R)hi[0] - r->o.x) * r->inv.x;
    const R t0y = ((R)lo[1] - r->o.y) * r->inv.y, t1y = ((R)hi[1] - r->o.y) * r->inv.y;
    const R t0z = ((R)lo[2] - r->o.z) * r->inv.z, t1z = ((R)hi[2] - r->o.z) * r->inv.z;
    R       tn  = RFMAX(RFMAX(RFMIN(t0x, t1x), RFMIN(t0y, t1y)), RFMAX(RFMIN(t0z, t1z), r->tmin));
    R       tf  = RFMIN(RFMIN(RFMAX(t0x, t1x), RFMAX(t0y, t1y)), RFMIN(RFMAX(t0z, t1z), tmax));
    tn          = tn * RC(0.99999);
    tf          = tf * RC(1.00001);
    *tnear      = tn;
    return tn <= tf;
}

typedef struct
{
    R        t, b1, b2;
    uint32_t inst, prim;
} SFX(hit);

/* Total order of candidate hits: smaller t, then instance, then primitive. */
static inline int SFX(better)(R t, uint32_t inst, uint32_t prim, const SFX(hit) * h)
{
    if (t != h->t)
        return t < h->t;
    if (inst != h->inst)
        return inst < h->inst;
    return prim < h->prim;
}

/* Transform a world ray into an instance's object space (affine, no
 * renormalisation, so t is preserved). */
static void SFX(to_object)(const orc_scene* s, uint32_t inst, const SFX(ray) * w, SFX(ray) * o)
{
#if ORC_IS_F32
    const float* m = s->inv[inst];
#else
    const double* m = s->invd[inst];
#endif
    o->o = SFX(mk)(RFMA((R)m[0], w->o.x, RFMA((R)m[1], w->o.y, RFMA((R)m[2], w->o.z, (R)m[3]))),
                   RFMA((R)m[4], w->o.x, RFMA((R)m[5], w->o.y, RFMA((R)m[6], w->o.z, (R)m[7]))),
                   RFMA((R)m[8], w->o.x, RFMA((R)m[9], w->o.y, RFMA((R)m[10], w->o.z, (R)m[11]))));
    o->d = SFX(mk)(RFMA((R)m[0], w->d.x, RFMA((R)m[1], w->d.y, (R)m[2] * w->d.z)),
                   RFMA((R)m[4], w->d.x, RFMA((R)m[5], w->d.y, (R)m[6] * w->d.z)),
                   RFMA((R)m[8], w->d.x, RFMA((R)m[9], w->d.y, (R)m[10] * w->d.z)));
    o->tmin = w->tmin;
    o->tmax = w->tmax;
    SFX(ray_setup)(o);
}

static inline void SFX(tri_verts)(const orc_mesh* me, uint32_t prim, V3* v0, V3* v1, V3* v2)
{
    const uint32_t* id = me->indices + 3 * (size_t)prim;
    *v0                = SFX(ldf)(me->vertices + 3 * (size_t)id[0]);
    *v1                = SFX(ldf)(me->vertices + 3 * (size_t)id[1]);
    *v2                = SFX(ldf)(me->vertices + 3 * (size_t)id[2]);
}

/* Push the children of an inner node, the nearer one last (popped first). */
static inline void SFX(push_children)(const orc_bvh* b, const orc_node* nd, const SFX(ray) * r, R tmax, uint32_t* st, int* sp)
{
    const orc_node* c0 = &b->nodes[nd->a];
    const orc_node* c1 = &b->nodes[nd->a + 1];
    R               t0, t1;
    const int       h0 = SFX(box_test)(r, c0->lo, c0->hi, tmax, &t0);
    const int       h1 = SFX(box_test)(r, c1->lo, c1->hi, tmax, &t1);
    if (h0 && h1)
    {
        if (t0 <= t1)
        {
            st[(*sp)++] = nd->a + 1;
            st[(*sp)++] = nd->a;
        }
        else
        {
            st[(*sp)++] = nd->a;
            st[(*sp)++] = nd->a + 1;
        }
    }
    else if (h0)
        st[(*sp)++] = nd->a;
    else if (h1)
        st[(*sp)++] = nd->a + 1;
}

/* Closest hit (any_hit = 0) or occlusion (any_hit = 1) through the oracle's
 * two-level BVH2.  Replaces OptiX traversal (optix.cu:64-74, :99-109). */
static int SFX(traverse)(const orc_scene* s, const SFX(ray) * wr, int any_hit, SFX(hit) * best)
{
    best->t    = wr->tmax;
    best->inst = UINT32_MAX;
    best->prim = UINT32_MAX;
    int      found = 0;
    uint32_t tstack[256];
    int      tsp = 0;
    if (s->tlas.nnodes == 0 || s->d.ninstances == 0)
        return 0;
    tstack[tsp++] = 0;
    while (tsp > 0)
    {
        const orc_node* nd = &s->tlas.nodes[tstack[--tsp]];
        R               tn;
        if (!SFX(box_test)(wr, nd->lo, nd->hi, best->t, &tn))
            continue;
        if (nd->n == 0)
        {
            SFX(push_children)(&s->tlas, nd, wr, best->t, tstack, &tsp);
            continue;
        }
        for (uint32_t k = 0; k < nd->n; ++k)
        {
            const uint32_t  inst = s->tlas.prims[nd->a + k];
            const orc_bvh*  bl   = &s->blas[s->instances[inst].mesh];
            const orc_mesh* me   = &s->meshes[s->instances[inst].mesh];
            SFX(ray)        orr;
            SFX(to_object)(s, inst, wr, &orr);
            if (bl->nnodes == 0 || me->nt == 0)
                continue;
            uint32_t bstack[256];
            int      bsp = 0;
            bstack[bsp++] = 0;
            while (bsp > 0)
            {
                const orc_node* bn = &bl->nodes[bstack[--bsp]];
                R               btn;
                if (!SFX(box_test)(&orr, bn->lo, bn->hi, best->t, &btn))
                    continue;
                if (bn->n == 0)
                {
                    SFX(push_children)(bl, bn, &orr, best->t, bstack, &bsp);
                    continue;
                }
                for (uint32_t j = 0; j < bn->n; ++j)
                {
                    const uint32_t prim = bl->prims[bn->a + j];
                    V3             v0, v1, v2;
                    SFX(tri_verts)(me, prim, &v0, &v1, &v2);
                    R t, b1, b2;
                    if (!SFX(tri_test)(&orr, v0, v1, v2, &t, &b1, &b2))
                        continue;
                    if (any_hit)
                        return 1;
                    if (SFX(better)(t, inst, prim, best))
                    {
                        best->t    = t;
                        best->b1   = b1;
                        best->b2   = b2;
                        best->inst = inst;
                        best->prim = prim;
                        found      = 1;
                    }
                }
            }
        }
    }
    return found;
}

/* Every triangle of every instance, no hierarchy. */
static int SFX(brute)(const orc_scene* s, const SFX(ray) * wr, int any_hit, SFX(hit) * best)
{
    best->t    = wr->tmax;
    best->inst = UINT32_MAX;
    best->prim = UINT32_MAX;
    int found  = 0;
    for (uint32_t inst = 0; inst < s->d.ninstances; ++inst)
    {
        const orc_mesh* me = &s->meshes[s->instances[inst].mesh];
        SFX(ray)        orr;
        SFX(to_object)(s, inst, wr, &orr);
        for (uint32_t prim = 0; prim < me->nt; ++prim)
        {
            V3 v0, v1, v2;
            SFX(tri_verts)(me, prim, &v0, &v1, &v2);
            R t, b1, b2;
            if (!SFX(tri_test)(&orr, v0, v1, v2, &t, &b1, &b2))
                continue;
            if (any_hit)
                return 1;
            if (SFX(better)(t, inst, prim, best))
            {
                best->t    = t;
                best->b1   = b1;
                best->b2   = b2;
                best->inst = inst;
                best->prim = prim;
                found      = 1;
            }
        }
    }
    return found;
}

/* ------------------------------------------------------------ shading */

typedef struct
{
    V3       p;    /* world position */
    V3       ng;   /* geometric normal, flipped to the shading side */
    SFX(frame) f;  /* s = dpdv, t = dpdu, n = Ns */
    R        u, v; /* surface uv */
    uint32_t material;
} SFX(surface);

/* Closest-hit epilogue: optix.cu:130-228.  Corrections: the degenerate-Ns
 * test happens before normalising (the reference tests after, where a zero
 * vector is already NaN), and the shading frame is built on Ns (D10:
 * raygeneration.cu:324-330 passes ns and ng in swapped order). */
static void SFX(surface_interaction)(const orc_scene* s, const SFX(hit) * h, SFX(surface) * si)
{
    const orc_instance* in = &s->instances[h->inst];
    const orc_mesh*     me = &s->meshes[in->mesh];
    const uint32_t*     id = me->indices + 3 * (size_t)h->prim;
    const R             b1 = h->b1, b2 = h->b2;
    const R             b0 = RC(1.0) - b1 - b2;
    const V3 p0 = SFX(ldf)(me->vertices + 3 * (size_t)id[0]);
    const V3 p1 = SFX(ldf)(me->vertices + 3 * (size_t)id[1]);
    const V3 p2 = SFX(ldf)(me->vertices + 3 * (size_t)id[2]);
    const V3 po = SFX(add)(SFX(add)(SFX(sclL)(b0, p0), SFX(sclL)(b1, p1)), SFX(sclL)(b2, p2));
    if (me->uvs)
    {
        const float* t0 = me->uvs + 2 * (size_t)id[0];
        const float* t1 = me->uvs + 2 * (size_t)id[1];
        const float* t2 = me->uvs + 2 * (size_t)id[2];
        si->u = b0 * (R)t0[0] + b1 * (R)t1[0] + b2 * (R)t2[0];
        si->v = b0 * (R)t0[1] + b1 * (R)t1[1] + b2 * (R)t2[1];
    }
    else
    {
        si->u = b0;
        si->v = b1;
    }
    const V3 ngo = SFX(cross)(SFX(sub)(p1, p0), SFX(sub)(p2, p0));
    V3       nso = ngo;
    if (me->normals)
    {
        const V3 n0 = SFX(ldf)(me->normals + 3 * (size_t)id[0]);
        const V3 n1 = SFX(ldf)(me->normals + 3 * (size_t)id[1]);
        const V3 n2 = SFX(ldf)(me->normals + 3 * (size_t)id[2]);
        nso         = SFX(add)(SFX(add)(SFX(sclL)(b0, n0), SFX(sclL)(b1, n1)), SFX(sclL)(b2, n2));
    }
    V3 tgo = SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
    if (me->tangents)
    {
        const V3 a0 = SFX(ldf)(me->tangents + 3 * (size_t)id[0]);
        const V3 a1 = SFX(ldf)(me->tangents + 3 * (size_t)id[1]);
        const V3 a2 = SFX(ldf)(me->tangents + 3 * (size_t)id[2]);
        tgo         = SFX(add)(SFX(add)(SFX(sclL)(b0, a0), SFX(sclL)(b1, a1)), SFX(sclL)(b2, a2));
    }
    /* object -> world: points and tangents by M, normals by inverse transpose */
    const float* M = in->xform;
#if ORC_IS_F32
    const float* I = s->inv[h->inst];
#else
    const double* I = s->invd[h->inst];
#endif
    si->p = SFX(mk)((R)M[0] * po.x + (R)M[1] * po.y + (R)M[2] * po.z + (R)M[3],
                    (R)M[4] * po.x + (R)M[5] * po.y + (R)M[6] * po.z + (R)M[7],
                    (R)M[8] * po.x + (R)M[9] * po.y + (R)M[10] * po.z + (R)M[11]);
#define ORC_NRM(v) SFX(mk)((R)I[0] * v.x + (R)I[4] * v.y + (R)I[8] * v.z, (R)I[1] * v.x + (R)I[5] * v.y + (R)I[9] * v.z, (R)I[2] * v.x + (R)I[6] * v.y + (R)I[10] * v.z)
    const V3 ngw = ORC_NRM(ngo);
    const V3 nsw = ORC_NRM(nso);
#undef ORC_NRM
    const V3 tgw = SFX(mk)((R)M[0] * tgo.x + (R)M[1] * tgo.y + (R)M[2] * tgo.z,
                           (R)M[4] * tgo.x + (R)M[5] * tgo.y + (R)M[6] * tgo.z,
                           (R)M[8] * tgo.x + (R)M[9] * tgo.y + (R)M[10] * tgo.z);
    V3 Ng = SFX(normalize)(ngw);
    V3 Ns = SFX(dot)(nsw, nsw) == RC(0.0) ? Ng : SFX(normalize)(nsw);
    if (SFX(dot)(Ng, Ns) < RC(0.0))
        Ng = SFX(neg)(Ng);
    V3 dpdu = tgw, dpdv;
    if (SFX(dot)(dpdu, dpdu) > RC(0.0))
    {
        dpdu = SFX(normalize)(dpdu);
        dpdv = SFX(cross)(Ns, dpdu);
        if (SFX(dot)(dpdv, dpdv) > RC(0.0))
        {
            dpdv = SFX(normalize)(dpdv);
            dpdu = SFX(cross)(dpdv, Ns);
        }
        else
            SFX(make_frame)(Ns, &dpdu, &dpdv);
    }
    else
        SFX(make_frame)(Ns, &dpdu, &dpdv);
    si->ng       = Ng;
    si->f.s      = dpdv;
    si->f.t      = dpdu;
    si->f.n      = Ns;
    si->material = in->material;
}

/* Material::fixNormalPerturbingShadingFrame (material.h:236-285): the
 * IRay pull-up of a normal-mapped frame on the side of wo. */
static void SFX(pull_up)(SFX(frame) * f, V3 ng, V3 wo)
{
    if (SFX(dot)(wo, ng) < RC(0.0))
    {
        f->n = SFX(neg)(f->n);
        ng   = SFX(neg)(ng);
    }
    if (!(SFX(dot)(wo, ng) > RC(0.0)))
        return;
    const V3 r = SFX(reflect)(wo, f->n);
    if (SFX(dot)(r, ng) > RC(0.0))
        return;
    const R a = RFABS(SFX(dot)(r, ng)), b = SFX(dot)(f->n, ng);
    if (!(b > RC(0.0)))
        return;
    f->n = SFX(normalize)(SFX(add)(SFX(normalize)(SFX(add)(r, SFX(scl)(SFX(scl)(f->n, a / b), (R)(float)1.1))), wo));
    f->t = SFX(normalize)(SFX(sub)(f->t, SFX(scl)(f->n, SFX(dot)(f->t, f->n))));
    f->s = SFX(normalize)(SFX(cross)(f->n, f->t));
    if (SFX(dot)(f->n, ng) < RC(0.0))
    {
        f->n = SFX(neg)(f->n);
        f->t = SFX(neg)(f->t);
    }
}

/* Texture-driven material inputs at a surface point: Material::getBSDF's
 * reflectance and roughness lookups (diffusemtl.h:22-31, metal.h:36-55;
 * roughness = first channel, squared by mtl_alpha) and the normal-mapped
 * frame of Material::getShadingFrame (material.h:124-219), taken in the
 * hit's tangent frame (dpdu, dpdv, Ns) (D10), then the pull-up on diffuse
 * and metal.  A normal-map texel with no direction (0.5, 0.5, 0.5) leaves
 * the frame unchanged. */
static void SFX(material_at)(const orc_scene* s, orc_material* m, SFX(surface) * si, V3 wo)
{
    R c[4];
    if (m->reflectance_tex)
    {
        const orc_texture* t = &s->d.textures[m->reflectance_tex - 1u];
        SFX(tex_eval)(t->rgba, t->w, t->h, si->u, si->v, c);
        m->reflectance[0] = (float)c[0];
        m->reflectance[1] = (float)c[1];
        m->reflectance[2] = (float)c[2];
    }
    if (m->roughness_tex)
    {
        const orc_texture* t = &s->d.textures[m->roughness_tex - 1u];
        SFX(tex_eval)(t->rgba, t->w, t->h, si->u, si->v, c);
        m->roughness = (float)c[0];
    }
    if (m->normal_tex)
    {
        const orc_texture* t = &s->d.textures[m->normal_tex - 1u];
        SFX(tex_eval)(t->rgba, t->w, t->h, si->u, si->v, c);
        const V3 nt = SFX(mk)(c[0] * RC(2.0) - RC(1.0), c[1] * RC(2.0) - RC(1.0), c[2] * RC(2.0) - RC(1.0));
        if (SFX(dot)(nt, nt) > RC(0.0))
        {
            V3 n  = SFX(normalize)(SFX(to_world)(&si->f, SFX(normalize)(nt)));
            V3 tm = SFX(sub)(si->f.t, SFX(scl)(n, SFX(dot)(si->f.t, n)));
            V3 sm;
            if (SFX(dot)(tm, tm) > RC(0.0))
            {
                tm = SFX(normalize)(tm);
                sm = SFX(normalize)(SFX(cross)(n, tm));
            }
            else
                SFX(make_frame)(n, &tm, &sm);
            if (SFX(dot)(n, si->ng) < RC(0.0))
            {
                n  = SFX(neg)(n);
                tm = SFX(neg)(tm);
            }
            si->f.s = sm;
            si->f.t = tm;
            si->f.n = n;
            if (m->type == ORC_MTL_DIFFUSE || m->type == ORC_MTL_METAL)
                SFX(pull_up)(&si->f, si->ng, wo);
        }
    }
}

/* ------------------------------------------------------------ integrator */

#define ORC_EXTRA_DRAW(depth, slot) (0x80000000u + 8u * (uint32_t)(depth) + (uint32_t)(slot))
#define ORC_SLOT_RR 0u
#define ORC_SLOT_LOBE 1u

static inline R SFX(draw)(uint32_t px, uint32_t py, uint32_t z, uint32_t k) { return (R)orc_draw(px, py, z, k); }

/* One camera path: generateCameraRays (raygeneration.cu:59-103), then the
 * bounce loop of WavefrontPathTracingIntegrator::render (path.cpp:141-211)
 * with evaluateEscapedRays (raygeneration.cu:105-136) and
 * evaluateMaterialsAndLightsPathTracing (raygeneration.cu:285-456), with:
 *  D1/D2 plain float64 sum per pixel (no running mean, no sRGB round trip);
 *  D3 escaped MIS contribution = T_new * (Le * w_bsdf);
 *  D8 NEE weight 1 for the directional emitter (BSDF rays never see it);
 *  D9 light pdf in MIS weights scaled by the selection probability 1/n;
 *  D11 the path continues whatever the dome pdf of the new direction;
 *  D13 Russian roulette from rr_depth (survival q = min(1, max(T)));
 *  D15 shadow rays use any-hit;
 *  D16 depth from the parameters. */
static void SFX(path)(const orc_scene* s, const orc_render_params* P, uint32_t px, uint32_t py, uint32_t sample,
                      R L[3], orc_counts* cnt)
{
    const uint32_t z   = (P->view << 20) | sample;
    const orc_dome* dm = &s->d.dome;
    const int has_dome = s->d.dome_index >= 0;
    const R   nl       = (R)s->d.nemitters;
    L[0] = L[1] = L[2] = RC(0.0);

    const R  u0 = SFX(draw)(px, py, z, 1u), u1 = SFX(draw)(px, py, z, 2u);
    const R  sx = ((R)px + u0) / (R)P->width;
    const R  sy = ((R)py + u1) / (R)P->height;
    const V3 d00 = SFX(ldf)(P->cam_d00), ddu = SFX(ldf)(P->cam_ddu), ddv = SFX(ldf)(P->cam_ddv);
    SFX(ray) ray;
    ray.o    = SFX(ldf)(P->cam_pos);
    ray.d    = SFX(normalize)(SFX(add)(SFX(add)(d00, SFX(sclL)(sx, ddu)), SFX(sclL)(sy, ddv)));
    ray.tmin = RC(0.0);
    ray.tmax = (R)FLT_MAX;
    V3 T      = SFX(mk)(RC(1.0), RC(1.0), RC(1.0));
    R  prev_pdf = RC(0.0);
    for (uint32_t depth = 1;; ++depth)
    {
        SFX(ray_setup)(&ray);
        SFX(hit) h;
        cnt->rays[depth]++;
        if (!SFX(traverse)(s, &ray, 0, &h))
        {
            if (has_dome)
            {
                const V3 le = SFX(dome_le)(dm, ray.d);
                if (depth == 1)
                {
                    L[0] += le.x;
                    L[1] += le.y;
                    L[2] += le.z;
                }
                else
                {
                    const R lp = SFX(dome_pdf)(dm, ray.d) / nl;
                    const R w  = prev_pdf / (prev_pdf + lp);
                    const V3 c = SFX(mul)(T, SFX(scl)(le, w));
                    L[0] += c.x;
                    L[1] += c.y;
                    L[2] += c.z;
                }
            }
            break;
        }
        cnt->hits[depth]++;
        if (depth >= P->max_depth)
            break;
        SFX(surface) si;
        SFX(surface_interaction)(s, &h, &si);
        const V3     wo   = SFX(neg)(ray.d);
        orc_material mloc = s->materials[si.material];
        SFX(material_at)(s, &mloc, &si, wo);
        const orc_material* mt  = &mloc;
        const V3            woL = SFX(to_local)(&si.f, wo);
        const uint32_t k0 = 2u + 4u * (depth - 1u);

        /* next-event estimation (raygeneration.cu:310-384) */
        {
            const R  lu = SFX(draw)(px, py, z, k0 + 1u), lv = SFX(draw)(px, py, z, k0 + 2u);
            V3       ldir = SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
            R        lpdf = RC(0.0);
            uint32_t which = 0;
            const V3 value = SFX(sample_emitters)(s, lu, lv, &ldir, &lpdf, &which);
            if (SFX(nonzero)(value) && lpdf > RC(0.0))
            {
                const V3 wiL = SFX(to_local)(&si.f, ldir);
                const V3 f   = SFX(bsdf_eval)(mt, wiL, woL);
                if (SFX(nonzero)(f))
                {
                    R weight = RC(1.0);
                    if (s->emitters[which].type == ORC_EMITTER_DOME)
                    {
                        const R bp = SFX(bsdf_pdf)(mt, wiL, woL);
                        const R lp = lpdf / nl;
                        weight     = lp / (lp + bp);
                    }
                    const R  cosv = RFABS(wiL.z);
                    const V3 Li   = SFX(mk)(T.x * value.x * f.x * cosv * weight / lpdf,
                                          T.y * value.y * f.y * cosv * weight / lpdf,
                                          T.z * value.z * f.z * cosv * weight / lpdf);
                    cnt->shadow[depth]++;
                    SFX(ray) sr;
                    sr.o    = si.p;
                    sr.d    = ldir;
                    sr.tmin = (R)(float)0.001;
                    sr.tmax = (R)FLT_MAX;
                    SFX(ray_setup)(&sr);
                    SFX(hit) sh;
                    if (!SFX(traverse)(s, &sr, 1, &sh))
                    {
                        cnt->visible[depth]++;
                        L[0] += Li.x;
                        L[1] += Li.y;
                        L[2] += Li.z;
                    }
                }
            }
        }

        /* BSDF sampling (raygeneration.cu:391-453) */
        const R bu = SFX(draw)(px, py, z, k0 + 3u), bv = SFX(draw)(px, py, z, k0 + 4u);
        const R ul = SFX(draw)(px, py, z, ORC_EXTRA_DRAW(depth, ORC_SLOT_LOBE));
        V3      wiL = SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
        R       bpdf = RC(0.0);
        const V3 f = SFX(bsdf_sample)(mt, woL, bu, bv, ul, &wiL, &bpdf);
        if (!SFX(nonzero)(f) || !(bpdf > RC(0.0)))
            break;
        const R cosv = RFABS(wiL.z);
        /* pathThroughput *= bsdfVal * |cos| / pdf  (raygeneration.cu:430) */
        T = SFX(mk)(T.x * (f.x * cosv / bpdf), T.y * (f.y * cosv / bpdf), T.z * (f.z * cosv / bpdf));
        ray.o    = si.p;
        ray.d    = SFX(to_world)(&si.f, wiL);
        ray.tmin = (R)(float)0.001;
        ray.tmax = (R)FLT_MAX;
        prev_pdf = bpdf;
        if (P->rr_depth > 0 && depth >= P->rr_depth)
        {
            const R q = RFMIN(RC(1.0), RFMAX(T.x, RFMAX(T.y, T.z)));
            const R ur = SFX(draw)(px, py, z, ORC_EXTRA_DRAW(depth, ORC_SLOT_RR));
            if (!(ur < q))
                break;
            T = SFX(mk)(T.x / q, T.y / q, T.z / q);
        }
    }
}

typedef struct
{
    const orc_scene*         s;
    const orc_render_params* p;
    uint32_t                 s0, s1;
    double*                  sum3;
    orc_counts*              per_worker;
    uint32_t                 y0;
    const uint32_t*          rows; /* NULL: row i is y0 + i */
} SFX(render_ctx);

static void SFX(render_row)(void* vctx, uint32_t i, uint32_t worker)
{
    SFX(render_ctx)* c = (SFX(render_ctx)*)vctx;
    const uint32_t   y = c->rows ? c->rows[i] : c->y0 + i;
    for (uint32_t x = 0; x < c->p->width; ++x)
    {
        double acc[3] = {0.0, 0.0, 0.0};
        for (uint32_t smp = c->s0; smp < c->s1; ++smp)
        {
            R L[3];
            SFX(path)(c->s, c->p, x, y, smp, L, &c->per_worker[worker]);
            acc[0] += (double)L[0];
            acc[1] += (double)L[1];
            acc[2] += (double)L[2];
        }
        double* o = c->sum3 + 3 * ((size_t)i * c->p->width + x);
        o[0] += acc[0];
        o[1] += acc[1];
        o[2] += acc[2];
    }
}

/* ------------------------------------------------------------ exports */

void SFX(orc_bsdf_eval)(const orc_material* m, const float* wi, const float* wo, uint32_t n, float* f3, float* pdf)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        const V3 a = SFX(ldf)(wi + 3 * (size_t)i), b = SFX(ldf)(wo + 3 * (size_t)i);
        const V3 f = SFX(bsdf_eval)(&m[i], a, b);
        f3[3 * i]     = (float)f.x;
        f3[3 * i + 1] = (float)f.y;
        f3[3 * i + 2] = (float)f.z;
        pdf[i]        = (float)SFX(bsdf_pdf)(&m[i], a, b);
    }
}

void SFX(orc_bsdf_sample)(const orc_material* m, const float* wo, const float* u3, uint32_t n, float* wi, float* f3, float* pdf)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        V3       w = SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
        R        p = RC(0.0);
        const V3 f = SFX(bsdf_sample)(&m[i], SFX(ldf)(wo + 3 * (size_t)i), (R)u3[3 * i], (R)u3[3 * i + 1], (R)u3[3 * i + 2], &w, &p);
        wi[3 * i]     = (float)w.x;
        wi[3 * i + 1] = (float)w.y;
        wi[3 * i + 2] = (float)w.z;
        f3[3 * i]     = (float)f.x;
        f3[3 * i + 1] = (float)f.y;
        f3[3 * i + 2] = (float)f.z;
        pdf[i]        = (float)p;
    }
}

void SFX(orc_warp_batch)(const float* u2, uint32_t n, float* disk2, float* hemi3)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        R x, y;
        SFX(concentric_disk)((R)u2[2 * i], (R)u2[2 * i + 1], &x, &y);
        disk2[2 * i]     = (float)x;
        disk2[2 * i + 1] = (float)y;
        const V3 h       = SFX(cosine_hemisphere)((R)u2[2 * i], (R)u2[2 * i + 1]);
        hemi3[3 * i]     = (float)h.x;
        hemi3[3 * i + 1] = (float)h.y;
        hemi3[3 * i + 2] = (float)h.z;
    }
}

void SFX(orc_ggx_batch)(float alpha, const float* m3, const float* v3, const float* u2, uint32_t n, float* D, float* G1,
                        float* ms3, float* mpdf)
{
    const R a = SFX(ggx_alpha)((R)alpha);
    for (uint32_t i = 0; i < n; ++i)
    {
        const V3 m = SFX(ldf)(m3 + 3 * (size_t)i), v = SFX(ldf)(v3 + 3 * (size_t)i);
        D[i]       = (float)SFX(ggx_D)(a, m);
        G1[i]      = (float)SFX(ggx_G1)(a, v, m);
        R        p;
        const V3 s = SFX(ggx_sample_all)(a, (R)u2[2 * i], (R)u2[2 * i + 1], &p);
        ms3[3 * i]     = (float)s.x;
        ms3[3 * i + 1] = (float)s.y;
        ms3[3 * i + 2] = (float)s.z;
        mpdf[i]        = (float)p;
    }
}

void SFX(orc_fresnel_batch)(const float* cos_i, uint32_t n, const float* eta3, const float* k3, float ior, float* fc3,
                            float* fd, float* cos_t)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        const R c = (R)cos_i[i];
        for (int ch = 0; ch < 3; ++ch)
            fc3[3 * i + ch] = (float)SFX(fresnel_conductor1)(c, (R)eta3[ch], (R)k3[ch]);
        R ct;
        fd[i]    = (float)SFX(fresnel_dielectric_ext)(c, &ct, (R)ior);
        cos_t[i] = (float)ct;
    }
}

/* convertFromLinearTosRGB (libkernel/base/math.h:114-122), one channel: no
 * clamp; the reference's exponent 0.41666. */
static R SFX(srgb_encode)(R x)
{
    return x < RC(0.0031308) ? x * RC(12.92) : RC(1.055) * RPOW(x, RC(0.41666)) - RC(0.055);
}

/* convertsRGBToLinear (math.h:131-140), clamped to [0, 1]. */
static R SFX(srgb_decode)(R s)
{
    if (s <= RC(0.0))
        return RC(0.0);
    if (s >= RC(1.0))
        return RC(1.0);
    if (s <= RC(0.04045))
        return s / RC(12.92);
    return RPOW((s + RC(0.055)) / RC(1.055), RC(2.4));
}

/* pull_up on n frames: frame9 = s, t, n per item (in and out). */
void SFX(orc_pull_up)(const float* frame9, const float* ng3, const float* wo3, uint32_t n, float* out9)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        const float* a = frame9 + 9 * (size_t)i;
        SFX(frame) f;
        f.s     = (V3){(R)a[0], (R)a[1], (R)a[2]};
        f.t     = (V3){(R)a[3], (R)a[4], (R)a[5]};
        f.n     = (V3){(R)a[6], (R)a[7], (R)a[8]};
        V3 ng   = {(R)ng3[3 * (size_t)i], (R)ng3[3 * (size_t)i + 1], (R)ng3[3 * (size_t)i + 2]};
        V3 wo   = {(R)wo3[3 * (size_t)i], (R)wo3[3 * (size_t)i + 1], (R)wo3[3 * (size_t)i + 2]};
        SFX(pull_up)(&f, ng, wo);
        float* o = out9 + 9 * (size_t)i;
        o[0] = (float)f.s.x, o[1] = (float)f.s.y, o[2] = (float)f.s.z;
        o[3] = (float)f.t.x, o[4] = (float)f.t.y, o[5] = (float)f.t.z;
        o[6] = (float)f.n.x, o[7] = (float)f.n.y, o[8] = (float)f.n.z;
    }
}

void SFX(orc_srgb)(const float* in, uint32_t n, int decode, float* out)
{
    for (uint32_t i = 0; i < n; ++i)
        out[i] = (float)(decode ? SFX(srgb_decode)((R)in[i]) : SFX(srgb_encode)((R)in[i]));
}

void SFX(orc_tex_eval)(const float* rgba, uint32_t w, uint32_t h, const float* uv, uint32_t n, float* out4)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        R c[4];
        SFX(tex_eval)(rgba, w, h, (R)uv[2 * i], (R)uv[2 * i + 1], c);
        for (int k = 0; k < 4; ++k)
            out4[4 * i + k] = (float)c[k];
    }
}

void SFX(orc_emitter_sample)(const orc_scene* s, const float* u2, uint32_t n, float* dir3, float* value3, float* pdf, uint32_t* which)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        V3       d = SFX(mk)(RC(0.0), RC(0.0), RC(0.0));
        R        p = RC(0.0);
        uint32_t w = 0;
        const V3 v = SFX(sample_emitters)(s, (R)u2[2 * i], (R)u2[2 * i + 1], &d, &p, &w);
        dir3[3 * i]       = (float)d.x;
        dir3[3 * i + 1]   = (float)d.y;
        dir3[3 * i + 2]   = (float)d.z;
        value3[3 * i]     = (float)v.x;
        value3[3 * i + 1] = (float)v.y;
        value3[3 * i + 2] = (float)v.z;
        pdf[i]            = (float)p;
        which[i]          = w;
    }
}

void SFX(orc_dome_eval)(const orc_scene* s, const float* dir3, uint32_t n, float* le3, float* pdf)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        const V3 d  = SFX(ldf)(dir3 + 3 * (size_t)i);
        const V3 le = SFX(dome_le)(&s->d.dome, d);
        le3[3 * i]     = (float)le.x;
        le3[3 * i + 1] = (float)le.y;
        le3[3 * i + 2] = (float)le.z;
        pdf[i]         = (float)SFX(dome_pdf)(&s->d.dome, d);
    }
}

/* rays8: o.xyz, tmin, d.xyz, tmax */
void SFX(orc_trace)(const orc_scene* s, const float* rays8, uint32_t n, int brute, float* t, float* b12, uint32_t* inst, uint32_t* prim)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        const float* r = rays8 + 8 * (size_t)i;
        SFX(ray)     ray;
        ray.o    = SFX(mk)((R)r[0], (R)r[1], (R)r[2]);
        ray.tmin = (R)r[3];
        ray.d    = SFX(mk)((R)r[4], (R)r[5], (R)r[6]);
        ray.tmax = (R)r[7];
        SFX(ray_setup)(&ray);
        SFX(hit) h;
        const int found = brute ? SFX(brute)(s, &ray, 0, &h) : SFX(traverse)(s, &ray, 0, &h);
        t[i]            = found ? (float)h.t : -1.0f;
        b12[2 * i]      = found ? (float)h.b1 : 0.0f;
        b12[2 * i + 1]  = found ? (float)h.b2 : 0.0f;
        inst[i]         = found ? h.inst : UINT32_MAX;
        prim[i]         = found ? h.prim : UINT32_MAX;
    }
}

void SFX(orc_occluded)(const orc_scene* s, const float* rays8, uint32_t n, int brute, uint8_t* occ)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        const float* r = rays8 + 8 * (size_t)i;
        SFX(ray)     ray;
        ray.o    = SFX(mk)((R)r[0], (R)r[1], (R)r[2]);
        ray.tmin = (R)r[3];
        ray.d    = SFX(mk)((R)r[4], (R)r[5], (R)r[6]);
        ray.tmax = (R)r[7];
        SFX(ray_setup)(&ray);
        SFX(hit) h;
        occ[i] = (uint8_t)(brute ? SFX(brute)(s, &ray, 1, &h) : SFX(traverse)(s, &ray, 1, &h));
    }
}

static void SFX(render_list)(const orc_scene* s, const orc_render_params* p, uint32_t y0, const uint32_t* rows,
                             uint32_t nrows, uint32_t s0, uint32_t s1, uint32_t nthreads, double* sum3,
                             orc_counts* counts)
{
    if (nthreads == 0)
        nthreads = 1;
    orc_counts* pw = (orc_counts*)calloc(nthreads, sizeof(orc_counts));
    SFX(render_ctx) c = {s, p, s0, s1, sum3, pw, y0, rows};
    orc_parallel_rows(nthreads, 0, nrows, SFX(render_row), &c);
    if (counts)
    {
        for (uint32_t t = 0; t < nthreads; ++t)
            for (int b = 0; b <= ORC_MAX_DEPTH; ++b)
            {
                counts->rays[b] += pw[t].rays[b];
                counts->hits[b] += pw[t].hits[b];
                counts->shadow[b] += pw[t].shadow[b];
                counts->visible[b] += pw[t].visible[b];
            }
    }
    free(pw);
}

/* Render rows [y0, y1) of the frame, samples [s0, s1); sum3 holds
 * (y1 - y0) * width RGB float64 sums (accumulated into). */
void SFX(orc_render)(const orc_scene* s, const orc_render_params* p, uint32_t y0, uint32_t y1, uint32_t s0, uint32_t s1,
                     uint32_t nthreads, double* sum3, orc_counts* counts)
{
    SFX(render_list)(s, p, y0, NULL, y1 - y0, s0, s1, nthreads, sum3, counts);
}

/* Render an explicit list of rows (a stratified sample of the frame); sum3
 * holds nrows * width RGB sums in list order. */
void SFX(orc_render_rows)(const orc_scene* s, const orc_render_params* p, const uint32_t* rows, uint32_t nrows,
                          uint32_t s0, uint32_t s1, uint32_t nthreads, double* sum3, orc_counts* counts)
{
    SFX(render_list)(s, p, 0, rows, nrows, s0, s1, nthreads, sum3, counts);
}
