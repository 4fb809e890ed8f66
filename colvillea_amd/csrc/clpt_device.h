/*
 * clpt_device.h -- per-lane shading math of the gfx950 path tracer.
 *
 * Every routine restates a reference routine (cited at each definition) in
 * float32 with the reference's evaluation order, compiled with
 * -ffp-contract=off and correctly rounded divide/sqrt, and with the
 * elementary functions of clpt_elem.h, so that for identical inputs the
 * results are reproducible bit for bit on the CPU.
 */
#ifndef CLPT_DEVICE_H
#define CLPT_DEVICE_H

#include "clpt_elem.h"
#include "clpt_scene.h"

#define CLPT_D __device__ __forceinline__
#define CLPT_PI_F 3.14159265358979323846f     /* M_PIf, libkernel/base/math.h:14-16 */
#define CLPT_INV_PI_F 0.318309886183790671538f /* M_1_PIf, math.h:20-22 */
#define CLPT_RAY_EPS ((float)0.001)           /* kRayTOffsetEpsilon, config.h:11 */
#define CLPT_FLT_MAX 3.402823466e+38f

struct f3
{
    float x, y, z;
};

CLPT_D f3   mk3(float x, float y, float z) { return f3{x, y, z}; }
CLPT_D f3   ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
CLPT_D f3   add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CLPT_D f3   sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CLPT_D f3   mul3(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
CLPT_D f3   scl3(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
CLPT_D f3   sclL3(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
CLPT_D f3   neg3(f3 a) { return f3{-a.x, -a.y, -a.z}; }
CLPT_D float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CLPT_D f3   cross3(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
/* OWL normalize form: v * (1 / sqrt(v.v)) */
CLPT_D f3   norm3(f3 v) { return scl3(v, 1.0f / sqrtf(dot3(v, v))); }
CLPT_D bool nonzero3(f3 v) { return v.x > 0.0f || v.y > 0.0f || v.z > 0.0f; }

/* ------------------------------------------------------------ display encoding */

/* convertFromLinearTosRGB (libkernel/base/math.h:114-122), per channel: no
 * clamp, so HDR values above 1 keep rising; the exponent is the reference's
 * 0.41666 (not 1/2.4).  powf is the device library's. */
CLPT_D float srgb_encode(float x) { return x < 0.0031308f ? x * 12.92f : 1.055f * powf(x, 0.41666f) - 0.055f; }

/* convertsRGBToLinear (math.h:131-140): clamped to [0, 1]. */
CLPT_D float srgb_decode(float s)
{
    if (s <= 0.0f)
        return 0.0f;
    if (s >= 1.0f)
        return 1.0f;
    if (s <= 0.04045f)
        return s / 12.92f;
    return powf((s + 0.055f) / 1.055f, 2.4f);
}

/* ------------------------------------------------------------ RNG */

/* pcg4d: libkernel/base/sampler.h:16-36 */
CLPT_D uint32_t pcg4d_x(uint32_t x, uint32_t y, uint32_t z, uint32_t w, uint4* all = nullptr)
{
    x = x * 1664525u + 1013904223u;
    y = y * 1664525u + 1013904223u;
    z = z * 1664525u + 1013904223u;
    w = w * 1664525u + 1013904223u;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
    x ^= x >> 16u;
    y ^= y >> 16u;
    z ^= z >> 16u;
    w ^= w >> 16u;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
    if (all)
        *all = make_uint4(x, y, z, w);
    return x;
}

/* uintToFloat: sampler.h:39-42 */
CLPT_D float u2f01(uint32_t x) { return __uint_as_float(0x3f800000u | (x >> 9)) - 1.0f; }

/* Draw k of the counter-based sampler (rand pre-increments w: sampler.h:51-55). */
CLPT_D float rng_draw(uint32_t px, uint32_t py, uint32_t z, uint32_t k) { return u2f01(pcg4d_x(px, py, z, k)); }

/* Draws outside the reference's layout (Russian roulette, dielectric lobe,
 * guiding) use counters >= 2^31 so the reference's w = 1.. stream is intact. */
#define CLPT_EXTRA_DRAW(depth, slot) (0x80000000u + 8u * (uint32_t)(depth) + (uint32_t)(slot))
#define CLPT_SLOT_RR 0u
#define CLPT_SLOT_LOBE 1u
#define CLPT_SLOT_GUIDE 2u   /* BSDF or learned proposal at a guided vertex */
#define CLPT_SLOT_LOBESEL 3u /* lobe of the learned mixture */

/* ------------------------------------------------------------ warps */

/* warp::squareToUniformConcentricDisk: warp.h:25-50 (phi in double, from M_PI). */
CLPT_D void concentric_disk(float ux, float uy, float* ox, float* oy)
{
    const float r1 = 2.0f * ux - 1.0f;
    const float r2 = 2.0f * uy - 1.0f;
    float       phi, r;
    if (r1 == 0.0f && r2 == 0.0f)
    {
        r = phi = 0.0f;
    }
    else if (r1 * r1 > r2 * r2)
    {
        r   = r1;
        phi = (float)((3.14159265358979323846 / 4.0) * (double)(r2 / r1));
    }
    else
    {
        r   = r2;
        phi = (float)((3.14159265358979323846 / 2.0) - (double)(r1 / r2) * (3.14159265358979323846 / 4.0));
    }
    float s, c;
    clpt_sincosf(phi, &s, &c);
    *ox = r * c;
    *oy = r * s;
}

/* warp::squareToCosineHemisphere: warp.h:58-64 (radicand clamped at 0). */
CLPT_D f3 cosine_hemisphere(float ux, float uy)
{
    float px, py;
    concentric_disk(ux, uy, &px, &py);
    return f3{px, py, sqrtf(fmaxf(0.0f, 1.0f - px * px - py * py))};
}

/* makeFrame: math.h:33-48 */
CLPT_D void make_frame(f3 a, f3* b, f3* c)
{
    if (fabsf(a.x) > fabsf(a.y))
    {
        const float inv = 1.0f / sqrtf(a.x * a.x + a.z * a.z);
        *c              = f3{a.z * inv, 0.0f, -a.x * inv};
    }
    else
    {
        const float inv = 1.0f / sqrtf(a.y * a.y + a.z * a.z);
        *c              = f3{0.0f, a.z * inv, -a.y * inv};
    }
    *b = cross3(*c, a);
}

/* Frame: frame.h:24-130 (s, t, n; local = (v.t, v.s, v.n)). */
struct Frame
{
    f3 s, t, n;
};
CLPT_D f3 to_local(const Frame& f, f3 v) { return f3{dot3(v, f.t), dot3(v, f.s), dot3(v, f.n)}; }
CLPT_D f3 to_world(const Frame& f, f3 v) { return add3(add3(sclL3(v.x, f.t), sclL3(v.y, f.s)), sclL3(v.z, f.n)); }

/* ------------------------------------------------------------ GGX */

/* hypot2: math.h:55-73 */
CLPT_D float hypot2(float a, float b)
{
    float r;
    if (fabsf(a) > fabsf(b))
    {
        r = b / a;
        r = fabsf(a) * sqrtf(1.0f + r * r);
    }
    else if (b != 0.0f)
    {
        r = a / b;
        r = fabsf(b) * sqrtf(1.0f + r * r);
    }
    else
        r = 0.0f;
    return r;
}

/* MicrofacetDistribution::D: microfacet.h:79-90 */
CLPT_D float ggx_D(float alpha, f3 m)
{
    if (m.z <= 0.0f)
        return 0.0f;
    const float cos2 = m.z * m.z;
    const float be   = (m.x * m.x + m.y * m.y) / (alpha * alpha * cos2);
    const float root = (1.0f + be) * cos2;
    return 1.0f / (CLPT_PI_F * alpha * alpha * root * root);
}

/* Frame::sinTheta: frame.h:80-84 */
CLPT_D float sin_theta(f3 v)
{
    const float t = 1.0f - v.z * v.z;
    return t <= 0.0f ? 0.0f : sqrtf(t);
}

/* MicrofacetDistribution::smithG1: microfacet.h:40-54 */
CLPT_D float ggx_G1(float alpha, f3 v, f3 m)
{
    if (dot3(v, m) * v.z <= 0.0f)
        return 0.0f;
    const float tan_theta = fabsf(sin_theta(v) / v.z);
    if (tan_theta == 0.0f)
        return 1.0f;
    const float root = alpha * tan_theta;
    return 2.0f / (1.0f + hypot2(1.0f, root));
}

CLPT_D float ggx_G(float alpha, f3 a, f3 b, f3 m) { return ggx_G1(alpha, a, m) * ggx_G1(alpha, b, m); }

/* MicrofacetDistribution::sampleAll: microfacet.h:92-121 */
CLPT_D f3 ggx_sample_all(float alpha, float ux, float uy, float* pdf)
{
    const float alpha_sqr = alpha * alpha;
    float       sin_phi, cos_phi;
    clpt_sincosf((2.0f * CLPT_PI_F) * uy, &sin_phi, &cos_phi);
    const float tan2 = alpha_sqr * ux / (1.0f - ux);
    const float cosm = 1.0f / sqrtf(1.0f + tan2);
    const float temp = 1.0f + tan2 / alpha_sqr;
    float       p    = CLPT_INV_PI_F / (alpha_sqr * cosm * cosm * cosm * temp * temp);
    if (p < 1e-20f)
        p = 0.0f;
    *pdf             = p;
    const float sinm = sqrtf(fmaxf(0.0f, 1.0f - cosm * cosm));
    return f3{sinm * cos_phi, sinm * sin_phi, cosm};
}

/* ------------------------------------------------------------ Fresnel */

/* fresnelConductor: math.h:172-193, per channel */
CLPT_D float fresnel_conductor1(float c, float eta, float k)
{
    const float cos2  = c * c;
    const float sin2  = 1.0f - cos2;
    const float sin4  = sin2 * sin2;
    const float temp1 = eta * eta - k * k - sin2;
    const float a2pb2 = sqrtf(temp1 * temp1 + k * k * eta * eta * 4.0f);
    const float a     = sqrtf((a2pb2 + temp1) * 0.5f);
    const float term1 = a2pb2 + cos2;
    const float term2 = a * (2.0f * c);
    const float rs2   = (term1 - term2) / (term1 + term2);
    const float term3 = a2pb2 * cos2 + sin4;
    const float term4 = term2 * sin2;
    const float rp2   = rs2 * (term3 - term4) / (term3 + term4);
    return 0.5f * (rp2 + rs2);
}

/* fresnelDielectricExt: math.h:255-287 */
CLPT_D float fresnel_dielectric_ext(float cos_i_, float* cos_t_, float eta)
{
    if (eta == 1.0f)
    {
        *cos_t_ = -cos_i_;
        return 0.0f;
    }
    const float scale = (cos_i_ > 0.0f) ? 1.0f / eta : eta;
    const float ct2   = 1.0f - (1.0f - cos_i_ * cos_i_) * (scale * scale);
    if (ct2 <= 0.0f)
    {
        *cos_t_ = 0.0f;
        return 1.0f;
    }
    const float cos_i = fabsf(cos_i_);
    const float cos_t = sqrtf(ct2);
    const float rs    = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    const float rp    = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    *cos_t_           = (cos_i_ > 0.0f) ? -cos_t : cos_t;
    return 0.5f * (rs * rs + rp * rp);
}

CLPT_D float fresnel_dielectric(float cos_i, float eta)
{
    float ct;
    return fresnel_dielectric_ext(cos_i, &ct, eta);
}

/* reflect / refract: math.h:299-329 */
CLPT_D f3 reflect3(f3 wo, f3 n) { return add3(neg3(wo), sclL3(2.0f * dot3(wo, n), n)); }
CLPT_D f3 refract3(f3 wi, f3 n, float eta, float cos_t)
{
    if (cos_t < 0.0f)
        eta = 1.0f / eta;
    return add3(sclL3(-eta, wi), sclL3(eta * dot3(wi, n) + cos_t, n));
}

/* ------------------------------------------------------------ BSDFs */

CLPT_D f3 conductor_F(const MaterialRec& m, float c)
{
    return mul3(f3{fresnel_conductor1(c, m.eta[0], m.k[0]), fresnel_conductor1(c, m.eta[1], m.k[1]),
                   fresnel_conductor1(c, m.eta[2], m.k[2])},
                f3{m.r[0], m.r[1], m.r[2]});
}

/* BSDF::eval (bsdf.h:48-62): SmoothDiffuse::eval smoothdiffuse.h:31-44,
 * RoughConductor::eval roughconductor.h:39-70, RoughDielectric::eval
 * roughdielectric.h:36-103 (whose names swap wi/wo).  wi = new direction,
 * wo = towards the previous vertex, both local. */
CLPT_D f3 bsdf_eval(const MaterialRec& m, f3 wi, f3 wo)
{
    const f3 zero{0.0f, 0.0f, 0.0f};
    if (m.type == 0u)
    {
        if (wo.z <= 0.0f || wi.z <= 0.0f)
            return zero;
        return scl3(f3{m.r[0], m.r[1], m.r[2]}, CLPT_INV_PI_F);
    }
    if (m.type == 1u)
    {
        if (wo.z <= 0.0f || wi.z <= 0.0f)
            return zero;
        const f3    h = norm3(add3(wo, wi));
        const float D = ggx_D(m.alpha, h);
        if (D == 0.0f)
            return zero;
        const f3    F     = conductor_F(m, dot3(wo, h));
        const float G     = ggx_G(m.alpha, wo, wi, h);
        const float denom = 4.0f * wo.z * wi.z;
        return f3{F.x * D * G / denom, F.y * D * G / denom, F.z * D * G / denom};
    }
    const f3 dwi = wo, dwo = wi;
    if (dwi.z == 0.0f)
        return zero;
    const float ior  = m.ior;
    const bool  refl = dwi.z * dwo.z > 0.0f;
    f3          H;
    if (refl)
        H = norm3(add3(dwo, dwi));
    else
    {
        const float eta = dwi.z > 0.0f ? ior / 1.0f : 1.0f / ior;
        H               = norm3(add3(dwi, scl3(dwo, eta)));
    }
    H             = scl3(H, H.z > 0.0f ? 1.0f : -1.0f);
    const float D = ggx_D(m.alpha, H);
    if (D == 0.0f)
        return zero;
    const float F = fresnel_dielectric(dot3(dwi, H), ior / 1.0f);
    const float G = ggx_G(m.alpha, dwi, dwo, H);
    float       value;
    if (refl)
        value = F * D * G / (4.0f * fabsf(dwi.z) * fabsf(dwo.z));
    else
    {
        const float eta = dwi.z > 0.0f ? ior / 1.0f : 1.0f / ior;
        const float sd  = dot3(dwi, H) + eta * dot3(dwo, H);
        value = fabsf((1.0f - F) * D * G * dot3(dwi, H) * dot3(dwo, H) / (sd * sd * fabsf(dwi.z) * fabsf(dwo.z)));
    }
    return f3{value, value, value};
}

/* BSDF::pdf (bsdf.h:64-78) in the corrected forms (D4 smoothdiffuse.h:56,
 * D5 roughconductor.h:89, D6 roughdielectric.h:149): each equals the density
 * that sample() reports for the same direction. */
CLPT_D float bsdf_pdf(const MaterialRec& m, f3 wi, f3 wo)
{
    if (m.type == 0u)
    {
        if (wo.z <= 0.0f || wi.z <= 0.0f)
            return 0.0f;
        return wi.z * CLPT_INV_PI_F;
    }
    if (m.type == 1u)
    {
        if (wo.z <= 0.0f || wi.z <= 0.0f)
            return 0.0f;
        const f3 h = norm3(add3(wo, wi));
        return ggx_D(m.alpha, h) * h.z / (4.0f * fabsf(dot3(wo, h)));
    }
    const f3    dwi = wo, dwo = wi;
    const float ior  = m.ior;
    const bool  refl = dwi.z * dwo.z > 0.0f;
    f3          H;
    float       dwh;
    if (refl)
    {
        H   = norm3(add3(dwo, dwi));
        dwh = 1.0f / (4.0f * dot3(dwo, H));
    }
    else
    {
        const float eta = dwi.z > 0.0f ? ior / 1.0f : 1.0f / ior;
        H               = norm3(add3(dwi, scl3(dwo, eta)));
        const float sd  = dot3(dwi, H) + eta * dot3(dwo, H);
        dwh             = (eta * eta * dot3(dwo, H)) / (sd * sd);
    }
    H             = scl3(H, H.z < 0.0f ? -1.0f : 1.0f);
    float prob    = ggx_D(m.alpha, H) * H.z;
    const float F = fresnel_dielectric(dot3(dwi, H), ior / 1.0f);
    prob *= refl ? F : (1.0f - F);
    return fabsf(prob * dwh);
}

/* BSDF::sample (bsdf.h:80-96).  Returns f (zero = no direction).  u_lobe is
 * the dielectric's independent reflect/refract draw (D7). */
CLPT_D f3 bsdf_sample(const MaterialRec& m, f3 wo, float ux, float uy, float u_lobe, f3* wi, float* pdf)
{
    const f3 zero{0.0f, 0.0f, 0.0f};
    if (m.type == 0u)
    {
        /* smoothdiffuse.h:59-79 */
        if (wo.z <= 0.0f)
            return zero;
        *wi  = cosine_hemisphere(ux, uy);
        *pdf = wi->z * CLPT_INV_PI_F;
        return scl3(f3{m.r[0], m.r[1], m.r[2]}, CLPT_INV_PI_F);
    }
    if (m.type == 1u)
    {
        /* roughconductor.h:97-137 */
        if (wo.z <= 0.0f)
            return zero;
        float    p;
        const f3 mf = ggx_sample_all(m.alpha, ux, uy, &p);
        if (p == 0.0f)
            return zero;
        const f3 w = reflect3(wo, mf);
        if (w.z <= 0.0f)
            return zero;
        p /= (4.0f * dot3(wo, mf));
        const float D = ggx_D(m.alpha, mf);
        if (D == 0.0f)
            return zero;
        const f3    F     = conductor_F(m, dot3(wo, mf));
        const float G     = ggx_G(m.alpha, wo, w, mf);
        const float denom = 4.0f * wo.z * w.z;
        *wi               = w;
        *pdf              = p;
        return f3{F.x * D * G / denom, F.y * D * G / denom, F.z * D * G / denom};
    }
    /* roughdielectric.h:158-251 */
    const f3    dwi = wo;
    const float ior = m.ior;
    float       mpdf;
    const f3    mf = ggx_sample_all(m.alpha, ux, uy, &mpdf);
    if (mpdf == 0.0f)
        return zero;
    float       p = mpdf;
    float       cos_t;
    const float F    = fresnel_dielectric_ext(dot3(dwi, mf), &cos_t, ior / 1.0f);
    bool        refl = true;
    if (u_lobe > F)
    {
        refl = false;
        p *= 1.0f - F;
    }
    else
        p *= F;
    const float D = ggx_D(m.alpha, mf);
    if (refl)
    {
        const f3 dwo = reflect3(dwi, mf);
        if (dwi.z * dwo.z <= 0.0f)
            return zero;
        const float G   = ggx_G(m.alpha, dwo, dwi, mf);
        const float dwh = fabsf(1.0f / (4.0f * dot3(dwo, mf)));
        p *= dwh;
        const float v = fabsf(F * D * G / (4.0f * dwo.z * dwi.z));
        *wi           = dwo;
        *pdf          = p;
        return f3{v, v, v};
    }
    if (cos_t == 0.0f)
        return zero;
    const f3 dwo = refract3(dwi, mf, ior / 1.0f, cos_t);
    if (dwo.z * dwi.z >= 0.0f)
        return zero;
    const float G   = ggx_G(m.alpha, dwo, dwi, mf);
    const float eta = dwi.z > 0.0f ? ior / 1.0f : 1.0f / ior;
    const float sd  = dot3(dwi, mf) + eta * dot3(dwo, mf);
    const float dwh = fabsf(dot3(dwo, mf) * eta * eta / (sd * sd));
    p *= dwh;
    const float v = fabsf((1.0f - F) * D * G * dot3(dwi, mf) * dot3(dwo, mf) / (sd * sd * fabsf(dwi.z) * fabsf(dwo.z)));
    *wi           = dwo;
    *pdf          = p;
    return f3{v, v, v};
}

/* ------------------------------------------------------------ textures */

/* Exact bilinear RGBA32F fetch, wrap addressing, normalized coordinates:
 * the renderer's ImageTexture2D::eval2D (imagetex2d.h:34-37; sampler state
 * src/core/nodes/textures/imagetex2d.cpp:50-66). */
CLPT_D float4 tex_eval(const float4* tex, uint32_t w, uint32_t h, float u, float v)
{
    const float uu = u - floorf(u);
    const float vv = v - floorf(v);
    const float x  = uu * (float)w - 0.5f;
    const float y  = vv * (float)h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    int         ix0 = (int)fx, iy0 = (int)fy;
    int         ix1 = ix0 + 1, iy1 = iy0 + 1;
    const int   W = (int)w, H = (int)h;
    ix0 = ix0 < 0 ? ix0 + W : (ix0 >= W ? ix0 - W : ix0);
    ix1 = ix1 < 0 ? ix1 + W : (ix1 >= W ? ix1 - W : ix1);
    iy0 = iy0 < 0 ? iy0 + H : (iy0 >= H ? iy0 - H : iy0);
    iy1 = iy1 < 0 ? iy1 + H : (iy1 >= H ? iy1 - H : iy1);
    const float4 t00 = tex[(size_t)iy0 * W + ix0];
    const float4 t10 = tex[(size_t)iy0 * W + ix1];
    const float4 t01 = tex[(size_t)iy1 * W + ix0];
    const float4 t11 = tex[(size_t)iy1 * W + ix1];
    const float  w00 = (1.0f - ax) * (1.0f - ay);
    const float  w10 = ax * (1.0f - ay);
    const float  w01 = (1.0f - ax) * ay;
    const float  w11 = ax * ay;
    return make_float4(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
                       w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
                       w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z,
                       w00 * t00.w + w10 * t10.w + w01 * t01.w + w11 * t11.w);
}

/* ------------------------------------------------------------ emitters */

/* upper_bound: math.h:84-96 */
CLPT_D uint32_t upper_bound_f(const float* a, uint32_t begin, uint32_t end, float value)
{
    while (begin < end)
    {
        const uint32_t mid = begin + (end - begin) / 2;
        if (a[mid] <= value)
            begin = mid + 1;
        else
            end = mid;
    }
    return begin;
}

/* Distribution1D::sample1D: hdridome.h:31-60 */
/* Cut points of a CDF search: cut[k] = upper_bound(cdf[begin, end), k / 256)
 * for k = 0..256.  upper_bound is monotone in the value, so for u in
 * [k/256, (k+1)/256) its result lies in [cut[k], cut[k+1]] and a search of
 * that range returns the same index as a search of the whole CDF. */
CLPT_D uint32_t upper_bound_cut(const float* a, const uint32_t* cut, float value)
{
    const uint32_t k = min((uint32_t)(value * (float)CLPT_CUTS), CLPT_CUTS - 1u); /* value in [0, 1) */
    return upper_bound_f(a, cut[k], cut[k + 1], value);
}

/* cut: the CDF's cut points, or NULL for a search of the whole CDF */
CLPT_D float sample1d(const float* f, const float* cdf, uint32_t N, float c, float u, uint32_t cdf_off, float* pdf,
                      uint32_t* off, const uint32_t* cut)
{
    if (c == 0.0f)
        return 0.0f;
    uint32_t o = cut ? upper_bound_cut(cdf, cut, u) : upper_bound_f(cdf, cdf_off, cdf_off + N + 1, u);
    o          = o - 1u;
    if (o < cdf_off)
        o = cdf_off;
    if (o > cdf_off + N - 1)
        o = cdf_off + N - 1;
    *pdf     = f[o] / c;
    float du = u - cdf[o];
    if (cdf[o + 1] - cdf[o] != 0.0f)
        du /= cdf[o + 1] - cdf[o];
    o -= cdf_off;
    *off = o;
    return ((float)o + du) / (float)N;
}

/* Distribution2D::sample2D: hdridome.h:96-123 */
CLPT_D void sample2d(const SceneView& sc, float ux, float uy, float* su, float* sv, float* pdf)
{
    const uint32_t nu = sc.dome_w, nv = sc.dome_h;
    float          pu = 0.0f, pvv = 0.0f;
    uint32_t       off = 0;
    *sv                = sample1d(sc.pv, sc.cdfv, nv, sc.pv[nv], uy, 0, &pvv, &off, sc.cutv);
    const float cv     = sc.pcond[(size_t)off * (nu + 1) + nu];
    *su = sample1d(sc.pcond, sc.cdfcond, nu, cv, ux, off * (nu + 1), &pu, &off,
                   sc.cutcond ? sc.cutcond + (size_t)off * (CLPT_CUTS + 1u) : nullptr);
    *pdf               = pu * pvv;
}

/* Distribution2D::pdf2D: hdridome.h:125-135 */
CLPT_D float pdf2d(const SceneView& sc, float u, float v)
{
    if (u < 0.0f || v < 0.0f)
        return 0.0f;
    uint32_t iu = (uint32_t)((float)sc.dome_w * u);
    uint32_t iv = (uint32_t)((float)sc.dome_h * v);
    if (iu > sc.dome_w - 1)
        iu = sc.dome_w - 1;
    if (iv > sc.dome_h - 1)
        iv = sc.dome_h - 1;
    return sc.pcond[(size_t)iv * (sc.dome_w + 1) + iu] / sc.pv[sc.dome_h];
}

/* cos(theta) of a direction for acos: clamped to [-1, 1], since a unit
 * vector rotated through a float32 frame can come out with |y| just above 1
 * (the reference passes dir.y unclamped and gets NaN there). */
CLPT_D float dome_cos_theta(f3 dir) { return fminf(1.0f, fmaxf(-1.0f, dir.y)); }

/* HDRIDome::evalEnvironment + directionToUVCoords: hdridome.h:291-297, 472-476 */
CLPT_D f3 dome_le(const SceneView& sc, f3 dir)
{
    const float  u = clpt_atan2f(dir.x, dir.z) * -CLPT_INV_PI_F * 0.5f + 0.5f;
    const float  v = clpt_acosf(dome_cos_theta(dir)) * CLPT_INV_PI_F;
    const float4 c = tex_eval(sc.dome_tex, sc.dome_w, sc.dome_h, u, v);
    return f3{c.x, c.y, c.z};
}

/* HDRIDome::pdfDirect: hdridome.h:248-279 */
CLPT_D float dome_pdf(const SceneView& sc, f3 dir)
{
    const float phi   = clpt_atan2f(dir.x, dir.z);
    const float theta = clpt_acosf(dome_cos_theta(dir));
    float       st, ct;
    clpt_sincosf(theta, &st, &ct);
    if (st == 0.0f)
        return 0.0f;
    const float u = phi * -CLPT_INV_PI_F * 0.5f + 0.5f;
    const float v = theta * CLPT_INV_PI_F;
    return pdf2d(sc, u, v) / (2.0f * CLPT_PI_F * CLPT_PI_F * st);
}

/* HDRIDome::sampleDirect: hdridome.h:186-228 */
CLPT_D f3 dome_sample(const SceneView& sc, float ux, float uy, f3* dir, float* pdf)
{
    float su, sv, p;
    sample2d(sc, ux, uy, &su, &sv, &p);
    const float theta = CLPT_PI_F * sv;
    const float phi   = -2.0f * CLPT_PI_F * (su - 0.5f);
    float       st, ct, sp, cp;
    clpt_sincosf(theta, &st, &ct);
    clpt_sincosf(phi, &sp, &cp);
    if (st == 0.0f)
    {
        *pdf = 0.0f;
        return f3{0.0f, 0.0f, 0.0f};
    }
    *dir           = f3{st * sp, ct, st * cp};
    *pdf           = p / (2.0f * CLPT_PI_F * CLPT_PI_F * st);
    const float4 c = tex_eval(sc.dome_tex, sc.dome_w, sc.dome_h, su, sv);
    return f3{c.x, c.y, c.z};
}

/* DirectionalEmitter::sampleDirect: directional.h:30-56 */
CLPT_D f3 sun_sample(const EmitterRec& e, float ux, float uy, f3* dir, float* pdf)
{
    const f3 rev = neg3(f3{e.dir[0], e.dir[1], e.dir[2]});
    float    dx, dy;
    concentric_disk(ux, uy, &dx, &dy);
    dx *= e.tan_r;
    dy *= e.tan_r;
    f3 ax, ay;
    make_frame(rev, &ax, &ay);
    *dir              = norm3(add3(add3(rev, scl3(ax, dx)), scl3(ay, dy)));
    const float tan_t = sqrtf(dx * dx + dy * dy);
    const float cos_t = 1.0f / sqrtf(1.0f + tan_t * tan_t);
    *pdf              = CLPT_INV_PI_F / (e.tan_r * e.tan_r * cos_t * cos_t * cos_t);
    return f3{e.intensity[0], e.intensity[1], e.intensity[2]};
}

/* LightSampler::sampleEmitterDirect: emitter.h:168-202 -- uniform pick with
 * sample reuse; the value is scaled by n. */
CLPT_D f3 sample_emitters(const SceneView& sc, float ux, float uy, f3* dir, float* pdf, uint32_t* which)
{
    const uint32_t n = sc.nemitters;
    *pdf             = 0.0f;
    if (n == 0)
        return f3{0.0f, 0.0f, 0.0f};
    uint32_t idx = (uint32_t)floorf(ux * (float)n);
    if (idx > n - 1)
        idx = n - 1;
    const float ux2 = ux * (float)n - (float)idx;
    *which          = idx;
    const EmitterRec e = sc.emitters[idx];
    f3               value;
    if (e.type == 1u)
        value = dome_sample(sc, ux2, uy, dir, pdf);
    else
        value = sun_sample(e, ux2, uy, dir, pdf);
    value = scl3(value, (float)n);
    if (*pdf == 0.0f)
        return f3{0.0f, 0.0f, 0.0f};
    return value;
}

/* ------------------------------------------------------------ geometry */

/* A ray set up for traversal.  The lower bound of t is not part of it: it
 * is the same for every ray of a launch and is passed alongside. */
struct RayW
{
    f3       o, d, inv;
    uint32_t k; /* the axes kx | ky << 8 | kz << 16 of the watertight test, and in bits
                   24..26 the signs of inv.x, inv.y, inv.z (1 = negative) */
    float    Sx, Sy, Sz;
};

CLPT_D float comp3(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

/* Watertight ray set-up (Woop, Benthin, Wald, JCGT 2(1), 2013). */
CLPT_D void ray_setup(RayW& r)
{
    const float ax = fabsf(r.d.x), ay = fabsf(r.d.y), az = fabsf(r.d.z);
    int         kz = 0;
    if (ay > ax)
        kz = 1;
    if (az > (kz == 0 ? ax : ay))
        kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (comp3(r.d, kz) < 0.0f)
    {
        const int t = kx;
        kx          = ky;
        ky          = t;
    }
    r.k            = (uint32_t)kx | ((uint32_t)ky << 8) | ((uint32_t)kz << 16);
    /* one correctly rounded division; the shear is a per-ray linear map
     * applied alike to every vertex, so rounding Sx, Sy through Sz keeps the
     * test watertight */
    const float dz = comp3(r.d, kz);
    r.Sz           = 1.0f / dz;
    r.Sx           = comp3(r.d, kx) * r.Sz;
    r.Sy           = comp3(r.d, ky) * r.Sz;
    /* box-slab reciprocals: only for culling, whose bounds are widened by
     * 1e-5 relative, so the hardware reciprocal (1 ulp) serves; the shear
     * constants above decide hits and stay correctly rounded */
    const float tiny = 1e-20f;
    r.inv = f3{__builtin_amdgcn_rcpf(fabsf(r.d.x) > tiny ? r.d.x : (r.d.x < 0.0f ? -tiny : tiny)),
               __builtin_amdgcn_rcpf(fabsf(r.d.y) > tiny ? r.d.y : (r.d.y < 0.0f ? -tiny : tiny)),
               __builtin_amdgcn_rcpf(fabsf(r.d.z) > tiny ? r.d.z : (r.d.z < 0.0f ? -tiny : tiny))};
    /* the near plane of each slab: lo where inv >= 0, hi where it is negative */
    r.k |= ((__float_as_uint(r.inv.x) >> 31) << 24) | ((__float_as_uint(r.inv.y) >> 31) << 25) |
           ((__float_as_uint(r.inv.z) >> 31) << 26);
}

/* Watertight two-sided triangle test; hit iff tmin <= t <= tmax. */
CLPT_D bool tri_test(const RayW& r, float tmin, float tmax, f3 v0, f3 v1, f3 v2, float* t, float* b1, float* b2)
{
    const f3    A = sub3(v0, r.o), B = sub3(v1, r.o), C = sub3(v2, r.o);
    const int   kx = (int)(r.k & 0xffu), ky = (int)((r.k >> 8) & 0xffu), kz = (int)((r.k >> 16) & 0xffu);
    const float Az0 = comp3(A, kz), Bz0 = comp3(B, kz), Cz0 = comp3(C, kz);
    const float Ax = comp3(A, kx) - r.Sx * Az0;
    const float Ay = comp3(A, ky) - r.Sy * Az0;
    const float Bx = comp3(B, kx) - r.Sx * Bz0;
    const float By = comp3(B, ky) - r.Sy * Bz0;
    const float Cx = comp3(C, kx) - r.Sx * Cz0;
    const float Cy = comp3(C, ky) - r.Sy * Cz0;
    float       U  = Cx * By - Cy * Bx;
    float       V  = Ax * Cy - Ay * Cx;
    float       W  = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f)
    {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return false;
    const float det = U + V + W;
    if (det == 0.0f)
        return false;
    const float Az = r.Sz * Az0, Bz = r.Sz * Bz0, Cz = r.Sz * Cz0;
    const float T  = U * Az + V * Bz + W * Cz;
    const float tt = T / det;
    if (!(tt >= tmin && tt <= tmax))
        return false;
    *t  = tt;
    *b1 = V / det;
    *b2 = W / det;
    return true;
}

/* Object-space ray of an instance (affine, no renormalisation), with
 * explicit fused multiply-adds in a fixed order (each correctly rounded,
 * so a host fmaf reproduces them).  The caller runs ray_setup on it. */
CLPT_D void ray_to_object(const float* m, const RayW& w, RayW& o)
{
    o.o = f3{fmaf(m[0], w.o.x, fmaf(m[1], w.o.y, fmaf(m[2], w.o.z, m[3]))),
             fmaf(m[4], w.o.x, fmaf(m[5], w.o.y, fmaf(m[6], w.o.z, m[7]))),
             fmaf(m[8], w.o.x, fmaf(m[9], w.o.y, fmaf(m[10], w.o.z, m[11])))};
    o.d = f3{fmaf(m[0], w.d.x, fmaf(m[1], w.d.y, m[2] * w.d.z)), fmaf(m[4], w.d.x, fmaf(m[5], w.d.y, m[6] * w.d.z)),
             fmaf(m[8], w.d.x, fmaf(m[9], w.d.y, m[10] * w.d.z))};
}

/* Total order of hits: t, then instance, then primitive. */
CLPT_D bool hit_better(float t, uint32_t inst, uint32_t prim, float bt, uint32_t binst, uint32_t bprim)
{
    if (t != bt)
        return t < bt;
    if (inst != binst)
        return inst < binst;
    return prim < bprim;
}

#endif
