/*
 * clpt_guide.h -- the learned directional proposal (device side).
 *
 * A query is 9 numbers: the hit position normalised to the scene box
 * ([0,1]^3), the shading normal and the outgoing direction (world space).
 * The encoding maps them to 256 features: for each of the 9 inputs and each
 * frequency f = 0..11 the pair sin(2^f pi v), cos(2^f pi v) (216 values),
 * then the 9 raw inputs, a constant 1 (the first layer's bias) and zeros.
 * The network is 256 -> 64 -> 64 -> 64 -> 64 -> 64 with ReLU after every
 * hidden layer and no biases (32 768 weights); 40 of the 64 outputs are
 * used: for lobe j, out[5j] is the mixture logit, out[5j+1..3] the
 * unnormalised mean direction and out[5j+4] the log concentration.  The
 * proposal is the mixture of 8 von Mises-Fisher lobes
 *     p(w) = sum_j pi_j kappa_j / (2 pi (1 - e^{-2 kappa_j})) e^{kappa_j (mu_j.w - 1)}
 * with pi = softmax(logits), mu_j = normalised means,
 * kappa_j = exp(clamp(s_j, -4, 8)).
 *
 * Weights and activations are bfloat16 with float32 accumulation on the
 * matrix cores (v_mfma_f32_32x32x16_bf16).  Lane layout of that
 * instruction (checked on gfx950): lane l holds A[l%32][8(l/32)+s] and
 * B[8(l/32)+s][l%32] for s = 0..7, and accumulator register r holds
 * D[8(r/4) + 4(l/32) + r%4][l%32].
 *
 * The fused forward pass computes the transposed product H^T = W^T X^T so
 * that every column (one query) stays in one lane pair: the accumulator of
 * one layer is already the B operand of the next once each hidden index k
 * is read as k with bits 2 and 3 exchanged.  The packed weight image
 * stores every layer as W^T rows (output unit) with that exchange applied
 * to the input index of layers 1..4, rows padded by 8 elements against LDS
 * bank conflicts.
 */
#ifndef CLPT_GUIDE_H
#define CLPT_GUIDE_H

#include "clpt_device.h"

#define CLPT_HD __host__ __device__

#define CLPT_G_DIMS 9
#define CLPT_G_FREQ 12
#define CLPT_G_IN 256
#define CLPT_G_HID 64
#define CLPT_G_NOUT 40
#define CLPT_G_LOBES 8
#define CLPT_G_LAYERS 5
#define CLPT_G_PARAMS (CLPT_G_IN * CLPT_G_HID + 4 * CLPT_G_HID * CLPT_G_HID) /* 32768 */
#define CLPT_G_S0 (CLPT_G_IN + 8)  /* padded row of layer 0 in the packed image */
#define CLPT_G_S1 (CLPT_G_HID + 8) /* padded row of layers 1..4 */
#define CLPT_G_IMG (CLPT_G_HID * CLPT_G_S0 + 4 * CLPT_G_HID * CLPT_G_S1) /* 35328 bf16 */
#define CLPT_G_KMIN -4.0f
#define CLPT_G_KMAX 8.0f

/* Offset of layer l in the float32 parameter vector; layer l is stored
 * row-major as W_l[in][out]. */
CLPT_HD inline uint32_t guide_layer_offset(int l) { return l == 0 ? 0u : (uint32_t)(CLPT_G_IN * CLPT_G_HID + (l - 1) * CLPT_G_HID * CLPT_G_HID); }
CLPT_HD inline uint32_t guide_layer_in(int l) { return l == 0 ? CLPT_G_IN : CLPT_G_HID; }
/* Offset of layer l in the packed bf16 image. */
CLPT_HD inline uint32_t guide_img_offset(int l) { return l == 0 ? 0u : (uint32_t)(CLPT_G_HID * CLPT_G_S0 + (l - 1) * CLPT_G_HID * CLPT_G_S1); }
/* Hidden index held by register slot k of the chained layout: bits 2, 3 exchanged. */
CLPT_HD inline uint32_t guide_perm(uint32_t k) { return (k & ~12u) | ((k & 4u) << 1) | ((k & 8u) >> 1); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float  f32x16 __attribute__((ext_vector_type(16)));
typedef float  f32x8 __attribute__((ext_vector_type(8)));

/* ------------------------------------------------------------ encoding */

/* The 8 features of group g (features 8g..8g+7).  Groups 0..26: input
 * d = g/3, frequencies 4(g%3)..4(g%3)+3 as (sin, cos) pairs; 27: raw
 * inputs 0..7; 28: raw input 8, 1, zeros; 29..31: zeros.  The sin/cos pair
 * of the lowest frequency of a group comes from the reduced argument
 * (exact in float32: scaling by 2^k and subtracting an even integer), the
 * next three by angle doubling. */
CLPT_D void guide_sincos_block(float v, float scale, float o[8])
{
    const float x = v * scale;                    /* exact */
    const float r = x - 2.0f * rintf(0.5f * x);   /* exact, in [-1, 1] */
    float       s, c;
    clpt_sincosf(CLPT_PI_F * r, &s, &c);
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        o[2 * k]     = s;
        o[2 * k + 1] = c;
        const float s2 = 2.0f * s * c;
        const float c2 = (c - s) * (c + s);
        s = s2;
        c = c2;
    }
}

template <int G>
CLPT_D void guide_group(const float v[9], float o[8])
{
    if constexpr (G < 27)
        guide_sincos_block(v[G / 3], (float)(1 << (4 * (G % 3))), o);
    else if constexpr (G == 27)
    {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            o[k] = v[k];
    }
    else if constexpr (G == 28)
    {
        o[0] = v[8];
        o[1] = 1.0f;
#pragma unroll
        for (int k = 2; k < 8; ++k)
            o[k] = 0.0f;
    }
    else
    {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            o[k] = 0.0f;
    }
}

CLPT_D bf16x8 guide_pack(const float o[8])
{
    /* a vector conversion: four packed v_cvt_pk_bf16_f32 */
    f32x8 v;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        v[k] = o[k];
    return __builtin_convertvector(v, bf16x8);
}

/* Group 2t+h of query v, h = lane half (0 or 1): one argument reduction per
 * lane where both candidate groups are sin/cos blocks. */
template <int T>
CLPT_D void guide_group_sel(const float v[9], bool h, float o[8])
{
    constexpr int G0 = 2 * T, G1 = 2 * T + 1;
    if constexpr (G1 < 27)
    {
        const float x  = h ? v[G1 / 3] : v[G0 / 3];
        const float sc = h ? (float)(1 << (4 * (G1 % 3))) : (float)(1 << (4 * (G0 % 3)));
        guide_sincos_block(x, sc, o);
    }
    else
    {
        float a[8], b[8];
        guide_group<G0>(v, a);
        guide_group<G1>(v, b);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            o[k] = h ? b[k] : a[k];
    }
}

/* ------------------------------------------------------------ vMF mixture */

struct VmfMix
{
    float w[CLPT_G_LOBES];
    f3    mu[CLPT_G_LOBES];
    float kappa[CLPT_G_LOBES];
    float norm[CLPT_G_LOBES]; /* pi_j kappa_j / (2 pi (1 - e^{-2 kappa_j})) */
};

CLPT_D void vmf_mixture(const float out[CLPT_G_NOUT], VmfMix& m)
{
    float amax = out[0];
#pragma unroll
    for (int j = 1; j < CLPT_G_LOBES; ++j)
        amax = fmaxf(amax, out[5 * j]);
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < CLPT_G_LOBES; ++j)
    {
        m.w[j] = clpt_expf(out[5 * j] - amax);
        sum += m.w[j];
    }
#pragma unroll
    for (int j = 0; j < CLPT_G_LOBES; ++j)
    {
        m.w[j] /= sum;
        const f3    r  = f3{out[5 * j + 1], out[5 * j + 2], out[5 * j + 3]};
        const float r2 = dot3(r, r);
        m.mu[j]        = r2 > 1e-12f ? scl3(r, 1.0f / sqrtf(r2)) : f3{0.0f, 0.0f, 1.0f};
        const float k  = clpt_expf(fminf(fmaxf(out[5 * j + 4], CLPT_G_KMIN), CLPT_G_KMAX));
        m.kappa[j]     = k;
        m.norm[j]      = m.w[j] * k / ((2.0f * CLPT_PI_F) * (1.0f - clpt_expf(-2.0f * k)));
    }
}

CLPT_D float vmf_mixture_pdf(const VmfMix& m, f3 w)
{
    float p = 0.0f;
#pragma unroll
    for (int j = 0; j < CLPT_G_LOBES; ++j)
        p += m.norm[j] * clpt_expf(m.kappa[j] * (dot3(m.mu[j], w) - 1.0f));
    return p;
}

/* Lobe by usel over the weights, then the lobe's direction by inversion of
 * its cosine distribution (Jakob 2012, "Numerically stable sampling of the
 * von Mises Fisher distribution on S^2"). */
CLPT_D f3 vmf_mixture_sample(const VmfMix& m, float usel, float u1, float u2)
{
    /* one pass of conditional moves (no indexed access into the lobes) */
    f3    mu    = m.mu[CLPT_G_LOBES - 1];
    float k     = m.kappa[CLPT_G_LOBES - 1];
    float acc   = 0.0f;
    bool  found = false;
#pragma unroll
    for (int i = 0; i < CLPT_G_LOBES - 1; ++i)
    {
        acc += m.w[i];
        const bool take = !found && usel < acc;
        mu.x            = take ? m.mu[i].x : mu.x;
        mu.y            = take ? m.mu[i].y : mu.y;
        mu.z            = take ? m.mu[i].z : mu.z;
        k               = take ? m.kappa[i] : k;
        found           = found || take;
    }
    const float e  = clpt_expf(-2.0f * k);
    float       ct = 1.0f + clpt_logf(u1 + (1.0f - u1) * e) / k;
    ct             = fminf(1.0f, fmaxf(-1.0f, ct));
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float       sp, cp;
    clpt_sincosf((2.0f * CLPT_PI_F) * u2, &sp, &cp);
    f3 t, b;
    make_frame(mu, &t, &b);
    return norm3(add3(add3(sclL3(st * cp, t), sclL3(st * sp, b)), sclL3(ct, mu)));
}

/* ------------------------------------------------------------ fused forward */

/* FP8 (OCP e4m3) inference, the precision ablation: weights of layer l are
 * stored as e4m3(W * ws_l), the input of layer l enters as e4m3(x * as_l)
 * (both saturated at +-448), and the float32 accumulator is scaled back by
 * 1 / (as_l ws_l).  scl[2l] = as_l, scl[2l+1] = 1 / (as_l ws_l).  The
 * instruction is v_mfma_f32_32x32x16_fp8_fp8, lane layout as for bf16 with
 * 8 bytes per lane. */
#define CLPT_G_FP8_MAX 448.0f

CLPT_D long guide_pack_fp8(const float o[8], float scale)
{
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        x[k] = fminf(fmaxf(o[k] * scale, -CLPT_G_FP8_MAX), CLPT_G_FP8_MAX);
    int lo = __builtin_amdgcn_cvt_pk_fp8_f32(x[0], x[1], 0, false);
    lo     = __builtin_amdgcn_cvt_pk_fp8_f32(x[2], x[3], lo, true);
    int hi = __builtin_amdgcn_cvt_pk_fp8_f32(x[4], x[5], 0, false);
    hi     = __builtin_amdgcn_cvt_pk_fp8_f32(x[6], x[7], hi, true);
    return (long)(uint32_t)lo | ((long)(uint32_t)hi << 32);
}

template <bool FP8>
struct GuideFrag
{
    typedef bf16x8 type;
};
template <>
struct GuideFrag<true>
{
    typedef long type;
};

template <bool FP8>
CLPT_D typename GuideFrag<FP8>::type guide_load_a(const void* W, uint32_t off)
{
    if constexpr (FP8)
        return *(const long*)((const uint8_t*)W + off);
    else
        return *(const bf16x8*)((const __bf16*)W + off);
}

template <bool FP8>
CLPT_D typename GuideFrag<FP8>::type guide_pack_b(const float o[8], float scale)
{
    if constexpr (FP8)
        return guide_pack_fp8(o, scale);
    else
        return guide_pack(o);
}

template <bool FP8>
CLPT_D f32x16 guide_mfma(typename GuideFrag<FP8>::type a, typename GuideFrag<FP8>::type b, f32x16 c)
{
    if constexpr (FP8)
        return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a, b, c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

CLPT_D f32x16 guide_zero16()
{
    f32x16 z;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        z[k] = 0.0f;
    return z;
}

/* Layer L >= 1: input = relu(previous accumulator) (times `in_scale`, the
 * previous layer's 1/(as ws) folded with this layer's as for FP8). */
template <bool FP8, int L>
CLPT_D void guide_hidden_layer(const void* W, f32x16 acc[2][2], uint32_t c, uint32_t h, float in_scale)
{
    typename GuideFrag<FP8>::type b[2][4];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int t = 0; t < 4; ++t)
        {
            float o[8];
#pragma unroll
            for (int s = 0; s < 8; ++s)
                o[s] = fmaxf(acc[t / 2][ct][8 * (t % 2) + s], 0.0f);
            b[ct][t] = guide_pack_b<FP8>(o, in_scale);
        }
    const uint32_t base = guide_img_offset(L);
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
            acc[rt][ct] = guide_zero16();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
        {
            const auto a = guide_load_a<FP8>(W, base + (32 * rt + c) * CLPT_G_S1 + 16 * t + 8 * h);
            acc[rt][0]   = guide_mfma<FP8>(a, b[0][t], acc[rt][0]);
            acc[rt][1]   = guide_mfma<FP8>(a, b[1][t], acc[rt][1]);
        }
}

/* Forward pass for the 64 queries of a wave (one per lane; every lane of
 * the wave must call it, with any query for idle lanes).  W is the packed
 * image (LDS or global): bf16 elements, or e4m3 bytes with the scales scl
 * (FP8).  On return out[] holds this lane's 40 outputs. */
template <bool FP8>
CLPT_D void guide_forward_wave_t(const void* W, const float* scl, const float q[9], float out[CLPT_G_NOUT])
{
    const uint32_t lane = __lane_id();
    const uint32_t c = lane & 31u, h = lane >> 5;
    float          qa[9], qb[9];
#pragma unroll
    for (int d = 0; d < 9; ++d)
    {
        qa[d] = __shfl(q[d], (int)c);
        qb[d] = __shfl(q[d], (int)(c + 32));
    }
    float sc[10];
#pragma unroll
    for (int k = 0; k < 10; ++k)
        sc[k] = FP8 ? scl[k] : 1.0f;
    f32x16 acc[2][2];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
            acc[rt][ct] = guide_zero16();
    const bool hb = h != 0;
#define CLPT_G_L0_STEP(T)                                                                                          \
    {                                                                                                              \
        float o0[8], o1[8];                                                                                        \
        guide_group_sel<T>(qa, hb, o0);                                                                            \
        guide_group_sel<T>(qb, hb, o1);                                                                            \
        const auto b0 = guide_pack_b<FP8>(o0, sc[0]), b1 = guide_pack_b<FP8>(o1, sc[0]);                           \
        _Pragma("unroll") for (int rt = 0; rt < 2; ++rt)                                                           \
        {                                                                                                          \
            const auto a = guide_load_a<FP8>(W, (32 * rt + c) * CLPT_G_S0 + 16 * (T) + 8 * h);                    \
            acc[rt][0]   = guide_mfma<FP8>(a, b0, acc[rt][0]);                                                     \
            acc[rt][1]   = guide_mfma<FP8>(a, b1, acc[rt][1]);                                                     \
        }                                                                                                          \
    }
    CLPT_G_L0_STEP(0) CLPT_G_L0_STEP(1) CLPT_G_L0_STEP(2) CLPT_G_L0_STEP(3) CLPT_G_L0_STEP(4) CLPT_G_L0_STEP(5)
    CLPT_G_L0_STEP(6) CLPT_G_L0_STEP(7) CLPT_G_L0_STEP(8) CLPT_G_L0_STEP(9) CLPT_G_L0_STEP(10) CLPT_G_L0_STEP(11)
    CLPT_G_L0_STEP(12) CLPT_G_L0_STEP(13) CLPT_G_L0_STEP(14) CLPT_G_L0_STEP(15)
#undef CLPT_G_L0_STEP
    guide_hidden_layer<FP8, 1>(W, acc, c, h, sc[1] * sc[2]);
    guide_hidden_layer<FP8, 2>(W, acc, c, h, sc[3] * sc[4]);
    guide_hidden_layer<FP8, 3>(W, acc, c, h, sc[5] * sc[6]);
    guide_hidden_layer<FP8, 4>(W, acc, c, h, sc[7] * sc[8]);
    /* lane l takes column tile h (query l): its own rows of that tile and
     * the partner lane's (l ^ 32) rows of the same tile */
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int r = 0; r < 16; ++r)
        {
            const int row0 = 32 * rt + 8 * (r / 4) + r % 4; /* half 0; half 1 is row0 + 4 */
            if (row0 >= CLPT_G_NOUT)
                continue;
            const float own  = (hb ? acc[rt][1][r] : acc[rt][0][r]) * sc[9];
            const float send = (hb ? acc[rt][0][r] : acc[rt][1][r]) * sc[9];
            const float recv = __shfl_xor(send, 32);
            out[row0]        = hb ? recv : own;
            if (row0 + 4 < CLPT_G_NOUT)
                out[row0 + 4] = hb ? own : recv;
        }
}

CLPT_D void guide_forward_wave(const __bf16* W, const float q[9], float out[CLPT_G_NOUT])
{
    guide_forward_wave_t<false>(W, nullptr, q, out);
}

/* The query of a surface point: position in the scene box, normal, wo. */
struct GuideBox
{
    float lo[3], inv_ext[3];
};

CLPT_D void guide_query(const GuideBox& b, f3 p, f3 n, f3 wo, float q[9])
{
    q[0] = (p.x - b.lo[0]) * b.inv_ext[0];
    q[1] = (p.y - b.lo[1]) * b.inv_ext[1];
    q[2] = (p.z - b.lo[2]) * b.inv_ext[2];
    q[3] = n.x;
    q[4] = n.y;
    q[5] = n.z;
    q[6] = wo.x;
    q[7] = wo.y;
    q[8] = wo.z;
}

#endif
