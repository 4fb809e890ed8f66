/*
 * guide.hip -- kernels of the learned directional proposal: batched
 * inference (the fused forward pass of clpt_guide.h), online training
 * (encoding, matrix-core GEMMs, mixture loss, Adam) and the gather that
 * turns the tracer's recorded path vertices into training samples.
 *
 * Training runs layer by layer over a batch of N samples kept in HBM:
 *   forward   A_{l+1} = relu(A_l W_l)           (GEMM "nt", N x 64 tiles)
 *   loss      G_5 = d(-sum_n w_n log p(wi_n)) / d out, scaled by 1 / sum w
 *   backward  dW_l = A_l^T G_{l+1}              (GEMM "tn", split over N)
 *             G_l  = (G_{l+1} W_l^T) . [A_l > 0]
 *   Adam      on the float32 master weights, then the bfloat16 copies.
 * Activations and gradients are stored as bfloat16, products accumulate in
 * float32 on v_mfma_f32_32x32x16_bf16; split-N partial weight gradients
 * are summed in a fixed order, so a step is deterministic.
 */
#include "clpt_guide.h"
#include "clpt_kernels.h"

namespace
{

__device__ __forceinline__ float lum3(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

/* Samples of training step slice [off, off + N) of the gathered buffer. */
/* Samples of Adam step `step` of `nsteps` on the gathered samples: rows
 * step, step + nsteps, step + 2 nsteps, ... of the first min(count,
 * N nsteps), so every step draws from all of the batch's paths. */
__device__ __forceinline__ uint32_t guide_count(const GuideCounters* gc, uint32_t step, uint32_t nsteps, uint32_t N)
{
    const uint32_t c = min(gc->count, N * nsteps);
    return c > step ? (c - step + nsteps - 1u) / nsteps : 0u;
}

/* Elementary functions of the training loss: the hardware base-2
 * exponential and logarithm, reciprocal and reciprocal square root (about
 * one ulp).  The loss and its gradient need no bit-exact agreement with the
 * CPU restatement, unlike the mixture the tracer samples. */
__device__ __forceinline__ float gl_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float gl_log(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

/* The terms of one vMF lobe for direction wi, from its 5 outputs o5
 * (logit, unnormalised mean, log concentration) and the logit maximum m:
 * a = e^{logit - m}, the normalised mean, kappa, e = e^{-2 kappa}, mu.wi,
 * 1/|raw mean| (0 for a vanishing mean, whose direction is +z) and
 * q = kappa / (2 pi (1 - e)) e^{kappa (mu.wi - 1)}, the lobe's density. */
struct GuideLobe
{
    float a, q, k, e, cmu, inv, mx, my, mz;
};

__device__ __forceinline__ void guide_lobe(const float* o5, float m, f3 wi, GuideLobe& L)
{
    L.a            = gl_exp(o5[0] - m);
    const float r2 = o5[1] * o5[1] + o5[2] * o5[2] + o5[3] * o5[3];
    const bool  ok = r2 > 1e-12f;
    L.inv          = ok ? __builtin_amdgcn_rsqf(r2) : 0.0f;
    L.mx           = ok ? o5[1] * L.inv : 0.0f;
    L.my           = ok ? o5[2] * L.inv : 0.0f;
    L.mz           = ok ? o5[3] * L.inv : 1.0f;
    L.k            = gl_exp(fminf(fmaxf(o5[4], CLPT_G_KMIN), CLPT_G_KMAX));
    L.e            = gl_exp(-2.0f * L.k);
    L.cmu          = L.mx * wi.x + L.my * wi.y + L.mz * wi.z;
    L.q            = L.k * __builtin_amdgcn_rcpf((2.0f * CLPT_PI_F) * (1.0f - L.e)) * gl_exp(L.k * (L.cmu - 1.0f));
}

/* Gradient of -wn log p w.r.t. the lobe's 5 outputs, for its mixture
 * weight pi and responsibility r = pi q / p. */
__device__ __forceinline__ void guide_lobe_grad(const float* o5, const GuideLobe& L, float pi, float r, float wn, f3 wi,
                                                float* g5)
{
    g5[0]          = -wn * (r - pi);
    const float s1 = -wn * r * L.k * L.inv;
    g5[1]          = s1 * (wi.x - L.cmu * L.mx);
    g5[2]          = s1 * (wi.y - L.cmu * L.my);
    g5[3]          = s1 * (wi.z - L.cmu * L.mz);
    const float dl = __builtin_amdgcn_rcpf(L.k) - 2.0f * L.e * __builtin_amdgcn_rcpf(1.0f - L.e) + (L.cmu - 1.0f);
    g5[4]          = o5[4] > CLPT_G_KMIN && o5[4] < CLPT_G_KMAX ? -wn * r * dl * L.k : 0.0f;
}

/* The weighted negative log-likelihood -wn log p(wi) of the vMF mixture
 * with network outputs o, and its gradient w.r.t. the 40 outputs (g). */
__device__ __forceinline__ float guide_nll_grad(const float o[CLPT_G_NOUT], f3 wi, float wn, float g[CLPT_G_NOUT])
{
    float m = o[0];
#pragma unroll
    for (int j = 1; j < CLPT_G_LOBES; ++j)
        m = fmaxf(m, o[5 * j]);
    GuideLobe L[CLPT_G_LOBES];
    float     S = 0.0f;
#pragma unroll
    for (int j = 0; j < CLPT_G_LOBES; ++j)
    {
        guide_lobe(o + 5 * j, m, wi, L[j]);
        S += L[j].a;
    }
    const float iS = __builtin_amdgcn_rcpf(S);
    float       p  = 0.0f;
#pragma unroll
    for (int j = 0; j < CLPT_G_LOBES; ++j)
        p += (L[j].a * iS) * L[j].q;
    const float ip = p > 0.0f ? __builtin_amdgcn_rcpf(p) : 0.0f;
#pragma unroll
    for (int j = 0; j < CLPT_G_LOBES; ++j)
    {
        const float pi = L[j].a * iS;
        guide_lobe_grad(o + 5 * j, L[j], pi, pi * L[j].q * ip, wn, wi, g + 5 * j);
    }
    return -wn * gl_log(fmaxf(p, 1e-30f));
}

} // namespace

/* ------------------------------------------------------------ inference */

template <bool FP8>
__global__ __launch_bounds__(256) void k_guide_infer(const void* img, const float* scl, const float* q9, uint32_t n,
                                                     float* out40)
{
    __shared__ uint4 w[CLPT_G_IMG * (FP8 ? 1 : 2) / 16];
    __shared__ float s[10];
    {
        const uint4* src = reinterpret_cast<const uint4*>(img);
        for (uint32_t k = threadIdx.x; k < CLPT_G_IMG * (FP8 ? 1 : 2) / 16; k += blockDim.x)
            w[k] = src[k];
        if (threadIdx.x < 10)
            s[threadIdx.x] = FP8 ? scl[threadIdx.x] : 1.0f;
        __syncthreads();
    }
    const uint32_t lane   = __lane_id();
    const uint32_t wave   = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t base = wave * 64u; base < n; base += nwaves * 64u)
    {
        const uint32_t i = base + lane;
        float          q[9];
#pragma unroll
        for (int d = 0; d < 9; ++d)
            q[d] = i < n ? q9[9 * (size_t)i + d] : 0.0f;
        float out[CLPT_G_NOUT];
        guide_forward_wave_t<FP8>(w, s, q, out);
        if (i < n)
#pragma unroll
            for (int k = 0; k < CLPT_G_NOUT; ++k)
                out40[CLPT_G_NOUT * (size_t)i + k] = out[k];
    }
}

/* The mixture of the learned proposal on given network outputs (parity
 * surface of the functions k_shade_guided uses): density at dir3, a sample
 * from u3 = (lobe, u1, u2) and the density at that sample. */
__global__ __launch_bounds__(256) void k_guide_mixture_batch(const float* out40, const float* dir3, const float* u3,
                                                             uint32_t n, float* pdf, float* smp3, float* smp_pdf)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        float o[CLPT_G_NOUT];
#pragma unroll
        for (int k = 0; k < CLPT_G_NOUT; ++k)
            o[k] = out40[CLPT_G_NOUT * (size_t)i + k];
        VmfMix m;
        vmf_mixture(o, m);
        pdf[i]      = vmf_mixture_pdf(m, f3{dir3[3 * (size_t)i], dir3[3 * (size_t)i + 1], dir3[3 * (size_t)i + 2]});
        const f3 w  = vmf_mixture_sample(m, u3[3 * (size_t)i], u3[3 * (size_t)i + 1], u3[3 * (size_t)i + 2]);
        smp3[3 * (size_t)i]     = w.x;
        smp3[3 * (size_t)i + 1] = w.y;
        smp3[3 * (size_t)i + 2] = w.z;
        smp_pdf[i]              = vmf_mixture_pdf(m, w);
    }
}

/* guide_query for given surface points (position, shading normal, wo). */
__global__ __launch_bounds__(256) void k_guide_query_batch(GuideBox box, const float* p3, const float* n3,
                                                           const float* wo3, uint32_t n, float* q9)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        float q[9];
        guide_query(box, ld3(p3 + 3 * (size_t)i), ld3(n3 + 3 * (size_t)i), ld3(wo3 + 3 * (size_t)i), q);
#pragma unroll
        for (int k = 0; k < 9; ++k)
            q9[9 * (size_t)i + k] = q[k];
    }
}

/* ------------------------------------------------------------ training */

/* Also the step's bookkeeping: per-block sums of the sample weights (fixed
 * tree order) into wpart, the sample count, and the cleared activation
 * maxima.  N is a multiple of 256 (one thread per sample row). */
__global__ __launch_bounds__(256) void k_guide_features(const GuideSample* s, GuideCounters* gc, uint32_t step,
                                                        uint32_t nsteps, uint32_t N, __bf16* X, double* wpart)
{
    __shared__ double wred[256];
    const uint32_t n   = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t cnt = guide_count(gc, step, nsteps, N);
    float          v[9];
    double         wt = 0.0;
    if (n < cnt)
    {
        const GuideSample g = s[(size_t)n * nsteps];
        v[0] = g.a.x; v[1] = g.a.y; v[2] = g.a.z; v[3] = g.a.w;
        v[4] = g.b.x; v[5] = g.b.y; v[6] = g.b.z; v[7] = g.b.w;
        v[8] = g.c.x;
        wt   = (double)g.d.x;
    }
    wred[threadIdx.x] = wt;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1)
    {
        if (threadIdx.x < st)
            wred[threadIdx.x] += wred[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        wpart[blockIdx.x] = wred[0];
        if (blockIdx.x == 0)
            gc->trained = cnt;
    }
    bf16x8* row = reinterpret_cast<bf16x8*>(X + (size_t)n * CLPT_G_IN);
    float   o[8];
#define CLPT_G_FEAT(G)                                                                                             \
    {                                                                                                              \
        if (n < cnt)                                                                                               \
            guide_group<G>(v, o);                                                                                  \
        else                                                                                                       \
            for (int k = 0; k < 8; ++k)                                                                            \
                o[k] = 0.0f;                                                                                       \
        row[G] = guide_pack(o);                                                                                    \
    }
    CLPT_G_FEAT(0) CLPT_G_FEAT(1) CLPT_G_FEAT(2) CLPT_G_FEAT(3) CLPT_G_FEAT(4) CLPT_G_FEAT(5) CLPT_G_FEAT(6)
    CLPT_G_FEAT(7) CLPT_G_FEAT(8) CLPT_G_FEAT(9) CLPT_G_FEAT(10) CLPT_G_FEAT(11) CLPT_G_FEAT(12) CLPT_G_FEAT(13)
    CLPT_G_FEAT(14) CLPT_G_FEAT(15) CLPT_G_FEAT(16) CLPT_G_FEAT(17) CLPT_G_FEAT(18) CLPT_G_FEAT(19)
    CLPT_G_FEAT(20) CLPT_G_FEAT(21) CLPT_G_FEAT(22) CLPT_G_FEAT(23) CLPT_G_FEAT(24) CLPT_G_FEAT(25)
    CLPT_G_FEAT(26) CLPT_G_FEAT(27) CLPT_G_FEAT(28) CLPT_G_FEAT(29) CLPT_G_FEAT(30) CLPT_G_FEAT(31)
#undef CLPT_G_FEAT
}

enum
{
    GEMM_RELU_BF16 = 0, /* C = bf16(max(y, 0)) */
    GEMM_F32       = 1, /* C = y */
    GEMM_MASK_BF16 = 2  /* C = bf16(mask > 0 ? y : 0) */
};

/* Y[n][j] = sum_k A[n][k] Bt[j][k]; A is N x K, Bt is J x K (both row-major
 * bfloat16), one 32 x 32 tile of Y per wave. */
template <int EPI>
__global__ __launch_bounds__(256) void k_guide_gemm_nt(const __bf16* __restrict__ A, const __bf16* __restrict__ Bt,
                                                       uint32_t N, uint32_t K, uint32_t J, void* C,
                                                       const __bf16* __restrict__ mask, float* amax_part)
{
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t jt   = J / 32u;
    const uint32_t tn = wave / jt, tj = wave % jt;
    /* the launch covers exactly N / 32 x J / 32 tiles (N a multiple of 1024,
     * J = 64), so every wave of a block reaches the block-wide max below */
    const uint32_t lane = __lane_id(), c = lane & 31u, h = lane >> 5;
    const uint32_t n0 = 32u * tn, j0 = 32u * tj;
    f32x16         acc = guide_zero16();
    const __bf16*  pa  = A + (size_t)(n0 + c) * K + 8u * h;
    const __bf16*  pb  = Bt + (size_t)(j0 + c) * K + 8u * h;
    /* K is 64 or 256: fragments of 4 k-steps are loaded before their
     * products, keeping 8 loads in flight per wave */
    for (uint32_t k0 = 0; k0 < K; k0 += 64u)
    {
        bf16x8 a[4], b[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
        {
            a[t] = *reinterpret_cast<const bf16x8*>(pa + k0 + 16u * t);
            b[t] = *reinterpret_cast<const bf16x8*>(pb + k0 + 16u * t);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[t], b[t], acc, 0, 0, 0);
    }
    float mx = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r)
    {
        const size_t idx = (size_t)(n0 + 8u * (r / 4) + 4u * h + (r % 4)) * J + j0 + c;
        const float  y   = acc[r];
        mx               = fmaxf(mx, y);
        if constexpr (EPI == GEMM_RELU_BF16)
            static_cast<__bf16*>(C)[idx] = (__bf16)fmaxf(y, 0.0f);
        else if constexpr (EPI == GEMM_F32)
            static_cast<float*>(C)[idx] = y;
        else
            static_cast<__bf16*>(C)[idx] = (float)mask[idx] > 0.0f ? (__bf16)y : (__bf16)0.0f;
    }
    if constexpr (EPI == GEMM_RELU_BF16)
    {
        /* the block's activation maximum (for the FP8 scales), stored per
         * block and reduced once per step by k_guide_loss */
        __shared__ float bmx[4];
        for (int off = 32; off > 0; off >>= 1)
            mx = fmaxf(mx, __shfl_xor(mx, off));
        if (lane == 0)
            bmx[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0)
            amax_part[blockIdx.x] = fmaxf(fmaxf(bmx[0], bmx[1]), fmaxf(bmx[2], bmx[3]));
    }
}

/* Partial weight gradient over one chunk of 1024 samples:
 * P[chunk][k][m] = sum_{n in chunk} A[n][k] G[n][m]; A is N x K, G is N x 64.
 * One block per (32-row k tile, 32-column m tile, chunk); each of its four
 * waves takes 256 of the chunk's samples, 64 per iteration, with the next
 * iteration's loads in flight during the current products.  The 64 x 32
 * tiles of A and G go through LDS (16-byte row writes, column reads) to
 * reach the MFMA operand layout, where a lane holds 8 consecutive samples.
 * The four waves' sums are added in a fixed order, so the result does not
 * depend on timing.  Written in the parameter layout (row k, column m) at
 * param_off. */
__global__ __launch_bounds__(256) void k_guide_gemm_tn(const __bf16* __restrict__ A, const __bf16* __restrict__ G,
                                                       uint32_t K, float* partial, uint32_t param_off)
{
    const uint32_t kt = K / 32u;
    const uint32_t tk = blockIdx.x % kt, rem = blockIdx.x / kt;
    const uint32_t tm = rem % 2u, ch = rem / 2u;
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id(), c = lane & 31u, h = lane >> 5;
    const uint32_t k0 = 32u * tk, m0 = 32u * tm;
    const uint32_t lr = lane >> 2, lc = 8u * (lane & 3u);
    __shared__ __align__(16) __bf16 ta[4][64][40], tg[4][64][40];
    const __bf16* pa = A + (size_t)(1024u * ch + 256u * w + lr) * K + k0 + lc;
    const __bf16* pg = G + (size_t)(1024u * ch + 256u * w + lr) * CLPT_G_HID + m0 + lc;
    bf16x8        va[4], vg[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
    {
        va[t] = *reinterpret_cast<const bf16x8*>(pa + (size_t)(16 * t) * K);
        vg[t] = *reinterpret_cast<const bf16x8*>(pg + (size_t)(16 * t) * CLPT_G_HID);
    }
    f32x16 acc = guide_zero16();
    for (uint32_t it = 0; it < 4u; ++it)
    {
#pragma unroll
        for (int t = 0; t < 4; ++t)
        {
            *reinterpret_cast<bf16x8*>(&ta[w][16 * t + lr][lc]) = va[t];
            *reinterpret_cast<bf16x8*>(&tg[w][16 * t + lr][lc]) = vg[t];
        }
        if (it + 1u < 4u)
        {
            const size_t n1 = (size_t)64u * (it + 1u);
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                va[t] = *reinterpret_cast<const bf16x8*>(pa + (n1 + 16u * t) * K);
                vg[t] = *reinterpret_cast<const bf16x8*>(pg + (n1 + 16u * t) * CLPT_G_HID);
            }
        }
        /* the LDS writes have landed (lgkmcnt 0; vector loads stay in flight) */
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 4; ++t)
        {
            bf16x8 a, b;
#pragma unroll
            for (int q = 0; q < 8; ++q)
            {
                a[q] = ta[w][16 * t + 8 * h + q][c];
                b[q] = tg[w][16 * t + 8 * h + q][c];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __builtin_amdgcn_wave_barrier();
    }
    /* the four waves' tiles, added in wave order */
    __syncthreads();
    float* red = reinterpret_cast<float*>(&ta[0][0][0]); /* 4 x 16 x 64 floats = 16 KB of the 20 KB */
#pragma unroll
    for (int r = 0; r < 16; ++r)
        red[(w * 16u + r) * 64u + lane] = acc[r];
    __syncthreads();
    float* P = partial + (size_t)ch * CLPT_G_PARAMS + param_off;
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        const uint32_t e = threadIdx.x + 256u * q, r = e / 64u, l = e % 64u;
        const float    v = ((red[r * 64u + l] + red[(16u + r) * 64u + l]) + red[(32u + r) * 64u + l]) +
                        red[(48u + r) * 64u + l];
        P[(size_t)(k0 + 8u * (r / 4u) + 4u * (l / 32u) + (r % 4u)) * CLPT_G_HID + m0 + (l % 32u)] = v;
    }
}

/* Gradient of -w/W log p(wi) w.r.t. the 40 outputs of each sample (the
 * vMF-mixture log-likelihood weighted by the sample's importance weight w
 * and normalised by the batch weight sum W); rows beyond the sample count
 * and output columns 40..63 are zero.  Per-block loss sums to loss_part. */
__global__ __launch_bounds__(256) void k_guide_loss(const float* __restrict__ out, const GuideSample* s,
                                                    GuideCounters* gc, uint32_t step, uint32_t nsteps, uint32_t N,
                                                    __bf16* G5, float* loss_part, const double* wpart,
                                                    const float* amax_part)
{
    __shared__ float  red[256];
    __shared__ double wred[256];
    const uint32_t    n   = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t    cnt = guide_count(gc, step, nsteps, N);
    /* the hidden activations' maxima of this step; a step without samples
     * keeps the previous ones (its activations are those of zero inputs) */
    if (blockIdx.x == 0 && cnt > 0u)
        for (int l = 0; l < 4; ++l)
        {
            float m = 0.0f;
            for (uint32_t k = threadIdx.x; k < N / 64u; k += blockDim.x)
                m = fmaxf(m, amax_part[(size_t)l * (N / 64u) + k]);
            red[threadIdx.x] = m;
            __syncthreads();
            for (uint32_t st = 128; st > 0; st >>= 1)
            {
                if (threadIdx.x < st)
                    red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + st]);
                __syncthreads();
            }
            if (threadIdx.x == 0)
                gc->amax[l] = __float_as_uint(red[0]);
            __syncthreads();
        }
    /* W: the per-block weight sums in a fixed order (the same in every block) */
    double wacc = 0.0;
    for (uint32_t k = threadIdx.x; k < N / 256u; k += blockDim.x)
        wacc += wpart[k];
    wred[threadIdx.x] = wacc;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1)
    {
        if (threadIdx.x < st)
            wred[threadIdx.x] += wred[threadIdx.x + st];
        __syncthreads();
    }
    const float wsum = (float)wred[0];
    if (n == 0)
        gc->wsum = wsum;
    float g[CLPT_G_HID];
#pragma unroll
    for (int k = CLPT_G_NOUT; k < CLPT_G_HID; ++k)
        g[k] = 0.0f;
    float loss = 0.0f;
    if (n < cnt && wsum > 0.0f)
    {
        const GuideSample smp = s[(size_t)n * nsteps];
        float             o[CLPT_G_NOUT];
#pragma unroll
        for (int k = 0; k < CLPT_G_NOUT; ++k)
            o[k] = out[(size_t)n * CLPT_G_HID + k];
        loss = guide_nll_grad(o, f3{smp.c.y, smp.c.z, smp.c.w}, smp.d.x / wsum, g);
    }
    else
#pragma unroll
        for (int k = 0; k < CLPT_G_NOUT; ++k)
            g[k] = 0.0f;
    if (n < N)
    {
        bf16x8* row = reinterpret_cast<bf16x8*>(G5 + (size_t)n * CLPT_G_HID);
#pragma unroll
        for (int q = 0; q < CLPT_G_HID / 8; ++q)
            row[q] = guide_pack(g + 8 * q);
    }
    red[threadIdx.x] = loss;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1)
    {
        if (threadIdx.x < st)
            red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        loss_part[blockIdx.x] = red[0];
}

/* bfloat16 copies of parameter p: W (row-major [in][out]), WT ([out][in])
 * and the packed inference image. */
__device__ __forceinline__ void guide_store_param(uint32_t p, float th, __bf16* W, __bf16* WT, __bf16* img)
{
    const uint32_t l   = p < (uint32_t)(CLPT_G_IN * CLPT_G_HID) ? 0u : 1u + (p - CLPT_G_IN * CLPT_G_HID) / (CLPT_G_HID * CLPT_G_HID);
    const uint32_t off = guide_layer_offset((int)l), in = guide_layer_in((int)l);
    const uint32_t k = (p - off) / CLPT_G_HID, m = (p - off) % CLPT_G_HID;
    const __bf16   b = (__bf16)th;
    W[p]                  = b;
    WT[off + m * in + k]  = b;
    if (l == 0)
        img[m * CLPT_G_S0 + k] = b;
    else
        img[guide_img_offset((int)l) + m * CLPT_G_S1 + guide_perm(k)] = b;
}

__global__ void k_guide_pack(const float* params, __bf16* W, __bf16* WT, __bf16* img)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < CLPT_G_PARAMS)
        guide_store_param(p, params[p], W, WT, img);
}

/* e4m3 image of the weights, one block per layer: ws_l = 448 / max|W_l|,
 * as_0 = 448 (|features| <= 1), as_l = 448 / (1.25 max activation of the
 * last training step) for the hidden inputs (2.0 before any training). */
__global__ __launch_bounds__(256) void k_guide_pack_fp8(const float* params, const GuideCounters* gc, uint8_t* img8,
                                                        float* scl)
{
    __shared__ float red[256];
    const int      l   = blockIdx.x;
    const uint32_t off = guide_layer_offset(l), in = guide_layer_in(l), cnt = in * CLPT_G_HID;
    float          mx  = 0.0f;
    for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x)
        mx = fmaxf(mx, fabsf(params[off + k]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1)
    {
        if (threadIdx.x < st)
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    const float ws = red[0] > 0.0f ? CLPT_G_FP8_MAX / red[0] : 1.0f;
    float       as = CLPT_G_FP8_MAX;
    if (l > 0)
    {
        const float am = __uint_as_float(gc->amax[l - 1]);
        as             = CLPT_G_FP8_MAX / (am > 0.0f ? 1.25f * am : 2.0f);
    }
    for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x)
    {
        const uint32_t i = k / CLPT_G_HID, m = k % CLPT_G_HID;
        const float    x = fminf(fmaxf(params[off + k] * ws, -CLPT_G_FP8_MAX), CLPT_G_FP8_MAX);
        const uint8_t  b = (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false) & 0xff);
        if (l == 0)
            img8[m * CLPT_G_S0 + i] = b;
        else
            img8[guide_img_offset(l) + m * CLPT_G_S1 + guide_perm(i)] = b;
    }
    if (threadIdx.x == 0)
    {
        scl[2 * l]     = as;
        scl[2 * l + 1] = 1.0f / (as * ws);
    }
}

/* Adam on the float32 master weights (gradient = fixed-order sum of the
 * split-N partials), then the bfloat16 copies.  No update when the batch
 * held no samples. */
__global__ void k_guide_adam(float* params, float* m1, float* m2, const float* partial, uint32_t nchunks,
                             GuideCounters* gc, float lr, float beta1, float beta2, float eps, __bf16* W, __bf16* WT,
                             __bf16* img, float* grad_out)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= CLPT_G_PARAMS || gc->trained == 0u)
        return;
    float g = 0.0f;
    for (uint32_t c = 0; c < nchunks; ++c)
        g += partial[(size_t)c * CLPT_G_PARAMS + p];
    if (grad_out)
        grad_out[p] = g;
    const double t  = (double)(gc->steps + 1u);
    const float  a  = beta1 * m1[p] + (1.0f - beta1) * g;
    const float  b  = beta2 * m2[p] + (1.0f - beta2) * g * g;
    m1[p]           = a;
    m2[p]           = b;
    const float  ah = (float)((double)a / (1.0 - pow((double)beta1, t)));
    const float  bh = (float)((double)b / (1.0 - pow((double)beta2, t)));
    const float  th = params[p] - lr * ah / (sqrtf(bh) + eps);
    params[p]       = th;
    guide_store_param(p, th, W, WT, img);
}

/* Close a step: loss of the batch, step counter; the last step of a batch
 * empties the sample buffer. */
__global__ __launch_bounds__(256) void k_guide_finish(GuideCounters* gc, const float* loss_part, uint32_t nparts,
                                                      int last)
{
    __shared__ float red[256];
    float            acc = 0.0f;
    for (uint32_t k = threadIdx.x; k < nparts; k += blockDim.x)
        acc += loss_part[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1)
    {
        if (threadIdx.x < st)
            red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        if (gc->trained)
        {
            gc->loss = red[0];
            gc->steps += 1u;
            gc->samples_total += gc->trained;
        }
        if (last)
            gc->count = 0u;
    }
}

/* ------------------------------------------------------------ fused training step */

/* One Adam step in two launches instead of one per GEMM.
 *
 * k_guide_train_fused: CLPT_GT_WAVES waves per block, 32 samples per wave,
 * CLPT_GT_S samples per pass; a block makes `iters` passes over consecutive
 * samples.  The forward image (W_l^T rows), the backward image (W_l rows,
 * layers 1..4, output index in the chained order) and a staging area stay in
 * LDS.  Per pass and wave, all in the chained register layout of the fused
 * forward pass (a sample per lane pair, units in registers):
 *   encode -> 5 forward GEMMs (activations kept as bf16 operand fragments)
 *   -> mixture loss and its gradient G_4 (the pair exchanges its halves of
 *   the 40 outputs) -> G_{l-1} = (W_l G_l) . [A_{l-1} > 0] for l = 4..1.
 * Weight gradients dW_l = A_{l-1} G_l^T contract over samples, so the
 * operands are staged in LDS as [unit][sample] (all waves' samples), and
 * wave w accumulates one 32 x 32 tile of every layer (of every 64-feature
 * chunk of layer 0), adding each pass to the block's gradient.  Sample weights
 * are normalised by the block's weight sum W_b; the block writes its
 * gradient, W_b, its loss and its activation maxima.
 *
 * k_guide_adam_fused: the gradient sum_b (W_b / W) P_b in block order, Adam,
 * the bfloat16 copies; its block 0 closes the step (loss, maxima, weight
 * sum).  The fused kernel advances the step counters.  Both kernels are
 * deterministic and use no atomics. */
#define CLPT_GT_WAVES 4
#define CLPT_GT_S (32 * CLPT_GT_WAVES)
#define CLPT_GT_SS (CLPT_GT_S + 8)
#define CLPT_GT_SCAL 5

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

/* One 32 x 32 tile of a weight gradient over the pass's samples: rows
 * 32 fr.. of staging array A times rows 32 orr.. of staging array B
 * (both [unit][sample]), added to the block's gradient.  The block's
 * gradient is reached through a buffer resource (the 32-bit buffer
 * operations move the float's bits): `lofs` is the lane's byte offset in
 * the tile, the uniform parts (block, layer, tile corner) are `sofs`, so
 * the addresses cost no registers per element.  guide_dw_load fetches the
 * tile's value after the earlier passes (zero on the first) ahead of the
 * barrier; guide_dw_tile adds the pass and stores, so a tile is summed over
 * the passes in pass order. */
__device__ __forceinline__ f32x16 guide_dw_load(__amdgpu_buffer_rsrc_t res, uint32_t lofs, uint32_t sofs, bool add)
{
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r)
        acc[r] = add ? __uint_as_float(
                           __builtin_amdgcn_raw_buffer_load_b32(res, lofs + (8u * (r / 4) + r % 4) * 256u, sofs, 0))
                     : 0.0f;
    return acc;
}

__device__ __forceinline__ void guide_dw_tile(const __bf16 (*A)[CLPT_GT_SS], const __bf16 (*B)[CLPT_GT_SS], uint32_t fr,
                                              uint32_t orr, uint32_t c, uint32_t h, __amdgpu_buffer_rsrc_t res,
                                              uint32_t lofs, uint32_t sofs, f32x16 acc)
{
#pragma unroll
    for (int kt = 0; kt < CLPT_GT_S / 16; ++kt)
    {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&A[32u * fr + c][16 * kt + 8u * h]);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(&B[32u * orr + c][16 * kt + 8u * h]);
        acc            = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(acc[r]), res, lofs + (8u * (r / 4) + r % 4) * 256u,
                                              sofs, 0);
}

/* Operand fragment of encoding group 2T + h. */
template <int T>
__device__ __forceinline__ bf16x8 guide_feature_frag(const float v[9], bool hb)
{
    float o[8];
    guide_group_sel<T>(v, hb, o);
    return guide_pack(o);
}

__global__ __launch_bounds__(256) void k_guide_train_fused(const GuideSample* __restrict__ s, GuideCounters* gc,
                                                           uint32_t step, uint32_t nsteps, uint32_t N, uint32_t iters,
                                                           const __bf16* __restrict__ img,
                                                           const __bf16* __restrict__ Wrm, float* __restrict__ partial,
                                                           double* __restrict__ wpart, float* __restrict__ spart)
{
    __shared__ __align__(16) __bf16 lw[CLPT_G_IMG];
    __shared__ __align__(16) __bf16 lb[4 * CLPT_G_HID * CLPT_G_S1];
    __shared__ __align__(16) __bf16 st[2][CLPT_G_HID][CLPT_GT_SS];
    __shared__ double               dred[CLPT_GT_WAVES];
    __shared__ float                fred[CLPT_GT_WAVES][CLPT_GT_SCAL];

    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = __lane_id(), c = lane & 31u, h = lane >> 5;
    const bool     hb  = h != 0u;
    const uint32_t cnt = guide_count(gc, step, nsteps, N);
    /* the step's samples (cnt of the N slots) spread over the blocks: the
     * fewest passes per block that cover them, at most the launch's `iters` */
    const uint32_t per   = CLPT_GT_S * gridDim.x;
    iters                = min(iters, max(1u, (cnt + per - 1u) / per));
    const uint32_t first = blockIdx.x * iters * CLPT_GT_S;
    const uint32_t stop  = min(first + iters * CLPT_GT_S, cnt);
    /* the step's counters (no block of this kernel reads them; the Adam
     * kernel that follows takes its bias correction from the new count) */
    if (blockIdx.x == 0u && tid == 0u)
    {
        gc->trained = cnt;
        if (cnt > 0u)
        {
            gc->steps += 1u;
            gc->samples_total += cnt;
        }
    }
    /* a block without samples contributes nothing (W_b = 0, which the
     * reduction skips) */
    if (first >= cnt)
    {
        if (tid < CLPT_GT_SCAL)
            spart[(size_t)blockIdx.x * CLPT_GT_SCAL + tid] = 0.0f;
        if (tid == 0u)
            wpart[blockIdx.x] = 0.0;
        return;
    }

    {
        /* every load of a thread issued before its LDS writes */
        constexpr uint32_t NV = CLPT_G_IMG / 8u, NW = NV / 256u; /* 17 full rounds and a partial one */
        const uint4*       src = reinterpret_cast<const uint4*>(img);
        uint4*             dst = reinterpret_cast<uint4*>(lw);
        uint4              wv[NW];
#pragma unroll
        for (uint32_t q = 0; q < NW; ++q)
            wv[q] = src[tid + 256u * q];
        const bool  tail = tid + 256u * NW < NV;
        const uint4 wt   = tail ? src[tid + 256u * NW] : uint4{};
#pragma unroll
        for (uint32_t q = 0; q < NW; ++q)
            dst[tid + 256u * q] = wv[q];
        if (tail)
            dst[tid + 256u * NW] = wt;
        /* lb[l-1][i][16t + 8u + s] = W_l[i][perm(16t + 8u + s)]: the two runs
         * of 4 consecutive outputs 16t + 4u and 16t + 8 + 4u */
        bf16x4 lo[8], hi[8];
#pragma unroll
        for (uint32_t q = 0; q < 8u; ++q)
        {
            const uint32_t k = tid + 256u * q, l = k / 512u, i = (k / 8u) % 64u, t = (k % 8u) >> 1, u = k & 1u;
            const __bf16*  row = Wrm + guide_layer_offset((int)l + 1) + i * CLPT_G_HID;
            lo[q]              = *reinterpret_cast<const bf16x4*>(row + 16u * t + 4u * u);
            hi[q]              = *reinterpret_cast<const bf16x4*>(row + 16u * t + 8u + 4u * u);
        }
#pragma unroll
        for (uint32_t q = 0; q < 8u; ++q)
        {
            const uint32_t k = tid + 256u * q, l = k / 512u, i = (k / 8u) % 64u, t = (k % 8u) >> 1, u = k & 1u;
            bf16x8         v;
            v[0] = lo[q][0]; v[1] = lo[q][1]; v[2] = lo[q][2]; v[3] = lo[q][3];
            v[4] = hi[q][0]; v[5] = hi[q][1]; v[6] = hi[q][2]; v[7] = hi[q][3];
            *reinterpret_cast<bf16x8*>(lb + (l * CLPT_G_HID + i) * CLPT_G_S1 + 16u * t + 8u * u) = v;
        }
    }
    /* W_b: the block's sample weights, summed in a fixed order */
    double wl = 0.0;
    for (uint32_t k = first + tid; k < stop; k += blockDim.x)
        wl += (double)s[(size_t)k * nsteps].d.x;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        wl += __shfl_xor(wl, off);
    if (lane == 0u)
        dred[w] = wl;
    __syncthreads();
    const double wb = ((dred[0] + dred[1]) + dred[2]) + dred[3];

    /* this wave's tile (input rows 32 fr.., output columns 32 orr..) of
     * every layer's gradient and of every 64-feature chunk of layer 0 */
    const uint32_t fr = w >> 1, orr = w & 1u;
    const __amdgpu_buffer_rsrc_t pres = __builtin_amdgcn_make_buffer_rsrc(
        partial + (size_t)blockIdx.x * CLPT_G_PARAMS, (short)0, CLPT_G_PARAMS * 4, 0x00020000);
    const uint32_t lofs  = ((32u * fr + 4u * h) * CLPT_G_HID + 32u * orr + c) * 4u;
    float          lsum = 0.0f, amx[4] = {0.0f, 0.0f, 0.0f, 0.0f};

    for (uint32_t it = 0; it < iters; ++it)
    {
        const uint32_t k    = first + it * CLPT_GT_S + 32u * w + c;
        const bool     live = k < cnt;
        float          v[9];
        f3             wi{0.0f, 0.0f, 1.0f};
        float          wn = 0.0f;
#pragma unroll
        for (int d = 0; d < 9; ++d)
            v[d] = 0.0f;
        if (live)
        {
            const GuideSample g = s[(size_t)k * nsteps];
            v[0] = g.a.x; v[1] = g.a.y; v[2] = g.a.z; v[3] = g.a.w;
            v[4] = g.b.x; v[5] = g.b.y; v[6] = g.b.z; v[7] = g.b.w;
            v[8] = g.c.x;
            wi   = f3{g.c.y, g.c.z, g.c.w};
            wn   = wb > 0.0 ? (float)((double)g.d.x / wb) : 0.0f;
        }
        /* forward, layer 0: features of group 2t + h, kept for the weight
         * gradient.  An idle sample (zero query) has zero loss gradient, so
         * it adds nothing to any weight gradient; it is left out of the
         * activation maxima. */
        bf16x8 xf[16];
        f32x16 acc[2] = {guide_zero16(), guide_zero16()};
#define CLPT_GT_L0(T)                                                                                              \
    {                                                                                                              \
        xf[T] = guide_feature_frag<T>(v, hb);                                                                \
        _Pragma("unroll") for (int rt = 0; rt < 2; ++rt)                                                           \
        {                                                                                                          \
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(lw + (32u * rt + c) * CLPT_G_S0 + 16u * (T) + 8u * h); \
            acc[rt]        = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xf[T], acc[rt], 0, 0, 0);                  \
        }                                                                                                          \
    }
        CLPT_GT_L0(0) CLPT_GT_L0(1) CLPT_GT_L0(2) CLPT_GT_L0(3) CLPT_GT_L0(4) CLPT_GT_L0(5) CLPT_GT_L0(6)
        CLPT_GT_L0(7) CLPT_GT_L0(8) CLPT_GT_L0(9) CLPT_GT_L0(10) CLPT_GT_L0(11) CLPT_GT_L0(12) CLPT_GT_L0(13)
        CLPT_GT_L0(14) CLPT_GT_L0(15)
#undef CLPT_GT_L0
        /* hidden layers; af[l] = relu(Z_l) as the operand fragments of layer l+1 */
        bf16x8 af[4][4];
#pragma unroll
        for (int l = 1; l <= 4; ++l)
        {
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                float o[8];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                {
                    o[q]       = fmaxf(acc[t / 2][8 * (t % 2) + q], 0.0f);
                    amx[l - 1] = fmaxf(amx[l - 1], live ? o[q] : 0.0f);
                }
                af[l - 1][t] = guide_pack(o);
            }
            acc[0] = guide_zero16();
            acc[1] = guide_zero16();
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int rt = 0; rt < 2; ++rt)
                {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(
                        lw + guide_img_offset(l) + (32u * rt + c) * CLPT_G_S1 + 16u * t + 8u * h);
                    acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, af[l - 1][t], acc[rt], 0, 0, 0);
                }
        }
        /* The loss, its 8 lobes split over the lane pair: half h takes lobes
         * 4h..4h+3, outputs 20h + t.  Units t with bit 2 clear are in this
         * lane's registers for both halves; the others come from the partner,
         * which holds them at the register of the other half's unit. */
        float ol[20];
#pragma unroll
        for (int t = 0; t < 20; ++t)
        {
            const float v0 = acc[t / 32][4 * ((t % 32) / 8) + t % 4];
            const float v1 = acc[(20 + t) / 32][4 * (((20 + t) % 32) / 8) + (20 + t) % 4];
            if (((t >> 2) & 1) == 0)
                ol[t] = hb ? v1 : v0;
            else
                ol[t] = __shfl_xor(hb ? v0 : v1, 32);
        }
        GuideLobe L[4];
        float     m = fmaxf(fmaxf(ol[0], ol[5]), fmaxf(ol[10], ol[15]));
        m           = fmaxf(m, __shfl_xor(m, 32));
#pragma unroll
        for (int j = 0; j < 4; ++j)
            guide_lobe(ol + 5 * j, m, wi, L[j]);
        /* sums over all 8 lobes, lobes 0..3 first in both halves */
        const float sa = ((L[0].a + L[1].a) + L[2].a) + L[3].a, so = __shfl_xor(sa, 32);
        const float iS = __builtin_amdgcn_rcpf(hb ? so + sa : sa + so);
        float       pa = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            pa += (L[j].a * iS) * L[j].q;
        const float po = __shfl_xor(pa, 32);
        const float p  = hb ? po + pa : pa + po;
        const float ip = p > 0.0f ? __builtin_amdgcn_rcpf(p) : 0.0f;
        float       gl[20];
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const float pi = L[j].a * iS;
            guide_lobe_grad(ol + 5 * j, L[j], pi, pi * L[j].q * ip, wn, wi, gl + 5 * j);
        }
        lsum += hb ? 0.0f : -wn * gl_log(fmaxf(p, 1e-30f));
        /* each half sends the gradients of units 4..7, 12..15 of its range */
        float gr[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            gr[q] = __shfl_xor(gl[q < 4 ? 4 + q : 8 + q], 32);
        /* G_4 fragments: element q of fragment t is unit 16t + 8(q/4) + 4h + q%4 */
        bf16x8 gf[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 8; ++q)
            {
                const int u0 = 16 * t + 8 * (q / 4) + q % 4, u1 = u0 + 4;
                const float g0 = u0 < 20 ? gl[u0 < 20 ? u0 : 0]
                               : (u0 >= 24 && u0 <= 27) ? gr[(u0 >= 24 && u0 <= 27) ? u0 - 24 : 0]
                               : (u0 >= 32 && u0 <= 35) ? gr[(u0 >= 32 && u0 <= 35) ? u0 - 28 : 0] : 0.0f;
                const float g1 = (u1 >= 20 && u1 < 40) ? gl[(u1 >= 20 && u1 < 40) ? u1 - 20 : 0]
                               : (u1 >= 4 && u1 <= 7)  ? gr[(u1 >= 4 && u1 <= 7) ? u1 - 4 : 0]
                               : (u1 >= 12 && u1 <= 15) ? gr[(u1 >= 12 && u1 <= 15) ? u1 - 8 : 0] : 0.0f;
                gf[t][q] = (__bf16)(hb ? g1 : g0);
            }
        /* backward through layers 4..1 */
#pragma unroll
        for (int l = 4; l >= 1; --l)
        {
            const f32x16 prev = guide_dw_load(pres, lofs, guide_layer_offset(l) * 4u, it > 0u);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int q = 0; q < 8; ++q)
                {
                    const uint32_t u = 16u * t + 8u * (q / 4) + 4u * h + q % 4;
                    st[0][u][32u * w + c] = af[l - 1][t][q];
                    st[1][u][32u * w + c] = gf[t][q];
                }
            f32x16 bacc[2] = {guide_zero16(), guide_zero16()};
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int rt = 0; rt < 2; ++rt)
                {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(
                        lb + ((l - 1) * CLPT_G_HID + 32u * rt + c) * CLPT_G_S1 + 16u * t + 8u * h);
                    bacc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, gf[t], bacc[rt], 0, 0, 0);
                }
            __syncthreads();
            guide_dw_tile(st[0], st[1], fr, orr, c, h, pres, lofs, guide_layer_offset(l) * 4u, prev);
            __syncthreads();
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                float o[8];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    o[q] = (float)af[l - 1][t][q] > 0.0f ? bacc[t / 2][8 * (t % 2) + q] : 0.0f;
                gf[t] = guide_pack(o);
            }
        }
        /* layer 0: G_0 once, the features chunk by chunk (64 features) */
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 8; ++q)
                st[1][16u * t + 8u * (q / 4) + 4u * h + q % 4][32u * w + c] = gf[t][q];
#pragma unroll
        for (int ch = 0; ch < 4; ++ch)
        {
            const f32x16 prev = guide_dw_load(pres, lofs, 64u * ch * CLPT_G_HID * 4u, it > 0u);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    st[0][16u * t + 8u * h + q][32u * w + c] = xf[4 * ch + t][q];
            __syncthreads();
            guide_dw_tile(st[0], st[1], fr, orr, c, h, pres, lofs, 64u * ch * CLPT_G_HID * 4u, prev);
            __syncthreads();
        }
    }

    /* the block's loss and maxima, in wave order */
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        lsum += __shfl_xor(lsum, off);
#pragma unroll
        for (int l = 0; l < 4; ++l)
            amx[l] = fmaxf(amx[l], __shfl_xor(amx[l], off));
    }
    if (lane == 0u)
    {
        fred[w][0] = lsum;
#pragma unroll
        for (int l = 0; l < 4; ++l)
            fred[w][1 + l] = amx[l];
    }
    __syncthreads();
    if (tid < CLPT_GT_SCAL)
    {
        float a = fred[0][tid];
#pragma unroll
        for (int q = 1; q < CLPT_GT_WAVES; ++q)
            a = tid == 0u ? a + fred[q][tid] : fmaxf(a, fred[q][tid]);
        spart[(size_t)blockIdx.x * CLPT_GT_SCAL + tid] = a;
        if (tid == 0u)
            wpart[blockIdx.x] = wb;
    }
}

/* The reduction over the blocks and Adam: CLPT_GA_P parameters per block,
 * each summed over the fused step's blocks by CLPT_GA_G groups of threads
 * (group q takes blocks q, q + CLPT_GA_G, ...), the groups then added in
 * order, so that many loads are in flight and the sum stays in a fixed
 * order. */
#define CLPT_GT_MAXBLK CLPT_GUIDE_FUSED_MAXBLK
#define CLPT_GA_P 32
#define CLPT_GA_G 8
__global__ __launch_bounds__(256) void k_guide_adam_fused(float* params, float* m1, float* m2,
                                                          const float* __restrict__ partial,
                                                          const double* __restrict__ wpart,
                                                          const float* __restrict__ spart, uint32_t nblk,
                                                          GuideCounters* gc, uint32_t step, uint32_t nsteps,
                                                          float lr, float beta1, float beta2, float eps,
                                                          __bf16* W, __bf16* WT, __bf16* img, float* grad_out)
{
    __shared__ double dred[256];
    __shared__ float  scale[CLPT_GT_MAXBLK];
    __shared__ float  gsum[CLPT_GA_G][CLPT_GA_P];
    __shared__ float  fred[256];
    const uint32_t    tid = threadIdx.x;
    double            a   = 0.0;
    for (uint32_t b = tid; b < nblk; b += blockDim.x)
        a += wpart[b];
    dred[tid] = a;
    __syncthreads();
    for (uint32_t k = 128; k > 0; k >>= 1)
    {
        if (tid < k)
            dred[tid] += dred[tid + k];
        __syncthreads();
    }
    const double wsum = dred[0];
    for (uint32_t b = tid; b < nblk; b += blockDim.x)
        scale[b] = wsum > 0.0 ? (float)(wpart[b] / wsum) : 0.0f;
    __syncthreads();
    const uint32_t cnt = gc->trained;
    const uint32_t j = tid % CLPT_GA_P, grp = tid / CLPT_GA_P;
    const uint32_t p = blockIdx.x * CLPT_GA_P + j;
    if (cnt > 0u)
    {
        const float* src = partial + p;
        float        g   = 0.0f;
        uint32_t     b   = grp;
        /* blocks without samples (W_b = 0) wrote no gradient: skipped */
        for (; b + 3u * CLPT_GA_G < nblk && scale[b + 3u * CLPT_GA_G] != 0.0f; b += 4u * CLPT_GA_G)
        {
            const float x0 = src[(size_t)b * CLPT_G_PARAMS];
            const float x1 = src[(size_t)(b + CLPT_GA_G) * CLPT_G_PARAMS];
            const float x2 = src[(size_t)(b + 2u * CLPT_GA_G) * CLPT_G_PARAMS];
            const float x3 = src[(size_t)(b + 3u * CLPT_GA_G) * CLPT_G_PARAMS];
            g              = scale[b] != 0.0f ? fmaf(scale[b], x0, g) : g;
            g              = scale[b + CLPT_GA_G] != 0.0f ? fmaf(scale[b + CLPT_GA_G], x1, g) : g;
            g              = scale[b + 2u * CLPT_GA_G] != 0.0f ? fmaf(scale[b + 2u * CLPT_GA_G], x2, g) : g;
            g              = fmaf(scale[b + 3u * CLPT_GA_G], x3, g);
        }
        for (; b < nblk; b += CLPT_GA_G)
            if (scale[b] != 0.0f)
                g = fmaf(scale[b], src[(size_t)b * CLPT_G_PARAMS], g);
        gsum[grp][j] = g;
    }
    __syncthreads();
    if (cnt > 0u && grp == 0u)
    {
        float g = gsum[0][j];
#pragma unroll
        for (int q = 1; q < CLPT_GA_G; ++q)
            g += gsum[q][j];
        if (grad_out)
            grad_out[p] = g;
        const double t  = (double)gc->steps; /* already counts this step */
        const float  u  = beta1 * m1[p] + (1.0f - beta1) * g;
        const float  v  = beta2 * m2[p] + (1.0f - beta2) * g * g;
        m1[p]           = u;
        m2[p]           = v;
        const float  uh = (float)((double)u / (1.0 - pow((double)beta1, t)));
        const float  vh = (float)((double)v / (1.0 - pow((double)beta2, t)));
        const float  th = params[p] - lr * uh / (sqrtf(vh) + eps);
        params[p]       = th;
        guide_store_param(p, th, W, WT, img);
    }
    /* block 0 closes the step: the loss (block losses weighted by W_b / W),
     * the activation maxima, the weight sum; the last step of a batch empties
     * the sample buffer (no block of this kernel reads the count) */
    if (blockIdx.x != 0u)
        return;
    for (int q = 0; q < CLPT_GT_SCAL; ++q)
    {
        float x = 0.0f;
        for (uint32_t b = tid; b < nblk; b += blockDim.x)
            x = q == 0 ? fmaf(scale[b], spart[(size_t)b * CLPT_GT_SCAL], x)
                       : fmaxf(x, spart[(size_t)b * CLPT_GT_SCAL + q]);
        fred[tid] = x;
        __syncthreads();
        for (uint32_t k = 128; k > 0; k >>= 1)
        {
            if (tid < k)
                fred[tid] = q == 0 ? fred[tid] + fred[tid + k] : fmaxf(fred[tid], fred[tid + k]);
            __syncthreads();
        }
        if (tid == 0u && cnt > 0u)
        {
            if (q == 0)
                gc->loss = fred[0];
            else
                gc->amax[q - 1] = __float_as_uint(fred[0]);
        }
        __syncthreads();
    }
    if (tid == 0u)
    {
        gc->wsum = (float)wsum;
        if (step + 1u == nsteps)
            gc->count = 0u;
    }
}

/* ------------------------------------------------------------ gather */

/* One thread per recorded path: walk its vertices from the deepest up,
 * keeping the sum S of the contributions that arrived after each vertex.
 * Vertex k yields the sample (query, sampled direction, weight) with weight
 * lum(S * R) / p, R = f cos / T_next per channel and p the density the
 * direction was drawn with: the single-sample estimate of
 * lum(L_i f cos) / p of the radiance that came back along it.
 *
 * The samples are appended in a fixed order -- by path slot, then from the
 * deepest vertex up -- so that training sees the same data in the same
 * order on every run: a counting pass, a scan of the per-block counts and a
 * writing pass that walks the paths again. */
template <class F>
__device__ __forceinline__ void guide_walk(const float4* V, const float4* E, uint32_t nslots, uint32_t slot,
                                           uint32_t maxd, F&& emit)
{
    const size_t pl = (size_t)(CLPT_MAXD + 1) * nslots;
    float        sr = 0.0f, sg = 0.0f, sb = 0.0f;
    for (uint32_t j = maxd; j >= 1u; --j)
    {
        const size_t rec = (size_t)j * nslots + slot;
        const float4 v3  = V[3 * pl + rec];
        float        w   = 0.0f;
        if (v3.x > 0.0f)
            w = lum3(sr * v3.y, sg * v3.z, sb * v3.w) / v3.x;
        if (w > 0.0f && w < 3.0e38f)
            emit(rec, pl, w);
        const float4 e = E[rec];
        sr += e.x;
        sg += e.y;
        sb += e.z;
    }
}

/* Exclusive prefix sum over the 256 threads of a block (fixed order). */
__device__ __forceinline__ uint32_t block_scan_256(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t wsum[4];
    const uint32_t      lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t            x    = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t y = __shfl_up(x, (unsigned)d);
        if (lane >= (uint32_t)d)
            x += y;
    }
    if (lane == 63u)
        wsum[w] = x;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t k = 0; k < w; ++k)
        before += wsum[k];
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return before + x - v;
}

__global__ __launch_bounds__(256) void k_guide_gather_count(const float4* V, const float4* E, uint32_t nslots,
                                                            uint32_t maxd, uint32_t* gcount, uint32_t* gblock)
{
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t       n    = 0;
    if (slot < nslots)
    {
        guide_walk(V, E, nslots, slot, maxd, [&](size_t, size_t, float) { ++n; });
        gcount[slot] = n;
    }
    uint32_t total;
    (void)block_scan_256(n, &total);
    if (threadIdx.x == 0)
        gblock[blockIdx.x] = total;
}

/* gblock[b] <- count before block b (starting at the samples already held),
 * in block order; the new total goes to gc->count.  One block of 256. */
__global__ __launch_bounds__(256) void k_guide_gather_scan(uint32_t* gblock, uint32_t nblocks, GuideCounters* gc)
{
    uint32_t base = gc->count;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += 256u)
    {
        const uint32_t b = b0 + threadIdx.x;
        const uint32_t v = b < nblocks ? gblock[b] : 0u;
        uint32_t       total;
        const uint32_t ex = block_scan_256(v, &total);
        if (b < nblocks)
            gblock[b] = base + ex;
        base += total;
    }
    if (threadIdx.x == 0)
        gc->count = base;
}

__global__ __launch_bounds__(256) void k_guide_gather_write(const float4* V, const float4* E, uint32_t nslots,
                                                            uint32_t maxd, const uint32_t* gcount,
                                                            const uint32_t* gblock, GuideSample* out, uint32_t cap)
{
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n    = slot < nslots ? gcount[slot] : 0u;
    uint32_t       total;
    uint32_t       idx = gblock[blockIdx.x] + block_scan_256(n, &total);
    if (slot >= nslots || n == 0u)
        return;
    guide_walk(V, E, nslots, slot, maxd, [&](size_t rec, size_t pl, float w) {
        if (idx < cap)
        {
            GuideSample g;
            g.a      = V[rec];
            g.b      = V[pl + rec];
            g.c      = V[2 * pl + rec];
            g.d      = make_float4(w, 0.0f, 0.0f, 0.0f);
            out[idx] = g;
        }
        ++idx;
    });
}

/* ------------------------------------------------------------ launchers */

namespace clpt
{
hipError_t launch_guide_infer(hipStream_t s, const __bf16* img, const uint8_t* img8, const float* scl, const float* q9,
                              uint32_t n, float* out40, uint32_t blocks)
{
    if (img8)
        hipLaunchKernelGGL(k_guide_infer<true>, dim3(blocks), dim3(256), 0, s, (const void*)img8, scl, q9, n, out40);
    else
        hipLaunchKernelGGL(k_guide_infer<false>, dim3(blocks), dim3(256), 0, s, (const void*)img, scl, q9, n, out40);
    return hipGetLastError();
}

hipError_t launch_guide_mixture_batch(hipStream_t s, const float* out40, const float* dir3, const float* u3, uint32_t n,
                                      float* pdf, float* smp3, float* smp_pdf)
{
    const uint32_t blocks = std::max(1u, std::min(4096u, (n + 255u) / 256u));
    hipLaunchKernelGGL(k_guide_mixture_batch, dim3(blocks), dim3(256), 0, s, out40, dir3, u3, n, pdf, smp3, smp_pdf);
    return hipGetLastError();
}

hipError_t launch_guide_query_batch(hipStream_t s, const float lo[3], const float inv_ext[3], const float* p3,
                                    const float* n3, const float* wo3, uint32_t n, float* q9)
{
    GuideBox box;
    for (int k = 0; k < 3; ++k)
    {
        box.lo[k]      = lo[k];
        box.inv_ext[k] = inv_ext[k];
    }
    const uint32_t blocks = std::max(1u, std::min(4096u, (n + 255u) / 256u));
    hipLaunchKernelGGL(k_guide_query_batch, dim3(blocks), dim3(256), 0, s, box, p3, n3, wo3, n, q9);
    return hipGetLastError();
}

hipError_t launch_guide_pack_fp8(hipStream_t s, const float* params, const GuideCounters* gc, uint8_t* img8,
                                 float* scl)
{
    hipLaunchKernelGGL(k_guide_pack_fp8, dim3(CLPT_G_LAYERS), dim3(256), 0, s, params, gc, img8, scl);
    return hipGetLastError();
}

hipError_t launch_guide_pack(hipStream_t s, const float* params, __bf16* W, __bf16* WT, __bf16* img)
{
    hipLaunchKernelGGL(k_guide_pack, dim3(CLPT_G_PARAMS / 256), dim3(256), 0, s, params, W, WT, img);
    return hipGetLastError();
}

hipError_t launch_guide_gather(hipStream_t s, const GuideView& gv, uint32_t maxd, GuideSample* out, uint32_t cap,
                               GuideCounters* gc)
{
    if (gv.nslots == 0 || maxd == 0)
        return hipSuccess;
    const uint32_t nb = (gv.nslots + 255) / 256;
    hipLaunchKernelGGL(k_guide_gather_count, dim3(nb), dim3(256), 0, s, gv.V, gv.E, gv.nslots, maxd, gv.gcount,
                       gv.gblock);
    hipLaunchKernelGGL(k_guide_gather_scan, dim3(1), dim3(256), 0, s, gv.gblock, nb, gc);
    hipLaunchKernelGGL(k_guide_gather_write, dim3(nb), dim3(256), 0, s, gv.V, gv.E, gv.nslots, maxd, gv.gcount,
                       gv.gblock, out, cap);
    return hipGetLastError();
}

uint32_t guide_fused_blocks(uint32_t N, uint32_t target, uint32_t* iters)
{
    const uint32_t passes = N / CLPT_GT_S;
    const uint32_t it     = std::max(1u, (passes + std::max(1u, target) - 1u) / std::max(1u, target));
    *iters                = it;
    return (passes + it - 1u) / it;
}

hipError_t launch_guide_train(hipStream_t s, const GuideTrainBufs& b, uint32_t N, float lr, float* grad_out,
                              uint32_t step, uint32_t nsteps)
{
    if (N % 1024u != 0u || N == 0u || step >= nsteps)
        return hipErrorInvalidValue;
    if (b.fused_blocks)
    {
        uint32_t       iters = 1;
        const uint32_t nblk  = guide_fused_blocks(N, b.fused_blocks, &iters);
        if (nblk > CLPT_GT_MAXBLK || nblk > b.part_cap)
            return hipErrorInvalidValue;
        const GuideSample* smp = b.samples + step;
        hipLaunchKernelGGL(k_guide_train_fused, dim3(nblk), dim3(64 * CLPT_GT_WAVES), 0, s, smp, b.gc, step, nsteps,
                           N, iters, (const __bf16*)b.img, (const __bf16*)b.W, b.partial, b.wpart, b.spart);
        hipLaunchKernelGGL(k_guide_adam_fused, dim3(CLPT_G_PARAMS / CLPT_GA_P), dim3(CLPT_GA_P * CLPT_GA_G), 0, s,
                           b.params, b.m1, b.m2,
                           b.partial, b.wpart, b.spart, nblk, b.gc, step, nsteps, lr, 0.9f, 0.999f, 1e-8f, b.W,
                           b.WT, b.img, grad_out);
        if (b.img8)
            hipLaunchKernelGGL(k_guide_pack_fp8, dim3(CLPT_G_LAYERS), dim3(256), 0, s, b.params, b.gc, b.img8, b.scl);
        return hipGetLastError();
    }
    const uint32_t     nt   = N / 32u;
    const GuideSample* smp  = b.samples + step; /* rows step + k nsteps */
    hipLaunchKernelGGL(k_guide_features, dim3(N / 256), dim3(256), 0, s, smp, b.gc, step, nsteps, N, b.X, b.wpart);
    /* forward: waves = tiles (N/32) x (64/32) */
    const __bf16* in = b.X;
    for (int l = 0; l < 4; ++l)
    {
        hipLaunchKernelGGL(k_guide_gemm_nt<GEMM_RELU_BF16>, dim3(nt * 2 / 4), dim3(256), 0, s, in,
                           b.WT + guide_layer_offset(l), N, guide_layer_in(l), (uint32_t)CLPT_G_HID,
                           (void*)b.act[l], (const __bf16*)nullptr, b.amax_part + (size_t)l * (N / 64u));
        in = b.act[l];
    }
    hipLaunchKernelGGL(k_guide_gemm_nt<GEMM_F32>, dim3(nt * 2 / 4), dim3(256), 0, s, in, b.WT + guide_layer_offset(4),
                       N, (uint32_t)CLPT_G_HID, (uint32_t)CLPT_G_HID, (void*)b.out, (const __bf16*)nullptr,
                       (float*)nullptr);
    hipLaunchKernelGGL(k_guide_loss, dim3(N / 256), dim3(256), 0, s, b.out, smp, b.gc, step, nsteps, N, b.grad[4],
                       b.loss_part, b.wpart, b.amax_part);
    /* backward: G[l] is the gradient w.r.t. the input of layer l+1 (N x 64) */
    const uint32_t chunk = 1024u, nch = N / chunk;
    for (int l = 4; l >= 0; --l)
    {
        const __bf16*  A  = l == 0 ? b.X : b.act[l - 1];
        const uint32_t K  = guide_layer_in(l);
        hipLaunchKernelGGL(k_guide_gemm_tn, dim3((K / 32u) * 2u * nch), dim3(256), 0, s, A, b.grad[l], K,
                           b.partial, guide_layer_offset(l));
        if (l > 0)
            hipLaunchKernelGGL(k_guide_gemm_nt<GEMM_MASK_BF16>, dim3(nt * 2 / 4), dim3(256), 0, s, b.grad[l],
                               b.W + guide_layer_offset(l), N, (uint32_t)CLPT_G_HID, (uint32_t)CLPT_G_HID,
                               (void*)b.grad[l - 1], b.act[l - 1], (float*)nullptr);
    }
    hipLaunchKernelGGL(k_guide_adam, dim3(CLPT_G_PARAMS / 256), dim3(256), 0, s, b.params, b.m1, b.m2, b.partial,
                       nch, b.gc, lr, 0.9f, 0.999f, 1e-8f, b.W, b.WT, b.img, grad_out);
    hipLaunchKernelGGL(k_guide_finish, dim3(1), dim3(256), 0, s, b.gc, b.loss_part, N / 256,
                       (int)(step + 1u == nsteps));
    if (b.img8)
        hipLaunchKernelGGL(k_guide_pack_fp8, dim3(CLPT_G_LAYERS), dim3(256), 0, s, b.params, b.gc, b.img8, b.scl);
    return hipGetLastError();
}
} // namespace clpt
