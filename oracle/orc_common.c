/*
 * orc_common.c -- variant-independent parts of the oracle: the counter-based
 * RNG, the instance-transform inverse, the HDRI sampling tables, the BVH the
 * oracle traverses, and a row-parallel driver.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "orc_internal.h"
#include "../colvillea_amd/csrc/clpt_elem.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ RNG */

/* pcg4d: include/colvillea/libkernel/base/sampler.h:16-36 (after Jarzynski &
 * Olano, "Hash Functions for GPU Rendering", JCGT 9(3), 2020). */
void orc_pcg4d(const uint32_t in[4], uint32_t out[4])
{
    uint32_t x = in[0] * 1664525u + 1013904223u;
    uint32_t y = in[1] * 1664525u + 1013904223u;
    uint32_t z = in[2] * 1664525u + 1013904223u;
    uint32_t w = in[3] * 1664525u + 1013904223u;
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
    out[0] = x;
    out[1] = y;
    out[2] = z;
    out[3] = w;
}

/* uintToFloat: sampler.h:39-42 -- 23 high bits as mantissa of [1,2), minus 1. */
float orc_uint_to_float(uint32_t x)
{
    return clpt_u2f(0x3f800000u | (x >> 9)) - 1.0f;
}

/* rand(): sampler.h:51-55.  The state is (px, py, frame, w); draw k uses
 * w = k (rand pre-increments w), so draw k is stateless in k. */
float orc_draw(uint32_t px, uint32_t py, uint32_t zlane, uint32_t k)
{
    uint32_t in[4] = {px, py, zlane, k}, out[4];
    orc_pcg4d(in, out);
    return orc_uint_to_float(out[0]);
}

/* Sampler::next1D repeated: seeds are (px, py, frame, w0); emits ndraws. */
void orc_rng_batch(const uint32_t* seeds4, uint32_t n, uint32_t ndraws, uint32_t* out_u32, float* out_f)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        uint32_t st[4] = {seeds4[4 * i], seeds4[4 * i + 1], seeds4[4 * i + 2], seeds4[4 * i + 3]};
        for (uint32_t d = 0; d < ndraws; ++d)
        {
            uint32_t o[4];
            st[3] += 1u;
            orc_pcg4d(st, o);
            if (out_u32)
                out_u32[(size_t)i * ndraws + d] = o[0];
            if (out_f)
                out_f[(size_t)i * ndraws + d] = orc_uint_to_float(o[0]);
        }
    }
}

/* The renderer's float32 elementary functions (clpt_elem.h), exposed so
 * the test suite can pin them against float64 libm: fn 0 sin, 1 cos,
 * 2 atan2(x, y), 3 acos, 4 asin, 5 exp, 6 log, 7 tan. */
void orc_elem_batch(int fn, const float* x, const float* y, uint32_t n, float* out)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        float r;
        switch (fn)
        {
            case 0: r = clpt_sinf(x[i]); break;
            case 1: r = clpt_cosf(x[i]); break;
            case 2: r = clpt_atan2f(x[i], y[i]); break;
            case 3: r = clpt_acosf(x[i]); break;
            case 4: r = clpt_asinf(x[i]); break;
            case 5: r = clpt_expf(x[i]); break;
            case 6: r = clpt_logf(x[i]); break;
            default: r = clpt_tanf(x[i]); break;
        }
        out[i] = r;
    }
}

/* ------------------------------------------------------ transform inverse */

/* Inverse of an affine 3x4 [A|t] by the adjugate, in float64, then rounded.
 * The renderer computes the identical expression (per-instance transforms
 * are an extension: the reference's TLAS uses identity, optix.cu:115). */
void orc_inverse3x4(const float m[12], float out[12], double outd[12])
{
    const double a00 = m[0], a01 = m[1], a02 = m[2], t0 = m[3];
    const double a10 = m[4], a11 = m[5], a12 = m[6], t1 = m[7];
    const double a20 = m[8], a21 = m[9], a22 = m[10], t2 = m[11];
    const double c00 = a11 * a22 - a12 * a21;
    const double c01 = a12 * a20 - a10 * a22;
    const double c02 = a10 * a21 - a11 * a20;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    double       r[12];
    r[0]  = c00 / det;
    r[1]  = (a02 * a21 - a01 * a22) / det;
    r[2]  = (a01 * a12 - a02 * a11) / det;
    r[4]  = c01 / det;
    r[5]  = (a00 * a22 - a02 * a20) / det;
    r[6]  = (a02 * a10 - a00 * a12) / det;
    r[8]  = c02 / det;
    r[9]  = (a01 * a20 - a00 * a21) / det;
    r[10] = (a00 * a11 - a01 * a10) / det;
    r[3]  = -(r[0] * t0 + r[1] * t1 + r[2] * t2);
    r[7]  = -(r[4] * t0 + r[5] * t1 + r[6] * t2);
    r[11] = -(r[8] * t0 + r[9] * t1 + r[10] * t2);
    for (int i = 0; i < 12; ++i)
    {
        if (out)
            out[i] = (float)r[i];
        if (outd)
            outd[i] = r[i];
    }
}

/* ------------------------------------------------------------ HDRI tables */

/* linearToLuminance: math.h:155-158 (double constants, result to float). */
static float orc_luminance(const float* texel)
{
    return (float)((double)texel[0] * 0.212671 + (double)texel[1] * 0.71516 + (double)texel[2] * 0.072169);
}

/* HDRIDome::prefiltering: hdridome.h:311-324.  The reference samples the
 * texture at the texel centre with bilinear filtering, which returns the
 * texel itself; the oracle reads the texel. */
static void orc_prefilter(const float* rgba, uint32_t w, uint32_t h, float* pcond)
{
    for (uint32_t v = 0; v < h; ++v)
    {
        const float sin_theta = clpt_sinf((3.14159265358979323846f * ((float)v + 0.5f)) / (float)h);
        for (uint32_t u = 0; u < w; ++u)
            pcond[(size_t)v * (w + 1) + u] = sin_theta * orc_luminance(rgba + 4 * ((size_t)v * w + u));
    }
}

/* Reference algorithm, float32 serial: preprocessPCondV (hdridome.h:333-385)
 * then preprocessPV (hdridome.h:387-432), including the zero-row and
 * all-black fall-backs. */
void orc_dome_tables_ref32(const float* rgba, uint32_t w, uint32_t h, float* pcond, float* cdfcond, float* pv, float* cdfv)
{
    orc_prefilter(rgba, w, h, pcond);
    const size_t W = w + 1;
    for (uint32_t v = 0; v < h; ++v)
    {
        float* cdf = cdfcond + (size_t)v * W;
        float* p   = pcond + (size_t)v * W;
        cdf[0]     = 0.0f;
        for (uint32_t i = 1; i < w + 1; ++i)
            cdf[i] = cdf[i - 1] + p[i - 1] / (float)w;
        const float c = cdf[w];
        p[w]          = c;
        if (c <= 0.0f)
        {
            for (uint32_t i = 0; i < w; ++i)
            {
                p[i]   = 0.0f;
                cdf[i] = (float)i / (float)w;
            }
            cdf[w] = 1.0f;
            p[w]   = 0.0f;
        }
        else
        {
            for (uint32_t i = 0; i < w + 1; ++i)
                cdf[i] /= c;
        }
        pv[v] = c;
    }
    /* preprocessPV: every job j in [0, h] reads the row integrals pv[0..h-1]
     * as written by preprocessPCondV. */
    float* rowint = (float*)malloc(sizeof(float) * h);
    memcpy(rowint, pv, sizeof(float) * h);
    for (uint32_t j = 0; j <= h; ++j)
    {
        float sum = 0.0f, c = 0.0f;
        for (uint32_t i = 1; i <= h; ++i)
        {
            if (i <= j)
                sum += rowint[i - 1];
            c += rowint[i - 1];
        }
        cdfv[j] = sum / (float)h;
        c /= (float)h;
        if (j == h)
            pv[h] = c;
        if (c <= 0.0f)
        {
            pv[j]   = 0.0f;
            cdfv[j] = (float)j / (float)h;
        }
        else
        {
            cdfv[j] /= c;
        }
    }
    free(rowint);
}

/* The renderer's definition: the same tables with every running sum and
 * integral accumulated in float64 (in this serial order) and rounded once. */
void orc_dome_tables_f64(const float* rgba, uint32_t w, uint32_t h, float* pcond, float* cdfcond, float* pv, float* cdfv)
{
    orc_prefilter(rgba, w, h, pcond);
    const size_t W      = w + 1;
    double*      cdfd   = (double*)malloc(sizeof(double) * W);
    double*      rowint = (double*)malloc(sizeof(double) * h);
    for (uint32_t v = 0; v < h; ++v)
    {
        float* cdf = cdfcond + (size_t)v * W;
        float* p   = pcond + (size_t)v * W;
        double acc = 0.0;
        cdfd[0]    = 0.0;
        for (uint32_t i = 1; i < w + 1; ++i)
        {
            acc += (double)p[i - 1] / (double)w;
            cdfd[i] = acc;
        }
        if (acc <= 0.0)
        {
            for (uint32_t i = 0; i < w; ++i)
            {
                p[i]   = 0.0f;
                cdf[i] = (float)i / (float)w;
            }
            cdf[w]    = 1.0f;
            p[w]      = 0.0f;
            rowint[v] = 0.0;
            pv[v]     = 0.0f;
        }
        else
        {
            for (uint32_t i = 0; i < w + 1; ++i)
                cdf[i] = (float)(cdfd[i] / acc);
            p[w]      = (float)acc;
            rowint[v] = acc;
            pv[v]     = (float)acc;
        }
    }
    double acc = 0.0;
    double* cdfvd = (double*)malloc(sizeof(double) * (h + 1));
    cdfvd[0]      = 0.0;
    for (uint32_t j = 1; j <= h; ++j)
    {
        acc += rowint[j - 1] / (double)h;
        cdfvd[j] = acc;
    }
    if (acc <= 0.0)
    {
        for (uint32_t j = 0; j <= h; ++j)
        {
            pv[j]   = 0.0f;
            cdfv[j] = (float)j / (float)h;
        }
    }
    else
    {
        for (uint32_t j = 0; j <= h; ++j)
            cdfv[j] = (float)(cdfvd[j] / acc);
        pv[h] = (float)acc;
    }
    free(cdfvd);
    free(rowint);
    free(cdfd);
}

/* ---------------------------------------------------------------- BVH */

typedef struct
{
    uint32_t node, start, end;
} orc_build_item;

/* Spatial-median BVH2 over primitive boxes, leaves of <= 4 primitives.  Any
 * conservative hierarchy gives the same closest hit (the oracle's tie rule
 * is total), so the oracle uses the simplest builder. */
static void orc_bvh_build(orc_bvh* out, uint32_t n, const float* plo, const float* phi)
{
    out->prims  = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    out->nodes  = (orc_node*)malloc(sizeof(orc_node) * (2 * (size_t)n + 1));
    out->nnodes = 1;
    float* cen  = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i)
    {
        out->prims[i] = i;
        for (int k = 0; k < 3; ++k)
            cen[3 * i + k] = 0.5f * (plo[3 * i + k] + phi[3 * i + k]);
    }
    /* pending items never exceed the leaf count, which is at most n */
    orc_build_item* stack = (orc_build_item*)malloc(sizeof(orc_build_item) * ((size_t)n + 2));
    int             sp    = 0;
    stack[sp++]           = (orc_build_item){0, 0, n};
    while (sp > 0)
    {
        orc_build_item it = stack[--sp];
        orc_node*      nd = &out->nodes[it.node];
        float          clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; ++k)
        {
            nd->lo[k] = INFINITY;
            nd->hi[k] = -INFINITY;
        }
        for (uint32_t i = it.start; i < it.end; ++i)
        {
            const uint32_t p = out->prims[i];
            for (int k = 0; k < 3; ++k)
            {
                nd->lo[k] = fminf(nd->lo[k], plo[3 * p + k]);
                nd->hi[k] = fmaxf(nd->hi[k], phi[3 * p + k]);
                clo[k]    = fminf(clo[k], cen[3 * p + k]);
                chi[k]    = fmaxf(chi[k], cen[3 * p + k]);
            }
        }
        const uint32_t cnt = it.end - it.start;
        if (cnt <= 4)
        {
            nd->a = it.start;
            nd->n = cnt;
            continue;
        }
        int axis = 0;
        for (int k = 1; k < 3; ++k)
            if (chi[k] - clo[k] > chi[axis] - clo[axis])
                axis = k;
        const float split = 0.5f * (clo[axis] + chi[axis]);
        uint32_t    i = it.start, j = it.end;
        while (i < j)
        {
            if (cen[3 * out->prims[i] + axis] < split)
                ++i;
            else
            {
                --j;
                uint32_t tmp  = out->prims[i];
                out->prims[i] = out->prims[j];
                out->prims[j] = tmp;
            }
        }
        uint32_t mid = i;
        if (mid == it.start || mid == it.end)
            mid = it.start + cnt / 2;
        const uint32_t left = out->nnodes;
        out->nnodes += 2;
        nd->a       = left;
        nd->n       = 0;
        stack[sp++] = (orc_build_item){left, it.start, mid};
        stack[sp++] = (orc_build_item){left + 1, mid, it.end};
    }
    free(stack);
    free(cen);
}

static void orc_bvh_free(orc_bvh* b)
{
    free(b->nodes);
    free(b->prims);
}

orc_scene* orc_scene_create(const orc_scene_desc* desc)
{
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->d         = *desc;
    s->meshes    = (orc_mesh*)malloc(sizeof(orc_mesh) * (desc->nmeshes ? desc->nmeshes : 1));
    memcpy(s->meshes, desc->meshes, sizeof(orc_mesh) * desc->nmeshes);
    s->instances = (orc_instance*)malloc(sizeof(orc_instance) * (desc->ninstances ? desc->ninstances : 1));
    memcpy(s->instances, desc->instances, sizeof(orc_instance) * desc->ninstances);
    s->materials = (orc_material*)malloc(sizeof(orc_material) * (desc->nmaterials ? desc->nmaterials : 1));
    memcpy(s->materials, desc->materials, sizeof(orc_material) * desc->nmaterials);
    s->emitters = (orc_emitter*)malloc(sizeof(orc_emitter) * (desc->nemitters ? desc->nemitters : 1));
    memcpy(s->emitters, desc->emitters, sizeof(orc_emitter) * desc->nemitters);
    s->d.meshes    = s->meshes;
    s->d.instances = s->instances;
    s->d.materials = s->materials;
    s->d.emitters  = s->emitters;
    s->sun_tan     = (float*)malloc(sizeof(float) * (desc->nemitters ? desc->nemitters : 1));
    for (uint32_t e = 0; e < desc->nemitters; ++e)
        s->sun_tan[e] = (float)tan((double)desc->emitters[e].angular_radius);

    /* BLAS per mesh over triangle boxes. */
    s->blas = (orc_bvh*)calloc(desc->nmeshes ? desc->nmeshes : 1, sizeof(orc_bvh));
    float* mlo = (float*)malloc(sizeof(float) * 3 * (desc->nmeshes ? desc->nmeshes : 1));
    float* mhi = (float*)malloc(sizeof(float) * 3 * (desc->nmeshes ? desc->nmeshes : 1));
    for (uint32_t m = 0; m < desc->nmeshes; ++m)
    {
        const orc_mesh* me = &desc->meshes[m];
        float*          lo = (float*)malloc(sizeof(float) * 3 * (me->nt ? me->nt : 1));
        float*          hi = (float*)malloc(sizeof(float) * 3 * (me->nt ? me->nt : 1));
        for (int k = 0; k < 3; ++k)
        {
            mlo[3 * m + k] = INFINITY;
            mhi[3 * m + k] = -INFINITY;
        }
        for (uint32_t t = 0; t < me->nt; ++t)
        {
            for (int k = 0; k < 3; ++k)
            {
                lo[3 * t + k] = INFINITY;
                hi[3 * t + k] = -INFINITY;
            }
            for (int c = 0; c < 3; ++c)
            {
                const float* v = me->vertices + 3 * (size_t)me->indices[3 * (size_t)t + c];
                for (int k = 0; k < 3; ++k)
                {
                    lo[3 * t + k] = fminf(lo[3 * t + k], v[k]);
                    hi[3 * t + k] = fmaxf(hi[3 * t + k], v[k]);
                }
            }
            for (int k = 0; k < 3; ++k)
            {
                mlo[3 * m + k] = fminf(mlo[3 * m + k], lo[3 * t + k]);
                mhi[3 * m + k] = fmaxf(mhi[3 * m + k], hi[3 * t + k]);
            }
        }
        orc_bvh_build(&s->blas[m], me->nt, lo, hi);
        free(lo);
        free(hi);
    }
    /* Instance inverses and world boxes (8 transformed corners, float64). */
    s->inv    = (float(*)[12])malloc(sizeof(float) * 12 * (desc->ninstances ? desc->ninstances : 1));
    s->invd   = (double(*)[12])malloc(sizeof(double) * 12 * (desc->ninstances ? desc->ninstances : 1));
    float* ilo = (float*)malloc(sizeof(float) * 3 * (desc->ninstances ? desc->ninstances : 1));
    float* ihi = (float*)malloc(sizeof(float) * 3 * (desc->ninstances ? desc->ninstances : 1));
    for (uint32_t i = 0; i < desc->ninstances; ++i)
    {
        const orc_instance* in = &desc->instances[i];
        orc_inverse3x4(in->xform, s->inv[i], s->invd[i]);
        const uint32_t m = in->mesh;
        for (int k = 0; k < 3; ++k)
        {
            ilo[3 * i + k] = INFINITY;
            ihi[3 * i + k] = -INFINITY;
        }
        for (int c = 0; c < 8; ++c)
        {
            const double p[3] = {(c & 1) ? mhi[3 * m] : mlo[3 * m], (c & 2) ? mhi[3 * m + 1] : mlo[3 * m + 1],
                                 (c & 4) ? mhi[3 * m + 2] : mlo[3 * m + 2]};
            for (int k = 0; k < 3; ++k)
            {
                const double q = in->xform[4 * k] * p[0] + in->xform[4 * k + 1] * p[1] + in->xform[4 * k + 2] * p[2] +
                    in->xform[4 * k + 3];
                /* widen by one float ulp each way so rounding never shrinks the box */
                ilo[3 * i + k] = fminf(ilo[3 * i + k], nextafterf((float)q, -INFINITY));
                ihi[3 * i + k] = fmaxf(ihi[3 * i + k], nextafterf((float)q, INFINITY));
            }
        }
    }
    orc_bvh_build(&s->tlas, desc->ninstances, ilo, ihi);
    free(ilo);
    free(ihi);
    free(mlo);
    free(mhi);
    return s;
}

void orc_scene_destroy(orc_scene* s)
{
    if (!s)
        return;
    for (uint32_t m = 0; m < s->d.nmeshes; ++m)
        orc_bvh_free(&s->blas[m]);
    orc_bvh_free(&s->tlas);
    free(s->blas);
    free(s->inv);
    free(s->invd);
    free(s->sun_tan);
    free(s->meshes);
    free(s->instances);
    free(s->materials);
    free(s->emitters);
    free(s);
}

/* -------------------------------------------------------- row parallelism */

typedef struct
{
    orc_row_fn fn;
    void*      ctx;
    uint32_t   next, end, worker;
} orc_pool;

typedef struct
{
    orc_pool* pool;
    uint32_t  id;
} orc_worker_arg;

static void* orc_worker(void* p)
{
    orc_worker_arg* a = (orc_worker_arg*)p;
    for (;;)
    {
        const uint32_t row = __atomic_fetch_add(&a->pool->next, 1u, __ATOMIC_RELAXED);
        if (row >= a->pool->end)
            break;
        a->pool->fn(a->pool->ctx, row, a->id);
    }
    return NULL;
}

void orc_parallel_rows(uint32_t nthreads, uint32_t y0, uint32_t y1, orc_row_fn fn, void* ctx)
{
    if (nthreads <= 1)
    {
        for (uint32_t y = y0; y < y1; ++y)
            fn(ctx, y, 0);
        return;
    }
    orc_pool        pool = {fn, ctx, y0, y1, 0};
    pthread_t*      th   = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    orc_worker_arg* args = (orc_worker_arg*)malloc(sizeof(orc_worker_arg) * nthreads);
    for (uint32_t t = 0; t < nthreads; ++t)
    {
        args[t] = (orc_worker_arg){&pool, t};
        pthread_create(&th[t], NULL, orc_worker, &args[t]);
    }
    for (uint32_t t = 0; t < nthreads; ++t)
        pthread_join(th[t], NULL);
    free(args);
    free(th);
}
