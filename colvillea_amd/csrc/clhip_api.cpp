/*
 * clhip_api.cpp -- the C ABI (include/clhip.h): context, scene upload,
 * BVH build, the bounce loop of Integrator::render and the batch parity
 * entry points.  Host code; the kernels are in kernels.hip.
 */
#include "../../include/clhip.h"
#include "bvh_build.h"
#include "clpt_guide.h"
#include "clpt_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace
{
thread_local std::string g_last_error;

struct DevBuf
{
    void*  p     = nullptr;
    size_t bytes = 0;
    DevBuf()     = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p     = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n)
    {
        if (n <= bytes && p)
            return hipSuccess;
        release();
        if (n == 0)
            n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess)
            bytes = n;
        else
            p = nullptr;
        return e;
    }
    template <typename T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

struct TexHost
{
    std::unique_ptr<DevBuf> buf;
    uint32_t                w = 0, h = 0;
};

struct MeshInfo
{
    uint32_t tri_base = 0, ntris = 0, nverts = 0, root = 0, depth = 0;
    float    lo[3]    = {0, 0, 0}, hi[3] = {0, 0, 0};
};

/* One mesh's tree (its own node and triangle numbering) and records. */
struct MeshBuild
{
    std::vector<Node4>    nodes4;
    std::vector<Node8>    nodes8;
    std::vector<TriRec>   tris;
    std::vector<float>    pos, nrm, tan, uv;
    std::vector<uint32_t> idx;
    uint32_t              flags = 0; /* MESH_HAS_* */
    MeshInfo              info;
};

struct TimedLaunch
{
    hipEvent_t a, b;
    int        kind; /* 0 trace, 1 shade, 2 shadow */
};

/* Pinned host buffer holding one render's batch tables until the stream
 * has copied them (event `done`). */
struct StagingSlot
{
    void*      p       = nullptr;
    size_t     cap     = 0;
    hipEvent_t done    = nullptr;
    bool       pending = false;
};

/* One instantiated batch graph and the launch arguments it was captured
 * with (kernel arguments are baked into the graph). */
struct GraphSlot
{
    std::string     key;
    hipGraphExec_t  exec = nullptr;
    hipGraph_t      graph = nullptr;
    uint64_t        used  = 0;
};
} // namespace

struct clhip_ctx
{
    int         device  = 0;
    hipStream_t stream  = nullptr;
    uint32_t    width   = 0, height = 0, itype = 0;
    uint32_t    num_cus = 0;
    std::string err;

    std::vector<TexHost> textures;

    /* geometry */
    uint32_t                 bvh_width = 4;
    std::vector<Node4>       blas_nodes4;
    std::vector<Node8>       blas_nodes8;
    std::vector<MeshInfo>    meshes;
    std::vector<MeshRec>     mesh_recs;
    std::vector<MeshBuild>   mesh_builds; /* per-mesh trees, reassembled on every BLAS (re)build */
    uint64_t                 blas_builds = 0; /* mesh trees built since creation */
    bool                     have_blas = false, have_tlas = false;
    DevBuf                   d_nodes, d_tris, d_meshes, d_pos, d_nrm, d_tan, d_uv, d_idx, d_inst;
    size_t                   node_bytes = 0, tri_bytes = 0; /* in use in d_nodes, d_tris */
    uint2*                   noalloc_stack = nullptr; /* this device's scratch of clhip_trace_rays_device */
    std::vector<InstanceRec> inst;
    std::vector<uint32_t>    inst_entity;
    uint32_t                 tlas_root = 0, tlas_depth = 0;
    uint32_t                 world_inst = 0xffffffffu; /* instance whose mesh subtrees the top tree holds */
    bool                     inst_dirty = true;

    /* materials / entities */
    std::vector<MaterialRec> mtl;
    std::vector<uint32_t>    entity_material;
    DevBuf                   d_mtl;
    DevBuf                   d_texrec; /* TexRec per texture, for material lookups */

    /* emitters */
    std::vector<EmitterRec> emitters;
    int32_t                 dome_index = -1, dome_tex = -1;
    uint32_t                dome_w = 0, dome_h = 0;
    DevBuf                  d_emitters, d_pcond, d_cdfcond, d_pv, d_cdfv, d_rowint, d_cutv, d_cutcond;

    /* camera / parameters (the single-view film) */
    ViewRec      cam{};
    bool         have_camera = false;
    clhip_params params{5u, 0u, 0u, 0u};

    /* render state */
    DevBuf   d_accum, d_out;
    DevBuf   d_q[2][3], d_hit, d_sq[3], d_pacc, d_gstack, d_bc, d_tot, d_ctr;
    size_t   qcap = 0;
    uint32_t spp  = 0;
    uint32_t blocks_trace = 0, blocks_shadow = 0, blocks_shade = 0;
    clhip_launch launch{};      /* as set; zero fields = defaults */
    uint32_t     def_trace = 0, def_shadow = 0;

    /* batch tables: device copies at stable addresses (grow only), filled
     * from a ring of pinned staging buffers on the stream */
    DevBuf                d_views, d_batches, d_entries, d_cursor;
    StagingSlot           staging[4];
    uint32_t              staging_next = 0;
    std::vector<GraphSlot> graphs;
    uint64_t              graph_clock = 0;

    /* multi-view film */
    DevBuf                d_vaccum;
    std::vector<ViewRec>  vviews;
    std::vector<uint32_t> vspp, vlast_batch;
    std::vector<hipEvent_t> batch_events;
    bool                  have_views = false;

    /* learned proposal */
    float       world_lo[3] = {0.0f, 0.0f, 0.0f}, world_hi[3] = {1.0f, 1.0f, 1.0f};
    clhip_guide guide{};
    bool        guide_ready = false;
    DevBuf      g_params, g_m1, g_m2, g_W, g_WT, g_img, g_gc, g_img8, g_scl;
    DevBuf      g_X, g_act[4], g_out, g_grad[5], g_loss, g_wpart, g_spart, g_amaxp, g_partial, g_samples;
    uint32_t    g_partcap = 0; /* blocks the per-block gradient buffers hold */
    uint32_t    g_N = 0;
    DevBuf      g_V, g_E, g_tag, g_gcount, g_gblock;
    uint32_t    g_slots = 0;
    size_t      g_tagcap = 0;
    uint32_t    blocks_guided = 0;
    uint64_t    g_issued = 0; /* training steps issued since the last reset */
    std::vector<TimedLaunch> train_timed;

    /* timing */
    bool                      timing = false;
    std::vector<TimedLaunch>  timed;
    std::vector<hipEvent_t>   event_pool;
    hipEvent_t                ev_begin = nullptr, ev_end = nullptr;
    bool                      have_render_events = false;
    uint64_t                  trace_launches = 0, shadow_launches = 0;
    uint64_t                  nbatches = 0, graph_replays = 0, graph_captures = 0;
};

namespace
{
int fail(clhip_ctx* ctx, int code, const char* fmt, ...)
{
    char    buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx)
        ctx->err = buf;
    g_last_error = buf;
    return code;
}

#define CLHIP_CHECK(ctx, expr)                                                                           \
    do                                                                                                   \
    {                                                                                                    \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return fail((ctx), CLHIP_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                       \
    } while (0)

template <typename T>
hipError_t upload(DevBuf& b, const std::vector<T>& v)
{
    hipError_t e = b.alloc(std::max<size_t>(v.size() * sizeof(T), 16));
    if (e != hipSuccess)
        return e;
    if (v.empty())
        return hipSuccess;
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

/* Inverse of an affine 3x4 by the adjugate in float64 (the oracle's
 * orc_inverse3x4 uses the identical expression). */
void inverse3x4(const float m[12], float out[12])
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
        out[i] = (float)r[i];
}

SceneView scene_view(const clhip_ctx* c)
{
    SceneView s{};
    s.nodes      = c->d_nodes.p;
    s.bvh_width  = c->bvh_width;
    s.tris       = c->d_tris.as<const TriRec>();
    s.node_bytes = (uint32_t)std::min<size_t>(c->node_bytes, 0xffffffffu);
    s.inst       = c->d_inst.as<const InstanceRec>();
    s.meshes     = c->d_meshes.as<const MeshRec>();
    s.pos        = c->d_pos.as<const float>();
    s.nrm        = c->d_nrm.as<const float>();
    s.tan        = c->d_tan.as<const float>();
    s.uv         = c->d_uv.as<const float>();
    s.idx        = c->d_idx.as<const uint32_t>();
    s.mtl        = c->d_mtl.as<const MaterialRec>();
    s.tex        = c->d_texrec.as<const TexRec>();
    s.ntex       = (uint32_t)c->textures.size();
    s.emitters   = c->d_emitters.as<const EmitterRec>();
    s.nemitters  = (uint32_t)c->emitters.size();
    s.dome_index = c->dome_index;
    if (c->dome_tex >= 0)
    {
        s.dome_tex = c->textures[c->dome_tex].buf->as<const float4>();
        s.dome_w   = c->dome_w;
        s.dome_h   = c->dome_h;
    }
    s.pcond     = c->d_pcond.as<const float>();
    s.cdfcond   = c->d_cdfcond.as<const float>();
    s.pv        = c->d_pv.as<const float>();
    s.cdfv      = c->d_cdfv.as<const float>();
    s.cutv      = c->d_cutv.as<const uint32_t>();
    s.cutcond   = c->d_cutcond.as<const uint32_t>();
    s.tlas_root  = c->tlas_root;
    s.world_inst = c->world_inst;
    s.ninst      = (uint32_t)c->inst.size();
    return s;
}

/* Resolve entity -> material for every instance and upload. */
int sync_instances(clhip_ctx* c)
{
    if (!c->have_tlas)
        return fail(c, CLHIP_E_STATE, "no TLAS: call clhip_build_tlas first");
    if (!c->inst_dirty)
        return CLHIP_OK;
    for (size_t i = 0; i < c->inst.size(); ++i)
    {
        const uint32_t e = c->inst_entity[i];
        uint32_t       m = 0;
        if (e < c->entity_material.size())
            m = c->entity_material[e];
        else if (c->entity_material.empty())
            m = e; /* no entity table: entity index doubles as material index */
        else
            return fail(c, CLHIP_E_INVALID, "instance %zu references entity %u of %zu", i, e,
                        c->entity_material.size());
        if (m >= c->mtl.size())
            return fail(c, CLHIP_E_STATE, "instance %zu needs material %u but only %zu are built", i, m, c->mtl.size());
        c->inst[i].material = m;
        c->inst[i].entity   = e;
    }
    CLHIP_CHECK(c, upload(c->d_inst, c->inst));
    c->inst_dirty = false;
    return CLHIP_OK;
}

/* Paths per batch: the launch setting, else 2^27.  A persistent traversal
 * launch finishes with its slowest lane, and the last bounces of a batch
 * hold few rays, so every batch pays a fixed cost whatever its size.  2^27
 * paths (64 spp of a 1080p film, 164 B of queues and path sums per path = 22.0 GB of the
 * 288 GB of HBM) make it small: 971 Msamples/s on the 1080p scene against
 * 949 at 2^26 and 913 at 2^25.  Buffers are sized by the batch a render
 * actually uses, so small films and few samples allocate less. */
uint32_t default_paths_per_batch(const clhip_ctx* c)
{
    return c->launch.paths_per_batch ? c->launch.paths_per_batch : (128u << 20);
}

int ensure_render_buffers(clhip_ctx* c, size_t paths)
{
    const size_t npix = (size_t)c->width * c->height;
    const size_t gthreads = (size_t)std::max(c->blocks_trace, c->blocks_shadow) * 256;
    if (paths > c->qcap || 3 * npix * sizeof(double) > c->d_accum.bytes ||
        gthreads * CLPT_GSTACK_ENTRIES * sizeof(uint2) > c->d_gstack.bytes)
        CLHIP_CHECK(c, hipStreamSynchronize(c->stream)); /* buffers about to move may be in use */
    CLHIP_CHECK(c, c->d_accum.alloc(3 * npix * sizeof(double)));
    CLHIP_CHECK(c, c->d_out.alloc(npix * sizeof(float4)));
    if (paths > c->qcap)
    {
        for (int q = 0; q < 2; ++q)
            for (int k = 0; k < 3; ++k)
                CLHIP_CHECK(c, c->d_q[q][k].alloc(paths * sizeof(float4)));
        CLHIP_CHECK(c, c->d_hit.alloc(paths * sizeof(float4)));
        CLHIP_CHECK(c, c->d_sq[0].alloc(paths * sizeof(float4)));
        CLHIP_CHECK(c, c->d_sq[1].alloc(paths * sizeof(float4)));
        CLHIP_CHECK(c, c->d_sq[2].alloc(paths * sizeof(float2)));
        CLHIP_CHECK(c, c->d_pacc.alloc(paths * 3 * sizeof(float)));
        c->qcap = paths;
    }
    CLHIP_CHECK(c, c->d_gstack.alloc(gthreads * CLPT_GSTACK_ENTRIES * sizeof(uint2)));
    if (!c->d_cursor.p)
        CLHIP_CHECK(c, c->d_cursor.alloc(256));
    if (!c->d_bc.p)
    {
        CLHIP_CHECK(c, c->d_bc.alloc(sizeof(BatchCounters)));
        CLHIP_CHECK(c, hipMemsetAsync(c->d_bc.p, 0, sizeof(BatchCounters), c->stream));
    }
    if (!c->d_tot.p)
    {
        CLHIP_CHECK(c, c->d_tot.alloc(sizeof(TotalCounters)));
        CLHIP_CHECK(c, hipMemsetAsync(c->d_tot.p, 0, sizeof(TotalCounters), c->stream));
    }
    if (!c->d_ctr.p)
    {
        CLHIP_CHECK(c, c->d_ctr.alloc(4 * sizeof(unsigned long long)));
        CLHIP_CHECK(c, hipMemsetAsync(c->d_ctr.p, 0, 4 * sizeof(unsigned long long), c->stream));
    }
    return CLHIP_OK;
}

hipEvent_t pool_event(clhip_ctx* c)
{
    if (!c->event_pool.empty())
    {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void recycle_events(clhip_ctx* c)
{
    for (auto& t : c->timed)
    {
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->timed.clear();
}

void recycle_batch_events(clhip_ctx* c)
{
    for (hipEvent_t e : c->batch_events)
        c->event_pool.push_back(e);
    c->batch_events.clear();
}

/* ---- learned proposal: buffers and state ---- */

bool guiding(const clhip_ctx* c) { return c->guide.enable != 0u && c->guide_ready; }

int guide_reset_weights(clhip_ctx* c, uint64_t seed)
{
    if (!c->g_params.p)
    {
        CLHIP_CHECK(c, c->g_params.alloc(CLPT_G_PARAMS * sizeof(float)));
        CLHIP_CHECK(c, c->g_m1.alloc(CLPT_G_PARAMS * sizeof(float)));
        CLHIP_CHECK(c, c->g_m2.alloc(CLPT_G_PARAMS * sizeof(float)));
        CLHIP_CHECK(c, c->g_W.alloc(CLPT_G_PARAMS * sizeof(__bf16)));
        CLHIP_CHECK(c, c->g_WT.alloc(CLPT_G_PARAMS * sizeof(__bf16)));
        CLHIP_CHECK(c, c->g_img.alloc(CLPT_G_IMG * sizeof(__bf16)));
        CLHIP_CHECK(c, c->g_gc.alloc(sizeof(GuideCounters)));
        CLHIP_CHECK(c, c->g_img8.alloc(CLPT_G_IMG));
        CLHIP_CHECK(c, c->g_scl.alloc(16 * sizeof(float)));
        CLHIP_CHECK(c, hipMemset(c->g_img.p, 0, CLPT_G_IMG * sizeof(__bf16)));
        CLHIP_CHECK(c, hipMemset(c->g_img8.p, 0, CLPT_G_IMG));
    }
    /* He-uniform weights from a SplitMix64 stream; the output layer is
     * scaled by 0.1 and its unused columns (40..63) are zero */
    std::vector<float> w(CLPT_G_PARAMS, 0.0f);
    uint64_t           st = seed;
    auto               next = [&st]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < CLPT_G_LAYERS; ++l)
    {
        const uint32_t in = guide_layer_in(l), off = guide_layer_offset(l);
        const double   a  = std::sqrt(6.0 / in) * (l == CLPT_G_LAYERS - 1 ? 0.1 : 1.0);
        for (uint32_t k = 0; k < in; ++k)
            for (uint32_t m = 0; m < (uint32_t)CLPT_G_HID; ++m)
            {
                const double u = next();
                if (l < CLPT_G_LAYERS - 1 || m < (uint32_t)CLPT_G_NOUT)
                    w[off + k * CLPT_G_HID + m] = (float)((2.0 * u - 1.0) * a);
            }
    }
    CLHIP_CHECK(c, hipStreamSynchronize(c->stream));
    CLHIP_CHECK(c, hipMemcpy(c->g_params.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    CLHIP_CHECK(c, hipMemset(c->g_m1.p, 0, CLPT_G_PARAMS * sizeof(float)));
    CLHIP_CHECK(c, hipMemset(c->g_m2.p, 0, CLPT_G_PARAMS * sizeof(float)));
    CLHIP_CHECK(c, hipMemset(c->g_gc.p, 0, sizeof(GuideCounters)));
    CLHIP_CHECK(c, clpt::launch_guide_pack(c->stream, c->g_params.as<const float>(), c->g_W.as<__bf16>(),
                                           c->g_WT.as<__bf16>(), c->g_img.as<__bf16>()));
    CLHIP_CHECK(c, clpt::launch_guide_pack_fp8(c->stream, c->g_params.as<const float>(), c->g_gc.as<GuideCounters>(),
                                               c->g_img8.as<uint8_t>(), c->g_scl.as<float>()));
    CLHIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->guide_ready = true;
    c->g_issued    = 0;
    return CLHIP_OK;
}

int guide_train_buffers(clhip_ctx* c, uint32_t N, uint32_t nsteps = 1)
{
    if ((size_t)N * nsteps > c->g_samples.bytes / sizeof(GuideSample))
    {
        CLHIP_CHECK(c, hipStreamSynchronize(c->stream));
        CLHIP_CHECK(c, c->g_samples.alloc((size_t)N * nsteps * sizeof(GuideSample)));
    }
    if (N <= c->g_N)
        return CLHIP_OK;
    CLHIP_CHECK(c, hipStreamSynchronize(c->stream));
    CLHIP_CHECK(c, c->g_X.alloc((size_t)N * CLPT_G_IN * sizeof(__bf16)));
    for (int l = 0; l < 4; ++l)
        CLHIP_CHECK(c, c->g_act[l].alloc((size_t)N * CLPT_G_HID * sizeof(__bf16)));
    for (int l = 0; l < 5; ++l)
        CLHIP_CHECK(c, c->g_grad[l].alloc((size_t)N * CLPT_G_HID * sizeof(__bf16)));
    CLHIP_CHECK(c, c->g_out.alloc((size_t)N * CLPT_G_HID * sizeof(float)));
    CLHIP_CHECK(c, c->g_loss.alloc((size_t)(N / 256) * sizeof(float)));
    /* per-block gradients: N / 1024 chunks per GEMM, or the fused step's
     * blocks (at most one per 128 samples and at most the target count, so
     * this also covers every smaller N) */
    const uint32_t nblk = std::max(N / 1024u, std::min(N / 128u, std::max(1u, c->num_cus)));
    CLHIP_CHECK(c, c->g_wpart.alloc((size_t)std::max(N / 256u, nblk) * sizeof(double)));
    CLHIP_CHECK(c, c->g_spart.alloc((size_t)nblk * 5u * sizeof(float)));
    CLHIP_CHECK(c, c->g_amaxp.alloc((size_t)4 * (N / 64) * sizeof(float)));
    CLHIP_CHECK(c, c->g_partial.alloc((size_t)nblk * CLPT_G_PARAMS * sizeof(float)));
    c->g_partcap = nblk;
    CLHIP_CHECK(c, c->g_samples.alloc((size_t)N * std::max(1u, nsteps) * sizeof(GuideSample)));
    c->g_N = N;
    return CLHIP_OK;
}

/* Recording buffers: training records for `slots` paths and a tag per
 * shadow-queue entry. */
int guide_record_buffers(clhip_ctx* c, uint32_t slots, size_t qcap)
{
    if (slots > c->g_slots)
    {
        CLHIP_CHECK(c, hipStreamSynchronize(c->stream));
        CLHIP_CHECK(c, c->g_V.alloc((size_t)slots * (CLPT_MAXD + 1) * 4 * sizeof(float4)));
        CLHIP_CHECK(c, c->g_E.alloc((size_t)slots * (CLPT_MAXD + 1) * sizeof(float4)));
        CLHIP_CHECK(c, c->g_gcount.alloc((size_t)slots * sizeof(uint32_t)));
        CLHIP_CHECK(c, c->g_gblock.alloc((size_t)((slots + 255) / 256) * sizeof(uint32_t)));
        c->g_slots = slots;
    }
    if (qcap > c->g_tagcap)
    {
        CLHIP_CHECK(c, hipStreamSynchronize(c->stream));
        CLHIP_CHECK(c, c->g_tag.alloc(qcap * sizeof(uint32_t)));
        c->g_tagcap = qcap;
    }
    return CLHIP_OK;
}

bool guide_recording(const clhip_ctx* c)
{
    return guiding(c) && c->guide.train && (c->guide.max_steps == 0u || c->g_issued < c->guide.max_steps);
}

GuideView guide_view(const clhip_ctx* c)
{
    GuideView g{};
    if (!guiding(c))
        return g;
    g.img    = c->g_img.as<const __bf16>();
    g.img8   = c->guide.precision == 1u ? c->g_img8.as<const uint8_t>() : nullptr;
    g.scl    = c->g_scl.as<const float>();
    g.c_bsdf = c->guide.bsdf_fraction;
    for (int k = 0; k < 3; ++k)
    {
        const float ext = c->world_hi[k] - c->world_lo[k];
        g.lo[k]         = c->world_lo[k];
        g.inv_ext[k]    = ext > 1e-20f ? 1.0f / ext : 1.0f;
    }
    if (guide_recording(c))
    {
        g.nslots = c->guide.record_paths;
        g.V      = c->g_V.as<float4>();
        g.E      = c->g_E.as<float4>();
        g.gcount = c->g_gcount.as<uint32_t>();
        g.gblock = c->g_gblock.as<uint32_t>();
    }
    return g;
}

clpt::GuideTrainBufs guide_bufs(clhip_ctx* c)
{
    clpt::GuideTrainBufs b{};
    b.samples = c->g_samples.as<GuideSample>();
    b.gc      = c->g_gc.as<GuideCounters>();
    b.X       = c->g_X.as<__bf16>();
    for (int l = 0; l < 4; ++l)
        b.act[l] = c->g_act[l].as<__bf16>();
    b.out = c->g_out.as<float>();
    for (int l = 0; l < 5; ++l)
        b.grad[l] = c->g_grad[l].as<__bf16>();
    b.loss_part = c->g_loss.as<float>();
    b.wpart     = c->g_wpart.as<double>();
    b.amax_part = c->g_amaxp.as<float>();
    b.partial   = c->g_partial.as<float>();
    b.params    = c->g_params.as<float>();
    b.m1        = c->g_m1.as<float>();
    b.m2        = c->g_m2.as<float>();
    b.W         = c->g_W.as<__bf16>();
    b.WT        = c->g_WT.as<__bf16>();
    b.img       = c->g_img.as<__bf16>();
    b.img8      = c->guide.precision == 1u ? c->g_img8.as<uint8_t>() : nullptr; /* FP8 image only when used */
    b.scl       = c->g_scl.as<float>();
    b.spart     = c->g_spart.as<float>();
    b.fused_blocks = c->guide.train_kernel == 0u ? c->num_cus : 0u;
    b.part_cap     = c->g_partcap;
    return b;
}

/* Grid sizes from the launch shape (zero = default). */
void apply_launch(clhip_ctx* c)
{
    const clhip_launch& l = c->launch;
    c->blocks_trace       = c->num_cus * (l.trace_blocks_per_cu ? l.trace_blocks_per_cu : c->def_trace);
    c->blocks_shadow      = c->num_cus * (l.shadow_blocks_per_cu ? l.shadow_blocks_per_cu : c->def_shadow);
    c->blocks_shade       = c->num_cus * (l.shade_blocks_per_cu ? l.shade_blocks_per_cu : 8u);
}

bool use_graph(const clhip_ctx* c) { return !c->timing && c->launch.graph != 2u; }

/* Paths of a sample in 16 x 16 pixel tiles by default: on the 1080p scene
 * 1219-1221 Msamples/s in pixel rows against 1224-1231 in tiles (traversal
 * and shading of the first bounces see more compact groups of rays). */
uint32_t launch_tile(const clhip_ctx* c) { return c->launch.tile ? c->launch.tile : 16u; }

/* WavefrontDirectLightingIntegrator (direct.cpp:84-201,
 * raygeneration.cu:139-282): at the first hit one light sample and one BSDF
 * sample, MIS-weighted against each other, the BSDF ray counting only if it
 * escapes to the dome.  That is the path tracer stopped after the second
 * segment without Russian roulette: the same kernels, RNG draws and MIS
 * weights (with the reference's defects corrected as in the path tracer). */
void direct_lighting_depth(const clhip_ctx* c, ViewRec& v)
{
    if (c->itype == CLHIP_INTEGRATOR_DIRECT_LIGHTING)
    {
        v.max_depth = 2u;
        v.rr_depth  = 0u;
    }
}

/* Copies the batch tables to their device arrays, in stream order: the
 * previous render may still read the old tables, so the copies are queued
 * behind it from pinned staging that stays untouched until they ran. */
int upload_tables(clhip_ctx* c, const std::vector<ViewRec>& views, const std::vector<BatchRec>& batches,
                  const std::vector<BatchEntry>& entries)
{
    const size_t vb = views.size() * sizeof(ViewRec), bb = batches.size() * sizeof(BatchRec),
                 eb = entries.size() * sizeof(BatchEntry);
    if (vb > c->d_views.bytes || bb > c->d_batches.bytes || eb > c->d_entries.bytes)
    {
        /* growth moves the arrays: nothing in flight may still read them */
        CLHIP_CHECK(c, hipStreamSynchronize(c->stream));
        CLHIP_CHECK(c, c->d_views.alloc(std::max<size_t>(2 * vb, 4096)));
        CLHIP_CHECK(c, c->d_batches.alloc(std::max<size_t>(2 * bb, 4096)));
        CLHIP_CHECK(c, c->d_entries.alloc(std::max<size_t>(2 * eb, 4096)));
    }
    StagingSlot& s = c->staging[c->staging_next++ % 4];
    if (s.pending)
    {
        CLHIP_CHECK(c, hipEventSynchronize(s.done));
        s.pending = false;
    }
    const size_t need = vb + bb + eb;
    if (need > s.cap)
    {
        if (s.p)
            CLHIP_CHECK(c, hipHostFree(s.p));
        s.p   = nullptr;
        s.cap = 0;
        CLHIP_CHECK(c, hipHostMalloc(&s.p, std::max<size_t>(2 * need, 4096), hipHostMallocDefault));
        s.cap = std::max<size_t>(2 * need, 4096);
    }
    if (!s.done)
        CLHIP_CHECK(c, hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    char* h = static_cast<char*>(s.p);
    std::memcpy(h, views.data(), vb);
    std::memcpy(h + vb, batches.data(), bb);
    std::memcpy(h + vb + bb, entries.data(), eb);
    CLHIP_CHECK(c, hipMemcpyAsync(c->d_views.p, h, vb, hipMemcpyHostToDevice, c->stream));
    CLHIP_CHECK(c, hipMemcpyAsync(c->d_batches.p, h + vb, bb, hipMemcpyHostToDevice, c->stream));
    CLHIP_CHECK(c, hipMemcpyAsync(c->d_entries.p, h + vb + bb, eb, hipMemcpyHostToDevice, c->stream));
    CLHIP_CHECK(c, hipEventRecord(s.done, c->stream));
    s.pending = true;
    return CLHIP_OK;
}

/* Everything the kernels of one batch take as arguments. */
struct BatchArgs
{
    SceneView      sc;
    BatchTable     bt;
    RayQueue       q[2];
    HitBuf         hb;
    ShadowQueue    sq;
    double*        accum;
    float*         pacc; /* escaped-ray radiance per path slot, folded into accum at the batch end */
    BatchCounters* bc;
    TotalCounters* tot;
    uint2*         gst;
    unsigned long long* ctr;
    uint32_t       max_depth, blocks_trace, blocks_shadow, blocks_shade, shade_threads;
    const float*   cam_o; /* the single view's camera position (device), or null */
    GuideView      gv;
    uint32_t       guided, blocks_guided;
    GuideSample*   gsamples; /* gather target when recording */
    GuideCounters* gc;
    uint32_t       gcap;
};

/* single_view: every batch holds one view (view 0), so camera rays share its
 * position and the origin plane of depth 1 is not written */
BatchArgs batch_args(clhip_ctx* c, double* accum, uint32_t max_depth, bool single_view)
{
    BatchArgs a{};
    a.sc = scene_view(c);
    a.bt = BatchTable{c->d_views.as<const ViewRec>(), c->d_batches.as<const BatchRec>(),
                      c->d_entries.as<const BatchEntry>(), c->d_cursor.as<uint32_t>(),
                      launch_tile(c)};
    for (int k = 0; k < 2; ++k)
    {
        a.q[k].o     = c->d_q[k][0].as<float4>();
        a.q[k].d     = c->d_q[k][1].as<float4>();
        a.q[k].T     = c->d_q[k][2].as<float4>();
        a.q[k].count = nullptr;
    }
    a.hb            = HitBuf{c->d_hit.as<float4>()};
    a.sq            = ShadowQueue{c->d_sq[0].as<float4>(), c->d_sq[1].as<float4>(), c->d_sq[2].as<float2>(), nullptr};
    a.accum         = accum;
    a.pacc          = c->d_pacc.as<float>();
    a.bc            = c->d_bc.as<BatchCounters>();
    a.tot           = c->d_tot.as<TotalCounters>();
    a.gst           = c->d_gstack.as<uint2>();
    a.ctr           = c->d_ctr.as<unsigned long long>();
    a.gv            = guide_view(c);
    a.guided        = guiding(c) ? 1u : 0u;
    a.blocks_guided = c->blocks_guided;
    if (a.guided)
        a.sq.tag = c->g_tag.as<uint32_t>();
    if (a.gv.nslots)
    {
        a.gsamples = c->g_samples.as<GuideSample>();
        a.gc       = c->g_gc.as<GuideCounters>();
        a.gcap     = c->guide.batch_size * c->guide.steps_per_batch;
    }
    a.max_depth     = max_depth;
    a.blocks_trace  = c->blocks_trace;
    a.blocks_shadow = c->blocks_shadow;
    a.blocks_shade  = c->blocks_shade;
    a.shade_threads = c->launch.shade_block ? c->launch.shade_block : 256u;
    a.cam_o         = single_view ? reinterpret_cast<const float*>(reinterpret_cast<const char*>(c->d_views.p) +
                                                                   offsetof(ViewRec, pos))
                                  : nullptr;
    return a;
}

/* The launches of one batch: camera rays, then per depth closest hit,
 * shading and (below the last depth) shadow rays, then the batch end that
 * folds the counters and advances the cursor.  With `timed`, HIP events
 * bracket every traversal and shading launch. */
int issue_batch(clhip_ctx* c, const BatchArgs& a, bool timed)
{
    CLHIP_CHECK(c, clpt::launch_camera(c->stream, a.bt, a.gv, a.q[0], a.bc, a.cam_o ? 0u : 1u, a.shade_threads,
                                       a.blocks_shade));
    for (uint32_t depth = 1; depth <= a.max_depth; ++depth)
    {
        const RayQueue& qin  = a.q[(depth - 1) & 1];
        const RayQueue& qout = a.q[depth & 1];
        for (int kind = 0; kind < 3; ++kind)
        {
            if (kind == 2 && depth == a.max_depth)
                break;
            TimedLaunch tl{};
            if (timed)
            {
                tl.a    = pool_event(c);
                tl.b    = pool_event(c);
                tl.kind = kind;
                CLHIP_CHECK(c, hipEventRecord(tl.a, c->stream));
            }
            if (kind == 0)
                CLHIP_CHECK(c, clpt::launch_trace(c->stream, a.sc, qin, bc_rays(a.bc, depth), depth, a.hb, a.gst, a.ctr,
                                                  depth == 1 ? a.cam_o : nullptr, a.blocks_trace));
            else if (kind == 1 && a.guided)
                CLHIP_CHECK(c, clpt::launch_shade_guided(c->stream, a.sc, a.bt, a.gv, depth, qin, a.hb, qout, a.sq,
                                                         a.pacc, a.bc, a.blocks_guided));
            else if (kind == 1)
                CLHIP_CHECK(c, clpt::launch_shade(c->stream, a.sc, a.bt, depth, qin, a.hb, qout, a.sq, a.pacc, a.bc,
                                                  a.shade_threads, a.blocks_shade));
            else
                CLHIP_CHECK(c, clpt::launch_shadow(c->stream, a.sc, a.sq, depth, a.accum, a.bc, a.gst, a.ctr,
                                                   a.gv.nslots ? a.gv.E : nullptr, a.blocks_shadow));
            if (timed)
            {
                CLHIP_CHECK(c, hipEventRecord(tl.b, c->stream));
                c->timed.push_back(tl);
            }
        }
    }
    CLHIP_CHECK(c, clpt::launch_film_fold(c->stream, a.bt, a.pacc, a.accum, a.blocks_shade));
    if (a.gv.nslots)
        CLHIP_CHECK(c, clpt::launch_guide_gather(c->stream, a.gv, a.max_depth, a.gsamples, a.gcap, a.gc));
    CLHIP_CHECK(c, clpt::launch_batch_end(c->stream, a.bc, a.tot, a.bt.cursor));
    return CLHIP_OK;
}

/* One training step on the samples gathered by the last batch (timed with
 * events for the statistics). */
int issue_train(clhip_ctx* c)
{
    TimedLaunch tl{};
    tl.a = pool_event(c);
    tl.b = pool_event(c);
    CLHIP_CHECK(c, hipEventRecord(tl.a, c->stream));
    const uint32_t ns = c->guide.steps_per_batch;
    for (uint32_t k = 0; k < ns; ++k)
        CLHIP_CHECK(c, clpt::launch_guide_train(c->stream, guide_bufs(c), c->guide.batch_size, c->guide.learning_rate,
                                                nullptr, k, ns));
    CLHIP_CHECK(c, hipEventRecord(tl.b, c->stream));
    c->train_timed.push_back(tl);
    c->g_issued += ns;
    return CLHIP_OK;
}

/* The instantiated graph of one batch for these arguments, captured on
 * first use.  At most four shapes are kept (least recently used goes). */
int batch_graph(clhip_ctx* c, const BatchArgs& a, hipGraphExec_t* out)
{
    const std::string key(reinterpret_cast<const char*>(&a), sizeof(a));
    ++c->graph_clock;
    for (GraphSlot& g : c->graphs)
        if (g.exec && g.key == key)
        {
            g.used = c->graph_clock;
            *out   = g.exec;
            return CLHIP_OK;
        }
    GraphSlot* slot = nullptr;
    if (c->graphs.size() < 4)
    {
        c->graphs.emplace_back();
        slot = &c->graphs.back();
    }
    else
    {
        slot = &c->graphs[0];
        for (GraphSlot& g : c->graphs)
            if (g.used < slot->used)
                slot = &g;
        /* the old graph may still be running */
        CLHIP_CHECK(c, hipStreamSynchronize(c->stream));
        if (slot->exec)
            (void)hipGraphExecDestroy(slot->exec);
        if (slot->graph)
            (void)hipGraphDestroy(slot->graph);
        slot->exec  = nullptr;
        slot->graph = nullptr;
    }
    CLHIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    const int  rc = issue_batch(c, a, false);
    hipGraph_t g  = nullptr;
    hipError_t e  = hipStreamEndCapture(c->stream, &g);
    if (rc)
    {
        if (g)
            (void)hipGraphDestroy(g);
        return rc;
    }
    CLHIP_CHECK(c, e);
    hipGraphExec_t x = nullptr;
    e                = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess)
    {
        (void)hipGraphDestroy(g);
        CLHIP_CHECK(c, e);
    }
    slot->key   = key;
    slot->graph = g;
    slot->exec  = x;
    slot->used  = c->graph_clock;
    ++c->graph_captures;
    *out = x;
    return CLHIP_OK;
}

/* Runs batches 0..nbatches-1 of the uploaded tables into `accum`.  Counters
 * and the render events cover this call; with `batch_events` an event is
 * recorded after every batch. */
int run_batches(clhip_ctx* c, uint32_t nbatches, uint32_t max_depth, double* accum, bool batch_events,
                bool single_view)
{
    CLHIP_CHECK(c, hipMemsetAsync(c->d_tot.p, 0, sizeof(TotalCounters), c->stream));
    CLHIP_CHECK(c, hipMemsetAsync(c->d_ctr.p, 0, 4 * sizeof(unsigned long long), c->stream));
    recycle_events(c);
    recycle_batch_events(c);
    c->trace_launches = c->shadow_launches = 0;
    c->nbatches = c->graph_replays = c->graph_captures = 0;
    if (!c->ev_begin)
    {
        CLHIP_CHECK(c, hipEventCreate(&c->ev_begin));
        CLHIP_CHECK(c, hipEventCreate(&c->ev_end));
    }
    for (auto& t : c->train_timed)
    {
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->train_timed.clear();
    if (guiding(c))
    {
        int rc = guide_record_buffers(c, c->guide.train ? c->guide.record_paths : 0u, c->qcap);
        if (!rc && c->guide.train)
            rc = guide_train_buffers(c, c->guide.batch_size, c->guide.steps_per_batch);
        if (rc)
            return rc;
    }
    CLHIP_CHECK(c, clpt::launch_set_cursor(c->stream, c->d_cursor.as<uint32_t>(), 0u));
    BatchArgs      a = batch_args(c, accum, max_depth, single_view);
    hipGraphExec_t x = nullptr;
    if (use_graph(c))
    {
        int rc = batch_graph(c, a, &x);
        if (rc)
            return rc;
    }
    CLHIP_CHECK(c, hipEventRecord(c->ev_begin, c->stream));
    for (uint32_t b = 0; b < nbatches; ++b)
    {
        /* the recording flag flips when training stops: re-derive the
         * arguments (and graph) per batch while guiding */
        if (guiding(c) && b > 0)
        {
            const BatchArgs a2 = batch_args(c, accum, max_depth, single_view);
            if (std::memcmp(&a2, &a, sizeof(a)) != 0)
            {
                a = a2;
                x = nullptr;
                if (use_graph(c))
                {
                    int rc = batch_graph(c, a, &x);
                    if (rc)
                        return rc;
                }
            }
        }
        if (x)
        {
            CLHIP_CHECK(c, hipGraphLaunch(x, c->stream));
            ++c->graph_replays;
        }
        else
        {
            int rc = issue_batch(c, a, c->timing);
            if (rc)
                return rc;
        }
        if (a.gv.nslots)
        {
            int rc = issue_train(c);
            if (rc)
                return rc;
        }
        if (batch_events)
        {
            hipEvent_t e = pool_event(c);
            CLHIP_CHECK(c, hipEventRecord(e, c->stream));
            c->batch_events.push_back(e);
        }
    }
    CLHIP_CHECK(c, hipEventRecord(c->ev_end, c->stream));
    c->have_render_events = true;
    c->nbatches           = nbatches;
    c->trace_launches     = (uint64_t)nbatches * max_depth;
    c->shadow_launches    = (uint64_t)nbatches * (max_depth - 1);
    return CLHIP_OK;
}

/* Scratch device buffer for the batch entry points. */
struct Tmp
{
    DevBuf b;
    template <typename T>
    hipError_t from_host(const T* h, size_t count)
    {
        hipError_t e = b.alloc(std::max<size_t>(count * sizeof(T), 16));
        if (e != hipSuccess || count == 0)
            return e;
        return hipMemcpy(b.p, h, count * sizeof(T), hipMemcpyHostToDevice);
    }
    hipError_t zeros(size_t bytes)
    {
        hipError_t e = b.alloc(std::max<size_t>(bytes, 16));
        if (e != hipSuccess)
            return e;
        return hipMemset(b.p, 0, std::max<size_t>(bytes, 16));
    }
    template <typename T>
    hipError_t to_host(T* h, size_t count)
    {
        if (count == 0)
            return hipSuccess;
        return hipMemcpy(h, b.p, count * sizeof(T), hipMemcpyDeviceToHost);
    }
};
/* Partial re-braiding (Benthin, Woop, Wald and Afra, "Improved Two-Level
 * BVHs using Partial Re-Braiding", HPG 2017) for one instance: the top
 * tree is built over the largest subtrees of that instance's mesh tree in
 * place of the instance's one box, so that its parts and the other
 * instances interleave in one hierarchy.  Only an instance whose transform
 * is the identity qualifies (its subtrees are valid in world space as they
 * are); traversal tests their triangles with the world ray and records
 * their hits under SceneView::world_inst.  The parts are found by opening
 * inner nodes from the mesh tree's root, largest surface area first, until
 * kRebraidParts are held; each keeps the box its parent stores for it (for
 * BVH8 the decoded box, which contains the subtree). */
constexpr size_t kRebraidParts = 64;

struct TopPrim
{
    uint32_t ref;
    float    lo[3], hi[3];
};

void child_box(const Node4& n, int c, float lo[3], float hi[3])
{
    lo[0] = n.lox[c];
    lo[1] = n.loy[c];
    lo[2] = n.loz[c];
    hi[0] = n.hix[c];
    hi[1] = n.hiy[c];
    hi[2] = n.hiz[c];
}

void child_box(const Node8& n, int c, float lo[3], float hi[3])
{
    for (int a = 0; a < 3; ++a)
    {
        lo[a] = clpt::node8_lo(n, a, c);
        hi[a] = clpt::node8_hi(n, a, c);
    }
}

template <typename NodeT, int W>
void rebraid_parts(const std::vector<NodeT>& nodes, uint32_t root, std::vector<TopPrim>& out)
{
    std::vector<TopPrim> parts;
    const auto           open = [&](uint32_t node) {
        const NodeT& n = nodes[node];
        for (int c = 0; c < W; ++c)
            if (n.child[c] != CLPT_CHILD_EMPTY)
            {
                TopPrim p;
                p.ref = n.child[c];
                child_box(n, c, p.lo, p.hi);
                parts.push_back(p);
            }
    };
    open(root);
    while (parts.size() < kRebraidParts)
    {
        size_t pick = parts.size();
        double best = -1.0;
        for (size_t k = 0; k < parts.size(); ++k)
        {
            if (parts[k].ref & CLPT_LEAF_BIT)
                continue;
            const double dx = (double)parts[k].hi[0] - parts[k].lo[0], dy = (double)parts[k].hi[1] - parts[k].lo[1],
                         dz = (double)parts[k].hi[2] - parts[k].lo[2];
            const double area = dx * dy + dy * dz + dz * dx;
            if (area > best)
            {
                best = area;
                pick = k;
            }
        }
        if (pick == parts.size())
            break; /* only leaves left */
        const uint32_t node = parts[pick].ref;
        parts[pick]         = parts.back();
        parts.pop_back();
        open(node);
    }
    out.insert(out.end(), parts.begin(), parts.end());
}
} // namespace

extern "C" {

const char* clhip_last_error(const clhip_ctx* ctx)
{
    if (ctx)
        return ctx->err.c_str();
    return g_last_error.c_str();
}

int clhip_create(uint32_t width, uint32_t height, uint32_t integrator_type, clhip_ctx** out)
{
    if (!out)
        return fail(nullptr, CLHIP_E_INVALID, "out is NULL");
    *out = nullptr;
    if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 30))
        return fail(nullptr, CLHIP_E_INVALID, "bad film size %ux%u", width, height);
    if (integrator_type != CLHIP_INTEGRATOR_PATH_TRACING && integrator_type != CLHIP_INTEGRATOR_DIRECT_LIGHTING)
        return fail(nullptr, CLHIP_E_INVALID, "unknown integrator type %u", integrator_type);
    int        ndev = 0;
    hipError_t e    = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return fail(nullptr, CLHIP_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    auto c   = std::make_unique<clhip_ctx>();
    c->width = width;
    c->height = height;
    c->itype  = integrator_type;
    e         = hipGetDevice(&c->device);
    if (e != hipSuccess)
        return fail(nullptr, CLHIP_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    int cus = 0;
    e       = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    if (e != hipSuccess || cus <= 0)
        cus = 256;
    c->num_cus = (uint32_t)cus;
    {
        /* persistent grids: every lane refills its own rays, so at most the
         * blocks that fit on the chip at once.  Measured on the 1080p scene
         * (MI355X, 64 spp in one batch): both traversals are fastest at full
         * residency; any hit at 2, 3, 4 and 6 blocks/CU gives 909, 977, 1015
         * and 1039 Msamples/s.  clhip_set_launch and the environment
         * variables override. */
        auto env_u = [](const char* name) {
            const char* v = std::getenv(name);
            return v ? (uint32_t)std::atoi(v) : 0u;
        };
        const uint32_t rt = clpt::resident_blocks_per_cu(0), rs = clpt::resident_blocks_per_cu(1);
        const uint32_t ot = env_u("CLPT_TRACE_BLOCKS_PER_CU"), os = env_u("CLPT_SHADOW_BLOCKS_PER_CU");
        c->def_trace  = ot ? ot : std::max(1u, rt);
        c->def_shadow = os ? os : std::max(1u, rs);
        c->blocks_guided = c->num_cus * std::max(1u, clpt::resident_blocks_per_cu(3));
        apply_launch(c.get());
    }
    e          = hipStreamCreate(&c->stream);
    if (e != hipSuccess)
        return fail(nullptr, CLHIP_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    /* the device address of clhip_trace_rays_device's scratch, looked up
     * once per device with that device current */
    e = clpt::noalloc_scratch(c->device, &c->noalloc_stack);
    if (e != hipSuccess)
    {
        (void)hipStreamDestroy(c->stream);
        return fail(nullptr, CLHIP_E_HIP, "scratch address on device %d: %s", c->device, hipGetErrorString(e));
    }
    c->cam.width  = width;
    c->cam.height = height;
    c->cam.npix   = width * height;
    *out         = c.release();
    return CLHIP_OK;
}

int clhip_destroy(clhip_ctx* ctx)
{
    if (!ctx)
        return CLHIP_OK;
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    recycle_events(ctx);
    recycle_batch_events(ctx);
    for (auto& t : ctx->train_timed)
    {
        ctx->event_pool.push_back(t.a);
        ctx->event_pool.push_back(t.b);
    }
    ctx->train_timed.clear();
    for (hipEvent_t e : ctx->event_pool)
        (void)hipEventDestroy(e);
    for (GraphSlot& g : ctx->graphs)
    {
        if (g.exec)
            (void)hipGraphExecDestroy(g.exec);
        if (g.graph)
            (void)hipGraphDestroy(g.graph);
    }
    for (StagingSlot& st : ctx->staging)
    {
        if (st.done)
            (void)hipEventDestroy(st.done);
        if (st.p)
            (void)hipHostFree(st.p);
    }
    if (ctx->ev_begin)
        (void)hipEventDestroy(ctx->ev_begin);
    if (ctx->ev_end)
        (void)hipEventDestroy(ctx->ev_end);
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CLHIP_OK;
}

/* Appends a texture and re-uploads the device table of all textures, which
 * material lookups index. */
static int add_texture(clhip_ctx* ctx, TexHost&& t, int32_t* texture_id)
{
    ctx->textures.push_back(std::move(t));
    *texture_id = (int32_t)ctx->textures.size() - 1;
    std::vector<TexRec> recs(ctx->textures.size());
    for (size_t i = 0; i < recs.size(); ++i)
    {
        recs[i].data = ctx->textures[i].buf->as<const float4>();
        recs[i].w    = ctx->textures[i].w;
        recs[i].h    = ctx->textures[i].h;
    }
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream)); /* the old table may be in use */
    CLHIP_CHECK(ctx, upload(ctx->d_texrec, recs));
    return CLHIP_OK;
}

int clhip_create_texture(clhip_ctx* ctx, const float* data, uint32_t width, uint32_t height, uint32_t channels,
                         int32_t* texture_id)
{
    if (!ctx || !data || !texture_id || width == 0 || height == 0 || (channels != 3 && channels != 4))
        return fail(ctx, CLHIP_E_INVALID, "clhip_create_texture: bad arguments");
    const size_t       n = (size_t)width * height;
    std::vector<float> rgba(4 * n);
    for (size_t i = 0; i < n; ++i)
    {
        rgba[4 * i]     = data[channels * i];
        rgba[4 * i + 1] = data[channels * i + 1];
        rgba[4 * i + 2] = data[channels * i + 2];
        rgba[4 * i + 3] = channels == 4 ? data[4 * i + 3] : 1.0f;
    }
    TexHost t;
    t.buf = std::make_unique<DevBuf>();
    t.w   = width;
    t.h   = height;
    CLHIP_CHECK(ctx, upload(*t.buf, rgba));
    return add_texture(ctx, std::move(t), texture_id);
}

int clhip_create_texture_u8(clhip_ctx* ctx, const uint8_t* data, uint32_t width, uint32_t height, uint32_t channels,
                            int srgb, int32_t* texture_id)
{
    if (!ctx || !data || !texture_id || width == 0 || height == 0 || (channels != 3 && channels != 4))
        return fail(ctx, CLHIP_E_INVALID, "clhip_create_texture_u8: bad arguments");
    if ((uint64_t)width * height > (1ull << 30))
        return fail(ctx, CLHIP_E_INVALID, "clhip_create_texture_u8: %ux%u texels, more than 2^30", width, height);
    const size_t         n = (size_t)width * height;
    std::vector<uint8_t> rgba(4 * n);
    for (size_t i = 0; i < n; ++i)
    {
        rgba[4 * i]     = data[channels * i];
        rgba[4 * i + 1] = data[channels * i + 1];
        rgba[4 * i + 2] = data[channels * i + 2];
        rgba[4 * i + 3] = channels == 4 ? data[4 * i + 3] : 255u;
    }
    Tmp raw;
    CLHIP_CHECK(ctx, raw.from_host(rgba.data(), rgba.size()));
    TexHost t;
    t.buf = std::make_unique<DevBuf>();
    t.w   = width;
    t.h   = height;
    CLHIP_CHECK(ctx, t.buf->alloc(n * sizeof(float4)));
    CLHIP_CHECK(ctx, clpt::launch_texture_u8(ctx->stream, raw.b.as<const uchar4>(), (uint32_t)n, srgb ? 1 : 0,
                                             t.buf->as<float4>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return add_texture(ctx, std::move(t), texture_id);
}

int clhip_set_bvh_width(clhip_ctx* ctx, uint32_t width)
{
    if (!ctx || (width != 4 && width != 8))
        return fail(ctx, CLHIP_E_INVALID, "clhip_set_bvh_width: width must be 4 or 8");
    if (width != ctx->bvh_width)
    {
        CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        ctx->bvh_width = width;
        ctx->have_blas = ctx->have_tlas = false; /* the trees must be built again */
        ctx->blas_nodes4.clear();
        ctx->blas_nodes8.clear();
        ctx->mesh_builds.clear();
    }
    return CLHIP_OK;
}

extern "C++" {
/* One mesh's tree built on its own: inner references count from its first
 * node, leaves from its first triangle; assemble_blas relocates them into
 * the shared arrays.  Kept per mesh so that an edit rebuilds one tree. */
static int build_mesh(clhip_ctx* ctx, const clhip_mesh& me, uint32_t m, MeshBuild& out)
{
    if (!me.vertices || (!me.indices && me.num_triangles))
        return fail(ctx, CLHIP_E_INVALID, "mesh %u: missing vertices or indices", m);
    const uint32_t nt = me.num_triangles, nv = me.num_vertices;
    if (nt >= CLPT_MAX_TRIS)
        return fail(ctx, CLHIP_E_UNSUPPORTED, "mesh %u: %u triangles exceed the 2^27 limit", m, nt);
    for (size_t i = 0; i < 3 * (size_t)nt; ++i)
        if (me.indices[i] >= nv)
            return fail(ctx, CLHIP_E_INVALID, "mesh %u: index %u out of range (%u vertices)", m, me.indices[i], nv);
    std::vector<float> lo(3 * (size_t)nt), hi(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k)
        {
            float a = me.vertices[3 * (size_t)me.indices[3 * (size_t)t] + k];
            float b = me.vertices[3 * (size_t)me.indices[3 * (size_t)t + 1] + k];
            float c = me.vertices[3 * (size_t)me.indices[3 * (size_t)t + 2] + k];
            lo[3 * (size_t)t + k] = std::min(a, std::min(b, c));
            hi[3 * (size_t)t + k] = std::max(a, std::max(b, c));
        }
    MeshBuild             mb;
    std::vector<uint32_t> order;
    const auto            leaf = [](uint32_t first, uint32_t count) { return CLPT_TRI_LEAF(first, count); };
    const clpt::BvhResult r    = ctx->bvh_width == 8
                                     ? clpt::build_bvh8(lo.data(), hi.data(), nt, CLPT_MAX_LEAF_TRIS, mb.nodes8, order, leaf)
                                     : clpt::build_bvh4(lo.data(), hi.data(), nt, CLPT_MAX_LEAF_TRIS, mb.nodes4, order, leaf);
    mb.tris.resize(nt);
    for (uint32_t i = 0; i < nt; ++i)
    {
        const uint32_t p  = order[i];
        const float*   v0 = me.vertices + 3 * (size_t)me.indices[3 * (size_t)p];
        const float*   v1 = me.vertices + 3 * (size_t)me.indices[3 * (size_t)p + 1];
        const float*   v2 = me.vertices + 3 * (size_t)me.indices[3 * (size_t)p + 2];
        float          pf;
        std::memcpy(&pf, &p, 4);
        mb.tris[i].v0p = make_float4(v0[0], v0[1], v0[2], pf);
        mb.tris[i].v1  = make_float4(v1[0], v1[1], v1[2], 0.0f);
        mb.tris[i].v2  = make_float4(v2[0], v2[1], v2[2], 0.0f);
    }
    mb.info.ntris  = nt;
    mb.info.nverts = nv;
    mb.info.root   = r.root;
    mb.info.depth  = r.max_depth;
    std::memcpy(mb.info.lo, r.lo, sizeof(mb.info.lo));
    std::memcpy(mb.info.hi, r.hi, sizeof(mb.info.hi));
    mb.flags = (me.normals ? MESH_HAS_NORMALS : 0u) | (me.tangents ? MESH_HAS_TANGENTS : 0u) |
        (me.uvs ? MESH_HAS_UVS : 0u);
    mb.pos.assign(me.vertices, me.vertices + 3 * (size_t)nv);
    if (me.normals)
        mb.nrm.assign(me.normals, me.normals + 3 * (size_t)nv);
    if (me.tangents)
        mb.tan.assign(me.tangents, me.tangents + 3 * (size_t)nv);
    if (me.uvs)
        mb.uv.assign(me.uvs, me.uvs + 2 * (size_t)nv);
    mb.idx.assign(me.indices, me.indices + 3 * (size_t)nt);
    out = std::move(mb);
    ++ctx->blas_builds;
    return CLHIP_OK;
}

/* A child reference of a mesh's own tree, moved to where its nodes and
 * triangles land in the shared arrays. */
static inline uint32_t relocate_ref(uint32_t c, uint32_t node_off, uint32_t tri_off)
{
    if (c == CLPT_CHILD_EMPTY)
        return c;
    if (c & CLPT_LEAF_BIT)
        return (c & ~(CLPT_MAX_TRIS - 1u)) | ((c & (CLPT_MAX_TRIS - 1u)) + tri_off);
    return c + node_off;
}

template <typename NodeT>
static void append_nodes(std::vector<NodeT>& dst, const std::vector<NodeT>& src, uint32_t tri_off)
{
    const uint32_t node_off = (uint32_t)dst.size();
    for (NodeT nd : src)
    {
        for (uint32_t& c : nd.child)
            c = relocate_ref(c, node_off, tri_off);
        dst.push_back(nd);
    }
}

/* The shared node, triangle and attribute arrays from the per-mesh builds,
 * uploaded; the TLAS must be built again (the roots moved). */
static int assemble_blas(clhip_ctx* ctx)
{
    const std::vector<MeshBuild>& mbs = ctx->mesh_builds;
    uint64_t                      total_tris = 0;
    bool                          any_n = false, any_t = false, any_uv = false;
    for (const MeshBuild& mb : mbs)
    {
        total_tris += mb.info.ntris;
        any_n |= (mb.flags & MESH_HAS_NORMALS) != 0u;
        any_t |= (mb.flags & MESH_HAS_TANGENTS) != 0u;
        any_uv |= (mb.flags & MESH_HAS_UVS) != 0u;
    }
    if (total_tris >= CLPT_MAX_TRIS)
        return fail(ctx, CLHIP_E_UNSUPPORTED, "%llu triangles exceed the 2^27 limit", (unsigned long long)total_tris);
    std::vector<Node4>    nodes4;
    std::vector<Node8>    nodes8;
    std::vector<TriRec>   tris;
    std::vector<MeshInfo> infos(mbs.size());
    std::vector<MeshRec>  recs(mbs.size());
    std::vector<float>    pos, nrm, tan, uv;
    std::vector<uint32_t> idx;
    tris.reserve(total_tris);
    for (size_t m = 0; m < mbs.size(); ++m)
    {
        const MeshBuild& mb       = mbs[m];
        const uint32_t   tri_base = (uint32_t)tris.size();
        const uint32_t   node_off = (uint32_t)(ctx->bvh_width == 8 ? nodes8.size() : nodes4.size());
        if (ctx->bvh_width == 8)
            append_nodes(nodes8, mb.nodes8, tri_base);
        else
            append_nodes(nodes4, mb.nodes4, tri_base);
        tris.insert(tris.end(), mb.tris.begin(), mb.tris.end());
        infos[m]          = mb.info;
        infos[m].tri_base = tri_base;
        infos[m].root     = relocate_ref(mb.info.root, node_off, tri_base);
        const uint32_t nv = mb.info.nverts;
        recs[m].idx_off   = (uint32_t)(idx.size() / 3);
        recs[m].vtx_off   = (uint32_t)(pos.size() / 3);
        recs[m].ntris     = mb.info.ntris;
        recs[m].flags     = mb.flags;
        pos.insert(pos.end(), mb.pos.begin(), mb.pos.end());
        if (any_n)
        {
            if (mb.flags & MESH_HAS_NORMALS)
                nrm.insert(nrm.end(), mb.nrm.begin(), mb.nrm.end());
            else
                nrm.resize(nrm.size() + 3 * (size_t)nv, 0.0f);
        }
        if (any_t)
        {
            if (mb.flags & MESH_HAS_TANGENTS)
                tan.insert(tan.end(), mb.tan.begin(), mb.tan.end());
            else
                tan.resize(tan.size() + 3 * (size_t)nv, 0.0f);
        }
        if (any_uv)
        {
            if (mb.flags & MESH_HAS_UVS)
                uv.insert(uv.end(), mb.uv.begin(), mb.uv.end());
            else
                uv.resize(uv.size() + 2 * (size_t)nv, 0.0f);
        }
        idx.insert(idx.end(), mb.idx.begin(), mb.idx.end());
    }
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream)); /* the old arrays may be in use */
    CLHIP_CHECK(ctx, upload(ctx->d_tris, tris));
    ctx->tri_bytes = tris.size() * sizeof(TriRec);
    CLHIP_CHECK(ctx, upload(ctx->d_meshes, recs));
    CLHIP_CHECK(ctx, upload(ctx->d_pos, pos));
    CLHIP_CHECK(ctx, upload(ctx->d_nrm, nrm));
    CLHIP_CHECK(ctx, upload(ctx->d_tan, tan));
    CLHIP_CHECK(ctx, upload(ctx->d_uv, uv));
    CLHIP_CHECK(ctx, upload(ctx->d_idx, idx));
    ctx->blas_nodes4 = std::move(nodes4);
    ctx->blas_nodes8 = std::move(nodes8);
    ctx->meshes      = std::move(infos);
    ctx->mesh_recs   = std::move(recs);
    ctx->have_blas   = true;
    ctx->have_tlas   = false; /* roots moved: the TLAS must be rebuilt */
    return CLHIP_OK;
}

} // extern "C++"

int clhip_build_blas(clhip_ctx* ctx, const clhip_mesh* meshes, uint32_t num_meshes)
{
    if (!ctx || (!meshes && num_meshes))
        return fail(ctx, CLHIP_E_INVALID, "clhip_build_blas: bad arguments");
    std::vector<MeshBuild> builds(num_meshes);
    for (uint32_t m = 0; m < num_meshes; ++m)
    {
        const int rc = build_mesh(ctx, meshes[m], m, builds[m]);
        if (rc)
            return rc;
    }
    ctx->mesh_builds = std::move(builds);
    return assemble_blas(ctx);
}

int clhip_build_blas_subset(clhip_ctx* ctx, const clhip_mesh* meshes, const uint32_t* mesh_index, uint32_t n)
{
    if (!ctx || (n && (!meshes || !mesh_index)))
        return fail(ctx, CLHIP_E_INVALID, "clhip_build_blas_subset: bad arguments");
    std::vector<MeshBuild> builds(n);
    size_t                 count = ctx->mesh_builds.size();
    for (uint32_t k = 0; k < n; ++k)
    {
        if (mesh_index[k] > count)
            return fail(ctx, CLHIP_E_INVALID, "mesh index %u: %zu meshes (and the next) exist", mesh_index[k], count);
        count = std::max<size_t>(count, (size_t)mesh_index[k] + 1);
        const int rc = build_mesh(ctx, meshes[k], mesh_index[k], builds[k]);
        if (rc)
            return rc;
    }
    if (count > ctx->mesh_builds.size())
        ctx->mesh_builds.resize(count);
    for (uint32_t k = 0; k < n; ++k)
        ctx->mesh_builds[mesh_index[k]] = std::move(builds[k]);
    return assemble_blas(ctx);
}

int clhip_build_tlas(clhip_ctx* ctx, const uint32_t* mesh_of_instance, const uint32_t* instance_ids,
                     const float* xforms, uint32_t num_instances)
{
    if (!ctx || (!mesh_of_instance && num_instances))
        return fail(ctx, CLHIP_E_INVALID, "clhip_build_tlas: bad arguments");
    if (!ctx->have_blas)
        return fail(ctx, CLHIP_E_STATE, "clhip_build_tlas before clhip_build_blas");
    if (num_instances >= CLPT_MAX_INSTANCES)
        return fail(ctx, CLHIP_E_UNSUPPORTED, "too many instances");
    std::vector<InstanceRec> inst(num_instances);
    std::vector<uint32_t>    ent(num_instances);
    std::vector<float>       lo(3 * (size_t)num_instances), hi(3 * (size_t)num_instances);
    uint32_t                 max_blas_depth = 0;
    static const float       id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    /* the re-braided instance: the most triangles among those whose
     * transform and inverse are the identity (by value: the computed
     * inverse holds -0 translations) */
    uint32_t                 flat = 0xffffffffu;
    for (uint32_t i = 0; i < num_instances; ++i)
    {
        const uint32_t m = mesh_of_instance[i];
        if (m >= ctx->meshes.size())
            return fail(ctx, CLHIP_E_INVALID, "instance %u: mesh %u of %zu", i, m, ctx->meshes.size());
        InstanceRec& r = inst[i];
        std::memcpy(r.xf, xforms ? xforms + 12 * (size_t)i : id, sizeof(r.xf));
        inverse3x4(r.xf, r.inv);
        for (int k = 0; k < 12; ++k)
            if (!std::isfinite(r.inv[k]))
                return fail(ctx, CLHIP_E_INVALID, "instance %u: singular transform", i);
        r.blas_root = ctx->meshes[m].root;
        r.mesh      = m;
        r.material  = 0;
        r.entity    = instance_ids ? instance_ids[i] : i;
        ent[i]      = r.entity;
        max_blas_depth = std::max(max_blas_depth, ctx->meshes[m].depth);
        const MeshInfo& mi = ctx->meshes[m];
        bool            identity = true;
        for (int k = 0; k < 12; ++k)
            identity = identity && r.xf[k] == id[k] && r.inv[k] == id[k];
        if (mi.ntris > 0 && identity && (flat == 0xffffffffu || mi.ntris > ctx->meshes[inst[flat].mesh].ntris))
            flat = i;
        for (int k = 0; k < 3; ++k)
        {
            lo[3 * (size_t)i + k] = INFINITY;
            hi[3 * (size_t)i + k] = -INFINITY;
        }
        if (mi.ntris == 0)
        {
            /* an empty mesh gets an empty box far from everything */
            for (int k = 0; k < 3; ++k)
                lo[3 * (size_t)i + k] = hi[3 * (size_t)i + k] = 0.0f;
            continue;
        }
        for (int c = 0; c < 8; ++c)
        {
            const double p[3] = {(c & 1) ? mi.hi[0] : mi.lo[0], (c & 2) ? mi.hi[1] : mi.lo[1],
                                 (c & 4) ? mi.hi[2] : mi.lo[2]};
            for (int k = 0; k < 3; ++k)
            {
                const double q = (double)r.xf[4 * k] * p[0] + (double)r.xf[4 * k + 1] * p[1] +
                    (double)r.xf[4 * k + 2] * p[2] + (double)r.xf[4 * k + 3];
                lo[3 * (size_t)i + k] = std::min(lo[3 * (size_t)i + k], std::nextafter((float)q, -INFINITY));
                hi[3 * (size_t)i + k] = std::max(hi[3 * (size_t)i + k], std::nextafter((float)q, INFINITY));
            }
        }
    }
    /* the top tree's primitives: every other instance's box, and the
     * re-braided instance's parts */
    std::vector<TopPrim> prims;
    prims.reserve(num_instances + kRebraidParts + 8);
    for (uint32_t i = 0; i < num_instances; ++i)
        if (i != flat)
        {
            TopPrim p;
            p.ref = CLPT_INST_LEAF(i);
            for (int k = 0; k < 3; ++k)
            {
                p.lo[k] = lo[3 * (size_t)i + k];
                p.hi[k] = hi[3 * (size_t)i + k];
            }
            prims.push_back(p);
        }
    if (flat != 0xffffffffu)
    {
        if (ctx->bvh_width == 8)
            rebraid_parts<Node8, 8>(ctx->blas_nodes8, inst[flat].blas_root, prims);
        else
            rebraid_parts<Node4, 4>(ctx->blas_nodes4, inst[flat].blas_root, prims);
    }
    const uint32_t     np = (uint32_t)prims.size();
    std::vector<float> plo(3 * (size_t)np), phi(3 * (size_t)np);
    for (uint32_t p = 0; p < np; ++p)
        for (int k = 0; k < 3; ++k)
        {
            plo[3 * (size_t)p + k] = prims[p].lo[k];
            phi[3 * (size_t)p + k] = prims[p].hi[k];
        }
    std::vector<uint32_t> order;
    const auto            leaf = [&order, &prims](uint32_t first, uint32_t) { return prims[order[first]].ref; };
    std::vector<Node4>    nodes4;
    std::vector<Node8>    nodes8;
    clpt::BvhResult       r;
    if (ctx->bvh_width == 8)
    {
        nodes8 = ctx->blas_nodes8;
        r      = clpt::build_bvh8(plo.data(), phi.data(), np, 1u, nodes8, order, leaf);
    }
    else
    {
        nodes4 = ctx->blas_nodes4;
        r      = clpt::build_bvh4(plo.data(), phi.data(), np, 1u, nodes4, order, leaf);
    }
    /* stack bound: each W-wide level leaves at most W - 1 entries behind,
     * plus the popped entry and a margin */
    const uint32_t need = (ctx->bvh_width - 1u) * (r.max_depth + max_blas_depth) + 4u;
    if (need > 128u)
        return fail(ctx, CLHIP_E_UNSUPPORTED, "tree too deep for the traversal stack (TLAS %u + BLAS %u levels)",
                    r.max_depth, max_blas_depth);
    /* traversal addresses nodes by 32-bit byte offsets */
    const size_t nbytes = ctx->bvh_width == 8 ? nodes8.size() * sizeof(Node8) : nodes4.size() * sizeof(Node4);
    if (nbytes > 0xffffffffull)
        return fail(ctx, CLHIP_E_UNSUPPORTED, "node array of %zu bytes exceeds 4 GiB", nbytes);
    if (ctx->bvh_width == 8)
        CLHIP_CHECK(ctx, upload(ctx->d_nodes, nodes8));
    else
        CLHIP_CHECK(ctx, upload(ctx->d_nodes, nodes4));
    ctx->node_bytes = ctx->bvh_width == 8 ? nodes8.size() * sizeof(Node8) : nodes4.size() * sizeof(Node4);
    /* the records go up now, so traversal alone (clhip_trace_rays_device)
     * needs no further upload; sync_instances fills in the materials */
    CLHIP_CHECK(ctx, upload(ctx->d_inst, inst));
    ctx->inst        = std::move(inst);
    ctx->inst_entity = std::move(ent);
    ctx->tlas_root   = r.root;
    ctx->tlas_depth  = r.max_depth;
    ctx->world_inst  = flat;
    /* the scene box (it normalises the learned proposal's positions) is the
     * union of the instance boxes, whatever parts the top tree holds */
    for (int k = 0; k < 3; ++k)
    {
        float wl = INFINITY, wh = -INFINITY;
        for (uint32_t i = 0; i < num_instances; ++i)
        {
            wl = std::min(wl, lo[3 * (size_t)i + k]);
            wh = std::max(wh, hi[3 * (size_t)i + k]);
        }
        ctx->world_lo[k] = num_instances ? wl : 0.0f;
        ctx->world_hi[k] = num_instances ? wh : 1.0f;
    }
    ctx->have_tlas   = true;
    ctx->inst_dirty  = true;
    return CLHIP_OK;
}

int clhip_build_materials(clhip_ctx* ctx, const clhip_material* materials, uint32_t n)
{
    if (!ctx || (!materials && n))
        return fail(ctx, CLHIP_E_INVALID, "clhip_build_materials: bad arguments");
    std::vector<MaterialRec> recs(n);
    for (uint32_t i = 0; i < n; ++i)
    {
        const clhip_material& m = materials[i];
        if (m.type > CLHIP_MTL_GLASS)
            return fail(ctx, CLHIP_E_INVALID, "material %u: unknown type %u", i, m.type);
        if (m.type == CLHIP_MTL_GLASS && m.interior_ior == 1.0f)
            return fail(ctx, CLHIP_E_INVALID, "material %u: glass IOR must differ from 1 (roughdielectric.h:28)", i);
        MaterialRec& r = recs[i];
        std::memset(&r, 0, sizeof(r));
        r.type = m.type;
        for (int k = 0; k < 3; ++k)
        {
            r.r[k]   = m.reflectance[k];
            r.eta[k] = m.eta[k];
            r.k[k]   = m.k[k];
        }
        const float rough = m.roughness;
        r.alpha           = std::fmax(rough * rough, 1e-4f); /* metal.h:52, microfacet.h:27 */
        r.ior             = m.interior_ior;
        const uint32_t nt = (uint32_t)ctx->textures.size();
        if (m.reflectance_tex > nt || m.roughness_tex > nt || m.normal_tex > nt)
            return fail(ctx, CLHIP_E_INVALID, "material %u: texture id beyond the %u created", i, nt);
        if (m.reflectance_tex && m.type == CLHIP_MTL_GLASS)
            return fail(ctx, CLHIP_E_INVALID, "material %u: glass has no reflectance texture (glass.h:61-67)", i);
        if (m.roughness_tex && m.type != CLHIP_MTL_METAL)
            return fail(ctx, CLHIP_E_INVALID, "material %u: only metal takes a roughness texture (metal.h:44-48)", i);
        r.tex_r     = m.reflectance_tex;
        r.tex_rough = m.roughness_tex;
        r.tex_n     = m.normal_tex;
    }
    CLHIP_CHECK(ctx, upload(ctx->d_mtl, recs));
    ctx->mtl        = std::move(recs);
    ctx->inst_dirty = true;
    return CLHIP_OK;
}

int clhip_build_entities(clhip_ctx* ctx, const uint32_t* material_index, uint32_t n)
{
    if (!ctx || (!material_index && n))
        return fail(ctx, CLHIP_E_INVALID, "clhip_build_entities: bad arguments");
    ctx->entity_material.assign(material_index, material_index + n);
    ctx->inst_dirty = true;
    return CLHIP_OK;
}

int clhip_build_emitters(clhip_ctx* ctx, const clhip_emitter* emitters, uint32_t n, int32_t dome_index)
{
    if (!ctx || (!emitters && n))
        return fail(ctx, CLHIP_E_INVALID, "clhip_build_emitters: bad arguments");
    if (dome_index >= (int32_t)n)
        return fail(ctx, CLHIP_E_INVALID, "dome index %d out of %u emitters", dome_index, n);
    std::vector<EmitterRec> recs(n);
    int32_t                 dome_tex = -1;
    for (uint32_t i = 0; i < n; ++i)
    {
        const clhip_emitter& e = emitters[i];
        EmitterRec&          r = recs[i];
        std::memset(&r, 0, sizeof(r));
        r.type = e.type;
        for (int k = 0; k < 3; ++k)
        {
            r.intensity[k] = e.intensity[k];
            r.dir[k]       = e.direction[k];
        }
        if (e.type == CLHIP_EMITTER_DIRECTIONAL)
        {
            r.tan_r = (float)std::tan((double)e.angular_radius);
            if (!(r.tan_r > 0.0f))
                return fail(ctx, CLHIP_E_INVALID, "emitter %u: angular radius must be in (0, pi/2)", i);
        }
        else if (e.type == CLHIP_EMITTER_HDRIDOME)
        {
            if (e.texture < 0 || e.texture >= (int32_t)ctx->textures.size())
                return fail(ctx, CLHIP_E_INVALID, "emitter %u: bad texture id %d", i, e.texture);
            if ((int32_t)i != dome_index)
                return fail(ctx, CLHIP_E_INVALID, "emitter %u is a dome but dome_index is %d", i, dome_index);
            dome_tex = e.texture;
        }
        else
            return fail(ctx, CLHIP_E_INVALID, "emitter %u: unknown type %u", i, e.type);
    }
    if (dome_index >= 0 && dome_tex < 0)
        return fail(ctx, CLHIP_E_INVALID, "dome_index %d is not an HDRI dome", dome_index);
    CLHIP_CHECK(ctx, upload(ctx->d_emitters, recs));
    ctx->emitters   = std::move(recs);
    ctx->dome_index = dome_index;
    ctx->dome_tex   = dome_tex;
    if (dome_tex >= 0)
    {
        const TexHost& t = ctx->textures[dome_tex];
        ctx->dome_w      = t.w;
        ctx->dome_h      = t.h;
        const size_t W   = (size_t)t.w + 1;
        CLHIP_CHECK(ctx, ctx->d_pcond.alloc(W * t.h * sizeof(float)));
        CLHIP_CHECK(ctx, ctx->d_cdfcond.alloc(W * t.h * sizeof(float)));
        CLHIP_CHECK(ctx, ctx->d_pv.alloc(((size_t)t.h + 1) * sizeof(float)));
        CLHIP_CHECK(ctx, ctx->d_cdfv.alloc(((size_t)t.h + 1) * sizeof(float)));
        CLHIP_CHECK(ctx, ctx->d_rowint.alloc((size_t)t.h * sizeof(double)));
        CLHIP_CHECK(ctx, ctx->d_cutv.alloc((CLPT_CUTS + 1u) * sizeof(uint32_t)));
        CLHIP_CHECK(ctx, ctx->d_cutcond.alloc((size_t)t.h * (CLPT_CUTS + 1u) * sizeof(uint32_t)));
        CLHIP_CHECK(ctx, clpt::launch_dome_tables(ctx->stream, t.buf->as<const float4>(), t.w, t.h,
                                                  ctx->d_pcond.as<float>(), ctx->d_cdfcond.as<float>(),
                                                  ctx->d_pv.as<float>(), ctx->d_cdfv.as<float>(),
                                                  ctx->d_rowint.as<double>(), ctx->d_cutv.as<uint32_t>(),
                                                  ctx->d_cutcond.as<uint32_t>()));
        CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return CLHIP_OK;
}

int clhip_set_camera(clhip_ctx* ctx, const float pos[3], const float d00[3], const float ddu[3], const float ddv[3])
{
    if (!ctx || !pos || !d00 || !ddu || !ddv)
        return fail(ctx, CLHIP_E_INVALID, "clhip_set_camera: bad arguments");
    /* an unchanged camera keeps the progressive film (path.h:73-80) */
    bool same = ctx->have_camera;
    for (int k = 0; k < 3 && same; ++k)
        same = ctx->cam.pos[k] == pos[k] && ctx->cam.d00[k] == d00[k] && ctx->cam.ddu[k] == ddu[k] &&
               ctx->cam.ddv[k] == ddv[k];
    if (same)
        return CLHIP_OK;
    for (int k = 0; k < 3; ++k)
    {
        ctx->cam.pos[k] = pos[k];
        ctx->cam.d00[k] = d00[k];
        ctx->cam.ddu[k] = ddu[k];
        ctx->cam.ddv[k] = ddv[k];
    }
    ctx->have_camera = true;
    return clhip_reset(ctx);
}

int clhip_set_params(clhip_ctx* ctx, const clhip_params* p)
{
    if (!ctx || !p)
        return fail(ctx, CLHIP_E_INVALID, "clhip_set_params: bad arguments");
    if (p->max_depth < 1 || p->max_depth > CLHIP_MAX_DEPTH)
        return fail(ctx, CLHIP_E_INVALID, "max_depth %u outside [1, %d]", p->max_depth, CLHIP_MAX_DEPTH);
    if (p->view >= (1u << 12))
        return fail(ctx, CLHIP_E_INVALID, "view %u outside [0, 4096)", p->view);
    /* the batch size does not change any sample, so only the estimator's
     * parameters restart the accumulation */
    const bool same = p->max_depth == ctx->params.max_depth && p->rr_depth == ctx->params.rr_depth &&
                      p->view == ctx->params.view;
    ctx->params = *p;
    return same ? CLHIP_OK : clhip_reset(ctx);
}

int clhip_reset(clhip_ctx* ctx)
{
    if (!ctx)
        return fail(nullptr, CLHIP_E_INVALID, "ctx is NULL");
    ctx->spp = 0;
    if (ctx->d_accum.p)
        CLHIP_CHECK(ctx, hipMemsetAsync(ctx->d_accum.p, 0, ctx->d_accum.bytes, ctx->stream));
    return CLHIP_OK;
}

int clhip_resize(clhip_ctx* ctx, uint32_t width, uint32_t height)
{
    if (!ctx || width == 0 || height == 0 || (uint64_t)width * height > (1ull << 30))
        return fail(ctx, CLHIP_E_INVALID, "clhip_resize: bad size");
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->width     = width;
    ctx->height    = height;
    ctx->cam.width  = width;
    ctx->cam.height = height;
    ctx->cam.npix   = width * height;
    ctx->d_accum.release();
    ctx->d_out.release();
    ctx->spp = 0;
    return CLHIP_OK;
}

int clhip_get_film_size(const clhip_ctx* ctx, uint32_t* width, uint32_t* height, uint32_t* spp)
{
    if (!ctx)
        return fail(nullptr, CLHIP_E_INVALID, "ctx is NULL");
    if (width)
        *width = ctx->width;
    if (height)
        *height = ctx->height;
    if (spp)
        *spp = ctx->spp;
    return CLHIP_OK;
}

int clhip_set_timing(clhip_ctx* ctx, int timing)
{
    if (!ctx)
        return fail(nullptr, CLHIP_E_INVALID, "ctx is NULL");
    ctx->timing = timing != 0;
    return CLHIP_OK;
}

int clhip_render(clhip_ctx* ctx, uint32_t spp)
{
    if (!ctx)
        return fail(nullptr, CLHIP_E_INVALID, "ctx is NULL");
    if (!ctx->have_camera)
        return fail(ctx, CLHIP_E_STATE, "no camera: call clhip_set_camera first");
    int rc = sync_instances(ctx);
    if (rc)
        return rc;
    if (spp == 0)
        return CLHIP_OK;
    if ((uint64_t)ctx->spp + spp > (1u << 20))
        return fail(ctx, CLHIP_E_INVALID, "%u + %u samples exceed the 2^20 sample indices of a view", ctx->spp, spp);
    const uint64_t npix = (uint64_t)ctx->width * ctx->height;
    uint32_t       spb  = ctx->params.samples_per_batch;
    if (spb == 0) /* whole samples of the film up to the batch's path budget */
        spb = (uint32_t)std::max<uint64_t>(1, default_paths_per_batch(ctx) / npix);
    spb = std::min(spb, spp);
    while ((uint64_t)spb * npix > 0xffffffffull / 2)
        spb = std::max(1u, spb / 2);
    const bool fresh = ctx->d_accum.p == nullptr;
    rc               = ensure_render_buffers(ctx, (size_t)spb * npix);
    if (rc)
        return rc;
    if (fresh || ctx->spp == 0)
        CLHIP_CHECK(ctx, hipMemsetAsync(ctx->d_accum.p, 0, ctx->d_accum.bytes, ctx->stream));

    /* one view; batch b holds samples [spp_done + b*spb, ...) of every pixel */
    ViewRec v     = ctx->cam;
    v.max_depth   = ctx->params.max_depth;
    v.rr_depth    = ctx->params.rr_depth;
    direct_lighting_depth(ctx, v);
    v.lane        = ctx->params.view;
    v.film_offset = 0;
    std::vector<ViewRec>    views{v};
    std::vector<BatchRec>   batches;
    std::vector<BatchEntry> entries;
    for (uint32_t done = 0; done < spp; done += spb)
    {
        const uint32_t nb = std::min(spb, spp - done);
        batches.push_back(BatchRec{(uint32_t)entries.size(), 1u, (uint32_t)(nb * npix), 0u});
        entries.push_back(BatchEntry{0u, 0u, ctx->spp + done, 0u});
    }
    rc = upload_tables(ctx, views, batches, entries);
    if (rc)
        return rc;
    rc = run_batches(ctx, (uint32_t)batches.size(), v.max_depth, ctx->d_accum.as<double>(), false, true);
    if (rc)
        return rc;
    ctx->spp += spp;
    return CLHIP_OK;
}

int clhip_set_launch(clhip_ctx* ctx, const clhip_launch* l)
{
    if (!ctx || !l)
        return fail(ctx, CLHIP_E_INVALID, "clhip_set_launch: bad arguments");
    const uint32_t rt = clpt::resident_blocks_per_cu(0), rs = clpt::resident_blocks_per_cu(1);
    if (l->trace_blocks_per_cu > std::max(1u, rt) || l->shadow_blocks_per_cu > std::max(1u, rs))
        return fail(ctx, CLHIP_E_INVALID,
                    "persistent grids hold at most %u (closest hit) and %u (any hit) resident blocks per CU", rt, rs);
    if (l->shade_blocks_per_cu > 64 || l->graph > 2 || l->order > 2)
        return fail(ctx, CLHIP_E_INVALID, "clhip_set_launch: field out of range");
    if (l->shade_block != 0 && l->shade_block != 64 && l->shade_block != 128 && l->shade_block != 256)
        return fail(ctx, CLHIP_E_INVALID, "shade_block %u: 64, 128 or 256 threads (0 = 256)", l->shade_block);
    if (l->tile > 32 || (l->tile & (l->tile - 1u)) != 0u)
        return fail(ctx, CLHIP_E_INVALID, "tile %u: 1 (rows) or a square tile side of 2, 4, 8, 16 or 32", l->tile);
    if (l->paths_per_batch && l->paths_per_batch < 4096)
        return fail(ctx, CLHIP_E_INVALID, "paths_per_batch %u below 4096", l->paths_per_batch);
    ctx->launch = *l;
    apply_launch(ctx);
    return CLHIP_OK;
}

int clhip_get_launch(const clhip_ctx* ctx, clhip_launch* out)
{
    if (!ctx || !out)
        return fail(nullptr, CLHIP_E_INVALID, "clhip_get_launch: bad arguments");
    out->trace_blocks_per_cu  = ctx->blocks_trace / ctx->num_cus;
    out->shadow_blocks_per_cu = ctx->blocks_shadow / ctx->num_cus;
    out->shade_blocks_per_cu  = ctx->blocks_shade / ctx->num_cus;
    out->paths_per_batch      = default_paths_per_batch(ctx);
    out->graph                = use_graph(ctx) ? 1u : 2u;
    out->order                = ctx->launch.order ? ctx->launch.order : 1u;
    out->shade_block          = ctx->launch.shade_block ? ctx->launch.shade_block : 256u;
    out->tile                 = launch_tile(ctx);
    return CLHIP_OK;
}

int clhip_render_views(clhip_ctx* ctx, const clhip_view* views, uint32_t nviews)
{
    if (!ctx || (!views && nviews))
        return fail(ctx, CLHIP_E_INVALID, "clhip_render_views: bad arguments");
    int rc = sync_instances(ctx);
    if (rc)
        return rc;
    uint64_t film = 0;
    uint32_t maxd = 1, maxpix = 1;
    for (uint32_t i = 0; i < nviews; ++i)
    {
        const clhip_view& w = views[i];
        if (w.width == 0 || w.height == 0 || (uint64_t)w.width * w.height > (1u << 26))
            return fail(ctx, CLHIP_E_INVALID, "view %u: bad film size %ux%u", i, w.width, w.height);
        if (w.spp == 0 || w.spp > (1u << 20))
            return fail(ctx, CLHIP_E_INVALID, "view %u: spp %u outside [1, 2^20]", i, w.spp);
        if (w.max_depth < 1 || w.max_depth > CLHIP_MAX_DEPTH)
            return fail(ctx, CLHIP_E_INVALID, "view %u: max_depth %u outside [1, %d]", i, w.max_depth, CLHIP_MAX_DEPTH);
        if (w.lane >= (1u << 12))
            return fail(ctx, CLHIP_E_INVALID, "view %u: lane %u outside [0, 4096)", i, w.lane);
        film += (uint64_t)w.width * w.height;
        maxd   = std::max(maxd, ctx->itype == CLHIP_INTEGRATOR_DIRECT_LIGHTING ? 2u : w.max_depth);
        maxpix = std::max(maxpix, w.width * w.height);
    }
    if (film > (1ull << 30))
        return fail(ctx, CLHIP_E_INVALID, "views hold %llu pixels, more than 2^30", (unsigned long long)film);

    ctx->vviews.assign(nviews, ViewRec{});
    ctx->vspp.assign(nviews, 0u);
    ctx->vlast_batch.assign(nviews, 0u);
    uint32_t off = 0;
    for (uint32_t i = 0; i < nviews; ++i)
    {
        const clhip_view& w = views[i];
        ViewRec&          v = ctx->vviews[i];
        v.width       = w.width;
        v.height      = w.height;
        v.npix        = w.width * w.height;
        v.max_depth   = w.max_depth;
        v.rr_depth    = w.rr_depth;
        v.lane        = w.lane;
        v.film_offset = off;
        direct_lighting_depth(ctx, v);
        for (int k = 0; k < 3; ++k)
        {
            v.pos[k] = w.pos[k];
            v.d00[k] = w.d00[k];
            v.ddu[k] = w.ddu[k];
            v.ddv[k] = w.ddv[k];
        }
        ctx->vspp[i] = w.spp;
        off += v.npix;
    }

    /* batches of whole samples: view i contributes samples [s0, s1) of all
     * its pixels to a batch; a batch holds at most `cap` paths */
    std::vector<uint32_t> order(nviews);
    for (uint32_t i = 0; i < nviews; ++i)
        order[i] = i;
    if (ctx->launch.order == 2u)
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            const uint64_t wx = (uint64_t)ctx->vviews[x].npix * ctx->vspp[x] * ctx->vviews[x].max_depth;
            const uint64_t wy = (uint64_t)ctx->vviews[y].npix * ctx->vspp[y] * ctx->vviews[y].max_depth;
            return wx < wy;
        });
    const uint32_t cap = std::max(maxpix, default_paths_per_batch(ctx));
    std::vector<BatchRec>   batches;
    std::vector<BatchEntry> entries;
    uint32_t                fill = cap;
    for (uint32_t i : order)
    {
        const uint32_t npix = ctx->vviews[i].npix;
        for (uint32_t s0 = 0; s0 < ctx->vspp[i];)
        {
            if (cap - fill < npix)
            {
                batches.push_back(BatchRec{(uint32_t)entries.size(), 0u, 0u, 0u});
                fill = 0;
            }
            const uint32_t ns = std::min((cap - fill) / npix, ctx->vspp[i] - s0);
            BatchRec&      b  = batches.back();
            entries.push_back(BatchEntry{i, fill, s0, 0u});
            b.nentries += 1;
            fill += ns * npix;
            b.npaths = fill;
            s0 += ns;
            ctx->vlast_batch[i] = (uint32_t)batches.size() - 1;
        }
    }

    /* queues for the largest batch actually built, not for `cap`: a few
     * small views need only their own paths */
    uint32_t most = 1;
    for (const BatchRec& b : batches)
        most = std::max(most, b.npaths);
    rc = ensure_render_buffers(ctx, most);
    if (rc)
        return rc;
    if (3 * (size_t)film * sizeof(double) > ctx->d_vaccum.bytes)
    {
        CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream)); /* the old film may still be written */
        CLHIP_CHECK(ctx, ctx->d_vaccum.alloc(std::max<size_t>(3 * (size_t)film * sizeof(double), 16)));
    }
    CLHIP_CHECK(ctx, hipMemsetAsync(ctx->d_vaccum.p, 0, 3 * (size_t)film * sizeof(double), ctx->stream));
    ctx->have_views = true;
    if (nviews == 0)
        return CLHIP_OK;
    rc = upload_tables(ctx, ctx->vviews, batches, entries);
    if (rc)
        return rc;
    return run_batches(ctx, (uint32_t)batches.size(), maxd, ctx->d_vaccum.as<double>(), true, false);
}

int clhip_view_sums(clhip_ctx* ctx, uint32_t view, double* rgb, size_t bytes)
{
    if (!ctx || !rgb)
        return fail(ctx, CLHIP_E_INVALID, "clhip_view_sums: bad arguments");
    if (!ctx->have_views || view >= ctx->vviews.size())
        return fail(ctx, CLHIP_E_INVALID, "view %u of %zu rendered", view, ctx->vviews.size());
    const ViewRec& v = ctx->vviews[view];
    if (bytes < (size_t)v.npix * 3 * sizeof(double))
        return fail(ctx, CLHIP_E_INVALID, "clhip_view_sums: buffer too small");
    CLHIP_CHECK(ctx, hipMemcpyAsync(rgb, ctx->d_vaccum.as<double>() + 3 * (size_t)v.film_offset,
                                    (size_t)v.npix * 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return CLHIP_OK;
}

int clhip_view_readback(clhip_ctx* ctx, uint32_t view, float* rgba, size_t bytes)
{
    if (!ctx || !rgba)
        return fail(ctx, CLHIP_E_INVALID, "clhip_view_readback: bad arguments");
    if (!ctx->have_views || view >= ctx->vviews.size())
        return fail(ctx, CLHIP_E_INVALID, "view %u of %zu rendered", view, ctx->vviews.size());
    const ViewRec& v = ctx->vviews[view];
    if (bytes < (size_t)v.npix * 16)
        return fail(ctx, CLHIP_E_INVALID, "clhip_view_readback: buffer too small");
    DevBuf out;
    CLHIP_CHECK(ctx, out.alloc((size_t)v.npix * 16));
    CLHIP_CHECK(ctx, clpt::launch_resolve(ctx->stream, ctx->d_vaccum.as<double>() + 3 * (size_t)v.film_offset, v.npix,
                                          ctx->vspp[view], 0, out.as<float4>()));
    CLHIP_CHECK(ctx, hipMemcpyAsync(rgba, out.p, (size_t)v.npix * 16, hipMemcpyDeviceToHost, ctx->stream));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return CLHIP_OK;
}

int clhip_view_latency(clhip_ctx* ctx, double* ms, uint32_t nviews)
{
    if (!ctx || !ms)
        return fail(ctx, CLHIP_E_INVALID, "clhip_view_latency: bad arguments");
    if (!ctx->have_views || nviews > ctx->vviews.size())
        return fail(ctx, CLHIP_E_INVALID, "asked for %u of %zu views", nviews, ctx->vviews.size());
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->batch_events.empty())
        return fail(ctx, CLHIP_E_STATE, "the last render was not a clhip_render_views");
    for (uint32_t i = 0; i < nviews; ++i)
    {
        float t = 0.0f;
        CLHIP_CHECK(ctx, hipEventElapsedTime(&t, ctx->ev_begin, ctx->batch_events[ctx->vlast_batch[i]]));
        ms[i] = t;
    }
    return CLHIP_OK;
}

int clhip_sync(clhip_ctx* ctx)
{
    if (!ctx)
        return fail(nullptr, CLHIP_E_INVALID, "ctx is NULL");
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return CLHIP_OK;
}

static int readback_film(clhip_ctx* ctx, float* rgba, size_t bytes, int srgb, const char* who)
{
    if (!ctx || !rgba)
        return fail(ctx, CLHIP_E_INVALID, "%s: bad arguments", who);
    const size_t npix = (size_t)ctx->width * ctx->height;
    if (bytes < npix * 4 * sizeof(float))
        return fail(ctx, CLHIP_E_INVALID, "%s: buffer of %zu bytes, need %zu", who, bytes, npix * 16);
    if (!ctx->d_accum.p)
    {
        /* nothing rendered: black, which both encodings map to 0 */
        std::memset(rgba, 0, npix * 16);
        for (size_t i = 0; i < npix; ++i)
            rgba[4 * i + 3] = 1.0f;
        return CLHIP_OK;
    }
    CLHIP_CHECK(ctx, clpt::launch_resolve(ctx->stream, ctx->d_accum.as<double>(), (uint32_t)npix, ctx->spp, srgb,
                                          ctx->d_out.as<float4>()));
    CLHIP_CHECK(ctx, hipMemcpyAsync(rgba, ctx->d_out.p, npix * 16, hipMemcpyDeviceToHost, ctx->stream));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return CLHIP_OK;
}

int clhip_readback(clhip_ctx* ctx, float* rgba, size_t bytes) { return readback_film(ctx, rgba, bytes, 0, "clhip_readback"); }

int clhip_readback_srgb(clhip_ctx* ctx, float* rgba, size_t bytes)
{
    return readback_film(ctx, rgba, bytes, 1, "clhip_readback_srgb");
}

int clhip_srgb_batch(clhip_ctx* ctx, const float* in, uint32_t n, int decode, float* out)
{
    if (!ctx || (n && (!in || !out)))
        return fail(ctx, CLHIP_E_INVALID, "clhip_srgb_batch: bad arguments");
    if (n == 0)
        return CLHIP_OK;
    Tmp a, b;
    CLHIP_CHECK(ctx, a.from_host(in, n));
    CLHIP_CHECK(ctx, b.zeros((size_t)n * sizeof(float)));
    CLHIP_CHECK(ctx, clpt::launch_srgb_batch(ctx->stream, a.b.as<const float>(), n, decode ? 1 : 0, b.b.as<float>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, b.to_host(out, n));
    return CLHIP_OK;
}

int clhip_readback_sums(clhip_ctx* ctx, double* rgb, size_t bytes)
{
    if (!ctx || !rgb)
        return fail(ctx, CLHIP_E_INVALID, "clhip_readback_sums: bad arguments");
    const size_t npix = (size_t)ctx->width * ctx->height;
    if (bytes < npix * 3 * sizeof(double))
        return fail(ctx, CLHIP_E_INVALID, "clhip_readback_sums: buffer too small");
    if (!ctx->d_accum.p)
    {
        std::memset(rgb, 0, npix * 3 * sizeof(double));
        return CLHIP_OK;
    }
    CLHIP_CHECK(ctx, hipMemcpyAsync(rgb, ctx->d_accum.p, npix * 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return CLHIP_OK;
}

int clhip_counters(clhip_ctx* ctx, clhip_stats* out)
{
    if (!ctx || !out)
        return fail(ctx, CLHIP_E_INVALID, "clhip_counters: bad arguments");
    std::memset(out, 0, sizeof(*out));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->d_tot.p)
    {
        TotalCounters      t;
        unsigned long long c[4];
        CLHIP_CHECK(ctx, hipMemcpy(&t, ctx->d_tot.p, sizeof(t), hipMemcpyDeviceToHost));
        CLHIP_CHECK(ctx, hipMemcpy(c, ctx->d_ctr.p, sizeof(c), hipMemcpyDeviceToHost));
        for (int b = 0; b <= CLHIP_MAX_DEPTH && b < CLPT_MAXD + 2; ++b)
        {
            out->rays[b]    = t.rays[b];
            out->hits[b]    = t.hits[b];
            out->shadow[b]  = t.shadow[b];
            out->visible[b] = t.visible[b];
        }
        out->trace_nodes  = c[0];
        out->trace_tris   = c[1];
        out->shadow_nodes = c[2];
        out->shadow_tris  = c[3];
    }
    out->trace_launches  = ctx->trace_launches;
    out->shadow_launches = ctx->shadow_launches;
    out->batches         = ctx->nbatches;
    out->graph_replays   = ctx->graph_replays;
    out->graph_captures  = ctx->graph_captures;
    out->queue_paths     = ctx->qcap;
    out->scene_bytes     = ctx->node_bytes + ctx->tri_bytes;
    out->blas_builds     = ctx->blas_builds;
    for (auto& tl : ctx->timed)
    {
        float ms = 0.0f;
        CLHIP_CHECK(ctx, hipEventElapsedTime(&ms, tl.a, tl.b));
        if (tl.kind == 0)
            out->trace_ms += ms;
        else if (tl.kind == 1)
            out->shade_ms += ms;
        else
            out->shadow_ms += ms;
    }
    if (ctx->have_render_events)
    {
        float ms = 0.0f;
        CLHIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
        out->render_ms = ms;
    }
    return CLHIP_OK;
}

/* ------------------------------------------------------------ batch APIs */

int clhip_rng_batch(clhip_ctx* ctx, const uint32_t* seeds4, uint32_t n, uint32_t ndraws, uint32_t* out_u32, float* out_f)
{
    if (!ctx || !seeds4 || !out_u32 || !out_f)
        return fail(ctx, CLHIP_E_INVALID, "clhip_rng_batch: bad arguments");
    Tmp s, ou, of;
    CLHIP_CHECK(ctx, s.from_host(seeds4, 4 * (size_t)n));
    CLHIP_CHECK(ctx, ou.zeros((size_t)n * ndraws * 4));
    CLHIP_CHECK(ctx, of.zeros((size_t)n * ndraws * 4));
    CLHIP_CHECK(ctx, clpt::launch_rng_batch(ctx->stream, s.b.as<const uint4>(), n, ndraws, ou.b.as<uint32_t>(), of.b.as<float>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, ou.to_host(out_u32, (size_t)n * ndraws));
    CLHIP_CHECK(ctx, of.to_host(out_f, (size_t)n * ndraws));
    return CLHIP_OK;
}

int clhip_camera_rays(clhip_ctx* ctx, const uint32_t* pixel_sample2, uint32_t n, float* out8)
{
    if (!ctx || !pixel_sample2 || !out8)
        return fail(ctx, CLHIP_E_INVALID, "clhip_camera_rays: bad arguments");
    if (!ctx->have_camera)
        return fail(ctx, CLHIP_E_STATE, "no camera");
    for (uint32_t i = 0; i < n; ++i)
        if (pixel_sample2[2 * i] >= ctx->cam.npix)
            return fail(ctx, CLHIP_E_INVALID, "pixel %u out of range", pixel_sample2[2 * i]);
    ViewRec fr = ctx->cam;
    fr.lane    = ctx->params.view;
    Tmp ps, out;
    CLHIP_CHECK(ctx, ps.from_host(pixel_sample2, 2 * (size_t)n));
    CLHIP_CHECK(ctx, out.zeros((size_t)n * 32));
    CLHIP_CHECK(ctx, clpt::launch_camera_batch(ctx->stream, fr, ps.b.as<const uint2>(), n, out.b.as<float4>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, out.to_host(out8, 8 * (size_t)n));
    return CLHIP_OK;
}

int clhip_bsdf_batch(clhip_ctx* ctx, const uint32_t* material, const float* wi3, const float* wo3, const float* u3,
                     uint32_t n, float* eval_f3, float* eval_pdf, float* smp_wi3, float* smp_f3, float* smp_pdf)
{
    if (!ctx || !material || !wi3 || !wo3 || !u3 || !eval_f3 || !eval_pdf || !smp_wi3 || !smp_f3 || !smp_pdf)
        return fail(ctx, CLHIP_E_INVALID, "clhip_bsdf_batch: bad arguments");
    for (uint32_t i = 0; i < n; ++i)
        if (material[i] >= ctx->mtl.size())
            return fail(ctx, CLHIP_E_INVALID, "item %u: material %u of %zu", i, material[i], ctx->mtl.size());
    const SceneView sc = scene_view(ctx);
    Tmp             m, a, b, u, ef, ep, sw, sf, sp;
    CLHIP_CHECK(ctx, m.from_host(material, n));
    CLHIP_CHECK(ctx, a.from_host(wi3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, b.from_host(wo3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, u.from_host(u3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, ef.zeros(12 * (size_t)n));
    CLHIP_CHECK(ctx, ep.zeros(4 * (size_t)n));
    CLHIP_CHECK(ctx, sw.zeros(12 * (size_t)n));
    CLHIP_CHECK(ctx, sf.zeros(12 * (size_t)n));
    CLHIP_CHECK(ctx, sp.zeros(4 * (size_t)n));
    CLHIP_CHECK(ctx, clpt::launch_bsdf_batch(ctx->stream, sc, m.b.as<const uint32_t>(), a.b.as<const float>(),
                                             b.b.as<const float>(), u.b.as<const float>(), n, ef.b.as<float>(),
                                             ep.b.as<float>(), sw.b.as<float>(), sf.b.as<float>(), sp.b.as<float>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, ef.to_host(eval_f3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, ep.to_host(eval_pdf, n));
    CLHIP_CHECK(ctx, sw.to_host(smp_wi3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, sf.to_host(smp_f3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, sp.to_host(smp_pdf, n));
    return CLHIP_OK;
}

int clhip_emitter_batch(clhip_ctx* ctx, const float* u2, uint32_t n, float* dir3, float* value3, float* pdf, uint32_t* which)
{
    if (!ctx || !u2 || !dir3 || !value3 || !pdf || !which)
        return fail(ctx, CLHIP_E_INVALID, "clhip_emitter_batch: bad arguments");
    const SceneView sc = scene_view(ctx);
    Tmp             u, d, v, p, w;
    CLHIP_CHECK(ctx, u.from_host(u2, 2 * (size_t)n));
    CLHIP_CHECK(ctx, d.zeros(12 * (size_t)n));
    CLHIP_CHECK(ctx, v.zeros(12 * (size_t)n));
    CLHIP_CHECK(ctx, p.zeros(4 * (size_t)n));
    CLHIP_CHECK(ctx, w.zeros(4 * (size_t)n));
    CLHIP_CHECK(ctx, clpt::launch_emitter_batch(ctx->stream, sc, u.b.as<const float>(), n, d.b.as<float>(),
                                                v.b.as<float>(), p.b.as<float>(), w.b.as<uint32_t>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, d.to_host(dir3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, v.to_host(value3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, p.to_host(pdf, n));
    CLHIP_CHECK(ctx, w.to_host(which, n));
    return CLHIP_OK;
}

int clhip_dome_eval_batch(clhip_ctx* ctx, const float* dir3, uint32_t n, float* le3, float* pdf)
{
    if (!ctx || !dir3 || !le3 || !pdf)
        return fail(ctx, CLHIP_E_INVALID, "clhip_dome_eval_batch: bad arguments");
    if (ctx->dome_tex < 0)
        return fail(ctx, CLHIP_E_STATE, "no HDRI dome");
    const SceneView sc = scene_view(ctx);
    Tmp             d, l, p;
    CLHIP_CHECK(ctx, d.from_host(dir3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, l.zeros(12 * (size_t)n));
    CLHIP_CHECK(ctx, p.zeros(4 * (size_t)n));
    CLHIP_CHECK(ctx, clpt::launch_dome_eval_batch(ctx->stream, sc, d.b.as<const float>(), n, l.b.as<float>(), p.b.as<float>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, l.to_host(le3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, p.to_host(pdf, n));
    return CLHIP_OK;
}

int clhip_dome_tables(clhip_ctx* ctx, float* pcond, float* cdfcond, float* pv, float* cdfv, uint32_t* w, uint32_t* h)
{
    if (!ctx)
        return fail(nullptr, CLHIP_E_INVALID, "ctx is NULL");
    if (ctx->dome_tex < 0)
        return fail(ctx, CLHIP_E_STATE, "no HDRI dome");
    if (w)
        *w = ctx->dome_w;
    if (h)
        *h = ctx->dome_h;
    const size_t W = (size_t)ctx->dome_w + 1;
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (pcond)
        CLHIP_CHECK(ctx, hipMemcpy(pcond, ctx->d_pcond.p, W * ctx->dome_h * 4, hipMemcpyDeviceToHost));
    if (cdfcond)
        CLHIP_CHECK(ctx, hipMemcpy(cdfcond, ctx->d_cdfcond.p, W * ctx->dome_h * 4, hipMemcpyDeviceToHost));
    if (pv)
        CLHIP_CHECK(ctx, hipMemcpy(pv, ctx->d_pv.p, ((size_t)ctx->dome_h + 1) * 4, hipMemcpyDeviceToHost));
    if (cdfv)
        CLHIP_CHECK(ctx, hipMemcpy(cdfv, ctx->d_cdfv.p, ((size_t)ctx->dome_h + 1) * 4, hipMemcpyDeviceToHost));
    return CLHIP_OK;
}

static int trace_common(clhip_ctx* ctx, const float* rays8, uint32_t n, int any, std::vector<float4>& out,
                        std::vector<uint32_t>& prim)
{
    int rc = sync_instances(ctx);
    if (rc)
        return rc;
    const SceneView sc = scene_view(ctx);
    const uint32_t  blocks = std::min<uint32_t>(ctx->num_cus * 8, std::max<uint32_t>(1, (n + 255) / 256));
    Tmp             r, o, p, g;
    CLHIP_CHECK(ctx, r.from_host(rays8, 8 * (size_t)n));
    CLHIP_CHECK(ctx, o.zeros(16 * (size_t)n));
    CLHIP_CHECK(ctx, p.zeros(4 * (size_t)n));
    CLHIP_CHECK(ctx, g.zeros((size_t)blocks * 256 * CLPT_GSTACK_ENTRIES * sizeof(uint2)));
    CLHIP_CHECK(ctx, clpt::launch_trace_batch(ctx->stream, sc, r.b.as<const float4>(), n, any, o.b.as<float4>(),
                                              p.b.as<uint32_t>(), g.b.as<uint2>(), blocks));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    out.resize(n);
    prim.resize(n);
    CLHIP_CHECK(ctx, o.to_host(out.data(), n));
    CLHIP_CHECK(ctx, p.to_host(prim.data(), n));
    return CLHIP_OK;
}

int clhip_trace_batch(clhip_ctx* ctx, const float* rays8, uint32_t n, float* t, float* b12, uint32_t* inst, uint32_t* prim)
{
    if (!ctx || !rays8 || !t || !b12 || !inst || !prim)
        return fail(ctx, CLHIP_E_INVALID, "clhip_trace_batch: bad arguments");
    std::vector<float4>   out;
    std::vector<uint32_t> pr;
    int                   rc = trace_common(ctx, rays8, n, 0, out, pr);
    if (rc)
        return rc;
    for (uint32_t i = 0; i < n; ++i)
    {
        t[i]           = out[i].x;
        b12[2 * i]     = out[i].y;
        b12[2 * i + 1] = out[i].z;
        uint32_t iu;
        std::memcpy(&iu, &out[i].w, 4);
        inst[i] = iu;
        prim[i] = pr[i];
    }
    return CLHIP_OK;
}

int clhip_occlusion_batch(clhip_ctx* ctx, const float* rays8, uint32_t n, uint8_t* occluded)
{
    if (!ctx || !rays8 || !occluded)
        return fail(ctx, CLHIP_E_INVALID, "clhip_occlusion_batch: bad arguments");
    std::vector<float4>   out;
    std::vector<uint32_t> pr;
    int                   rc = trace_common(ctx, rays8, n, 1, out, pr);
    if (rc)
        return rc;
    for (uint32_t i = 0; i < n; ++i)
        occluded[i] = out[i].x != 0.0f ? 1u : 0u;
    return CLHIP_OK;
}

int clhip_trace_rays_device(clhip_ctx* ctx, const float* rays8, uint32_t n, int any, float* hit4, uint32_t* prim)
{
    if (!ctx || (n && (!rays8 || !hit4 || !prim)))
        return fail(ctx, CLHIP_E_INVALID, "clhip_trace_rays_device: bad arguments");
    if (n > CLHIP_TRACE_RAYS_DEVICE_MAX)
        return fail(ctx, CLHIP_E_INVALID, "clhip_trace_rays_device: %u rays, at most %u per call", n,
                    (uint32_t)CLHIP_TRACE_RAYS_DEVICE_MAX);
    if (!ctx->have_tlas)
        return fail(ctx, CLHIP_E_STATE, "no TLAS: call clhip_build_tlas first");
    if (!ctx->noalloc_stack)
        return fail(ctx, CLHIP_E_STATE, "no scratch address for device %d", ctx->device);
    if (n == 0)
        return CLHIP_OK;
    int dev = -1;
    CLHIP_CHECK(ctx, hipGetDevice(&dev));
    if (dev != ctx->device)
        return fail(ctx, CLHIP_E_STATE, "device %d is current, the context's is %d", dev, ctx->device);
    CLHIP_CHECK(ctx, clpt::launch_trace_noalloc(ctx->stream, scene_view(ctx), reinterpret_cast<const float4*>(rays8),
                                                n, any ? 1 : 0, reinterpret_cast<float4*>(hit4), prim,
                                                ctx->noalloc_stack));
    return CLHIP_OK;
}

int clhip_copy_probe(clhip_ctx* ctx, size_t bytes, uint32_t iters, double* ms_per_copy)
{
    if (!ctx || bytes < 16 || iters == 0)
        return fail(ctx, CLHIP_E_INVALID, "clhip_copy_probe: bad arguments");
    const size_t n = bytes / 16;
    DevBuf       a, b;
    CLHIP_CHECK(ctx, a.alloc(n * 16));
    CLHIP_CHECK(ctx, b.alloc(n * 16));
    CLHIP_CHECK(ctx, hipMemsetAsync(a.p, 0, n * 16, ctx->stream));
    hipEvent_t e0, e1;
    CLHIP_CHECK(ctx, hipEventCreate(&e0));
    CLHIP_CHECK(ctx, hipEventCreate(&e1));
    const uint32_t blocks = ctx->num_cus * 32;
    CLHIP_CHECK(ctx, clpt::launch_copy_probe(ctx->stream, a.as<const float4>(), b.as<float4>(), n, blocks));
    CLHIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
    for (uint32_t i = 0; i < iters; ++i)
        CLHIP_CHECK(ctx, clpt::launch_copy_probe(ctx->stream, a.as<const float4>(), b.as<float4>(), n, blocks));
    CLHIP_CHECK(ctx, hipEventRecord(e1, ctx->stream));
    CLHIP_CHECK(ctx, hipEventSynchronize(e1));
    float ms = 0.0f;
    CLHIP_CHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_copy)
        *ms_per_copy = (double)ms / iters;
    return CLHIP_OK;
}

int clhip_instance_inverses(clhip_ctx* ctx, float* out12, uint32_t n)
{
    if (!ctx || !out12)
        return fail(ctx, CLHIP_E_INVALID, "clhip_instance_inverses: bad arguments");
    if (n > ctx->inst.size())
        return fail(ctx, CLHIP_E_INVALID, "asked for %u of %zu instances", n, ctx->inst.size());
    for (uint32_t i = 0; i < n; ++i)
        std::memcpy(out12 + 12 * (size_t)i, ctx->inst[i].inv, 12 * sizeof(float));
    return CLHIP_OK;
}

/* ------------------------------------------------------------ learned proposal */

int clhip_set_guide(clhip_ctx* ctx, const clhip_guide* g)
{
    if (!ctx || !g)
        return fail(ctx, CLHIP_E_INVALID, "clhip_set_guide: bad arguments");
    clhip_guide v = *g;
    if (v.bsdf_fraction == 0.0f)
        v.bsdf_fraction = 0.5f;
    if (!(v.bsdf_fraction > 0.0f && v.bsdf_fraction <= 1.0f))
        return fail(ctx, CLHIP_E_INVALID, "bsdf_fraction %g outside (0, 1]: the BSDF must keep a positive share",
                    (double)v.bsdf_fraction);
    if (v.record_paths == 0)
        v.record_paths = 32768u;
    if (v.batch_size == 0)
        v.batch_size = 65536u;
    if (v.batch_size % 1024u != 0u || v.batch_size > (1u << 22))
        return fail(ctx, CLHIP_E_INVALID, "batch_size %u must be a multiple of 1024 up to 2^22", v.batch_size);
    if (v.record_paths > (1u << 22))
        return fail(ctx, CLHIP_E_INVALID, "record_paths %u above 2^22", v.record_paths);
    if (v.learning_rate == 0.0f)
        v.learning_rate = 1e-3f;
    if (v.steps_per_batch == 0)
        v.steps_per_batch = 1;
    if (v.steps_per_batch > 256)
        return fail(ctx, CLHIP_E_INVALID, "steps_per_batch %u above 256", v.steps_per_batch);
    if (v.precision > 1u)
        return fail(ctx, CLHIP_E_INVALID, "precision %u: 0 (bfloat16) or 1 (FP8)", v.precision);
    if (v.train_kernel > 1u)
        return fail(ctx, CLHIP_E_INVALID, "train_kernel %u: 0 (fused) or 1 (per GEMM)", v.train_kernel);
    if (!(v.learning_rate > 0.0f))
        return fail(ctx, CLHIP_E_INVALID, "learning_rate must be positive");
    const bool fp8_on = v.precision == 1u && ctx->guide.precision != 1u;
    ctx->guide        = v;
    if (v.enable && !ctx->guide_ready)
    {
        int rc = guide_reset_weights(ctx, 1u);
        if (rc)
            return rc;
    }
    if (fp8_on && ctx->guide_ready) /* refresh the e4m3 image for the current weights */
    {
        CLHIP_CHECK(ctx, clpt::launch_guide_pack_fp8(ctx->stream, ctx->g_params.as<const float>(),
                                                     ctx->g_gc.as<GuideCounters>(), ctx->g_img8.as<uint8_t>(),
                                                     ctx->g_scl.as<float>()));
        CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    /* training buffers now rather than inside the first render */
    if (v.enable && v.train)
    {
        int rc = guide_train_buffers(ctx, v.batch_size, v.steps_per_batch);
        if (!rc)
            rc = guide_record_buffers(ctx, v.record_paths, ctx->qcap);
        if (rc)
            return rc;
    }
    return CLHIP_OK;
}

int clhip_guide_reset(clhip_ctx* ctx, uint64_t seed)
{
    if (!ctx)
        return fail(nullptr, CLHIP_E_INVALID, "ctx is NULL");
    return guide_reset_weights(ctx, seed);
}

int clhip_guide_get_weights(clhip_ctx* ctx, float* w, uint32_t n)
{
    if (!ctx || !w || n != CLPT_G_PARAMS)
        return fail(ctx, CLHIP_E_INVALID, "clhip_guide_get_weights: need %d floats", CLPT_G_PARAMS);
    if (!ctx->guide_ready)
        return fail(ctx, CLHIP_E_STATE, "no weights: call clhip_guide_reset or clhip_set_guide first");
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, hipMemcpy(w, ctx->g_params.p, CLPT_G_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
    return CLHIP_OK;
}

int clhip_guide_set_weights(clhip_ctx* ctx, const float* w, uint32_t n)
{
    if (!ctx || !w || n != CLPT_G_PARAMS)
        return fail(ctx, CLHIP_E_INVALID, "clhip_guide_set_weights: need %d floats", CLPT_G_PARAMS);
    if (!ctx->guide_ready)
    {
        int rc = guide_reset_weights(ctx, 1u);
        if (rc)
            return rc;
    }
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, hipMemcpy(ctx->g_params.p, w, CLPT_G_PARAMS * sizeof(float), hipMemcpyHostToDevice));
    CLHIP_CHECK(ctx, clpt::launch_guide_pack(ctx->stream, ctx->g_params.as<const float>(), ctx->g_W.as<__bf16>(),
                                             ctx->g_WT.as<__bf16>(), ctx->g_img.as<__bf16>()));
    CLHIP_CHECK(ctx, clpt::launch_guide_pack_fp8(ctx->stream, ctx->g_params.as<const float>(),
                                                 ctx->g_gc.as<GuideCounters>(), ctx->g_img8.as<uint8_t>(),
                                                 ctx->g_scl.as<float>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return CLHIP_OK;
}

int clhip_guide_fp8_scales(clhip_ctx* ctx, float* scl)
{
    if (!ctx || !scl)
        return fail(ctx, CLHIP_E_INVALID, "clhip_guide_fp8_scales: bad arguments");
    if (!ctx->guide_ready)
        return fail(ctx, CLHIP_E_STATE, "no weights: call clhip_guide_reset or clhip_set_guide first");
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, hipMemcpy(scl, ctx->g_scl.p, 10 * sizeof(float), hipMemcpyDeviceToHost));
    return CLHIP_OK;
}

int clhip_guide_records(clhip_ctx* ctx, float* V, float* E, float* samples13, uint32_t cap, uint32_t info[3])
{
    if (!ctx || !info)
        return fail(ctx, CLHIP_E_INVALID, "clhip_guide_records: bad arguments");
    /* the records are laid out for the current record_paths (<= capacity) */
    const uint32_t slots = ctx->g_slots ? std::min<uint32_t>(ctx->guide.record_paths, (uint32_t)ctx->g_slots) : 0u;
    info[0]              = slots;
    info[1]              = CLPT_MAXD + 1u;
    info[2]              = (uint32_t)(ctx->g_samples.bytes / sizeof(GuideSample));
    if (!V && !E && !samples13)
        return CLHIP_OK;
    if (slots == 0)
        return fail(ctx, CLHIP_E_STATE, "clhip_guide_records: nothing recorded yet");
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const size_t recs = (size_t)(CLPT_MAXD + 1) * slots;
    if (V)
        CLHIP_CHECK(ctx, hipMemcpy(V, ctx->g_V.p, 4 * recs * sizeof(float4), hipMemcpyDeviceToHost));
    if (E)
        CLHIP_CHECK(ctx, hipMemcpy(E, ctx->g_E.p, recs * sizeof(float4), hipMemcpyDeviceToHost));
    if (samples13 && cap)
    {
        const uint32_t           n = std::min(cap, info[2]);
        std::vector<GuideSample> hs(n);
        CLHIP_CHECK(ctx, hipMemcpy(hs.data(), ctx->g_samples.p, n * sizeof(GuideSample), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i)
        {
            const float src[13] = {hs[i].a.x, hs[i].a.y, hs[i].a.z, hs[i].a.w, hs[i].b.x, hs[i].b.y, hs[i].b.z,
                                   hs[i].b.w, hs[i].c.x, hs[i].c.y, hs[i].c.z, hs[i].c.w, hs[i].d.x};
            std::memcpy(samples13 + 13 * (size_t)i, src, sizeof(src));
        }
    }
    return CLHIP_OK;
}

int clhip_guide_stats_get(clhip_ctx* ctx, clhip_guide_stats* out)
{
    if (!ctx || !out)
        return fail(ctx, CLHIP_E_INVALID, "clhip_guide_stats_get: bad arguments");
    std::memset(out, 0, sizeof(*out));
    if (!ctx->guide_ready)
        return CLHIP_OK;
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    GuideCounters g;
    CLHIP_CHECK(ctx, hipMemcpy(&g, ctx->g_gc.p, sizeof(g), hipMemcpyDeviceToHost));
    out->steps      = g.steps;
    out->samples    = g.samples_total;
    out->last_count = g.trained;
    out->last_loss  = g.loss;
    out->issued     = ctx->g_issued;
    out->training   = guide_recording(ctx) ? 1u : 0u;
    for (auto& t : ctx->train_timed)
    {
        float ms = 0.0f;
        CLHIP_CHECK(ctx, hipEventElapsedTime(&ms, t.a, t.b));
        out->train_ms += ms;
    }
    return CLHIP_OK;
}

int clhip_guide_infer(clhip_ctx* ctx, const float* q9, uint32_t n, float* out40)
{
    if (!ctx || !q9 || !out40)
        return fail(ctx, CLHIP_E_INVALID, "clhip_guide_infer: bad arguments");
    if (!ctx->guide_ready)
        return fail(ctx, CLHIP_E_STATE, "no weights: call clhip_guide_reset or clhip_set_guide first");
    if (n == 0)
        return CLHIP_OK;
    Tmp q, o;
    CLHIP_CHECK(ctx, q.from_host(q9, 9 * (size_t)n));
    CLHIP_CHECK(ctx, o.zeros((size_t)n * CLPT_G_NOUT * sizeof(float)));
    const uint32_t blocks = std::min<uint32_t>(ctx->num_cus * 2u, (n + 255u) / 256u);
    CLHIP_CHECK(ctx, clpt::launch_guide_infer(ctx->stream, ctx->g_img.as<const __bf16>(),
                                              ctx->guide.precision == 1u ? ctx->g_img8.as<const uint8_t>() : nullptr,
                                              ctx->g_scl.as<const float>(), q.b.as<const float>(), n, o.b.as<float>(),
                                              blocks));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, o.to_host(out40, (size_t)n * CLPT_G_NOUT));
    return CLHIP_OK;
}

int clhip_guide_mixture_batch(clhip_ctx* ctx, const float* out40, const float* dir3, const float* u3, uint32_t n,
                              float* pdf, float* smp_dir3, float* smp_pdf)
{
    if (!ctx || (n && (!out40 || !dir3 || !u3 || !pdf || !smp_dir3 || !smp_pdf)))
        return fail(ctx, CLHIP_E_INVALID, "clhip_guide_mixture_batch: bad arguments");
    if (n == 0)
        return CLHIP_OK;
    Tmp o, d, u, p, sd, sp;
    CLHIP_CHECK(ctx, o.from_host(out40, (size_t)n * CLPT_G_NOUT));
    CLHIP_CHECK(ctx, d.from_host(dir3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, u.from_host(u3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, p.zeros((size_t)n * sizeof(float)));
    CLHIP_CHECK(ctx, sd.zeros(3 * (size_t)n * sizeof(float)));
    CLHIP_CHECK(ctx, sp.zeros((size_t)n * sizeof(float)));
    CLHIP_CHECK(ctx, clpt::launch_guide_mixture_batch(ctx->stream, o.b.as<const float>(), d.b.as<const float>(),
                                                      u.b.as<const float>(), n, p.b.as<float>(), sd.b.as<float>(),
                                                      sp.b.as<float>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, p.to_host(pdf, n));
    CLHIP_CHECK(ctx, sd.to_host(smp_dir3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, sp.to_host(smp_pdf, n));
    return CLHIP_OK;
}

int clhip_guide_query_batch(clhip_ctx* ctx, const float* p3, const float* n3, const float* wo3, uint32_t n, float* q9,
                            float box6[6])
{
    if (!ctx || (n && (!p3 || !n3 || !wo3 || !q9)))
        return fail(ctx, CLHIP_E_INVALID, "clhip_guide_query_batch: bad arguments");
    if (!ctx->have_tlas)
        return fail(ctx, CLHIP_E_STATE, "no TLAS: the query box is the scene's bounds");
    float lo[3], inv[3];
    for (int k = 0; k < 3; ++k)
    {
        const float ext = ctx->world_hi[k] - ctx->world_lo[k];
        lo[k]           = ctx->world_lo[k];
        inv[k]          = ext > 1e-20f ? 1.0f / ext : 1.0f;
        if (box6)
        {
            box6[k]     = ctx->world_lo[k];
            box6[3 + k] = ctx->world_hi[k];
        }
    }
    if (n == 0)
        return CLHIP_OK;
    Tmp a, b, c, q;
    CLHIP_CHECK(ctx, a.from_host(p3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, b.from_host(n3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, c.from_host(wo3, 3 * (size_t)n));
    CLHIP_CHECK(ctx, q.zeros(9 * (size_t)n * sizeof(float)));
    CLHIP_CHECK(ctx, clpt::launch_guide_query_batch(ctx->stream, lo, inv, a.b.as<const float>(), b.b.as<const float>(),
                                                    c.b.as<const float>(), n, q.b.as<float>()));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CLHIP_CHECK(ctx, q.to_host(q9, 9 * (size_t)n));
    return CLHIP_OK;
}

int clhip_guide_train(clhip_ctx* ctx, const float* s13, uint32_t n, float* grad, float* loss)
{
    if (!ctx || (!s13 && n))
        return fail(ctx, CLHIP_E_INVALID, "clhip_guide_train: bad arguments");
    if (!ctx->guide_ready)
        return fail(ctx, CLHIP_E_STATE, "no weights: call clhip_guide_reset or clhip_set_guide first");
    if (n > (1u << 22))
        return fail(ctx, CLHIP_E_INVALID, "at most 2^22 samples per step");
    const uint32_t N = std::max(1024u, (n + 1023u) / 1024u * 1024u);
    int            rc = guide_train_buffers(ctx, N);
    if (rc)
        return rc;
    std::vector<GuideSample> hs(n);
    for (uint32_t i = 0; i < n; ++i)
    {
        const float* x = s13 + 13 * (size_t)i;
        hs[i].a        = make_float4(x[0], x[1], x[2], x[3]);
        hs[i].b        = make_float4(x[4], x[5], x[6], x[7]);
        hs[i].c        = make_float4(x[8], x[9], x[10], x[11]);
        hs[i].d        = make_float4(x[12], 0.0f, 0.0f, 0.0f);
    }
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (n)
        CLHIP_CHECK(ctx, hipMemcpy(ctx->g_samples.p, hs.data(), n * sizeof(GuideSample), hipMemcpyHostToDevice));
    CLHIP_CHECK(ctx, hipMemcpy(ctx->g_gc.p, &n, sizeof(uint32_t), hipMemcpyHostToDevice)); /* GuideCounters::count */
    Tmp g;
    if (grad)
        CLHIP_CHECK(ctx, g.zeros(CLPT_G_PARAMS * sizeof(float)));
    CLHIP_CHECK(ctx, clpt::launch_guide_train(ctx->stream, guide_bufs(ctx), N,
                                              ctx->guide.learning_rate > 0.0f ? ctx->guide.learning_rate : 1e-3f,
                                              grad ? g.b.as<float>() : nullptr, 0u, 1u));
    CLHIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (grad)
        CLHIP_CHECK(ctx, g.to_host(grad, CLPT_G_PARAMS));
    if (loss)
    {
        GuideCounters gc;
        CLHIP_CHECK(ctx, hipMemcpy(&gc, ctx->g_gc.p, sizeof(gc), hipMemcpyDeviceToHost));
        *loss = gc.loss;
    }
    return CLHIP_OK;
}

} // extern "C"

