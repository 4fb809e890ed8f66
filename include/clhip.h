/*
 * clhip.h -- C ABI of libclpt_hip.so, the MI355X (gfx950) wavefront path
 * tracer that replaces Colvillea's rendering path.
 *
 * Drop-in boundary: core::Integrator (include/colvillea/librender/
 * integrator.h:34-103), created by Integrator::createIntegrator
 * (src/core/librender/integrator.cpp:10-22) and driven by RenderEngine
 * (src/core/librender/renderengine.cpp:27-121).  Each entry point below names
 * the reference method it replaces.  The scene-side inputs are plain POD
 * mirrors of the reference's kernel value types (kernel::Material,
 * kernel::Emitter, kernel::Entity, TriangleMesh) so that a reference Scene
 * can feed them unchanged.
 *
 * Conventions:
 *   - every function returns 0 on success or a negative CLHIP_E_* code and
 *     records a message readable with clhip_last_error(); nothing exits the
 *     process (the reference aborts on CUDA errors, cudacommon.h:13-26);
 *   - inputs are host pointers, copied before the call returns; the context
 *     owns every device allocation;
 *   - one context = one HIP stream; a context is not thread-safe (the
 *     reference is single-threaded as well).
 */
#ifndef CLHIP_H
#define CLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clhip_ctx clhip_ctx;

enum
{
    CLHIP_OK            = 0,
    CLHIP_E_INVALID     = -1, /* bad argument */
    CLHIP_E_HIP         = -2, /* HIP runtime error (no device, out of memory, ...) */
    CLHIP_E_STATE       = -3, /* call order: e.g. render before build */
    CLHIP_E_UNSUPPORTED = -4
};

/* IntegratorType (integrator.h:22-32) */
enum
{
    CLHIP_INTEGRATOR_NONE            = 0,
    CLHIP_INTEGRATOR_DIRECT_LIGHTING = 1,
    CLHIP_INTEGRATOR_PATH_TRACING    = 2
};

/* TriangleMesh (include/colvillea/nodes/shapes/trianglemesh.h) and the
 * device record TriMesh (libkernel/shapes/trimesh.h:16-32). */
typedef struct
{
    const float*    vertices; /* 3 * num_vertices */
    const uint32_t* indices;  /* 3 * num_triangles */
    const float*    normals;  /* 3 * num_vertices, or NULL */
    const float*    tangents; /* 3 * num_vertices, or NULL */
    const float*    uvs;      /* 2 * num_vertices, or NULL */
    uint32_t        num_vertices;
    uint32_t        num_triangles;
} clhip_mesh;

/* MaterialType (libkernel/base/material.h:24-37) */
enum
{
    CLHIP_MTL_DIFFUSE = 0,
    CLHIP_MTL_METAL   = 1,
    CLHIP_MTL_GLASS   = 2
};

/* kernel::Material payloads: DiffuseMtl (materials/diffusemtl.h:13-17),
 * MetalMtl (metal.h:60-65), GlassMtl (glass.h:34-37).  GGX alpha is
 * roughness^2 as in the reference (metal.h:52, glass.h:26).
 * Textures (clhip_create_texture ids, stored as id + 1 so that 0 = none)
 * replace the constant where set, as Material::getBSDF does
 * (diffusemtl.h:22-31, metal.h:36-55): reflectance_tex for the diffuse
 * reflectance or the metal specular reflectance (rgb), roughness_tex for
 * the metal roughness (first channel, squared after the lookup).
 * normal_tex is the tangent-space normal map of Material::setNormalmap
 * (material.h:81-84, getShadingFrameImpl 169-219, with the pull-up fix of
 * 236-285 on diffuse and metal).  Lookups are bilinear with wrap at the
 * surface uv (imagetex2d.h:34-37). */
typedef struct
{
    uint32_t type;
    float    reflectance[3]; /* diffuse reflectance, or metal specular reflectance */
    float    roughness;      /* metal, glass */
    float    eta[3];         /* metal */
    float    k[3];           /* metal */
    float    interior_ior;   /* glass (exterior IOR is 1) */
    uint32_t reflectance_tex; /* texture id + 1, 0 = none (diffuse, metal) */
    uint32_t roughness_tex;   /* texture id + 1, 0 = none (metal) */
    uint32_t normal_tex;      /* texture id + 1, 0 = none (any type) */
} clhip_material;

/* EmitterType (libkernel/base/emitter.h:16-26) */
enum
{
    CLHIP_EMITTER_DIRECTIONAL = 0,
    CLHIP_EMITTER_HDRIDOME    = 1
};

/* DirectionalEmitter (emitters/directional.h:70-73); HDRIDome takes a
 * texture created with clhip_create_texture (hdridome.h:163-175). */
typedef struct
{
    uint32_t type;
    float    intensity[3];   /* directional: colour * intensity (radiance) */
    float    direction[3];   /* directional: emitted direction (sun -> scene) */
    float    angular_radius; /* directional: radians */
    int32_t  texture;        /* HDRI dome: texture id */
} clhip_emitter;

/* Per-render parameters (the reference hard-codes depth 5, path.cpp:135). */
typedef struct
{
    uint32_t max_depth;        /* number of path segments; NEE at vertices 1..max_depth-1 */
    uint32_t rr_depth;         /* Russian roulette from this vertex depth (0 = off) */
    uint32_t view;             /* RNG view lane: z = (view << 20) | sample index */
    uint32_t samples_per_batch;/* samples of every pixel in flight together (0 = as many whole samples
                                  as fit the launch's paths_per_batch, default 2^27 paths) */
} clhip_params;

/* Exact counters of the last clhip_render (sum over its samples).  Index b
 * is the path depth (1..max_depth). */
#define CLHIP_MAX_DEPTH 32
typedef struct
{
    uint64_t rays[CLHIP_MAX_DEPTH + 1];    /* closest-hit rays traced at depth b */
    uint64_t hits[CLHIP_MAX_DEPTH + 1];    /* of which hit geometry */
    uint64_t shadow[CLHIP_MAX_DEPTH + 1];  /* shadow rays spawned at vertex depth b */
    uint64_t visible[CLHIP_MAX_DEPTH + 1]; /* of which unoccluded */
    uint64_t trace_nodes;                  /* BVH4 nodes fetched by closest-hit traversal */
    uint64_t trace_tris;                   /* triangles tested by closest-hit traversal */
    uint64_t shadow_nodes;                 /* BVH4 nodes fetched by any-hit traversal */
    uint64_t shadow_tris;                  /* triangles tested by any-hit traversal */
    uint64_t trace_launches;               /* closest-hit kernel launches */
    uint64_t shadow_launches;              /* any-hit kernel launches */
    double   trace_ms;                     /* summed closest-hit kernel time (events; when timing is on) */
    double   shade_ms;                     /* summed shading kernel time */
    double   shadow_ms;                    /* summed any-hit kernel time */
    double   render_ms;                    /* wall time of the last clhip_render on the stream */
    uint64_t batches;                      /* queue batches of the last render */
    uint64_t graph_replays;                /* of which replayed from a captured hipGraph */
    uint64_t graph_captures;               /* hipGraph captures made by the last render */
    uint64_t queue_paths;                  /* paths the context's queues hold (the largest batch so far) */
    uint64_t scene_bytes;                  /* bytes of the tree's node and triangle records (TLAS and BLASes) */
    uint64_t blas_builds;                  /* mesh trees (BLASes) built since the context was created */
} clhip_stats;

/* ---- context: Integrator::createIntegrator (integrator.cpp:10-22) and
 *      WavefrontPathTracingIntegrator ctor/dtor (path.cpp:19-40) ---- */
int         clhip_create(uint32_t width, uint32_t height, uint32_t integrator_type, clhip_ctx** out);
int         clhip_destroy(clhip_ctx* ctx);
const char* clhip_last_error(const clhip_ctx* ctx); /* ctx may be NULL: last global error */

/* core::ImageTexture2D ctor (src/core/nodes/textures/imagetex2d.cpp:9-69):
 * RGB or RGBA float32 data, stored as RGBA32F, wrap + bilinear. */
int clhip_create_texture(clhip_ctx* ctx, const float* data, uint32_t width, uint32_t height, uint32_t channels,
                         int32_t* texture_id);

/* core::ImageTexture2D ctor for ImageTextureChannelFormat::RGBAU8 images
 * (imagetex2d.cpp:13-15, 50-66): 8-bit RGB or RGBA read as c / 255
 * (cudaReadModeNormalizedFloat); srgb != 0 marks an sRGB image
 * (Image::getIsSRGB, texDesc.sRGB at :65), whose colour channels are
 * decoded to linear per texel, before bilinear filtering, with
 * convertsRGBToLinear (libkernel/base/math.h:131-140).  Stored as RGBA32F. */
int clhip_create_texture_u8(clhip_ctx* ctx, const uint8_t* data, uint32_t width, uint32_t height, uint32_t channels,
                            int srgb, int32_t* texture_id);

/* Tree width of the acceleration structures built from now on: 4 (default;
 * 128-byte nodes with float boxes) or 8 (96-byte nodes with child boxes
 * quantised to bytes).  No reference counterpart (OptiX builds its own).
 * Results do not depend on it; changing it requires building the BLAS and
 * TLAS again. */
int clhip_set_bvh_width(clhip_ctx* ctx, uint32_t width);

/* Integrator::buildBLAS (integrator.h:47; OptiXDevice::buildOptiXAccelBLASes,
 * optixdevice.cpp:136-192): one BVH4 per mesh. */
int clhip_build_blas(clhip_ctx* ctx, const clhip_mesh* meshes, uint32_t num_meshes);

/* The per-mesh form of Integrator::buildBLAS as RenderEngine drives it
 * (renderengine.cpp:64-80 with Scene::collectTriangleMeshForBLASBuilding,
 * scene.cpp:186-210): only the meshes whose BLAS needs a rebuild are passed,
 * meshes[k] being mesh number mesh_index[k]; the other meshes keep their
 * trees.  An index one past the current meshes appends a mesh.  The TLAS
 * must then be built again (clhip_build_tlas), as after clhip_build_blas. */
int clhip_build_blas_subset(clhip_ctx* ctx, const clhip_mesh* meshes, const uint32_t* mesh_index, uint32_t n);

/* Integrator::buildTLAS (integrator.h:49-50; optixdevice.cpp:194-233): one
 * instance per entity.  instance_ids[i] is the entity index of instance i
 * (the reference passes entity indices, scene.cpp:232-236).  xforms is NULL
 * (identity, as in the reference) or num_instances row-major 3x4
 * object-to-world matrices (an extension).  The top tree holds the 64
 * largest subtrees of the identity-transformed instance with the most
 * triangles in place of that instance's one box (partial re-braiding);
 * results do not depend on it. */
int clhip_build_tlas(clhip_ctx* ctx, const uint32_t* mesh_of_instance, const uint32_t* instance_ids,
                     const float* xforms, uint32_t num_instances);

/* Integrator::buildMaterials (integrator.h:52; path.cpp:53-57) */
int clhip_build_materials(clhip_ctx* ctx, const clhip_material* materials, uint32_t num_materials);

/* Integrator::buildGeometryEntities (integrator.h:54; path.cpp:59-63):
 * kernel::Entity{materialIndex} (libkernel/base/entity.h:24-31). */
int clhip_build_entities(clhip_ctx* ctx, const uint32_t* material_index, uint32_t num_entities);

/* Integrator::buildEmitters (integrator.h:56-58; path.cpp:65-82): copies the
 * emitters and, when dome_index >= 0, builds the HDRI sampling tables on the
 * GPU (float64 prefix sums; replaces hdridome.cu:15-51). */
int clhip_build_emitters(clhip_ctx* ctx, const clhip_emitter* emitters, uint32_t num_emitters, int32_t dome_index);

/* Integrator::updateCamera (integrator.h:80; core::Camera, librender/
 * nodebase/camera.h:8-27).  Resets the progressive accumulation when the
 * camera differs from the current one (path.h:73-80). */
int clhip_set_camera(clhip_ctx* ctx, const float pos[3], const float d00[3], const float ddu[3], const float ddv[3]);
/* Path depth, Russian-roulette depth, RNG view lane and batch size.  A
 * change of depth, RR depth or lane restarts the accumulation; the batch
 * size alone does not (it changes no sample). */
int clhip_set_params(clhip_ctx* ctx, const clhip_params* params);

/* Integrator::render (integrator.h:61; path.cpp:84-231): adds `spp`
 * progressive samples per pixel.  Asynchronous on the context stream. */
int clhip_render(clhip_ctx* ctx, uint32_t spp);
int clhip_sync(clhip_ctx* ctx);

/* Integrator::resize / resetIterationIndex / getFilmSize
 * (integrator.h:64,75,77; path.cpp:233-250) */
int clhip_resize(clhip_ctx* ctx, uint32_t width, uint32_t height);
int clhip_reset(clhip_ctx* ctx);
int clhip_get_film_size(const clhip_ctx* ctx, uint32_t* width, uint32_t* height, uint32_t* spp);

/* Integrator::readbackFramebuffer (integrator.h:73; path.cpp:271-279):
 * linear HDR RGBA float32 = per-pixel float64 sum / spp (replaces the
 * running mean and sRGB round trip, raygeneration.cu:75-81, 498-506). */
int clhip_readback(clhip_ctx* ctx, float* rgba, size_t bytes);
/* Integrator::readbackFramebuffer as the reference returns it: the film
 * resolved as above and then encoded per channel by convertFromLinearTosRGB
 * (math.h:114-122, applied by postprocessing, raygeneration.cu:498-506),
 * alpha 1.  Only this display copy is encoded; the float64 film stays
 * linear, so progressive rendering continues unaffected (the reference
 * decodes its buffer again at raygeneration.cu:80-81). */
int clhip_readback_srgb(clhip_ctx* ctx, float* rgba, size_t bytes);
/* The raw float64 RGB sums (3 * width * height doubles). */
int clhip_readback_sums(clhip_ctx* ctx, double* rgb, size_t bytes);

/* Counters of the last render (queue counts, node/triangle fetches, kernel
 * times).  timing != 0 records HIP events around every traversal and
 * shading launch from the next render on. */
int clhip_counters(clhip_ctx* ctx, clhip_stats* out);
int clhip_set_timing(clhip_ctx* ctx, int timing);

/* ---- batch entry points on the same kernels (parity surfaces) ---- */

/* Sampler (libkernel/base/sampler.h:16-78): for each seed (px, py, frame, w0)
 * emit ndraws successive next1D() draws (u32 hash word and float). */
int clhip_rng_batch(clhip_ctx* ctx, const uint32_t* seeds4, uint32_t n, uint32_t ndraws, uint32_t* out_u32,
                    float* out_f);

/* generateCameraRays (raygeneration.cu:59-103) for `n` (pixel, sample)
 * pairs of the current film/camera: out8 = o.xyz, 0, d.xyz, 0 per ray. */
int clhip_camera_rays(clhip_ctx* ctx, const uint32_t* pixel_sample2, uint32_t n, float* out8);

/* BSDF::eval/pdf/sample (libkernel/base/bsdf.h:48-96) for materials of the
 * current material table: per item material index, local wi (eval), wo,
 * u3 = (u.x, u.y, lobe draw).  Outputs: eval f3 + pdf at (wi, wo); sampled
 * wi3, f3, pdf. */
int clhip_bsdf_batch(clhip_ctx* ctx, const uint32_t* material, const float* wi3, const float* wo3, const float* u3,
                     uint32_t n, float* eval_f3, float* eval_pdf, float* smp_wi3, float* smp_f3, float* smp_pdf);

/* LightSampler::sampleEmitterDirect (emitter.h:168-202) with the current
 * emitters: per u2 -> direction, selection-scaled value, pdf, emitter index. */
int clhip_emitter_batch(clhip_ctx* ctx, const float* u2, uint32_t n, float* dir3, float* value3, float* pdf,
                        uint32_t* which);

/* HDRIDome::evalEnvironment + pdfDirect (hdridome.h:248-297). */
int clhip_dome_eval_batch(clhip_ctx* ctx, const float* dir3, uint32_t n, float* le3, float* pdf);

/* convertFromLinearTosRGB (decode = 0) or convertsRGBToLinear (decode != 0)
 * (math.h:114-140) on n single channels. */
int clhip_srgb_batch(clhip_ctx* ctx, const float* in, uint32_t n, int decode, float* out);

/* The HDRI tables built by clhip_build_emitters: (w+1)*h, (w+1)*h, h+1, h+1. */
int clhip_dome_tables(clhip_ctx* ctx, float* pcond, float* cdfcond, float* pv, float* cdfv, uint32_t* w, uint32_t* h);

/* Closest hit (optix.cu:38-75 + closest-hit/miss programs) for rays8 =
 * o.xyz, tmin, d.xyz, tmax.  Miss: t = -1, inst = prim = 0xffffffff. */
int clhip_trace_batch(clhip_ctx* ctx, const float* rays8, uint32_t n, float* t, float* b12, uint32_t* inst,
                      uint32_t* prim);

/* Any hit (optix.cu:77-110 with terminate-on-first-hit). */
int clhip_occlusion_batch(clhip_ctx* ctx, const float* rays8, uint32_t n, uint8_t* occluded);

/* Closest hit (any = 0) or any hit (any != 0) for n rays already in device
 * memory, writing device memory, with no allocation of any kind: for
 * callers that cannot allocate (code under HIP graph capture, serving loops
 * that trace on their own schedule, hosts whose allocator policy forbids
 * hipMalloc).  No counterpart in the reference.  rays8 = n x (o.xyz, tmin,
 * d.xyz, tmax); hit4 = n x (t or -1, b1, b2, instance id as uint bits) for
 * closest hit, or (1 if occluded else 0, 0, 0, 0) for any hit; prim = n
 * primitive ids (0xffffffff on a miss; 0 for any hit).  Results equal
 * clhip_trace_batch / clhip_occlusion_batch.  Asynchronous on the context
 * stream (clhip_sync waits); the buffers must stay valid until then.
 * Traversal stacks spill into a scratch array fixed in the library (58 MiB
 * per device, held until the process ends), shared by every call on that
 * device: calls that overlap on one GPU race.  At most
 * CLHIP_TRACE_RAYS_DEVICE_MAX rays per call; the TLAS must be built. */
#define CLHIP_TRACE_RAYS_DEVICE_MAX 0x80000000u
int clhip_trace_rays_device(clhip_ctx* ctx, const float* rays8, uint32_t n, int any, float* hit4, uint32_t* prim);

/* World-to-object matrices the context computed for its instances
 * (num_instances * 12 floats). */
int clhip_instance_inverses(clhip_ctx* ctx, float* out12, uint32_t n);

/* ---- batched multi-view rendering.  No reference counterpart: the
 *      reference renders one camera per Integrator::render (path.cpp:84-231).
 *      Many views of the current scene go through one sequence of
 *      persistent launches; paths of several views share each queue batch.
 *      Callers: view servers, probe/lightmap bakers, multi-camera capture. ---- */
typedef struct
{
    uint32_t width, height; /* film of this view */
    uint32_t spp;           /* samples per pixel, 1..2^20 */
    uint32_t max_depth;     /* 1..CLHIP_MAX_DEPTH */
    uint32_t rr_depth;      /* Russian roulette from this depth (0 = off) */
    uint32_t lane;          /* RNG view lane < 4096: z = (lane << 20) | sample */
    float    pos[3], d00[3], ddu[3], ddv[3]; /* camera as in clhip_set_camera */
} clhip_view;

/* Launch shape of the bounce loop.  Zero in a field selects the default.
 * Results do not depend on any field (up to float64 summation order). */
typedef struct
{
    uint32_t trace_blocks_per_cu;  /* persistent closest-hit grid (256-thread blocks per CU) */
    uint32_t shadow_blocks_per_cu; /* persistent any-hit grid */
    uint32_t shade_blocks_per_cu;  /* camera and shading grids */
    uint32_t paths_per_batch;      /* paths in flight per batch (0 = 2^27; clhip_render_views raises it to the
                                      largest view; clhip_render uses whole samples of the film) */
    uint32_t graph;                /* 1: replay a captured hipGraph per batch; 2: plain launches */
    uint32_t order;                /* clhip_render_views: 1 = given order, 2 = least work first */
    uint32_t shade_block;          /* threads per block of the camera and shading kernels: 64, 128 or 256 (0 = 256) */
    uint32_t tile;                 /* path order within a sample of a view: 1 = pixel rows, or square tiles of
                                      2, 4, 8, 16 or 32 pixels on a side, row by row within a tile (0 = 16) */
} clhip_launch;

int clhip_set_launch(clhip_ctx* ctx, const clhip_launch* launch);
int clhip_get_launch(const clhip_ctx* ctx, clhip_launch* out); /* effective values */

/* Renders `spp` samples of every view into a film of its own (cleared on
 * each call; the single-view film is untouched).  Asynchronous. */
int clhip_render_views(clhip_ctx* ctx, const clhip_view* views, uint32_t num_views);
/* View v of the last clhip_render_views: RGBA float32 = sums / spp, or the
 * raw float64 RGB sums. */
int clhip_view_readback(clhip_ctx* ctx, uint32_t view, float* rgba, size_t bytes);
int clhip_view_sums(clhip_ctx* ctx, uint32_t view, double* rgb, size_t bytes);
/* Milliseconds from the start of the last clhip_render_views until the
 * batch holding view v's last path finished, per view (synchronises). */
int clhip_view_latency(clhip_ctx* ctx, double* ms, uint32_t num_views);

/* ---- learned directional proposal (path guiding).  No reference
 *      counterpart.  A 256-64-64-64-64-64 MLP (bfloat16 on the matrix
 *      cores) maps (position in the scene box, normal, outgoing direction)
 *      to a mixture of 8 von Mises-Fisher lobes; at diffuse and conductor
 *      vertices the continuation is drawn from
 *      bsdf_fraction * p_bsdf + (1 - bsdf_fraction) * p_mlp and weighted by
 *      that mixture density (one-sample MIS), which also enters the MIS
 *      weights against HDRI sampling.  Training is online: recorded paths
 *      give (query, direction, lum(L f cos) / p) samples, and Adam steps on
 *      the negative log-likelihood follow every batch. ---- */
typedef struct
{
    uint32_t enable;        /* 0: plain tracer */
    float    bsdf_fraction; /* (0, 1]: probability of BSDF sampling at guided vertices */
    uint32_t train;         /* 1: one training step after every batch */
    uint32_t record_paths;  /* paths per batch that record training vertices (0 = 32768) */
    uint32_t batch_size;    /* samples per training step, multiple of 1024 (0 = 65536) */
    float    learning_rate; /* Adam (0 = 1e-3) */
    uint32_t max_steps;     /* stop training after this many steps (0 = never) */
    uint32_t precision;     /* inference: 0 = bfloat16, 1 = FP8 e4m3 with per-layer scales (ablation) */
    uint32_t steps_per_batch; /* Adam steps after each batch, each on its own batch_size samples (0 = 1) */
    uint32_t train_kernel;  /* 0: fused step (one kernel from encoding to weight gradients, one for the
                               reduction and Adam); 1: one launch per GEMM (the comparison path) */
} clhip_guide;

typedef struct
{
    uint64_t steps;         /* Adam steps taken since the last reset */
    uint64_t samples;       /* samples trained on since the last reset */
    uint32_t last_count;    /* samples in the last step */
    float    last_loss;     /* weighted negative log-likelihood of the last step (per unit weight) */
    double   train_ms;      /* summed device time of the training steps of the last render */
    uint64_t issued;        /* Adam steps issued since the last reset (steps counts those with samples) */
    uint32_t training;      /* 1 while the next batch records and trains (enabled, train, below max_steps) */
} clhip_guide_stats;

int clhip_set_guide(clhip_ctx* ctx, const clhip_guide* guide);
/* Fresh weights (He-uniform from `seed`, output layer scaled by 0.1) and
 * Adam state. */
int clhip_guide_reset(clhip_ctx* ctx, uint64_t seed);
/* The float32 master weights (32768 values, layer l row-major [in][out]). */
int clhip_guide_get_weights(clhip_ctx* ctx, float* weights, uint32_t n);
int clhip_guide_set_weights(clhip_ctx* ctx, const float* weights, uint32_t n);
int clhip_guide_stats_get(clhip_ctx* ctx, clhip_guide_stats* out);
/* The FP8 scales (10 floats: per layer the activation scale and the
 * accumulator multiplier) of the current e4m3 image. */
int clhip_guide_fp8_scales(clhip_ctx* ctx, float* scales10);
/* Network outputs (40 per query) for n queries of 9 floats, at the
 * inference precision of clhip_set_guide. */
int clhip_guide_infer(clhip_ctx* ctx, const float* q9, uint32_t n, float* out40);
/* Test surfaces of the functions the guided shading kernel uses.  The
 * mixture (softmax weights, normalised means, kappa = exp(clamp(s, -4, 8)))
 * of n rows of 40 network outputs: its density at dir3, a direction drawn
 * with u3 = (lobe draw, u1, u2) and the density there.  And the network
 * query of n surface points: (position - lo) / (hi - lo) per axis over the
 * scene's bounds (box6 = lo[3], hi[3], optional), shading normal, wo. */
int clhip_guide_mixture_batch(clhip_ctx* ctx, const float* out40, const float* dir3, const float* u3, uint32_t n,
                              float* pdf, float* smp_dir3, float* smp_pdf);
int clhip_guide_query_batch(clhip_ctx* ctx, const float* p3, const float* n3, const float* wo3, uint32_t n, float* q9,
                            float box6[6]);
/* One Adam step on n samples of 13 floats (query 9, direction 3, weight);
 * grad (32768 floats, optional) receives the summed gradient. */
int clhip_guide_train(clhip_ctx* ctx, const float* samples13, uint32_t n, float* grad, float* loss);
/* Test surface (no reference counterpart): the training records of the
 * last recorded batch and the samples gathered from them, in gather order.
 * info = {recorded paths S, depth planes D, sample capacity}.  V holds
 * 4 x D x S float4 (plane p, depth j, path s at ((p D + j) S + s)), E holds
 * D x S float4 (path contributions per depth), samples13 the first `cap`
 * gathered samples (query 9, direction 3, weight).  NULL buffers: info only. */
int clhip_guide_records(clhip_ctx* ctx, float* V, float* E, float* samples13, uint32_t cap, uint32_t info[3]);

/* ---- image files (ImageUtils, src/delegate/imageutil.cpp:13-139, without
 *      stb or FreeImage).  Images are RGBA float32, row 0 at the top.
 *      Loaders: call with rgba = NULL to get width and height, then with a
 *      buffer of at least 16 * w * h bytes.  Errors: clhip_image_last_error. */
const char* clhip_image_last_error(void);
/* Radiance RGBE .hdr: flat or run-length scanlines, "-Y H +X W"; value =
 * c * 2^(e-136), alpha 1 (the reference loader's conventions). */
int clhip_image_load_hdr(const char* path, float* rgba, size_t bytes, uint32_t* w, uint32_t* h);
int clhip_image_save_hdr(const char* path, const float* rgba, uint32_t w, uint32_t h);
/* PFM: little-endian float32 RGB, rows bottom to top. */
int clhip_image_load_pfm(const char* path, float* rgba, size_t bytes, uint32_t* w, uint32_t* h);
int clhip_image_save_pfm(const char* path, const float* rgba, uint32_t w, uint32_t h);
/* OpenEXR: uncompressed scanline FLOAT RGBA (the reference saves EXR via
 * FreeImage, imageutil.cpp:31-79). */
int clhip_image_save_exr(const char* path, const float* rgba, uint32_t w, uint32_t h);

/* ---- mesh import (MeshImporter, src/delegate/meshimporter.cpp:41-121,
 *      without assimp): Wavefront OBJ, polygons fanned into triangles, one
 *      vertex per distinct (v, vt, vn) corner, area-weighted smooth normals
 *      when the file has none, flipped texture v, tangents from the UV
 *      parametrisation (Gram-Schmidt against the normal).  Call with
 *      vertices = NULL to get the counts, then with arrays of 3 * nv floats
 *      (vertices, normals, tangents), 2 * nv (uvs) and 3 * nt indices;
 *      tangents and uvs are written only when *has_uvs. */
const char* clhip_mesh_last_error(void);
int clhip_mesh_load_obj(const char* path, float* vertices, float* normals, float* tangents, float* uvs,
                        uint32_t* indices, uint32_t* num_vertices, uint32_t* num_triangles, uint32_t* has_uvs);

/* Measurement helper with no reference counterpart: `iters` launches of a
 * plain 16-byte streaming copy of `bytes` bytes; returns ms per copy.  Used
 * as the known-traffic kernel for HBM byte counters and the copy ceiling. */
int clhip_copy_probe(clhip_ctx* ctx, size_t bytes, uint32_t iters, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
