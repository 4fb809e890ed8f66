/*
 * oracle.h -- CPU restatement of Colvillea's wavefront path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * renderer (colvillea_amd/, libclpt_hip.so) is compared against.  Only the
 * test suite, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it; the renderer never links, imports or falls back to it.
 *
 * Every routine restates a reference routine (file:line under the
 * reference tree cited at each definition in orc_real.h / orc_common.c),
 * with the reference's defects D1-D16 corrected in the same way as the
 * renderer does (named at each site).  Two numeric variants are built from
 * one source (orc_real.h):
 *   *_f32 : float32, -ffp-contract=off, the renderer's own elementary
 *           functions (colvillea_amd/csrc/clpt_elem.h) -> bit-comparable to
 *           the GPU for identical inputs;
 *   *_f64 : float64 with libm -> the image-level reference (relMSE).
 *
 * The reference itself cannot be compiled here (it needs OWL's vec.h and
 * CUDA headers that the image lacks), so this restatement is pinned by
 * closed forms, invariants, independent textbook formulas and brute force
 * in tests/test_oracle_*.py.
 */
#ifndef CLPT_ORACLE_H
#define CLPT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    const float*    vertices; /* 3*nv */
    const uint32_t* indices;  /* 3*nt */
    const float*    normals;  /* 3*nv or NULL */
    const float*    tangents; /* 3*nv or NULL */
    const float*    uvs;      /* 2*nv or NULL */
    uint32_t        nv, nt;
} orc_mesh;

typedef struct
{
    uint32_t mesh, material;
    float    xform[12]; /* object->world, row-major 3x4 */
} orc_instance;

enum
{
    ORC_MTL_DIFFUSE = 0,
    ORC_MTL_METAL   = 1,
    ORC_MTL_GLASS   = 2
};

typedef struct
{
    uint32_t type;
    float    reflectance[3]; /* diffuse reflectance / metal specular reflectance */
    float    roughness;      /* metal, glass: GGX alpha = roughness^2 */
    float    eta[3], k[3];   /* metal complex IOR */
    float    ior;            /* glass interior IOR (exterior = 1) */
    uint32_t reflectance_tex, roughness_tex, normal_tex; /* texture index + 1, 0 = none */
} orc_material;

/* An RGBA float32 texture (row 0 at v = 0) for material lookups. */
typedef struct
{
    const float* rgba;
    uint32_t     w, h;
} orc_texture;

enum
{
    ORC_EMITTER_DIRECTIONAL = 0,
    ORC_EMITTER_DOME        = 1
};

typedef struct
{
    uint32_t type;
    float    intensity[3];
    float    direction[3]; /* emitted direction (from the sun towards the scene) */
    float    angular_radius;
} orc_emitter;

typedef struct
{
    const float* rgba; /* w*h*4 */
    uint32_t     w, h;
    const float* pcond;   /* (w+1)*h */
    const float* cdfcond; /* (w+1)*h */
    const float* pv;      /* h+1 */
    const float* cdfv;    /* h+1 */
} orc_dome;

typedef struct
{
    const orc_mesh*     meshes;
    uint32_t            nmeshes;
    const orc_instance* instances;
    uint32_t            ninstances;
    const orc_material* materials;
    uint32_t            nmaterials;
    const orc_emitter*  emitters;
    uint32_t            nemitters;
    int32_t             dome_index; /* index into emitters, -1 = no dome */
    orc_dome            dome;
    const orc_texture*  textures; /* material textures (kept alive by the caller) */
    uint32_t            ntextures;
} orc_scene_desc;

typedef struct
{
    uint32_t width, height;
    float    cam_pos[3], cam_d00[3], cam_ddu[3], cam_ddv[3];
    uint32_t max_depth, rr_depth;
    uint32_t view; /* RNG z lane = (view << 20) | sample */
} orc_render_params;

/* Per-bounce counters: index b = path depth (1..max_depth). */
#define ORC_MAX_DEPTH 64
typedef struct
{
    uint64_t rays[ORC_MAX_DEPTH + 1];    /* closest-hit rays traced at depth b */
    uint64_t hits[ORC_MAX_DEPTH + 1];    /* of which hit geometry */
    uint64_t shadow[ORC_MAX_DEPTH + 1];  /* shadow rays spawned at vertex depth b */
    uint64_t visible[ORC_MAX_DEPTH + 1]; /* of which unoccluded */
} orc_counts;

typedef struct orc_scene orc_scene;

/* ---- RNG (sampler.h:16-55) ---- */
void  orc_pcg4d(const uint32_t in[4], uint32_t out[4]);
float orc_uint_to_float(uint32_t x);
float orc_draw(uint32_t px, uint32_t py, uint32_t zlane, uint32_t k);
void  orc_rng_batch(const uint32_t* seeds4, uint32_t n, uint32_t ndraws, uint32_t* out_u32, float* out_f);

/* ---- the renderer's elementary functions, for pinning against libm ---- */
void orc_elem_batch(int fn, const float* x, const float* y, uint32_t n, float* out);

/* ---- instance transform inverse (shared definition with the renderer) ---- */
void orc_inverse3x4(const float m[12], float out[12], double outd[12]);

/* ---- HDRI tables ---- */
void orc_dome_tables_ref32(const float* rgba, uint32_t w, uint32_t h, float* pcond, float* cdfcond, float* pv, float* cdfv);
void orc_dome_tables_f64(const float* rgba, uint32_t w, uint32_t h, float* pcond, float* cdfcond, float* pv, float* cdfv);

/* ---- scene ---- */
orc_scene* orc_scene_create(const orc_scene_desc* desc);
void       orc_scene_destroy(orc_scene* s);

/* ---- per-variant batch entry points (R = f32 | f64) ---- */
#define ORC_DECLARE_VARIANT(SFX)                                                                                      \
    void orc_bsdf_eval_##SFX(const orc_material* m, const float* wi, const float* wo, uint32_t n, float* f3,          \
                             float* pdf);                                                                             \
    void orc_bsdf_sample_##SFX(const orc_material* m, const float* wo, const float* u3, uint32_t n, float* wi,       \
                               float* f3, float* pdf);                                                                \
    void orc_warp_batch_##SFX(const float* u2, uint32_t n, float* disk2, float* hemi3);                              \
    void orc_ggx_batch_##SFX(float alpha, const float* m3, const float* v3, const float* u2, uint32_t n, float* D,    \
                             float* G1, float* msample3, float* mpdf);                                                \
    void orc_fresnel_batch_##SFX(const float* cos_i, uint32_t n, const float* eta3, const float* k3, float ior,       \
                                 float* fc3, float* fd, float* cos_t);                                                \
    void orc_tex_eval_##SFX(const float* rgba, uint32_t w, uint32_t h, const float* uv, uint32_t n, float* out4);    \
    void orc_srgb_##SFX(const float* in, uint32_t n, int decode, float* out);                                        \
    void orc_pull_up_##SFX(const float* frame9, const float* ng3, const float* wo3, uint32_t n, float* out9);        \
    void orc_emitter_sample_##SFX(const orc_scene* s, const float* u2, uint32_t n, float* dir3, float* value3,        \
                                  float* pdf, uint32_t* which);                                                       \
    void orc_dome_eval_##SFX(const orc_scene* s, const float* dir3, uint32_t n, float* le3, float* pdf);             \
    void orc_trace_##SFX(const orc_scene* s, const float* rays8, uint32_t n, int brute, float* t, float* b12,        \
                         uint32_t* inst, uint32_t* prim);                                                             \
    void orc_occluded_##SFX(const orc_scene* s, const float* rays8, uint32_t n, int brute, uint8_t* occ);            \
    void orc_render_##SFX(const orc_scene* s, const orc_render_params* p, uint32_t y0, uint32_t y1, uint32_t s0,     \
                          uint32_t s1, uint32_t nthreads, double* sum3, orc_counts* counts);                  \
    void orc_render_rows_##SFX(const orc_scene* s, const orc_render_params* p, const uint32_t* rows, uint32_t nrows, \
                               uint32_t s0, uint32_t s1, uint32_t nthreads, double* sum3, orc_counts* counts);

ORC_DECLARE_VARIANT(f32)
ORC_DECLARE_VARIANT(f64)

#ifdef __cplusplus
}
#endif
#endif
