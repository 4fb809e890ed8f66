/*
 * clpt_scene.h -- device-resident scene and wavefront layout (host/device
 * PODs).  Everything lives in HBM as flat arrays addressed by 32-bit
 * indices; all per-ray state is structure-of-arrays in float4 planes so a
 * wave64 reads and writes whole 1 KiB lines.
 */
#ifndef CLPT_SCENE_H
#define CLPT_SCENE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* ---- BVH4 node: 128 bytes, one L2 line pair, loaded as 8 dwordx4.
 * Child boxes are stored per axis (SoA inside the node). ---- */
struct alignas(128) Node4
{
    float    lox[4], hix[4];
    float    loy[4], hiy[4];
    float    loz[4], hiz[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 is two 64-byte halves");

/* ---- BVH8 node: 96 bytes, read as 6 dwordx4.  Child boxes are stored
 * relative to the node's lower corner p, quantised to bytes per axis with a
 * power-of-two scale 2^(e - 127): child c spans
 *   [p + qlo[a][c] * 2^(e_a - 127), p + qhi[a][c] * 2^(e_a - 127)]
 * on axis a, evaluated in float32 exactly as written (a product of a byte
 * and a power of two is exact; the sum rounds once).  The builder picks the
 * bytes so that the decoded box contains the child's true box.  Empty slots
 * have child = CLPT_CHILD_EMPTY. ---- */
struct alignas(16) Node8
{
    float    p[3];
    uint8_t  e[3];
    uint8_t  pad;
    uint8_t  qlo[3][8];
    uint8_t  qhi[3][8];
    uint32_t child[8];
};
static_assert(sizeof(Node8) == 96, "Node8 is six 16-byte words");

/* Child reference encoding. */
#define CLPT_CHILD_EMPTY 0xffffffffu
#define CLPT_LEAF_BIT 0x80000000u
#define CLPT_INST_BIT 0x40000000u
/* triangle leaf: LEAF | (count-1) << 27 | first triangle (27 bits) */
#define CLPT_TRI_LEAF(first, count) (CLPT_LEAF_BIT | ((uint32_t)((count)-1) << 27) | (uint32_t)(first))
#define CLPT_INST_LEAF(idx) (CLPT_LEAF_BIT | CLPT_INST_BIT | (uint32_t)(idx))
/* Mesh-tree leaves hold at most 2 triangles: the closest-hit kernel tests
 * one triangle per step beside its node visits, and 2 measured slightly
 * faster than 4 (1 costs 4% more node visits). */
#define CLPT_MAX_LEAF_TRIS 2u

/* Traversal stack: CLPT_LDS_STACK entries per lane in LDS, the rest (up to
 * CLPT_STACK_TOTAL) in a global overflow area of (ref, t) pairs. */
#define CLPT_LDS_STACK 12
#define CLPT_STACK_TOTAL 128
#define CLPT_GSTACK_ENTRIES (CLPT_STACK_TOTAL - CLPT_LDS_STACK)
#define CLPT_MAX_TRIS (1u << 27)
#define CLPT_MAX_INSTANCES 0x3ffffff0u

/* Triangle in leaf order: v0.xyz + original primitive index (bits), v1, v2. */
struct TriRec
{
    float4 v0p;
    float4 v1;
    float4 v2;
};

struct InstanceRec
{
    float    xf[12];  /* object -> world */
    float    inv[12]; /* world -> object */
    uint32_t blas_root;
    uint32_t mesh;
    uint32_t material;
    uint32_t entity;
};

enum : uint32_t
{
    MESH_HAS_NORMALS  = 1u,
    MESH_HAS_TANGENTS = 2u,
    MESH_HAS_UVS      = 4u
};

struct MeshRec
{
    uint32_t idx_off; /* first index (x3) in the global index array */
    uint32_t vtx_off; /* first vertex in the global vertex arrays */
    uint32_t ntris;
    uint32_t flags;
};

/* Material: type, reflectance, GGX alpha = max(roughness^2, 1e-4), eta, k,
 * ior, and texture ids + 1 (0 = none) for the reflectance, the roughness
 * and the normal map. */
struct alignas(16) MaterialRec
{
    uint32_t type;
    float    r[3];
    float    alpha;
    float    eta[3];
    float    k[3];
    float    ior;
    uint32_t tex_r, tex_rough, tex_n, pad;
};
static_assert(sizeof(MaterialRec) == 64, "MaterialRec is four 16-byte words");

/* An RGBA32F texture: texels row-major, row 0 at v = 0. */
struct alignas(16) TexRec
{
    const float4* data;
    uint32_t      w, h;
};

struct alignas(16) EmitterRec
{
    uint32_t type;
    float    intensity[3];
    float    dir[3];
    float    tan_r; /* tan(angular radius) rounded once on the host */
};

/* Buckets of the HDRI CDF search cut points (SceneView::cutv, cutcond). */
#define CLPT_CUTS 1024u

/* Everything a kernel needs to find the scene; passed by value. */
struct SceneView
{
    const void*        nodes; /* Node4 or Node8 array (the tree width is a kernel template argument) */
    const TriRec*      tris;
    const InstanceRec* inst;
    const MeshRec*     meshes;
    const float*       pos; /* xyz per vertex */
    const float*       nrm;
    const float*       tan;
    const float*       uv;
    const uint32_t*    idx;
    const MaterialRec* mtl;
    const TexRec*      tex; /* material textures by id */
    uint32_t           ntex;
    const EmitterRec*  emitters;
    uint32_t           nemitters;
    int32_t            dome_index;
    const float4*      dome_tex;
    uint32_t           dome_w, dome_h;
    const float*       pcond;
    const float*       cdfcond;
    const float*       pv;
    const float*       cdfv;
    const uint32_t*    cutv;    /* CLPT_CUTS + 1 search cut points of cdfv */
    const uint32_t*    cutcond; /* CLPT_CUTS + 1 per row of cdfcond */
    uint32_t           tlas_root;
    uint32_t           world_inst; /* the instance whose mesh subtrees the top tree holds (its
                                      transform is the identity), or 0xffffffff */
    uint32_t           ninst;
    uint32_t           bvh_width; /* 4 (Node4) or 8 (Node8) */
    uint32_t           node_bytes; /* size of the node array (below 4 GiB) */
};

/* One camera view: film size, path depth, RNG view lane, the view's offset
 * in the float64 film, and the camera basis (core::Camera, camera.h:8-27). */
struct alignas(16) ViewRec
{
    uint32_t width, height, npix, max_depth;
    uint32_t rr_depth, lane, film_offset, pad;
    float    pos[4], d00[4], ddu[4], ddv[4];
};

/* A batch is a list of (view, first path, first sample) entries; path p of
 * the batch belongs to the last entry with path_begin <= p, local index
 * l = p - path_begin maps to pixel l % npix and sample sample_base + l / npix.
 * The batch to run is read from *cursor, which k_batch_end advances, so a
 * captured batch graph replays unchanged. */
struct BatchEntry
{
    uint32_t view, path_begin, sample_base, pad;
};

struct BatchRec
{
    uint32_t entry_off, nentries, npaths, pad;
};

struct BatchTable
{
    const ViewRec*    views;
    const BatchRec*   batches;
    const BatchEntry* entries;
    uint32_t*         cursor;
    uint32_t          tile; /* 1: paths of a sample in pixel rows; T > 1: in T x T tiles */
};

/* A ray queue: 3 float4 planes + count (device). */
struct RayQueue
{
    float4*   o;  /* origin.xyz, w unused (read by traversal only) */
    float4*   d;  /* direction.xyz, path id (bits) */
    float4*   T;  /* throughput.xyz, previous BSDF pdf (read by shading only; not
                     written for camera paths, whose throughput is 1 and pdf 0) */
    uint32_t* count;
};

/* Closest hit of each ray: b1, b2, instance (bits; CLPT_NO_HIT for a miss),
 * original primitive index (bits).  Shading needs no hit distance. */
#define CLPT_NO_HIT 0xffffffffu
struct HitBuf
{
    float4* h;
};

struct ShadowQueue
{
    float4*   o; /* origin.xyz, Lo.r */
    float4*   d; /* direction.xyz, Lo.g */
    float2*   x; /* Lo.b, pixel (bits) */
    uint32_t* count;
    uint32_t* tag; /* training record of the shadow ray (guided shading only; else NULL) */
};

/* ---- learned directional proposal ---- */

/* A training sample: query (9), sampled direction (3), weight. */
struct GuideSample
{
    float4 a; /* q0..q3 */
    float4 b; /* q4..q7 */
    float4 c; /* q8, wi.xyz */
    float4 d; /* weight, unused */
};

/* Device state of online training. */
struct GuideCounters
{
    uint32_t           count;   /* samples gathered since the last step (may exceed the buffer) */
    uint32_t           trained; /* samples in the last step */
    uint32_t           steps;   /* Adam steps taken */
    float              wsum;    /* weight sum of the last step */
    float              loss;    /* weighted negative log-likelihood of the last step (per unit weight) */
    uint32_t           pad;
    unsigned long long samples_total;
    uint32_t           amax[4]; /* max hidden activation of each layer in the last step (float bits) */
};

/* What the guided shading kernels need.  Training records, per vertex depth
 * j and recorded path (slot), at rec = j * nslots + slot: V is 4 planes of
 * float4, plane k at k * (CLPT_MAXD + 1) * nslots (query q0..q8, the
 * sampled direction, then p_mix and R = f cos / T_next per channel; p_mix =
 * 0 marks no record), and E[rec] the contribution (rgb) that arrived at
 * depth j (escape of ray j or the shadow ray of vertex j). */
struct GuideView
{
    const __bf16*  img;  /* packed bf16 weights */
    const uint8_t* img8; /* packed e4m3 weights (FP8 inference), else NULL */
    const float*   scl;  /* FP8 scales (10 floats) */
    float         lo[3], inv_ext[3];
    float         c_bsdf; /* probability of BSDF sampling at guided vertices */
    uint32_t      nslots; /* recorded paths per batch (0 = no recording) */
    float4*       V;
    float4*       E;
    uint32_t*     gcount; /* gather scratch: samples per recorded path */
    uint32_t*     gblock; /* gather scratch: per-256-path block counts, then offsets */
};

/* Per-batch exact counters (device): queue counts per depth and traversal
 * work.  Layout of the u32 array: [0..MAXD] ray counts (set by the producer
 * of each queue), [MAXD+1 ..] hits, shadows, visible. */
#define CLPT_MAXD 32
/* Per-batch queue counts.  The rays of depth d + 1 and the shadow rays of
 * depth d are interleaved (rs[2 d] = shadow rays of depth d, rs[2 d + 1] =
 * rays of depth d + 1), so that shading reserves both of its output queues
 * with one 64-bit atomic on &rs[2 d]. */
struct alignas(8) BatchCounters
{
    uint32_t rs[2 * (CLPT_MAXD + 2)];
    uint32_t hits[CLPT_MAXD + 2];
    uint32_t visible[CLPT_MAXD + 2];
};

/* rays of depth d >= 1, shadow rays of depth d */
__host__ __device__ inline uint32_t* bc_rays(BatchCounters* bc, uint32_t d) { return &bc->rs[2 * d - 1]; }
__host__ __device__ inline uint32_t* bc_shadow(BatchCounters* bc, uint32_t d) { return &bc->rs[2 * d]; }
__host__ __device__ inline unsigned long long* bc_push(BatchCounters* bc, uint32_t d)
{
    return reinterpret_cast<unsigned long long*>(&bc->rs[2 * d]);
}

struct TotalCounters
{
    unsigned long long rays[CLPT_MAXD + 2];
    unsigned long long hits[CLPT_MAXD + 2];
    unsigned long long shadow[CLPT_MAXD + 2];
    unsigned long long visible[CLPT_MAXD + 2];
    unsigned long long trace_nodes, trace_tris, shadow_nodes, shadow_tris;
};

#endif
