/*
 * kernels.hip -- gfx950 wavefront path-tracing kernels.
 *
 * Pipeline of one batch of paths (replaces the 30 synchronous launches per
 * spp of WavefrontPathTracingIntegrator::render, path.cpp:84-231):
 *   k_camera                         generateCameraRays (raygeneration.cu:59-103)
 *   for depth = 1..max_depth:
 *     k_trace   closest hit          OptiX primaryRay raygen + closest-hit/miss (optix.cu:38-75, 112-262)
 *     k_shade   escape + NEE + BSDF  evaluateEscapedRays (raygeneration.cu:105-136) and
 *                                    evaluateMaterialsAndLightsPathTracing (:285-456)
 *     k_shadow  any hit + accumulate OptiX shadowRay raygen + miss (optix.cu:77-110, 265-284)
 *   k_batch_end                      per-batch counters -> totals
 * All kernels are grid-stride loops over device-side queue counts, so a
 * batch needs no host synchronisation (the reference resets counters with
 * <<<1,1>>> launches and synchronises after each kernel, cudadevice.h:185-219).
 * Queue pushes are wave64-aggregated: one atomic per wave (ballot + mbcnt),
 * replacing one atomicAdd per item (soa.h:39-49).
 */
#include "clpt_device.h"
#include "clpt_guide.h"
#include "clpt_kernels.h"

#include <mutex>


/* ------------------------------------------------------------ wave utils */

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

/* Reserves slots in two queues with one atomic per wave: the wave's count
 * for the first queue goes to the low word of a 64-bit counter and for the
 * second to the high word (a count stays below 2^32, so the low word never
 * carries); each lane gets its slot in each queue. */
__device__ __forceinline__ void wave_push2(unsigned long long* counter, bool want_a, bool want_b, uint32_t& slot_a,
                                           uint32_t& slot_b)
{
    const unsigned long long ma = __ballot(want_a), mb = __ballot(want_b);
    slot_a = slot_b = 0u;
    if ((ma | mb) == 0ull)
        return;
    const unsigned long long active = __ballot(true);
    const uint32_t           leader = (uint32_t)(__ffsll((long long)active) - 1);
    unsigned long long       base   = 0ull;
    if (lane_id() == leader)
        base = atomicAdd(counter, (unsigned long long)__popcll(ma) | ((unsigned long long)__popcll(mb) << 32));
    const uint32_t lo = __shfl((uint32_t)base, (int)leader), hi = __shfl((uint32_t)(base >> 32), (int)leader);
    const unsigned long long below = (1ull << lane_id()) - 1ull;
    slot_a                         = lo + (uint32_t)__popcll(ma & below);
    slot_b                         = hi + (uint32_t)__popcll(mb & below);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void wave_count(uint32_t* counter, bool pred)
{
    const unsigned long long mask = __ballot(pred);
    if (mask != 0ull && lane_id() == (uint32_t)(__ffsll((long long)mask) - 1))
        atomicAdd(counter, (uint32_t)__popcll(mask));
}

/* A path's radiance from its escaped ray, kept by shading in a float32
 * triple per path slot of the batch (the value is a float32 product, so
 * nothing is rounded) and folded into the float64 film at the batch's end
 * (k_film_fold): depth 1 writes every path's slot (its escape radiance or
 * zero), and a path that escapes later writes its slot then.  An escape ends
 * the path, so a slot takes at most one nonzero value and is written by one
 * lane at a time: plain stores, no read, no atomics on the film. */
__device__ __forceinline__ void path_set(float* pacc, uint32_t path, f3 c)
{
    float* a = pacc + 3 * (size_t)path;
    a[0]     = c.x;
    a[1]     = c.y;
    a[2]     = c.z;
}

__device__ __forceinline__ void accum_add(double* acc, uint32_t pixel, f3 c)
{
    unsafeAtomicAdd(acc + 3 * (size_t)pixel + 0, (double)c.x);
    unsafeAtomicAdd(acc + 3 * (size_t)pixel + 1, (double)c.y);
    unsafeAtomicAdd(acc + 3 * (size_t)pixel + 2, (double)c.z);
}

/* ------------------------------------------------------------ traversal */

/* Per-lane short stack of (child reference, entry distance): the first
 * CLPT_LDS_STACK entries live in LDS laid out [entry][lane] (a wave touches
 * 64 consecutive dwords per access), deeper entries in a global overflow area
 * laid out [entry][global thread].  Entries whose entry distance is beyond the
 * current closest hit are dropped when popped, without fetching them. */
struct TraverseStack
{
    uint32_t* lds_ref;
    float*    lds_t;
    uint2*    gmem;
    uint32_t  tid, gid, gstride;
    int       sp;
    __device__ __forceinline__ void push(uint32_t ref, float t)
    {
        if (sp < CLPT_LDS_STACK)
        {
            lds_ref[sp * blockDim.x + tid] = ref;
            lds_t[sp * blockDim.x + tid]   = t;
        }
        else if (sp < CLPT_STACK_TOTAL)
            gmem[(size_t)(sp - CLPT_LDS_STACK) * gstride + gid] = make_uint2(ref, __float_as_uint(t));
        /* deeper pushes are dropped: clhip_build_tlas rejects trees whose
         * depth could reach CLPT_STACK_TOTAL */
        ++sp;
    }
    __device__ __forceinline__ void pop(uint32_t& ref, float& t)
    {
        --sp;
        if (sp < CLPT_LDS_STACK)
        {
            ref = lds_ref[sp * blockDim.x + tid];
            t   = lds_t[sp * blockDim.x + tid];
        }
        else if (sp < CLPT_STACK_TOTAL)
        {
            const uint2 e = gmem[(size_t)(sp - CLPT_LDS_STACK) * gstride + gid];
            ref           = e.x;
            t             = __uint_as_float(e.y);
        }
        else
        {
            ref = CLPT_CHILD_EMPTY;
            t   = CLPT_FLT_MAX;
        }
    }
};

#define CLPT_DECLARE_STACK(st, gstack)                                                                                 \
    __shared__ uint32_t lds_ref_[CLPT_LDS_STACK * 256];                                                                \
    __shared__ float    lds_t_[CLPT_LDS_STACK * 256];                                                                  \
    TraverseStack       st;                                                                                            \
    st.lds_ref = lds_ref_;                                                                                             \
    st.lds_t   = lds_t_;                                                                                               \
    st.gmem    = (gstack);                                                                                             \
    st.tid     = threadIdx.x;                                                                                          \
    st.gid     = blockIdx.x * blockDim.x + threadIdx.x;                                                                \
    st.gstride = gridDim.x * blockDim.x;                                                                               \
    st.sp      = 0

/* The rays of a wave's share of a persistent launch -- rays
 * w + k * stride + j (w = the wave's first global thread, j < 64) -- handed
 * to whichever lanes finish first: one ballot per refill and a wave-uniform
 * cursor, no atomics.  A lane that draws short rays takes more of them, so
 * the wave's lanes finish together instead of waiting for the lane with the
 * most work. */
struct WavePool
{
    uint32_t base, stride, next;
    __device__ __forceinline__ uint32_t index(uint32_t m) const { return base + (m >> 6) * stride + (m & 63u); }
    /* the lanes in `mask` (ballot of the finished lanes) take the next rays */
    __device__ __forceinline__ uint32_t take(unsigned long long mask)
    {
        const uint32_t m = next + (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
        return index(m);
    }
};

struct HitRec
{
    float    t, b1, b2;
    uint32_t inst, prim;
};

__device__ __forceinline__ void cswap(float& ta, uint32_t& ra, float& tb, uint32_t& rb)
{
    const bool  sw = tb < ta;
    const float t0 = sw ? tb : ta, t1 = sw ? ta : tb;
    const uint32_t r0 = sw ? rb : ra, r1 = sw ? ra : rb;
    ta = t0;
    tb = t1;
    ra = r0;
    rb = r1;
}

/* One inner node: test the children's boxes against the ray (slab test,
 * padded by 1e-5 relative so that rounding never drops a box), push the
 * hits farther than the nearest (nearest last) with their entry distance,
 * and continue at the nearest.  Returns false when no child was hit. */
template <int W>
__device__ __forceinline__ bool visit_node(const SceneView& sc, const RayW& r, float tmin, float tmax,
                                           TraverseStack& st, uint32_t& ref);

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <>
__device__ __forceinline__ bool visit_node<4>(const SceneView& sc, const RayW& r, float tmin, float tmax,
                                              TraverseStack& st, uint32_t& ref)
{
    /* Each axis's near plane is lo where the ray's reciprocal is >= 0 and hi
     * where it is negative, so the loads fetch the near and far plane of
     * every slab directly (offsets from the ray's sign bits) and no min/max
     * orders them: t = (plane - o) * inv is the same product as before, and
     * multiplying by a positive (negative) factor keeps (reverses) order.
     * The loads are buffer loads at 32-bit byte offsets into the node array
     * (the host keeps it under 4 GiB), so each plane's address is one add. */
    const __amdgpu_buffer_rsrc_t res =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(sc.nodes), (short)0, sc.node_bytes, 0x00020000);
    const uint32_t nb = ref * (uint32_t)sizeof(Node4);
    const uint32_t sx = ((r.k >> 24) & 1u) * 16u, sy = ((r.k >> 25) & 1u) * 16u, sz = ((r.k >> 26) & 1u) * 16u;
    const u32x4    nx4 = __builtin_amdgcn_raw_buffer_load_b128(res, nb + sx, 0, 0);
    const u32x4    fx4 = __builtin_amdgcn_raw_buffer_load_b128(res, nb - sx + 16u, 0, 0);
    const u32x4    ny4 = __builtin_amdgcn_raw_buffer_load_b128(res, nb + sy + 32u, 0, 0);
    const u32x4    fy4 = __builtin_amdgcn_raw_buffer_load_b128(res, nb - sy + 48u, 0, 0);
    const u32x4    nz4 = __builtin_amdgcn_raw_buffer_load_b128(res, nb + sz + 64u, 0, 0);
    const u32x4    fz4 = __builtin_amdgcn_raw_buffer_load_b128(res, nb - sz + 80u, 0, 0);
    const u32x4    ch  = __builtin_amdgcn_raw_buffer_load_b128(res, nb + 96u, 0, 0);
    float          tn[4];
    uint32_t       cr[4]  = {ch.x, ch.y, ch.z, ch.w};
    const float    n_x[4] = {__uint_as_float(nx4.x), __uint_as_float(nx4.y), __uint_as_float(nx4.z), __uint_as_float(nx4.w)};
    const float    f_x[4] = {__uint_as_float(fx4.x), __uint_as_float(fx4.y), __uint_as_float(fx4.z), __uint_as_float(fx4.w)};
    const float    n_y[4] = {__uint_as_float(ny4.x), __uint_as_float(ny4.y), __uint_as_float(ny4.z), __uint_as_float(ny4.w)};
    const float    f_y[4] = {__uint_as_float(fy4.x), __uint_as_float(fy4.y), __uint_as_float(fy4.z), __uint_as_float(fy4.w)};
    const float    n_z[4] = {__uint_as_float(nz4.x), __uint_as_float(nz4.y), __uint_as_float(nz4.z), __uint_as_float(nz4.w)};
    const float    f_z[4] = {__uint_as_float(fz4.x), __uint_as_float(fz4.y), __uint_as_float(fz4.z), __uint_as_float(fz4.w)};
#pragma unroll
    for (int c = 0; c < 4; ++c)
    {
        const float tnx = (n_x[c] - r.o.x) * r.inv.x, tfx = (f_x[c] - r.o.x) * r.inv.x;
        const float tny = (n_y[c] - r.o.y) * r.inv.y, tfy = (f_y[c] - r.o.y) * r.inv.y;
        const float tnz = (n_z[c] - r.o.z) * r.inv.z, tfz = (f_z[c] - r.o.z) * r.inv.z;
        float tnear = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        float tfar  = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        tnear       = tnear * 0.99999f;
        tfar        = tfar * 1.00001f;
        /* an empty child (inverted infinite box) gives tnear = +inf > tfar */
        tn[c]       = tnear <= tfar ? tnear : CLPT_FLT_MAX;
    }
    /* nearest first: a 5-comparator sorting network, fixed indices */
    cswap(tn[0], cr[0], tn[1], cr[1]);
    cswap(tn[2], cr[2], tn[3], cr[3]);
    cswap(tn[0], cr[0], tn[2], cr[2]);
    cswap(tn[1], cr[1], tn[3], cr[3]);
    cswap(tn[1], cr[1], tn[2], cr[2]);
    if (tn[3] < CLPT_FLT_MAX)
        st.push(cr[3], tn[3]);
    if (tn[2] < CLPT_FLT_MAX)
        st.push(cr[2], tn[2]);
    if (tn[1] < CLPT_FLT_MAX)
        st.push(cr[1], tn[1]);
    if (tn[0] < CLPT_FLT_MAX)
    {
        ref = cr[0];
        return true;
    }
    return false;
}

template <>
__device__ __forceinline__ bool visit_node<8>(const SceneView& sc, const RayW& r, float tmin, float tmax,
                                              TraverseStack& st, uint32_t& ref)
{
    const uint4* np = reinterpret_cast<const uint4*>(static_cast<const Node8*>(sc.nodes) + ref);
    const uint4  w0 = np[0], q0 = np[1], q1 = np[2], q2 = np[3], c0 = np[4], c1 = np[5];
    /* node origin and per-axis power-of-two scales */
    const float px = __uint_as_float(w0.x), py = __uint_as_float(w0.y), pz = __uint_as_float(w0.z);
    const float sx = __uint_as_float((w0.w & 0xffu) << 23);
    const float sy = __uint_as_float(((w0.w >> 8) & 0xffu) << 23);
    const float sz = __uint_as_float(((w0.w >> 16) & 0xffu) << 23);
    /* byte planes: lo x (q0.xy), lo y (q0.zw), lo z (q1.xy), hi x (q1.zw), hi y (q2.xy), hi z (q2.zw);
     * the near plane of each axis is lo where the ray's reciprocal is >= 0
     * and hi where it is negative (ray sign bits), so no min/max orders the
     * two slab distances: the same products, a monotone factor */
    const bool     nx = ((r.k >> 24) & 1u) != 0u, ny = ((r.k >> 25) & 1u) != 0u, nz = ((r.k >> 26) & 1u) != 0u;
    const uint32_t axw[2] = {nx ? q1.z : q0.x, nx ? q1.w : q0.y}, bxw[2] = {nx ? q0.x : q1.z, nx ? q0.y : q1.w};
    const uint32_t ayw[2] = {ny ? q2.x : q0.z, ny ? q2.y : q0.w}, byw[2] = {ny ? q0.z : q2.x, ny ? q0.w : q2.y};
    const uint32_t azw[2] = {nz ? q2.z : q1.x, nz ? q2.w : q1.y}, bzw[2] = {nz ? q1.x : q2.z, nz ? q1.y : q2.w};
    uint32_t       cr[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    float          tn[8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        /* plane = p + q s with q a byte and s a normal power of two (or 0):
         * q s is exact, so the fused multiply-add rounds once, as the add
         * after an exact product does -- the same bits, one instruction */
        const int   wi = c >> 2, sh = 8 * (c & 3);
        const float ax = __builtin_fmaf((float)((axw[wi] >> sh) & 0xffu), sx, px);
        const float ay = __builtin_fmaf((float)((ayw[wi] >> sh) & 0xffu), sy, py);
        const float az = __builtin_fmaf((float)((azw[wi] >> sh) & 0xffu), sz, pz);
        const float bx = __builtin_fmaf((float)((bxw[wi] >> sh) & 0xffu), sx, px);
        const float by = __builtin_fmaf((float)((byw[wi] >> sh) & 0xffu), sy, py);
        const float bz = __builtin_fmaf((float)((bzw[wi] >> sh) & 0xffu), sz, pz);
        const float tnx = (ax - r.o.x) * r.inv.x, tfx = (bx - r.o.x) * r.inv.x;
        const float tny = (ay - r.o.y) * r.inv.y, tfy = (by - r.o.y) * r.inv.y;
        const float tnz = (az - r.o.z) * r.inv.z, tfz = (bz - r.o.z) * r.inv.z;
        float tnear = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        float tfar  = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        tnear       = tnear * 0.99999f;
        tfar        = tfar * 1.00001f;
        tn[c]       = (cr[c] != CLPT_CHILD_EMPTY && tnear <= tfar) ? tnear : CLPT_FLT_MAX;
    }
    /* nearest first: a 19-comparator sorting network for 8 (fixed indices) */
    cswap(tn[0], cr[0], tn[2], cr[2]);
    cswap(tn[1], cr[1], tn[3], cr[3]);
    cswap(tn[4], cr[4], tn[6], cr[6]);
    cswap(tn[5], cr[5], tn[7], cr[7]);
    cswap(tn[0], cr[0], tn[4], cr[4]);
    cswap(tn[1], cr[1], tn[5], cr[5]);
    cswap(tn[2], cr[2], tn[6], cr[6]);
    cswap(tn[3], cr[3], tn[7], cr[7]);
    cswap(tn[0], cr[0], tn[1], cr[1]);
    cswap(tn[2], cr[2], tn[3], cr[3]);
    cswap(tn[4], cr[4], tn[5], cr[5]);
    cswap(tn[6], cr[6], tn[7], cr[7]);
    cswap(tn[2], cr[2], tn[4], cr[4]);
    cswap(tn[3], cr[3], tn[5], cr[5]);
    cswap(tn[1], cr[1], tn[4], cr[4]);
    cswap(tn[3], cr[3], tn[6], cr[6]);
    cswap(tn[1], cr[1], tn[2], cr[2]);
    cswap(tn[3], cr[3], tn[4], cr[4]);
    cswap(tn[5], cr[5], tn[6], cr[6]);
#pragma unroll
    for (int c = 7; c >= 1; --c)
        if (tn[c] < CLPT_FLT_MAX)
            st.push(cr[c], tn[c]);
    if (tn[0] < CLPT_FLT_MAX)
    {
        ref = cr[0];
        return true;
    }
    return false;
}

/* Two-level BVH traversal (W = 4 or 8 wide) as a per-lane state machine:
 * step() does up to one node visit and one triangle test.  A leaf found
 * during traversal is held while the lane keeps visiting inner nodes of
 * the same tree, and its triangles are tested one per step alongside, so
 * a lane makes progress on both kinds of work in one pass of the wave
 * (a second leaf, an instance entry or the return to the top tree wait
 * until the held leaf is done).  It reports when the ray is finished, so a lane whose ray
 * ends can start its next ray at once.  The closest hit does not depend on
 * the visiting order (ties are broken by (t, instance, primitive)).
 * ANY = occlusion: stop at the first hit (D15: optix.cu:106 traces shadow
 * rays with closest-hit flags).  Inside an instance the object-space ray
 * replaces the world ray.  The top tree may also hold subtrees of the mesh
 * tree of one instance whose transform is the identity
 * (SceneView::world_inst), whose triangles are tested with the world ray
 * itself: the world ray is taken with every -0 component made +0, and for
 * such a ray the instance's stored inverse (ones on the diagonal, zeros of
 * either sign elsewhere) reproduces every component exactly, so those
 * triangles see the bits an instance entry would give them.  The stack depth at instance entry marks where the
 * instance's entries begin: popping an entry below it leaves the instance,
 * and the world ray is restored as soon as no leaf of the instance is held
 * (a held leaf's triangles still need the object-space ray).
 *
 * A new ray's set-up (reciprocals and the watertight shear) is not done
 * when the ray is taken but at its first step, in the same block as an
 * instance's object-space ray: both are rare per lane and few lanes of a
 * wave need either in one pass, so the wave issues the shared set-up once
 * per pass instead of once for each kind. */
template <bool ANY, int W>
struct Tracer
{
    RayW     rw; /* the world ray */
    RayW     r;  /* the ray of the current tree (world, or an instance's object space) */
    uint32_t cur_inst;
    HitRec   best;
    uint32_t ref;
    uint32_t leaf_next, leaf_left;
    int      inst_sp; /* stack depth at instance entry, 0 in the top tree, -1 once an
                         entry of the top tree was popped inside an instance, -2 until
                         the world ray is set up */
    bool     has_ref;
    bool     found;

    __device__ __forceinline__ void start(const SceneView& sc, TraverseStack& st, f3 o, f3 d, float t1)
    {
        /* x + 0 turns -0 into +0 and keeps every other value */
        rw.o      = f3{o.x + 0.0f, o.y + 0.0f, o.z + 0.0f};
        rw.d      = f3{d.x + 0.0f, d.y + 0.0f, d.z + 0.0f};
        cur_inst  = sc.world_inst;
        best.t    = t1;
        best.b1   = 0.0f;
        best.b2   = 0.0f;
        best.inst = 0xffffffffu;
        best.prim = 0xffffffffu;
        found     = false;
        st.sp     = 0;
        ref       = sc.tlas_root;
        has_ref   = true;
        leaf_next = 0;
        leaf_left = 0;
        inst_sp   = -2;
    }

    /* Returns true when the ray is finished; hits need t >= tmin. */
    __device__ __forceinline__ bool step(const SceneView& sc, TraverseStack& st, float tmin, uint32_t& nnodes,
                                         uint32_t& ntris)
    {
        /* node unit */
        if (has_ref)
        {
            const bool held  = leaf_left > 0u;
            const bool enter = inst_sp >= -1 && !held && (ref & (CLPT_LEAF_BIT | CLPT_INST_BIT)) ==
                                                             (CLPT_LEAF_BIT | CLPT_INST_BIT);
            if (inst_sp == -1 && !held && !enter)
            {
                /* back in the top tree: the world ray again */
                r        = rw;
                cur_inst = sc.world_inst;
                inst_sp  = 0;
            }
            if (enter || inst_sp == -2)
            {
                if (enter)
                {
                    /* instance leaf: continue in the instance's BLAS */
                    cur_inst              = ref & ~(CLPT_LEAF_BIT | CLPT_INST_BIT);
                    const InstanceRec* ip = sc.inst + cur_inst;
                    float              m[12];
#pragma unroll
                    for (int k = 0; k < 12; ++k)
                        m[k] = ip->inv[k];
                    ray_to_object(m, rw, r);
                    inst_sp = st.sp;
                    ref     = ip->blas_root;
                }
                else
                {
                    r.o = rw.o;
                    r.d = rw.d;
                }
                ray_setup(r);
                if (inst_sp == -2)
                {
                    rw      = r;
                    inst_sp = 0;
                }
            }
            if (inst_sp < 0)
            {
                /* wait until the held leaf is done */
            }
            else if ((ref & (CLPT_LEAF_BIT | CLPT_INST_BIT)) == CLPT_LEAF_BIT)
            {
                if (!held)
                {
                    leaf_next = ref & (CLPT_MAX_TRIS - 1u);
                    leaf_left = ((ref >> 27) & 7u) + 1u;
                    has_ref   = false;
                }
            }
            else if ((ref & CLPT_LEAF_BIT) == 0u)
            {
                ++nnodes;
                has_ref = visit_node<W>(sc, r, tmin, best.t, st, ref);
            }
            /* else an instance leaf waiting for the held leaf (or, at a
             * ray's first step, for the next pass) */
        }
        /* a second inner node when the first led to one: the blocks below
         * (triangle, pops) and above (set-ups) are shared by two visits */
        if (has_ref && inst_sp >= 0 && (ref & CLPT_LEAF_BIT) == 0u)
        {
            ++nnodes;
            has_ref = visit_node<W>(sc, r, tmin, best.t, st, ref);
        }
        /* triangle unit: one triangle of the held leaf */
        if (leaf_left > 0u)
        {
            const TriRec tr = sc.tris[leaf_next];
            ++ntris;
            ++leaf_next;
            --leaf_left;
            float tt, b1, b2;
            if (tri_test(r, tmin, best.t, f3{tr.v0p.x, tr.v0p.y, tr.v0p.z}, f3{tr.v1.x, tr.v1.y, tr.v1.z},
                         f3{tr.v2.x, tr.v2.y, tr.v2.z}, &tt, &b1, &b2))
            {
                if (ANY)
                {
                    found = true;
                    return true;
                }
                const uint32_t prim = __float_as_uint(tr.v0p.w);
                if (hit_better(tt, cur_inst, prim, best.t, best.inst, best.prim))
                {
                    best.t    = tt;
                    best.b1   = b1;
                    best.b2   = b2;
                    best.inst = cur_inst;
                    best.prim = prim;
                    found     = true;
                }
            }
        }
        if (!has_ref)
        {
            /* the next entry that can still hold a closer hit */
            while (st.sp > 0)
            {
                float t;
                st.pop(ref, t);
                if (st.sp < inst_sp)
                    inst_sp = -1;
                if (t <= best.t)
                {
                    has_ref = true;
                    break;
                }
            }
            if (!has_ref && leaf_left == 0u)
                return true;
        }
        return false;
    }
};

/* ------------------------------------------------------------ shading */

struct Surface
{
    f3       p, ng;
    Frame    f;
    float    u, v;
    uint32_t material;
};

/* Closest-hit epilogue (optix.cu:130-228) with the renderer's corrections:
 * degenerate Ns tested before normalising; frame on Ns (D10). */
__device__ __forceinline__ void surface_interaction(const SceneView& sc, const HitRec& h, Surface& si)
{
    const InstanceRec* ip = sc.inst + h.inst;
    const MeshRec      me = sc.meshes[ip->mesh];
    const uint32_t     i0 = sc.idx[3 * (size_t)(me.idx_off + h.prim) + 0] + me.vtx_off;
    const uint32_t     i1 = sc.idx[3 * (size_t)(me.idx_off + h.prim) + 1] + me.vtx_off;
    const uint32_t     i2 = sc.idx[3 * (size_t)(me.idx_off + h.prim) + 2] + me.vtx_off;
    const float        b1 = h.b1, b2 = h.b2;
    const float        b0 = 1.0f - b1 - b2;
    const f3 p0 = ld3(sc.pos + 3 * (size_t)i0), p1 = ld3(sc.pos + 3 * (size_t)i1), p2 = ld3(sc.pos + 3 * (size_t)i2);
    const f3 po = add3(add3(sclL3(b0, p0), sclL3(b1, p1)), sclL3(b2, p2));
    if (me.flags & MESH_HAS_UVS)
    {
        const float* t0 = sc.uv + 2 * (size_t)i0;
        const float* t1 = sc.uv + 2 * (size_t)i1;
        const float* t2 = sc.uv + 2 * (size_t)i2;
        si.u            = b0 * t0[0] + b1 * t1[0] + b2 * t2[0];
        si.v            = b0 * t0[1] + b1 * t1[1] + b2 * t2[1];
    }
    else
    {
        si.u = b0;
        si.v = b1;
    }
    const f3 ngo = cross3(sub3(p1, p0), sub3(p2, p0));
    f3       nso = ngo;
    if (me.flags & MESH_HAS_NORMALS)
        nso = add3(add3(sclL3(b0, ld3(sc.nrm + 3 * (size_t)i0)), sclL3(b1, ld3(sc.nrm + 3 * (size_t)i1))),
                   sclL3(b2, ld3(sc.nrm + 3 * (size_t)i2)));
    f3 tgo{0.0f, 0.0f, 0.0f};
    if (me.flags & MESH_HAS_TANGENTS)
        tgo = add3(add3(sclL3(b0, ld3(sc.tan + 3 * (size_t)i0)), sclL3(b1, ld3(sc.tan + 3 * (size_t)i1))),
                   sclL3(b2, ld3(sc.tan + 3 * (size_t)i2)));
    float M[12], I[12];
#pragma unroll
    for (int k = 0; k < 12; ++k)
    {
        M[k] = ip->xf[k];
        I[k] = ip->inv[k];
    }
    si.p = f3{M[0] * po.x + M[1] * po.y + M[2] * po.z + M[3], M[4] * po.x + M[5] * po.y + M[6] * po.z + M[7],
              M[8] * po.x + M[9] * po.y + M[10] * po.z + M[11]};
    const f3 ngw = f3{I[0] * ngo.x + I[4] * ngo.y + I[8] * ngo.z, I[1] * ngo.x + I[5] * ngo.y + I[9] * ngo.z,
                      I[2] * ngo.x + I[6] * ngo.y + I[10] * ngo.z};
    const f3 nsw = f3{I[0] * nso.x + I[4] * nso.y + I[8] * nso.z, I[1] * nso.x + I[5] * nso.y + I[9] * nso.z,
                      I[2] * nso.x + I[6] * nso.y + I[10] * nso.z};
    const f3 tgw = f3{M[0] * tgo.x + M[1] * tgo.y + M[2] * tgo.z, M[4] * tgo.x + M[5] * tgo.y + M[6] * tgo.z,
                      M[8] * tgo.x + M[9] * tgo.y + M[10] * tgo.z};
    f3 Ng = norm3(ngw);
    f3 Ns = dot3(nsw, nsw) == 0.0f ? Ng : norm3(nsw);
    if (dot3(Ng, Ns) < 0.0f)
        Ng = neg3(Ng);
    f3 dpdu = tgw, dpdv;
    if (dot3(dpdu, dpdu) > 0.0f)
    {
        dpdu = norm3(dpdu);
        dpdv = cross3(Ns, dpdu);
        if (dot3(dpdv, dpdv) > 0.0f)
        {
            dpdv = norm3(dpdv);
            dpdu = cross3(dpdv, Ns);
        }
        else
            make_frame(Ns, &dpdu, &dpdv);
    }
    else
        make_frame(Ns, &dpdu, &dpdv);
    si.ng       = Ng;
    si.f.s      = dpdv;
    si.f.t      = dpdu;
    si.f.n      = Ns;
    si.material = ip->material;
}

/* The pull-up fix of a normal-mapped frame (Material::
 * fixNormalPerturbingShadingFrame, material.h:236-285): on the side of wo,
 * if wo mirrored about the shading normal would go below the geometric
 * surface, tilt the normal to the half vector of wo and that mirror
 * direction lifted by 1.1 times its depth below the surface. */
__device__ __forceinline__ void pull_up(Frame& f, f3 ng, f3 wo)
{
    if (dot3(wo, ng) < 0.0f)
    {
        f.n = neg3(f.n);
        ng  = neg3(ng);
    }
    if (!(dot3(wo, ng) > 0.0f))
        return;
    const f3 r = reflect3(wo, f.n);
    if (dot3(r, ng) > 0.0f)
        return;
    const float a = fabsf(dot3(r, ng)), b = dot3(f.n, ng);
    if (!(b > 0.0f))
        return;
    f.n = norm3(add3(norm3(add3(r, scl3(scl3(f.n, a / b), 1.1f))), wo));
    f.t = norm3(sub3(f.t, scl3(f.n, dot3(f.t, f.n))));
    f.s = norm3(cross3(f.n, f.t));
    if (dot3(f.n, ng) < 0.0f)
    {
        f.n = neg3(f.n);
        f.t = neg3(f.t);
    }
}

/* Texture-driven inputs of a material at a surface point: the reflectance
 * and the roughness lookups of Material::getBSDF (diffusemtl.h:22-31,
 * metal.h:36-55; roughness = first channel, squared, alpha >= 1e-4), and
 * the normal-mapped frame of Material::getShadingFrame (material.h:124-219):
 * n = normalize(2 tex - 1) in the tangent frame (dpdu, dpdv, Ns) of the hit
 * (Ns rather than Ng: D10), t = dpdu made orthogonal to n, s = n x t, turned
 * to the side of Ng; then the pull-up fix on diffuse and metal.  A normal
 * map texel of (0.5, 0.5, 0.5) -- no direction -- leaves the frame as is. */
__device__ __forceinline__ void material_at(const SceneView& sc, MaterialRec& m, Surface& si, f3 wo)
{
    if (m.tex_r)
    {
        const TexRec t = sc.tex[m.tex_r - 1u];
        const float4 c = tex_eval(t.data, t.w, t.h, si.u, si.v);
        m.r[0]         = c.x;
        m.r[1]         = c.y;
        m.r[2]         = c.z;
    }
    if (m.tex_rough)
    {
        const TexRec t = sc.tex[m.tex_rough - 1u];
        const float  x = tex_eval(t.data, t.w, t.h, si.u, si.v).x;
        m.alpha        = fmaxf(x * x, 1e-4f);
    }
    if (m.tex_n)
    {
        const TexRec t  = sc.tex[m.tex_n - 1u];
        const float4 c  = tex_eval(t.data, t.w, t.h, si.u, si.v);
        const f3     nt = f3{c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f};
        if (dot3(nt, nt) > 0.0f)
        {
            f3 n  = norm3(to_world(si.f, norm3(nt)));
            f3 tm = sub3(si.f.t, scl3(n, dot3(si.f.t, n)));
            f3 sm;
            if (dot3(tm, tm) > 0.0f)
            {
                tm = norm3(tm);
                sm = norm3(cross3(n, tm));
            }
            else
                make_frame(n, &tm, &sm);
            if (dot3(n, si.ng) < 0.0f)
            {
                n  = neg3(n);
                tm = neg3(tm);
            }
            si.f.s = sm;
            si.f.t = tm;
            si.f.n = n;
            if (m.type == 0u || m.type == 1u)
                pull_up(si.f, si.ng, wo);
        }
    }
}

/* ------------------------------------------------------------ kernels */

/* generateCameraRays (raygeneration.cu:83-94): seed (px, py, lane<<20 | s,
 * 0), two draws of box-filter jitter, d = normalize(d00 + u ddu + v ddv). */
__device__ __forceinline__ void camera_ray(const ViewRec& v, uint32_t px, uint32_t py, uint32_t sample, f3* o, f3* d)
{
    const uint32_t z  = (v.lane << 20) | sample;
    const float    u0 = rng_draw(px, py, z, 1u), u1 = rng_draw(px, py, z, 2u);
    const float    sx = ((float)px + u0) / (float)v.width;
    const float    sy = ((float)py + u1) / (float)v.height;
    const f3 d00 = ld3(v.d00), ddu = ld3(v.ddu), ddv = ld3(v.ddv);
    *o           = ld3(v.pos);
    *d           = norm3(add3(add3(d00, sclL3(sx, ddu)), sclL3(sy, ddv)));
}

/* Pixel number q (0 .. w h - 1) of a sample's paths -> the row-major pixel
 * index, with the paths ordered by square tiles of side T: bands of T pixel
 * rows (the last may be lower), tiles of T columns within a band (the last
 * may be narrower), pixels row by row within a tile.  A bijection of
 * [0, w h), so every pixel gets every sample in any order. */
__device__ __forceinline__ uint32_t tile_pixel(uint32_t q, uint32_t w, uint32_t h, uint32_t T)
{
    const uint32_t band = w * T;
    const uint32_t r    = q / band;
    const uint32_t q1   = q - r * band;
    const uint32_t th   = min(T, h - r * T);
    const uint32_t c    = q1 / (T * th);
    const uint32_t s    = q1 - c * T * th;
    const uint32_t tw   = min(T, w - c * T);
    return (r * T + s / tw) * w + c * T + s % tw;
}

/* Path p of batch b -> its view, pixel and sample index. */
__device__ __forceinline__ void locate_path(const BatchTable& bt, const BatchRec& b, uint32_t p, uint32_t& view,
                                            uint32_t& pixel, uint32_t& sample)
{
    uint32_t lo = 0, hi = b.nentries;
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (bt.entries[b.entry_off + mid].path_begin <= p)
            lo = mid;
        else
            hi = mid;
    }
    const BatchEntry e = bt.entries[b.entry_off + lo];
    const uint32_t   n = bt.views[e.view].npix;
    const uint32_t   l = p - e.path_begin;
    view               = e.view;
    pixel              = l % n;
    sample             = e.sample_base + l / n;
    if (bt.tile > 1u)
    {
        const ViewRec& v = bt.views[e.view];
        pixel            = tile_pixel(pixel, v.width, v.height, bt.tile);
    }
}

/* Paths of a batch that record training vertices: one per block of
 * tstride consecutive paths, at an offset hashed from the block and the
 * batch (jittered-stratified over the batch's pixels and samples). */
__device__ __forceinline__ uint32_t guide_tstride(const GuideView& gv, uint32_t npaths)
{
    return gv.nslots ? max(1u, (npaths + gv.nslots - 1u) / gv.nslots) : 0u;
}

__device__ __forceinline__ bool guide_records(uint32_t path, uint32_t tstr, uint32_t batch)
{
    if (!tstr)
        return false;
    uint32_t h = (path / tstr) * 0x9E3779B1u ^ (batch * 0x85EBCA77u + 0x165667B1u);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return path % tstr == h % tstr;
}

/* write_o = 0: the batch's paths all start at one camera position, which
 * traversal takes from the view (the origin plane is not written) */
__global__ __launch_bounds__(256) void k_camera(BatchTable bt, GuideView gv, RayQueue q, BatchCounters* bc,
                                                uint32_t write_o)
{
    const uint32_t gid    = blockIdx.x * blockDim.x + threadIdx.x;
    const BatchRec b      = bt.batches[*bt.cursor];
    const uint32_t npaths = b.npaths;
    const uint32_t tstr   = guide_tstride(gv, npaths);
    if (gid == 0)
    {
        for (int k = 0; k < CLPT_MAXD + 2; ++k)
        {
            bc->rs[2 * k]     = 0u;
            bc->rs[2 * k + 1] = 0u;
            bc->hits[k]       = 0u;
            bc->visible[k]    = 0u;
        }
        *bc_rays(bc, 1) = npaths;
    }
    for (uint32_t i = gid; i < npaths; i += gridDim.x * blockDim.x)
    {
        uint32_t vi, pixel, sample;
        locate_path(bt, b, i, vi, pixel, sample);
        const ViewRec v = bt.views[vi];
        f3            o, d;
        camera_ray(v, pixel % v.width, pixel / v.width, sample, &o, &d);
        if (write_o)
            q.o[i] = make_float4(o.x, o.y, o.z, 0.0f);
        q.d[i] = make_float4(d.x, d.y, d.z, __uint_as_float(i));
        /* the throughput plane is not written: a camera path's throughput is
         * 1, and shading at depth 1 takes it as such */
        if (guide_records(i, tstr, *bt.cursor))
        {
            const uint32_t slot = i / tstr;
            const size_t   pl   = (size_t)(CLPT_MAXD + 1) * gv.nslots;
            const float4   z    = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (uint32_t j = 1; j <= v.max_depth; ++j)
            {
                const size_t rec = (size_t)j * gv.nslots + slot;
                gv.V[3 * pl + rec] = z;
                gv.E[rec]          = z;
            }
        }
    }
}

template <int W>
__global__ __launch_bounds__(256, 5) void k_trace(SceneView sc, RayQueue q, const uint32_t* count, uint32_t depth,
                                               HitBuf hb, uint2* gstack, unsigned long long* ctr,
                                               const float* cam_o)
{
    CLPT_DECLARE_STACK(st, gstack);
    unsigned long long sn = 0, stt = 0; /* the wave's node visits and triangle tests */
    const uint32_t     n    = *count;
    const float        tmin = depth == 1 ? 0.0f : CLPT_RAY_EPS;
    /* camera rays of a single view: every origin is the camera position
     * (cam_o, the view's pos), not read from the origin plane */
    const f3       co     = cam_o ? f3{cam_o[0], cam_o[1], cam_o[2]} : f3{0.0f, 0.0f, 0.0f};
    WavePool       pool{st.gid - lane_id(), st.gstride, 64u};
    uint32_t       i      = st.gid;
    bool           active = i < n;
    Tracer<false, W> tr;
    if (active)
    {
        const float4 d = q.d[i];
        const float4 o = cam_o ? make_float4(co.x, co.y, co.z, 0.0f) : q.o[i];
        tr.start(sc, st, f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, CLPT_FLT_MAX);
    }
    while (__any(active))
    {
        bool     fin = false;
        uint32_t vn = 0, vt = 0;
        if (active && tr.step(sc, st, tmin, vn, vt))
        {
            hb.h[i] = make_float4(tr.best.b1, tr.best.b2, __uint_as_float(tr.found ? tr.best.inst : CLPT_NO_HIT),
                                  __uint_as_float(tr.best.prim));
            fin        = true;
        }
        sn += (unsigned long long)__popcll(__ballot(vn != 0u)) + (unsigned long long)__popcll(__ballot(vn > 1u));
        stt += (unsigned long long)__popcll(__ballot(vt != 0u));
        const unsigned long long mask = __ballot(fin);
        if (mask)
        {
            if (fin)
            {
                i      = pool.take(mask);
                active = i < n;
                if (active)
                {
                    const float4 d = q.d[i];
                    const float4 o = cam_o ? make_float4(co.x, co.y, co.z, 0.0f) : q.o[i];
                    tr.start(sc, st, f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, CLPT_FLT_MAX);
                }
            }
            pool.next += (uint32_t)__popcll(mask);
        }
    }
    if (lane_id() == 0)
    {
        atomicAdd(ctr + 0, sn);
        atomicAdd(ctr + 1, stt);
    }
}

/* Rays per thread in one class-grouping window of the shading kernel. */
#define CLPT_SHADE_GROUP 4u

__global__ __launch_bounds__(256) void k_shade(SceneView sc, BatchTable bt, uint32_t depth, RayQueue qin, HitBuf hb,
                                               RayQueue qout, ShadowQueue sq, float* pacc, BatchCounters* bc)
{
    const uint32_t n      = *bc_rays(bc, depth);
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool     dome   = sc.dome_index >= 0;
    const float    nl     = (float)sc.nemitters;
    const BatchRec b      = bt.batches[*bt.cursor];
    /* The block takes CLPT_SHADE_GROUP x blockDim.x consecutive rays at a
     * time and hands them to its threads grouped by class (escaped, or the
     * hit material's type), blockDim.x at a time, so a wave mostly runs one
     * BSDF's code instead of every type present among its lanes; the wider
     * the window, the fewer waves straddle two classes.  Results do not
     * depend on the order (the RNG is keyed by path and sample); only
     * float64 film sums and queue slots are taken in another order. */
    __shared__ uint32_t order[CLPT_SHADE_GROUP * 256]; /* blockDim.x <= 256 (launch bounds) */
    __shared__ uint32_t ncls[8];
    /* the window's hit records and directions, read once by the class pass
     * (four rays per thread in flight) and taken from LDS by the shading */
    __shared__ float4   sh_h[CLPT_SHADE_GROUP * 256];
    __shared__ float4   sh_d[CLPT_SHADE_GROUP * 256];
    const uint32_t      win     = CLPT_SHADE_GROUP * blockDim.x;
    const uint32_t      wstride = gridDim.x * win;
    /* (window, part): every thread takes the same sequence, so the barriers
     * of a window's sort are reached by the whole block */
    for (uint32_t wbase = blockIdx.x * win, part = 0; wbase < n;)
    {
        if (part == 0u)
        {
            uint32_t cls[CLPT_SHADE_GROUP], slot[CLPT_SHADE_GROUP];
            if (threadIdx.x < 8u)
                ncls[threadIdx.x] = 0u;
            __syncthreads();
#pragma unroll
            for (uint32_t g = 0; g < CLPT_SHADE_GROUP; ++g)
            {
                const uint32_t w  = g * blockDim.x + threadIdx.x;
                const uint32_t i0 = wbase + w;
                cls[g]            = 7u; /* past the end */
                if (i0 < n)
                {
                    const float4 hh = hb.h[i0];
                    sh_h[w]         = hh;
                    sh_d[w]         = qin.d[i0];
                    cls[g] = __float_as_uint(hh.z) != CLPT_NO_HIT ? 1u + sc.mtl[sc.inst[__float_as_uint(hh.z)].material].type : 0u;
                }
            }
#pragma unroll
            for (uint32_t g = 0; g < CLPT_SHADE_GROUP; ++g)
                slot[g] = atomicAdd(&ncls[cls[g]], 1u);
            __syncthreads();
#pragma unroll
            for (uint32_t g = 0; g < CLPT_SHADE_GROUP; ++g)
            {
                uint32_t first = 0u;
                for (uint32_t c = 0; c < cls[g]; ++c)
                    first += ncls[c];
                order[first + slot[g]] = wbase + g * blockDim.x + threadIdx.x;
            }
            __syncthreads();
        }
        const uint32_t i  = order[part * blockDim.x + threadIdx.x];
        const uint32_t wi = i - wbase; /* the ray's place in the window */
        if (++part == CLPT_SHADE_GROUP)
        {
            /* the next window's sort starts with a barrier, so no thread
             * overwrites order while another still reads it */
            part  = 0u;
            wbase = n - wbase > wstride ? wbase + wstride : n;
        }
        if (i >= n)
            continue;
        const float4   d4 = sh_d[wi];
        /* throughput and, in w, the previous BSDF pdf; the origin plane is
         * traversal's only */
        const float4   T4 = depth == 1u ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : qin.T[i];
        const float4   h4 = sh_h[wi];
        const uint32_t path = __float_as_uint(d4.w);
        uint32_t       vi, vpix, sample;
        locate_path(bt, b, path, vi, vpix, sample);
        const ViewRec& fr    = bt.views[vi];
        const uint32_t pixel = fr.film_offset + vpix;
        const uint32_t px = vpix % fr.width, py = vpix / fr.width;
        const uint32_t z  = (fr.lane << 20) | sample;
        const f3       dir{d4.x, d4.y, d4.z};
        f3             T{T4.x, T4.y, T4.z};
        const bool     hit = __float_as_uint(h4.z) != CLPT_NO_HIT;
        bool           want_shadow = false, want_next = false;
        f3             Li{0.0f, 0.0f, 0.0f}, ldir{0.0f, 0.0f, 0.0f}, ndir{0.0f, 0.0f, 0.0f}, p{0.0f, 0.0f, 0.0f};
        float          npdf = 0.0f;
        /* escaped ray (raygeneration.cu:105-136, D3 corrected) */
        f3 esc{0.0f, 0.0f, 0.0f};
        if (!hit && dome)
        {
            const f3 le = dome_le(sc, dir);
            if (depth == 1)
                esc = le;
            else
            {
                const float lp = dome_pdf(sc, dir) / nl;
                const float w  = T4.w / (T4.w + lp);
                esc            = mul3(T, scl3(le, w));
            }
        }
        if (depth == 1u || (!hit && dome))
            path_set(pacc, path, esc);
        if (hit && depth < fr.max_depth)
        {
            HitRec hr;
            hr.t    = 0.0f; /* shading needs no distance */
            hr.b1   = h4.x;
            hr.b2   = h4.y;
            hr.inst = __float_as_uint(h4.z);
            hr.prim = __float_as_uint(h4.w);
            Surface si;
            surface_interaction(sc, hr, si);
            p                    = si.p;
            const f3    wo = neg3(dir);
            MaterialRec mt = sc.mtl[si.material];
            material_at(sc, mt, si, wo);
            const f3          woL = to_local(si.f, wo);
            const uint32_t    k0  = 2u + 4u * (depth - 1u);
            /* next-event estimation (raygeneration.cu:310-384; D8, D9) */
            {
                const float lu = rng_draw(px, py, z, k0 + 1u), lv = rng_draw(px, py, z, k0 + 2u);
                float       lpdf  = 0.0f;
                uint32_t    which = 0;
                const f3    value = sample_emitters(sc, lu, lv, &ldir, &lpdf, &which);
                if (nonzero3(value) && lpdf > 0.0f)
                {
                    const f3 wiL = to_local(si.f, ldir);
                    const f3 f   = bsdf_eval(mt, wiL, woL);
                    if (nonzero3(f))
                    {
                        float weight = 1.0f;
                        if (sc.emitters[which].type == 1u)
                        {
                            const float bp = bsdf_pdf(mt, wiL, woL);
                            const float lp = lpdf / nl;
                            weight         = lp / (lp + bp);
                        }
                        const float cosv = fabsf(wiL.z);
                        Li          = f3{T.x * value.x * f.x * cosv * weight / lpdf, T.y * value.y * f.y * cosv * weight / lpdf,
                                T.z * value.z * f.z * cosv * weight / lpdf};
                        want_shadow = true;
                    }
                }
            }
            /* BSDF sampling (raygeneration.cu:391-453; D7, D11, D13) */
            const float bu = rng_draw(px, py, z, k0 + 3u), bv = rng_draw(px, py, z, k0 + 4u);
            const float ul = rng_draw(px, py, z, CLPT_EXTRA_DRAW(depth, CLPT_SLOT_LOBE));
            f3          wiL{0.0f, 0.0f, 0.0f};
            float       bpdf = 0.0f;
            const f3    f    = bsdf_sample(mt, woL, bu, bv, ul, &wiL, &bpdf);
            if (nonzero3(f) && bpdf > 0.0f)
            {
                const float cosv = fabsf(wiL.z);
                T         = f3{T.x * (f.x * cosv / bpdf), T.y * (f.y * cosv / bpdf), T.z * (f.z * cosv / bpdf)};
                ndir      = to_world(si.f, wiL);
                npdf      = bpdf;
                want_next = true;
                if (fr.rr_depth > 0u && depth >= fr.rr_depth)
                {
                    const float q  = fminf(1.0f, fmaxf(T.x, fmaxf(T.y, T.z)));
                    const float ur = rng_draw(px, py, z, CLPT_EXTRA_DRAW(depth, CLPT_SLOT_RR));
                    if (!(ur < q))
                        want_next = false;
                    else
                        T = f3{T.x / q, T.y / q, T.z / q};
                }
            }
        }
        /* converged: wave-aggregated queue pushes */
        wave_count(&bc->hits[depth], hit);
        uint32_t ss, ns;
        wave_push2(bc_push(bc, depth), want_shadow, want_next, ss, ns);
        if (want_shadow)
        {
            sq.o[ss] = make_float4(p.x, p.y, p.z, Li.x);
            sq.d[ss] = make_float4(ldir.x, ldir.y, ldir.z, Li.y);
            sq.x[ss] = make_float2(Li.z, __uint_as_float(pixel));
        }
        if (want_next)
        {
            qout.o[ns] = make_float4(p.x, p.y, p.z, 0.0f);
            qout.d[ns] = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float(path));
            qout.T[ns] = make_float4(T.x, T.y, T.z, npdf);
        }
    }
}

/* Shading with the learned proposal (guided vertices: diffuse and
 * conductor surfaces).  The continuation direction is drawn from the
 * one-sample mixture c_bsdf * p_bsdf + (1 - c_bsdf) * p_guide; its density
 * p_mix replaces the BSDF pdf everywhere the plain kernel uses it (path
 * throughput f cos / p_mix, the escape MIS weight and the next-event MIS
 * weight against the HDRI), so with c_bsdf > 0 the estimator stays
 * unbiased.  The network runs once per wave of 64 rays (all lanes take part;
 * its weights sit in LDS).  Paths selected for training record their
 * guided vertices and the contributions that came back. */
/* At least 2 waves per SIMD: the LDS image allows 2 blocks per CU, and the
 * register budget is held to match (accumulators in VGPRs, no spills).
 * FP8 selects the e4m3 inference ablation. */
template <bool FP8>
__global__ __launch_bounds__(256, 2) void k_shade_guided(SceneView sc, BatchTable bt, GuideView gv, uint32_t depth,
                                                         RayQueue qin, HitBuf hb, RayQueue qout, ShadowQueue sq,
                                                         float* pacc, BatchCounters* bc)
{
    __shared__ uint4 wlds[CLPT_G_IMG * (FP8 ? 1 : 2) / 16];
    __shared__ float wscl[10];
    {
        const uint4* src = reinterpret_cast<const uint4*>(FP8 ? (const void*)gv.img8 : (const void*)gv.img);
        for (uint32_t k = threadIdx.x; k < CLPT_G_IMG * (FP8 ? 1 : 2) / 16; k += blockDim.x)
            wlds[k] = src[k];
        if (threadIdx.x < 10)
            wscl[threadIdx.x] = FP8 ? gv.scl[threadIdx.x] : 1.0f;
        __syncthreads();
    }
    const uint32_t n      = *bc_rays(bc, depth);
    const bool     dome   = sc.dome_index >= 0;
    const float    nl     = (float)sc.nemitters;
    const BatchRec b      = bt.batches[*bt.cursor];
    const uint32_t tstr   = guide_tstride(gv, b.npaths);
    const float    cb     = gv.c_bsdf;
    __shared__ uint32_t order[CLPT_SHADE_GROUP * 256]; /* blockDim.x <= 256 (launch bounds) */
    __shared__ uint32_t ncls[8];
    const uint32_t      win     = CLPT_SHADE_GROUP * blockDim.x;
    const uint32_t      wstride = gridDim.x * win;
    /* rays handed to threads grouped by class over windows of
     * CLPT_SHADE_GROUP x blockDim.x rays, as in k_shade: a wave then mostly
     * holds one material, and waves without a guided vertex skip the
     * network.  Every thread takes the same (window, part) sequence and
     * runs the body (the network needs the whole wave). */
    for (uint32_t wbase = blockIdx.x * win, part = 0; wbase < n;)
    {
        if (part == 0u)
        {
            uint32_t cls[CLPT_SHADE_GROUP], slot[CLPT_SHADE_GROUP];
            if (threadIdx.x < 8u)
                ncls[threadIdx.x] = 0u;
            __syncthreads();
#pragma unroll
            for (uint32_t g = 0; g < CLPT_SHADE_GROUP; ++g)
            {
                const uint32_t i0 = wbase + g * blockDim.x + threadIdx.x;
                cls[g]            = 7u; /* past the end */
                if (i0 < n)
                {
                    const float4 hh = hb.h[i0];
                    cls[g] = __float_as_uint(hh.z) != CLPT_NO_HIT ? 1u + sc.mtl[sc.inst[__float_as_uint(hh.z)].material].type : 0u;
                }
            }
#pragma unroll
            for (uint32_t g = 0; g < CLPT_SHADE_GROUP; ++g)
                slot[g] = atomicAdd(&ncls[cls[g]], 1u);
            __syncthreads();
#pragma unroll
            for (uint32_t g = 0; g < CLPT_SHADE_GROUP; ++g)
            {
                uint32_t first = 0u;
                for (uint32_t c = 0; c < cls[g]; ++c)
                    first += ncls[c];
                order[first + slot[g]] = wbase + g * blockDim.x + threadIdx.x;
            }
            __syncthreads();
        }
        const uint32_t i = order[part * blockDim.x + threadIdx.x];
        if (++part == CLPT_SHADE_GROUP)
        {
            part  = 0u;
            wbase = n - wbase > wstride ? wbase + wstride : n;
        }
        const bool active = i < n;
        float4 d4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), T4 = d4, h4 = make_float4(0.0f, 0.0f, __uint_as_float(CLPT_NO_HIT), 0.0f);
        if (active)
        {
            d4 = qin.d[i];
            T4 = depth == 1u ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : qin.T[i];
            h4 = hb.h[i];
        }
        const uint32_t path = __float_as_uint(d4.w);
        uint32_t       vi = 0, vpix = 0, sample = 0;
        if (active)
            locate_path(bt, b, path, vi, vpix, sample);
        const ViewRec& fr     = bt.views[vi];
        const uint32_t pixel  = fr.film_offset + vpix;
        const uint32_t px = vpix % fr.width, py = vpix / fr.width;
        const uint32_t z      = (fr.lane << 20) | sample;
        const f3       dir{d4.x, d4.y, d4.z};
        f3             T{T4.x, T4.y, T4.z};
        const bool     hit    = active && __float_as_uint(h4.z) != CLPT_NO_HIT;
        const bool     vertex = hit && depth < fr.max_depth;
        const bool     train  = active && guide_records(path, tstr, *bt.cursor);
        /* record (depth, slot) of a training path; planes of V are (CLPT_MAXD+1) * nslots apart */
        const size_t   rec    = train ? (size_t)depth * gv.nslots + path / tstr : 0;
        const size_t   pl     = (size_t)(CLPT_MAXD + 1) * gv.nslots;
        Surface        si{};
        MaterialRec    mt{};
        if (vertex)
        {
            HitRec hr;
            hr.t    = 0.0f; /* shading needs no distance */
            hr.b1   = h4.x;
            hr.b2   = h4.y;
            hr.inst = __float_as_uint(h4.z);
            hr.prim = __float_as_uint(h4.w);
            surface_interaction(sc, hr, si);
            mt = sc.mtl[si.material];
            material_at(sc, mt, si, neg3(dir));
        }
        const bool guided = vertex && (mt.type == 0u || mt.type == 1u);
        const f3   wo     = neg3(dir);
        float      q[9];
        {
            GuideBox box;
            for (int k = 0; k < 3; ++k)
            {
                box.lo[k]      = gv.lo[k];
                box.inv_ext[k] = gv.inv_ext[k];
            }
            guide_query(box, si.p, si.f.n, wo, q);
        }
        VmfMix mix;
        if (__ballot(guided) != 0ull) /* wave-uniform */
        {
            float out[CLPT_G_NOUT];
            guide_forward_wave_t<FP8>(wlds, wscl, q, out);
            vmf_mixture(out, mix);
        }
        bool  want_shadow = false, want_next = false;
        f3    Li{0.0f, 0.0f, 0.0f}, ldir{0.0f, 0.0f, 0.0f}, ndir{0.0f, 0.0f, 0.0f};
        float npdf = 0.0f;
        /* escaped ray: the path's slot as in k_shade */
        f3 esc{0.0f, 0.0f, 0.0f};
        if (active && !hit && dome)
        {
            const f3 le = dome_le(sc, dir);
            if (depth == 1)
                esc = le;
            else
            {
                const float lp = dome_pdf(sc, dir) / nl;
                const float w  = T4.w / (T4.w + lp);
                esc            = mul3(T, scl3(le, w));
            }
            if (train)
                gv.E[rec] = make_float4(esc.x, esc.y, esc.z, 0.0f);
        }
        if (active && (depth == 1u || (!hit && dome)))
            path_set(pacc, path, esc);
        if (vertex)
        {
            const f3       woL = to_local(si.f, wo);
            const uint32_t k0  = 2u + 4u * (depth - 1u);
            {
                const float lu = rng_draw(px, py, z, k0 + 1u), lv = rng_draw(px, py, z, k0 + 2u);
                float       lpdf  = 0.0f;
                uint32_t    which = 0;
                const f3    value = sample_emitters(sc, lu, lv, &ldir, &lpdf, &which);
                if (nonzero3(value) && lpdf > 0.0f)
                {
                    const f3 wiL = to_local(si.f, ldir);
                    const f3 f   = bsdf_eval(mt, wiL, woL);
                    if (nonzero3(f))
                    {
                        float weight = 1.0f;
                        if (sc.emitters[which].type == 1u)
                        {
                            float ps = bsdf_pdf(mt, wiL, woL);
                            if (guided)
                                ps = cb * ps + (1.0f - cb) * vmf_mixture_pdf(mix, ldir);
                            const float lp = lpdf / nl;
                            weight         = lp / (lp + ps);
                        }
                        const float cosv = fabsf(wiL.z);
                        Li          = f3{T.x * value.x * f.x * cosv * weight / lpdf, T.y * value.y * f.y * cosv * weight / lpdf,
                                T.z * value.z * f.z * cosv * weight / lpdf};
                        want_shadow = true;
                    }
                }
            }
            const float bu = rng_draw(px, py, z, k0 + 3u), bv = rng_draw(px, py, z, k0 + 4u);
            const float ul = rng_draw(px, py, z, CLPT_EXTRA_DRAW(depth, CLPT_SLOT_LOBE));
            f3          wiL{0.0f, 0.0f, 0.0f}, f{0.0f, 0.0f, 0.0f};
            float       pdf = 0.0f;
            if (!guided)
                f = bsdf_sample(mt, woL, bu, bv, ul, &wiL, &pdf);
            else
            {
                const float ut = rng_draw(px, py, z, CLPT_EXTRA_DRAW(depth, CLPT_SLOT_GUIDE));
                f3          wi;
                float       bp = 0.0f;
                if (ut < cb)
                {
                    f  = bsdf_sample(mt, woL, bu, bv, ul, &wiL, &bp);
                    wi = to_world(si.f, wiL);
                }
                else
                {
                    const float us = rng_draw(px, py, z, CLPT_EXTRA_DRAW(depth, CLPT_SLOT_LOBESEL));
                    wi             = vmf_mixture_sample(mix, us, bu, bv);
                    wiL            = to_local(si.f, wi);
                    f              = bsdf_eval(mt, wiL, woL);
                    bp             = bsdf_pdf(mt, wiL, woL);
                }
                pdf = cb * bp + (1.0f - cb) * vmf_mixture_pdf(mix, wi);
            }
            if (nonzero3(f) && pdf > 0.0f)
            {
                const float cosv = fabsf(wiL.z);
                const f3    fc{f.x * cosv, f.y * cosv, f.z * cosv};
                T         = f3{T.x * (fc.x / pdf), T.y * (fc.y / pdf), T.z * (fc.z / pdf)};
                ndir      = to_world(si.f, wiL);
                npdf      = pdf;
                want_next = true;
                if (fr.rr_depth > 0u && depth >= fr.rr_depth)
                {
                    const float qq = fminf(1.0f, fmaxf(T.x, fmaxf(T.y, T.z)));
                    const float ur = rng_draw(px, py, z, CLPT_EXTRA_DRAW(depth, CLPT_SLOT_RR));
                    if (!(ur < qq))
                        want_next = false;
                    else
                        T = f3{T.x / qq, T.y / qq, T.z / qq};
                }
                if (train && guided && want_next)
                {
                    gv.V[rec]          = make_float4(q[0], q[1], q[2], q[3]);
                    gv.V[pl + rec]     = make_float4(q[4], q[5], q[6], q[7]);
                    gv.V[2 * pl + rec] = make_float4(q[8], ndir.x, ndir.y, ndir.z);
                    gv.V[3 * pl + rec] = make_float4(pdf, T.x > 0.0f ? fc.x / T.x : 0.0f,
                                                     T.y > 0.0f ? fc.y / T.y : 0.0f, T.z > 0.0f ? fc.z / T.z : 0.0f);
                }
            }
        }
        wave_count(&bc->hits[depth], hit);
        uint32_t ss, ns;
        wave_push2(bc_push(bc, depth), want_shadow, want_next, ss, ns);
        if (want_shadow)
        {
            sq.o[ss]   = make_float4(si.p.x, si.p.y, si.p.z, Li.x);
            sq.d[ss]   = make_float4(ldir.x, ldir.y, ldir.z, Li.y);
            sq.x[ss]   = make_float2(Li.z, __uint_as_float(pixel));
            sq.tag[ss] = train ? (uint32_t)rec : 0xffffffffu;
        }
        if (want_next)
        {
            qout.o[ns] = make_float4(si.p.x, si.p.y, si.p.z, 0.0f);
            qout.d[ns] = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float(path));
            qout.T[ns] = make_float4(T.x, T.y, T.z, npdf);
        }
    }
}

template <int W>
__global__ __launch_bounds__(256, 6) void k_shadow(SceneView sc, ShadowQueue sq, uint32_t depth, double* accum,
                                                BatchCounters* bc, uint2* gstack, unsigned long long* ctr,
                                                float4* guide_E)
{
    CLPT_DECLARE_STACK(st, gstack);
    unsigned long long sn = 0, stt = 0, sv = 0; /* the wave's node visits, triangle tests, unoccluded rays */
    const uint32_t     n      = *bc_shadow(bc, depth);
    WavePool       pool{st.gid - lane_id(), st.gstride, 64u};
    uint32_t       i      = st.gid;
    bool           active = i < n;
    Tracer<true, W> tr;
    /* the ray's radiance and pixel, read with the ray so that a visible ray's
     * film update needs no load when it finishes */
    f3       lo{0.0f, 0.0f, 0.0f};
    uint32_t lpix = 0;
    if (active)
    {
        const float4 o = sq.o[i], d = sq.d[i];
        const float2 x = sq.x[i];
        lo             = f3{o.w, d.w, x.x};
        lpix           = __float_as_uint(x.y);
        tr.start(sc, st, f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, CLPT_FLT_MAX);
    }
    while (__any(active))
    {
        bool     fin = false, vis = false;
        uint32_t vn = 0, vt = 0;
        if (active && tr.step(sc, st, CLPT_RAY_EPS, vn, vt))
        {
            if (!tr.found)
            {
                accum_add(accum, lpix, lo);
                vis = true;
                if (sq.tag)
                {
                    const uint32_t tg = sq.tag[i];
                    if (tg != 0xffffffffu)
                        guide_E[tg] = make_float4(lo.x, lo.y, lo.z, 0.0f);
                }
            }
            fin = true;
        }
        sn += (unsigned long long)__popcll(__ballot(vn != 0u)) + (unsigned long long)__popcll(__ballot(vn > 1u));
        stt += (unsigned long long)__popcll(__ballot(vt != 0u));
        sv += (unsigned long long)__popcll(__ballot(vis));
        const unsigned long long mask = __ballot(fin);
        if (mask)
        {
            if (fin)
            {
                i      = pool.take(mask);
                active = i < n;
                if (active)
                {
                    const float4 o = sq.o[i], d = sq.d[i];
                    const float2 x = sq.x[i];
                    lo             = f3{o.w, d.w, x.x};
                    lpix           = __float_as_uint(x.y);
                    tr.start(sc, st, f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, CLPT_FLT_MAX);
                }
            }
            pool.next += (uint32_t)__popcll(mask);
        }
    }
    if (lane_id() == 0)
    {
        atomicAdd(ctr + 2, sn);
        atomicAdd(ctr + 3, stt);
        if (sv)
            atomicAdd(&bc->visible[depth], (uint32_t)sv);
    }
}

/* The batch's path sums (path_set) into the film: for each entry, one thread
 * per pixel adds the sums of the entry's samples of that pixel, in sample
 * order.  A view has one entry per batch, so each film pixel has one writer:
 * plain float64 adds, and the film's sums do not depend on scheduling. */
__global__ __launch_bounds__(256) void k_film_fold(BatchTable bt, const float* pacc, double* accum)
{
    const BatchRec b      = bt.batches[*bt.cursor];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t en = 0; en < b.nentries; ++en)
    {
        const BatchEntry e   = bt.entries[b.entry_off + en];
        const ViewRec&   v   = bt.views[e.view];
        const uint32_t   end = en + 1u < b.nentries ? bt.entries[b.entry_off + en + 1u].path_begin : b.npaths;
        for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < v.npix; l += stride)
        {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (uint32_t q = e.path_begin + l; q < end; q += v.npix)
            {
                s0 += (double)pacc[3 * (size_t)q + 0];
                s1 += (double)pacc[3 * (size_t)q + 1];
                s2 += (double)pacc[3 * (size_t)q + 2];
            }
            const uint32_t pix = v.film_offset + (bt.tile > 1u ? tile_pixel(l, v.width, v.height, bt.tile) : l);
            accum[3 * (size_t)pix + 0] += s0;
            accum[3 * (size_t)pix + 1] += s1;
            accum[3 * (size_t)pix + 2] += s2;
        }
    }
}

__global__ void k_batch_end(const BatchCounters* bc, TotalCounters* tot, uint32_t* cursor)
{
    const int b = threadIdx.x;
    if (b < CLPT_MAXD + 2)
    {
        tot->rays[b] += b >= 1 ? bc->rs[2 * b - 1] : 0u;
        tot->hits[b] += bc->hits[b];
        tot->shadow[b] += bc->rs[2 * b];
        tot->visible[b] += bc->visible[b];
    }
    if (b == 0)
        *cursor += 1u;
}

/* Sets the batch cursor (a kernel rather than a memset, so that it may sit
 * in a captured graph). */
__global__ void k_set_cursor(uint32_t* cursor, uint32_t value)
{
    if (threadIdx.x == 0)
        *cursor = value;
}

/* sums / spp per pixel; with srgb the display copy is encoded as the
 * reference's postprocessing kernel does (raygeneration.cu:498-506), while
 * the float64 film stays linear. */
__global__ void k_resolve(const double* accum, uint32_t npix, uint32_t spp, int srgb, float4* out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x)
    {
        const double inv = spp ? 1.0 / (double)spp : 0.0;
        float        r = (float)(accum[3 * (size_t)i] * inv), g = (float)(accum[3 * (size_t)i + 1] * inv),
              b = (float)(accum[3 * (size_t)i + 2] * inv);
        if (srgb)
        {
            r = srgb_encode(r);
            g = srgb_encode(g);
            b = srgb_encode(b);
        }
        out[i] = make_float4(r, g, b, 1.0f);
    }
}

/* The display functions on n values (parity surface). */
__global__ void k_srgb_batch(const float* in, uint32_t n, int decode, float* out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = decode ? srgb_decode(in[i]) : srgb_encode(in[i]);
}

/* An 8-bit texture as the texture unit reads it with
 * cudaReadModeNormalizedFloat and, for sRGB images, sRGB decoding of the
 * colour channels (imagetex2d.cpp:50-66): c / 255, then convertsRGBToLinear
 * before any filtering; alpha is not decoded. */
__global__ void k_texture_u8(const uchar4* in, uint32_t n, int srgb, float4* out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uchar4 c = in[i];
        float        r = (float)c.x / 255.0f, g = (float)c.y / 255.0f, b = (float)c.z / 255.0f;
        if (srgb)
        {
            r = srgb_decode(r);
            g = srgb_decode(g);
            b = srgb_decode(b);
        }
        out[i] = make_float4(r, g, b, (float)c.w / 255.0f);
    }
}

/* ------------------------------------------------------------ HDRI tables */

/* prefiltering (hdridome.h:311-324): sin(theta_v) * luminance(texel). */
__global__ void k_dome_prefilter(const float4* tex, uint32_t w, uint32_t h, float* pcond)
{
    const uint32_t total = w * h;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x)
    {
        const uint32_t v  = i / w, u = i % w;
        const float    st = clpt_sinf((CLPT_PI_F * ((float)v + 0.5f)) / (float)h);
        const float4   c  = tex[i];
        const float    lum = (float)((double)c.x * 0.212671 + (double)c.y * 0.71516 + (double)c.z * 0.072169);
        pcond[(size_t)v * (w + 1) + u] = st * lum;
    }
}

/* Block-wide exclusive scan of one double per thread (256 threads). */
__device__ double block_exclusive_scan(double v, double* sh, double* total)
{
    const uint32_t t    = threadIdx.x;
    const uint32_t lane = t & 63u, wave = t >> 6;
    double         x    = v;
    for (int off = 1; off < 64; off <<= 1)
    {
        const double y = __shfl_up(x, off);
        if (lane >= (uint32_t)off)
            x += y;
    }
    if (lane == 63u)
        sh[wave] = x;
    __syncthreads();
    double wave_off = 0.0, tot = 0.0;
    for (uint32_t k = 0; k < (blockDim.x >> 6); ++k)
    {
        if (k < wave)
            wave_off += sh[k];
        tot += sh[k];
    }
    __syncthreads();
    *total = tot;
    return wave_off + x - v;
}

/* preprocessPCondV (hdridome.h:333-385) in float64: one block per row;
 * each thread scans a contiguous chunk, then a block scan joins them. */
__global__ __launch_bounds__(256) void k_dome_rows(uint32_t w, uint32_t h, float* pcond, float* cdfcond, float* pv,
                                                    double* rowint)
{
    __shared__ double sh[8];
    const uint32_t    v     = blockIdx.x;
    float*            p     = pcond + (size_t)v * (w + 1);
    float*            cdf   = cdfcond + (size_t)v * (w + 1);
    const uint32_t    chunk = (w + blockDim.x - 1) / blockDim.x;
    const uint32_t    b     = threadIdx.x * chunk;
    const uint32_t    e     = min(b + chunk, w);
    double            local = 0.0;
    for (uint32_t i = b; i < e; ++i)
        local += (double)p[i] / (double)w;
    double       total;
    const double off = block_exclusive_scan(local, sh, &total);
    if (total <= 0.0)
    {
        for (uint32_t i = b; i < e; ++i)
        {
            p[i]   = 0.0f;
            cdf[i] = (float)i / (float)w;
        }
        if (threadIdx.x == 0)
        {
            cdf[w]    = 1.0f;
            p[w]      = 0.0f;
            rowint[v] = 0.0;
            pv[v]     = 0.0f;
        }
        return;
    }
    double acc = off;
    if (threadIdx.x == 0)
        cdf[0] = 0.0f;
    for (uint32_t i = b; i < e; ++i)
    {
        acc += (double)p[i] / (double)w;
        cdf[i + 1] = (float)(acc / total);
    }
    if (threadIdx.x == 0)
    {
        p[w]      = (float)total;
        rowint[v] = total;
        pv[v]     = (float)total;
    }
}

/* preprocessPV (hdridome.h:387-432) in float64: one block. */
__global__ __launch_bounds__(256) void k_dome_marginal(uint32_t h, float* pv, float* cdfv, const double* rowint)
{
    __shared__ double sh[8];
    const uint32_t    chunk = (h + blockDim.x - 1) / blockDim.x;
    const uint32_t    b     = threadIdx.x * chunk;
    const uint32_t    e     = min(b + chunk, h);
    double            local = 0.0;
    for (uint32_t j = b; j < e; ++j)
        local += rowint[j] / (double)h;
    double       total;
    const double off = block_exclusive_scan(local, sh, &total);
    if (total <= 0.0)
    {
        for (uint32_t j = b; j < e; ++j)
        {
            pv[j]   = 0.0f;
            cdfv[j] = (float)j / (float)h;
        }
        if (threadIdx.x == 0)
        {
            pv[h]   = 0.0f;
            cdfv[h] = 1.0f;
        }
        return;
    }
    double acc = off;
    if (threadIdx.x == 0)
        cdfv[0] = 0.0f;
    for (uint32_t j = b; j < e; ++j)
    {
        acc += rowint[j] / (double)h;
        cdfv[j + 1] = (float)(acc / total);
    }
    if (threadIdx.x == 0)
        pv[h] = (float)total;
}

/* Search cut points of the row CDFs (blocks 0..h-1) and of the marginal
 * (block h): cut[k] = upper_bound over the CDF's entries of k / CLPT_CUTS. */
__global__ __launch_bounds__(256) void k_dome_cuts(uint32_t w, uint32_t h, const float* cdfcond, const float* cdfv,
                                                    uint32_t* cutv, uint32_t* cutcond)
{
    const uint32_t r = blockIdx.x;
    for (uint32_t k = threadIdx.x; k <= CLPT_CUTS; k += blockDim.x)
    {
        const float v = (float)k / (float)CLPT_CUTS;
        if (r < h)
            cutcond[(size_t)r * (CLPT_CUTS + 1u) + k] =
                upper_bound_f(cdfcond, r * (w + 1u), r * (w + 1u) + w + 1u, v);
        else
            cutv[k] = upper_bound_f(cdfv, 0u, h + 1u, v);
    }
}

/* ------------------------------------------------------------ batch kernels */

__global__ void k_rng_batch(const uint4* seeds, uint32_t n, uint32_t ndraws, uint32_t* out_u, float* out_f)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint4 s = seeds[i];
        for (uint32_t d = 0; d < ndraws; ++d)
        {
            const uint32_t x            = pcg4d_x(s.x, s.y, s.z, s.w + d + 1u);
            out_u[(size_t)i * ndraws + d] = x;
            out_f[(size_t)i * ndraws + d] = u2f01(x);
        }
    }
}

__global__ void k_camera_batch(ViewRec fr, const uint2* ps, uint32_t n, float4* out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint2 q = ps[i];
        f3          o, d;
        camera_ray(fr, q.x % fr.width, q.x / fr.width, q.y, &o, &d);
        out[2 * i]     = make_float4(o.x, o.y, o.z, 0.0f);
        out[2 * i + 1] = make_float4(d.x, d.y, d.z, 0.0f);
    }
}

__global__ void k_bsdf_batch(SceneView sc, const uint32_t* mtl, const float* wi3, const float* wo3, const float* u3,
                             uint32_t n, float* ef3, float* epdf, float* swi3, float* sf3, float* spdf)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const MaterialRec m  = sc.mtl[mtl[i]];
        const f3          wi = ld3(wi3 + 3 * (size_t)i), wo = ld3(wo3 + 3 * (size_t)i);
        const f3          f  = bsdf_eval(m, wi, wo);
        ef3[3 * i]           = f.x;
        ef3[3 * i + 1]       = f.y;
        ef3[3 * i + 2]       = f.z;
        epdf[i]              = bsdf_pdf(m, wi, wo);
        f3       w{0.0f, 0.0f, 0.0f};
        float    p  = 0.0f;
        const f3 fs = bsdf_sample(m, wo, u3[3 * i], u3[3 * i + 1], u3[3 * i + 2], &w, &p);
        swi3[3 * i]     = w.x;
        swi3[3 * i + 1] = w.y;
        swi3[3 * i + 2] = w.z;
        sf3[3 * i]      = fs.x;
        sf3[3 * i + 1]  = fs.y;
        sf3[3 * i + 2]  = fs.z;
        spdf[i]         = p;
    }
}

__global__ void k_emitter_batch(SceneView sc, const float* u2, uint32_t n, float* dir3, float* val3, float* pdf,
                                uint32_t* which)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        f3       d{0.0f, 0.0f, 0.0f};
        float    p = 0.0f;
        uint32_t w = 0;
        const f3 v = sample_emitters(sc, u2[2 * i], u2[2 * i + 1], &d, &p, &w);
        dir3[3 * i]     = d.x;
        dir3[3 * i + 1] = d.y;
        dir3[3 * i + 2] = d.z;
        val3[3 * i]     = v.x;
        val3[3 * i + 1] = v.y;
        val3[3 * i + 2] = v.z;
        pdf[i]          = p;
        which[i]        = w;
    }
}

__global__ void k_dome_eval_batch(SceneView sc, const float* dir3, uint32_t n, float* le3, float* pdf)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const f3 d  = ld3(dir3 + 3 * (size_t)i);
        const f3 le = dome_le(sc, d);
        le3[3 * i]     = le.x;
        le3[3 * i + 1] = le.y;
        le3[3 * i + 2] = le.z;
        pdf[i]         = dome_pdf(sc, d);
    }
}

template <int W>
__global__ __launch_bounds__(256) void k_trace_batch(SceneView sc, const float4* rays, uint32_t n, int any,
                                                     float4* out, uint32_t* prim, uint2* gstack)
{
    CLPT_DECLARE_STACK(st, gstack);
    uint32_t nn = 0, nt = 0;
    for (uint32_t i = st.gid; i < n; i += st.gstride)
    {
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        const f3     ro{a.x, a.y, a.z}, rd{b.x, b.y, b.z};
        if (any)
        {
            Tracer<true, W> tr;
            tr.start(sc, st, ro, rd, b.w);
            while (!tr.step(sc, st, a.w, nn, nt))
            {
            }
            out[i]  = make_float4(tr.found ? 1.0f : 0.0f, 0.0f, 0.0f, 0.0f);
            prim[i] = 0u;
        }
        else
        {
            Tracer<false, W> tr;
            tr.start(sc, st, ro, rd, b.w);
            while (!tr.step(sc, st, a.w, nn, nt))
            {
            }
            out[i]  = make_float4(tr.found ? tr.best.t : -1.0f, tr.found ? tr.best.b1 : 0.0f,
                                 tr.found ? tr.best.b2 : 0.0f, __uint_as_float(tr.found ? tr.best.inst : 0xffffffffu));
            prim[i] = tr.found ? tr.best.prim : 0xffffffffu;
        }
    }
}

/* Streaming copy (16-byte non-temporal loads and stores, four loads in
 * flight per lane before their stores): the known-traffic kernel for HBM
 * byte counters and the measured copy ceiling. */
typedef float f4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ src4, float4* __restrict__ dst4, size_t n)
{
    const f4v*   src    = reinterpret_cast<const f4v*>(src4);
    f4v*         dst    = reinterpret_cast<f4v*>(dst4);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t       i      = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride)
    {
        const f4v a = __builtin_nontemporal_load(src + i);
        const f4v b = __builtin_nontemporal_load(src + i + stride);
        const f4v c = __builtin_nontemporal_load(src + i + 2 * stride);
        const f4v d = __builtin_nontemporal_load(src + i + 3 * stride);
        __builtin_nontemporal_store(a, dst + i);
        __builtin_nontemporal_store(b, dst + i + stride);
        __builtin_nontemporal_store(c, dst + i + 2 * stride);
        __builtin_nontemporal_store(d, dst + i + 3 * stride);
    }
    for (; i < n; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

/* ------------------------------------------------------------ allocation-free tracing */

/* The traversal stacks' overflow area of the allocation-free trace entry
 * point: CLPT_NOALLOC_BLOCKS persistent blocks of 256 lanes, each lane
 * CLPT_GSTACK_ENTRIES (ref, t) pairs, laid out [entry][global lane] as the
 * trace kernel addresses it (58 MiB).  Each device holds its own copy for
 * the life of the process; every call on a device shares it, and it needs
 * no clearing (an entry is always pushed before it is popped). */
static __device__ __align__(256) uint2 scratch[(size_t)CLPT_NOALLOC_BLOCKS * 256u * CLPT_GSTACK_ENTRIES];

namespace clpt
{
/* The device address of `scratch` on `device` (which must be current),
 * looked up once per device.  Host code uses this address, never the
 * array's name, which in host code is only a stand-in. */
hipError_t noalloc_scratch(int device, uint2** out)
{
    static std::mutex mu;
    static uint2*     addr[CLPT_MAX_DEVICES] = {};
    if (device < 0 || device >= CLPT_MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device])
    {
        void*      p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        addr[device] = static_cast<uint2*>(p);
    }
    *out = addr[device];
    return hipSuccess;
}

/* The batch trace kernel on the fixed grid the scratch is laid out for. */
hipError_t launch_trace_noalloc(hipStream_t s, const SceneView& sc, const float4* rays, uint32_t n, int any,
                                float4* out, uint32_t* prim, uint2* scratch_addr)
{
    if (sc.bvh_width == 8)
        hipLaunchKernelGGL(k_trace_batch<8>, dim3(CLPT_NOALLOC_BLOCKS), dim3(256), 0, s, sc, rays, n, any, out, prim,
                           scratch_addr);
    else
        hipLaunchKernelGGL(k_trace_batch<4>, dim3(CLPT_NOALLOC_BLOCKS), dim3(256), 0, s, sc, rays, n, any, out, prim,
                           scratch_addr);
    return hipGetLastError();
}
} // namespace clpt

/* ------------------------------------------------------------ launchers */

namespace clpt
{
static inline uint32_t grid_for(uint64_t items, uint32_t block, uint32_t max_blocks)
{
    uint64_t g = (items + block - 1) / block;
    if (g < 1)
        g = 1;
    if (g > max_blocks)
        g = max_blocks;
    return (uint32_t)g;
}

hipError_t launch_camera(hipStream_t s, const BatchTable& bt, const GuideView& gv, const RayQueue& q, BatchCounters* bc,
                         uint32_t write_o, uint32_t threads, uint32_t blocks)
{
    hipLaunchKernelGGL(k_camera, dim3(blocks), dim3(threads), 0, s, bt, gv, q, bc, write_o);
    return hipGetLastError();
}

hipError_t launch_trace(hipStream_t s, const SceneView& sc, const RayQueue& q, const uint32_t* count, uint32_t depth,
                        const HitBuf& hb, uint2* gstack, unsigned long long* ctr, const float* cam_o, uint32_t blocks)
{
    if (sc.bvh_width == 8)
        hipLaunchKernelGGL(k_trace<8>, dim3(blocks), dim3(256), 0, s, sc, q, count, depth, hb, gstack, ctr, cam_o);
    else
        hipLaunchKernelGGL(k_trace<4>, dim3(blocks), dim3(256), 0, s, sc, q, count, depth, hb, gstack, ctr, cam_o);
    return hipGetLastError();
}

hipError_t launch_shade(hipStream_t s, const SceneView& sc, const BatchTable& bt, uint32_t depth, const RayQueue& qin,
                        const HitBuf& hb, const RayQueue& qout, const ShadowQueue& sq, float* pacc, BatchCounters* bc,
                        uint32_t threads, uint32_t blocks)
{
    hipLaunchKernelGGL(k_shade, dim3(blocks), dim3(threads), 0, s, sc, bt, depth, qin, hb, qout, sq, pacc, bc);
    return hipGetLastError();
}

hipError_t launch_shade_guided(hipStream_t s, const SceneView& sc, const BatchTable& bt, const GuideView& gv,
                               uint32_t depth, const RayQueue& qin, const HitBuf& hb, const RayQueue& qout,
                               const ShadowQueue& sq, float* pacc, BatchCounters* bc, uint32_t blocks)
{
    if (gv.img8)
        hipLaunchKernelGGL(k_shade_guided<true>, dim3(blocks), dim3(256), 0, s, sc, bt, gv, depth, qin, hb, qout, sq,
                           pacc, bc);
    else
        hipLaunchKernelGGL(k_shade_guided<false>, dim3(blocks), dim3(256), 0, s, sc, bt, gv, depth, qin, hb, qout, sq,
                           pacc, bc);
    return hipGetLastError();
}

hipError_t launch_shadow(hipStream_t s, const SceneView& sc, const ShadowQueue& sq, uint32_t depth, double* accum,
                         BatchCounters* bc, uint2* gstack, unsigned long long* ctr, float4* guide_E, uint32_t blocks)
{
    if (sc.bvh_width == 8)
        hipLaunchKernelGGL(k_shadow<8>, dim3(blocks), dim3(256), 0, s, sc, sq, depth, accum, bc, gstack, ctr, guide_E);
    else
        hipLaunchKernelGGL(k_shadow<4>, dim3(blocks), dim3(256), 0, s, sc, sq, depth, accum, bc, gstack, ctr, guide_E);
    return hipGetLastError();
}

hipError_t launch_film_fold(hipStream_t s, const BatchTable& bt, const float* pacc, double* accum, uint32_t blocks)
{
    hipLaunchKernelGGL(k_film_fold, dim3(blocks), dim3(256), 0, s, bt, pacc, accum);
    return hipGetLastError();
}

hipError_t launch_batch_end(hipStream_t s, const BatchCounters* bc, TotalCounters* tot, uint32_t* cursor)
{
    hipLaunchKernelGGL(k_batch_end, dim3(1), dim3(64), 0, s, bc, tot, cursor);
    return hipGetLastError();
}

hipError_t launch_set_cursor(hipStream_t s, uint32_t* cursor, uint32_t value)
{
    hipLaunchKernelGGL(k_set_cursor, dim3(1), dim3(64), 0, s, cursor, value);
    return hipGetLastError();
}

hipError_t launch_resolve(hipStream_t s, const double* accum, uint32_t npix, uint32_t spp, int srgb, float4* out)
{
    hipLaunchKernelGGL(k_resolve, dim3(grid_for(npix, 256, 4096)), dim3(256), 0, s, accum, npix, spp, srgb, out);
    return hipGetLastError();
}

hipError_t launch_srgb_batch(hipStream_t s, const float* in, uint32_t n, int decode, float* out)
{
    hipLaunchKernelGGL(k_srgb_batch, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, in, n, decode, out);
    return hipGetLastError();
}

hipError_t launch_texture_u8(hipStream_t s, const uchar4* in, uint32_t n, int srgb, float4* out)
{
    hipLaunchKernelGGL(k_texture_u8, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, in, n, srgb, out);
    return hipGetLastError();
}

hipError_t launch_dome_tables(hipStream_t s, const float4* tex, uint32_t w, uint32_t h, float* pcond, float* cdfcond,
                              float* pv, float* cdfv, double* rowint, uint32_t* cutv, uint32_t* cutcond)
{
    hipLaunchKernelGGL(k_dome_prefilter, dim3(grid_for((uint64_t)w * h, 256, 8192)), dim3(256), 0, s, tex, w, h, pcond);
    hipLaunchKernelGGL(k_dome_rows, dim3(h), dim3(256), 0, s, w, h, pcond, cdfcond, pv, rowint);
    hipLaunchKernelGGL(k_dome_marginal, dim3(1), dim3(256), 0, s, h, pv, cdfv, rowint);
    hipLaunchKernelGGL(k_dome_cuts, dim3(h + 1), dim3(256), 0, s, w, h, cdfcond, cdfv, cutv, cutcond);
    return hipGetLastError();
}

hipError_t launch_rng_batch(hipStream_t s, const uint4* seeds, uint32_t n, uint32_t ndraws, uint32_t* ou, float* of)
{
    hipLaunchKernelGGL(k_rng_batch, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, seeds, n, ndraws, ou, of);
    return hipGetLastError();
}

hipError_t launch_camera_batch(hipStream_t s, const ViewRec& fr, const uint2* ps, uint32_t n, float4* out)
{
    hipLaunchKernelGGL(k_camera_batch, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, fr, ps, n, out);
    return hipGetLastError();
}

hipError_t launch_bsdf_batch(hipStream_t s, const SceneView& sc, const uint32_t* mtl, const float* wi3,
                             const float* wo3, const float* u3, uint32_t n, float* ef3, float* epdf, float* swi3,
                             float* sf3, float* spdf)
{
    hipLaunchKernelGGL(k_bsdf_batch, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, sc, mtl, wi3, wo3, u3, n, ef3, epdf,
                       swi3, sf3, spdf);
    return hipGetLastError();
}

hipError_t launch_emitter_batch(hipStream_t s, const SceneView& sc, const float* u2, uint32_t n, float* dir3,
                                float* val3, float* pdf, uint32_t* which)
{
    hipLaunchKernelGGL(k_emitter_batch, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, sc, u2, n, dir3, val3, pdf, which);
    return hipGetLastError();
}

hipError_t launch_dome_eval_batch(hipStream_t s, const SceneView& sc, const float* dir3, uint32_t n, float* le3,
                                  float* pdf)
{
    hipLaunchKernelGGL(k_dome_eval_batch, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, sc, dir3, n, le3, pdf);
    return hipGetLastError();
}

uint32_t resident_blocks_per_cu(int which)
{
    int n = 0;
    hipError_t e;
    if (which == 0)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace<4>, 256, 0);
    else if (which == 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_shadow<4>, 256, 0);
    else if (which == 2)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_shade, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_shade_guided<false>, 256, 0);
    return (e == hipSuccess && n > 0) ? (uint32_t)n : 4u;
}

hipError_t launch_copy_probe(hipStream_t s, const float4* src, float4* dst, size_t n, uint32_t blocks)
{
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_trace_batch(hipStream_t s, const SceneView& sc, const float4* rays, uint32_t n, int any, float4* out,
                              uint32_t* prim, uint2* gstack, uint32_t blocks)
{
    if (sc.bvh_width == 8)
        hipLaunchKernelGGL(k_trace_batch<8>, dim3(blocks), dim3(256), 0, s, sc, rays, n, any, out, prim, gstack);
    else
        hipLaunchKernelGGL(k_trace_batch<4>, dim3(blocks), dim3(256), 0, s, sc, rays, n, any, out, prim, gstack);
    return hipGetLastError();
}
} // namespace clpt
