/* clpt_kernels.h -- host-side launchers of the kernels in kernels.hip. */
#ifndef CLPT_KERNELS_H
#define CLPT_KERNELS_H

#include "clpt_scene.h"

namespace clpt
{
/* `threads` per block (64, 128 or 256) for the camera and plain shading
 * kernels; write_o = 0 leaves the origin plane unwritten for a batch of one
 * view, whose closest-hit launch at depth 1 then takes cam_o (the view's
 * camera position in device memory) as every ray's origin */
hipError_t launch_camera(hipStream_t s, const BatchTable& bt, const GuideView& gv, const RayQueue& q, BatchCounters* bc,
                         uint32_t write_o, uint32_t threads, uint32_t blocks);
hipError_t launch_trace(hipStream_t s, const SceneView& sc, const RayQueue& q, const uint32_t* count, uint32_t depth,
                        const HitBuf& hb, uint2* gstack, unsigned long long* ctr, const float* cam_o, uint32_t blocks);
hipError_t launch_shade(hipStream_t s, const SceneView& sc, const BatchTable& bt, uint32_t depth, const RayQueue& qin,
                        const HitBuf& hb, const RayQueue& qout, const ShadowQueue& sq, float* pacc, BatchCounters* bc,
                        uint32_t threads, uint32_t blocks);
/* Shading with the learned proposal mixed into the continuation (weights in LDS). */
hipError_t launch_shade_guided(hipStream_t s, const SceneView& sc, const BatchTable& bt, const GuideView& gv,
                               uint32_t depth, const RayQueue& qin, const HitBuf& hb, const RayQueue& qout,
                               const ShadowQueue& sq, float* pacc, BatchCounters* bc, uint32_t blocks);
hipError_t launch_shadow(hipStream_t s, const SceneView& sc, const ShadowQueue& sq, uint32_t depth, double* accum,
                         BatchCounters* bc, uint2* gstack, unsigned long long* ctr, float4* guide_E, uint32_t blocks);
/* Escaped-ray radiance of the batch's paths (3 floats per path slot, written
 * by shading) into the float64 film, one thread per pixel and view. */
hipError_t launch_film_fold(hipStream_t s, const BatchTable& bt, const float* pacc, double* accum, uint32_t blocks);
hipError_t launch_batch_end(hipStream_t s, const BatchCounters* bc, TotalCounters* tot, uint32_t* cursor);
hipError_t launch_set_cursor(hipStream_t s, uint32_t* cursor, uint32_t value);
hipError_t launch_resolve(hipStream_t s, const double* accum, uint32_t npix, uint32_t spp, int srgb, float4* out);
hipError_t launch_srgb_batch(hipStream_t s, const float* in, uint32_t n, int decode, float* out);
hipError_t launch_texture_u8(hipStream_t s, const uchar4* in, uint32_t n, int srgb, float4* out);
hipError_t launch_dome_tables(hipStream_t s, const float4* tex, uint32_t w, uint32_t h, float* pcond, float* cdfcond,
                              float* pv, float* cdfv, double* rowint, uint32_t* cutv, uint32_t* cutcond);
hipError_t launch_rng_batch(hipStream_t s, const uint4* seeds, uint32_t n, uint32_t ndraws, uint32_t* ou, float* of);
hipError_t launch_camera_batch(hipStream_t s, const ViewRec& fr, const uint2* ps, uint32_t n, float4* out);
hipError_t launch_bsdf_batch(hipStream_t s, const SceneView& sc, const uint32_t* mtl, const float* wi3,
                             const float* wo3, const float* u3, uint32_t n, float* ef3, float* epdf, float* swi3,
                             float* sf3, float* spdf);
hipError_t launch_emitter_batch(hipStream_t s, const SceneView& sc, const float* u2, uint32_t n, float* dir3,
                                float* val3, float* pdf, uint32_t* which);
hipError_t launch_dome_eval_batch(hipStream_t s, const SceneView& sc, const float* dir3, uint32_t n, float* le3,
                                  float* pdf);
/* Blocks of 256 threads resident per CU for k_trace (0), k_shadow (1), k_shade (2),
 * k_shade_guided (3). */
uint32_t   resident_blocks_per_cu(int which);
hipError_t launch_copy_probe(hipStream_t s, const float4* src, float4* dst, size_t n, uint32_t blocks);
hipError_t launch_trace_batch(hipStream_t s, const SceneView& sc, const float4* rays, uint32_t n, int any, float4* out,
                              uint32_t* prim, uint2* gstack, uint32_t blocks);

/* learned proposal (guide.hip) */
struct GuideTrainBufs
{
    GuideSample*   samples;
    GuideCounters* gc;
    __bf16*        X;        /* N x 256 encoded inputs */
    __bf16*        act[4];   /* N x 64 hidden activations */
    float*         out;      /* N x 64 outputs */
    __bf16*        grad[5];  /* N x 64 gradients w.r.t. each layer's output */
    float*         loss_part;
    double*        wpart;    /* (N / 256) per-block sample-weight sums */
    float*         amax_part; /* 4 x (N / 64) per-block hidden-activation maxima */
    float*         partial;  /* (N / 1024) x 32768 split weight gradients */
    float*         params;
    float*         m1;
    float*         m2;
    __bf16*        W;
    __bf16*        WT;
    __bf16*        img;
    uint8_t*       img8; /* FP8 image refreshed after every step when not NULL */
    float*         scl;
    float*         spart;        /* fused step: per-block loss and activation maxima (5 floats) */
    uint32_t       fused_blocks; /* fused step with about this many blocks; 0 = one launch per GEMM */
    uint32_t       part_cap;     /* blocks that partial, wpart and spart hold */
};
/* Blocks and passes per block of the fused step for N samples and a target
 * block count; the per-block buffers (partial, wpart, spart) need that many
 * entries, at most CLPT_GUIDE_FUSED_MAXBLK. */
#define CLPT_GUIDE_FUSED_MAXBLK 1024u
uint32_t guide_fused_blocks(uint32_t N, uint32_t target, uint32_t* iters);
/* bf16 inference, or FP8 when img8 is not NULL (scales scl) */
/* The allocation-free trace entry point: its fixed grid, the per-device
 * address of its scratch (the stacks' overflow area), and the launch. */
#define CLPT_NOALLOC_BLOCKS 256u
#define CLPT_MAX_DEVICES 64
hipError_t noalloc_scratch(int device, uint2** out);
hipError_t launch_trace_noalloc(hipStream_t s, const SceneView& sc, const float4* rays, uint32_t n, int any,
                                float4* out, uint32_t* prim, uint2* scratch_addr);
hipError_t launch_guide_mixture_batch(hipStream_t s, const float* out40, const float* dir3, const float* u3, uint32_t n,
                                      float* pdf, float* smp3, float* smp_pdf);
hipError_t launch_guide_query_batch(hipStream_t s, const float lo[3], const float inv_ext[3], const float* p3,
                                    const float* n3, const float* wo3, uint32_t n, float* q9);
hipError_t launch_guide_infer(hipStream_t s, const __bf16* img, const uint8_t* img8, const float* scl, const float* q9,
                              uint32_t n, float* out40, uint32_t blocks);
hipError_t launch_guide_pack_fp8(hipStream_t s, const float* params, const GuideCounters* gc, uint8_t* img8,
                                 float* scl);
hipError_t launch_guide_pack(hipStream_t s, const float* params, __bf16* W, __bf16* WT, __bf16* img);
/* training samples of the recorded paths, appended in path-slot order */
hipError_t launch_guide_gather(hipStream_t s, const GuideView& gv, uint32_t maxd, GuideSample* out, uint32_t cap,
                               GuideCounters* gc);
/* Adam step `step` of `nsteps` on the gathered samples: rows step,
 * step + nsteps, ... of b.samples, up to N of them (count in b.gc; N a
 * multiple of 1024), so that each step draws from the whole batch.
 * The last step empties the buffer.  grad_out (optional) receives the
 * summed gradient. */
hipError_t launch_guide_train(hipStream_t s, const GuideTrainBufs& b, uint32_t N, float lr, float* grad_out,
                              uint32_t step, uint32_t nsteps);
} // namespace clpt

#endif
