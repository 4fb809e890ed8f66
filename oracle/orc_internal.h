/* orc_internal.h -- data shared by the two numeric variants of the oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#ifndef CLPT_ORC_INTERNAL_H
#define CLPT_ORC_INTERNAL_H

#include "oracle.h"

typedef struct
{
    float    lo[3], hi[3];
    uint32_t a; /* inner: index of the left child (right = a + 1); leaf: first entry in prims */
    uint32_t n; /* 0 for an inner node, primitive count for a leaf */
} orc_node;

typedef struct
{
    orc_node* nodes;
    uint32_t  nnodes;
    uint32_t* prims;
} orc_bvh;

struct orc_scene
{
    orc_scene_desc d;
    orc_mesh*      meshes;
    orc_instance*  instances;
    orc_material*  materials;
    orc_emitter*   emitters;
    orc_bvh*       blas; /* one per mesh */
    orc_bvh        tlas; /* over instance world boxes */
    float (*inv)[12];    /* world->object, float (the renderer's values) */
    double (*invd)[12];  /* world->object, float64 (the f64 reference) */
    float* sun_tan;      /* per emitter: tan(angular radius) as float */
};

typedef void (*orc_row_fn)(void* ctx, uint32_t row, uint32_t worker);
void orc_parallel_rows(uint32_t nthreads, uint32_t y0, uint32_t y1, orc_row_fn fn, void* ctx);

#endif
