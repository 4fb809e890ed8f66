/* bvh_build.h -- host BVH builders (binned SAH BVH2, collapsed to 4-wide
 * with float boxes, or to 8-wide with child boxes quantised to bytes). */
#ifndef CLPT_BVH_BUILD_H
#define CLPT_BVH_BUILD_H

#include "clpt_scene.h"

#include <cstdint>
#include <functional>
#include <vector>

namespace clpt
{
struct BvhResult
{
    uint32_t root      = 0; /* index into the node array (absolute) */
    uint32_t max_depth = 0; /* inner-node levels of the wide tree */
    float    lo[3]     = {0, 0, 0};
    float    hi[3]     = {0, 0, 0};
    double   sah_cost  = 0.0;
};

/* Build over n primitive boxes (lo/hi: 3 floats each).  Nodes are appended
 * to `nodes`; `order` receives the primitive permutation (leaf order);
 * leaf_ref(first, count) encodes a leaf of `count` primitives starting at
 * position `first` of `order`.  Leaves hold at most leaf_max primitives. */
BvhResult build_bvh4(const float* lo, const float* hi, uint32_t n, uint32_t leaf_max, std::vector<Node4>& nodes,
                     std::vector<uint32_t>& order, const std::function<uint32_t(uint32_t, uint32_t)>& leaf_ref);
BvhResult build_bvh8(const float* lo, const float* hi, uint32_t n, uint32_t leaf_max, std::vector<Node8>& nodes,
                     std::vector<uint32_t>& order, const std::function<uint32_t(uint32_t, uint32_t)>& leaf_ref);

/* Decoded box of child c on axis a, with the device's float32 operations. */
float node8_lo(const Node8& nd, int a, int c);
float node8_hi(const Node8& nd, int a, int c);
} // namespace clpt

#endif
