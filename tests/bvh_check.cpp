/* Host check of the BVH4 and BVH8 builders (colvillea_amd/csrc/bvh_build.cpp):
 * every child box (decoded, for BVH8) contains the exact bounds of its
 * subtree, empty BVH4 slots hold inverted infinite boxes, every leaf holds
 * at most leaf_max primitives, and every primitive appears in exactly one
 * leaf.  Prints "ok" or the first failures. */
#include "../colvillea_amd/csrc/bvh_build.h"

#include <cstdio>
#include <random>
#include <vector>

using namespace clpt;

static int      fail_count = 0;
static uint32_t leaf_max   = 4;

/* the primitives of a leaf: marks them seen and returns their exact bounds */
static void leaf_bounds(const std::vector<uint32_t>& order, const float* lo, const float* hi, uint32_t ref,
                        float out_lo[3], float out_hi[3], std::vector<int>& seen)
{
    const uint32_t first = ref & (CLPT_MAX_TRIS - 1u), count = ((ref >> 27) & 7u) + 1u;
    if (count > leaf_max && fail_count++ < 5)
        std::printf("leaf of %u primitives, at most %u allowed\n", count, leaf_max);
    for (uint32_t k = 0; k < count; ++k)
    {
        const uint32_t p = order[first + k];
        seen[p]++;
        for (int a = 0; a < 3; ++a)
        {
            out_lo[a] = std::min(out_lo[a], lo[3 * p + a]);
            out_hi[a] = std::max(out_hi[a], hi[3 * p + a]);
        }
    }
}

/* exact bounds of the subtree under ref (leaf refs encode first/count) */
static void bounds(const std::vector<Node8>& nodes, const std::vector<uint32_t>& order, const float* lo,
                   const float* hi, uint32_t ref, float out_lo[3], float out_hi[3], std::vector<int>& seen)
{
    for (int a = 0; a < 3; ++a)
    {
        out_lo[a] = 3.0e38f;
        out_hi[a] = -3.0e38f;
    }
    if (ref & CLPT_LEAF_BIT)
    {
        leaf_bounds(order, lo, hi, ref, out_lo, out_hi, seen);
        return;
    }
    const Node8& nd = nodes[ref];
    for (int c = 0; c < 8; ++c)
    {
        if (nd.child[c] == CLPT_CHILD_EMPTY)
            continue;
        float cl[3], ch[3];
        bounds(nodes, order, lo, hi, nd.child[c], cl, ch, seen);
        for (int a = 0; a < 3; ++a)
        {
            if (!(node8_lo(nd, a, c) <= cl[a] && node8_hi(nd, a, c) >= ch[a]))
            {
                if (fail_count++ < 5)
                    std::printf("child box not conservative: node %u child %d axis %d: [%.9g, %.9g] vs [%.9g, %.9g]\n",
                                ref, c, a, node8_lo(nd, a, c), node8_hi(nd, a, c), cl[a], ch[a]);
            }
            out_lo[a] = std::min(out_lo[a], cl[a]);
            out_hi[a] = std::max(out_hi[a], ch[a]);
        }
    }
}

static void bounds4(const std::vector<Node4>& nodes, const std::vector<uint32_t>& order, const float* lo,
                    const float* hi, uint32_t ref, float out_lo[3], float out_hi[3], std::vector<int>& seen)
{
    for (int a = 0; a < 3; ++a)
    {
        out_lo[a] = 3.0e38f;
        out_hi[a] = -3.0e38f;
    }
    if (ref & CLPT_LEAF_BIT)
    {
        leaf_bounds(order, lo, hi, ref, out_lo, out_hi, seen);
        return;
    }
    const Node4& nd = nodes[ref];
    for (int c = 0; c < 4; ++c)
    {
        const float bl[3] = {nd.lox[c], nd.loy[c], nd.loz[c]}, bh[3] = {nd.hix[c], nd.hiy[c], nd.hiz[c]};
        if (nd.child[c] == CLPT_CHILD_EMPTY)
        {
            for (int a = 0; a < 3; ++a)
                if (!(bl[a] == INFINITY && bh[a] == -INFINITY) && fail_count++ < 5)
                    std::printf("empty slot %d of node %u: box not inverted infinite\n", c, ref);
            continue;
        }
        float cl[3], ch[3];
        bounds4(nodes, order, lo, hi, nd.child[c], cl, ch, seen);
        for (int a = 0; a < 3; ++a)
        {
            if (!(bl[a] <= cl[a] && bh[a] >= ch[a]) && fail_count++ < 5)
                std::printf("child box does not contain its subtree: node %u child %d axis %d\n", ref, c, a);
            out_lo[a] = std::min(out_lo[a], cl[a]);
            out_hi[a] = std::max(out_hi[a], ch[a]);
        }
    }
}

static void run(uint32_t n, float centre, float spread, float size, uint32_t seed)
{
    std::mt19937                          g(seed);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f), s(0.0f, 1.0f);
    std::vector<float>                    lo(3 * n), hi(3 * n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a)
        {
            const float c = centre + spread * u(g);
            const float e = (i % 17 == 0) ? 0.0f : size * s(g); /* some degenerate boxes */
            lo[3 * i + a] = c - e;
            hi[3 * i + a] = c + e;
        }
    const auto leaf = [](uint32_t first, uint32_t count) { return CLPT_TRI_LEAF(first, count); };
    /* the library's mesh trees (CLPT_MAX_LEAF_TRIS), instance trees (1) and a wider cap */
    for (const uint32_t lm : {(uint32_t)CLPT_MAX_LEAF_TRIS, 1u, 4u})
    {
        leaf_max = lm;
        std::vector<Node8>    nodes8;
        std::vector<Node4>    nodes4;
        std::vector<uint32_t> order8, order4;
        const BvhResult       r8 = build_bvh8(lo.data(), hi.data(), n, lm, nodes8, order8, leaf);
        const BvhResult       r4 = build_bvh4(lo.data(), hi.data(), n, lm, nodes4, order4, leaf);
        std::vector<int>      seen8(n, 0), seen4(n, 0);
        float                 bl[3], bh[3];
        bounds(nodes8, order8, lo.data(), hi.data(), r8.root, bl, bh, seen8);
        bounds4(nodes4, order4, lo.data(), hi.data(), r4.root, bl, bh, seen4);
        for (uint32_t i = 0; i < n; ++i)
            if ((seen8[i] != 1 || seen4[i] != 1) && fail_count++ < 5)
                std::printf("primitive %u seen %d times (BVH8), %d times (BVH4), leaf cap %u\n", i, seen8[i], seen4[i], lm);
    }
}

int main()
{
    run(1, 0.0f, 1.0f, 0.1f, 1);
    run(7, 0.0f, 1.0f, 0.1f, 2);
    run(5000, 0.0f, 10.0f, 0.05f, 3);
    run(20000, 1.0e4f, 3.0f, 0.01f, 4);   /* far from the origin, small extent */
    run(20000, -3.0e-3f, 1.0e-3f, 1e-6f, 5); /* tiny */
    run(3000, 0.0f, 1.0e6f, 5.0e5f, 6);   /* huge, overlapping */
    std::printf(fail_count ? "FAILED %d\n" : "ok\n", fail_count);
    return fail_count ? 1 : 0;
}
