/*
 * bvh_build.cpp -- host BVH builder replacing OptiX's BLAS/TLAS build
 * (OptiXDevice::buildOptiXAccelBLASes/TLAS, optixdevice.cpp:136-233).
 *
 * Binned SAH (32 bins per axis) produces a binary tree, which is collapsed
 * to a 4- or 8-wide tree by repeatedly opening the child with the largest
 * surface area.  BVH4 nodes are 128-byte Node4 records with float boxes
 * (eight 16-byte loads); BVH8 nodes are 96-byte Node8 records with child
 * boxes quantised to bytes against the node's box, in the spirit of Ylitie,
 * Karras and Laine, "Efficient Incoherent Ray Traversal on GPUs Through
 * Compressed Wide BVHs", HPG 2017 (six 16-byte loads).
 */
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace clpt
{
namespace
{
struct Node2
{
    float    lo[3], hi[3];
    uint32_t left = 0, right = 0; /* inner */
    uint32_t first = 0, count = 0; /* leaf: count > 0 */
};

struct Box
{
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void  grow(const float* l, const float* h)
    {
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
    void grow(const Box& b) { grow(b.lo, b.hi); }
    double area() const
    {
        if (!(hi[0] >= lo[0]))
            return 0.0;
        const double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

constexpr int      kBins         = 32;
constexpr uint32_t kMaxSahDepth  = 48; /* deeper subtrees split at the median */
constexpr double   kTraverseCost = 1.0;

struct Builder
{
    const float*            plo;
    const float*            phi;
    std::vector<float>      cen;
    std::vector<uint32_t>&  order;
    std::vector<Node2>      n2;
    uint32_t                leaf_max;

    Builder(const float* l, const float* h, uint32_t n, uint32_t lm, std::vector<uint32_t>& ord) :
        plo(l), phi(h), order(ord), leaf_max(lm)
    {
        cen.resize(3 * (size_t)n);
        for (uint32_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k)
                cen[3 * (size_t)i + k] = 0.5f * (l[3 * (size_t)i + k] + h[3 * (size_t)i + k]);
        order.resize(n);
        std::iota(order.begin(), order.end(), 0u);
    }

    void build(uint32_t n)
    {
        struct Item
        {
            uint32_t node, start, end, depth;
        };
        n2.reserve(2 * (size_t)n + 1);
        n2.emplace_back();
        std::vector<Item> stack{{0u, 0u, n, 0u}};
        while (!stack.empty())
        {
            const Item it = stack.back();
            stack.pop_back();
            Box bb, cb;
            for (uint32_t i = it.start; i < it.end; ++i)
            {
                const uint32_t p = order[i];
                bb.grow(plo + 3 * (size_t)p, phi + 3 * (size_t)p);
                cb.grow(&cen[3 * (size_t)p], &cen[3 * (size_t)p]);
            }
            Node2& nd = n2[it.node];
            std::memcpy(nd.lo, bb.lo, sizeof(nd.lo));
            std::memcpy(nd.hi, bb.hi, sizeof(nd.hi));
            const uint32_t cnt = it.end - it.start;
            if (cnt <= 1 || (cnt <= leaf_max && it.depth >= kMaxSahDepth))
            {
                nd.first = it.start;
                nd.count = cnt;
                continue;
            }
            /* binned SAH */
            int    best_axis = -1, best_bin = -1;
            double best_cost = INFINITY;
            if (it.depth < kMaxSahDepth)
            {
                for (int ax = 0; ax < 3; ++ax)
                {
                    const float ext = cb.hi[ax] - cb.lo[ax];
                    if (!(ext > 0.0f))
                        continue;
                    Box      bin_box[kBins];
                    uint32_t bin_cnt[kBins] = {};
                    const float scale = (float)kBins / ext;
                    for (uint32_t i = it.start; i < it.end; ++i)
                    {
                        const uint32_t p = order[i];
                        int b = (int)((cen[3 * (size_t)p + ax] - cb.lo[ax]) * scale);
                        b     = std::min(std::max(b, 0), kBins - 1);
                        bin_cnt[b]++;
                        bin_box[b].grow(plo + 3 * (size_t)p, phi + 3 * (size_t)p);
                    }
                    double   left_area[kBins];
                    uint32_t left_cnt[kBins];
                    Box      acc;
                    uint32_t c = 0;
                    for (int b = 0; b < kBins; ++b)
                    {
                        acc.grow(bin_box[b]);
                        c += bin_cnt[b];
                        left_area[b] = acc.area();
                        left_cnt[b]  = c;
                    }
                    Box      racc;
                    uint32_t rc = 0;
                    for (int b = kBins - 1; b > 0; --b)
                    {
                        racc.grow(bin_box[b]);
                        rc += bin_cnt[b];
                        if (left_cnt[b - 1] == 0 || rc == 0)
                            continue;
                        const double cost = left_area[b - 1] * left_cnt[b - 1] + racc.area() * rc;
                        if (cost < best_cost)
                        {
                            best_cost = cost;
                            best_axis = ax;
                            best_bin  = b - 1;
                        }
                    }
                }
            }
            const double area = bb.area();
            if (cnt <= leaf_max)
            {
                const double split_cost = kTraverseCost + (area > 0.0 ? best_cost / area : 0.0);
                if (best_axis < 0 || (double)cnt <= split_cost)
                {
                    nd.first = it.start;
                    nd.count = cnt;
                    continue;
                }
            }
            uint32_t mid = it.start;
            if (best_axis >= 0)
            {
                const int   ax    = best_axis;
                const float scale = (float)kBins / (cb.hi[ax] - cb.lo[ax]);
                const float lo    = cb.lo[ax];
                const int   sb    = best_bin;
                auto        pm    = std::partition(order.begin() + it.start, order.begin() + it.end, [&](uint32_t p) {
                    int b = (int)((cen[3 * (size_t)p + ax] - lo) * scale);
                    b     = std::min(std::max(b, 0), kBins - 1);
                    return b <= sb;
                });
                mid = (uint32_t)(pm - order.begin());
            }
            if (mid == it.start || mid == it.end)
            {
                int ax = 0;
                for (int k = 1; k < 3; ++k)
                    if (cb.hi[k] - cb.lo[k] > cb.hi[ax] - cb.lo[ax])
                        ax = k;
                mid = it.start + cnt / 2;
                std::nth_element(order.begin() + it.start, order.begin() + mid, order.begin() + it.end,
                                 [&](uint32_t a, uint32_t b) { return cen[3 * (size_t)a + ax] < cen[3 * (size_t)b + ax]; });
            }
            const uint32_t l = (uint32_t)n2.size();
            n2.emplace_back();
            n2.emplace_back();
            n2[it.node].left  = l;
            n2[it.node].right = l + 1;
            n2[it.node].count = 0;
            stack.push_back({l, it.start, mid, it.depth + 1});
            stack.push_back({l + 1, mid, it.end, it.depth + 1});
        }
    }
};

double node_area(const Node2& n)
{
    Box b;
    b.grow(n.lo, n.hi);
    return b.area();
}

float pow2_scale(uint8_t e)
{
    uint32_t u = (uint32_t)e << 23;
    float    f;
    std::memcpy(&f, &u, 4);
    return f;
}
} // namespace

float node8_lo(const Node8& nd, int a, int c)
{
    const volatile float t = (float)nd.qlo[a][c] * pow2_scale(nd.e[a]); /* exact */
    return nd.p[a] + t;
}
float node8_hi(const Node8& nd, int a, int c)
{
    const volatile float t = (float)nd.qhi[a][c] * pow2_scale(nd.e[a]);
    return nd.p[a] + t;
}

namespace
{
/* Quantise the child boxes of a node against its own box [nlo, nhi]: per
 * axis the smallest power-of-two scale with 255 steps covering the extent,
 * then per child the largest lower byte and smallest upper byte whose
 * decoded bounds (float32, as the device computes them) still contain the
 * child; the scale grows if a child cannot be covered. */
void quantise(Node8& nd, const float nlo[3], const float nhi[3], const float (*clo)[3], const float (*chi)[3], int nk)
{
    for (int a = 0; a < 3; ++a)
    {
        nd.p[a]          = nlo[a];
        const double ext = (double)nhi[a] - (double)nlo[a];
        int          e   = 127;
        if (ext > 0.0)
            e = 127 + (int)std::ceil(std::log2(ext / 255.0));
        e = std::min(std::max(e, 1), 254);
        for (;; ++e)
        {
            nd.e[a]   = (uint8_t)e;
            const double sc = (double)pow2_scale((uint8_t)e);
            bool         ok = true;
            for (int c = 0; c < nk && ok; ++c)
            {
                double ql = std::floor(((double)clo[c][a] - nlo[a]) / sc);
                double qh = std::ceil(((double)chi[c][a] - nlo[a]) / sc);
                ql        = std::min(std::max(ql, 0.0), 255.0);
                qh        = std::min(std::max(qh, 0.0), 255.0);
                nd.qlo[a][c] = (uint8_t)ql;
                nd.qhi[a][c] = (uint8_t)qh;
                while (node8_lo(nd, a, c) > clo[c][a] && nd.qlo[a][c] > 0)
                    --nd.qlo[a][c];
                while (node8_hi(nd, a, c) < chi[c][a] && nd.qhi[a][c] < 255)
                    ++nd.qhi[a][c];
                ok = node8_lo(nd, a, c) <= clo[c][a] && node8_hi(nd, a, c) >= chi[c][a];
            }
            if (ok || e >= 254)
                break;
        }
        for (int c = nk; c < 8; ++c)
        {
            nd.qlo[a][c] = 255;
            nd.qhi[a][c] = 0;
        }
    }
}

template <typename NodeT, int W>
struct Collapser
{
    const std::vector<Node2>&                               n2;
    std::vector<NodeT>&                                     nodes;
    const std::function<uint32_t(uint32_t, uint32_t)>&     leaf_ref;
    uint32_t                                                max_depth = 0;
    double                                                  sah       = 0.0;
    double                                                  root_area = 1.0;

    /* The node for BVH2 node idx: its children are found by repeatedly
     * opening the internal child of largest surface area until W remain. */
    uint32_t make(uint32_t idx, uint32_t depth)
    {
        max_depth = std::max(max_depth, depth);
        uint32_t kids[W];
        int      nk = 0;
        if (n2[idx].count > 0)
            kids[nk++] = idx;
        else
        {
            kids[nk++] = n2[idx].left;
            kids[nk++] = n2[idx].right;
        }
        while (nk < W)
        {
            int    pick = -1;
            double pa   = -1.0;
            for (int k = 0; k < nk; ++k)
            {
                if (n2[kids[k]].count > 0)
                    continue;
                const double a = node_area(n2[kids[k]]);
                if (a > pa)
                {
                    pa   = a;
                    pick = k;
                }
            }
            if (pick < 0)
                break;
            const uint32_t open = kids[pick];
            kids[pick]          = n2[open].left;
            kids[nk++]          = n2[open].right;
        }
        const uint32_t me = (uint32_t)nodes.size();
        nodes.emplace_back();
        NodeT nd;
        std::memset(&nd, 0, sizeof(nd));
        float    clo[W][3], chi[W][3];
        uint32_t ref[W];
        for (int c = 0; c < W; ++c)
            ref[c] = CLPT_CHILD_EMPTY;
        sah += kTraverseCost * node_area(n2[idx]) / root_area;
        for (int c = 0; c < nk; ++c)
        {
            const Node2& k = n2[kids[c]];
            for (int a = 0; a < 3; ++a)
            {
                clo[c][a] = k.lo[a];
                chi[c][a] = k.hi[a];
            }
            if (k.count > 0)
            {
                ref[c] = leaf_ref(k.first, k.count);
                sah += (double)k.count * node_area(k) / root_area;
            }
            else
                ref[c] = make(kids[c], depth + 1);
        }
        fill(nd, n2[idx], clo, chi, ref, nk);
        nodes[me] = nd;
        return me;
    }

    static void fill(Node4& nd, const Node2&, const float (*clo)[3], const float (*chi)[3], const uint32_t* ref, int nk)
    {
        for (int c = 0; c < 4; ++c)
        {
            nd.child[c] = ref[c];
            if (c >= nk)
            {
                /* an empty child's box is inverted and infinite: each slab's
                 * near distance is +inf and its far distance -inf for every
                 * ray, so the slab test itself rejects it */
                nd.lox[c] = nd.loy[c] = nd.loz[c] = INFINITY;
                nd.hix[c] = nd.hiy[c] = nd.hiz[c] = -INFINITY;
                continue;
            }
            nd.lox[c] = clo[c][0];
            nd.loy[c] = clo[c][1];
            nd.loz[c] = clo[c][2];
            nd.hix[c] = chi[c][0];
            nd.hiy[c] = chi[c][1];
            nd.hiz[c] = chi[c][2];
        }
    }
    static void fill(Node8& nd, const Node2& self, const float (*clo)[3], const float (*chi)[3], const uint32_t* ref,
                     int nk)
    {
        for (int c = 0; c < 8; ++c)
            nd.child[c] = ref[c];
        quantise(nd, self.lo, self.hi, clo, chi, nk);
    }
};

template <typename NodeT, int W>
BvhResult build_wide(const float* lo, const float* hi, uint32_t n, uint32_t leaf_max, std::vector<NodeT>& nodes,
                     std::vector<uint32_t>& order, const std::function<uint32_t(uint32_t, uint32_t)>& leaf_ref)
{
    BvhResult res;
    Builder   b(lo, hi, n, leaf_max, order);
    if (n == 0)
    {
        /* an empty tree: one node with no children */
        NodeT nd;
        std::memset(&nd, 0, sizeof(nd));
        for (int c = 0; c < W; ++c)
            nd.child[c] = CLPT_CHILD_EMPTY;
        res.root = (uint32_t)nodes.size();
        nodes.push_back(nd);
        return res;
    }
    b.build(n);
    Collapser<NodeT, W> c{b.n2, nodes, leaf_ref};
    c.root_area   = std::max(node_area(b.n2[0]), 1e-30);
    res.root      = c.make(0, 1);
    res.max_depth = c.max_depth;
    res.sah_cost  = c.sah;
    std::memcpy(res.lo, b.n2[0].lo, sizeof(res.lo));
    std::memcpy(res.hi, b.n2[0].hi, sizeof(res.hi));
    return res;
}
} // namespace

BvhResult build_bvh4(const float* lo, const float* hi, uint32_t n, uint32_t leaf_max, std::vector<Node4>& nodes,
                     std::vector<uint32_t>& order, const std::function<uint32_t(uint32_t, uint32_t)>& leaf_ref)
{
    return build_wide<Node4, 4>(lo, hi, n, leaf_max, nodes, order, leaf_ref);
}

BvhResult build_bvh8(const float* lo, const float* hi, uint32_t n, uint32_t leaf_max, std::vector<Node8>& nodes,
                     std::vector<uint32_t>& order, const std::function<uint32_t(uint32_t, uint32_t)>& leaf_ref)
{
    return build_wide<Node8, 8>(lo, hi, n, leaf_max, nodes, order, leaf_ref);
}
} // namespace clpt
