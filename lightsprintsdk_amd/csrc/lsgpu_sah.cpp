// lsgpu_sah.cpp -- host-side SAH rebuild of the LBVH's inner nodes.
//
// The GPU LBVH (Morton order, Karras split at the highest differing bit) makes leaves of
// at most leaf_size consecutive triangles in Morton order; its inner nodes follow the
// Morton bits, not the geometry, so rays visit more nodes than a surface-area-heuristic
// tree would.  This pass keeps the leaves (their triangle ranges and their inflated boxes)
// and rebuilds the inner nodes over them top-down with a binned SAH (32 bins per axis on
// leaf-box centroids, cost = area(left)·n(left) + area(right)·n(right), the cheapest split
// over the three axes; equal centroids split the range in half).  The tree replaces the
// reference's BSP build (IntersectBsp.h, createAndSaveBsp), which is SAH-driven as well.
//
// Node layout as lsg_collider_download: 16 floats per inner node, child 0 box
// (lo.x hi.x lo.y hi.y lo.z hi.z), child 1 box, child refs as int bits, two spare.
// Output nodes are numbered in preorder (root = 0, left subtree before right), so the
// result depends only on the input tree.
#include "lsgpu.h"
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace lsg {

namespace {

constexpr int kBins = 32;

struct Box {
    float b[6];
    void empty() { b[0] = b[2] = b[4] = 3.0e38f; b[1] = b[3] = b[5] = -3.0e38f; }
    void grow(const float* o)
    {
        for (int a = 0; a < 3; a++) {
            b[2 * a] = std::min(b[2 * a], o[2 * a]);
            b[2 * a + 1] = std::max(b[2 * a + 1], o[2 * a + 1]);
        }
    }
    float area() const
    {
        const float dx = b[1] - b[0], dy = b[3] - b[2], dz = b[5] - b[4];
        return dx * dy + dy * dz + dz * dx;
    }
};

struct Leaf { int32_t ref; float b[6]; float c[3]; };

int32_t ref_of(const float* q, int k) { int32_t r; memcpy(&r, q + 12 + k, 4); return r; }

} // namespace

// Rebuilds the inner nodes of the BVH2 (nodes, root) over its leaves.  out must hold
// (leaves - 1) * 16 floats (at most nnodes * 16).  Returns 0 and the new root, node count
// and depth (inner nodes on the longest root path, root = 1); returns 1 when the tree has
// fewer than two leaves (nothing to rebuild), 2 on a malformed tree.
int sah_rebuild(const float* nodes, uint32_t nnodes, int32_t root, float* out, uint32_t* out_nodes, int32_t* out_root,
                uint32_t* out_depth)
{
    if (root < 0 || !nnodes) return 1;
    // leaves in depth-first order, left child first
    std::vector<Leaf> L;
    L.reserve(nnodes + 1);
    std::vector<int32_t> st{root};
    uint64_t visited = 0;
    while (!st.empty()) {
        const int32_t i = st.back(); st.pop_back();
        if ((uint32_t)i >= nnodes || ++visited > nnodes) return 2;
        const float* q = nodes + 16 * (size_t)i;
        for (int k = 0; k < 2; k++) {
            const int32_t r = ref_of(q, k);
            if (r >= 0) continue;
            Leaf f;
            f.ref = r;
            memcpy(f.b, q + 6 * k, 24);
            for (int a = 0; a < 3; a++) f.c[a] = (f.b[2 * a] + f.b[2 * a + 1]) * 0.5f;
            L.push_back(f);
        }
        for (int k = 1; k >= 0; k--)               // right pushed first: left is taken first
            if (ref_of(q, k) >= 0) st.push_back(ref_of(q, k));
    }
    const uint32_t n = (uint32_t)L.size();
    if (n < 2) return 1;
    std::vector<uint32_t> ids(n), tmp(n);
    for (uint32_t k = 0; k < n; k++) ids[k] = k;
    struct Task { uint32_t begin, end; int32_t parent; int side; uint32_t depth; };
    std::vector<Task> work{{0u, n, -1, 0, 1u}};
    uint32_t count = 0, depth = 0;
    while (!work.empty()) {
        const Task t = work.back(); work.pop_back();
        Box box; box.empty();
        for (uint32_t k = t.begin; k < t.end; k++) box.grow(L[ids[k]].b);
        int32_t ref;
        if (t.end - t.begin == 1) ref = L[ids[t.begin]].ref;
        else ref = (int32_t)count++;
        if (t.parent >= 0) {
            float* p = out + 16 * (size_t)t.parent;
            memcpy(p + 6 * t.side, box.b, 24);
            memcpy(p + 12 + t.side, &ref, 4);
        }
        if (t.end - t.begin == 1) continue;
        depth = std::max(depth, t.depth);
        float* q = out + 16 * (size_t)ref;
        q[14] = 0.f; q[15] = 0.f;
        // centroid bounds
        float clo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, chi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
        for (uint32_t k = t.begin; k < t.end; k++)
            for (int a = 0; a < 3; a++) { clo[a] = std::min(clo[a], L[ids[k]].c[a]); chi[a] = std::max(chi[a], L[ids[k]].c[a]); }
        float best = 0.f;
        int bax = -1, bk = -1;
        for (int a = 0; a < 3; a++) {
            if (!(chi[a] > clo[a])) continue;
            const float s = (float)kBins / (chi[a] - clo[a]);
            uint32_t cnt[kBins] = {};
            Box bb[kBins];
            for (int k = 0; k < kBins; k++) bb[k].empty();
            for (uint32_t k = t.begin; k < t.end; k++) {
                const Leaf& f = L[ids[k]];
                const int j = std::min((int)((f.c[a] - clo[a]) * s), kBins - 1);
                cnt[j]++;
                bb[j].grow(f.b);
            }
            float ra[kBins];
            uint32_t rc[kBins];
            Box acc; acc.empty();
            uint32_t c = 0;
            for (int k = kBins - 1; k >= 0; k--) {
                acc.grow(bb[k].b); c += cnt[k];
                ra[k] = c ? acc.area() : 0.f; rc[k] = c;
            }
            acc.empty(); c = 0;
            for (int k = 0; k < kBins - 1; k++) {
                acc.grow(bb[k].b); c += cnt[k];
                if (!c || !rc[k + 1]) continue;
                const float cost = acc.area() * (float)c + ra[k + 1] * (float)rc[k + 1];
                if (bax < 0 || cost < best) { best = cost; bax = a; bk = k; }
            }
        }
        uint32_t mid;
        if (bax < 0) mid = t.begin + (t.end - t.begin) / 2;
        else {
            // stable partition: bins <= bk to the left
            const float s = (float)kBins / (chi[bax] - clo[bax]);
            uint32_t l = t.begin, r = 0;
            for (uint32_t k = t.begin; k < t.end; k++) {
                const Leaf& f = L[ids[k]];
                const int j = std::min((int)((f.c[bax] - clo[bax]) * s), kBins - 1);
                if (j <= bk) ids[l++] = ids[k];
                else tmp[r++] = ids[k];
            }
            memcpy(&ids[l], tmp.data(), 4 * (size_t)r);
            mid = l;
        }
        // right task pushed first: the left subtree is numbered first (preorder)
        work.push_back({mid, t.end, ref, 1, t.depth + 1});
        work.push_back({t.begin, mid, ref, 0, t.depth + 1});
    }
    *out_nodes = count;
    *out_root = 0;
    *out_depth = depth;
    return 0;
}

} // namespace lsg

extern "C" int lsg_bvh_rebuild_sah(const float* nodes, uint32_t num_nodes, int32_t root, float* out_nodes,
                                   uint32_t* out_num_nodes, int32_t* out_root, uint32_t* out_depth)
{
    if (!nodes || !out_nodes || !out_num_nodes || !out_root || !out_depth) return LSG_ERR_INVALID;
    const int st = lsg::sah_rebuild(nodes, num_nodes, root, out_nodes, out_num_nodes, out_root, out_depth);
    if (st == 2) return LSG_ERR_INVALID;
    if (st == 1) {
        // nothing to rebuild: the input tree is the result
        memcpy(out_nodes, nodes, 64 * (size_t)num_nodes);
        *out_num_nodes = num_nodes; *out_root = root; *out_depth = 0;
    }
    return LSG_OK;
}
