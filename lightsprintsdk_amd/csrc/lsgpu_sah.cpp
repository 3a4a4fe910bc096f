// lsgpu_sah.cpp -- host-side SAH rebuild of the LBVH's inner nodes.
//
// The GPU LBVH (Morton order, Karras split at the highest differing bit) makes leaves of
// at most leaf_size consecutive triangles in Morton order; its inner nodes follow the
// Morton bits, not the geometry, so rays visit more nodes than a surface-area-heuristic
// tree would.  This pass keeps the leaves (their triangle ranges and their inflated boxes)
// and rebuilds the inner nodes over them top-down with a binned SAH (32 bins per axis on
// leaf-box centroids, as many as leaves below 32, cost = area(left)·n(left) + area(right)·n(right), the cheapest split
// over the three axes; equal centroids split the range in half).  The tree replaces the
// reference's BSP/kd build (createAndSaveBsp, RRCollider/bsp.cpp:1192), whose kd split
// search weighs box surface areas as well (find_best_root_kd, bsp.cpp:388-574).
//
// The leaves enter in the order of their triangles (the LBVH's Morton order); the build
// only reorders them by stable partitions, so the result depends on the leaves alone.
//
// Node layout as lsg_collider_download: 16 floats per inner node, child 0 box
// (lo.x hi.x lo.y hi.y lo.z hi.z), child 1 box, child refs as int bits, two spare.
// Output nodes are numbered in preorder (root = 0, left subtree before right), so the
// result depends only on the input tree.
#include "lsgpu.h"
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
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

// a range of leaves [begin, end) waiting for its node: its parent slot and side, depth,
// the union of its leaf boxes and the bounds of its leaf centroids
struct Task {
    uint32_t begin, end;
    int32_t parent;
    int side;
    uint32_t depth;
    Box box;
    float clo[3], chi[3];
};

struct Builder {
    Leaf* L;
    Leaf* tmp;
    float* out;

    static void bounds(const Leaf* f, uint32_t n, Task& t)
    {
        t.box.empty();
        for (int a = 0; a < 3; a++) { t.clo[a] = 3.0e38f; t.chi[a] = -3.0e38f; }
        for (uint32_t k = 0; k < n; k++) {
            t.box.grow(f[k].b);
            for (int a = 0; a < 3; a++) { t.clo[a] = std::min(t.clo[a], f[k].c[a]); t.chi[a] = std::max(t.chi[a], f[k].c[a]); }
        }
    }

    // Builds the subtree of t; its inner nodes take the numbers next, next+1, ... in
    // preorder.  Tasks of fewer than `defer` leaves (defer > 0) are not built here: each
    // gets the block of numbers its subtree will use (m leaves make m - 1 inner nodes),
    // and is appended to `later`, to be built with the same numbers by another call.
    void run(const Task& root, uint32_t next, uint32_t defer, std::vector<std::pair<Task, uint32_t>>* later, uint32_t& depth)
    {
        std::vector<Task> work{root};
        while (!work.empty()) {
            const Task t = work.back(); work.pop_back();
            const uint32_t m = t.end - t.begin;
            int32_t ref;
            if (m == 1) ref = L[t.begin].ref;
            else {
                ref = (int32_t)next;
                next += (defer && m < defer && t.parent >= 0) ? m - 1 : 1;
            }
            if (t.parent >= 0) {
                float* p = out + 16 * (size_t)t.parent;
                memcpy(p + 6 * t.side, t.box.b, 24);
                memcpy(p + 12 + t.side, &ref, 4);
            }
            if (m == 1) continue;
            if (defer && m < defer && t.parent >= 0) {
                Task c = t;
                c.parent = -1;
                later->push_back({c, (uint32_t)ref});
                continue;
            }
            depth = std::max(depth, t.depth);
            float* q = out + 16 * (size_t)ref;
            q[14] = 0.f; q[15] = 0.f;
            // bins on the three axes in one pass (as many bins as leaves below 32)
            const int nb = (int)std::min<uint32_t>(kBins, std::max<uint32_t>(4u, m));
            float sc[3];
            bool ok[3];
            uint32_t cnt[3][kBins] = {};
            Box bb[3][kBins];
            for (int a = 0; a < 3; a++) {
                ok[a] = t.chi[a] > t.clo[a];
                sc[a] = ok[a] ? (float)nb / (t.chi[a] - t.clo[a]) : 0.f;
                for (int k = 0; k < nb; k++) bb[a][k].empty();
            }
            for (uint32_t k = t.begin; k < t.end; k++) {
                const Leaf& f = L[k];
                for (int a = 0; a < 3; a++) {
                    if (!ok[a]) continue;
                    const int j = std::min((int)((f.c[a] - t.clo[a]) * sc[a]), nb - 1);
                    cnt[a][j]++;
                    bb[a][j].grow(f.b);
                }
            }
            float best = 0.f;
            int bax = -1, bk = -1;
            for (int a = 0; a < 3; a++) {
                if (!ok[a]) continue;
                float ra[kBins];
                uint32_t rc[kBins];
                Box acc; acc.empty();
                uint32_t c = 0;
                for (int k = nb - 1; k >= 0; k--) {
                    acc.grow(bb[a][k].b); c += cnt[a][k];
                    ra[k] = c ? acc.area() : 0.f; rc[k] = c;
                }
                acc.empty(); c = 0;
                for (int k = 0; k < nb - 1; k++) {
                    acc.grow(bb[a][k].b); c += cnt[a][k];
                    if (!c || !rc[k + 1]) continue;
                    const float cost = acc.area() * (float)c + ra[k + 1] * (float)rc[k + 1];
                    if (bax < 0 || cost < best) { best = cost; bax = a; bk = k; }
                }
            }
            Task lt, rt;
            uint32_t mid;
            if (bax < 0) {
                // all centroids equal: halves of the range
                mid = t.begin + m / 2;
                bounds(L + t.begin, mid - t.begin, lt);
                bounds(L + mid, t.end - mid, rt);
            } else {
                // stable partition, bins <= bk to the left, child bounds on the way
                lt.box.empty(); rt.box.empty();
                for (int a = 0; a < 3; a++) { lt.clo[a] = rt.clo[a] = 3.0e38f; lt.chi[a] = rt.chi[a] = -3.0e38f; }
                uint32_t l = t.begin, r = 0;
                for (uint32_t k = t.begin; k < t.end; k++) {
                    const Leaf f = L[k];
                    const int j = std::min((int)((f.c[bax] - t.clo[bax]) * sc[bax]), nb - 1);
                    Task& d = j <= bk ? lt : rt;
                    d.box.grow(f.b);
                    for (int a = 0; a < 3; a++) { d.clo[a] = std::min(d.clo[a], f.c[a]); d.chi[a] = std::max(d.chi[a], f.c[a]); }
                    if (j <= bk) L[l++] = f;
                    else tmp[t.begin + r++] = f;
                }
                memcpy(L + l, tmp + t.begin, sizeof(Leaf) * (size_t)r);
                mid = l;
            }
            lt.begin = t.begin; lt.end = mid; lt.parent = ref; lt.side = 0; lt.depth = t.depth + 1;
            rt.begin = mid; rt.end = t.end; rt.parent = ref; rt.side = 1; rt.depth = t.depth + 1;
            // right task pushed first: the left subtree is numbered first (preorder)
            work.push_back(rt);
            work.push_back(lt);
        }
    }
};

} // namespace

// Rebuilds the inner nodes of the BVH2 (nodes, root) over its leaves.  out must hold
// (leaves - 1) * 16 floats (at most nnodes * 16).  Returns 0 and the new root, node count
// and depth (inner nodes on the longest root path, root = 1); returns 1 when the tree has
// fewer than two leaves (nothing to rebuild), 2 on a malformed tree.  Large trees build
// their lower subtrees on several threads; every subtree keeps the numbers a sequential
// build gives it, so the result does not depend on the thread count.
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
    // leaves in the order of their triangles (the LBVH's Morton order): a leaf's start
    // index in the leaf-ordered triangles is unique
    std::sort(L.begin(), L.end(), [](const Leaf& a, const Leaf& b) {
        return (((uint32_t)a.ref >> 4) & 0x7ffffffu) < (((uint32_t)b.ref >> 4) & 0x7ffffffu);
    });
    std::vector<Leaf> tmp(n);
    Task t0;
    t0.begin = 0; t0.end = n; t0.parent = -1; t0.side = 0; t0.depth = 1;
    Builder::bounds(L.data(), n, t0);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint32_t defer = (hw > 1 && n >= 16384) ? std::max<uint32_t>(2048, n / 512) : 0;
    std::vector<std::pair<Task, uint32_t>> later;
    uint32_t depth = 0;
    Builder B{L.data(), tmp.data(), out};
    B.run(t0, 0, defer, &later, depth);
    if (!later.empty()) {
        std::atomic<size_t> next_job{0};
        std::vector<uint32_t> depths(hw, 0);
        std::vector<std::thread> pool;
        for (unsigned w = 0; w < hw; w++)
            pool.emplace_back([&, w]() {
                for (size_t j; (j = next_job.fetch_add(1)) < later.size();) {
                    const Task& t = later[j].first;
                    Builder b{L.data(), tmp.data(), out};
                    b.run(t, later[j].second, 0, nullptr, depths[w]);
                }
            });
        for (auto& th : pool) th.join();
        for (uint32_t d : depths) depth = std::max(depth, d);
    }
    *out_nodes = n - 1;
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
