// lsg_host_tree.h -- host mirror of a liblsgpu collider for single-ray queries.
//
// RRCollider::intersect (include/Lightsprint/RRCollider.h:197-228) must stay
// callable from any host thread with an arbitrary RRCollisionHandler: the solver's
// radiosity shots and user code issue single rays (rrcore.cpp:874).  This header
// keeps a host copy of the GPU-built tree and triangle records and traverses it
// with the same arithmetic as the GPU kernel (Box::intersect pretest,
// plane distance, SRLNP test), so single-ray and batched results are identical.
// It depends only on lsgpu.h, not on the reference's headers.
#pragma once
#include "lsgpu.h"
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace lsg_host {

struct Hit { int32_t tri; float t, u, v; bool front; uint32_t slot; };

class Tree {
public:
    bool load(const lsg_collider* c)
    {
        lsg_collider_info info;
        if (lsg_collider_info_get(c, &info)) return false;
        root = info.root_ref; nvalid = info.num_valid;
        memcpy(bmin, info.box_min, sizeof bmin); memcpy(bmax, info.box_max, sizeof bmax);
        nodes.resize(16 * (size_t)info.num_nodes);
        plane.resize(4 * (size_t)nvalid); sv.resize(4 * (size_t)nvalid); ev.resize(4 * (size_t)nvalid); meta.resize(2 * (size_t)nvalid);
        if (lsg_collider_download(c, nodes.data(), nullptr)) return false;
        if (lsg_collider_download_triangles(c, plane.data(), sv.data(), ev.data(), meta.data())) return false;
        return true;
    }
    size_t bytes() const { return 4 * (nodes.size() + plane.size() + sv.size() + ev.size() + meta.size()); }
    const float* tri_plane(uint32_t slot) const { return &plane[4 * (size_t)slot]; }

    // Box::intersect, RRCollider/geometry.cpp:38-86 (scalar path)
    static bool box_intersect(const float* mn, const float* mx, const float* o, const float* d, float rmin, float rmax, float& lo, float& hi)
    {
        auto mi = [](float a, float b) { return a < b ? a : b; };
        auto ma = [](float a, float b) { return a > b ? a : b; };
        float l1 = (mn[0] - o[0]) / d[0], l2 = (mx[0] - o[0]) / d[0];
        float lmin = mi(l1, l2), lmax = ma(l1, l2);
        l1 = (mn[1] - o[1]) / d[1]; l2 = (mx[1] - o[1]) / d[1];
        lmin = ma(mi(l1, l2), lmin); lmax = mi(ma(l1, l2), lmax);
        l1 = (mn[2] - o[2]) / d[2]; l2 = (mx[2] - o[2]) / d[2];
        lmin = ma(mi(l1, l2), lmin); lmax = mi(ma(l1, l2), lmax);
        lmin = ma(rmin, lmin); lmax = mi(rmax, lmax);
        lo = lmin; hi = lmax;
        return (lmax >= 0.f) & (lmax >= lmin);
    }

    // plane distance (IntersectBspFast.cpp:349-352) + SRLNP test (:189-252); t must lie in [lo,hi]
    bool test(uint32_t p, const float* o, const float* d, float lo, float hi, Hit& h) const
    {
        const float* n = &plane[4 * (size_t)p];
        float ndd = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
        float ndo = o[0] * n[0] + o[1] * n[1] + o[2] * n[2] + n[3];
        float t = -ndo / ndd;
        if (!(t >= lo && t <= hi)) return false;
        float pt[3] = { o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t };
        int byte = (int)meta[2 * (size_t)p + 1];
        int m = byte % 3, q = byte / 3;
        int x = m, y = (m + 1) % 3, z = q == 0 ? 2 : (q == 1 ? 0 : 1);
        const float* s = &sv[4 * (size_t)p];
        const float* e = &ev[4 * (size_t)p];
        float v = ((pt[y] - s[1]) * e[0] - (pt[x] - s[0]) * e[1]) * s[3];
        if (v < 0 || v > 1) return false;
        float u = (pt[z] - s[2] - e[2] * v) * e[3];
        if (u < 0 || u + v > 1) return false;
        if (byte > 8) return false;
        h.tri = (int32_t)meta[2 * (size_t)p]; h.t = t; h.u = u; h.v = v; h.front = ndd < 0; h.slot = p;
        return true;
    }

    // Visits every leaf triangle whose node boxes the ray's [lo,bound()] interval enters.
    template <class F>
    void traverse(const float* o, const float* d, float lo, F&& on_tri, const float& bound) const
    {
        if (!nvalid) return;
        float idir[3];
        for (int j = 0; j < 3; j++) {
            float dj = std::fabs(d[j]) < 1e-30f ? std::copysign(1e-30f, d[j]) : d[j];
            idir[j] = 1.0f / dj;
        }
        int32_t stack[128]; int sp = 0;
        int32_t ref = root;
        for (;;) {
            if (ref >= 0) {
                const float* nd = &nodes[16 * (size_t)ref];
                float tn[2], tf[2];
                for (int c = 0; c < 2; c++) {
                    const float* b = nd + 6 * c;
                    float a0 = (b[0] - o[0]) * idir[0], b0 = (b[1] - o[0]) * idir[0];
                    float a1 = (b[2] - o[1]) * idir[1], b1 = (b[3] - o[1]) * idir[1];
                    float a2 = (b[4] - o[2]) * idir[2], b2 = (b[5] - o[2]) * idir[2];
                    tn[c] = std::fmax(std::fmax(std::fmin(a0, b0), std::fmin(a1, b1)), std::fmax(std::fmin(a2, b2), lo));
                    tf[c] = std::fmin(std::fmin(std::fmax(a0, b0), std::fmax(a1, b1)), std::fmin(std::fmax(a2, b2), bound));
                }
                int32_t r0, r1; memcpy(&r0, nd + 12, 4); memcpy(&r1, nd + 13, 4);
                bool h0 = tn[0] <= tf[0], h1 = tn[1] <= tf[1];
                if (h0 && h1) {
                    bool sw = tn[1] < tn[0];
                    if (sp < 128) stack[sp++] = sw ? r0 : r1;
                    ref = sw ? r1 : r0;
                    continue;
                }
                if (h0) { ref = r0; continue; }
                if (h1) { ref = r1; continue; }
            } else {
                uint32_t st = ((uint32_t)ref >> 4) & 0x7ffffffu, cnt = ((uint32_t)ref & 15u) + 1;
                for (uint32_t k = 0; k < cnt; k++) on_tri(st + k);
            }
            if (!sp) break;
            ref = stack[--sp];
        }
    }

    // closest hit, ties to the lowest triangle index (same rule as the GPU kernel)
    bool closest(const float* o, const float* d, float rmin, float rmax, Hit& best) const
    {
        float lo, hi;
        best.tri = -1;
        if (!nvalid || !box_intersect(bmin, bmax, o, d, rmin, rmax, lo, hi)) return false;
        float bound = hi;
        traverse(o, d, lo, [&](uint32_t p) {
            Hit h;
            if (test(p, o, d, lo, bound, h) && (best.tri < 0 || h.t < bound || h.tri < best.tri)) { best = h; bound = h.t; }
        }, bound);
        return best.tri >= 0;
    }

    // every hit in the box-clipped interval, nearest first (ties by triangle index):
    // the order in which the reference's IT_BSP techniques offer hits to a
    // collision handler (RRCollider.h:57)
    bool all_hits(const float* o, const float* d, float rmin, float rmax, std::vector<Hit>& out) const
    {
        out.clear();
        float lo, hi;
        if (!nvalid || !box_intersect(bmin, bmax, o, d, rmin, rmax, lo, hi)) return false;
        const float bound = hi;
        traverse(o, d, lo, [&](uint32_t p) { Hit h; if (test(p, o, d, lo, bound, h)) out.push_back(h); }, bound);
        std::sort(out.begin(), out.end(), [](const Hit& a, const Hit& b) { return a.t < b.t || (a.t == b.t && a.tri < b.tri); });
        return true;
    }

    int32_t root = 0;
    uint32_t nvalid = 0;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    std::vector<float> nodes, plane, sv, ev;
    std::vector<uint32_t> meta;
};

} // namespace lsg_host
