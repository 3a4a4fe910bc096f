// lsgpu_trace.h -- closest-hit BVH2 traversal for one ray per lane (gfx950).
//
// Near-child-first stack traversal over 64-byte nodes whose child boxes are
// inflated by a DELTA_BSP-sized margin, so node culling is conservative and the
// result depends only on the per-triangle SRLNP test (IntersectBspFast.cpp:189-252)
// and the plane distance (:349-352): among accepted triangles with distance in
// [box-clipped rayLengthMin, current best], the smallest distance wins, ties to the
// lowest triangle index -- independent of traversal order.
//
// A Filter plays the role of RRCollisionHandler::collides (RRCollider.h:59): it
// sees each candidate (triangle, side, distance) and may reject it, and it may end
// the search early or narrow the interval once a hit is accepted.
#pragma once
#include "lsgpu_internal.h"

namespace lsg {

struct TraceResult { int tri; float t, u, v; bool front; uint32_t slot; };

struct NoFilter {
    __device__ bool accept(uint32_t, bool, float) const { return true; }
    // after accepting a hit at t: true = stop searching
    __device__ bool done(float) const { return false; }
    // after accepting a hit: upper bound of the interval still worth searching
    __device__ float clamp(float best) const { return best; }
};

// Child-box slab tests of one node.  noi = -o * idir per ray, so each plane distance
// is one fma: (b - o) * idir = b * idir - o * idir.  The fused form differs from the
// two-step one by about |o * idir| ulp, i.e. |o| * 6e-8 in distance, while every child
// box is inflated by DELTA_BSP = max|coordinate| * 1e-5; culling stays conservative and
// the hit, decided only by the triangle test, is unchanged.
__device__ inline void slab2(const Node& nd, const float3 noi, const float3 idir, float tlo, float thi,
                             float& n0, float& f0, float& n1, float& f1)
{
    float a, b, ta, tb, tc, td, te, tf;
    a = __fmaf_rn(nd.q0.x, idir.x, noi.x); b = __fmaf_rn(nd.q0.y, idir.x, noi.x); ta = fminf(a, b); tb = fmaxf(a, b);
    a = __fmaf_rn(nd.q0.z, idir.y, noi.y); b = __fmaf_rn(nd.q0.w, idir.y, noi.y); tc = fminf(a, b); td = fmaxf(a, b);
    a = __fmaf_rn(nd.q1.x, idir.z, noi.z); b = __fmaf_rn(nd.q1.y, idir.z, noi.z); te = fminf(a, b); tf = fmaxf(a, b);
    n0 = fmaxf(fmaxf(ta, tc), fmaxf(te, tlo)); f0 = fminf(fminf(tb, td), fminf(tf, thi));
    a = __fmaf_rn(nd.q1.z, idir.x, noi.x); b = __fmaf_rn(nd.q1.w, idir.x, noi.x); ta = fminf(a, b); tb = fmaxf(a, b);
    a = __fmaf_rn(nd.q2.x, idir.y, noi.y); b = __fmaf_rn(nd.q2.y, idir.y, noi.y); tc = fminf(a, b); td = fmaxf(a, b);
    a = __fmaf_rn(nd.q2.z, idir.z, noi.z); b = __fmaf_rn(nd.q2.w, idir.z, noi.z); te = fminf(a, b); tf = fmaxf(a, b);
    n1 = fmaxf(fmaxf(ta, tc), fmaxf(te, tlo)); f1 = fminf(fminf(tb, td), fminf(tf, thi));
}

__device__ inline float3 safe_inv(float3 d)
{
    float x = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
    float y = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
    float z = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
    return make_float3(1.0f / x, 1.0f / y, 1.0f / z);
}

// Closest accepted hit of one ray.  stk = this lane's first stack slot (LDS),
// successive slots `stride` ints apart.  Returns false on stack overflow (cannot
// happen when STACK > tree depth, which the host guarantees).
//
// Loop shape: "while-while" with speculative traversal (Aila & Laine, HPG 2009).
// Lanes descend inner nodes; a lane that reaches a leaf parks it and keeps
// descending, until every lane of the wave holds a parked leaf; then the wave tests
// triangles together.  Inner-node and triangle work thus run with full lane masks
// instead of interleaving per iteration.  The result does not depend on the visiting
// order (see the header), so it is the same as any other order's.
template <int STACK, bool STATS, class Filter>
__device__ inline bool trace_closest(const Filter& flt, const ColliderView& c, const float3 o, const float3 d,
                                     float rmin, float rmax, int32_t* stk, int stride, TraceResult& res,
                                     uint32_t& nvis, uint32_t& ntest)
{
    constexpr int32_t kDone = 0x7fffffff;          // never a node index
    res.tri = -1;
    float lo, hi;
    if (!c.nvalid || !ref_box_intersect(c.bmin, c.bmax, o, d, rmin, rmax, lo, hi)) return true;
    const float3 idir = safe_inv(d);
    const float3 noi = make_float3(-o.x * idir.x, -o.y * idir.y, -o.z * idir.z);
    float best = hi, bound = hi;
    int bi = -1; float bu = 0.f, bv = 0.f; bool bf = false; uint32_t bslot = 0;
    int32_t ref = c.root;          // next node: >= 0 inner, < 0 leaf, kDone finished
    int32_t leaf = 0;              // parked leaf (< 0) or none (0)
    int sp = 0;
    bool ok = true;
    auto pop = [&]() -> int32_t { return sp ? stk[(--sp) * stride] : kDone; };
    while (ref != kDone || leaf < 0) {
        // inner nodes, parking the first leaf met
        while (ref >= 0 && ref != kDone) {
            if (STATS) nvis++;
            const Node nd = c.nodes[ref];
            float n0, f0, n1, f1;
            slab2(nd, noi, idir, lo, bound, n0, f0, n1, f1);
            const int32_t r0 = __float_as_int(nd.q3.x), r1 = __float_as_int(nd.q3.y);
            const bool h0 = n0 <= f0, h1 = n1 <= f1;
            if (h0 && h1) {
                const bool swap = n1 < n0;
                if (sp < STACK) stk[(sp++) * stride] = swap ? r0 : r1; else ok = false;
                ref = swap ? r1 : r0;
            } else if (h0) ref = r0;
            else if (h1) ref = r1;
            else ref = pop();
            if (ref < 0 && leaf == 0) { leaf = ref; ref = pop(); }
            if (!__any(leaf == 0)) break;                // every active lane holds a leaf
        }
        // a leaf reached outside the loop above (the root itself, or a leaf next on
        // the stack while none was parked) is parked here, so every pass progresses
        if (leaf == 0 && ref < 0) { leaf = ref; ref = pop(); }
        // triangles of the parked leaf, then of leaves met directly on the stack
        while (leaf < 0) {
            const uint32_t st = leaf_start(leaf), cnt = leaf_count(leaf);
            bool stop = false;
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t p = st + k;
                if (STATS) ntest++;
                const float4 pl = c.plane[p];
                float ndd;
                const float t = plane_distance(pl, o, d, ndd);
                if (!(t >= lo && t <= bound)) continue;
                const uint2 mt = c.meta[p];
                if (!(bi < 0 || t < best || (int)mt.x < bi)) continue;
                float u, v;
                if (!srlnp_test(c.sv[p], c.ev[p], (int)mt.y, o, d, t, u, v)) continue;
                const bool fr = ndd < 0;
                if (!flt.accept(mt.x, fr, t)) continue;
                best = t; bi = (int)mt.x; bu = u; bv = v; bf = fr; bslot = p;
                bound = flt.clamp(best);
                if (flt.done(t)) { stop = true; break; }
            }
            if (stop) { ref = kDone; leaf = 0; break; }
            leaf = 0;
            if (ref < 0) { leaf = ref; ref = pop(); }
        }
    }
    if (bi >= 0) { res.tri = bi; res.t = best; res.u = bu; res.v = bv; res.front = bf; res.slot = bslot; }
    return ok;
}

// ---- resumable form ---------------------------------------------------------------------
// The same traversal with its state kept by the caller, so a lane can leave the loop
// before its ray is finished and continue later: trace_begin sets a ray up,
// trace_resume runs it until it is finished (returns true) or until no more than
// `yield_at` lanes of the wave are still tracing (returns false, state kept).  Used by
// the closest-hit kernel to refill lanes whose rays have finished.
struct TraceState {
    float3 o, d, idir, noi;
    float lo, best, bound, bu, bv;
    int32_t ref, leaf, bi;
    int sp;
    bool bf;
};

// false when the ray misses the scene box (then it is finished, no hit)
__device__ inline bool trace_begin(TraceState& S, const ColliderView& c, const float3 o, const float3 d, float rmin, float rmax)
{
    float lo, hi;
    S.bi = -1;
    if (!c.nvalid || !ref_box_intersect(c.bmin, c.bmax, o, d, rmin, rmax, lo, hi)) return false;
    S.o = o; S.d = d;
    S.idir = safe_inv(d);
    S.noi = make_float3(-o.x * S.idir.x, -o.y * S.idir.y, -o.z * S.idir.z);
    S.lo = lo; S.best = hi; S.bound = hi;
    S.bu = 0.f; S.bv = 0.f; S.bf = false;
    S.ref = c.root; S.leaf = 0; S.sp = 0;
    return true;
}

template <int STACK, class Filter>
__device__ inline bool trace_resume(TraceState& S, const Filter& flt, const ColliderView& c, int32_t* stk, int stride,
                                    uint32_t yield_at, bool& ok)
{
    constexpr int32_t kDone = 0x7fffffff;
    auto pop = [&]() -> int32_t { return S.sp ? stk[(--S.sp) * stride] : kDone; };
    while (S.ref != kDone || S.leaf < 0) {
        if ((uint32_t)__popcll(__ballot(1)) <= yield_at) return false;     // enough lanes are free
        while (S.ref >= 0 && S.ref != kDone) {
            const Node nd = c.nodes[S.ref];
            float n0, f0, n1, f1;
            slab2(nd, S.noi, S.idir, S.lo, S.bound, n0, f0, n1, f1);
            const int32_t r0 = __float_as_int(nd.q3.x), r1 = __float_as_int(nd.q3.y);
            const bool h0 = n0 <= f0, h1 = n1 <= f1;
            if (h0 && h1) {
                const bool swap = n1 < n0;
                if (S.sp < STACK) stk[(S.sp++) * stride] = swap ? r0 : r1; else ok = false;
                S.ref = swap ? r1 : r0;
            } else if (h0) S.ref = r0;
            else if (h1) S.ref = r1;
            else S.ref = pop();
            if (S.ref < 0 && S.leaf == 0) { S.leaf = S.ref; S.ref = pop(); }
            if (!__any(S.leaf == 0)) break;
        }
        if (S.leaf == 0 && S.ref < 0) { S.leaf = S.ref; S.ref = pop(); }
        while (S.leaf < 0) {
            const uint32_t st = leaf_start(S.leaf), cnt = leaf_count(S.leaf);
            bool stop = false;
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t p = st + k;
                const float4 pl = c.plane[p];
                float ndd;
                const float t = plane_distance(pl, S.o, S.d, ndd);
                if (!(t >= S.lo && t <= S.bound)) continue;
                const uint2 mt = c.meta[p];
                if (!(S.bi < 0 || t < S.best || (int)mt.x < S.bi)) continue;
                float u, v;
                if (!srlnp_test(c.sv[p], c.ev[p], (int)mt.y, S.o, S.d, t, u, v)) continue;
                const bool fr = ndd < 0;
                if (!flt.accept(mt.x, fr, t)) continue;
                S.best = t; S.bi = (int)mt.x; S.bu = u; S.bv = v; S.bf = fr;
                S.bound = flt.clamp(S.best);
                if (flt.done(t)) { stop = true; break; }
            }
            if (stop) { S.ref = kDone; S.leaf = 0; break; }
            S.leaf = 0;
            if (S.ref < 0) { S.leaf = S.ref; S.ref = pop(); }
        }
    }
    return true;
}

} // namespace lsg
