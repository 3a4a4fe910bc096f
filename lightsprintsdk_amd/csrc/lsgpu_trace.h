// lsgpu_trace.h -- closest-hit BVH4 traversal for one ray per lane (gfx950).
//
// Near-child-first stack traversal over 128-byte BVH4 nodes whose child boxes are
// inflated by a DELTA_BSP-sized margin, so node culling is conservative and the
// result depends only on the per-triangle SRLNP test (IntersectBspFast.cpp:189-252)
// and the plane distance (:349-352): among accepted triangles with distance in
// [box-clipped rayLengthMin, current best], the smallest distance wins, ties to the
// lowest triangle index -- independent of traversal order.
//
// A Filter plays the role of RRCollisionHandler::collides (RRCollider.h:59): it
// sees each candidate (triangle, side, distance, hitPoint2d) and may reject it, and it
// may end the search early or narrow the interval once a hit is accepted.
#pragma once
#include "lsgpu_internal.h"

namespace lsg {

struct TraceResult { int tri; float t, u, v; bool front; uint32_t slot; };

struct NoFilter {
    __device__ bool accept(uint32_t, bool, float, float, float) const { return true; }
    // after accepting a hit at t: true = stop searching
    __device__ bool done(float) const { return false; }
    // after accepting a hit: upper bound of the interval still worth searching
    __device__ float clamp(float best) const { return best; }
};

// One BVH4 node: the four child-box slab tests, then the children that are hit, nearest
// first.  noi = -o * idir per ray, so each plane distance is one fma: (b - o) * idir =
// b * idir - o * idir.  The fused form differs from the two-step one by about |o * idir|
// ulp, i.e. |o| * 6e-8 in distance, while every child box is inflated by DELTA_BSP =
// max|coordinate| * 1e-5; culling stays conservative and the hit, decided only by the
// triangle test, is unchanged.  The nearest child is returned as the next ref, the
// others go on the stack farthest first.  No child hit: the next ref is popped.  The
// tree is collapsed so that the stack never holds more than STACK-1 entries.
__device__ inline void cswap(float& ka, int32_t& ra, float& kb, int32_t& rb)
{
    const bool sw = kb < ka;
    const float k = sw ? kb : ka; kb = sw ? ka : kb; ka = k;
    const int32_t r = sw ? rb : ra; rb = sw ? ra : rb; ra = r;
}

template <int STACK, class Pop, bool ORDERED = true>
__device__ inline int32_t visit4(const Node4* __restrict__ nodes, int32_t ref, const float3 noi, const float3 idir, float tlo,
                                 float thi, int32_t* stk, int stride, int& sp, bool& ok, Pop pop)
{
    const Node4* np = nodes + ref;
    float4 tn, tf;
    // the node's seven 16-byte loads are issued together, before any use: left to itself
    // the scheduler issues four, waits for them, then the other three (two memory
    // round trips per node visit)
    const float4 n0 = np->q[0], n1 = np->q[1], n2 = np->q[2], n3 = np->q[3], n4 = np->q[4], n5 = np->q[5], n6 = np->q[6];
    __builtin_amdgcn_sched_barrier(0);
    {
        const float4 l = n0, h = n1;
        float a, b;
        a = __fmaf_rn(l.x, idir.x, noi.x); b = __fmaf_rn(h.x, idir.x, noi.x); tn.x = fminf(a, b); tf.x = fmaxf(a, b);
        a = __fmaf_rn(l.y, idir.x, noi.x); b = __fmaf_rn(h.y, idir.x, noi.x); tn.y = fminf(a, b); tf.y = fmaxf(a, b);
        a = __fmaf_rn(l.z, idir.x, noi.x); b = __fmaf_rn(h.z, idir.x, noi.x); tn.z = fminf(a, b); tf.z = fmaxf(a, b);
        a = __fmaf_rn(l.w, idir.x, noi.x); b = __fmaf_rn(h.w, idir.x, noi.x); tn.w = fminf(a, b); tf.w = fmaxf(a, b);
    }
    {
        const float4 l = n2, h = n3;
        float a, b;
        a = __fmaf_rn(l.x, idir.y, noi.y); b = __fmaf_rn(h.x, idir.y, noi.y); tn.x = fmaxf(tn.x, fminf(a, b)); tf.x = fminf(tf.x, fmaxf(a, b));
        a = __fmaf_rn(l.y, idir.y, noi.y); b = __fmaf_rn(h.y, idir.y, noi.y); tn.y = fmaxf(tn.y, fminf(a, b)); tf.y = fminf(tf.y, fmaxf(a, b));
        a = __fmaf_rn(l.z, idir.y, noi.y); b = __fmaf_rn(h.z, idir.y, noi.y); tn.z = fmaxf(tn.z, fminf(a, b)); tf.z = fminf(tf.z, fmaxf(a, b));
        a = __fmaf_rn(l.w, idir.y, noi.y); b = __fmaf_rn(h.w, idir.y, noi.y); tn.w = fmaxf(tn.w, fminf(a, b)); tf.w = fminf(tf.w, fmaxf(a, b));
    }
    {
        const float4 l = n4, h = n5;
        float a, b;
        a = __fmaf_rn(l.x, idir.z, noi.z); b = __fmaf_rn(h.x, idir.z, noi.z); tn.x = fmaxf(tn.x, fminf(a, b)); tf.x = fminf(tf.x, fmaxf(a, b));
        a = __fmaf_rn(l.y, idir.z, noi.z); b = __fmaf_rn(h.y, idir.z, noi.z); tn.y = fmaxf(tn.y, fminf(a, b)); tf.y = fminf(tf.y, fmaxf(a, b));
        a = __fmaf_rn(l.z, idir.z, noi.z); b = __fmaf_rn(h.z, idir.z, noi.z); tn.z = fmaxf(tn.z, fminf(a, b)); tf.z = fminf(tf.z, fmaxf(a, b));
        a = __fmaf_rn(l.w, idir.z, noi.z); b = __fmaf_rn(h.w, idir.z, noi.z); tn.w = fmaxf(tn.w, fminf(a, b)); tf.w = fminf(tf.w, fmaxf(a, b));
    }
    const float4 rq = n6;
    int32_t r0 = __float_as_int(rq.x), r1 = __float_as_int(rq.y), r2 = __float_as_int(rq.z), r3 = __float_as_int(rq.w);
    constexpr float kInf = __builtin_huge_valf();
    float k0 = fmaxf(tn.x, tlo), k1 = fmaxf(tn.y, tlo), k2 = fmaxf(tn.z, tlo), k3 = fmaxf(tn.w, tlo);
    k0 = (r0 != kNoChild && k0 <= fminf(tf.x, thi)) ? k0 : kInf;
    k1 = (r1 != kNoChild && k1 <= fminf(tf.y, thi)) ? k1 : kInf;
    k2 = (r2 != kNoChild && k2 <= fminf(tf.z, thi)) ? k2 : kInf;
    k3 = (r3 != kNoChild && k3 <= fminf(tf.w, thi)) ? k3 : kInf;
    auto push = [&](int32_t r) { if (sp < STACK) stk[(sp++) * stride] = r; else ok = false; };
    if constexpr (!ORDERED) {
        // any order: continue with the last hit child in slot order, push the others
        int32_t nxt = 0;
        bool have = false;
        auto take = [&](float k, int32_t r) { if (k < kInf) { if (have) push(nxt); nxt = r; have = true; } };
        take(k0, r0); take(k1, r1); take(k2, r2); take(k3, r3);
        return have ? nxt : pop();
    }
    cswap(k0, r0, k1, r1); cswap(k2, r2, k3, r3); cswap(k0, r0, k2, r2); cswap(k1, r1, k3, r3); cswap(k1, r1, k2, r2);
    if (k3 < kInf) push(r3);
    if (k2 < kInf) push(r2);
    if (k1 < kInf) push(r1);
    return k0 < kInf ? r0 : pop();
}

// visit4 with each child's entry and exit distances from its near and far planes only
// (same values, about 20 fewer VALU operations per visit, a dozen more registers)
template <int STACK, class Pop, bool ORDERED = true>
__device__ inline int32_t visit4_nf(const Node4* __restrict__ nodes, int32_t ref, const float3 noi, const float3 idir, float tlo,
                                 float thi, int32_t* stk, int stride, int& sp, bool& ok, Pop pop)
{
    const Node4* np = nodes + ref;
    // near and far planes by the signs of the ray direction: lo planes sit at byte
    // offsets 0/32/64 of the node and hi planes 16 bytes after them, so per axis the lane
    // loads its near plane first and the far one at the other offset.  near <= far per
    // axis (fma is monotonic and idir never 0), so these are the min/max of the two plane
    // distances without computing both; the hoisted offsets cost two adds per visit.
    const uint32_t ox = idir.x < 0.f ? 16u : 0u, oy = idir.y < 0.f ? 48u : 32u, oz = idir.z < 0.f ? 80u : 64u;
    const char* nb = (const char*)np;
    // the node's seven 16-byte loads are issued together, before any use: left to itself
    // the scheduler issues four, waits for them, then the other three (two memory
    // round trips per node visit)
    const float4 nx = *(const float4*)(nb + ox), fx = *(const float4*)(nb + (ox ^ 16u));
    const float4 ny = *(const float4*)(nb + oy), fy = *(const float4*)(nb + (oy ^ 16u));
    const float4 nz = *(const float4*)(nb + oz), fz = *(const float4*)(nb + (oz ^ 16u));
    const float4 n6 = np->q[6];
    __builtin_amdgcn_sched_barrier(0);
    auto enter = [&](float x, float y, float z) {
        return fmaxf(fmaxf(__fmaf_rn(x, idir.x, noi.x), __fmaf_rn(y, idir.y, noi.y)), fmaxf(__fmaf_rn(z, idir.z, noi.z), tlo));
    };
    auto leave = [&](float x, float y, float z) {
        return fminf(fminf(__fmaf_rn(x, idir.x, noi.x), __fmaf_rn(y, idir.y, noi.y)), fminf(__fmaf_rn(z, idir.z, noi.z), thi));
    };
    const float4 rq = n6;
    int32_t r0 = __float_as_int(rq.x), r1 = __float_as_int(rq.y), r2 = __float_as_int(rq.z), r3 = __float_as_int(rq.w);
    constexpr float kInf = __builtin_huge_valf();
    float k0 = enter(nx.x, ny.x, nz.x), k1 = enter(nx.y, ny.y, nz.y), k2 = enter(nx.z, ny.z, nz.z), k3 = enter(nx.w, ny.w, nz.w);
    k0 = (r0 != kNoChild && k0 <= leave(fx.x, fy.x, fz.x)) ? k0 : kInf;
    k1 = (r1 != kNoChild && k1 <= leave(fx.y, fy.y, fz.y)) ? k1 : kInf;
    k2 = (r2 != kNoChild && k2 <= leave(fx.z, fy.z, fz.z)) ? k2 : kInf;
    k3 = (r3 != kNoChild && k3 <= leave(fx.w, fy.w, fz.w)) ? k3 : kInf;
    auto push = [&](int32_t r) { if (sp < STACK) stk[(sp++) * stride] = r; else ok = false; };
    if constexpr (!ORDERED) {
        // any order: continue with the last hit child in slot order, push the others
        int32_t nxt = 0;
        bool have = false;
        auto take = [&](float k, int32_t r) { if (k < kInf) { if (have) push(nxt); nxt = r; have = true; } };
        take(k0, r0); take(k1, r1); take(k2, r2); take(k3, r3);
        return have ? nxt : pop();
    }
    cswap(k0, r0, k1, r1); cswap(k2, r2, k3, r3); cswap(k0, r0, k2, r2); cswap(k1, r1, k3, r3); cswap(k1, r1, k2, r2);
    if (k3 < kInf) push(r3);
    if (k2 < kInf) push(r2);
    if (k1 < kInf) push(r1);
    return k0 < kInf ? r0 : pop();
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
// ORDERED = false (shadow rays): children are visited in any order.  A shadow ray only
// asks whether a blocker lies below rugDistance and, if not, whether any blocker lies in
// range; after the first hit its range shrinks below rugDistance and the search there is
// exhaustive, so the answer does not depend on the order, and the sort is saved.
// NF: visit4_nf (fewer operations) or visit4 (fewer registers, for callers near the
// register cap).
template <int STACK, bool STATS, class Filter, bool ORDERED = true, bool NF = true>
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
    int32_t ref = c.root4;         // next node: >= 0 inner, < 0 leaf, kDone finished
    int32_t leaf = 0;              // parked leaf (< 0) or none (0)
    int sp = 0;
    bool ok = true;
    auto pop = [&]() -> int32_t { return sp ? stk[(--sp) * stride] : kDone; };
    while (ref != kDone || leaf < 0) {
        // inner nodes, parking the first leaf met
        while (ref >= 0 && ref != kDone) {
            if (STATS) nvis++;
            if constexpr (NF) ref = visit4_nf<STACK, decltype(pop), ORDERED>(c.nodes4, ref, noi, idir, lo, bound, stk, stride, sp, ok, pop);
            else ref = visit4<STACK, decltype(pop), ORDERED>(c.nodes4, ref, noi, idir, lo, bound, stk, stride, sp, ok, pop);
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
                const TriRest* tr = c.rest + p;
                const uint2 mt = tr->meta;
                if (!(bi < 0 || t < best || (int)mt.x < bi)) continue;
                float u, v;
                if (!srlnp_test(tr->sv, tr->ev, (int)mt.y, o, d, t, u, v)) continue;
                const bool fr = ndd < 0;
                if (!flt.accept(mt.x, fr, t, u, v)) continue;
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
    S.ref = c.root4; S.leaf = 0; S.sp = 0;
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
            S.ref = visit4<STACK>(c.nodes4, S.ref, S.noi, S.idir, S.lo, S.bound, stk, stride, S.sp, ok, pop);
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
                const TriRest* tr = c.rest + p;
                const uint2 mt = tr->meta;
                if (!(S.bi < 0 || t < S.best || (int)mt.x < S.bi)) continue;
                float u, v;
                if (!srlnp_test(tr->sv, tr->ev, (int)mt.y, S.o, S.d, t, u, v)) continue;
                const bool fr = ndd < 0;
                if (!flt.accept(mt.x, fr, t, u, v)) continue;
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
