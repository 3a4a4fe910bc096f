// lsgpu_collider.hip -- GPU LBVH build and closest-hit traversal for gfx950.
//
// Replaces the reference's CPU BSP/kd collider (IntersectBspFast, bsp.cpp) behind
// RRCollider::create / intersect.  Build: mesh box -> per-triangle SRLNP precompute
// + Morton key -> LSD radix sort (LDS histograms, wave-ballot stable ranking) ->
// Karras radix tree -> bottom-up refit with arrival counters -> 64-byte BVH2 nodes
// with up to leaf_size triangles per leaf, triangles stored in leaf order.
// Traversal: one ray per lane, near-child-first, per-lane stack in LDS, the
// reference's SRLNP triangle test, closest hit with ties to the lowest index.
#include "lsgpu_internal.h"
#include "lsgpu_trace.h"
#include <vector>
#include <chrono>
#include <cstring>
#include <cmath>

namespace lsg {

static inline void dfree(void* p) { if (p) (void)hipFree(p); }

// ---------------------------------------------------------------------------------------
// build kernels

// RRMesh::getAABB (RRMesh/RRMesh.cpp:164-194): min/max over finite vertex coordinates.
__global__ __launch_bounds__(256) void k_vertex_aabb(const float* __restrict__ v, uint32_t nv, float* __restrict__ partial)
{
    __shared__ float s[6][256];
    float mn[3] = {1e37f, 1e37f, 1e37f}, mx[3] = {-1e37f, -1e37f, -1e37f};
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < nv; i += gridDim.x * 256)
        for (int j = 0; j < 3; j++) {
            float c = v[3 * (size_t)i + j];
            if (isfinite(c)) { mn[j] = rminf(mn[j], c); mx[j] = rmaxf(mx[j], c); }
        }
    for (int j = 0; j < 3; j++) { s[j][threadIdx.x] = mn[j]; s[3 + j][threadIdx.x] = mx[j]; }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int j = 0; j < 3; j++) {
                s[j][threadIdx.x] = rminf(s[j][threadIdx.x], s[j][threadIdx.x + w]);
                s[3 + j][threadIdx.x] = rmaxf(s[3 + j][threadIdx.x], s[3 + j][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__device__ inline uint32_t expand10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ inline uint32_t quant10(float c, float mn, float scale)
{
    float q = (c - mn) * scale;
    if (!(q > 0)) return 0;
    if (q >= 1023.f) return 1023;
    return (uint32_t)q;
}

// TriangleSRLNP::setGeometry (RRCollider/IntersectBspFast.cpp:130-187), bit-exact,
// plus the triangle box and its Morton key.  Degenerate triangles (byte 10) get key
// 0xffffffff and sort behind every valid one (valid keys use 30 bits).
__global__ __launch_bounds__(256) void k_tri_setup(const float* __restrict__ verts, const uint32_t* __restrict__ idx, uint32_t ntri,
                                                   float3 mn, float3 scale, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                   float4* __restrict__ plane, float4* __restrict__ sv, float4* __restrict__ ev,
                                                   uint32_t* __restrict__ bytes, float4* __restrict__ blo, float4* __restrict__ bhi,
                                                   uint32_t* __restrict__ nvalid)
{
    uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntri) return;
    const float* a = verts + 3 * (size_t)idx[3 * (size_t)t + 0];
    const float* b = verts + 3 * (size_t)idx[3 * (size_t)t + 1];
    const float* c = verts + 3 * (size_t)idx[3 * (size_t)t + 2];
    float s[3], r[3], l[3], ir[3];
    for (int i = 0; i < 3; i++) { s[i] = a[i]; r[i] = b[i] - a[i]; l[i] = c[i] - a[i]; }
    for (int i = 0; i < 3; i++) ir[i] = 1 / r[i];
    float rxy = l[1] * r[0] - l[0] * r[1];
    float ryz = l[2] * r[1] - l[1] * r[2];
    float rzx = l[0] * r[2] - l[2] * r[0];
    int byte; float ireal;
    if (fabsf(rxy) >= fabsf(ryz)) {
        if (fabsf(rxy) >= fabsf(rzx)) { byte = 0; ireal = 1 / rxy; }
        else                          { byte = 2; ireal = 1 / rzx; }
    } else {
        if (fabsf(ryz) >= fabsf(rzx)) { byte = 1; ireal = 1 / ryz; }
        else                          { byte = 2; ireal = 1 / rzx; }
    }
    if (fabsf(r[0]) >= fabsf(r[1])) { if (fabsf(r[0]) >= fabsf(r[2])) byte += 3; }
    else { if (fabsf(r[1]) >= fabsf(r[2])) byte += 6; }
    float nx = r[1] * l[2] - r[2] * l[1];
    float ny = -r[0] * l[2] + r[2] * l[0];
    float nz = r[0] * l[1] - r[1] * l[0];
    float len = sqrtf(nx * nx + ny * ny + nz * nz);
    nx = nx / len; ny = ny / len; nz = nz / len;
    float nw = -(s[0] * nx + s[1] * ny + s[2] * nz);
    if (!(isfinite(nx) && isfinite(ny) && isfinite(nz))) byte = 10;
    int x, y, z;
    srlnp_axes(byte, x, y, z);
    plane[t] = make_float4(nx, ny, nz, nw);
    sv[t] = make_float4(s[x], s[y], s[z], ireal);
    ev[t] = make_float4(r[x], r[y], l[z], ir[z]);
    bytes[t] = (uint32_t)byte;
    float lo[3], hi[3];
    for (int j = 0; j < 3; j++) {
        lo[j] = rminf(rminf(a[j], b[j]), c[j]);
        hi[j] = rmaxf(rmaxf(a[j], b[j]), c[j]);
    }
    // oracle: l = a<b?a:b; l = l<c?l:c  (same as rminf chain)
    blo[t] = make_float4(lo[0], lo[1], lo[2], 0.f);
    bhi[t] = make_float4(hi[0], hi[1], hi[2], 0.f);
    uint32_t key = 0xffffffffu;
    if (byte <= 8) {
        uint32_t qx = quant10((lo[0] + hi[0]) * 0.5f, mn.x, scale.x);
        uint32_t qy = quant10((lo[1] + hi[1]) * 0.5f, mn.y, scale.y);
        uint32_t qz = quant10((lo[2] + hi[2]) * 0.5f, mn.z, scale.z);
        key = (expand10(qx) << 2) | (expand10(qy) << 1) | expand10(qz);
        atomicAdd(nvalid, 1u);
    }
    keys[t] = key;
    vals[t] = t;
}

// ---- LSD radix sort, 8-bit digits, tiles of 2048 keys per 256-thread block ------------
constexpr int RS_THREADS = 256, RS_ITEMS = 8, RS_TILE = RS_THREADS * RS_ITEMS;

__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift, uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_ITEMS; k++) {
        uint32_t i = base + k * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of m counters in place, one block of 1024 threads
__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                    uint32_t n, int shift, const uint32_t* __restrict__ offs)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[4][256];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    run[tid] = offs[tid * gridDim.x + blockIdx.x];
    for (int w = 0; w < 4; w++) wcnt[w][tid] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_ITEMS; k++) {
        uint32_t i = base + k * RS_THREADS + tid;
        bool valid = i < n;
        uint32_t key = valid ? kin[i] : 0u;
        uint32_t dig = (key >> shift) & 255u;
        uint64_t m = __ballot(valid);
        for (int bit = 0; bit < 8; bit++) {
            uint64_t bal = __ballot((dig >> bit) & 1u);
            m &= ((dig >> bit) & 1u) ? bal : ~bal;
        }
        uint64_t lower = m & ((1ull << lane) - 1ull);
        uint32_t rank = (uint32_t)__popcll(lower);
        if (valid && lower == 0) wcnt[wave][dig] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[dig] + rank;
            for (uint32_t w = 0; w < wave; w++) pos += wcnt[w][dig];
            kout[pos] = key;
            vout[pos] = vin[i];
        }
        __syncthreads();
        run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        for (int w = 0; w < 4; w++) wcnt[w][tid] = 0;
        __syncthreads();
    }
}

// ---- Karras radix tree (T. Karras, "Maximizing parallelism in the construction of
// BVHs, octrees and k-d trees", HPG 2012) -------------------------------------------------
__device__ inline int kdelta(const uint32_t* __restrict__ k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ keys, int n, int* __restrict__ first, int* __restrict__ last,
                                                int* __restrict__ split, int* __restrict__ parent_node, int* __restrict__ parent_leaf)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1) return;
    int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) > 0 ? 1 : -1;
    int dmin = kdelta(keys, n, i, i - d);
    int lmax = 2;
    while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(keys, n, i, j);
    int s = 0, t = l;
    do { t = (t + 1) >> 1; if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t; } while (t > 1);
    int g = i + s * d + (d < 0 ? d : 0);
    int f = min(i, j), la = max(i, j);
    first[i] = f; last[i] = la; split[i] = g;
    if (f == g) parent_leaf[g] = i; else parent_node[g] = i;
    if (la == g + 1) parent_leaf[g + 1] = i; else parent_node[g + 1] = i;
    if (i == 0) parent_node[0] = -1;
}

// Bottom-up union of child boxes, one tree level per launch: launch k computes the nodes
// whose inner children were finished by earlier launches (done[c] in 1..k-1; leaf children
// are always ready) and marks them done[i] = k.  A child finished in the same launch reads
// as 0 or k, so a node only reads boxes written before its launch began: no fences, no
// arrival counters (a device-scope fence per climb step writes back and invalidates a
// whole L2).  The root, node 0, is finished at launch height + 1.
__global__ __launch_bounds__(256) void k_refit_level(int n, const int* __restrict__ first, const int* __restrict__ last,
                                                     const int* __restrict__ split, const float4* __restrict__ plo,
                                                     const float4* __restrict__ phi, float4* __restrict__ nlo, float4* __restrict__ nhi,
                                                     uint32_t* __restrict__ done, uint32_t k)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1 || done[i]) return;
    const int g = split[i];
    const bool la = first[i] == g, lb = last[i] == g + 1;
    if (!la) { const uint32_t d = done[g]; if (d == 0 || d >= k) return; }
    if (!lb) { const uint32_t d = done[g + 1]; if (d == 0 || d >= k) return; }
    float4 al, ah, bl, bh;
    if (la) { al = plo[g]; ah = phi[g]; } else { al = nlo[g]; ah = nhi[g]; }
    if (lb) { bl = plo[g + 1]; bh = phi[g + 1]; } else { bl = nlo[g + 1]; bh = nhi[g + 1]; }
    nlo[i] = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.f);
    nhi[i] = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.f);
    done[i] = k;
}

__global__ __launch_bounds__(256) void k_write_nodes(int n, uint32_t leafmax, float margin, const int* __restrict__ first, const int* __restrict__ last,
                                                     const int* __restrict__ split, const int* __restrict__ parent_node,
                                                     const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                     const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                                     Node* __restrict__ nodes, uint32_t* __restrict__ depth)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1) return;
    int g = split[i];
    float b[2][6]; int32_t ref[2];
    for (int c = 0; c < 2; c++) {
        int cf = c == 0 ? first[i] : g + 1;
        int cl = c == 0 ? g : last[i];
        int cnt = cl - cf + 1;
        float4 lo, hi;
        if (cnt == 1) { lo = plo[cf]; hi = phi[cf]; }
        else { lo = nlo[c == 0 ? g : g + 1]; hi = nhi[c == 0 ? g : g + 1]; }
        b[c][0] = lo.x - margin; b[c][1] = hi.x + margin;
        b[c][2] = lo.y - margin; b[c][3] = hi.y + margin;
        b[c][4] = lo.z - margin; b[c][5] = hi.z + margin;
        ref[c] = ((uint32_t)cnt <= leafmax) ? make_leaf((uint32_t)cf, (uint32_t)cnt) : (c == 0 ? g : g + 1);
    }
    Node nd;
    nd.q0 = make_float4(b[0][0], b[0][1], b[0][2], b[0][3]);
    nd.q1 = make_float4(b[0][4], b[0][5], b[1][0], b[1][1]);
    nd.q2 = make_float4(b[1][2], b[1][3], b[1][4], b[1][5]);
    nd.q3 = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), 0.f, 0.f);
    nodes[i] = nd;
    if ((uint32_t)(last[i] - first[i] + 1) > leafmax) {
        uint32_t dd = 1; int p = parent_node[i];
        while (p >= 0) { dd++; p = parent_node[p]; }
        atomicMax(depth, dd);
    }
}

__global__ __launch_bounds__(256) void k_write_tris(uint32_t n, const uint32_t* __restrict__ order, const float4* __restrict__ plane_u,
                                                    const float4* __restrict__ sv_u, const float4* __restrict__ ev_u,
                                                    const uint32_t* __restrict__ bytes, const float4* __restrict__ blo_u,
                                                    const float4* __restrict__ bhi_u, float4* __restrict__ plane, TriRest* __restrict__ rest,
                                                    float4* __restrict__ plo, float4* __restrict__ phi, uint32_t* __restrict__ slot)
{
    uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    uint32_t t = order[p];
    slot[t] = p;
    plane[p] = plane_u[t];
    TriRest r;
    r.sv = sv_u[t]; r.ev = ev_u[t];
    r.meta = make_uint2(t, bytes[t]); r.pad = make_uint2(0u, 0u);
    rest[p] = r;
    plo[p] = blo_u[t]; phi[p] = bhi_u[t];
}

// ---------------------------------------------------------------------------------------
// traversal

// Closest hit of a device ray batch.  The waves of a launch take rays from one queue (a
// counter in device memory): whenever kRefill of a wave's lanes are free, the wave takes that
// many consecutive rays with one atomic add and hands them to the free lanes in lane order,
// while the other lanes keep tracing (trace_resume yields and keeps their state).  No wave
// idles while rays remain, whichever rays turn out long.  Rays that miss the scene box are
// finished at once.
constexpr uint32_t kRefill = 24;
constexpr uint32_t kRayQueues = 256;      // counters per collider: launches in flight at once

__global__ void k_ray_queue_reset(unsigned long long* q) { if (threadIdx.x == 0) *q = 0ull; }

template <int STACK>
__global__ __launch_bounds__(256) void k_trace_closest(ColliderView c, const float4* __restrict__ ro, const float4* __restrict__ rd,
                                                       uint32_t n, int4* __restrict__ hits, uint8_t* __restrict__ front,
                                                       uint32_t* __restrict__ overflow, unsigned long long* __restrict__ queue)
{
    __shared__ int32_t stk[STACK * 256];
    const uint32_t lane = threadIdx.x & 63u;
    bool busy = false, ok = true;
    bool done = false;              // the queue is empty (wave-uniform)
    uint32_t ray = 0;
    TraceState S;
    for (;;) {
        const uint64_t bm = __ballot(busy);
        const uint32_t nfree = 64u - (uint32_t)__popcll(bm);
        if (!done && (nfree >= kRefill || !bm)) {
            unsigned long long q0 = 0;
            if (lane == 0) q0 = atomicAdd(queue, (unsigned long long)nfree);
            q0 = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(q0 >> 32)) << 32)
                 | (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)q0);
            if (q0 >= n || n - q0 <= nfree) done = true;
            const uint32_t rank = (uint32_t)__popcll(~bm & (lane ? (~0ull >> (64 - lane)) : 0ull));   // free lanes below
            if (!busy && q0 < n && rank < n - q0) {
                ray = (uint32_t)q0 + rank;
                const float4 a = ro[ray], b = rd[ray];
                if (trace_begin(S, c, make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), a.w, b.w)) busy = true;
                else {
                    hits[ray] = make_int4(-1, 0, 0, 0);
                    if (front) front[ray] = 0;
                }
            }
            continue;               // lanes whose rays missed the box may take more at once
        }
        if (!bm) break;             // queue empty and nothing in flight
        if (!busy) continue;
        const uint32_t yield_at = done ? 0u : 64u - kRefill;
        if (!trace_resume<STACK>(S, NoFilter(), c, stk + threadIdx.x, 256, yield_at, ok)) continue;
        busy = false;
        hits[ray] = S.bi >= 0 ? make_int4(S.bi, __float_as_int(S.best), __float_as_int(S.bu), __float_as_int(S.bv))
                              : make_int4(-1, 0, 0, 0);
        if (front) front[ray] = (S.bi >= 0 && S.bf) ? 1 : 0;
    }
    if (!ok && overflow) atomicOr(overflow, 1u);
}

template <int STACK>
__global__ __launch_bounds__(256) void k_trace_stats(ColliderView c, const float4* __restrict__ ro, const float4* __restrict__ rd,
                                                     uint32_t n, unsigned long long* __restrict__ counters)
{
    __shared__ int32_t stk[STACK * 256];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 a = ro[i], b = rd[i];
    TraceResult r; uint32_t nv = 0, nt = 0;
    trace_closest<STACK, true>(NoFilter(), c, make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), a.w, b.w, stk + threadIdx.x, 256, r, nv, nt);
    atomicAdd(&counters[0], (unsigned long long)nv);
    atomicAdd(&counters[1], (unsigned long long)nt);
}

// RRRay-shaped results for lsg_collider_intersect_rrrays: (tri,t,u,v) + plane of the hit.
template <int STACK>
__global__ __launch_bounds__(256) void k_trace_full(ColliderView c, const float4* __restrict__ ro, const float4* __restrict__ rd,
                                                    uint32_t n, int4* __restrict__ hits, float4* __restrict__ planes,
                                                    uint8_t* __restrict__ front, uint32_t* __restrict__ overflow)
{
    __shared__ int32_t stk[STACK * 256];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 a = ro[i], b = rd[i];
    TraceResult r; uint32_t nv = 0, nt = 0;
    bool ok = trace_closest<STACK, false>(NoFilter(), c, make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), a.w, b.w, stk + threadIdx.x, 256, r, nv, nt);
    if (!ok) atomicOr(overflow, 1u);
    if (r.tri >= 0) {
        hits[i] = make_int4(r.tri, __float_as_int(r.t), __float_as_int(r.u), __float_as_int(r.v));
        planes[i] = c.plane[r.slot];
        front[i] = r.front ? 1 : 0;
    } else {
        hits[i] = make_int4(-1, 0, 0, 0);
        front[i] = 0;
    }
}

// ---------------------------------------------------------------------------------------
// host side

template <class T> static int dalloc(T** p, size_t n, uint64_t* acc = nullptr)
{
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) { *p = nullptr; return hip_fail(e, "hipMalloc", __LINE__); }
    if (acc) *acc += sizeof(T) * (n ? n : 1);
    return LSG_OK;
}

static void free_impl(ColliderImpl& c)
{
    dfree(c.d_verts); dfree(c.d_idx); dfree(c.d_nodes); dfree(c.d_nodes4);
    dfree(c.d_plane); dfree(c.d_rest); dfree(c.d_slot); dfree(c.d_queue);
    c = ColliderImpl();
}

static inline unsigned blocks(uint64_t n, unsigned per = 256) { return (unsigned)((n + per - 1) / per); }

// BVH2 -> BVH4, on the host after the GPU build.  Every BVH4 node starts from one BVH2
// node's two children and opens the inner child of largest surface area, up to four
// children, while the traversal stack stays bounded: a BVH2 subtree of height H needs
// at most H stack entries, so with P entries already on the stack a node may take the
// children ch when P + |ch| - 1 + max H(inner child) <= budget.  With budget =
// STACK - 1 >= BVH2 depth this holds at the root, and by construction below it.
static int build(ColliderImpl& C, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, uint32_t leafmax,
                 bool sah, const volatile int* aborting)
{
    auto t0 = std::chrono::steady_clock::now();
    int st;
    C.ntri = nt; C.nverts = nv; C.leaf = leafmax;
    if ((st = dalloc(&C.d_verts, 3 * (size_t)nv, &C.bytes))) return st;
    if ((st = dalloc(&C.d_idx, 3 * (size_t)nt, &C.bytes))) return st;
    LSG_HIP(hipMemcpy(C.d_verts, verts, 12 * (size_t)nv, hipMemcpyHostToDevice));
    LSG_HIP(hipMemcpy(C.d_idx, idx, 12 * (size_t)nt, hipMemcpyHostToDevice));

    // mesh box (RRMesh::getAABB) then Box::init inflation, DELTA_BSP-sized node margin
    float mn[3] = {1e37f, 1e37f, 1e37f}, mx[3] = {-1e37f, -1e37f, -1e37f};
    if (nv) {
        unsigned nb = std::min(blocks(nv), 1024u);
        float* d_part; if ((st = dalloc(&d_part, 6 * (size_t)nb))) return st;
        k_vertex_aabb<<<nb, 256>>>(C.d_verts, nv, d_part);
        std::vector<float> part(6 * (size_t)nb);
        hipError_t e = hipMemcpy(part.data(), d_part, part.size() * 4, hipMemcpyDeviceToHost);
        dfree(d_part);
        if (e != hipSuccess) return hip_fail(e, "aabb", __LINE__);
        for (unsigned b = 0; b < nb; b++)
            for (int j = 0; j < 3; j++) { mn[j] = rminf(mn[j], part[6 * b + j]); mx[j] = rmaxf(mx[j], part[6 * b + 3 + j]); }
    }
    for (int j = 0; j < 3; j++) if (mn[j] > mx[j]) mn[j] = mx[j] = 0;
    float scale[3];
    for (int j = 0; j < 3; j++) { float e = mx[j] - mn[j]; scale[j] = e > 0 ? 1024.0f / e : 0.0f; }
    float imn[3] = {mn[0], mn[1], mn[2]}, imx[3] = {mx[0], mx[1], mx[2]};
    {
        float d = (imx[0] - imn[0] + imx[1] - imn[1] + imx[2] - imn[2]) * 1e-5f;   // geometry.cpp:26
        for (int j = 0; j < 3; j++) { imn[j] -= d; imx[j] += d; }
    }
    {
        // IntersectLinear.cpp:133-139 DELTA_BSP on the inflated box
        float tx = std::max(fabsf(imx[0]), fabsf(imn[0])), ty = std::max(fabsf(imx[1]), fabsf(imn[1])), tz = std::max(fabsf(imx[2]), fabsf(imn[2]));
        float m = tx > ty ? (tx > tz ? tx : tz) : (ty > tz ? ty : tz);
        if (m == 0 || std::isnan(m)) m = 1;
        C.margin = m * 1e-5f;
    }
    memcpy(C.bmin, imn, sizeof imn); memcpy(C.bmax, imx, sizeof imx);
    if (aborting && *aborting) return fail(LSG_ERR_ABORTED, "collider build aborted");

    // per-triangle precompute + keys
    uint32_t *d_keys = nullptr, *d_vals = nullptr, *d_k2 = nullptr, *d_v2 = nullptr, *d_bytes = nullptr, *d_nvalid = nullptr;
    float4 *d_plane_u = nullptr, *d_sv_u = nullptr, *d_ev_u = nullptr, *d_blo_u = nullptr, *d_bhi_u = nullptr;
    uint32_t *d_hist = nullptr, *d_offs = nullptr, *d_stmp = nullptr;
    int *d_first = nullptr, *d_last = nullptr, *d_split = nullptr, *d_pn = nullptr, *d_pl = nullptr;
    float4 *d_plo = nullptr, *d_phi = nullptr, *d_nlo = nullptr, *d_nhi = nullptr;
    uint32_t *d_arrive = nullptr, *d_depth = nullptr;
    auto cleanup = [&]() {
        dfree(d_keys); dfree(d_vals); dfree(d_k2); dfree(d_v2); dfree(d_bytes); dfree(d_nvalid);
        dfree(d_plane_u); dfree(d_sv_u); dfree(d_ev_u); dfree(d_blo_u); dfree(d_bhi_u); dfree(d_hist); dfree(d_offs); dfree(d_stmp);
        dfree(d_first); dfree(d_last); dfree(d_split); dfree(d_pn); dfree(d_pl);
        dfree(d_plo); dfree(d_phi); dfree(d_nlo); dfree(d_nhi); dfree(d_arrive); dfree(d_depth);
    };
#define CK(x) do { if ((st = (x))) { cleanup(); return st; } } while (0)
#define CKH(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); return hip_fail(e_, #x, __LINE__); } } while (0)
    const size_t N = nt ? nt : 1;
    CK(dalloc(&d_keys, N)); CK(dalloc(&d_vals, N)); CK(dalloc(&d_k2, N)); CK(dalloc(&d_v2, N));
    CK(dalloc(&d_bytes, N)); CK(dalloc(&d_nvalid, 1));
    CK(dalloc(&d_plane_u, N)); CK(dalloc(&d_sv_u, N)); CK(dalloc(&d_ev_u, N)); CK(dalloc(&d_blo_u, N)); CK(dalloc(&d_bhi_u, N));
    CKH(hipMemset(d_nvalid, 0, 4));
    if (nt)
        k_tri_setup<<<blocks(nt), 256>>>(C.d_verts, C.d_idx, nt, make_float3(mn[0], mn[1], mn[2]),
                                         make_float3(scale[0], scale[1], scale[2]), d_keys, d_vals, d_plane_u, d_sv_u, d_ev_u,
                                         d_bytes, d_blo_u, d_bhi_u, d_nvalid);
    CKH(hipGetLastError());
    uint32_t nvalid = 0;
    CKH(hipMemcpy(&nvalid, d_nvalid, 4, hipMemcpyDeviceToHost));
    C.nvalid = nvalid;
    if (aborting && *aborting) { cleanup(); return fail(LSG_ERR_ABORTED, "collider build aborted"); }

    // sort (key, triangle) pairs: 4 passes of 8 bits
    if (nt > 1) {
        unsigned nb = blocks(nt, RS_TILE);
        CK(dalloc(&d_hist, 256 * (size_t)nb));
        CK(dalloc(&d_offs, 256 * (size_t)nb));
        CK(dalloc(&d_stmp, scan_tmp_words(256 * nb)));
        uint32_t *ki = d_keys, *vi = d_vals, *ko = d_k2, *vo = d_v2;
        for (int shift = 0; shift < 32; shift += 8) {
            k_rs_hist<<<nb, 256>>>(ki, nt, shift, d_hist);
            // digit-major counts -> each (digit, block)'s first output slot, by the
            // device-wide scan
            CK(scan_exclusive_u32(d_hist, d_offs, 256 * nb, nullptr, d_stmp, nullptr));
            k_rs_scatter<<<nb, 256>>>(ki, vi, ko, vo, nt, shift, d_offs);
            std::swap(ki, ko); std::swap(vi, vo);
        }
        CKH(hipGetLastError());
        // after 4 passes results are back in d_keys/d_vals
    }
    const uint32_t n = nvalid;
    CK(dalloc(&C.d_plane, n, &C.bytes)); CK(dalloc(&C.d_rest, n, &C.bytes));
    CK(dalloc(&C.d_slot, nt, &C.bytes));
    CKH(hipMemset(C.d_slot, 0, sizeof(uint32_t) * (size_t)(nt ? nt : 1)));   // degenerate triangles: never hit
    CK(dalloc(&C.d_queue, kRayQueues, &C.bytes));
    CK(dalloc(&d_plo, n)); CK(dalloc(&d_phi, n));
    if (n) k_write_tris<<<blocks(n), 256>>>(n, d_vals, d_plane_u, d_sv_u, d_ev_u, d_bytes, d_blo_u, d_bhi_u,
                                            C.d_plane, C.d_rest, d_plo, d_phi, C.d_slot);
    CKH(hipGetLastError());
    C.nnodes = n > 1 ? n - 1 : 0;
    CK(dalloc(&C.d_nodes, C.nnodes ? C.nnodes : 1, &C.bytes));
    C.depth = 0;
    if (n == 0) C.root = 0;
    else if (n <= leafmax) C.root = make_leaf(0, n);
    else {
        const int ni = (int)n - 1;
        CK(dalloc(&d_first, ni)); CK(dalloc(&d_last, ni)); CK(dalloc(&d_split, ni)); CK(dalloc(&d_pn, ni)); CK(dalloc(&d_pl, n));
        CK(dalloc(&d_nlo, ni)); CK(dalloc(&d_nhi, ni)); CK(dalloc(&d_arrive, ni)); CK(dalloc(&d_depth, 1));
        CKH(hipMemset(d_arrive, 0, 4 * (size_t)ni));
        CKH(hipMemset(d_depth, 0, 4));
        k_karras<<<blocks(ni), 256>>>(d_keys, (int)n, d_first, d_last, d_split, d_pn, d_pl);
        // one launch per level until the root is done (checked every 8 launches; launches
        // after that find every node done)
        for (uint32_t k = 1;; k++) {
            k_refit_level<<<blocks(ni), 256>>>((int)n, d_first, d_last, d_split, d_plo, d_phi, d_nlo, d_nhi, d_arrive, k);
            if (k % 8 == 0) {
                uint32_t root_done = 0;
                CKH(hipMemcpy(&root_done, d_arrive, 4, hipMemcpyDeviceToHost));
                if (root_done) break;
                if (k > 4096) { cleanup(); return fail(LSG_ERR_INVALID, "refit: no progress"); }
            }
        }
        k_write_nodes<<<blocks(ni), 256>>>((int)n, leafmax, C.margin, d_first, d_last, d_split, d_pn, d_plo, d_phi, d_nlo, d_nhi,
                                           C.d_nodes, d_depth);
        CKH(hipGetLastError());
        CKH(hipMemcpy(&C.depth, d_depth, 4, hipMemcpyDeviceToHost));
        C.root = 0;
        if (sah) {
            // inner nodes rebuilt over the same leaves by SAH on the GPU (the host's
            // lsg_bvh_rebuild_sah gives the same tree); kept unless the new tree would need
            // the larger traversal stack
            Node* s2 = nullptr;
            uint32_t nn = 0, d = 0;
            st = sah_rebuild_device(d_first, d_last, d_split, (int)n, leafmax, C.d_nodes, &s2, &nn, &d);
            if (st > 1) { cleanup(); return st; }
            if (st == 0) {
                const int have = stack_size_for_depth(C.depth), want = stack_size_for_depth(d);
                if (want && (!have || want <= have)) {
                    dfree(C.d_nodes);
                    C.bytes += sizeof(Node) * (uint64_t)nn;
                    C.bytes -= sizeof(Node) * (uint64_t)(C.nnodes ? C.nnodes : 1);
                    C.d_nodes = s2;
                    C.nnodes = nn; C.root = 0; C.depth = d;
                } else dfree(s2);
            }
            st = 0;
        }
    }
    CKH(hipDeviceSynchronize());
    cleanup();
#undef CK
#undef CKH
    if (!stack_size_for_depth(C.depth))
        return fail(LSG_ERR_LIMIT, "BVH depth %u exceeds the traversal stack (63)", C.depth);
    if (C.root < 0 || !C.nnodes) {
        C.root4 = C.root; C.nnodes4 = 0; C.stack_need = 0;
        if ((st = dalloc(&C.d_nodes4, 1, &C.bytes))) return st;
    } else if ((st = collapse4_device(C, (uint32_t)stack_size_for_depth(C.depth) - 1))) return st;
    C.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return LSG_OK;
}

// ---- tree files --------------------------------------------------------------------------
// A built collider on disk: header, then the leaf order of the valid triangles (uint32 each),
// the BVH2 nodes (64 B each) and the BVH4 nodes (128 B each).  Loading uploads the mesh,
// repeats the per-triangle setup and writes the triangles in the stored leaf order, and
// uploads both trees: the Morton sort, Karras tree, refit, SAH rebuild and collapse are
// skipped.  The header holds the mesh's SHA-1 (lsg_mesh_hash) and a 64-bit checksum of the
// payload, so a file of another mesh, an older format or a damaged file is refused.
struct TreeFileHeader {
    char magic[8];
    uint32_t version, header_bytes;
    uint8_t mesh_hash[20];
    uint64_t payload_sum;
    uint32_t nverts, ntri, leaf, flags, nvalid, nnodes, nnodes4, depth, stack_need;
    int32_t root, root4;
    float bmin[3], bmax[3], margin;
    uint64_t payload_bytes;
};
static const char kTreeMagic[8] = {'L', 'S', 'G', 'B', 'V', 'H', '4', 0};

// checksum of the payload: four interleaved FNV-1a style lanes over 64-bit words, folded
// (a damaged or truncated file is caught without hashing hundreds of MB cryptographically)
static uint64_t payload_checksum(const uint8_t* p, size_t n)
{
    const uint64_t prime = 0x100000001b3ull;
    uint64_t h[4] = {0xcbf29ce484222325ull, 0x84222325cbf29ce4ull, 0x9e3779b97f4a7c15ull, 0xc2b2ae3d27d4eb4full};
    size_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int k = 0; k < 4; k++) {
            uint64_t w;
            memcpy(&w, p + i + 8 * k, 8);
            h[k] = (h[k] ^ w) * prime;
        }
    uint64_t r = h[0] ^ (h[1] * 3) ^ (h[2] * 5) ^ (h[3] * 7);
    for (; i < n; i++) r = (r ^ p[i]) * prime;
    return (r ^ (uint64_t)n) * prime;
}

static int load_tree(ColliderImpl& C, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, const TreeFileHeader& h,
                     const uint32_t* order, const Node* nodes, const Node4* nodes4)
{
    auto t0 = std::chrono::steady_clock::now();
    int st;
    C.ntri = nt; C.nverts = nv; C.leaf = h.leaf;
    memcpy(C.bmin, h.bmin, sizeof C.bmin); memcpy(C.bmax, h.bmax, sizeof C.bmax); C.margin = h.margin;
    if ((st = dalloc(&C.d_verts, 3 * (size_t)nv, &C.bytes))) return st;
    if ((st = dalloc(&C.d_idx, 3 * (size_t)nt, &C.bytes))) return st;
    LSG_HIP(hipMemcpy(C.d_verts, verts, 12 * (size_t)nv, hipMemcpyHostToDevice));
    LSG_HIP(hipMemcpy(C.d_idx, idx, 12 * (size_t)nt, hipMemcpyHostToDevice));
    uint32_t *d_keys = nullptr, *d_vals = nullptr, *d_bytes = nullptr, *d_nvalid = nullptr, *d_order = nullptr;
    float4 *d_plane_u = nullptr, *d_sv_u = nullptr, *d_ev_u = nullptr, *d_blo_u = nullptr, *d_bhi_u = nullptr, *d_plo = nullptr, *d_phi = nullptr;
    auto cleanup = [&]() {
        dfree(d_keys); dfree(d_vals); dfree(d_bytes); dfree(d_nvalid); dfree(d_order);
        dfree(d_plane_u); dfree(d_sv_u); dfree(d_ev_u); dfree(d_blo_u); dfree(d_bhi_u); dfree(d_plo); dfree(d_phi);
    };
#define CK(x) do { if ((st = (x))) { cleanup(); return st; } } while (0)
#define CKH(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); return hip_fail(e_, #x, __LINE__); } } while (0)
    const size_t N = nt ? nt : 1;
    CK(dalloc(&d_keys, N)); CK(dalloc(&d_vals, N)); CK(dalloc(&d_bytes, N)); CK(dalloc(&d_nvalid, 1));
    CK(dalloc(&d_plane_u, N)); CK(dalloc(&d_sv_u, N)); CK(dalloc(&d_ev_u, N)); CK(dalloc(&d_blo_u, N)); CK(dalloc(&d_bhi_u, N));
    CKH(hipMemset(d_nvalid, 0, 4));
    if (nt)
        k_tri_setup<<<blocks(nt), 256>>>(C.d_verts, C.d_idx, nt, make_float3(h.bmin[0], h.bmin[1], h.bmin[2]), make_float3(0.f, 0.f, 0.f),
                                         d_keys, d_vals, d_plane_u, d_sv_u, d_ev_u, d_bytes, d_blo_u, d_bhi_u, d_nvalid);
    CKH(hipGetLastError());
    uint32_t nvalid = 0;
    CKH(hipMemcpy(&nvalid, d_nvalid, 4, hipMemcpyDeviceToHost));
    if (nvalid != h.nvalid) { cleanup(); return fail(LSG_ERR_INVALID, "tree file: %u valid triangles stored, %u in the mesh", h.nvalid, nvalid); }
    const uint32_t n = nvalid;
    C.nvalid = n;
    CK(dalloc(&d_order, n)); CK(dalloc(&d_plo, n)); CK(dalloc(&d_phi, n));
    if (n) CKH(hipMemcpy(d_order, order, 4 * (size_t)n, hipMemcpyHostToDevice));
    CK(dalloc(&C.d_plane, n, &C.bytes)); CK(dalloc(&C.d_rest, n, &C.bytes));
    CK(dalloc(&C.d_slot, nt, &C.bytes));
    CKH(hipMemset(C.d_slot, 0, sizeof(uint32_t) * N));
    CK(dalloc(&C.d_queue, kRayQueues, &C.bytes));
    if (n) k_write_tris<<<blocks(n), 256>>>(n, d_order, d_plane_u, d_sv_u, d_ev_u, d_bytes, d_blo_u, d_bhi_u,
                                            C.d_plane, C.d_rest, d_plo, d_phi, C.d_slot);
    CKH(hipGetLastError());
    C.nnodes = h.nnodes; C.root = h.root; C.depth = h.depth;
    CK(dalloc(&C.d_nodes, C.nnodes ? C.nnodes : 1, &C.bytes));
    if (C.nnodes) CKH(hipMemcpy(C.d_nodes, nodes, sizeof(Node) * (size_t)C.nnodes, hipMemcpyHostToDevice));
    C.nnodes4 = h.nnodes4; C.root4 = h.root4; C.stack_need = h.stack_need;
    CK(dalloc(&C.d_nodes4, C.nnodes4 ? C.nnodes4 : 1, &C.bytes));
    if (C.nnodes4) CKH(hipMemcpy(C.d_nodes4, nodes4, sizeof(Node4) * (size_t)C.nnodes4, hipMemcpyHostToDevice));
    CKH(hipDeviceSynchronize());
    cleanup();
#undef CK
#undef CKH
    C.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return LSG_OK;
}

// reads and checks a tree file for the mesh; 0 = header and payload ready
static int read_tree_file(const char* path, const uint8_t mesh_hash[20], uint32_t nv, uint32_t nt, TreeFileHeader& h,
                          std::vector<uint8_t>& payload)
{
    FILE* f = fopen(path, "rb");
    if (!f) return fail(LSG_ERR_INVALID, "tree file %s: cannot open", path);
    const bool head = fread(&h, sizeof h, 1, f) == 1;
    bool ok = head && !memcmp(h.magic, kTreeMagic, 8) && h.version == LSG_TREE_VERSION && h.header_bytes == sizeof h;
    if (ok && (memcmp(h.mesh_hash, mesh_hash, 20) || h.nverts != nv || h.ntri != nt)) {
        fclose(f);
        return fail(LSG_ERR_INVALID, "tree file %s: built for another mesh", path);
    }
    const uint64_t want = 4ull * h.nvalid + sizeof(Node) * (uint64_t)h.nnodes + sizeof(Node4) * (uint64_t)h.nnodes4;
    ok = ok && h.payload_bytes == want && h.nvalid <= nt;
    if (ok) {
        payload.resize(want);
        ok = !want || fread(payload.data(), 1, want, f) == want;
        uint8_t extra;
        ok = ok && fread(&extra, 1, 1, f) == 0;      // nothing after the payload
    }
    fclose(f);
    if (!ok) return fail(LSG_ERR_INVALID, "tree file %s: not a tree file of format %u, or truncated", path, LSG_TREE_VERSION);
    if (payload_checksum(payload.data(), payload.size()) != h.payload_sum)
        return fail(LSG_ERR_INVALID, "tree file %s: damaged (payload checksum)", path);
    return LSG_OK;
}

static int load_from_file(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, const uint8_t hash[20], const char* path,
                          uint32_t want_leaf, uint32_t want_flags, lsg_collider** out)
{
    TreeFileHeader h;
    std::vector<uint8_t> payload;
    int st = read_tree_file(path, hash, nv, nt, h, payload);
    if (st) return st;
    if ((want_leaf && h.leaf != want_leaf) || h.flags != want_flags)
        return fail(LSG_ERR_INVALID, "tree file %s: built with leaf size %u flags %u", path, h.leaf, h.flags);
    if (!stack_size_for_depth(h.depth) || h.stack_need + 1 > (uint32_t)stack_size_for_depth(h.depth))
        return fail(LSG_ERR_INVALID, "tree file %s: depth %u does not fit the traversal stack", path, h.depth);
    const uint32_t* order = (const uint32_t*)payload.data();
    for (uint32_t i = 0; i < h.nvalid; i++)
        if (order[i] >= nt) return fail(LSG_ERR_INVALID, "tree file %s: leaf order names triangle %u", path, order[i]);
    const Node* nodes = (const Node*)(payload.data() + 4ull * h.nvalid);
    const Node4* nodes4 = (const Node4*)(payload.data() + 4ull * h.nvalid + sizeof(Node) * (uint64_t)h.nnodes);
    lsg_collider* c = new (std::nothrow) lsg_collider;
    if (!c) return fail(LSG_ERR_OOM, "lsg_collider_load: host allocation failed");
    if ((st = load_tree(c->impl, verts, nv, idx, nt, h, order, nodes, nodes4))) { free_impl(c->impl); delete c; return st; }
    *out = c;
    return LSG_OK;
}

static int save_to_file(const lsg_collider* c, const char* path, uint32_t flags, uint64_t min_bytes)
{
    const ColliderImpl& C = c->impl;
    TreeFileHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, kTreeMagic, 8);
    h.version = LSG_TREE_VERSION; h.header_bytes = sizeof h;
    h.nverts = C.nverts; h.ntri = C.ntri; h.leaf = C.leaf; h.flags = flags; h.nvalid = C.nvalid;
    h.nnodes = C.nnodes; h.nnodes4 = C.nnodes4; h.depth = C.depth; h.stack_need = C.stack_need;
    h.root = C.root; h.root4 = C.root4;
    memcpy(h.bmin, C.bmin, sizeof h.bmin); memcpy(h.bmax, C.bmax, sizeof h.bmax); h.margin = C.margin;
    h.payload_bytes = 4ull * C.nvalid + sizeof(Node) * (uint64_t)C.nnodes + sizeof(Node4) * (uint64_t)C.nnodes4;
    if (sizeof h + h.payload_bytes < min_bytes) return LSG_OK;    // small trees build faster than a file loads
    std::vector<uint8_t> payload(h.payload_bytes);
    if (C.nvalid) {
        std::vector<TriRest> m(C.nvalid);
        LSG_HIP(hipMemcpy(m.data(), C.d_rest, sizeof(TriRest) * (size_t)C.nvalid, hipMemcpyDeviceToHost));
        uint32_t* order = (uint32_t*)payload.data();
        for (uint32_t i = 0; i < C.nvalid; i++) order[i] = m[i].meta.x;
    }
    if (C.nnodes) LSG_HIP(hipMemcpy(payload.data() + 4ull * C.nvalid, C.d_nodes, sizeof(Node) * (size_t)C.nnodes, hipMemcpyDeviceToHost));
    if (C.nnodes4)
        LSG_HIP(hipMemcpy(payload.data() + 4ull * C.nvalid + sizeof(Node) * (size_t)C.nnodes, C.d_nodes4, sizeof(Node4) * (size_t)C.nnodes4,
                          hipMemcpyDeviceToHost));
    // the mesh hash from the device copy of the mesh
    std::vector<float> v(3 * (size_t)C.nverts);
    std::vector<uint32_t> ix(3 * (size_t)C.ntri);
    if (C.nverts) LSG_HIP(hipMemcpy(v.data(), C.d_verts, 12 * (size_t)C.nverts, hipMemcpyDeviceToHost));
    if (C.ntri) LSG_HIP(hipMemcpy(ix.data(), C.d_idx, 12 * (size_t)C.ntri, hipMemcpyDeviceToHost));
    sha1(v.data(), 4 * v.size(), ix.data(), 4 * ix.size(), h.mesh_hash);
    h.payload_sum = payload_checksum(payload.data(), payload.size());
    // written under a temporary name and renamed, so a reader never sees half a file
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return fail(LSG_ERR_INVALID, "tree file %s: cannot create", path);
    bool ok = fwrite(&h, sizeof h, 1, f) == 1 && (payload.empty() || fwrite(payload.data(), 1, payload.size(), f) == payload.size());
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path)) {
        remove(tmp.c_str());
        return fail(LSG_ERR_INVALID, "tree file %s: write failed", path);
    }
    return LSG_OK;
}

} // namespace lsg

using namespace lsg;

extern "C" {

int lsg_collider_save(const lsg_collider* c, const char* path)
{
    if (!c || !path) return fail(LSG_ERR_INVALID, "lsg_collider_save: null argument");
    return save_to_file(c, path, c->impl.flags, 0);
}

int lsg_collider_load(const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntri, const char* path, lsg_collider** out)
{
    if (!out) return fail(LSG_ERR_INVALID, "lsg_collider_load: out is null");
    *out = nullptr;
    if (!path || (!verts && nverts) || (!indices && ntri)) return fail(LSG_ERR_INVALID, "lsg_collider_load: null argument");
    if (!device_ready()) return fail(LSG_ERR_NO_DEVICE, "lsg_collider_load: no gfx950 device");
    uint8_t hash[20];
    sha1(verts, 12 * (size_t)nverts, indices, 12 * (size_t)ntri, hash);
    TreeFileHeader h;
    std::vector<uint8_t> payload;
    int st = read_tree_file(path, hash, nverts, ntri, h, payload);
    if (st) return st;
    return load_from_file(verts, nverts, indices, ntri, hash, path, 0, h.flags, out);
}

int lsg_collider_create_cached(const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntri, const lsg_build_params* params,
                               const volatile int* aborting, const char* cache_location, lsg_collider** out, int32_t* from_cache)
{
    if (from_cache) *from_cache = 0;
    if (!out) return fail(LSG_ERR_INVALID, "lsg_collider_create_cached: out is null");
    *out = nullptr;
    const bool caching = !(cache_location && !strcmp(cache_location, "*"));
    std::string path;
    const uint32_t want_leaf = params ? params->leaf_size : 0u;
    const uint32_t want_flags = params ? (params->flags & LSG_BUILD_LBVH_ONLY) : 0u;
    if (caching && (verts || !nverts) && (indices || !ntri)) {
        uint8_t hash[20];
        sha1(verts, 12 * (size_t)nverts, indices, 12 * (size_t)ntri, hash);
        path = cache_file_name(hash, LSG_TREE_VERSION, cache_location, ".lsgbvh");
        if (device_ready() && load_from_file(verts, nverts, indices, ntri, hash, path.c_str(), want_leaf, want_flags, out) == LSG_OK) {
            if (from_cache) *from_cache = 1;
            return LSG_OK;
        }
    }
    int st = lsg_collider_create(verts, nverts, indices, ntri, params, aborting, out);
    if (st || !caching) return st;
    // a cache that cannot be written leaves the collider as built (the reference ignores a failed save)
    (void)save_to_file(*out, path.c_str(), want_flags, 1024);
    return LSG_OK;
}

int lsg_collider_create(const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntri,
                        const lsg_build_params* params, const volatile int* aborting, lsg_collider** out)
{
    if (!out) return fail(LSG_ERR_INVALID, "lsg_collider_create: out is null");
    *out = nullptr;
    if ((!verts && nverts) || (!indices && ntri)) return fail(LSG_ERR_INVALID, "lsg_collider_create: null mesh arrays");
    if (ntri >= (1u << 27)) return fail(LSG_ERR_LIMIT, "lsg_collider_create: %u triangles exceeds 2^27", ntri);
    for (uint64_t i = 0; i < 3ull * ntri; i++)
        if (indices[i] >= nverts) return fail(LSG_ERR_INVALID, "lsg_collider_create: index %u out of range at %llu", indices[i], (unsigned long long)i);
    const bool chosen = params && params->leaf_size;
    uint32_t leaf = chosen ? params->leaf_size : 2;
    if (leaf < 1 || leaf > 16) return fail(LSG_ERR_INVALID, "lsg_collider_create: leaf_size %u not in 1..16", leaf);
    if (!device_ready()) return fail(LSG_ERR_NO_DEVICE, "lsg_collider_create: no gfx950 device");
    lsg_collider* c = new (std::nothrow) lsg_collider;
    if (!c) return fail(LSG_ERR_OOM, "lsg_collider_create: host allocation failed");
    const bool sah = !(params && (params->flags & LSG_BUILD_LBVH_ONLY));
    int st = build(c->impl, verts, nverts, indices, ntri, leaf, sah, aborting);
    // Default leaf size: 2 triangles per leaf traverses fastest while the tree fits the
    // 32-entry traversal stack; a deeper tree would need the 64-entry stack, whose LDS
    // halves the waves per CU, so larger leaves are taken instead (up to 8).
    while (!st && !chosen && c->impl.depth + 1 > 32 && leaf < 8) {
        free_impl(c->impl);
        c->impl = ColliderImpl();
        leaf *= 2;
        st = build(c->impl, verts, nverts, indices, ntri, leaf, sah, aborting);
    }
    if (st) { free_impl(c->impl); delete c; return st; }
    c->impl.flags = sah ? 0u : (uint32_t)LSG_BUILD_LBVH_ONLY;
    *out = c;
    return LSG_OK;
}

void lsg_collider_destroy(lsg_collider* c)
{
    if (!c) return;
    free_impl(c->impl);
    delete c;
}

int lsg_collider_info_get(const lsg_collider* c, lsg_collider_info* info)
{
    if (!c || !info) return fail(LSG_ERR_INVALID, "lsg_collider_info_get: null argument");
    const ColliderImpl& C = c->impl;
    info->num_triangles = C.ntri; info->num_valid = C.nvalid; info->num_nodes = C.nnodes;
    info->root_ref = C.root; info->depth = C.depth; info->leaf_size = C.leaf;
    for (int j = 0; j < 3; j++) { info->box_min[j] = C.bmin[j]; info->box_max[j] = C.bmax[j]; }
    info->node_margin = C.margin; info->device_bytes = C.bytes; info->build_ms = C.build_ms;
    info->num_nodes4 = C.nnodes4; info->stack_need = C.stack_need;
    return LSG_OK;
}

int lsg_collider_download(const lsg_collider* c, float* nodes, uint32_t* prim)
{
    if (!c) return fail(LSG_ERR_INVALID, "lsg_collider_download: null collider");
    const ColliderImpl& C = c->impl;
    if (nodes && C.nnodes) LSG_HIP(hipMemcpy(nodes, C.d_nodes, sizeof(Node) * (size_t)C.nnodes, hipMemcpyDeviceToHost));
    if (prim && C.nvalid) {
        std::vector<TriRest> m(C.nvalid);
        LSG_HIP(hipMemcpy(m.data(), C.d_rest, sizeof(TriRest) * (size_t)C.nvalid, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < C.nvalid; i++) prim[i] = m[i].meta.x;
    }
    return LSG_OK;
}

int lsg_collider_download_triangles(const lsg_collider* c, float* plane, float* sv, float* ev, uint32_t* meta)
{
    if (!c) return fail(LSG_ERR_INVALID, "lsg_collider_download_triangles: null collider");
    const ColliderImpl& C = c->impl;
    const size_t n = C.nvalid;
    if (!n) return LSG_OK;
    if (plane) LSG_HIP(hipMemcpy(plane, C.d_plane, 16 * n, hipMemcpyDeviceToHost));
    if (sv || ev || meta) {
        std::vector<TriRest> m(n);
        LSG_HIP(hipMemcpy(m.data(), C.d_rest, sizeof(TriRest) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            if (sv) memcpy(sv + 4 * i, &m[i].sv, 16);
            if (ev) memcpy(ev + 4 * i, &m[i].ev, 16);
            if (meta) { meta[2 * i] = m[i].meta.x; meta[2 * i + 1] = m[i].meta.y; }
        }
    }
    return LSG_OK;
}

int lsg_collider_intersect_device(const lsg_collider* c, const void* ro, const void* rd, uint32_t n, void* hits, void* front, void* stream)
{
    if (!c || (!n)) return c ? LSG_OK : fail(LSG_ERR_INVALID, "lsg_collider_intersect_device: null collider");
    if (!ro || !rd || !hits) return fail(LSG_ERR_INVALID, "lsg_collider_intersect_device: null buffer");
    const ColliderImpl& C = c->impl;
    int stack = stack_size_for_depth(C.depth);
    if (!stack) return fail(LSG_ERR_LIMIT, "tree too deep");
    uint32_t* d_over = nullptr;   // stack >= depth+1 by construction; the kernel guard only prevents overrun
    hipStream_t s = (hipStream_t)stream;
    // a counter of this collider's ring (launches in flight on one collider take different
    // ones), zeroed on the launch's stream; waves enough to fill every CU at 4 per SIMD
    unsigned long long* q = C.d_queue + (__atomic_fetch_add(&C.queue_next, 1u, __ATOMIC_RELAXED) % kRayQueues);
    k_ray_queue_reset<<<1, 64, 0, s>>>(q);
    // resident blocks per CU as the kernel's registers and LDS allow, queried once
    static int cus = 0, per32 = 0, per64 = 0;
    if (!cus) {
        int v = 0;
        cus = (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, 0) == hipSuccess && v > 0) ? v : 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per32, k_trace_closest<32>, 256, 0) != hipSuccess || per32 < 1) per32 = 4;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per64, k_trace_closest<64>, 256, 0) != hipSuccess || per64 < 1) per64 = 2;
    }
    const unsigned resident = (unsigned)cus * (unsigned)(stack <= 32 ? per32 : per64);
    const unsigned nb = std::max(1u, std::min(blocks(n, 256u * 4u), resident));
    if (stack <= 32)
        k_trace_closest<32><<<nb, 256, 0, s>>>(C.view(), (const float4*)ro, (const float4*)rd, n, (int4*)hits, (uint8_t*)front, d_over, q);
    else
        k_trace_closest<64><<<nb, 256, 0, s>>>(C.view(), (const float4*)ro, (const float4*)rd, n, (int4*)hits, (uint8_t*)front, d_over, q);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

int lsg_collider_trace_stats(const lsg_collider* c, const void* ro, const void* rd, uint32_t n, double* node_visits, double* tri_tests)
{
    if (!c || !ro || !rd) return fail(LSG_ERR_INVALID, "lsg_collider_trace_stats: null argument");
    const ColliderImpl& C = c->impl;
    unsigned long long* d_cnt = nullptr;
    LSG_HIP(hipMalloc(&d_cnt, 16));
    LSG_HIP(hipMemset(d_cnt, 0, 16));
    if (n) {
        if (stack_size_for_depth(C.depth) <= 32)
            k_trace_stats<32><<<blocks(n), 256>>>(C.view(), (const float4*)ro, (const float4*)rd, n, d_cnt);
        else
            k_trace_stats<64><<<blocks(n), 256>>>(C.view(), (const float4*)ro, (const float4*)rd, n, d_cnt);
    }
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemcpy(h, d_cnt, 16, hipMemcpyDeviceToHost);
    dfree(d_cnt);
    if (e != hipSuccess) return hip_fail(e, "stats", __LINE__);
    if (node_visits) *node_visits = (double)h[0];
    if (tri_tests) *tri_tests = (double)h[1];
    return LSG_OK;
}

int lsg_collider_intersect_rrrays(const lsg_collider* c, void* rays, uint32_t n, uint8_t* hit_mask, uint32_t* num_hits)
{
    if (!c || (!rays && n)) return fail(LSG_ERR_INVALID, "lsg_collider_intersect_rrrays: null argument");
    if (num_hits) *num_hits = 0;
    if (!n) return LSG_OK;
    const ColliderImpl& C = c->impl;
    uint8_t* R = (uint8_t*)rays;
    std::vector<float4> ho(n), hd(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint8_t* r = R + (size_t)i * LSG_RRRAY_SIZE;
        float o[4], d[4], mn, mx;
        memcpy(o, r + LSG_RRRAY_ORIGIN, 16); memcpy(d, r + LSG_RRRAY_DIR, 16);
        memcpy(&mn, r + LSG_RRRAY_LENGTH_MIN, 4); memcpy(&mx, r + LSG_RRRAY_LENGTH_MAX, 4);
        ho[i] = make_float4(o[0], o[1], o[2], mn);
        hd[i] = make_float4(d[0], d[1], d[2], mx);
    }
    float4 *d_o = nullptr, *d_d = nullptr, *d_pl = nullptr; int4* d_h = nullptr; uint8_t* d_f = nullptr; uint32_t* d_over = nullptr;
    auto cleanup = [&]() { dfree(d_o); dfree(d_d); dfree(d_pl); dfree(d_h); dfree(d_f); dfree(d_over); };
    int st;
#define CK(x) do { if ((st = (x))) { cleanup(); return st; } } while (0)
#define CKH(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); return hip_fail(e_, #x, __LINE__); } } while (0)
    CK(dalloc(&d_o, n)); CK(dalloc(&d_d, n)); CK(dalloc(&d_pl, n)); CK(dalloc(&d_h, n)); CK(dalloc(&d_f, n)); CK(dalloc(&d_over, 1));
    CKH(hipMemcpy(d_o, ho.data(), 16 * (size_t)n, hipMemcpyHostToDevice));
    CKH(hipMemcpy(d_d, hd.data(), 16 * (size_t)n, hipMemcpyHostToDevice));
    CKH(hipMemset(d_over, 0, 4));
    if (stack_size_for_depth(C.depth) <= 32)
        k_trace_full<32><<<blocks(n), 256>>>(C.view(), d_o, d_d, n, d_h, d_pl, d_f, d_over);
    else
        k_trace_full<64><<<blocks(n), 256>>>(C.view(), d_o, d_d, n, d_h, d_pl, d_f, d_over);
    CKH(hipGetLastError());
    std::vector<int4> hh(n); std::vector<float4> hp(n); std::vector<uint8_t> hf(n); uint32_t over = 0;
    CKH(hipMemcpy(hh.data(), d_h, 16 * (size_t)n, hipMemcpyDeviceToHost));
    CKH(hipMemcpy(hp.data(), d_pl, 16 * (size_t)n, hipMemcpyDeviceToHost));
    CKH(hipMemcpy(hf.data(), d_f, n, hipMemcpyDeviceToHost));
    CKH(hipMemcpy(&over, d_over, 4, hipMemcpyDeviceToHost));
    cleanup();
#undef CK
#undef CKH
    if (over) return fail(LSG_ERR_LIMIT, "traversal stack overflow");
    uint32_t hitsn = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint8_t* r = R + (size_t)i * LSG_RRRAY_SIZE;
        bool hit = hh[i].x >= 0;
        if (hit_mask) hit_mask[i] = hit ? 1 : 0;
        if (!hit) continue;
        hitsn++;
        uint32_t flags; memcpy(&flags, r + LSG_RRRAY_FLAGS, 4);
        float t, u, v; memcpy(&t, &hh[i].y, 4); memcpy(&u, &hh[i].z, 4); memcpy(&v, &hh[i].w, 4);
        if (flags & LSG_FILL_DISTANCE) memcpy(r + LSG_RRRAY_HIT_DISTANCE, &t, 4);
        if (flags & LSG_FILL_TRIANGLE) { uint32_t tri = (uint32_t)hh[i].x; memcpy(r + LSG_RRRAY_HIT_TRIANGLE, &tri, 4); }
        if (flags & LSG_FILL_POINT2D) { float uv[2] = {u, v}; memcpy(r + LSG_RRRAY_HIT_POINT2D, uv, 8); }
        if (flags & LSG_FILL_PLANE) { float pl[4] = {hp[i].x, hp[i].y, hp[i].z, hp[i].w}; memcpy(r + LSG_RRRAY_HIT_PLANE, pl, 16); }
        if (flags & LSG_FILL_POINT3D) {
            // IntersectLinear.cpp:40-45 update_hitPoint3d
            float o[4], d[4]; memcpy(o, r + LSG_RRRAY_ORIGIN, 16); memcpy(d, r + LSG_RRRAY_DIR, 16);
            float p[3] = { o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t };
            memcpy(r + LSG_RRRAY_HIT_POINT3D, p, 12);
        }
        if (flags & LSG_FILL_SIDE) { bool fs = hf[i] != 0; memcpy(r + LSG_RRRAY_HIT_FRONTSIDE, &fs, 1); }
    }
    if (num_hits) *num_hits = hitsn;
    return LSG_OK;
}

} // extern "C"
