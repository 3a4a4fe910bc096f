// lsgpu_internal.h -- shared host/device definitions of liblsgpu (gfx950).
//
// All float arithmetic that decides a hit is written in the reference's
// operation order and compiled with -ffp-contract=off and correctly rounded
// division/sqrt, so per-triangle accept/reject and hit distance are bit-identical
// to LightsprintCore's IT_BSP_FASTER triangle test (IntersectBspFast.cpp:189-252).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include "lsgpu.h"

namespace lsg {

// ---- error plumbing ------------------------------------------------------------------
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, int line);
bool device_ready();   // a gfx950 device is present and selected

#define LSG_HIP(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return ::lsg::hip_fail(e_, #call, __LINE__); } while (0)

// ---- tree layout ---------------------------------------------------------------------
// BVH2 node, 64 bytes, four float4 loads:
//   q0 = (lo0.x, hi0.x, lo0.y, hi0.y)   q1 = (lo0.z, hi0.z, lo1.x, hi1.x)
//   q2 = (lo1.y, hi1.y, lo1.z, hi1.z)   q3 = (ref0, ref1, 0, 0) as int bits
// ref >= 0: internal node index; ref < 0: leaf, start = (ref>>4)&0x7ffffff,
// count = (ref&15)+1 triangles in leaf order.
struct alignas(16) Node { float4 q0, q1, q2, q3; };

// BVH4 node, 128 bytes (one cache line), collapsed from the BVH2 above and used by
// every GPU traversal:
//   q[0] = lo.x of children 0..3   q[1] = hi.x   q[2] = lo.y   q[3] = hi.y
//   q[4] = lo.z                    q[5] = hi.z   q[6] = child refs (int bits)
// Refs as in the BVH2 (>= 0 inner BVH4 node, < 0 leaf); kNoChild marks an unused slot.
struct alignas(128) Node4 { float4 q[8]; };
constexpr int32_t kNoChild = 0x7ffffffe;

// Triangle in leaf order, 48 bytes + 8 bytes meta (TriangleSRLNP,
// IntersectBspFast.h:33-49, with the projection-selected components pre-picked):
//   plane = n4;  sv = (s[x], s[y], s[z], intersectReal);  ev = (r[x], r[y], l[z], ir[z])
//   meta  = (original triangle index, intersectByte)
struct TriSoA { float4* plane; float4* sv; float4* ev; uint2* meta; };

constexpr uint32_t kLeafBit = 0x80000000u;
__host__ __device__ inline bool ref_is_leaf(int32_t r) { return r < 0; }
__host__ __device__ inline uint32_t leaf_start(int32_t r) { return ((uint32_t)r >> 4) & 0x7ffffffu; }
__host__ __device__ inline uint32_t leaf_count(int32_t r) { return ((uint32_t)r & 15u) + 1u; }
__host__ __device__ inline int32_t make_leaf(uint32_t start, uint32_t count) { return (int32_t)(kLeafBit | (start << 4) | (count - 1u)); }

// projection axes for intersectByte (IntersectBspFast.cpp:202-204): x,y from byte%3, z from byte/3
__host__ __device__ inline void srlnp_axes(int byte, int& x, int& y, int& z)
{
    int m = byte % 3, d = byte / 3;
    x = m; y = (m + 1) % 3; z = (d == 0) ? 2 : (d == 1 ? 0 : 1);
    if (byte > 8) { x = 2; y = 0; z = 1; }
}

__device__ inline float sel3(float a, float b, float c, int i) { return i == 0 ? a : (i == 1 ? b : c); }

// Box::intersect, RRCollider/geometry.cpp:38-86 (scalar path; minf/maxf are the
// reference's ternaries, not fminf, so NaN behaviour matches).
__host__ __device__ inline float rminf(float a, float b) { return a < b ? a : b; }
__host__ __device__ inline float rmaxf(float a, float b) { return a > b ? a : b; }
__device__ inline bool ref_box_intersect(const float3 mn, const float3 mx, const float3 o, const float3 d,
                                         float rmin, float rmax, float& lo, float& hi)
{
    float l1 = (mn.x - o.x) / d.x, l2 = (mx.x - o.x) / d.x;
    float lmin = rminf(l1, l2), lmax = rmaxf(l1, l2);
    l1 = (mn.y - o.y) / d.y; l2 = (mx.y - o.y) / d.y;
    lmin = rmaxf(rminf(l1, l2), lmin); lmax = rminf(rmaxf(l1, l2), lmax);
    l1 = (mn.z - o.z) / d.z; l2 = (mx.z - o.z) / d.z;
    lmin = rmaxf(rminf(l1, l2), lmin); lmax = rminf(rmaxf(l1, l2), lmax);
    lmin = rmaxf(rmin, lmin); lmax = rminf(rmax, lmax);
    lo = lmin; hi = lmax;
    return (lmax >= 0.f) & (lmax >= lmin);
}

// Plane distance (IntersectBspFast.cpp:349-352) of one triangle.
__device__ inline float plane_distance(const float4 n, const float3 o, const float3 d, float& ndotdir)
{
    ndotdir = d.x * n.x + d.y * n.y + d.z * n.z;
    float ndoto = o.x * n.x + o.y * n.y + o.z * n.z + n.w;
    return -ndoto / ndotdir;
}

// intersect_triangleSRLNP (IntersectBspFast.cpp:189-252) at distance t already in range.
__device__ inline bool srlnp_test(const float4 sv, const float4 ev, int byte, const float3 o, const float3 d,
                                  float t, float& u, float& v)
{
    float px = o.x + d.x * t, py = o.y + d.y * t, pz = o.z + d.z * t;   // IntersectLinear.cpp:40-45
    int x, y, z;
    srlnp_axes(byte, x, y, z);
    float pX = sel3(px, py, pz, x), pY = sel3(px, py, pz, y), pZ = sel3(px, py, pz, z);
    v = ((pY - sv.y) * ev.x - (pX - sv.x) * ev.y) * sv.w;
    if (v < 0 || v > 1) return false;
    u = (pZ - sv.z - ev.z * v) * ev.w;
    if (u < 0 || u + v > 1) return false;
    return byte <= 8;
}

// ---- device-side collider view ---------------------------------------------------------
// The rest of a triangle's test record in leaf order, 48 bytes: projected edge data and
// (original index, intersectByte), read together once a ray passes the plane test.
struct alignas(16) TriRest { float4 sv, ev; uint2 meta; uint2 pad; };

struct ColliderView {
    const Node* nodes;
    const Node4* nodes4;
    int32_t root4;
    const float4* plane; const TriRest* rest;
    int32_t root;
    uint32_t nvalid;
    float3 bmin, bmax;     // inflated mesh box
};

struct ColliderImpl {
    uint32_t ntri = 0, nverts = 0, nvalid = 0, nnodes = 0, depth = 0, leaf = 4;
    uint32_t nnodes4 = 0, stack_need = 0;   // BVH4 nodes; worst-case traversal stack entries
    int32_t root4 = 0;
    Node4* d_nodes4 = nullptr;
    int32_t root = 0;
    float bmin[3], bmax[3], margin = 0;
    float* d_verts = nullptr; uint32_t* d_idx = nullptr;
    Node* d_nodes = nullptr;
    float4* d_plane = nullptr; TriRest* d_rest = nullptr;
    uint32_t* d_slot = nullptr;   // leaf-order slot of each original triangle
    uint32_t flags = 0;           // lsg_build_params flags the tree was built with
    unsigned long long* d_queue = nullptr;   // ray-queue counters of closest-hit launches (a ring)
    mutable uint32_t queue_next = 0;   // next counter a launch takes (host atomic)
    uint64_t bytes = 0;
    double build_ms = 0;
    ColliderView view() const
    {
        ColliderView v;
        v.nodes = d_nodes; v.plane = d_plane; v.rest = d_rest;
        v.nodes4 = d_nodes4; v.root4 = root4;
        v.root = root; v.nvalid = nvalid;
        v.bmin = make_float3(bmin[0], bmin[1], bmin[2]);
        v.bmax = make_float3(bmax[0], bmax[1], bmax[2]);
        return v;
    }
};

int stack_size_for_depth(uint32_t depth);   // 0 when too deep

// ---- device-wide primitives (lsgpu_prim.hip) --------------------------------------------
// words of scratch scan_exclusive_u32 needs for n values
size_t scan_tmp_words(uint32_t n);
// out[i] = in[0] + ... + in[i-1] (in != out); *total (device, optional) = the sum of all
int scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, uint32_t* tmp, hipStream_t st);
size_t sort_tmp_words(uint32_t n);
// stable sort of n (key, value) pairs by the low `bits` bits of the keys; the result is
// in keys/vals, keys_tmp/vals_tmp are n-word scratch
int sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n, int bits,
                   uint32_t* tmp, hipStream_t st);
// SHA-1 of the concatenation a || b (lsgpu_hash.cpp)
void sha1(const void* a, size_t na, const void* b, size_t nb, uint8_t out[20]);
// RRHash::getFileName's cache file name for a hash (lsgpu_hash.cpp)
std::string cache_file_name(const uint8_t hash[20], uint32_t version, const char* location, const char* extension);
// The same rebuild on the GPU (lsgpu_sah.hip) from the LBVH's Karras arrays and nodes:
// *out_nodes (device, *out_count nodes, root 0) and its depth; 0 = rebuilt, 1 = fewer than
// two leaves, else an LSG_ERR_* code
int sah_rebuild_device(const int* d_first, const int* d_last, const int* d_split, int nprim, uint32_t leafmax, const Node* d_nodes,
                       Node** out_nodes, uint32_t* out_count, uint32_t* out_depth);
// The BVH4 of C's BVH2 on the GPU (lsgpu_sah.hip): C.d_nodes4, nnodes4, root4, stack_need
struct ColliderImpl;
int collapse4_device(ColliderImpl& C, uint32_t budget);
// SAH rebuild of a BVH2's inner nodes over its leaves (lsgpu_sah.cpp): 0 = rebuilt,
// 1 = fewer than two leaves, 2 = malformed tree
int sah_rebuild(const float* nodes, uint32_t nnodes, int32_t root, float* out, uint32_t* out_nodes, int32_t* out_root,
                uint32_t* out_depth);

} // namespace lsg

struct lsg_collider { lsg::ColliderImpl impl; };
