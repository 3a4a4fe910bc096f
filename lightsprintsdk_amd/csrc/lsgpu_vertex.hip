// lsgpu_vertex.hip -- per-vertex outputs (gfx950): the final gather of one texel per
// triangle (gatherPerTrianglePhysical, gather.cpp:843-952) and its smoothing into the
// solver's stitched vertices (updateVertexBufferFromPerTriangleDataPhysical,
// vertexBuffer.cpp:262-290).
//
// Stitching (host, once per mesh).  lsg_stitch_vertices groups mesh vertices as
// RRMesh::createOptimizedVertices does (RRLessVerticesFilter, RRMeshLessVertices.h:38-230):
// every used vertex gets the normal of its corners (their common normal, or their
// area-weighted sum once two differ, normalised); vertices are visited in (position,
// normal) order and each joins the most recently created group whose representative lies
// within the weld distance on every axis and whose normal is within the angle, else
// starts a group.  With a weld distance of 0 only identical positions can join, and they
// are adjacent in that order, so the search stops at the first representative with
// another position instead of walking back over every vertex of the same x.
//
// Smoothing plan (device).  A corner of a triangle the solver keeps carries
// power = corner angle x triangle area (IVertex::insert, interpol.cpp:51-73, called from
// Object::buildTopIVertices, interpol.cpp:125-172; the angle at the stitched positions,
// the area of the triangle as Triangle::setGeometry computes it, rrcore.cpp:104-146).  The
// corners are sorted by group with a stable radix sort, so each group lists its corners
// in triangle order, and one lane per group sums data[t] * power and divides by the summed
// power (IVertex::getVertexDataFromTriangleData, RRStaticSolver.cpp:130-145), in binary32
// in that order.
#include "lsgpu_internal.h"
#include "lsgpu_gather.h"
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

struct lsg_vertex_plan {
    const lsg_scene* scene = nullptr;
    uint32_t ntri = 0, ngroups = 0;
    uint32_t* d_key = nullptr;        // group of each counted corner (sorted), UINT32_MAX = not counted
    uint32_t* d_corner = nullptr;     // corner index in group order
    uint32_t* d_begin = nullptr;      // first sorted corner of each group (ngroups + 1)
    float* d_power = nullptr;         // power per corner (corner order)
    std::vector<uint32_t> h_key;      // group per corner (corner order), UINT32_MAX = not counted
    std::vector<float> h_power;
    lsg_texel* d_ttex = nullptr;      // one texel per triangle
    lsg_subtexel* d_tsub = nullptr;
    float4* d_tri[LSG_NUM_LAYERS] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

namespace lsg {
namespace {

// ---- binary32 geometry as the solver evaluates it ------------------------------------------
struct V3 { float x, y, z; };
__host__ __device__ inline V3 sub3(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline float size2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__host__ __device__ inline float size(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__host__ __device__ inline V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// acos by the Abramowitz & Stegun 4.4.45 polynomial, as the solver's fast_acos
// (geometry_v.h:114-129): sqrt(1-x) (1.5707288 - 0.2121144 x + 0.0742610 x^2 - 0.0187293 x^3)
__host__ __device__ inline float as_acos(float x)
{
    const float r = sqrtf(1.0f - x);
    float p = -0.0187293f;
    p *= x; p += 0.0742610f;
    p *= x; p -= 0.2121144f;
    p *= x; p += 1.5707288f;
    return p * r;
}

// angle between two edges (angleBetween, geometry_v.cpp:95-106)
__host__ __device__ inline float edge_angle(V3 a, V3 b)
{
    const float d = dot3(divs(a, size(a)), divs(b, size(b)));
    return as_acos(fmaxf(fminf(d, 1.0f), -1.0f));
}

// the area Triangle::setGeometry keeps for a triangle, or 0 when it drops it (rrcore.cpp:
// 112-146 with ignoreSmallerAngle = ignoreSmallerArea = 0)
__host__ __device__ inline float solver_area(V3 v0, V3 v1, V3 v2)
{
    const V3 s1 = sub3(v1, v0), s2 = sub3(v2, v0);
    const V3 n = {s1.y * s2.z - s1.z * s2.y, -s1.x * s2.z + s1.z * s2.x, s1.x * s2.y - s1.y * s2.x};
    const V3 q = divs(n, size(n));
    if (!(isfinite(q.x) && isfinite(q.y) && isfinite(q.z))) return 0.f;
    const float rs = size(s1), ls = size(s2);
    if (!(rs > 0) || !(ls > 0)) return 0.f;
    const float psqr = size2(sub3(divs(s1, rs), divs(s2, ls)));
    const float sina = sqrtf(psqr * (1 - psqr / 4));
    const float a0 = sina / 2 * rs * ls;
    if (psqr <= 0 || psqr >= 4 || 1 - psqr / 4 <= 0 || sina <= 0 || a0 <= 0) return 0.f;
    const float c3 = size(sub3(s2, s1));
    // smallest angle from the side lengths (minAngle, rrcore.cpp:96-102), called as (lsize, rsize, |s2-s1|)
    const float la = ls, lb = rs, lc = c3;
    const float A = as_acos((lb * lb + lc * lc - la * la) / (2 * lb * lc));
    const float B = as_acos((la * la + lc * lc - lb * lb) / (2 * la * lc));
    const float C = as_acos((la * la + lb * lb - lc * lc) / (2 * la * lb));
    const float mn = fminf(fminf(A, B), C);
    if (!isfinite(a0) || !(mn > 0)) return 0.f;
    const float a = size2(s1), b = size2(s2), c = size2(sub3(s2, s1));
    const float area = sqrtf(2 * b * c + 2 * c * a + 2 * a * b - a * a - b * b - c * c) / 4;
    if (!isfinite(area) || !(area > 0)) return 0.f;
    return area;
}

__device__ inline V3 ld3(const float* p, uint32_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// power and sort key of every corner
__global__ __launch_bounds__(256) void k_corner_power(uint32_t ntri, const float* __restrict__ verts, const uint32_t* __restrict__ idx,
                                                      const uint32_t* __restrict__ group, const float* __restrict__ gpos,
                                                      uint32_t* __restrict__ key, uint32_t* __restrict__ corner, float* __restrict__ power)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntri) return;
    const uint32_t g0 = group[3 * t], g1 = group[3 * t + 1], g2 = group[3 * t + 2];
    const float area = solver_area(ld3(verts, idx[3 * t]), ld3(verts, idx[3 * t + 1]), ld3(verts, idx[3 * t + 2]));
    // a triangle the solver dropped, or one whose corners stitching joined, has no corners
    const bool keep = area > 0 && g0 != 0xffffffffu && g1 != 0xffffffffu && g2 != 0xffffffffu && g0 != g1 && g0 != g2 && g1 != g2;
    const uint32_t g[3] = {g0, g1, g2};
    V3 p[3];
    if (keep) { p[0] = ld3(gpos, g0); p[1] = ld3(gpos, g1); p[2] = ld3(gpos, g2); }
    for (int k = 0; k < 3; k++) {
        float w = 0.f;
        if (keep) w = edge_angle(sub3(p[(k + 1) % 3], p[k]), sub3(p[(k + 2) % 3], p[k])) * area;
        key[3 * t + k] = keep ? g[k] : 0xffffffffu;
        corner[3 * t + k] = 3 * t + k;
        power[3 * t + k] = w;
    }
}

// first sorted corner of every group: begin[g] = position of the first key >= g (uncounted
// corners sort last as key ngroups); position i starts the groups in (key[i-1], key[i]]
__global__ __launch_bounds__(256) void k_group_begin(const uint32_t* __restrict__ key, uint32_t n, uint32_t ngroups, uint32_t* __restrict__ begin)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i > n) return;
    auto at = [&](uint32_t j) { const uint32_t k = key[j]; return k == 0xffffffffu ? (int64_t)ngroups : (int64_t)k; };
    const int64_t prev = i ? at(i - 1) : -1;
    const int64_t cur = i < n ? at(i) : (int64_t)ngroups;
    for (int64_t g = prev + 1; g <= cur; g++) begin[g] = i;
}

__global__ __launch_bounds__(256) void k_smooth(const uint32_t* __restrict__ begin, const uint32_t* __restrict__ corner,
                                                const float* __restrict__ power, const float4* __restrict__ data, uint32_t ngroups,
                                                float4* __restrict__ out)
{
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= ngroups) return;
    float x = 0.f, y = 0.f, z = 0.f, top = 0.f;
    for (uint32_t i = begin[g], e = begin[g + 1]; i < e; i++) {
        const uint32_t c = corner[i];
        const float w = power[c];
        const float4 d = data[c / 3];
        x += d.x * w; y += d.y * w; z += d.z * w;
        top += w;
    }
    out[g] = top != 0.f ? make_float4(x / top, y / top, z / top, 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_triangle_texels(uint32_t ntri, lsg_texel* __restrict__ tex, lsg_subtexel* __restrict__ sub)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntri) return;
    tex[t] = lsg_texel{t, t, 1u | (255u << 16), 0x80000000u | t};
    lsg_subtexel s;
    s.triangle = t;
    s.uv[0] = 0.f; s.uv[1] = 0.f; s.uv[2] = 1.f; s.uv[3] = 0.f; s.uv[4] = 0.f; s.uv[5] = 1.f;
    s.area = 1.f;
    sub[t] = s;
}

inline unsigned nblk(uint64_t n) { return (unsigned)((n + 255) / 256); }

template <class T> void vfree(T*& p) { if (p) (void)hipFree(p); p = nullptr; }

void free_plan(lsg_vertex_plan* P)
{
    vfree(P->d_key); vfree(P->d_corner); vfree(P->d_begin); vfree(P->d_power);
    vfree(P->d_ttex); vfree(P->d_tsub);
    for (auto& d : P->d_tri) vfree(d);
}

// ---- stitching (host) -----------------------------------------------------------------------
struct StitchVert { float p[3]; float n[3]; float area; bool used; };

inline bool lex_less(const StitchVert& a, const StitchVert& b)
{
    for (int i = 0; i < 3; i++) { if (a.p[i] < b.p[i]) return true; if (a.p[i] > b.p[i]) return false; }
    for (int i = 0; i < 3; i++) { if (a.n[i] < b.n[i]) return true; if (a.n[i] > b.n[i]) return false; }
    return false;
}

// triangle area of the mesh (RRMesh::getTriangleArea, RRMesh.cpp:112-132)
inline float mesh_area(const float* a, const float* b, const float* c)
{
    auto d2 = [](const float* u, const float* v) {
        const float x = v[0] - u[0], y = v[1] - u[1], z = v[2] - u[2];
        return x * x + y * y + z * z;
    };
    const float A = d2(a, b), B = d2(a, c), C = d2(b, c);
    const float d = 2 * B * C + 2 * C * A + 2 * A * B - A * A - B * B - C * C;
    return d > 0 ? sqrtf(d) * 0.25f : 0.f;
}

} // namespace
} // namespace lsg

using namespace lsg;

extern "C" {

int lsg_stitch_vertices(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri, const float* corner_normals,
                        float weld_distance, float max_smooth_angle, uint32_t* vertex_group, uint32_t* group_rep, uint32_t* ngroups)
{
    if (!verts || !idx || !corner_normals || !vertex_group || !ngroups) return fail(LSG_ERR_INVALID, "lsg_stitch_vertices: null argument");
    std::vector<StitchVert> v(nverts);
    for (uint32_t i = 0; i < nverts; i++) {
        v[i] = StitchVert{{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]}, {0.f, 0.f, 0.f}, 0.f, false};
    }
    for (uint32_t t = 0; t < ntri; t++) {
        for (int j = 0; j < 3; j++)
            if (idx[3 * t + j] >= nverts) return fail(LSG_ERR_INVALID, "lsg_stitch_vertices: triangle %u indexes vertex %u of %u", t, idx[3 * t + j], nverts);
        const float area = mesh_area(verts + 3 * idx[3 * t], verts + 3 * idx[3 * t + 1], verts + 3 * idx[3 * t + 2]);
        for (int j = 0; j < 3; j++) {
            StitchVert& s = v[idx[3 * t + j]];
            const float* n = corner_normals + 9 * (size_t)t + 3 * j;
            if (s.area >= 0) {
                if (s.area == 0 || (n[0] == s.n[0] && n[1] == s.n[1] && n[2] == s.n[2])) {
                    s.n[0] = n[0]; s.n[1] = n[1]; s.n[2] = n[2];
                    s.area += area;
                } else {      // the first differing normal: from now on an area-weighted sum
                    for (int c = 0; c < 3; c++) s.n[c] = s.n[c] * s.area + n[c] * area;
                    s.area = -1;
                }
            } else {
                for (int c = 0; c < 3; c++) s.n[c] += n[c] * area;
            }
            s.used = true;
        }
    }
    for (auto& s : v) {
        const float len = sqrtf(s.n[0] * s.n[0] + s.n[1] * s.n[1] + s.n[2] * s.n[2]);
        if (len) { s.n[0] /= len; s.n[1] /= len; s.n[2] /= len; }
        else { s.n[0] = 1; s.n[1] = 0; s.n[2] = 0; }
    }
    const bool merging = weld_distance >= 0 && max_smooth_angle >= 0;
    std::vector<uint32_t> order(nverts);
    std::iota(order.begin(), order.end(), 0u);
    if (merging)
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return lex_less(v[a], v[b]); });
    const bool identical = max_smooth_angle == 0;
    const float min_dot = (float)std::cos((double)max_smooth_angle);
    std::vector<uint32_t> rep;
    rep.reserve(nverts);
    for (uint32_t d : order) {
        const StitchVert& D = v[d];
        vertex_group[d] = 0xffffffffu;
        if (!D.used) continue;
        uint32_t hit = 0xffffffffu;
        if (merging)
            for (size_t u = rep.size(); u--;) {
                const StitchVert& U = v[rep[u]];
                if (!(std::fabs(D.p[0] - U.p[0]) <= weld_distance)) break;
                if (weld_distance == 0 && (D.p[1] != U.p[1] || D.p[2] != U.p[2])) break;   // left the run of this position
                if (!(std::fabs(D.p[1] - U.p[1]) <= weld_distance && std::fabs(D.p[2] - U.p[2]) <= weld_distance)) continue;
                const bool same = identical ? (D.n[0] == U.n[0] && D.n[1] == U.n[1] && D.n[2] == U.n[2])
                                            : D.n[0] * U.n[0] + D.n[1] * U.n[1] + D.n[2] * U.n[2] >= min_dot;
                if (same) { hit = (uint32_t)u; break; }
            }
        if (hit == 0xffffffffu) {
            hit = (uint32_t)rep.size();
            rep.push_back(d);
        }
        vertex_group[d] = hit;
    }
    *ngroups = (uint32_t)rep.size();
    if (group_rep) std::copy(rep.begin(), rep.end(), group_rep);
    return LSG_OK;
}

// Unwrap seams (UnwrapSeams, RRBuffer.cpp:536-625).  Vertices are stitched as the
// reference stitches them for this purpose (createOptimizedVertices(0, RR_DEG2RAD(1), 0,
// NULL): equal positions, normals within one degree, uvs ignored).  Each stitched vertex
// lists the triangles at it in triangle order, a triangle once per corner there (a
// counting sort of the corners).  Then for every triangle t1 and edge e = (corner e,
// corner e+1) = (a, b), the triangles t2 > t1 listed at b are matched against the list at
// a by one cursor that only moves forward, as the reference's merge does (so a triangle
// listed twice at b matches twice); a match whose uvs at a and b differ from t1's is a
// seam.  The order of the seams, and the smoothing that reads them in that order, are
// therefore the reference's.
int lsg_unwrap_seams(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri, const float* corner_normals,
                     const float* tri_uv, lsg_unwrap_seam* seams, uint32_t cap, uint32_t* nseams)
{
    if (!verts || !idx || !corner_normals || !tri_uv || !nseams || (cap && !seams))
        return fail(LSG_ERR_INVALID, "lsg_unwrap_seams: null argument");
    std::vector<uint32_t> group(nverts);
    uint32_t ng = 0;
    const float one_degree = 1 * ((float)3.14159265358979323846 / 180);
    if (int st = lsg_stitch_vertices(verts, nverts, idx, ntri, corner_normals, 0.f, one_degree, group.data(), nullptr, &ng)) return st;
    std::vector<uint32_t> head(ng + 1, 0), at(3 * (size_t)ntri);
    for (size_t c = 0; c < 3 * (size_t)ntri; c++) head[group[idx[c]] + 1]++;
    for (uint32_t g = 0; g < ng; g++) head[g + 1] += head[g];
    {
        std::vector<uint32_t> put(head.begin(), head.end() - 1);
        for (size_t c = 0; c < 3 * (size_t)ntri; c++) at[put[group[idx[c]]]++] = (uint32_t)(c / 3);
    }
    auto corner_group = [&](uint32_t t, int k) { return group[idx[3 * (size_t)t + k]]; };
    uint32_t n = 0;
    for (uint32_t t1 = 0; t1 < ntri; t1++) {
        const float* uv1 = tri_uv + 6 * (size_t)t1;
        for (int e = 0; e < 3; e++) {
            const int f = (e + 1) % 3;
            const uint32_t a = corner_group(t1, e), b = corner_group(t1, f);
            const uint32_t* la = at.data() + head[a];
            const uint32_t na = head[a + 1] - head[a];
            uint32_t cur = 0;
            for (uint32_t j = head[b]; j < head[b + 1]; j++) {
                const uint32_t t2 = at[j];
                if (t2 <= t1) continue;
                while (cur + 1 < na && la[cur] < t2) cur++;
                if (la[cur] != t2) continue;
                const int ka = corner_group(t2, 0) == a ? 0 : (corner_group(t2, 1) == a ? 1 : 2);
                const int kb = corner_group(t2, 0) == b ? 0 : (corner_group(t2, 1) == b ? 1 : 2);
                const float* uv2 = tri_uv + 6 * (size_t)t2;
                lsg_unwrap_seam sm;
                sm.edge1[0] = uv1[2 * e]; sm.edge1[1] = uv1[2 * e + 1]; sm.edge1[2] = uv1[2 * f]; sm.edge1[3] = uv1[2 * f + 1];
                sm.edge2[0] = uv2[2 * ka]; sm.edge2[1] = uv2[2 * ka + 1]; sm.edge2[2] = uv2[2 * kb]; sm.edge2[3] = uv2[2 * kb + 1];
                const bool same = sm.edge1[0] == sm.edge2[0] && sm.edge1[1] == sm.edge2[1] && sm.edge1[2] == sm.edge2[2] &&
                                  sm.edge1[3] == sm.edge2[3];
                if (same) continue;
                if (n < cap) seams[n] = sm;
                n++;
            }
        }
    }
    *nseams = n;
    return LSG_OK;
}

int lsg_vertex_plan_create(const lsg_scene* scene, const uint32_t* corner_group, const float* group_positions, uint32_t ngroups,
                           lsg_vertex_plan** out)
{
    if (!scene || !corner_group || !out || (ngroups && !group_positions)) return fail(LSG_ERR_INVALID, "lsg_vertex_plan_create: null argument");
    if (!device_ready()) return fail(LSG_ERR_NO_DEVICE, "lsg_vertex_plan_create: no gfx950 device");
    const ColliderImpl& C = scene->impl.col->impl;
    const uint32_t nt = C.ntri;
    for (size_t i = 0; i < 3 * (size_t)nt; i++)
        if (corner_group[i] != 0xffffffffu && corner_group[i] >= ngroups)
            return fail(LSG_ERR_INVALID, "lsg_vertex_plan_create: corner %zu names group %u of %u", i, corner_group[i], ngroups);
    lsg_vertex_plan* P = new lsg_vertex_plan;
    P->scene = scene;
    P->ntri = nt;
    P->ngroups = ngroups;
    const uint32_t nc = 3 * nt;
    uint32_t *d_group = nullptr, *ktmp = nullptr, *vtmp = nullptr, *tmp = nullptr;
    float* d_gpos = nullptr;
    auto bail = [&](int e) { vfree(d_group); vfree(ktmp); vfree(vtmp); vfree(tmp); vfree(d_gpos); free_plan(P); delete P; return e; };
    hipError_t he;
#define VP_HIP(call) do { if ((he = (call)) != hipSuccess) return bail(hip_fail(he, #call, __LINE__)); } while (0)
    const size_t ncb = 4 * (size_t)(nc ? nc : 1);
    VP_HIP(hipMalloc(&P->d_key, ncb)); VP_HIP(hipMalloc(&P->d_corner, ncb)); VP_HIP(hipMalloc(&P->d_power, ncb));
    VP_HIP(hipMalloc(&P->d_begin, 4 * ((size_t)ngroups + 1)));
    VP_HIP(hipMalloc(&d_group, ncb)); VP_HIP(hipMalloc(&ktmp, ncb)); VP_HIP(hipMalloc(&vtmp, ncb));
    VP_HIP(hipMalloc(&d_gpos, 12 * (size_t)(ngroups ? ngroups : 1)));
    VP_HIP(hipMalloc(&tmp, 4 * sort_tmp_words(nc)));
    VP_HIP(hipMemcpy(d_group, corner_group, 4 * (size_t)nc, hipMemcpyHostToDevice));
    if (ngroups) VP_HIP(hipMemcpy(d_gpos, group_positions, 12 * (size_t)ngroups, hipMemcpyHostToDevice));
    if (nt) {
        k_corner_power<<<nblk(nt), 256>>>(nt, C.d_verts, C.d_idx, d_group, d_gpos, P->d_key, P->d_corner, P->d_power);
        VP_HIP(hipGetLastError());
        P->h_key.resize(nc); P->h_power.resize(nc);
        VP_HIP(hipMemcpy(P->h_key.data(), P->d_key, 4 * (size_t)nc, hipMemcpyDeviceToHost));
        VP_HIP(hipMemcpy(P->h_power.data(), P->d_power, 4 * (size_t)nc, hipMemcpyDeviceToHost));
        int e = sort_pairs_u32(P->d_key, P->d_corner, ktmp, vtmp, nc, 32, tmp, nullptr);
        if (e) return bail(e);
    }
    k_group_begin<<<nblk((uint64_t)nc + 1), 256>>>(P->d_key, nc, ngroups, P->d_begin);
    VP_HIP(hipGetLastError());
    VP_HIP(hipDeviceSynchronize());
#undef VP_HIP
    vfree(d_group); vfree(ktmp); vfree(vtmp); vfree(tmp); vfree(d_gpos);
    *out = P;
    return LSG_OK;
}

void lsg_vertex_plan_destroy(lsg_vertex_plan* plan)
{
    if (!plan) return;
    free_plan(plan);
    delete plan;
}

int lsg_vertex_plan_corners(const lsg_vertex_plan* plan, uint32_t* corner_group, float* power)
{
    if (!plan) return fail(LSG_ERR_INVALID, "lsg_vertex_plan_corners: null plan");
    const size_t nc = 3 * (size_t)plan->ntri;
    if (corner_group && nc) std::copy(plan->h_key.begin(), plan->h_key.end(), corner_group);
    if (power && nc) std::copy(plan->h_power.begin(), plan->h_power.end(), power);
    return LSG_OK;
}

int lsg_vertex_plan_smooth(const lsg_vertex_plan* plan, const float* tri_data, float* group_out, void* stream)
{
    if (!plan || !tri_data || !group_out) return fail(LSG_ERR_INVALID, "lsg_vertex_plan_smooth: null argument");
    if (!plan->ngroups) return LSG_OK;
    k_smooth<<<nblk(plan->ngroups), 256, 0, (hipStream_t)stream>>>(plan->d_begin, plan->d_corner, plan->d_power, (const float4*)tri_data,
                                                                    plan->ngroups, (float4*)group_out);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

int lsg_vertex_plan_gather(lsg_vertex_plan* plan, const lsg_gather_params* params, uint32_t layer_mask, float* const* group_out,
                           void* stream)
{
    if (!plan || !params || !group_out) return fail(LSG_ERR_INVALID, "lsg_vertex_plan_gather: null argument");
    if (!layer_mask || layer_mask >= (1u << LSG_NUM_LAYERS)) return fail(LSG_ERR_INVALID, "lsg_vertex_plan_gather: layer mask 0x%x", layer_mask);
    for (int l = 0; l < LSG_NUM_LAYERS; l++)
        if ((layer_mask >> l & 1u) && !group_out[l]) return fail(LSG_ERR_INVALID, "lsg_vertex_plan_gather: no output for layer %d", l);
    const uint32_t nt = plan->ntri;
    if (!nt) return LSG_OK;
    hipStream_t st = (hipStream_t)stream;
    if (!plan->d_ttex) {
        LSG_HIP(hipMalloc(&plan->d_ttex, sizeof(lsg_texel) * nt));
        LSG_HIP(hipMalloc(&plan->d_tsub, sizeof(lsg_subtexel) * nt));
        k_triangle_texels<<<nblk(nt), 256, 0, st>>>(nt, plan->d_ttex, plan->d_tsub);
        LSG_HIP(hipGetLastError());
    }
    float* tri[LSG_NUM_LAYERS] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int l = 0; l < LSG_NUM_LAYERS; l++)
        if (layer_mask >> l & 1u) {
            if (!plan->d_tri[l]) LSG_HIP(hipMalloc(&plan->d_tri[l], 16 * (size_t)nt));
            tri[l] = (float*)plan->d_tri[l];
        }
    lsg_scene* s = const_cast<lsg_scene*>(plan->scene);
    int e = layer_mask == 1u ? lsg_gather_device(s, plan->d_ttex, nt, plan->d_tsub, nt, params, tri[0], stream)
                             : lsg_gather_layers_device(s, plan->d_ttex, nt, plan->d_tsub, nt, params, tri, stream);
    if (e) return e;
    for (int l = 0; l < LSG_NUM_LAYERS; l++)
        if (layer_mask >> l & 1u)
            if ((e = lsg_vertex_plan_smooth(plan, tri[l], group_out[l], stream))) return e;
    return LSG_OK;
}

int lsg_vertex_plan_triangle_data(const lsg_vertex_plan* plan, uint32_t layer, float* rgba)
{
    if (!plan || !rgba || layer >= LSG_NUM_LAYERS) return fail(LSG_ERR_INVALID, "lsg_vertex_plan_triangle_data: bad argument");
    if (!plan->d_tri[layer]) return fail(LSG_ERR_INVALID, "lsg_vertex_plan_triangle_data: layer %u was never gathered", layer);
    LSG_HIP(hipMemcpy(rgba, plan->d_tri[layer], 16 * (size_t)plan->ntri, hipMemcpyDeviceToHost));
    return LSG_OK;
}

} // extern "C"
