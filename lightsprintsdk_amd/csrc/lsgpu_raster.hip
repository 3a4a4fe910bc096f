// lsgpu_raster.hip -- texel rasterisation of one lightmap on the GPU (gfx950).
//
// Produces what enumerateTexelsPartial (RRSolver/lightmap.cpp:83-404, steps 1-2) gives the
// final gather for one object: every texel the object's unwrap covers, the pieces
// (SubTexels, gather.h:60-65) of each triangle inside it, and the texel's coverage flags
// (gather.h:68-80) -- by a design of this library:
//   1. each triangle's candidate texels are the cells of its bounding box in texel units;
//   2. one lane per (triangle, candidate) clips the triangle against the texel's square
//      (four axis-aligned half-planes in texel-local coordinates), and counts the fan
//      pieces and coverage flags of the intersection polygon;
//   3. non-empty pairs are compacted in candidate order (triangle-major) and stable-sorted
//      by texel index, so a texel's pieces come in triangle order and texels row-major;
//   4. runs of one texel become texel records; every pair writes its pieces (corner
//      coordinates in triangle space, area) at its scanned offset.
// Kept from the reference because they are its interface: the flag bits, the 0.03-texel
// edge tolerance of the edge flags (lightmap.cpp:292), the quadrant rule for an edge that
// crosses both texel axes (lightmap.cpp:302-311), and areaInMapSpace = half the piece's
// area in uv units (the reference's triangle-space area times the triangle's map area).
#include "lsgpu_internal.h"
#include <cmath>
#include <vector>

namespace lsg {

namespace {

enum : uint32_t { QX0Y0 = 1, QX1Y0 = 2, QX0Y1 = 4, QX1Y1 = 8, EX0 = 16, EX1 = 32, EY0 = 64, EY1 = 128 };

// a triangle's uv corners, its candidate box of texels and its index in the mesh
struct TriBox { float2 u0, u1, u2; uint32_t x0, y0, w, id; };

__global__ __launch_bounds__(256) void k_raster_boxes(const float* __restrict__ uv, const uint32_t* __restrict__ ids, uint32_t ntri,
                                                      uint32_t W, uint32_t H, TriBox* __restrict__ tb, uint32_t* __restrict__ cnt)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntri) return;
    const float* u = uv + 6 * (size_t)t;
    const float fw = (float)W, fh = (float)H;
    const float2 p0 = make_float2(u[0] * fw, u[1] * fh), p1 = make_float2(u[2] * fw, u[3] * fh), p2 = make_float2(u[4] * fw, u[5] * fh);
    const float2 e1 = make_float2(p1.x - p0.x, p1.y - p0.y), e2 = make_float2(p2.x - p0.x, p2.y - p0.y);
    const float cr = e1.x * e2.y - e1.y * e2.x;
    TriBox b;
    b.u0 = make_float2(u[0], u[1]); b.u1 = make_float2(u[2], u[3]); b.u2 = make_float2(u[4], u[5]);
    b.id = ids[t];
    b.x0 = b.y0 = b.w = 0;
    uint32_t n = 0;
    const float lx = fminf(p0.x, fminf(p1.x, p2.x)), hx = fmaxf(p0.x, fmaxf(p1.x, p2.x));
    const float ly = fminf(p0.y, fminf(p1.y, p2.y)), hy = fmaxf(p0.y, fmaxf(p1.y, p2.y));
    if (cr != 0.f && isfinite(cr) && isfinite(lx) && isfinite(hx) && isfinite(ly) && isfinite(hy) && hx >= 0.f && hy >= 0.f &&
        lx < fw && ly < fh) {
        const uint32_t x0 = (uint32_t)floorf(fmaxf(lx, 0.f)), y0 = (uint32_t)floorf(fmaxf(ly, 0.f));
        const uint32_t x1 = min(W, (uint32_t)floorf(fminf(hx, fw)) + 1u), y1 = min(H, (uint32_t)floorf(fminf(hy, fh)) + 1u);
        if (x1 > x0 && y1 > y0) { b.x0 = x0; b.y0 = y0; b.w = x1 - x0; n = (x1 - x0) * (y1 - y0); }
    }
    tb[t] = b;
    cnt[t] = n;
}

// The triangle (q0, q1, q2, texel-local) clipped to the unit square by the half-planes
// x >= 0, x <= 1, y >= 0, y <= 1 in turn; vertices on a side count as inside, crossings are
// placed exactly on the side.  Consecutive repeats are dropped.  Returns the vertex count
// (0 when fewer than three distinct vertices or no area remain).
__device__ inline int texel_polygon(float2 q0, float2 q1, float2 q2, float2* out)
{
    float2 a[8], b[8];
    a[0] = q0; a[1] = q1; a[2] = q2;
    int n = 3;
#pragma unroll
    for (int side = 0; side < 4; side++) {
        const bool ax = side < 2;                  // clip on x or on y
        const float lim = (side & 1) ? 1.f : 0.f;
        const float sg = (side & 1) ? -1.f : 1.f;  // inside when sg * (c - lim) >= 0
        int m = 0;
        for (int i = 0; i < n; i++) {
            const float2 p = a[i == 0 ? n - 1 : i - 1], c = a[i];
            const float dp = sg * ((ax ? p.x : p.y) - lim), dc = sg * ((ax ? c.x : c.y) - lim);
            if ((dp >= 0.f) != (dc >= 0.f)) {
                const float s = dp / (dp - dc);
                float2 x = make_float2(p.x + (c.x - p.x) * s, p.y + (c.y - p.y) * s);
                if (ax) x.x = lim; else x.y = lim;
                b[m++] = x;
            }
            if (dc >= 0.f) b[m++] = c;
        }
        n = 0;
        for (int i = 0; i < m; i++)
            if (n == 0 || b[i].x != a[n - 1].x || b[i].y != a[n - 1].y) a[n++] = b[i];
        if (n > 1 && a[n - 1].x == a[0].x && a[n - 1].y == a[0].y) n--;
        if (n < 3) return 0;
    }
    float twice = 0.f;
    for (int i = 1; i + 1 < n; i++)
        twice += (a[i].x - a[0].x) * (a[i + 1].y - a[0].y) - (a[i].y - a[0].y) * (a[i + 1].x - a[0].x);
    if (twice == 0.f) return 0;
    for (int i = 0; i < n; i++) out[i] = a[i];
    return n;
}

// Coverage flags of a texel's polygon (texel-local, square [0,1]^2): edge bits for vertices
// within 0.03 texel of a side, quadrant bits for vertices strictly inside a quadrant around
// the centre, and for an edge that changes both the x and the y side of the centre, the two
// quadrants below or above (by the sign of its line's height at the centre's x) and the two
// left or right (by the sign of where it crosses the centre's height).
__device__ inline uint32_t coverage_flags(const float2* v, int n)
{
    constexpr float kErr = 0.03f;
    uint32_t f = 0;
    float px = v[n - 1].x - 0.5f, py = v[n - 1].y - 0.5f;
    for (int i = 0; i < n; i++) {
        const float x = v[i].x - 0.5f, y = v[i].y - 0.5f;
        f |= (x < -0.5f + kErr ? EX0 : 0u) | (x > 0.5f - kErr ? EX1 : 0u) | (y < -0.5f + kErr ? EY0 : 0u) | (y > 0.5f - kErr ? EY1 : 0u);
        if (x != 0.f && y != 0.f) f |= x < 0.f ? (y < 0.f ? QX0Y0 : QX0Y1) : (y < 0.f ? QX1Y0 : QX1Y1);
        if ((px < 0.f) != (x < 0.f) && (py < 0.f) != (y < 0.f)) {
            const float slope = (y - py) / (x - px);
            const float at0 = y - slope * x;              // the line's height at the centre's x
            f |= at0 < 0.f ? (QX0Y0 | QX1Y0) : (QX0Y1 | QX1Y1);
            f |= -at0 / slope < 0.f ? (QX0Y0 | QX0Y1) : (QX1Y0 | QX1Y1);
        }
        px = x; py = y;
    }
    return f;
}

// candidate c -> triangle (largest t with off[t] <= c), texel (x, y)
__device__ inline uint32_t candidate(const uint32_t* off, uint32_t ntri, const TriBox* tb, uint32_t c, uint32_t& x, uint32_t& y)
{
    uint32_t lo = 0, hi = ntri;
    while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (off[mid] <= c) lo = mid; else hi = mid; }
    const TriBox& b = tb[lo];
    const uint32_t k = c - off[lo];
    x = b.x0 + k % b.w;
    y = b.y0 + k / b.w;
    return lo;
}

// the triangle's corners relative to texel (x, y), each u * W - x with one rounding (a
// fused multiply-add), so they keep the precision of small local coordinates
__device__ inline float2 local_corner(float2 u, float fw, float fh, float ox, float oy)
{
    return make_float2(__fmaf_rn(u.x, fw, -ox), __fmaf_rn(u.y, fh, -oy));
}

__device__ inline int clip_candidate(const TriBox& b, uint32_t x, uint32_t y, float fw, float fh, float2* poly, float2& q0)
{
    const float ox = (float)x, oy = (float)y;
    q0 = local_corner(b.u0, fw, fh, ox, oy);
    return texel_polygon(q0, local_corner(b.u1, fw, fh, ox, oy), local_corner(b.u2, fw, fh, ox, oy), poly);
}

// per candidate: (pieces << 8 | flags), and 1 where the pair is not empty
__global__ __launch_bounds__(256) void k_raster_count(const uint32_t* __restrict__ off, uint32_t ntri, const TriBox* __restrict__ tb,
                                                      uint32_t C, float fw, float fh, uint16_t* __restrict__ info,
                                                      uint32_t* __restrict__ nz)
{
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= C) return;
    uint32_t x, y;
    const uint32_t t = candidate(off, ntri, tb, c, x, y);
    float2 poly[8], q0;
    const int n = clip_candidate(tb[t], x, y, fw, fh, poly, q0);
    info[c] = n ? (uint16_t)(((uint32_t)(n - 2) << 8) | coverage_flags(poly, n)) : (uint16_t)0;
    nz[c] = n ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_raster_compact(const uint32_t* __restrict__ off, uint32_t ntri, const TriBox* __restrict__ tb,
                                                        uint32_t C, uint32_t W, const uint32_t* __restrict__ nz,
                                                        const uint32_t* __restrict__ pos, uint32_t* __restrict__ keys,
                                                        uint32_t* __restrict__ vals)
{
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= C || !nz[c]) return;
    uint32_t x, y;
    candidate(off, ntri, tb, c, x, y);
    keys[pos[c]] = x + y * W;
    vals[pos[c]] = c;
}

// per sorted pair: its piece count, and 1 where a run of one texel starts
__global__ __launch_bounds__(256) void k_raster_runs(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t np,
                                                     const uint16_t* __restrict__ info, uint32_t* __restrict__ pieces,
                                                     uint32_t* __restrict__ head)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    pieces[i] = (uint32_t)info[vals[i]] >> 8;
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_raster_texels(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t np,
                                                       const uint16_t* __restrict__ info, const uint32_t* __restrict__ head,
                                                       const uint32_t* __restrict__ ord, const uint32_t* __restrict__ soff,
                                                       uint32_t object_id, uint32_t out_base, lsg_texel* __restrict__ texels,
                                                       uint32_t* __restrict__ too_many)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np || !head[i]) return;
    const uint32_t key = keys[i];
    uint32_t count = 0, flags = 0;
    for (uint32_t j = i; j < np && keys[j] == key; j++) {
        const uint32_t q = info[vals[j]];
        count += q >> 8;
        flags |= q & 255u;
    }
    if (count > 0xffffu) { atomicMax(too_many, count); count = 0xffffu; }
    lsg_texel T;
    T.out_index = out_base + key;
    T.sub_begin = soff[i];
    T.sub_count_flags = count | (flags << 16);
    T.key = object_id * 0x9E3779B1u + key;      // stable per-texel seed id
    texels[ord[i]] = T;
}

__global__ __launch_bounds__(256) void k_raster_emit(const uint32_t* __restrict__ off, uint32_t ntri, const TriBox* __restrict__ tb,
                                                     const uint32_t* __restrict__ vals, uint32_t np, const uint32_t* __restrict__ soff,
                                                     float fw, float fh, float area_scale, lsg_subtexel* __restrict__ subs)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    uint32_t x, y;
    const uint32_t t = candidate(off, ntri, tb, vals[i], x, y);
    const TriBox b = tb[t];
    float2 poly[8], q0;
    const int n = clip_candidate(b, x, y, fw, fh, poly, q0);
    // triangle-space coordinates: p - corner 0 = a e1 + b e2, with the edges from the
    // texel-local corners
    const float2 q1 = local_corner(b.u1, fw, fh, (float)x, (float)y), q2 = local_corner(b.u2, fw, fh, (float)x, (float)y);
    const float2 e1 = make_float2(q1.x - q0.x, q1.y - q0.y), e2 = make_float2(q2.x - q0.x, q2.y - q0.y);
    const float inv = 1.f / (e1.x * e2.y - e1.y * e2.x);
    float2 tc[8];
    for (int k = 0; k < n; k++) {
        const float dx = poly[k].x - q0.x, dy = poly[k].y - q0.y;
        tc[k] = make_float2((dx * e2.y - dy * e2.x) * inv, (e1.x * dy - e1.y * dx) * inv);
    }
    lsg_subtexel* o = subs + soff[i];
    for (int k = 1; k + 1 < n; k++) {
        const float cr = (poly[k].x - poly[0].x) * (poly[k + 1].y - poly[0].y) - (poly[k].y - poly[0].y) * (poly[k + 1].x - poly[0].x);
        lsg_subtexel s;
        s.triangle = b.id;
        s.uv[0] = tc[0].x; s.uv[1] = tc[0].y;
        s.uv[2] = tc[k].x; s.uv[3] = tc[k].y;
        s.uv[4] = tc[k + 1].x; s.uv[5] = tc[k + 1].y;
        s.area = fabsf(cr) * area_scale;
        o[k - 1] = s;
    }
}

inline unsigned blocks(uint64_t n) { return (unsigned)((n + 255) / 256); }

struct DevArena {
    std::vector<void*> ptrs;
    ~DevArena() { for (void* p : ptrs) (void)hipFree(p); }
    template <class T> int get(T** p, size_t n)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, sizeof(T) * (n ? n : 1));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
        ptrs.push_back(q);
        *p = (T*)q;
        return LSG_OK;
    }
};

} // namespace

// One object's texels and pieces on the device.  counts_only: stop once the counts are
// known (texels/subs untouched).  Otherwise texels/subs are device arrays of at least the
// counts a counting call returned.  Synchronous on st.
int raster_object(const float* d_uv, const uint32_t* d_ids, uint32_t ntri, uint32_t W, uint32_t H, uint32_t object_id,
                  uint32_t out_base, bool counts_only, lsg_texel* texels, uint32_t cap_t, lsg_subtexel* subs, uint32_t cap_s,
                  uint32_t* ntex, uint32_t* nsub, hipStream_t st)
{
    *ntex = 0; *nsub = 0;
    if (!ntri) return LSG_OK;
    DevArena A;
    int e;
    TriBox* tb; uint32_t *cnt, *off, *stmp, *tot;
    if ((e = A.get(&tb, ntri)) || (e = A.get(&cnt, ntri)) || (e = A.get(&off, (size_t)ntri + 1)) || (e = A.get(&tot, 4)))
        return e;
    if ((e = A.get(&stmp, scan_tmp_words(ntri)))) return e;
    k_raster_boxes<<<blocks(ntri), 256, 0, st>>>(d_uv, d_ids, ntri, W, H, tb, cnt);
    if ((e = scan_exclusive_u32(cnt, off, ntri, tot, stmp, st))) return e;
    // candidates: 64-bit sum check on the host (each box is at most W*H <= 2^28 cells)
    std::vector<uint32_t> hc(ntri);
    LSG_HIP(hipMemcpyAsync(hc.data(), cnt, 4 * (size_t)ntri, hipMemcpyDeviceToHost, st));
    LSG_HIP(hipStreamSynchronize(st));
    uint64_t C64 = 0;
    for (uint32_t t = 0; t < ntri; t++) C64 += hc[t];
    if (C64 >= (1ull << 31)) return fail(LSG_ERR_LIMIT, "lsg_rasterize_texels: %llu candidate texels", (unsigned long long)C64);
    const uint32_t C = (uint32_t)C64;
    if (!C) return LSG_OK;
    uint16_t* info; uint32_t *nz, *pos, *ctmp;
    if ((e = A.get(&info, C)) || (e = A.get(&nz, C)) || (e = A.get(&pos, C)) || (e = A.get(&ctmp, scan_tmp_words(C)))) return e;
    k_raster_count<<<blocks(C), 256, 0, st>>>(off, ntri, tb, C, (float)W, (float)H, info, nz);
    if ((e = scan_exclusive_u32(nz, pos, C, tot, ctmp, st))) return e;
    uint32_t np = 0;
    LSG_HIP(hipMemcpyAsync(&np, tot, 4, hipMemcpyDeviceToHost, st));
    LSG_HIP(hipStreamSynchronize(st));
    if (!np) return LSG_OK;
    uint32_t *keys, *vals, *k2, *v2, *sorttmp, *pieces, *head, *soff, *ord, *ptmp;
    if ((e = A.get(&keys, np)) || (e = A.get(&vals, np)) || (e = A.get(&k2, np)) || (e = A.get(&v2, np)) ||
        (e = A.get(&sorttmp, sort_tmp_words(np))) || (e = A.get(&pieces, np)) || (e = A.get(&head, np)) ||
        (e = A.get(&soff, np)) || (e = A.get(&ord, np)) || (e = A.get(&ptmp, scan_tmp_words(np))))
        return e;
    k_raster_compact<<<blocks(C), 256, 0, st>>>(off, ntri, tb, C, W, nz, pos, keys, vals);
    int bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)W * H) bits++;
    if ((e = sort_pairs_u32(keys, vals, k2, v2, np, bits, sorttmp, st))) return e;
    k_raster_runs<<<blocks(np), 256, 0, st>>>(keys, vals, np, info, pieces, head);
    if ((e = scan_exclusive_u32(pieces, soff, np, tot, ptmp, st))) return e;
    if ((e = scan_exclusive_u32(head, ord, np, tot + 1, ptmp, st))) return e;
    uint32_t h[2] = {0, 0};
    LSG_HIP(hipMemcpyAsync(h, tot, 8, hipMemcpyDeviceToHost, st));
    LSG_HIP(hipStreamSynchronize(st));
    *nsub = h[0]; *ntex = h[1];
    if (counts_only) return LSG_OK;
    if (!texels || !subs || cap_t < h[1] || cap_s < h[0])
        return fail(LSG_ERR_LIMIT, "lsg_rasterize_texels: need %u texels / %u subtexels", h[1], h[0]);
    LSG_HIP(hipMemsetAsync(tot + 2, 0, 4, st));
    k_raster_texels<<<blocks(np), 256, 0, st>>>(keys, vals, np, info, head, ord, soff, object_id, out_base, texels, tot + 2);
    const float area_scale = (float)(0.25 / ((double)W * (double)H));
    k_raster_emit<<<blocks(np), 256, 0, st>>>(off, ntri, tb, vals, np, soff, (float)W, (float)H, area_scale, subs);
    uint32_t worst = 0;
    LSG_HIP(hipMemcpyAsync(&worst, tot + 2, 4, hipMemcpyDeviceToHost, st));
    LSG_HIP(hipStreamSynchronize(st));
    if (worst) return fail(LSG_ERR_LIMIT, "lsg_rasterize_texels: a texel has %u subtexels (at most 65535)", worst);
    return LSG_OK;
}

} // namespace lsg

using namespace lsg;

extern "C" int lsg_rasterize_texels(const float* tri_uv, const uint32_t* tri_ids, uint32_t ntri, uint32_t width, uint32_t height,
                                    uint32_t object_id, lsg_texel* texels, uint32_t* ntexels, lsg_subtexel* subs, uint32_t* nsubs,
                                    uint32_t out_base)
{
    if (!ntexels || !nsubs || (ntri && (!tri_uv || !tri_ids))) return fail(LSG_ERR_INVALID, "lsg_rasterize_texels: null argument");
    if (!width || !height || (uint64_t)width * height > (1ull << 28))
        return fail(LSG_ERR_INVALID, "lsg_rasterize_texels: bad map size %ux%u", width, height);
    if (!device_ready()) return fail(LSG_ERR_NO_DEVICE, "lsg_rasterize_texels: no gfx950 device");
    const uint32_t cap_t = *ntexels, cap_s = *nsubs;
    const bool counts_only = !texels && !subs;
    if (!counts_only && (!texels || !subs)) return fail(LSG_ERR_INVALID, "lsg_rasterize_texels: give both output arrays");
    hipStream_t st = nullptr;
    float* d_uv = nullptr;
    uint32_t* d_ids = nullptr;
    lsg_texel* d_t = nullptr;
    lsg_subtexel* d_s = nullptr;
    auto done = [&](int r) {
        if (d_uv) (void)hipFree(d_uv);
        if (d_ids) (void)hipFree(d_ids);
        if (d_t) (void)hipFree(d_t);
        if (d_s) (void)hipFree(d_s);
        return r;
    };
    hipError_t e;
    if ((e = hipMalloc(&d_uv, 24 * (size_t)(ntri ? ntri : 1))) != hipSuccess) return done(hip_fail(e, "hipMalloc", __LINE__));
    if ((e = hipMalloc(&d_ids, 4 * (size_t)(ntri ? ntri : 1))) != hipSuccess) return done(hip_fail(e, "hipMalloc", __LINE__));
    if (ntri && (e = hipMemcpy(d_uv, tri_uv, 24 * (size_t)ntri, hipMemcpyHostToDevice)) != hipSuccess) return done(hip_fail(e, "h2d", __LINE__));
    if (ntri && (e = hipMemcpy(d_ids, tri_ids, 4 * (size_t)ntri, hipMemcpyHostToDevice)) != hipSuccess) return done(hip_fail(e, "h2d", __LINE__));
    uint32_t nt = 0, ns = 0;
    int r = raster_object(d_uv, d_ids, ntri, width, height, object_id, out_base, true, nullptr, 0, nullptr, 0, &nt, &ns, st);
    if (r) return done(r);
    *ntexels = nt; *nsubs = ns;
    if (counts_only) return done(LSG_OK);
    if (cap_t < nt || cap_s < ns) return done(fail(LSG_ERR_LIMIT, "lsg_rasterize_texels: need %u texels / %u subtexels", nt, ns));
    if ((e = hipMalloc(&d_t, sizeof(lsg_texel) * (size_t)(nt ? nt : 1))) != hipSuccess) return done(hip_fail(e, "hipMalloc", __LINE__));
    if ((e = hipMalloc(&d_s, sizeof(lsg_subtexel) * (size_t)(ns ? ns : 1))) != hipSuccess) return done(hip_fail(e, "hipMalloc", __LINE__));
    r = raster_object(d_uv, d_ids, ntri, width, height, object_id, out_base, false, d_t, nt, d_s, ns, &nt, &ns, st);
    if (r) return done(r);
    if (nt && (e = hipMemcpy(texels, d_t, sizeof(lsg_texel) * (size_t)nt, hipMemcpyDeviceToHost)) != hipSuccess) return done(hip_fail(e, "d2h", __LINE__));
    if (ns && (e = hipMemcpy(subs, d_s, sizeof(lsg_subtexel) * (size_t)ns, hipMemcpyDeviceToHost)) != hipSuccess) return done(hip_fail(e, "d2h", __LINE__));
    return done(LSG_OK);
}
