// lsgpu_gather.hip -- per-texel lightmap final gather on gfx950.
//
// GPU statement of processTexel (RRSolver/gather.cpp:600-836) for opaque
// Lambertian scenes: hemisphere gather of exitance (GatheredIrradianceHemisphere,
// :164-343) plus shadow rays to point/directional lights
// (GatheredIrradianceLights, :356-586) with the reliability loop of :652-786 and
// the collision-handler rules of RRCollisionHandlerFinalGathering
// (RRStaticSolver/pathtracer.h:117-252).
//
// Work mapping: one 64-lane wavefront per texel (at most 16 or 32 shooters per
// series: a group of 16 or 32 lanes per texel).  The texel's shooters (1024 at
// quality 1024) are dealt to lanes 64 at a time; every lane casts the hemisphere
// ray of its shooter and, on the shooters the reference's interleave rule picks,
// one shadow ray per light.  Hemisphere directions come from the reference's
// HomogenousFiller2 sequence: its rejection sampling is serial, so the wave
// tests 64 candidate indices at once and hands accepted ones to lanes in order
// with a ballot + popcount rank.  Sums are reduced across the wave with shuffles.
//
// The reference's rand() draws (filler seed gather.h:102, texel jitter
// gather.cpp:742-749) become a seeded counter hash; the CPU oracle uses the same
// rule.  Each shooter's subtexel follows the reference's own area walk
// (gather.cpp:645-647, 669, 680-688), which is sequential over a texel's shooters:
// a pre-pass (k_subtexel_walk, one lane per texel) records where the first series
// enters each subtexel, and later series (or texels the record cannot describe)
// step the walk in the kernel, group-uniform, as the shooters are handed out.
//
// A hemisphere hit reads the flat exitance of the side it sees and, where the scene
// has per-corner indirect irradiance, adds its interpolation at the hit's hitPoint2d
// times the albedo (pathtracer.cpp:326-348, rrcore.cpp:262-273).
#include "lsgpu_internal.h"
#include "lsgpu_trace.h"
#include "lsgpu_gather.h"
#include <atomic>
#include <cmath>
#include <vector>
#include <cstring>

namespace lsg {

constexpr int kFeatEnv = 1, kFeatMats = 2;

// the kernel instantiation a launch needs
static inline int features(const GatherArgs& A)
{
    return (A.sc.env_type ? kFeatEnv : 0) | (A.sc.pm.tri_mat ? kFeatMats : 0);
}



__device__ inline uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// HomogenousFiller2::GetTrianglePoint (gather.cpp:66-82)
__device__ inline void filler_point(uint32_t n, float& x, float& y)
{
    x = 0.f; y = 0.f;
    float dist = 1.f;
    while (n) {
        const uint32_t k = n & 3u;
        const float dx = (k == 2u) ? 0.86602540378444f : ((k == 3u) ? -0.86602540378444f : 0.f);
        const float dy = (k == 0u) ? 0.f : ((k == 1u) ? 1.f : -0.5f);
        const float dd = (k == 0u) ? -0.5f : 0.5f;
        x += dist * dx;
        y += dist * dy;
        dist *= dd;
        n >>= 2;
    }
}

// seeded replacement of gather.cpp:742-749
__device__ inline void jitter(uint32_t key, uint32_t i, float& u, float& v)
{
    uint32_t a = lowbias32(key + 0x9E3779B9u * (2u * i + 1u)) >> 8;
    uint32_t b = lowbias32(key + 0x9E3779B9u * (2u * i + 2u)) >> 8;
    const uint32_t M = 0xFFFFFFu;
    if (a + b > M) { a = M - a; b = M - b; }
    u = (float)a / (float)M;
    v = (float)b / (float)M;
}

// shots among the first s shooters of a series of n with r rounds, r < n
// (closed form of the counter rule in gather.cpp:763,779: shot k happens at
// shooter floor(k*(n-1)/r)+1)
__device__ inline uint32_t shots_before(uint32_t s, uint32_t n, uint32_t r)
{
    if (s == 0) return 0;
    uint64_t c = ((uint64_t)(s - 1) * r + (n - 2)) / (n - 1);
    return (uint32_t)(c < r ? c : r);
}

__device__ inline uint64_t lanemask_lt(uint32_t lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }

// position of the k-th (0-based) set bit of m (k < popcount(m))
__device__ inline uint32_t nth_set_bit(uint64_t m, uint32_t k)
{
    uint32_t lo = 0;
    for (uint32_t step = 32; step; step >>= 1) {
        uint32_t p = lo + step;
        uint64_t below = p >= 64 ? m : (m & ((1ull << p) - 1ull));
        if ((uint32_t)__popcll(below) <= k) lo = p;
    }
    return lo;
}

// sums over aligned groups of G lanes (G a power of two <= 64; 64 = the whole wave).
// Lanes with nothing to add hold 0, so for a series of at most 16 shooters the three
// widths give the same bits.
template <int G> __device__ inline float group_sum(float v)
{
    for (int o = G / 2; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
template <int G> __device__ inline uint32_t group_sum_u(uint32_t v)
{
    for (int o = G / 2; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// RRCollisionHandlerFinalGathering::collides (pathtracer.h:117-242) for opaque
// materials: mode 0 = hemisphere (accept the rendered side), 1 = shadow ray
// (any shadow caster, both sides [#45]).
template <bool MATS = false>
struct GatherFilterT {
    const float* verts; const uint32_t* idx; const uint8_t* side;
    uint32_t shooter; float near_t; int mode; float rug_below;
    PMView pm;              // MATS: point materials of the scene
    uint32_t qpoint, qinterp;   // MATS: the quality this ray kind compares with minimalQualityForPointMaterials
    __device__ bool same_location(uint32_t tri) const
    {
        const uint32_t* si = idx + 3 * (size_t)shooter;
        const uint32_t* hi = idx + 3 * (size_t)tri;
        for (int i = 0; i < 3; i++) {
            const float* hv = verts + 3 * (size_t)hi[i];
            bool eq = false;
            for (int k = 0; k < 3; k++) {
                const float* sv = verts + 3 * (size_t)si[k];
                eq |= (hv[0] == sv[0]) & (hv[1] == sv[1]) & (hv[2] == sv[2]);
            }
            if (!eq) return false;
        }
        return true;
    }
    __device__ bool accept(uint32_t tri, bool front, float t, float u, float v) const
    {
        if (tri == shooter) return false;                              // pathtracer.h:133
        if (t < near_t && same_location(tri)) return false;             // pathtracer.h:139-166
        const uint32_t m = side[tri];
        const bool ok = mode == 0 ? ((m & (front ? 1u : 2u)) != 0) : ((m & 4u) != 0);
        if constexpr (MATS) {
            // a keyed texel that is fully transparent lets the ray through (pathtracer.h:194-226,
            // RRObject.cpp:225-275)
            const uint32_t mi = ok ? pm.tri_mat[tri] : 0xffffffffu;
            if (mi != 0xffffffffu) {
                const lsg_point_material& mat = pm.mats[mi];
                if (qpoint > mat.min_quality && mat.transmittance_texture >= 0
                    && pm_transparent(pm, mat, pm_uv(pm, tri, u, v), qinterp > mat.min_quality))
                    return false;
            }
        }
        return ok;
    }
    // shadow rays only need to know whether a blocker lies below rugDistance
    __device__ bool done(float t) const { return mode == 1 && t <= rug_below; }
    __device__ float clamp(float best) const { return mode == 1 ? fminf(best, rug_below) : best; }
};
using GatherFilter = GatherFilterT<false>;

// the filter of a hemisphere (mode 0) or shadow (mode 1) ray of shooter tri; point
// materials are read at quality (hemisphere, nearest texel) or 2 x quality (shadow rays,
// bilinear when quality / 10 exceeds the material's minimum), gather.cpp:175-180, 366-370
template <int FEAT>
__device__ inline GatherFilterT<(FEAT & kFeatMats) != 0> gather_filter(const GatherArgs& A, uint32_t tri, float near_t, int mode,
                                                                       float rug_below)
{
    GatherFilterT<(FEAT & kFeatMats) != 0> f;
    f.verts = A.sc.verts; f.idx = A.sc.idx; f.side = A.sc.side;
    f.shooter = tri; f.near_t = near_t; f.mode = mode; f.rug_below = rug_below;
    if constexpr ((FEAT & kFeatMats) != 0) {
        f.pm = A.sc.pm;
        f.qpoint = mode == 0 ? A.quality : 2 * A.quality;
        f.qinterp = mode == 0 ? 0u : A.quality / 10;
    }
    return f;
}

// The shooter of subtexel j: seeded jitter -> origin o on the subtexel's triangle tri,
// and the normalised interpolated vertex normal at the subtexel centroid
// (gather.cpp:690-752).  FRAME also returns the skewed-normalised tangent and bitangent
// there (gather.cpp:731-732), which directional layers use.
template <bool FRAME = false>
__device__ inline void shooter_at(const GatherArgs& A, const lsg_subtexel* ts, uint32_t j, uint32_t key, uint32_t jitter_index,
                                  float3& o, float& nx, float& ny, float& nz, uint32_t& tri, float3* tg = nullptr,
                                  float3* bt = nullptr)
{
    const lsg_subtexel st = ts[j];
    tri = st.triangle;
    const uint32_t i0 = A.sc.idx[3 * (size_t)tri], i1 = A.sc.idx[3 * (size_t)tri + 1], i2 = A.sc.idx[3 * (size_t)tri + 2];
    const float* V = A.sc.verts;
    const float3 v0 = make_float3(V[3 * (size_t)i0], V[3 * (size_t)i0 + 1], V[3 * (size_t)i0 + 2]);
    const float3 v1 = make_float3(V[3 * (size_t)i1], V[3 * (size_t)i1 + 1], V[3 * (size_t)i1 + 2]);
    const float3 v2 = make_float3(V[3 * (size_t)i2], V[3 * (size_t)i2 + 1], V[3 * (size_t)i2 + 2]);
    // gather.cpp:727-737 basis from interpolated vertex normals at the subtexel centroid
    const float cxs = (st.uv[0] + st.uv[2] + st.uv[4]) * 0.333333333f;
    const float cys = (st.uv[1] + st.uv[3] + st.uv[5]) * 0.333333333f;
    const float w = 1 - cxs - cys;
    const float* NV = A.sc.vnormal;
    nx = NV[3 * (size_t)i0] * w + NV[3 * (size_t)i1] * cxs + NV[3 * (size_t)i2] * cys;
    ny = NV[3 * (size_t)i0 + 1] * w + NV[3 * (size_t)i1 + 1] * cxs + NV[3 * (size_t)i2 + 1] * cys;
    nz = NV[3 * (size_t)i0 + 2] * w + NV[3 * (size_t)i1 + 2] * cxs + NV[3 * (size_t)i2 + 2] * cys;
    const float nlen = sqrtf(nx * nx + ny * ny + nz * nz);
    nx = nx / nlen; ny = ny / nlen; nz = nz / nlen;
    if constexpr (FRAME) {
        const float* F[2] = {A.sc.vtangent, A.sc.vbitangent};
        float3* out[2] = {tg, bt};
        for (int k = 0; k < 2; k++) {
            const float* T = F[k];
            const float x = T[3 * (size_t)i0] * w + T[3 * (size_t)i1] * cxs + T[3 * (size_t)i2] * cys;
            const float y = T[3 * (size_t)i0 + 1] * w + T[3 * (size_t)i1 + 1] * cxs + T[3 * (size_t)i2 + 1] * cys;
            const float z = T[3 * (size_t)i0 + 2] * w + T[3 * (size_t)i1 + 2] * cxs + T[3 * (size_t)i2 + 2] * cys;
            const float len = sqrtf(x * x + y * y + z * z);
            *out[k] = make_float3(x / len, y / len, z / len);
        }
    }
    float ju, jv;
    jitter(key, jitter_index, ju, jv);
    const float ux = st.uv[0] + (st.uv[2] - st.uv[0]) * ju + (st.uv[4] - st.uv[0]) * jv;
    const float uy = st.uv[1] + (st.uv[3] - st.uv[1]) * ju + (st.uv[5] - st.uv[1]) * jv;
    o = make_float3(v0.x + (v1.x - v0.x) * ux + (v2.x - v0.x) * uy,
                    v0.y + (v1.y - v0.y) * ux + (v2.y - v0.y) * uy,
                    v0.z + (v1.z - v0.z) * ux + (v2.z - v0.z) * uy);
}

// The normal at subtexel st's centroid (gather.cpp:727-737) and the hemisphere basis built
// from it (RRMathPrivate.h:48-53), then the triangle's three vertices: kFrame floats, the
// same for every ray of the subtexel.
__device__ inline void subtexel_frame(const GatherArgs& A, const lsg_subtexel& st, float* f)
{
    const uint32_t tri = st.triangle;
    const uint32_t i0 = A.sc.idx[3 * (size_t)tri], i1 = A.sc.idx[3 * (size_t)tri + 1], i2 = A.sc.idx[3 * (size_t)tri + 2];
    const float cxs = (st.uv[0] + st.uv[2] + st.uv[4]) * 0.333333333f;
    const float cys = (st.uv[1] + st.uv[3] + st.uv[5]) * 0.333333333f;
    const float w = 1 - cxs - cys;
    const float* NV = A.sc.vnormal;
    float nx = NV[3 * (size_t)i0] * w + NV[3 * (size_t)i1] * cxs + NV[3 * (size_t)i2] * cys;
    float ny = NV[3 * (size_t)i0 + 1] * w + NV[3 * (size_t)i1 + 1] * cxs + NV[3 * (size_t)i2 + 1] * cys;
    float nz = NV[3 * (size_t)i0 + 2] * w + NV[3 * (size_t)i1 + 2] * cxs + NV[3 * (size_t)i2 + 2] * cys;
    const float nlen = sqrtf(nx * nx + ny * ny + nz * nz);
    nx = nx / nlen; ny = ny / nlen; nz = nz / nlen;
    float ox, oy, oz;
    if (fabsf(nx) > fabsf(nz)) { ox = -ny; oy = nx; oz = 0.f; } else { ox = 0.f; oy = -nz; oz = ny; }
    const float olen = sqrtf(ox * ox + oy * oy + oz * oz);
    const float tx = ox / olen, ty = oy / olen, tz = oz / olen;
    f[0] = nx; f[1] = ny; f[2] = nz;
    f[3] = tx; f[4] = ty; f[5] = tz;
    f[6] = ny * tz - nz * ty; f[7] = -nx * tz + nz * tx; f[8] = nx * ty - ny * tx;
    const float* V = A.sc.verts;
    for (int k = 0; k < 3; k++) {
        f[9 + k] = V[3 * (size_t)i0 + k]; f[12 + k] = V[3 * (size_t)i1 + k]; f[15 + k] = V[3 * (size_t)i2 + k];
    }
}

// subtexels per texel whose frames a wave keeps in LDS (later ones are computed per ray)
constexpr uint32_t kSubCache = 8;
// floats per cached subtexel: normal, tangent, bitangent, three vertices
constexpr uint32_t kFrame = 18;

// shooter_at with the subtexel's frame read from the wave's LDS cache (kSubCache frames of
// kFrame floats, frame k at cache[k * kFrame]) when j is one of the first kSubCache
// subtexels.  Origin and normal equal shooter_at's bit for bit.
__device__ inline void shooter_cached(const GatherArgs& A, const lsg_subtexel* ts, uint32_t j, uint32_t key,
                                      uint32_t jitter_index, const float* cache, float3& o, float& nx, float& ny, float& nz,
                                      uint32_t& tri)
{
    if (j >= kSubCache) {
        shooter_at(A, ts, j, key, jitter_index, o, nx, ny, nz, tri);
        return;
    }
    const lsg_subtexel st = ts[j];
    tri = st.triangle;
    const float* c = cache + kFrame * j;
    const float3 v0 = make_float3(c[9], c[10], c[11]);
    const float3 v1 = make_float3(c[12], c[13], c[14]);
    const float3 v2 = make_float3(c[15], c[16], c[17]);
    nx = c[0]; ny = c[1]; nz = c[2];
    float ju, jv;
    jitter(key, jitter_index, ju, jv);
    const float ux = st.uv[0] + (st.uv[2] - st.uv[0]) * ju + (st.uv[4] - st.uv[0]) * jv;
    const float uy = st.uv[1] + (st.uv[3] - st.uv[1]) * ju + (st.uv[5] - st.uv[1]) * jv;
    o = make_float3(v0.x + (v1.x - v0.x) * ux + (v2.x - v0.x) * uy,
                    v0.y + (v1.y - v0.y) * ux + (v2.y - v0.y) * uy,
                    v0.z + (v1.z - v0.z) * ux + (v2.z - v0.z) * uy);
}

// ---- subtexel selection (gather.cpp:645-647, 669, 680-688) -------------------------------
// The reference keeps an iterator over the texel's subtexels, starting at the first with an
// area accumulator of minus its area.  A series of n shooters steps by areaMax/(n+0.91); each
// shooter adds the step, then moves the iterator on (cyclically) and subtracts the new
// subtexel's area while the accumulator is positive.  Both carry from series to series.
struct SubWalk {
    float accu;
    uint32_t it;
    __device__ void init(const lsg_subtexel* ts) { it = 0; accu = -ts[0].area; }
    __device__ uint32_t next(const lsg_subtexel* ts, uint32_t m, float step)
    {
        accu += step;
        while (accu > 0) {
            it = it + 1 == m ? 0u : it + 1;
            accu -= ts[it].area;
        }
        return it;
    }
};
__host__ __device__ inline float walk_step(float area_max, uint32_t n) { return area_max / ((float)n + 0.91f); }

// Record of a texel's first series (k_subtexel_walk), 32 bytes: sixteen 16-bit fields,
// field 0 = 1 when the record holds the series, fields 1..15 = the shooter at which the walk
// enters subtexels 1..15 (0xffff: never in the first series).  The first series walks
// forward from subtexel 0 and does not wrap (n steps of areaMax/(n+0.91) cover less than
// the texel), so its shooter s sits in subtexel #{i >= 1 : enter_i <= s}.
struct WalkRec { uint4 a, b; };
constexpr uint32_t kWalkRecMax = 16;   // subtexels a record describes

__device__ inline uint32_t walk_field(const WalkRec& R, uint32_t i)
{
    const uint4& q = i < 8 ? R.a : R.b;
    const uint32_t k = i & 7u;
    const uint32_t w = k < 2 ? q.x : (k < 4 ? q.y : (k < 6 ? q.z : q.w));
    return (k & 1u) ? (w >> 16) : (w & 0xffffu);
}
__device__ inline bool walk_valid(const WalkRec& R) { return (R.a.x & 0xffffu) == 1u; }
// shooter s of the first series from the record in memory (fields read one by one, so the
// 32-byte record is not held in registers)
__device__ inline uint32_t walk_pick(const WalkRec* R, uint32_t m, uint32_t s)
{
    const uint16_t* f = (const uint16_t*)R;
    uint32_t j = 0;
    for (uint32_t i = 1; i < m; i++) j += f[i] <= s ? 1u : 0u;
    return j;
}
__device__ inline bool walk_has(const WalkRec* R) { return ((const uint16_t*)R)[0] == 1u; }

// The subtexel of shooter g0 + k of a texel (k < cnt), by stepping the walk from the
// texel's first shooter: the first series has n1 shooters, every later one nrest (after
// the first series only light rounds are shot again, gather.cpp:652-661, and every such
// series has the same size).  The loop is the same for every lane of the group; lane k
// keeps shooter g0 + k.  Used where no record applies: later series and texels of more
// than kWalkRecMax subtexels.
__device__ inline uint32_t walk_replay(const lsg_subtexel* ts, uint32_t m, float area_max, uint32_t n1, uint32_t nrest,
                                       uint32_t g0, uint32_t cnt, uint32_t k)
{
    SubWalk w;
    w.init(ts);
    const float s1 = walk_step(area_max, n1), s2 = walk_step(area_max, nrest);
    uint32_t g = 0, mine = 0;
    for (; g < g0; g++) w.next(ts, m, g < n1 ? s1 : s2);
    for (uint32_t i = 0; i < cnt; i++, g++) {
        const uint32_t j = w.next(ts, m, g < n1 ? s1 : s2);
        if (i == k) mine = j;
    }
    return mine;
}

// One lane per texel: the reference's walk over the first series of n1 shooters.  Texels
// of more than kWalkRecMax subtexels, series beyond 65535 shooters, or a walk that wraps
// (float rounding at very high n) get no record and replay the walk in the gather kernel.
__global__ __launch_bounds__(256) void k_subtexel_walk(const lsg_texel* __restrict__ texels, uint32_t ntex,
                                                       const lsg_subtexel* __restrict__ subs, uint32_t n1,
                                                       WalkRec* __restrict__ W)
{
    const uint32_t ti = blockIdx.x * 256 + threadIdx.x;
    if (ti >= ntex) return;
    const lsg_texel T = texels[ti];
    const uint32_t m = T.sub_count_flags & 0xffffu;
    WalkRec out = {make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u)};
    if (m >= 2 && m <= kWalkRecMax && n1 >= 1 && n1 <= 65535u) {
        const lsg_subtexel* ts = subs + T.sub_begin;
        float area_max = 0.f;
        for (uint32_t j = 0; j < m; j++) area_max += ts[j].area;
        const float step = walk_step(area_max, n1);
        uint32_t f[kWalkRecMax];
        for (uint32_t i = 0; i < kWalkRecMax; i++) f[i] = 0xffffu;
        f[0] = 1u;
        SubWalk w;
        w.init(ts);
        uint32_t cur = 0;
        bool ok = true;
        for (uint32_t s = 0; s < n1; s++) {
            const uint32_t j = w.next(ts, m, step);
            if (j < cur) { ok = false; break; }
            while (cur < j) f[++cur] = s;
        }
        if (ok) {
            out.a = make_uint4(f[0] | (f[1] << 16), f[2] | (f[3] << 16), f[4] | (f[5] << 16), f[6] | (f[7] << 16));
            out.b = make_uint4(f[8] | (f[9] << 16), f[10] | (f[11] << 16), f[12] | (f[13] << 16), f[14] | (f[15] << 16));
        }
    }
    W[ti] = out;
}

// Exitance a hemisphere ray reads at its hit: the flat value of the side it sees (direct
// light times albedo plus emittance, pathtracer.cpp:235-262, 306-316) plus, on sides that
// take it, albedo * (I0 (1-u-v) + I1 u + I2 v) of the triangle's corner indirect irradiance
// at the hit's hitPoint2d (pathtracer.cpp:326-348, Triangle::getPointMeasure
// rrcore.cpp:262-273).  (u, v) is recomputed by the hit's own triangle test on the
// recorded distance, bit-identical to the traversal's.
template <bool MATS = false>
__device__ inline float4 hit_exitance(const SceneView& S, int tri, bool front, const float3 o, const float3 d, float t,
                                      uint32_t quality = 0)
{
    float4 e = S.exitance[2 * (size_t)tri + (front ? 0 : 1)];
    const bool ind = S.ind && (S.side[tri] & (front ? LSG_SIDE_FRONT_INDIRECT : LSG_SIDE_BACK_INDIRECT));
    const lsg_point_material* mat = nullptr;
    if constexpr (MATS) {
        const uint32_t mi = S.pm.tri_mat[tri];
        if (mi != 0xffffffffu) mat = S.pm.mats + mi;
    }
    if (!ind && !mat) return e;
    const bool point = mat && quality > mat->min_quality;
    float u = 0.f, v = 0.f;
    if (ind || point) {
        const TriRest* tr = S.col.rest + S.slot[tri];
        srlnp_test(tr->sv, tr->ev, (int)tr->meta.y, o, d, t, u, v);
    }
    const float w = 1 - u - v;
    float ix = 0.f, iy = 0.f, iz = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
    if (ind) {
        const float4 c0 = S.ind[3 * (size_t)tri], c1 = S.ind[3 * (size_t)tri + 1], c2 = S.ind[3 * (size_t)tri + 2];
        ix = c0.x * w + c1.x * u + c2.x * v; iy = c0.y * w + c1.y * u + c2.y * v; iz = c0.z * w + c1.z * u + c2.z * v;
        ax = c0.w; ay = c1.w; az = c2.w;
    }
    if (!mat) {
        e.x = e.x + ix * ax; e.y = e.y + iy * ay; e.z = e.z + iz * az;
        return e;
    }
    // a triangle with a material: its table entry holds the side's direct irradiance, and
    // the hit composes the exitance as getIncidentRadiance does (pathtracer.cpp:235-262,
    // 306-316, 326-348): direct x diffuse colour, + emittance on a side that emits, +
    // indirect x diffuse colour; the colours are the texel's where the quality reads
    // point materials, else the triangle's
    float2 muv = make_float2(0.f, 0.f);
    if (point) muv = pm_uv(S.pm, (uint32_t)tri, u, v);
    float rx = mat->albedo[0], ry = mat->albedo[1], rz = mat->albedo[2];
    if (point && mat->diffuse_texture >= 0) {
        const float4 c = pm_diffuse(S.pm, *mat, muv);
        rx = c.x; ry = c.y; rz = c.z;
    }
    e.x = e.x * rx; e.y = e.y * ry; e.z = e.z * rz;
    if (mat->emit_sides & (front ? 1u : 2u)) {
        if (point && mat->emissive_texture >= 0) {
            const float4 c = pm_emissive(S.pm, *mat, muv);
            const float k = mat->emittance_multiplier;
            e.x = e.x + c.x * k; e.y = e.y + c.y * k; e.z = e.z + c.z * k;
        } else {
            e.x = e.x + mat->emittance[0]; e.y = e.y + mat->emittance[1]; e.z = e.z + mat->emittance[2];
        }
    }
    if (ind) { e.x = e.x + ix * rx; e.y = e.y + iy * ry; e.z = e.z + iz * rz; }
    return e;
}

// asin in binary32 by Abramowitz & Stegun 4.4.46 (|error| <= 2e-8 in exact arithmetic):
// asin(x) = pi/2 - sqrt(1-x) * (a0 + a1 x + ... + a7 x^7) for 0 <= x <= 1, odd.  Plain
// float multiplies and adds in a fixed order, so the GPU and the CPU statement round
// alike.
__device__ inline float asin_as(float x)
{
    const float ax = fminf(fabsf(x), 1.0f);      // |x| may round past 1 near the poles
    float p = -0.0012624911f;
    p = p * ax + 0.0066700901f;
    p = p * ax - 0.0170881256f;
    p = p * ax + 0.0308918810f;
    p = p * ax - 0.0501743046f;
    p = p * ax + 0.0889789874f;
    p = p * ax - 0.2145988016f;
    p = p * ax + 1.5707963050f;
    const float r = 1.57079632679489662f - sqrtf(1.0f - ax) * p;
    return x < 0 ? -r : r;
}

// Nearest environment texel in direction d, RRBufferInMemory::getElementAtDirection
// (RRBufferInMemory.cpp:410-455).  Cube maps follow the OpenGL face convention: major
// axis m picks the face (+x,-x,+y,-y,+z,-z; ties go to x, then y), and the face
// coordinates are (sc/|m|+1)/2, (tc/|m|+1)/2 with
//   +x: sc=-z tc=-y   -x: sc=+z tc=-y   +y: sc=+x tc=+z
//   -y: sc=+x tc=-z   +z: sc=+x tc=-y   -z: sc=-x tc=-y.
// Panoramas map the yaw of (x, z) to columns and asin(y) to rows, in float as the
// reference, with asin_as for asin.
__device__ inline float4 env_texel(const float4* env, uint32_t type, uint32_t w, uint32_t h, const float3& d)
{
    uint32_t idx;
    if (type == LSG_ENV_CUBE) {
        const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
        const uint32_t axis = (ax >= ay && ax >= az) ? 0u : ((ay >= ax && ay >= az) ? 1u : 2u);
        const float m = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
        const uint32_t face = 2 * axis + (m < 0 ? 1u : 0u);
        const float am = fabsf(m);
        float sc, tc;
        switch (face) {
            case 0: sc = -d.z; tc = -d.y; break;
            case 1: sc = d.z; tc = -d.y; break;
            case 2: sc = d.x; tc = d.z; break;
            case 3: sc = d.x; tc = -d.z; break;
            case 4: sc = d.x; tc = -d.y; break;
            default: sc = -d.x; tc = -d.y; break;
        }
        int c0 = (int)((sc / am + 1) * (0.5f * (float)w));
        int c1 = (int)((tc / am + 1) * (0.5f * (float)w));
        c0 = c0 < 0 ? 0 : (c0 > (int)w - 1 ? (int)w - 1 : c0);
        c1 = c1 < 0 ? 0 : (c1 > (int)w - 1 ? (int)w - 1 : c1);
        idx = (uint32_t)c0 + (uint32_t)c1 * w + face * w * w;
    } else {
        const float pi = 3.14159265358979323846f;
        const float d2 = d.x * d.x + d.z * d.z;
        uint32_t c0 = 0;
        if (d2 != 0.f) {
            const float s = d.x / sqrtf(d2);
            float angle = asin_as(s);
            if (d.z < 0) angle = pi - angle;
            c0 = (uint32_t)fmaxf(0.f, (angle * (-0.5f / pi) + 0.75f) * (float)w) % w;   // >= 0 for the cast
        }
        const float len = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
        const float el = asin_as(d.y / len);
        const uint32_t c1 = (uint32_t)fmaxf(0.f, (el * (1.0f / pi) + 0.5f) * (float)h) % h;
        idx = c0 + c1 * w;
    }
    return env[idx];
}

// RRMatrix3x4::transformDirection (RRMath.cpp:212-219) by a row-major 3x3
__device__ inline float3 rotate3(const float* m, const float3& a)
{
    return make_float3(a.x * m[0] + a.y * m[1] + a.z * m[2], a.x * m[3] + a.y * m[4] + a.z * m[5],
                       a.x * m[6] + a.y * m[7] + a.z * m[8]);
}

// What a ray that leaves the scene in direction d sees (before the multiplier): the
// environment's texel, or for a blend (RRBufferBlend::getElementAtDirection,
// RRBufferBlend.cpp:49-66) color0 * (1 - blend) + color1 * blend of the two environments
// looked up in their rotated directions, an absent one giving 0
__device__ inline float4 env_at(const SceneView& S, const float3& d)
{
    if (S.env_type != kEnvBlend) return env_texel(S.env, S.env_type, S.env_w, S.env_h, d);
    float4 c0 = make_float4(0.f, 0.f, 0.f, 0.f), c1 = c0;
    if (S.env0_type) c0 = env_texel(S.env, S.env0_type, S.env_w, S.env_h, rotate3(S.env_rot, d));
    if (S.env1_type) c1 = env_texel(S.env1, S.env1_type, S.env1_w, S.env1_h, rotate3(S.env_rot + 9, d));
    const float k0 = 1 - S.env_blend, k1 = S.env_blend;
    return make_float4(c0.x * k0 + c1.x * k1, c0.y * k0 + c1.y * k1, c0.z * k0 + c1.z * k1, c0.w * k0 + c1.w * k1);
}

// g_lightmapDirections[i] (gather.cpp:43-49) in the skewed frame (t, b, n), normalised
// (gather.cpp:292-293, 458-459)
__device__ inline float3 layer_dir(int i, const float3& t, const float3& b, const float3& n)
{
    float g0 = 0.f, g1 = 0.f, g2 = 1.f;
    if (i == 1) { g1 = sqrtf(6.0f) / 3; g2 = 1 / sqrtf(3.0f); }
    else if (i >= 2) { g0 = (i == 2 ? -1 : 1) / sqrtf(2.0f); g1 = -1 / sqrtf(6.0f); g2 = 1 / sqrtf(3.0f); }
    const float x = t.x * g0 + b.x * g1 + n.x * g2;
    const float y = t.y * g0 + b.y * g1 + n.y * g2;
    const float z = t.z * g0 + b.z * g1 + n.z * g2;
    const float len = sqrtf(x * x + y * y + z * z);
    return make_float3(x / len, y / len, z / len);
}

// processTexel for texel ti by an aligned group of G lanes of one wavefront (G = 64:
// the whole wave; 16 or 32 for low ray counts, several texels per wave); lane = the
// lane in the wave; mystk = this lane's first LDS stack slot, slots `stride` ints
// apart; lq = this lane's slot of the group's light queue (2 slots, `stride` apart).
// Every group runs its own texel: control flow is uniform within a group.
//
// Light shots: the shooters that also cast shadow rays (about one in ten at quality
// 1024, gather.cpp:763,779) are queued in LDS as their chunks pass and traced G at a
// time, so shadow rays run with full lane masks instead of a few lanes alongside each
// chunk's hemisphere rays.  A shooter's shadow rays see the same origin, normal and
// lights as before; only which lane adds them to the sums changes.
//
// LAYERS: every output layer of updateLightmap (lightmap.cpp:806-820) -- the lightmap,
// with A.dirs the three directional lightmaps weighted per layer direction
// (gather.cpp:284-305 hemisphere, :453-481 lights), and the bent normal
// (gather.cpp:306, :482, :809-830), into A.layer[0..4].
//
// FEAT: kFeatEnv = hemisphere rays that leave the scene sample the environment;
// kFeatMats = point materials (textures and keyed transparency read at hits).  Each is a
// separate instantiation, so that scenes without them keep the smaller register footprint.
// direction-binned hemisphere order (gather_texel): series of at most kPermMaxRays rays,
// whose candidates fit kPermWords words of 64 (1024 rays need about 1700 candidates, and any
// 2048 consecutive filler indices hold more than 1100 points of the disk); per wave
// kPermScr words of LDS: the words' acceptance masks, the columns' bits, the key starts and
// the words' prefix counts
constexpr uint32_t kPermWords = 32, kPermMaxRays = 1024;
constexpr uint32_t kPermScr = 2 * kPermWords + 64 + (64 + kPermWords) / 2;

// The acceptance pass of a direction-binned series of n hemisphere rays whose candidates
// start at f0: word w holds the acceptance of candidates f0+64w .. f0+64w+63 (only the first
// n accepted count), column c the acceptance of candidates f0+64w+c over w.  Columns are
// ordered by the key of their candidates' index mod 64 (base-4 digits, coarsest first: the
// filler's coarse cells first), and the key starts give each key's first position in the
// binned order.  Returns false (nothing usable) when the candidates do not fit kPermWords
// words; else sets filler past the series as the sequential hand-out leaves it.
__device__ inline bool perm_prepare(uint32_t* pscr, uint32_t f0, uint32_t n, uint32_t lane, uint32_t& filler)
{
    uint64_t* const pmask = (uint64_t*)pscr;
    uint32_t* const pcol = pscr + 2 * kPermWords;
    uint16_t* const pstart = (uint16_t*)(pscr + 2 * kPermWords + 64);
    uint16_t* const pprefix = pstart + 64;
    uint32_t acc = 0, W = 0, colbits = 0, lastw_full = 0, lastbit = 0;
    __builtin_amdgcn_wave_barrier();
    while (acc < n && W < kPermWords) {
        float cx, cy;
        filler_point(f0 + 64u * W + lane, cx, cy);
        const bool ok = !(cx * cx + cy * cy >= 1.f);
        uint64_t am = __ballot(ok);
        const uint32_t na = (uint32_t)__popcll(am);
        lastw_full = acc + na == n;
        if (acc + na > n) {
            lastbit = nth_set_bit(am, n - acc - 1u);
            am &= lastbit == 63u ? ~0ull : ((1ull << (lastbit + 1u)) - 1ull);
        }
        if (lane == 0) { pmask[W] = am; pprefix[W] = (uint16_t)acc; }
        colbits |= (uint32_t)((am >> lane) & 1ull) << W;
        acc += (uint32_t)__popcll(am);
        W++;
    }
    if (acc != n) return false;
    filler = lastw_full ? f0 + 64u * W : f0 + 64u * (W - 1u) + lastbit + 1u;
    const uint32_t c = (f0 + lane) & 63u;
    const uint32_t key = ((c & 3u) << 4) | (((c >> 2) & 3u) << 2) | ((c >> 4) & 3u);
    pcol[lane] = colbits;
    pstart[key] = (uint16_t)__popc(colbits);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t kc = pstart[lane];                           // count of key = lane
    uint32_t incl = kc;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, o);
        if (lane >= o) incl += v;
    }
    __builtin_amdgcn_wave_barrier();
    pstart[lane] = (uint16_t)(incl - kc);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return true;
}

// position P (< n) of the binned order: the shooter (its rank among the accepted
// candidates, as the sequential hand-out ranks them) and the candidate's filler index
__device__ inline void perm_pick(const uint32_t* pscr, uint32_t f0, uint32_t P, uint32_t& shooter, uint32_t& cand)
{
    const uint64_t* const pmask = (const uint64_t*)pscr;
    const uint32_t* const pcol = pscr + 2 * kPermWords;
    const uint16_t* const pstart = (const uint16_t*)(pscr + 2 * kPermWords + 64);
    const uint16_t* const pprefix = pstart + 64;
    uint32_t k = 0;
    for (uint32_t step = 32; step; step >>= 1)
        if (pstart[k + step] <= P) k += step;
    const uint32_t t = P - pstart[k];
    const uint32_t cm = (k >> 4) | (((k >> 2) & 3u) << 2) | ((k & 3u) << 4);
    const uint32_t col = (cm - f0) & 63u;
    const uint32_t w = nth_set_bit((uint64_t)pcol[col], t);
    shooter = pprefix[w] + (uint32_t)__popcll(pmask[w] & lanemask_lt(col));
    cand = f0 + 64u * w + col;
}

template <int STACK, bool STATS, int G = 64, bool LAYERS = false, int FEAT = 0>
__device__ void gather_texel(const GatherArgs& A, uint32_t ti, uint32_t wlane, int32_t* mystk, uint32_t* lq, uint16_t* lqj,
                             int stride, float* subc = nullptr, uint32_t* pscr = nullptr)
{
    // one texel per wave: the frames of its first subtexels are kept in the wave's LDS
    // (when the caller gives one)
    const bool kCache = G == 64 && !LAYERS && subc != nullptr;
    static_assert(G == 16 || G == 32 || G == 64, "group width");
    const uint32_t lane = wlane & (uint32_t)(G - 1);               // lane in the group
    const uint32_t gbase = wlane & ~(uint32_t)(G - 1);             // group's first lane in the wave
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const lsg_texel T = A.texels[ti];
    const uint32_t m = T.sub_count_flags & 0xffffu, flags = (T.sub_count_flags >> 16) & 0xffu;
    const uint32_t hemi_rays = A.hemi_rays, rounds = A.rounds, nl = A.nlights;
    const uint32_t light_rays = nl * rounds;
    if (!m || (!hemi_rays && !light_rays)) {
        if (lane == 0) {
            if constexpr (LAYERS) {
                for (int i = 0; i < 5; i++)
                    if (A.layer[i]) A.layer[i][T.out_index] = make_float4(0.f, 0.f, 0.f, 0.f);
                return;
            }
            A.out[T.out_index] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (A.parts) { A.parts[2 * (size_t)ti] = make_float4(0.f, 0.f, 0.f, 0.f); A.parts[2 * (size_t)ti + 1] = make_float4(0.f, 0.f, 0.f, 0.f); }
        }
        return;
    }
    const lsg_subtexel* ts = A.subs + T.sub_begin;
    const uint32_t key = lowbias32(A.seed ^ lowbias32(T.key));
    uint32_t filler = key & 0xffffu;           // wave-uniform
    float area_max = 0.f;
    for (uint32_t j = 0; j < m; j++) area_max += ts[j].area;
    const float near_t = 1000 * A.tmin;
    const float rug_below = __int_as_float(__float_as_int(A.rug) - 1);   // largest float < rugDistance (> 0)
    const uint32_t n1 = first_series(A);
    if (kCache) {
        __builtin_amdgcn_wave_barrier();
        if (lane < min(m, kSubCache)) subtexel_frame(A, ts[lane], subc + kFrame * lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    float hr = 0.f, hg = 0.f, hb = 0.f, lr = 0.f, lg = 0.f, lb = 0.f;
    // LAYERS: hemisphere and light sums of the four lightmap layers, and the bent normal
    constexpr int NL = LAYERS ? 12 : 1;
    float hl[NL], ll[NL], bn[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < NL; k++) { hl[k] = 0.f; ll[k] = 0.f; }
    const bool dirs = LAYERS && A.dirs;
    uint32_t stH[3] = {0, 0, 0}, stS[3] = {0, 0, 0};
    uint32_t relH = 0, relL = 0, unrelL = 0, shotRounds = 0;        // wave totals (uniform)
    uint32_t shooter_base = 0;
    for (;;) {
        const bool shootH = (relH < hemi_rays || relH * 10 < hemi_rays) && hemi_rays > 0;
        const bool shootL = (relL + unrelL < light_rays || relL * 10 < light_rays) && unrelL < light_rays * 100;
        if (!shootH && !shootL) break;
        const uint32_t nH = shootH ? hemi_rays : 0, nL = shootL ? rounds : 0;
        const uint32_t n = nH > nL ? nH : nL;
        const bool allH = !shootL || hemi_rays >= rounds;
        const bool allL = !shootH || rounds >= hemi_rays;
        uint32_t cRelH = 0, cRelL = 0, cUnrelL = 0, cRounds = 0;      // this lane's counts in the series
        uint32_t* const lq0 = lq - lane;                                // the group's queue
        uint16_t* const lqj0 = lqj - lane;                              // and the queued shooters' subtexels
        uint32_t qhead = 0, qtail = 0;                                  // group-uniform
        // the shadow rays of queued light shooters qhead .. qhead+cnt-1, one per lane
        auto light_batch = [&](uint32_t cnt) {
            __builtin_amdgcn_wave_barrier();
            if (lane < cnt) {
                const uint32_t e = qhead + lane;
                const uint32_t qi = (e & (uint32_t)(G - 1)) + ((e / (uint32_t)G) & 1u) * stride;
                const uint32_t s = lq0[qi], sj = lqj0[qi];
                float3 o; float nx, ny, nz; uint32_t tri;
                float3 tg, bt;
                if (kCache) shooter_cached(A, ts, sj, key, shooter_base + s, subc, o, nx, ny, nz, tri);
                else shooter_at<LAYERS>(A, ts, sj, key, shooter_base + s, o, nx, ny, nz, tri, &tg, &bt);
                float3 ld[LAYERS ? 4 : 1];
                if (dirs) {
#pragma unroll
                    for (int i = 0; i < (LAYERS ? 4 : 1); i++) ld[i] = layer_dir(i, tg, bt, make_float3(nx, ny, nz));
                }
                auto flt = gather_filter<FEAT>(A, tri, near_t, 1, rug_below);
                for (uint32_t li = 0; li < nl; li++) {
                    const lsg_light L = A.lights[li];
                    float dx, dy, dz;
                    if (L.type == 1) { dx = -L.direction[0]; dy = -L.direction[1]; dz = -L.direction[2]; }
                    else { dx = L.position[0] - o.x; dy = L.position[1] - o.y; dz = L.position[2] - o.z; }
                    float dsz = sqrtf(dx * dx + dy * dy + dz * dz);
                    dx = dx / dsz; dy = dy / dsz; dz = dz / dsz;
                    if (L.type == 1) dsz = dsz * A.locality;
                    const float ni = dx * nx + dy * ny + dz * nz;            // gather.cpp:429
                    if (ni <= 0 || !isfinite(ni)) { cRelL++; continue; }
                    bool blocked = false; float bt = 0.f;
                    if (L.cast_shadows) {
                        TraceResult r;
                        uint32_t nv_ = 0, nt_ = 0;
                        trace_closest<STACK, STATS, decltype(flt), false>(flt, A.sc.col, o, make_float3(dx, dy, dz), A.tmin, dsz, mystk, stride, r, nv_, nt_);
                        if (STATS) { stS[0]++; stS[1] += nv_; stS[2] += nt_; }
                        blocked = r.tri >= 0; bt = r.t;
                    }
                    if (!blocked) {
                        // RRLight::getIrradiance (RRLight.cpp:345-362), REALISTIC attenuation
                        float att = 1.f;
                        if (L.type == 0) {
                            const float ex = o.x - L.position[0], ey = o.y - L.position[1], ez = o.z - L.position[2];
                            att = 1 / (0.000000000137289f + (ex * ex + ey * ey + ez * ez));
                        }
                        if constexpr (LAYERS) {
                            const float ka = att * A.light_mult;
                            if (dirs) {
#pragma unroll
                                for (int i = 0; i < 4; i++) {
                                    const float n2 = dx * ld[i].x + dy * ld[i].y + dz * ld[i].z;   // gather.cpp:460
                                    if (n2 > 0) {
                                        const float k = ka * n2;
                                        ll[3 * i] += L.color[0] * k; ll[3 * i + 1] += L.color[1] * k; ll[3 * i + 2] += L.color[2] * k;
                                    }
                                }
                            } else {
                                const float kl = ka * ni;      // the lightmap-only sums, as below
                                ll[0] += L.color[0] * kl; ll[1] += L.color[1] * kl; ll[2] += L.color[2] * kl;
                            }
                            const float w = (fabsf(L.color[0] * ka) + fabsf(L.color[1] * ka) + fabsf(L.color[2] * ka))
                                          * 0.33333333333333f * ni;   // gather.cpp:482
                            bn[0] += dx * w; bn[1] += dy * w; bn[2] += dz * w;
                        } else {
                            const float kl = att * A.light_mult * ni;
                            lr += L.color[0] * kl; lg += L.color[1] * kl; lb += L.color[2] * kl;
                        }
                    }
                    // a blocker nearer than rugDistance makes the ray unreliable, any other
                    // outcome is reliable (gather.cpp:483-511); counted arithmetically so
                    // both counters stay in registers
                    const uint32_t unrel = (blocked && bt < A.rug) ? 1u : 0u;
                    cUnrelL += unrel;
                    cRelL += 1u - unrel;
                }
                cRounds++;
            }
            __builtin_amdgcn_wave_barrier();
        };
        // one hemisphere ray: shooter s of the series (origin in subtexel myj, jitter index
        // shooter_base + s) in the direction of filler point (x, y)
        auto shoot_h = [&](uint32_t s, uint32_t myj, float x, float y) {
            float3 o; float nx, ny, nz; uint32_t tri;
            float3 tg, bt;
            float tx, ty, tz, bx, by, bz;
            if (kCache) shooter_cached(A, ts, myj, key, shooter_base + s, subc, o, nx, ny, nz, tri);
            else shooter_at<LAYERS>(A, ts, myj, key, shooter_base + s, o, nx, ny, nz, tri, &tg, &bt);
            const uint32_t sj = kCache ? myj : kSubCache;
            if (kCache && sj < kSubCache) {
                const float* f = subc + kFrame * sj;
                tx = f[3]; ty = f[4]; tz = f[5]; bx = f[6]; by = f[7]; bz = f[8];
            } else {
                float ox, oy, oz;                                           // RRMathPrivate.h:48-53
                if (fabsf(nx) > fabsf(nz)) { ox = -ny; oy = nx; oz = 0.f; } else { ox = 0.f; oy = -nz; oz = ny; }
                const float olen = sqrtf(ox * ox + oy * oy + oz * oz);
                tx = ox / olen; ty = oy / olen; tz = oz / olen;
                bx = ny * tz - nz * ty; by = -nx * tz + nz * tx; bz = nx * ty - ny * tx;
            }
            auto flt = gather_filter<FEAT>(A, tri, near_t, 0, rug_below);
            const float dist = x * x + y * y;
            const float cosa = sqrtf(1 - dist);                          // gather.cpp:125-126
            const float3 d = make_float3(nx * cosa + tx * x + bx * y, ny * cosa + ty * x + by * y, nz * cosa + tz * x + bz * y);
            TraceResult r;
            uint32_t nv_ = 0, nt_ = 0;
            trace_closest<STACK, STATS>(flt, A.sc.col, o, d, A.tmin, A.locality, mystk, stride, r, nv_, nt_);
            if (STATS) { stH[0]++; stH[1] += nv_; stH[2] += nt_; }
            // a hit reads the exitance of the side it sees; a ray that leaves the scene
            // returns environment * multiplier, or nothing (pathtracer.cpp:156-173)
            auto add = [&](const float4& e) {
                if constexpr (LAYERS) {
                    if (dirs) {
                        // gather.cpp:284-305: dirs near the normal are picked more often,
                        // so each layer weighs by its own cosine over the normal's
                        const float ni1 = d.x * nx + d.y * ny + d.z * nz;
                        if (ni1 > 0) {
                            const float inv = 1 / ni1;
#pragma unroll
                            for (int i = 0; i < 4; i++) {
                                const float3 l = layer_dir(i, tg, bt, make_float3(nx, ny, nz));
                                const float n2 = d.x * l.x + d.y * l.y + d.z * l.z;
                                if (n2 > 0) {
                                    const float k = n2 * inv;
                                    hl[3 * i] += e.x * k; hl[3 * i + 1] += e.y * k; hl[3 * i + 2] += e.z * k;
                                }
                            }
                        }
                    } else { hl[0] += e.x; hl[1] += e.y; hl[2] += e.z; }
                    const float w = (fabsf(e.x) + fabsf(e.y) + fabsf(e.z)) * 0.33333333333333f;   // gather.cpp:306
                    bn[0] += d.x * w; bn[1] += d.y * w; bn[2] += d.z * w;
                } else {
                    hr += e.x; hg += e.y; hb += e.z;
                }
            };
            if (r.tri >= 0) add(hit_exitance<(FEAT & kFeatMats) != 0>(A.sc, r.tri, r.front, o, d, r.t, A.quality));
            else if constexpr ((FEAT & kFeatEnv) != 0) {
                const float4 v = env_at(A.sc, d);
                add(make_float4(v.x * A.env_mult, v.y * A.env_mult, v.z * A.env_mult, 0.f));
            }
            cRelH++;
        };
        // direction-binned order: the series' hemisphere rays traced 64 at a time in the
        // order of their filler cell (coarse base-4 digits of the candidate index first), so a
        // wave's lanes trace neighbouring directions; every shooter keeps the filler point the
        // sequential hand-out gives it, only the order of the sums changes
        const bool permH = pscr != nullptr && G == 64 && shootH && allH && shooter_base == 0 && n == hemi_rays && n > 64u
                           && n <= kPermMaxRays && (m == 1 || (A.walk && walk_has(A.walk + ti)));
        bool permDone = false;
        if (permH) {
            const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)filler);   // wave-uniform
            if (perm_prepare(pscr, f0, n, lane, filler)) {
                permDone = true;
                for (uint32_t base = 0; base < n; base += 64u) {
                    const uint32_t P = base + lane;
                    if (P < n) {
                        uint32_t sh, cand;
                        perm_pick(pscr, f0, P, sh, cand);
                        float x, y;
                        filler_point(cand, x, y);
                        const uint32_t myj = m > 1 ? walk_pick(A.walk + ti, m, sh) : 0u;
                        shoot_h(sh, myj, x, y);
                    }
                }
                // the series' light shooters, in order
                for (uint32_t base = 0; base < n; base += 64u) {
                    const uint32_t sq = base + lane;
                    const bool doL = sq < n && shootL && (allL || shots_before(sq + 1, n, rounds) > shots_before(sq, n, rounds));
                    const uint64_t lm = __ballot(doL);
                    if (doL) {
                        const uint32_t e = qtail + (uint32_t)__popcll(lm & lanemask_lt(lane));
                        const uint32_t qi = (e & 63u) + ((e / 64u) & 1u) * stride;
                        lq0[qi] = sq;
                        lqj0[qi] = (uint16_t)(m > 1 ? walk_pick(A.walk + ti, m, sq) : 0u);
                    }
                    qtail += (uint32_t)__popcll(lm);
                    while (qtail - qhead >= 64u) { light_batch(64u); qhead += 64u; }
                }
            }
        }
        if (!permDone) for (uint32_t base = 0; base < n; base += G) {
            const uint32_t s = base + lane;
            const bool active = s < n;
            bool doH = false, doL = false;
            if (active && shootH) doH = allH || shots_before(s + 1, n, hemi_rays) > shots_before(s, n, hemi_rays);
            if (active && shootL) doL = allL || shots_before(s + 1, n, rounds) > shots_before(s, n, rounds);
            // hand out filler points: hemisphere shooters of this chunk take the next
            // accepted filler indices in lane order (GetCirclePoint, gather.cpp:85-90)
            const uint64_t hmask = (__ballot(doH) >> gbase) & gmask;
            const uint32_t need = (uint32_t)__popcll(hmask);
            const uint32_t myrank = (uint32_t)__popcll(hmask & lanemask_lt(lane));
            float my_x = 0.f, my_y = 0.f;   // the lane's accepted filler point (hand-out below)
            uint32_t got = 0;
            while (got < need) {
                const uint32_t cand = filler + lane;
                float cx, cy;
                filler_point(cand, cx, cy);
                const bool ok = !(cx * cx + cy * cy >= 1.f);
                const uint64_t am = (__ballot(ok) >> gbase) & gmask;
                const uint32_t na = (uint32_t)__popcll(am);
                const uint32_t want = doH ? myrank - got : 64u;
                const uint32_t src = (doH && myrank >= got && want < na) ? nth_set_bit(am, want) : lane;
                // the accepted point itself travels with it (no second filler_point call)
                const float vx = __shfl(cx, (int)(gbase + src)), vy = __shfl(cy, (int)(gbase + src));
                if (doH && myrank >= got && want < na) { my_x = vx; my_y = vy; }
                if (got + na <= need) { got += na; filler += (uint32_t)G; }
                else {
                    const uint32_t last = nth_set_bit(am, need - got - 1u);
                    filler = (uint32_t)__shfl((int)cand, (int)(gbase + last)) + 1u;
                    got = need;
                }
            }
            // the chunk's subtexels (every shooter steps the walk, shooting or not): from
            // the first series' record, else by replaying the walk
            uint32_t myj = 0;
            if (m > 1) {
                if (A.walk && shooter_base == 0 && walk_has(A.walk + ti)) myj = walk_pick(A.walk + ti, m, s);
                else myj = walk_replay(ts, m, area_max, n1, n, shooter_base + base, min((uint32_t)G, n - base), lane);
            }
            {
                const uint64_t lm = (__ballot(doL) >> gbase) & gmask;
                if (doL) {
                    const uint32_t e = qtail + (uint32_t)__popcll(lm & lanemask_lt(lane));
                    const uint32_t qi = (e & (uint32_t)(G - 1)) + ((e / (uint32_t)G) & 1u) * stride;
                    lq0[qi] = s;
                    lqj0[qi] = (uint16_t)myj;
                }
                qtail += (uint32_t)__popcll(lm);
            }
            if (doH) shoot_h(s, myj, my_x, my_y);
            while (qtail - qhead >= (uint32_t)G) { light_batch((uint32_t)G); qhead += (uint32_t)G; }
        }
        while (qtail != qhead) {
            const uint32_t cnt = min((uint32_t)G, qtail - qhead);
            light_batch(cnt);
            qhead += cnt;
        }
        relH += group_sum_u<G>(cRelH); relL += group_sum_u<G>(cRelL); unrelL += group_sum_u<G>(cUnrelL);
        shotRounds += group_sum_u<G>(cRounds);
        shooter_base += n;
    }
    if (STATS) {
        for (int k = 0; k < 3; k++) {
            atomicAdd(&A.stats[k], (unsigned long long)stH[k]);
            atomicAdd(&A.stats[3 + k], (unsigned long long)stS[k]);
        }
    }
    if constexpr (LAYERS) {
#pragma unroll
        for (int k = 0; k < NL; k++) { hl[k] = group_sum<G>(hl[k]); ll[k] = group_sum<G>(ll[k]); }
#pragma unroll
        for (int k = 0; k < 3; k++) bn[k] = group_sum<G>(bn[k]);
        if (lane == 0) {
            const bool useH = hemi_rays && relH, useL = light_rays && relL;
            const float f = useH ? 1 / (float)relH : 0.f;
            const float alpha = (useH || useL) ? flags / 255.f : 0.f;   // gather.cpp:795-803
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (!A.layer[i]) continue;
                float c[3] = {0.f, 0.f, 0.f};
                for (int k = 0; k < 3; k++) {
                    if (useH) c[k] += hl[3 * i + k] * f;
                    if (useL) c[k] += ll[3 * i + k] / (float)shotRounds;
                }
                A.layer[i][T.out_index] = make_float4(c[0], c[1], c[2], alpha);
            }
            if (A.layer[4]) {
                // gather.cpp:809-830: normalised sum, stored as (n+1)/2
                float x = bn[0], y = bn[1], z = bn[2];
                if (x != 0.f || y != 0.f || z != 0.f) {
                    const float len = sqrtf(x * x + y * y + z * z);
                    x = x / len; y = y / len; z = z / len;
                }
                A.layer[4][T.out_index] = make_float4((x + 1) * 0.5f, (y + 1) * 0.5f, (z + 1) * 0.5f, alpha);
            }
        }
        return;
    }
    hr = group_sum<G>(hr); hg = group_sum<G>(hg); hb = group_sum<G>(hb);
    lr = group_sum<G>(lr); lg = group_sum<G>(lg); lb = group_sum<G>(lb);
    if (lane == 0) {
        float r = 0.f, g = 0.f, b = 0.f;
        float4 ph = make_float4(0.f, 0.f, 0.f, 0.f), pl = make_float4(0.f, 0.f, 0.f, 0.f);
        bool reliable = false;
        if (hemi_rays && relH) {
            const float f = 1 / (float)relH;
            ph = make_float4(hr * f, hg * f, hb * f, 1.f);
            r += ph.x; g += ph.y; b += ph.z; reliable = true;
        }
        if (light_rays && relL) {
            pl = make_float4(lr / (float)shotRounds, lg / (float)shotRounds, lb / (float)shotRounds, 1.f);
            r += pl.x; g += pl.y; b += pl.z; reliable = true;
        }
        A.out[T.out_index] = make_float4(r, g, b, reliable ? flags / 255.f : 0.f);
        if (A.parts) { A.parts[2 * (size_t)ti] = ph; A.parts[2 * (size_t)ti + 1] = pl; }
    }
}

// processTexel by one wave with refill (at most 64 lanes per series chunk replaced by a
// pool): lanes take the series' shooters as they become free (in the direction-binned
// order where gather_texel uses it, else in sequence), trace their hemisphere rays with the
// resumable traversal, and the wave hands out the next shooters whenever kGatherRefill
// lanes are free, instead of waiting for the slowest ray of every 64 (a sweep over 8-48 on
// the batched bake picked 32).  Shooter s still gets the filler point that sequential order
// gives it.
// Light shooters are queued as before; their shadow rays run only while no hemisphere
// ray is in flight (the lanes' LDS stacks are free then): when the queue holds a full
// batch, refills pause until every lane is done, and the rest at the series' end.
constexpr uint32_t kGatherRefill = 32;

template <int STACK, int FEAT>
__device__ void gather_texel_refill(const GatherArgs& A, uint32_t ti, uint32_t lane, int32_t* mystk, uint32_t* lq, uint16_t* lqj,
                                    int stride, float* subc, uint32_t* pscr)
{
    const lsg_texel T = A.texels[ti];
    const uint32_t m = T.sub_count_flags & 0xffffu, flags = (T.sub_count_flags >> 16) & 0xffu;
    const uint32_t hemi_rays = A.hemi_rays, rounds = A.rounds, nl = A.nlights;
    const uint32_t light_rays = nl * rounds;
    if (!m || (!hemi_rays && !light_rays)) {
        if (lane == 0) {
            A.out[T.out_index] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (A.parts) { A.parts[2 * (size_t)ti] = make_float4(0.f, 0.f, 0.f, 0.f); A.parts[2 * (size_t)ti + 1] = make_float4(0.f, 0.f, 0.f, 0.f); }
        }
        return;
    }
    const lsg_subtexel* ts = A.subs + T.sub_begin;
    const uint32_t key = lowbias32(A.seed ^ lowbias32(T.key));
    uint32_t filler = key & 0xffffu;           // wave-uniform
    float area_max = 0.f;
    for (uint32_t j = 0; j < m; j++) area_max += ts[j].area;
    const float near_t = 1000 * A.tmin;
    const float rug_below = __int_as_float(__float_as_int(A.rug) - 1);   // largest float < rugDistance (> 0)
    const uint32_t n1 = first_series(A);
    // the frames of the texel's first subtexels, once per texel, in the wave's LDS
    __builtin_amdgcn_wave_barrier();
    if (lane < min(m, kSubCache)) subtexel_frame(A, ts[lane], subc + kFrame * lane);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float hr = 0.f, hg = 0.f, hb = 0.f, lr = 0.f, lg = 0.f, lb = 0.f;
    uint32_t relH = 0, relL = 0, unrelL = 0, shotRounds = 0, shooter_base = 0;
    for (;;) {
        const bool shootH = (relH < hemi_rays || relH * 10 < hemi_rays) && hemi_rays > 0;
        const bool shootL = (relL + unrelL < light_rays || relL * 10 < light_rays) && unrelL < light_rays * 100;
        if (!shootH && !shootL) break;
        const uint32_t nH = shootH ? hemi_rays : 0, nL = shootL ? rounds : 0;
        const uint32_t n = nH > nL ? nH : nL;
        const bool allH = !shootL || hemi_rays >= rounds;
        const bool allL = !shootH || rounds >= hemi_rays;
        uint32_t cRelH = 0, cRelL = 0, cUnrelL = 0, cRounds = 0;
        uint32_t* const lq0 = lq - lane;
        uint16_t* const lqj0 = lqj - lane;
        uint32_t qhead = 0, qtail = 0;             // wave-uniform
        auto light_batch = [&](uint32_t cnt) {
            __builtin_amdgcn_wave_barrier();
            if (lane < cnt) {
                const uint32_t e = qhead + lane;
                const uint32_t qi = (e & 63u) + ((e / 64u) & 1u) * stride;
                const uint32_t s = lq0[qi], sj = lqj0[qi];
                float3 o; float nx, ny, nz; uint32_t tri;
                shooter_cached(A, ts, sj, key, shooter_base + s, subc, o, nx, ny, nz, tri);
                auto flt = gather_filter<FEAT>(A, tri, near_t, 1, rug_below);
                for (uint32_t li = 0; li < nl; li++) {
                    const lsg_light L = A.lights[li];
                    float dx, dy, dz;
                    if (L.type == 1) { dx = -L.direction[0]; dy = -L.direction[1]; dz = -L.direction[2]; }
                    else { dx = L.position[0] - o.x; dy = L.position[1] - o.y; dz = L.position[2] - o.z; }
                    float dsz = sqrtf(dx * dx + dy * dy + dz * dz);
                    dx = dx / dsz; dy = dy / dsz; dz = dz / dsz;
                    if (L.type == 1) dsz = dsz * A.locality;
                    const float ni = dx * nx + dy * ny + dz * nz;            // gather.cpp:429
                    if (ni <= 0 || !isfinite(ni)) { cRelL++; continue; }
                    bool blocked = false; float bt = 0.f;
                    if (L.cast_shadows) {
                        TraceResult r;
                        uint32_t nv_ = 0, nt_ = 0;
                        // min/max visit: this kernel keeps a hemisphere ray's state alongside
                        trace_closest<STACK, false, decltype(flt), false, false>(flt, A.sc.col, o, make_float3(dx, dy, dz), A.tmin, dsz, mystk, stride, r, nv_, nt_);
                        blocked = r.tri >= 0; bt = r.t;
                    }
                    if (!blocked) {
                        // RRLight::getIrradiance (RRLight.cpp:345-362), REALISTIC attenuation
                        float att = 1.f;
                        if (L.type == 0) {
                            const float ex = o.x - L.position[0], ey = o.y - L.position[1], ez = o.z - L.position[2];
                            att = 1 / (0.000000000137289f + (ex * ex + ey * ey + ez * ez));
                        }
                        const float kl = att * A.light_mult * ni;
                        lr += L.color[0] * kl; lg += L.color[1] * kl; lb += L.color[2] * kl;
                    }
                    const uint32_t unrel = (blocked && bt < A.rug) ? 1u : 0u;    // gather.cpp:483-511
                    cUnrelL += unrel;
                    cRelL += 1u - unrel;
                }
                cRounds++;
            }
            __builtin_amdgcn_wave_barrier();
        };
        // hemisphere rays in the direction-binned order (as gather_texel): free lanes take
        // the next positions of that order
        uint32_t f0 = 0;
        bool perm = false;
        if (shootH && allH && shooter_base == 0 && n == hemi_rays && n > 64u && n <= kPermMaxRays
            && (m == 1 || (A.walk && walk_has(A.walk + ti)))) {
            f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)filler);
            perm = perm_prepare(pscr, f0, n, lane, filler);
        }
        uint32_t next = 0;                         // next shooter to hand out (wave-uniform)
        bool busy = false, ok = true;
        uint32_t stri = 0;                         // the lane's shooter triangle
        TraceState S;
        for (;;) {
            const uint64_t bm = __ballot(busy);
            const bool qfull = qtail - qhead >= 64u;
            if (next >= n && !bm) break;
            if (next < n && !qfull && (64u - (uint32_t)__popcll(bm) >= kGatherRefill || !bm)) {
                // free lanes take the next shooters, in lane order
                const uint64_t fm = ~bm;
                const uint32_t rank = (uint32_t)__popcll(fm & lanemask_lt(lane));
                const uint32_t take = min((uint32_t)__popcll(fm), n - next);
                const bool mine = !busy && rank < take;
                uint32_t s = next + rank;
                float my_x = 0.f, my_y = 0.f;
                bool doH = false, doL = false;
                if (perm) {
                    // binned order: position s gives the shooter and its filler point
                    if (mine) {
                        uint32_t cand;
                        perm_pick(pscr, f0, s, s, cand);
                        filler_point(cand, my_x, my_y);
                        doH = true;
                        if (shootL) doL = allL || shots_before(s + 1, n, rounds) > shots_before(s, n, rounds);
                    }
                } else {
                    if (mine && shootH) doH = allH || shots_before(s + 1, n, hemi_rays) > shots_before(s, n, hemi_rays);
                    if (mine && shootL) doL = allL || shots_before(s + 1, n, rounds) > shots_before(s, n, rounds);
                }
                // filler points in shooter order (GetCirclePoint, gather.cpp:85-90)
                const uint64_t hmask = __ballot(doH);
                const uint32_t need = perm ? 0u : (uint32_t)__popcll(hmask);
                const uint32_t myrank = (uint32_t)__popcll(hmask & lanemask_lt(lane));
                uint32_t got = 0;
                while (got < need) {
                    const uint32_t cand = filler + lane;
                    float cx, cy;
                    filler_point(cand, cx, cy);
                    const bool acc = !(cx * cx + cy * cy >= 1.f);
                    const uint64_t am = __ballot(acc);
                    const uint32_t na = (uint32_t)__popcll(am);
                    const uint32_t want = doH ? myrank - got : 64u;
                    const uint32_t src = (doH && myrank >= got && want < na) ? nth_set_bit(am, want) : lane;
                    const float vx = __shfl(cx, (int)src), vy = __shfl(cy, (int)src);
                    if (doH && myrank >= got && want < na) { my_x = vx; my_y = vy; }
                    if (got + na <= need) { got += na; filler += 64u; }
                    else {
                        const uint32_t last = nth_set_bit(am, need - got - 1u);
                        filler = (uint32_t)__shfl((int)cand, (int)last) + 1u;
                        got = need;
                    }
                }
                // the handed-out shooters' subtexels: from the first series' record, else
                // by replaying the walk
                uint32_t myj = 0;
                if (m > 1) {
                    if (A.walk && shooter_base == 0 && walk_has(A.walk + ti)) myj = walk_pick(A.walk + ti, m, s);
                    else myj = walk_replay(ts, m, area_max, n1, n, shooter_base + next, take, mine ? rank : 64u);
                }
                {
                    const uint64_t lm = __ballot(doL);
                    if (doL) {
                        const uint32_t e = qtail + (uint32_t)__popcll(lm & lanemask_lt(lane));
                        const uint32_t qi = (e & 63u) + ((e / 64u) & 1u) * stride;
                        lq0[qi] = s;
                        lqj0[qi] = (uint16_t)myj;
                    }
                    qtail += (uint32_t)__popcll(lm);
                }
                next += take;
                if (doH) {
                    float3 o; float nx, ny, nz; uint32_t tri;
                    shooter_cached(A, ts, myj, key, shooter_base + s, subc, o, nx, ny, nz, tri);
                    const uint32_t sj = myj;
                    float tx, ty, tz, bx, by, bz;
                    if (sj < kSubCache) {
                        const float* fr = subc + kFrame * sj;
                        tx = fr[3]; ty = fr[4]; tz = fr[5]; bx = fr[6]; by = fr[7]; bz = fr[8];
                    } else {
                        float ox, oy, oz;                                           // RRMathPrivate.h:48-53
                        if (fabsf(nx) > fabsf(nz)) { ox = -ny; oy = nx; oz = 0.f; } else { ox = 0.f; oy = -nz; oz = ny; }
                        const float olen = sqrtf(ox * ox + oy * oy + oz * oz);
                        tx = ox / olen; ty = oy / olen; tz = oz / olen;
                        bx = ny * tz - nz * ty; by = -nx * tz + nz * tx; bz = nx * ty - ny * tx;
                    }
                    const float x = my_x, y = my_y;
                    const float dist = x * x + y * y;
                    const float cosa = sqrtf(1 - dist);                          // gather.cpp:125-126
                    const float3 d = make_float3(nx * cosa + tx * x + bx * y, ny * cosa + ty * x + by * y, nz * cosa + tz * x + bz * y);
                    stri = tri;
                    if (trace_begin(S, A.sc.col, o, d, A.tmin, A.locality)) busy = true;
                    else {
                        // the ray leaves the scene box: environment or nothing (pathtracer.cpp:156-173)
                        if constexpr ((FEAT & kFeatEnv) != 0) {
                            const float4 v = env_at(A.sc, d);
                            hr += v.x * A.env_mult; hg += v.y * A.env_mult; hb += v.z * A.env_mult;
                        }
                        cRelH++;
                    }
                }
            }
            if (busy) {
                // trace until the ray ends or enough lanes are free to hand out more shooters
                const bool refill = next < n && qtail - qhead < 64u;
                const uint32_t yield_at = refill ? 64u - kGatherRefill : 0u;
                auto flt = gather_filter<FEAT>(A, stri, near_t, 0, rug_below);
                if (trace_resume<STACK>(S, flt, A.sc.col, mystk, stride, yield_at, ok)) {
                    busy = false;
                    // a hit reads the exitance of the side it sees; a ray that leaves the
                    // scene returns environment * multiplier, or nothing
                    if (S.bi >= 0) {
                        const float4 e = hit_exitance<(FEAT & kFeatMats) != 0>(A.sc, S.bi, S.bf, S.o, S.d, S.best, A.quality);
                        hr += e.x; hg += e.y; hb += e.z;
                    } else if constexpr ((FEAT & kFeatEnv) != 0) {
                        const float4 v = env_at(A.sc, S.d);
                        hr += v.x * A.env_mult; hg += v.y * A.env_mult; hb += v.z * A.env_mult;
                    }
                    cRelH++;
                }
            }
            // a full light batch runs once no hemisphere ray is in flight
            if (!__any(busy) && qtail - qhead >= 64u) { light_batch(64u); qhead += 64u; }
        }
        while (qtail != qhead) {
            const uint32_t cnt = min(64u, qtail - qhead);
            light_batch(cnt);
            qhead += cnt;
        }
        relH += group_sum_u<64>(cRelH); relL += group_sum_u<64>(cRelL); unrelL += group_sum_u<64>(cUnrelL);
        shotRounds += group_sum_u<64>(cRounds);
        shooter_base += n;
    }
    hr = group_sum<64>(hr); hg = group_sum<64>(hg); hb = group_sum<64>(hb);
    lr = group_sum<64>(lr); lg = group_sum<64>(lg); lb = group_sum<64>(lb);
    if (lane == 0) {
        float r = 0.f, g = 0.f, b = 0.f;
        float4 ph = make_float4(0.f, 0.f, 0.f, 0.f), pl = make_float4(0.f, 0.f, 0.f, 0.f);
        bool reliable = false;
        if (hemi_rays && relH) {
            const float f = 1 / (float)relH;
            ph = make_float4(hr * f, hg * f, hb * f, 1.f);
            r += ph.x; g += ph.y; b += ph.z; reliable = true;
        }
        if (light_rays && relL) {
            pl = make_float4(lr / (float)shotRounds, lg / (float)shotRounds, lb / (float)shotRounds, 1.f);
            r += pl.x; g += pl.y; b += pl.z; reliable = true;
        }
        A.out[T.out_index] = make_float4(r, g, b, reliable ? flags / 255.f : 0.f);
        if (A.parts) { A.parts[2 * (size_t)ti] = ph; A.parts[2 * (size_t)ti + 1] = pl; }
    }
}

// processTexel for texel ti by one lane: the low ray counts of the cheap gather (at most
// 32 shooters per series).  Texels of one quality walk the same shooter schedule, so the
// lanes of a wave trace their hemisphere rays and their few shadow rays in step, with no
// lanes idle while a texel's shadow rays run; nothing is shared between lanes.  The rays
// are those of gather_texel (the same filler indices in shooter order, the same origins
// and shots), summed per lane instead of across a group.
template <int STACK, int FEAT>
__device__ void gather_texel_lane(const GatherArgs& A, uint32_t ti, int32_t* stk, int stride)
{
    const lsg_texel T = A.texels[ti];
    const uint32_t m = T.sub_count_flags & 0xffffu, flags = (T.sub_count_flags >> 16) & 0xffu;
    const uint32_t hemi_rays = A.hemi_rays, rounds = A.rounds, nl = A.nlights;
    const uint32_t light_rays = nl * rounds;
    if (!m || (!hemi_rays && !light_rays)) {
        A.out[T.out_index] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (A.parts) { A.parts[2 * (size_t)ti] = make_float4(0.f, 0.f, 0.f, 0.f); A.parts[2 * (size_t)ti + 1] = make_float4(0.f, 0.f, 0.f, 0.f); }
        return;
    }
    const lsg_subtexel* ts = A.subs + T.sub_begin;
    const uint32_t key = lowbias32(A.seed ^ lowbias32(T.key));
    uint32_t filler = key & 0xffffu;
    float area_max = 0.f;
    for (uint32_t j = 0; j < m; j++) area_max += ts[j].area;
    const float near_t = 1000 * A.tmin;
    const float rug_below = __int_as_float(__float_as_int(A.rug) - 1);   // largest float < rugDistance (> 0)
    float hr = 0.f, hg = 0.f, hb = 0.f, lr = 0.f, lg = 0.f, lb = 0.f;
    uint32_t relH = 0, relL = 0, unrelL = 0, shotRounds = 0, shooter_base = 0;
    SubWalk wk;                                // the lane walks its own texel's shooters in order
    wk.init(ts);
    for (;;) {
        const bool shootH = (relH < hemi_rays || relH * 10 < hemi_rays) && hemi_rays > 0;
        const bool shootL = (relL + unrelL < light_rays || relL * 10 < light_rays) && unrelL < light_rays * 100;
        if (!shootH && !shootL) break;
        const uint32_t nH = shootH ? hemi_rays : 0, nL = shootL ? rounds : 0;
        const uint32_t n = nH > nL ? nH : nL;
        const bool allH = !shootL || hemi_rays >= rounds;
        const bool allL = !shootH || rounds >= hemi_rays;
        const float astep = walk_step(area_max, n);
        for (uint32_t s = 0; s < n; s++) {
            const uint32_t j = m > 1 ? wk.next(ts, m, astep) : 0u;      // every shooter steps the walk
            const bool doH = shootH && (allH || shots_before(s + 1, n, hemi_rays) > shots_before(s, n, hemi_rays));
            const bool doL = shootL && (allL || shots_before(s + 1, n, rounds) > shots_before(s, n, rounds));
            if (!doH && !doL) continue;
            float3 o; float nx, ny, nz; uint32_t tri;
            shooter_at(A, ts, j, key, shooter_base + s, o, nx, ny, nz, tri);
            if (doH) {
                float ox, oy, oz;                                               // RRMathPrivate.h:48-53
                if (fabsf(nx) > fabsf(nz)) { ox = -ny; oy = nx; oz = 0.f; } else { ox = 0.f; oy = -nz; oz = ny; }
                const float olen = sqrtf(ox * ox + oy * oy + oz * oz);
                const float tx = ox / olen, ty = oy / olen, tz = oz / olen;
                const float bx = ny * tz - nz * ty, by = -nx * tz + nz * tx, bz = nx * ty - ny * tx;
                float x, y;
                for (;;) {                                                      // GetCirclePoint, gather.cpp:85-90
                    filler_point(filler++, x, y);
                    if (!(x * x + y * y >= 1.f)) break;
                }
                const float dist = x * x + y * y;
                const float cosa = sqrtf(1 - dist);                          // gather.cpp:125-126
                const float3 d = make_float3(nx * cosa + tx * x + bx * y, ny * cosa + ty * x + by * y, nz * cosa + tz * x + bz * y);
                auto flt = gather_filter<FEAT>(A, tri, near_t, 0, rug_below);
                TraceResult r;
                uint32_t nv_ = 0, nt_ = 0;
                // min/max visit: measured faster here (the near/far form raised this kernel
                // from 109 to 122 VGPRs and the 10M-triangle city's cheap gathers slowed)
                trace_closest<STACK, false, decltype(flt), true, false>(flt, A.sc.col, o, d, A.tmin, A.locality, stk, stride, r, nv_, nt_);
                if (r.tri >= 0) {
                    const float4 e = hit_exitance<(FEAT & kFeatMats) != 0>(A.sc, r.tri, r.front, o, d, r.t, A.quality);
                    hr += e.x; hg += e.y; hb += e.z;
                } else if constexpr ((FEAT & kFeatEnv) != 0) {
                    const float4 v = env_at(A.sc, d);
                    hr += v.x * A.env_mult; hg += v.y * A.env_mult; hb += v.z * A.env_mult;
                }
                relH++;
            }
            if (doL) {
                auto flt = gather_filter<FEAT>(A, tri, near_t, 1, rug_below);
                for (uint32_t li = 0; li < nl; li++) {
                    const lsg_light L = A.lights[li];
                    float dx, dy, dz;
                    if (L.type == 1) { dx = -L.direction[0]; dy = -L.direction[1]; dz = -L.direction[2]; }
                    else { dx = L.position[0] - o.x; dy = L.position[1] - o.y; dz = L.position[2] - o.z; }
                    float dsz = sqrtf(dx * dx + dy * dy + dz * dz);
                    dx = dx / dsz; dy = dy / dsz; dz = dz / dsz;
                    if (L.type == 1) dsz = dsz * A.locality;
                    const float ni = dx * nx + dy * ny + dz * nz;            // gather.cpp:429
                    if (ni <= 0 || !isfinite(ni)) { relL++; continue; }
                    bool blocked = false; float bt = 0.f;
                    if (L.cast_shadows) {
                        TraceResult r;
                        uint32_t nv_ = 0, nt_ = 0;
                        trace_closest<STACK, false, decltype(flt), false, false>(flt, A.sc.col, o, make_float3(dx, dy, dz), A.tmin, dsz, stk, stride, r, nv_, nt_);
                        blocked = r.tri >= 0; bt = r.t;
                    }
                    if (!blocked) {
                        // RRLight::getIrradiance (RRLight.cpp:345-362), REALISTIC attenuation
                        float att = 1.f;
                        if (L.type == 0) {
                            const float ex = o.x - L.position[0], ey = o.y - L.position[1], ez = o.z - L.position[2];
                            att = 1 / (0.000000000137289f + (ex * ex + ey * ey + ez * ez));
                        }
                        const float kl = att * A.light_mult * ni;
                        lr += L.color[0] * kl; lg += L.color[1] * kl; lb += L.color[2] * kl;
                    }
                    const uint32_t unrel = (blocked && bt < A.rug) ? 1u : 0u;    // gather.cpp:483-511
                    unrelL += unrel;
                    relL += 1u - unrel;
                }
                shotRounds++;
            }
        }
        shooter_base += n;
    }
    float r = 0.f, g = 0.f, b = 0.f;
    float4 ph = make_float4(0.f, 0.f, 0.f, 0.f), pl = make_float4(0.f, 0.f, 0.f, 0.f);
    bool reliable = false;
    if (hemi_rays && relH) {
        const float f = 1 / (float)relH;
        ph = make_float4(hr * f, hg * f, hb * f, 1.f);
        r += ph.x; g += ph.y; b += ph.z; reliable = true;
    }
    if (light_rays && relL) {
        pl = make_float4(lr / (float)shotRounds, lg / (float)shotRounds, lb / (float)shotRounds, 1.f);
        r += pl.x; g += pl.y; b += pl.z; reliable = true;
    }
    A.out[T.out_index] = make_float4(r, g, b, reliable ? flags / 255.f : 0.f);
    if (A.parts) { A.parts[2 * (size_t)ti] = ph; A.parts[2 * (size_t)ti + 1] = pl; }
}

// launches of at least this many rounds of 64-texel waves per wave slot use k_gather_lanes
constexpr uint32_t kLaneRounds = 16;

// low ray counts without output layers: one texel per lane, 256 per block
template <int STACK, int FEAT = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_gather_lanes(GatherArgs A)
{
    __shared__ int32_t stk[STACK * 256];
    const uint32_t tid = threadIdx.x;
    const uint32_t ti = blockIdx.x * 256 + tid;
    if (ti >= A.ntex) return;
    gather_texel_lane<STACK, FEAT>(A, ti, stk + tid, 256);
}

// One texel per wave, 4 per block.  The plain and environment kernels are capped at 128
// VGPRs (4 waves per SIMD, measured faster than 3 with more registers); the layered ones
// hold 12 more sums per lane and are left uncapped rather than spill.
template <int STACK, bool STATS, bool LAYERS, int FEAT>
__device__ inline void gather_block(const GatherArgs& A)
{
    __shared__ int32_t stk[STACK * 256];
    __shared__ uint32_t lq[2 * 256];
    __shared__ uint16_t lqj[2 * 256];
    __shared__ float subc[4 * kFrame * kSubCache];
    __shared__ uint32_t pscr[4 * kPermScr];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    // wave-uniform: kept in a scalar register, the texel record is read by scalar loads
    const uint32_t ti = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (tid >> 6)));
    if (ti >= A.ntex) return;                 // whole wavefront leaves together
    gather_texel<STACK, STATS, 64, LAYERS, FEAT>(A, ti, lane, stk + tid, lq + tid, lqj + tid, 256,
                                                subc + (tid >> 6) * kFrame * kSubCache, pscr + (tid >> 6) * kPermScr);
}

template <int STACK, bool STATS, int FEAT = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_gather(GatherArgs A)
{
    gather_block<STACK, STATS, false, FEAT>(A);
}

template <int STACK, int FEAT>
__global__ __launch_bounds__(256) void k_gather_layers(GatherArgs A)
{
    gather_block<STACK, false, true, FEAT>(A);
}

// low ray counts (at most G shooters per series): 256/G texels per block, G lanes each
template <int STACK, int G, bool LAYERS, int FEAT>
__device__ inline void gather_groups(const GatherArgs& A)
{
    __shared__ int32_t stk[STACK * 256];
    __shared__ uint32_t lq[2 * 256];
    __shared__ uint16_t lqj[2 * 256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t ti = blockIdx.x * (256u / G) + tid / G;
    if (ti >= A.ntex) return;                 // the whole group leaves together
    gather_texel<STACK, false, G, LAYERS, FEAT>(A, ti, lane, stk + tid, lq + tid, lqj + tid, 256);
}

template <int STACK, int G, int FEAT = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_gather_grouped(GatherArgs A)
{
    gather_groups<STACK, G, false, FEAT>(A);
}

template <int STACK, int G, int FEAT>
__global__ __launch_bounds__(256) void k_gather_grouped_layers(GatherArgs A)
{
    gather_groups<STACK, G, true, FEAT>(A);
}

__global__ void k_queue_reset(uint32_t* counters, uint32_t n)
{
    if (threadIdx.x < n) counters[threadIdx.x] = 0;
}

// Capped at 128 VGPRs it runs 4 waves per SIMD, measured faster than 3 with more
// registers; the chunked form fits, the refill form (REFILL) spills 44 bytes.
template <int STACK, int WAVES, int FEAT = 0, bool REFILL = false>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(4))) void k_gather_persistent(GatherArgs A, WorkQueue Q)
{
    __shared__ int32_t stk[STACK * 64 * WAVES];
    __shared__ uint32_t lq[2 * 64 * WAVES];
    __shared__ uint16_t lqj[2 * 64 * WAVES];
    __shared__ float subc[WAVES * kFrame * kSubCache];
    __shared__ uint32_t pscr[WAVES * kPermScr];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t home = Q.nqueues > 1 ? blockIdx.x % Q.nqueues : 0;
    for (uint32_t qi = 0; qi < Q.nqueues; qi++) {
        const uint32_t q = (home + qi) % Q.nqueues;
        const uint32_t t0 = (uint32_t)((uint64_t)Q.ntiles * q / Q.nqueues);
        const uint32_t t1 = (uint32_t)((uint64_t)Q.ntiles * (q + 1) / Q.nqueues);
        for (;;) {
            uint32_t k = 0;
            if (lane == 0) k = atomicAdd(&Q.counters[q], 1u);
            k = (uint32_t)__builtin_amdgcn_readfirstlane((int)k);   // lane 0's tile, scalar
            const uint32_t tile = t0 + k;
            if (tile >= t1) break;           // this queue is exhausted for every lane of the wave
            const uint32_t a = tile * Q.tile, b = min(A.ntex, a + Q.tile);
            for (uint32_t ti = a; ti < b; ti++)
                if constexpr (REFILL) gather_texel_refill<STACK, FEAT>(A, ti, lane, stk + tid, lq + tid, lqj + tid, 64 * WAVES,
                                                                     subc + (tid >> 6) * kFrame * kSubCache,
                                                                     pscr + (tid >> 6) * kPermScr);
                else gather_texel<STACK, false, 64, false, FEAT>(A, ti, lane, stk + tid, lq + tid, lqj + tid, 64 * WAVES,
                                                                subc + (tid >> 6) * kFrame * kSubCache,
                                                                pscr + (tid >> 6) * kPermScr);
        }
    }
}

static inline void dfree(void* p) { if (p) (void)hipFree(p); }

// TangentBasis::buildBasisFromNormal (RRMesh.cpp:70-75) at every vertex: the frame
// RRMeshArrays::getTriangleNormals gives a mesh without tangents (RRMeshArrays.cpp:297-305)
__global__ __launch_bounds__(256) void k_default_frames(const float* __restrict__ vn, uint32_t nv, float* __restrict__ vt,
                                                        float* __restrict__ vb)
{
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    const float nx = vn[3 * (size_t)v], ny = vn[3 * (size_t)v + 1], nz = vn[3 * (size_t)v + 2];
    float ox, oy, oz;                                                    // RRMathPrivate.h:48-53
    if (fabsf(nx) > fabsf(nz)) { ox = -ny; oy = nx; oz = 0.f; } else { ox = 0.f; oy = -nz; oz = ny; }
    const float len = sqrtf(ox * ox + oy * oy + oz * oz);
    const float tx = ox / len, ty = oy / len, tz = oz / len;
    vt[3 * (size_t)v] = tx; vt[3 * (size_t)v + 1] = ty; vt[3 * (size_t)v + 2] = tz;
    vb[3 * (size_t)v] = ny * tz - nz * ty;                              // RRMathPrivate.h:55-58
    vb[3 * (size_t)v + 1] = -nx * tz + nz * tx;
    vb[3 * (size_t)v + 2] = nx * ty - ny * tx;
}

static int frames_alloc(SceneImpl& S)
{
    const uint32_t nv = S.col->impl.nverts;
    if (!S.d_vtan) LSG_HIP(hipMalloc(&S.d_vtan, 12 * (size_t)(nv ? nv : 1)));
    if (!S.d_vbit) LSG_HIP(hipMalloc(&S.d_vbit, 12 * (size_t)(nv ? nv : 1)));
    return LSG_OK;
}

int scene_frames(lsg_scene* s, hipStream_t st)
{
    SceneImpl& S = s->impl;
    if (S.d_vtan && S.d_vbit) return LSG_OK;
    const int e = frames_alloc(S);
    if (e) return e;
    const uint32_t nv = S.col->impl.nverts;
    if (nv) k_default_frames<<<(nv + 255) / 256, 256, 0, st>>>(S.d_vnormal, nv, S.d_vtan, S.d_vbit);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

int gather_args(lsg_scene* s, const lsg_texel* texels, uint32_t ntexels, const lsg_subtexel* subs,
                const lsg_gather_params* p, float* out_rgba, GatherArgs& A, int& stack)
{
    const ColliderImpl& C = s->impl.col->impl;
    stack = stack_size_for_depth(C.depth);
    if (!stack) return fail(LSG_ERR_LIMIT, "tree too deep");
    if (!(p->ray_length_min >= 0) || !(p->rug_distance > 0) || !(p->locality > 0))
        return fail(LSG_ERR_INVALID, "gather: ray_length_min >= 0, rug_distance > 0 and locality > 0 required");
    A.sc.col = C.view();
    A.sc.verts = C.d_verts; A.sc.idx = C.d_idx;
    A.sc.vnormal = s->impl.d_vnormal; A.sc.side = s->impl.d_side; A.sc.exitance = s->impl.d_exit;
    A.sc.vtangent = s->impl.d_vtan; A.sc.vbitangent = s->impl.d_vbit;
    // the environment is sampled only with a nonzero multiplier (gather.cpp:149); the first
    // environment decides the hemisphere gather (gather.cpp:149, 197), the blend of both
    // what a ray that leaves sees (pathtracer.cpp:29-34)
    const SceneImpl& SI = s->impl;
    const bool env = SI.env_type != LSG_ENV_NONE && p->environment_multiplier != 0;
    const bool blend = SI.env_blended && (SI.env_type != LSG_ENV_NONE || SI.env1_type != LSG_ENV_NONE) && p->environment_multiplier != 0;
    A.sc.env = SI.d_env; A.sc.env_type = blend ? kEnvBlend : (env ? SI.env_type : 0u);
    A.sc.env_w = SI.env_w; A.sc.env_h = SI.env_h;
    A.sc.env1 = SI.d_env1; A.sc.env0_type = SI.env_type; A.sc.env1_type = SI.env1_type;
    A.sc.env1_w = SI.env1_w; A.sc.env1_h = SI.env1_h;
    for (int i = 0; i < 18; i++) A.sc.env_rot[i] = SI.env_rot[i];
    A.sc.env_blend = SI.env_blend;
    A.env_mult = p->environment_multiplier;
    for (int i = 0; i < 5; i++) A.layer[i] = nullptr;
    A.dirs = 0; A.layered = 0;
    A.texels = texels; A.ntex = ntexels; A.subs = subs; A.out = (float4*)out_rgba; A.stats = nullptr; A.parts = nullptr;
    A.quality = p->quality > 1 ? p->quality : 1;      // gather.cpp:865
    A.hemi_rays = (p->hemisphere || env) ? (p->quality > 1 ? p->quality : 1) : 0;    // gather.cpp:197
    A.rounds = (p->light_multiplier != 0 && s->impl.nlights) ? (p->light_rounds ? p->light_rounds : p->quality / 10 + 1) : 0; // gather.cpp:396
    A.tmin = p->ray_length_min; A.locality = p->locality; A.rug = p->rug_distance;
    A.light_mult = p->light_multiplier; A.seed = p->seed;
    A.lights = s->impl.d_lights; A.nlights = s->impl.nlights;
    A.sc.ind = s->impl.d_ind; A.sc.slot = C.d_slot;
    A.sc.pm.tri_mat = s->impl.d_tri_mat; A.sc.pm.tri_uv = s->impl.d_tri_uv; A.sc.pm.mats = s->impl.d_mats;
    A.sc.pm.tex = s->impl.d_tex; A.sc.pm.texels = s->impl.d_texels;
    if (s->impl.d_tri_mat) {
        const int e = check_point_quality(s->impl, A.quality);
        if (e) return e;
    }
    A.walk = nullptr; A.walk_buf = nullptr; A.walk_cap = 0;
    A.pq = nullptr;
    {
        SceneImpl& SP = s->impl;
        if (!SP.d_pq) LSG_HIP(hipMalloc(&SP.d_pq, sizeof(uint32_t) * 8 * kQueueSlots));
        A.pq = SP.d_pq + 8 * (__atomic_fetch_add(&SP.pq_next, 1u, __ATOMIC_RELAXED) % kQueueSlots);
    }
    return LSG_OK;
}

int scene_walk_scratch(lsg_scene* s, uint32_t n, GatherArgs& A)
{
    SceneImpl& S = s->impl;
    if (n > S.walk_cap) {
        dfree(S.d_walk);
        S.d_walk = nullptr; S.walk_cap = 0;
        LSG_HIP(hipMalloc(&S.d_walk, 2 * sizeof(uint4) * (size_t)n));
        S.walk_cap = n;
    }
    A.walk_buf = S.d_walk; A.walk_cap = S.walk_cap;
    return LSG_OK;
}

template <int FEAT, bool REFILL>
static void launch_persistent(const GatherArgs& A, int stack, const WorkQueue& Q, uint32_t waves_per_block, uint32_t blocks,
                              hipStream_t st)
{
    if (waves_per_block == 8) {
        if (stack <= 32) k_gather_persistent<32, 8, FEAT, REFILL><<<blocks, 512, 0, st>>>(A, Q);
        else k_gather_persistent<64, 8, FEAT, REFILL><<<blocks, 512, 0, st>>>(A, Q);
    } else if (waves_per_block == 2) {
        if (stack <= 32) k_gather_persistent<32, 2, FEAT, REFILL><<<blocks, 128, 0, st>>>(A, Q);
        else k_gather_persistent<64, 2, FEAT, REFILL><<<blocks, 128, 0, st>>>(A, Q);
    } else {
        if (stack <= 32) k_gather_persistent<32, 4, FEAT, REFILL><<<blocks, 256, 0, st>>>(A, Q);
        else k_gather_persistent<64, 4, FEAT, REFILL><<<blocks, 256, 0, st>>>(A, Q);
    }
}

int launch_gather_persistent(const GatherArgs& A, int stack, const WorkQueue& Q, uint32_t waves_per_block, uint32_t blocks,
                             bool refill, hipStream_t st)
{
    k_queue_reset<<<1, 64, 0, st>>>(Q.counters, Q.nqueues);
    GatherArgs B = A;
    if (B.walk_buf && B.ntex <= B.walk_cap && B.ntex) {
        k_subtexel_walk<<<(B.ntex + 255) / 256, 256, 0, st>>>(B.texels, B.ntex, B.subs, first_series(B), (WalkRec*)B.walk_buf);
        B.walk = (const WalkRec*)B.walk_buf;
    }
    // shooters per series above 32: the texel's lanes may take shooters as they free up
    refill = refill && std::max(B.hemi_rays, B.rounds) > 32u;
    if (B.sc.pm.tri_mat) {
        if (B.sc.env_type) {
            if (refill) launch_persistent<kFeatEnv | kFeatMats, true>(B, stack, Q, waves_per_block, blocks, st);
            else launch_persistent<kFeatEnv | kFeatMats, false>(B, stack, Q, waves_per_block, blocks, st);
        } else {
            if (refill) launch_persistent<kFeatMats, true>(B, stack, Q, waves_per_block, blocks, st);
            else launch_persistent<kFeatMats, false>(B, stack, Q, waves_per_block, blocks, st);
        }
    } else if (B.sc.env_type) {
        if (refill) launch_persistent<kFeatEnv, true>(B, stack, Q, waves_per_block, blocks, st);
        else launch_persistent<kFeatEnv, false>(B, stack, Q, waves_per_block, blocks, st);
    } else {
        if (refill) launch_persistent<0, true>(B, stack, Q, waves_per_block, blocks, st);
        else launch_persistent<0, false>(B, stack, Q, waves_per_block, blocks, st);
    }
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

static int launch_gather_one(const GatherArgs& A, int stack, hipStream_t st);

// very large texel lists go out in launches of at most kLaunchTexels texels at 1024
// shooters per series (a few seconds of work each), proportionally more at fewer
// shooters (fewer launch drains for low-quality gathers), at most 2^25 texels (64
// threads per texel stay below 2^31 threads per launch)
constexpr uint32_t kLaunchTexels = 1u << 22;
constexpr uint32_t kPersistentMinTexels = 16384;   // smaller launches stay one wave per texel

uint32_t gather_chunk(const GatherArgs& A)
{
    const uint32_t series = std::max(std::max(A.hemi_rays, A.rounds), 1u);
    return (uint32_t)std::min<uint64_t>((uint64_t)kLaunchTexels * std::max(1u, 1024u / series), 1u << 25);
}

int launch_gather(const GatherArgs& A, int stack, hipStream_t st)
{
    const uint32_t chunk = gather_chunk(A);
    for (uint64_t base = 0; base < A.ntex; base += chunk) {
        GatherArgs B = A;
        B.texels = A.texels + base;
        B.ntex = (uint32_t)std::min<uint64_t>(chunk, A.ntex - base);
        if (A.parts) B.parts = A.parts + 2 * (size_t)base;
        const int e = launch_gather_one(B, stack, st);
        if (e) return e;
    }
    return LSG_OK;
}

// compute units of the current device (cached per device)
static uint32_t device_cus()
{
    static std::atomic<uint32_t> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    uint32_t c = cached[dev].load(std::memory_order_relaxed);
    if (!c) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        c = (uint32_t)n;
        cached[dev].store(c, std::memory_order_relaxed);
    }
    return c;
}

// one texel per lane makes each wave G times longer: worth it only when the launch
// fills every wave slot (4 per SIMD, 4 SIMDs per CU) many times over, else the last waves
// run on an almost idle chip (733k texels: 20% slower, 11.5M texels: 13% faster than groups)
static bool use_lanes(const GatherArgs& A, uint32_t ntex)
{
    const uint32_t series = std::max(A.hemi_rays, A.rounds);
    return !A.layered && series <= 32 && (uint64_t)ntex >= (uint64_t)kLaneRounds * 64u * 16u * device_cus();
}

// the walk records the kernels of one launch read (none for the one-texel-per-lane kernel,
// whose lanes walk their own texels)
int walk_records(lsg_scene* s, GatherArgs& A)
{
    const uint32_t n = std::min(A.ntex, gather_chunk(A));
    if (!n || use_lanes(A, n)) return LSG_OK;
    return scene_walk_scratch(s, n, A);
}

template <bool LAYERS, int FEAT>
static void launch_gather_kernels(GatherArgs A, int stack, hipStream_t st)
{
    // shooters per series = max(hemisphere rays, light rounds); a group of G >= that
    // many lanes does the same work as a whole wave with idle lanes
    const uint32_t series = std::max(A.hemi_rays, A.rounds);
    const bool lanes = !LAYERS && use_lanes(A, A.ntex);
    if (!lanes && A.walk_buf && A.ntex <= A.walk_cap) {
        k_subtexel_walk<<<(A.ntex + 255) / 256, 256, 0, st>>>(A.texels, A.ntex, A.subs, first_series(A), (WalkRec*)A.walk_buf);
        A.walk = (const WalkRec*)A.walk_buf;
    }
    if (series <= 32) {
        const uint32_t G = series <= 16 ? 16u : 32u;
        const unsigned grid = (unsigned)(((uint64_t)A.ntex * G + 255) / 256);
        if constexpr (LAYERS) {
            if (G == 16) {
                if (stack <= 32) k_gather_grouped_layers<32, 16, FEAT><<<grid, 256, 0, st>>>(A);
                else k_gather_grouped_layers<64, 16, FEAT><<<grid, 256, 0, st>>>(A);
            } else {
                if (stack <= 32) k_gather_grouped_layers<32, 32, FEAT><<<grid, 256, 0, st>>>(A);
                else k_gather_grouped_layers<64, 32, FEAT><<<grid, 256, 0, st>>>(A);
            }
        } else if (lanes) {
            const unsigned lgrid = (A.ntex + 255) / 256;
            if (stack <= 32) k_gather_lanes<32, FEAT><<<lgrid, 256, 0, st>>>(A);
            else k_gather_lanes<64, FEAT><<<lgrid, 256, 0, st>>>(A);
        } else if (G == 16) {
            if (stack <= 32) k_gather_grouped<32, 16, FEAT><<<grid, 256, 0, st>>>(A);
            else k_gather_grouped<64, 16, FEAT><<<grid, 256, 0, st>>>(A);
        } else {
            if (stack <= 32) k_gather_grouped<32, 32, FEAT><<<grid, 256, 0, st>>>(A);
            else k_gather_grouped<64, 32, FEAT><<<grid, 256, 0, st>>>(A);
        }
        return;
    }
    const unsigned grid = (A.ntex + 3) / 4;
    if constexpr (LAYERS) {
        if (stack <= 32) k_gather_layers<32, FEAT><<<grid, 256, 0, st>>>(A);
        else k_gather_layers<64, FEAT><<<grid, 256, 0, st>>>(A);
    } else {
        if (stack <= 32) k_gather<32, false, FEAT><<<grid, 256, 0, st>>>(A);
        else k_gather<64, false, FEAT><<<grid, 256, 0, st>>>(A);
    }
}

static int launch_gather_one(const GatherArgs& A, int stack, hipStream_t st)
{
    if (!A.ntex) return LSG_OK;
    // many texels with more than 32 hemisphere rays each: the persistent kernel (texel
    // tiles from 8 per-XCD queues, direction-binned hemisphere order; bit-identical to
    // k_gather), with the bake plan's default shape.  Gathers of light rays only (a first
    // gather) stay one wave per texel, which is faster for them.
    if (!A.layered && A.pq && !A.parts && !A.stats && A.ntex >= kPersistentMinTexels && A.hemi_rays > 32u) {
        WorkQueue Q;
        Q.counters = A.pq; Q.nqueues = 8; Q.ntiles = A.ntex; Q.tile = 1;
        return launch_gather_persistent(A, stack, Q, 4, device_cus() * 4, false, st);
    }
    switch (features(A)) {
    case 0: if (A.layered) launch_gather_kernels<true, 0>(A, stack, st); else launch_gather_kernels<false, 0>(A, stack, st); break;
    case kFeatEnv:
        if (A.layered) launch_gather_kernels<true, kFeatEnv>(A, stack, st); else launch_gather_kernels<false, kFeatEnv>(A, stack, st);
        break;
    case kFeatMats:
        if (A.layered) launch_gather_kernels<true, kFeatMats>(A, stack, st); else launch_gather_kernels<false, kFeatMats>(A, stack, st);
        break;
    default:
        if (A.layered) launch_gather_kernels<true, kFeatEnv | kFeatMats>(A, stack, st);
        else launch_gather_kernels<false, kFeatEnv | kFeatMats>(A, stack, st);
        break;
    }
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

} // namespace lsg


using namespace lsg;

extern "C" {

int lsg_scene_set_exitance(lsg_scene* s, const float* exitance)
{
    if (!s || !exitance) return fail(LSG_ERR_INVALID, "lsg_scene_set_exitance: null argument");
    const uint32_t nt = s->impl.col->impl.ntri;
    LSG_HIP(hipMemcpy(s->impl.d_exit, exitance, 32 * (size_t)nt, hipMemcpyHostToDevice));
    return LSG_OK;
}

int lsg_scene_get_exitance(const lsg_scene* s, float* exitance)
{
    if (!s || !exitance) return fail(LSG_ERR_INVALID, "lsg_scene_get_exitance: null argument");
    const uint32_t nt = s->impl.col->impl.ntri;
    LSG_HIP(hipDeviceSynchronize());
    LSG_HIP(hipMemcpy(exitance, s->impl.d_exit, 32 * (size_t)nt, hipMemcpyDeviceToHost));
    return LSG_OK;
}

int lsg_scene_set_indirect(lsg_scene* s, const float* corners, const float* albedo, const uint8_t* sides)
{
    if (!s) return fail(LSG_ERR_INVALID, "lsg_scene_set_indirect: null scene");
    SceneImpl& S = s->impl;
    const uint32_t nt = S.col->impl.ntri;
    std::vector<uint8_t> side(S.h_side);
    for (uint32_t t = 0; t < nt; t++) side[t] &= (uint8_t)~(LSG_SIDE_FRONT_INDIRECT | LSG_SIDE_BACK_INDIRECT);
    if (!corners) {
        dfree(S.d_ind); S.d_ind = nullptr;
        if (nt) LSG_HIP(hipMemcpy(S.d_side, side.data(), nt, hipMemcpyHostToDevice));
        S.h_side = side;
        return LSG_OK;
    }
    if (!albedo) return fail(LSG_ERR_INVALID, "lsg_scene_set_indirect: corners without albedo");
    std::vector<float4> rec(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; t++) {
        for (int k = 0; k < 3; k++)
            rec[3 * (size_t)t + k] = make_float4(corners[9 * (size_t)t + 3 * k], corners[9 * (size_t)t + 3 * k + 1],
                                                 corners[9 * (size_t)t + 3 * k + 2], albedo[3 * (size_t)t + k]);
        const uint8_t b = sides ? sides[t] : 3u;
        side[t] |= (uint8_t)(((b & 1u) ? LSG_SIDE_FRONT_INDIRECT : 0u) | ((b & 2u) ? LSG_SIDE_BACK_INDIRECT : 0u));
    }
    if (!S.d_ind) LSG_HIP(hipMalloc(&S.d_ind, sizeof(float4) * 3 * (size_t)(nt ? nt : 1)));
    if (nt) {
        LSG_HIP(hipMemcpy(S.d_ind, rec.data(), sizeof(float4) * 3 * (size_t)nt, hipMemcpyHostToDevice));
        LSG_HIP(hipMemcpy(S.d_side, side.data(), nt, hipMemcpyHostToDevice));
    }
    S.h_side = side;
    return LSG_OK;
}

int lsg_scene_set_lights(lsg_scene* s, const lsg_light* lights, uint32_t n)
{
    if (!s || (n && !lights)) return fail(LSG_ERR_INVALID, "lsg_scene_set_lights: null argument");
    if (n > (uint32_t)kMaxLights) return fail(LSG_ERR_LIMIT, "lsg_scene_set_lights: %u lights, at most %d", n, kMaxLights);
    for (uint32_t i = 0; i < n; i++)
        if (lights[i].type != 0 && lights[i].type != 1) return fail(LSG_ERR_INVALID, "light %u: type must be 0 (point) or 1 (directional)", i);
    if (n) LSG_HIP(hipMemcpy(s->impl.d_lights, lights, sizeof(lsg_light) * n, hipMemcpyHostToDevice));
    s->impl.nlights = n;
    return LSG_OK;
}

int lsg_scene_create(const lsg_collider* col, const float* vertex_normals, const uint8_t* side_bits, const float* exitance,
                     const lsg_light* lights, uint32_t nlights, lsg_scene** out)
{
    if (!out) return fail(LSG_ERR_INVALID, "lsg_scene_create: out is null");
    *out = nullptr;
    if (!col || !vertex_normals || !side_bits || !exitance) return fail(LSG_ERR_INVALID, "lsg_scene_create: null argument");
    lsg_scene* s = new (std::nothrow) lsg_scene;
    if (!s) return fail(LSG_ERR_OOM, "lsg_scene_create: host allocation");
    SceneImpl& S = s->impl;
    S.col = col;
    const uint32_t nv = col->impl.nverts, nt = col->impl.ntri;
    auto bail = [&](int st) { dfree(S.d_vnormal); dfree(S.d_side); dfree(S.d_exit); dfree(S.d_lights); delete s; return st; };
    hipError_t e;
    if ((e = hipMalloc(&S.d_lights, sizeof(lsg_light) * kMaxLights)) != hipSuccess) return bail(hip_fail(e, "hipMalloc", __LINE__));
    if ((e = hipMalloc(&S.d_vnormal, 12 * (size_t)(nv ? nv : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc", __LINE__));
    if ((e = hipMalloc(&S.d_side, (size_t)(nt ? nt : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc", __LINE__));
    if ((e = hipMalloc(&S.d_exit, 32 * (size_t)(nt ? nt : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc", __LINE__));
    if ((e = hipMemcpy(S.d_vnormal, vertex_normals, 12 * (size_t)nv, hipMemcpyHostToDevice)) != hipSuccess) return bail(hip_fail(e, "h2d", __LINE__));
    if ((e = hipMemcpy(S.d_side, side_bits, (size_t)nt, hipMemcpyHostToDevice)) != hipSuccess) return bail(hip_fail(e, "h2d", __LINE__));
    S.h_side.assign(side_bits, side_bits + nt);
    if ((e = hipMemcpy(S.d_exit, exitance, 32 * (size_t)nt, hipMemcpyHostToDevice)) != hipSuccess) return bail(hip_fail(e, "h2d", __LINE__));
    int st = lsg_scene_set_lights(s, lights, nlights);
    if (st) return bail(st);
    *out = s;
    return LSG_OK;
}

void lsg_scene_destroy(lsg_scene* s)
{
    if (!s) return;
    dfree(s->impl.d_vnormal); dfree(s->impl.d_side); dfree(s->impl.d_exit); dfree(s->impl.d_lights);
    dfree(s->impl.d_vtan); dfree(s->impl.d_vbit); dfree(s->impl.d_env); dfree(s->impl.d_env1); dfree(s->impl.d_pq); dfree(s->impl.d_lmaps);
    dfree(s->impl.d_ind); dfree(s->impl.d_walk);
    free_point_materials(s->impl);
    delete s;
}

int lsg_scene_set_environment(lsg_scene* s, int32_t type, uint32_t width, uint32_t height, const float* rgb)
{
    if (!s) return fail(LSG_ERR_INVALID, "lsg_scene_set_environment: null scene");
    SceneImpl& S = s->impl;
    if (type == LSG_ENV_NONE) {
        dfree(S.d_env); S.d_env = nullptr; S.env_type = 0; S.env_w = S.env_h = 0;
        S.env_version++;
        return LSG_OK;
    }
    if (type != LSG_ENV_CUBE && type != LSG_ENV_PANORAMA) return fail(LSG_ERR_INVALID, "lsg_scene_set_environment: type %d", type);
    if (!rgb || !width || !height) return fail(LSG_ERR_INVALID, "lsg_scene_set_environment: empty environment");
    if (type == LSG_ENV_CUBE && width != height) return fail(LSG_ERR_INVALID, "lsg_scene_set_environment: cube faces must be square");
    const uint64_t n = (uint64_t)width * height * (type == LSG_ENV_CUBE ? 6u : 1u);
    if (n > (1ull << 28)) return fail(LSG_ERR_LIMIT, "lsg_scene_set_environment: %llu texels", (unsigned long long)n);
    std::vector<float4> t((size_t)n);
    for (size_t i = 0; i < (size_t)n; i++) t[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.f);
    float4* d = nullptr;
    LSG_HIP(hipMalloc(&d, 16 * (size_t)n));
    hipError_t e = hipMemcpy(d, t.data(), 16 * (size_t)n, hipMemcpyHostToDevice);
    if (e != hipSuccess) { dfree(d); return hip_fail(e, "h2d", __LINE__); }
    dfree(S.d_env);
    S.d_env = d; S.env_type = (uint32_t)type; S.env_w = width; S.env_h = height;
    S.env_version++;
    return LSG_OK;
}

// RRMatrix3x4::rotationByYawPitchRoll((yaw, 0, 0)) (RRMath.cpp:457-470): the reference's
// double expressions, rounded to float entry by entry
static void yaw_rotation(float yaw, float* m)
{
    const double pitch = 0, roll = 0;
    const double c1 = cos((double)yaw), c2 = cos(pitch), c3 = cos(roll);
    const double s1 = sin((double)yaw), s2 = sin(pitch), s3 = sin(roll);
    m[0] = (float)(c1 * c3 + s1 * s2 * s3); m[1] = (float)(c3 * s1 * s2 - c1 * s3); m[2] = (float)(c2 * s1);
    m[3] = (float)(c2 * s3); m[4] = (float)(c2 * c3); m[5] = (float)(-s2);
    m[6] = (float)(c1 * s2 * s3 - c3 * s1); m[7] = (float)(s1 * s3 + c1 * c3 * s2); m[8] = (float)(c1 * c2);
}

int lsg_scene_set_environment_blend(lsg_scene* s, int32_t type1, uint32_t width1, uint32_t height1, const float* rgb1,
                                    float angle0, float angle1, float blend_factor)
{
    if (!s) return fail(LSG_ERR_INVALID, "lsg_scene_set_environment_blend: null scene");
    if (!std::isfinite(angle0) || !std::isfinite(angle1) || !std::isfinite(blend_factor))
        return fail(LSG_ERR_INVALID, "lsg_scene_set_environment_blend: angles and blend factor must be finite");
    SceneImpl& S = s->impl;
    float4* d = nullptr;
    if (type1 != LSG_ENV_NONE) {
        if (type1 != LSG_ENV_CUBE && type1 != LSG_ENV_PANORAMA) return fail(LSG_ERR_INVALID, "lsg_scene_set_environment_blend: type %d", type1);
        if (!rgb1 || !width1 || !height1) return fail(LSG_ERR_INVALID, "lsg_scene_set_environment_blend: empty environment");
        if (type1 == LSG_ENV_CUBE && width1 != height1) return fail(LSG_ERR_INVALID, "lsg_scene_set_environment_blend: cube faces must be square");
        const uint64_t n = (uint64_t)width1 * height1 * (type1 == LSG_ENV_CUBE ? 6u : 1u);
        if (n > (1ull << 28)) return fail(LSG_ERR_LIMIT, "lsg_scene_set_environment_blend: %llu texels", (unsigned long long)n);
        std::vector<float4> t((size_t)n);
        for (size_t i = 0; i < (size_t)n; i++) t[i] = make_float4(rgb1[3 * i], rgb1[3 * i + 1], rgb1[3 * i + 2], 0.f);
        LSG_HIP(hipMalloc(&d, 16 * (size_t)n));
        hipError_t e = hipMemcpy(d, t.data(), 16 * (size_t)n, hipMemcpyHostToDevice);
        if (e != hipSuccess) { dfree(d); return hip_fail(e, "h2d", __LINE__); }
    }
    dfree(S.d_env1);
    S.d_env1 = d;
    S.env1_type = d ? (uint32_t)type1 : 0u; S.env1_w = d ? width1 : 0u; S.env1_h = d ? height1 : 0u;
    // RRBuffer::createEnvironmentBlend (RRBufferBlend.cpp:21-30): no blend without a second
    // environment and a rotation of the first
    S.env_blended = d != nullptr || angle0 != 0;
    yaw_rotation(angle0, S.env_rot);
    yaw_rotation(angle1, S.env_rot + 9);
    S.env_blend = blend_factor;
    S.env_version++;
    return LSG_OK;
}

int lsg_scene_set_tangents(lsg_scene* s, const float* tangents, const float* bitangents)
{
    if (!s) return fail(LSG_ERR_INVALID, "lsg_scene_set_tangents: null scene");
    if (!tangents != !bitangents) return fail(LSG_ERR_INVALID, "lsg_scene_set_tangents: give both tangents and bitangents, or neither");
    SceneImpl& S = s->impl;
    if (!tangents) {
        dfree(S.d_vtan); dfree(S.d_vbit); S.d_vtan = S.d_vbit = nullptr;
        const int e = scene_frames(s, nullptr);
        if (e) return e;
        LSG_HIP(hipStreamSynchronize(nullptr));
        return LSG_OK;
    }
    const int e = frames_alloc(S);
    if (e) return e;
    const size_t nb = 12 * (size_t)S.col->impl.nverts;
    if (nb) {
        LSG_HIP(hipMemcpy(S.d_vtan, tangents, nb, hipMemcpyHostToDevice));
        LSG_HIP(hipMemcpy(S.d_vbit, bitangents, nb, hipMemcpyHostToDevice));
    }
    return LSG_OK;
}

int lsg_gather_layers_device(lsg_scene* s, const lsg_texel* texels, uint32_t ntexels, const lsg_subtexel* subs, uint32_t nsubs,
                             const lsg_gather_params* p, float* const* layers, void* stream)
{
    if (!s || !p || !layers) return fail(LSG_ERR_INVALID, "lsg_gather_layers_device: null argument");
    bool any = false;
    for (int i = 0; i < 5; i++) any |= layers[i] != nullptr;
    if (!any) return fail(LSG_ERR_INVALID, "lsg_gather_layers_device: no output layer");
    if (!ntexels) return LSG_OK;
    if (!texels || !subs || !nsubs) return fail(LSG_ERR_INVALID, "lsg_gather_layers_device: null buffer");
    hipStream_t st = (hipStream_t)stream;
    int e = scene_frames(s, st);
    if (e) return e;
    GatherArgs A;
    int stack = 0;
    if ((e = gather_args(s, texels, ntexels, subs, p, layers[0], A, stack))) return e;
    for (int i = 0; i < 5; i++) A.layer[i] = (float4*)layers[i];
    A.dirs = (layers[1] || layers[2] || layers[3]) ? 1u : 0u;       // lightmap.cpp:820
    A.layered = 1;
    if ((e = walk_records(s, A))) return e;
    return launch_gather(A, stack, st);
}

static int gather_launch(lsg_scene* s, const lsg_texel* texels, uint32_t ntexels, const lsg_subtexel* subs, uint32_t nsubs,
                         const lsg_gather_params* p, float* out_rgba, void* stream, unsigned long long* stats)
{
    if (!s || !p) return fail(LSG_ERR_INVALID, "lsg_gather_device: null argument");
    if (!ntexels) return LSG_OK;
    if (!texels || !subs || !out_rgba || !nsubs) return fail(LSG_ERR_INVALID, "lsg_gather_device: null buffer");
    GatherArgs A;
    int stack = 0;
    int st = gather_args(s, texels, ntexels, subs, p, out_rgba, A, stack);
    if (st) return st;
    if (!stats) {
        if ((st = walk_records(s, A))) return st;
        return launch_gather(A, stack, (hipStream_t)stream);
    }
    A.stats = stats;
    const unsigned grid = (ntexels + 3) / 4;
    hipStream_t hs = (hipStream_t)stream;
    if ((st = scene_walk_scratch(s, ntexels, A))) return st;
    k_subtexel_walk<<<(ntexels + 255) / 256, 256, 0, hs>>>(texels, ntexels, subs, first_series(A), (WalkRec*)A.walk_buf);
    A.walk = (const WalkRec*)A.walk_buf;
    switch (features(A)) {
    case 0:
        if (stack <= 32) k_gather<32, true, 0><<<grid, 256, 0, hs>>>(A); else k_gather<64, true, 0><<<grid, 256, 0, hs>>>(A);
        break;
    case kFeatEnv:
        if (stack <= 32) k_gather<32, true, kFeatEnv><<<grid, 256, 0, hs>>>(A); else k_gather<64, true, kFeatEnv><<<grid, 256, 0, hs>>>(A);
        break;
    case kFeatMats:
        if (stack <= 32) k_gather<32, true, kFeatMats><<<grid, 256, 0, hs>>>(A); else k_gather<64, true, kFeatMats><<<grid, 256, 0, hs>>>(A);
        break;
    default:
        if (stack <= 32) k_gather<32, true, kFeatEnv | kFeatMats><<<grid, 256, 0, hs>>>(A);
        else k_gather<64, true, kFeatEnv | kFeatMats><<<grid, 256, 0, hs>>>(A);
        break;
    }
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

int lsg_gather_device(lsg_scene* s, const lsg_texel* texels, uint32_t ntexels, const lsg_subtexel* subs, uint32_t nsubs,
                      const lsg_gather_params* p, float* out_rgba, void* stream)
{
    return gather_launch(s, texels, ntexels, subs, nsubs, p, out_rgba, stream, nullptr);
}

int lsg_gather_stats(lsg_scene* s, const lsg_texel* texels, uint32_t ntexels, const lsg_subtexel* subs, uint32_t nsubs,
                     const lsg_gather_params* p, float* out_rgba, double* stats6)
{
    if (!stats6) return fail(LSG_ERR_INVALID, "lsg_gather_stats: null stats");
    unsigned long long* d = nullptr;
    LSG_HIP(hipMalloc(&d, 6 * sizeof(unsigned long long)));
    hipError_t e = hipMemset(d, 0, 6 * sizeof(unsigned long long));
    int st = e == hipSuccess ? gather_launch(s, texels, ntexels, subs, nsubs, p, out_rgba, nullptr, d) : hip_fail(e, "memset", __LINE__);
    unsigned long long h[6] = {0, 0, 0, 0, 0, 0};
    if (!st && (e = hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost)) != hipSuccess) st = hip_fail(e, "d2h", __LINE__);
    dfree(d);
    for (int k = 0; k < 6; k++) stats6[k] = (double)h[k];
    return st;
}

} // extern "C"
