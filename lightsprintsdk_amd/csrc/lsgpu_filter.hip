// lsgpu_filter.hip -- lightmap filter chain and writeback on gfx950.
//
// After the gather, RRSolver::updateLightmap (RRSolver/lightmap.cpp:852-876) runs
// on every RGBA temp lightmap: lightmapSmooth(sigma) -> lightmapGrowForBilinear-
// Interpolation -> lightmapGrow(spread) -> lightmapFillBackground(bg) ->
// copyElementsTo(user buffer).  These kernels restate RRBuffer.cpp:630-975 for a
// batch of lightmaps at once (one grid over every texel of every map), with the
// same float operation order, so a CPU statement reproduces them bit for bit.
// Alpha carries texelFlags/255 (gather.h:79-80) exactly as in the reference.
// FilteringParameters::wrap is honoured where the reference reads it, and the
// unwrap seams of an object (UnwrapSeams, RRBuffer.cpp:529-626, found on the host by
// lsg_unwrap_seams) are smoothed after every blur pass (k_seam_pass).
#include "lsgpu_internal.h"
#include "lsgpu_filter.h"
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace lsg {

enum { QUADRANT_X0Y0 = 1, QUADRANT_X1Y0 = 2, QUADRANT_X0Y1 = 4, QUADRANT_X1Y1 = 8,
       EDGE_X0 = 16, EDGE_X1 = 32, EDGE_Y0 = 64, EDGE_Y1 = 128 };   // gather.h:68-78

__device__ inline uint32_t to_flags(float a) { return (uint32_t)(a * 255 + 0.5f); }  // FLOAT_TO_TEXELFLAGS

// map lookup: texel t -> (map index) by binary search over the prefix of texel counts
__device__ inline uint32_t find_map(const FilterMap* maps, uint32_t nmaps, uint32_t t)
{
    uint32_t lo = 0, hi = nmaps;
    while (hi - lo > 1) { uint32_t mid = (lo + hi) >> 1; if (maps[mid].first <= t) lo = mid; else hi = mid; }
    return lo;
}

// lightmapSmooth step 1 (RRBuffer.cpp:659-697): per-texel blur flags from the edge bits
// of center and neighbours (with wrap, neighbours across the opposite border), then
// alpha := alpha>0 (:700-703)
__global__ __launch_bounds__(256) void k_smooth_prepare(const FilterMap* __restrict__ maps, uint32_t nmaps, uint32_t total,
                                                        const float4* __restrict__ rgba, float4* __restrict__ src, uint8_t* __restrict__ bf,
                                                        int wrap)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const FilterMap M = maps[find_map(maps, nmaps, t)];
    const uint32_t k = t - M.first, w = M.width, h = M.height;
    const int i = (int)(k % w), j = (int)(k / w);
    const float4* img = rgba + M.base;
    const uint32_t cf = to_flags(img[k].w);
    uint32_t bit = 1, flags = 0;
    for (int jj = -1; jj < 2; jj++)
        for (int ii = -1; ii < 2; ii++)
            if (ii || jj) {
                int x = i + ii, y = j + jj;
                if (wrap) { x = (int)(((uint32_t)x + w) % w); y = (int)(((uint32_t)y + h) % h); }
                if (x >= 0 && x < (int)w && y >= 0 && y < (int)h) {
                    const uint32_t nf = to_flags(img[x + y * w].w);
                    if ((ii >= 0 || ((cf & EDGE_X0) && (nf & EDGE_X1))) && (ii <= 0 || ((cf & EDGE_X1) && (nf & EDGE_X0)))
                        && (jj >= 0 || ((cf & EDGE_Y0) && (nf & EDGE_Y1))) && (jj <= 0 || ((cf & EDGE_Y1) && (nf & EDGE_Y0))))
                        flags |= bit;
                }
                bit *= 2;
            }
    bf[t] = (uint8_t)flags;
    const float4 c = img[k];
    src[t] = make_float4(c.x, c.y, c.z, c.w > 0 ? 1.f : 0.f);
}

__device__ inline float4 f4add(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ inline float4 f4mul(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }

// one blur pass (RRBuffer.cpp:716-732); neighbour addresses wrap as in the reference,
// neighbours outside the map have flag 0 and weight 0
__global__ __launch_bounds__(256) void k_smooth_pass(const FilterMap* __restrict__ maps, uint32_t nmaps, uint32_t total,
                                                     const float4* __restrict__ src, float4* __restrict__ dst,
                                                     const uint8_t* __restrict__ bf, float4 kernel)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const FilterMap M = maps[find_map(maps, nmaps, t)];
    const uint32_t k = t - M.first, w = M.width, h = M.height;
    const uint32_t i = k % w, j = k / w;
    const float4* s = src + M.first;
    float4 c = s[k];
    if (c.w != 0) {
        const uint32_t f = bf[t];
        const uint32_t im = (i - 1 + w) % w, ip = (i + 1) % w, jm = (j - 1 + h) % h, jp = (j + 1) % h;
        c = f4mul(c, kernel.x);
        c = f4add(c, f4mul(s[im + jm * w], (f & 1) ? kernel.z : 0.f));
        c = f4add(c, f4mul(s[i + jm * w], (f & 2) ? kernel.y : 0.f));
        c = f4add(c, f4mul(s[ip + jm * w], (f & 4) ? kernel.z : 0.f));
        c = f4add(c, f4mul(s[im + j * w], (f & 8) ? kernel.y : 0.f));
        c = f4add(c, f4mul(s[ip + j * w], (f & 16) ? kernel.y : 0.f));
        c = f4add(c, f4mul(s[im + jp * w], (f & 32) ? kernel.z : 0.f));
        c = f4add(c, f4mul(s[i + jp * w], (f & 64) ? kernel.y : 0.f));
        c = f4add(c, f4mul(s[ip + jp * w], (f & 128) ? kernel.z : 0.f));
        const float a = c.w;
        c = make_float4(c.x / a, c.y / a, c.z / a, c.w / a);
    }
    dst[M.first + k] = c;
}

// RRBuffer.cpp:780-784 copy back (alpha preserved), then
// lightmapGrowForBilinearInterpolation (RRBuffer.cpp:804-863): empty texels (alpha <=
// 0.002) take the weighted mean of neighbours whose quadrant flags face them.
// Texels filled here get alpha 0.001 = flags 0, so the reference's in-place update
// reads the same values as this out-of-place one.
__global__ __launch_bounds__(256) void k_smooth_finish(const FilterMap* __restrict__ maps, uint32_t nmaps, uint32_t total,
                                                       float4* __restrict__ rgba, const float4* __restrict__ sm)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const FilterMap M = maps[find_map(maps, nmaps, t)];
    const uint32_t k = t - M.first;
    const float4 s = sm[t];
    if (s.w > 0) { float4& o = rgba[M.base + k]; o = make_float4(s.x, s.y, s.z, o.w); }
}

__global__ __launch_bounds__(256) void k_grow_bilinear(const FilterMap* __restrict__ maps, uint32_t nmaps, uint32_t total,
                                                       const float4* __restrict__ rgba, float4* __restrict__ out, int wrap)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const FilterMap M = maps[find_map(maps, nmaps, t)];
    const uint32_t k = t - M.first, w = M.width, h = M.height;
    const int i = (int)(k % w), j = (int)(k / w);
    const float4* img = rgba + M.base;
    float4 c = img[k];
    if (!(c.w > 0.002f)) {
        float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int jj = -1; jj < 2; jj++)
            for (int ii = -1; ii < 2; ii++) {
                int x = i + ii, y = j + jj;
                if (wrap) { x = (int)(((uint32_t)x + w) % w); y = (int)(((uint32_t)y + h) % h); }
                else if (x < 0 || x >= (int)w || y < 0 || y >= (int)h) continue;
                const float4 n = img[x + y * w];
                const uint32_t tf = to_flags(n.w);
                if ((ii < 1 && jj < 1 && (tf & QUADRANT_X1Y1)) || (ii > -1 && jj > -1 && (tf & QUADRANT_X0Y0))
                    || (ii < 1 && jj > -1 && (tf & QUADRANT_X1Y0)) || (ii > -1 && jj < 1 && (tf & QUADRANT_X0Y1))) {
                    const float wgt = (ii && jj) ? 1.f : 1.6f;
                    sum = f4add(sum, make_float4(n.x * wgt, n.y * wgt, n.z * wgt, 1.f * wgt));
                }
            }
        if (sum.w != 0) c = make_float4(sum.x / sum.w, sum.y / sum.w, sum.z / sum.w, 0.001f);
    }
    out[t] = c;
}

// lightmapGrow init (RRBuffer.cpp:889-893)
__global__ __launch_bounds__(256) void k_grow_init(uint32_t total, const float4* __restrict__ img, float4* __restrict__ g)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const float4 c = img[t];
    g[t] = c.w > 0 ? make_float4(c.x, c.y, c.z, 1.f) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// lightmapGrow by frontiers.  The reference's step (RRBuffer.cpp:897-947) changes exactly
// the empty texels that have a filled neighbour in the previous state, and those are the
// still-empty texels that count a texel filled one step earlier among their neighbours.
// So each step visits only its frontier: every frontier texel's value is computed from
// the previous state (1.6 x the 4 edge neighbours + the 4 corner neighbours, divided by
// its alpha) and written, and the empty texels that count it as a neighbour are queued
// once (mark) as the next frontier.  Grids are fixed and sizes live in device memory, so
// a step captures into a graph; a step with an empty frontier does nothing, as the
// reference stops there.
constexpr unsigned kFrontBlocks = 2048;

__device__ inline void grow_neighbours(const FilterMap& M, uint32_t k, int& x, int& y, int& xp, int& xm, int& yp, int& ym)
{
    const int w = (int)M.width, h = (int)M.height;
    x = (int)(k % M.width); y = (int)(k / M.width);
    xp = x + 1 < w - 1 ? x + 1 : w - 1; xm = x - 1 > 0 ? x - 1 : 0;
    yp = y + 1 < h - 1 ? y + 1 : h - 1; ym = y - 1 > 0 ? y - 1 : 0;
}

// The 8 neighbours lightmapGrow averages, in its order: x+1, x-1, y+1, y-1, then
// (+1,+1), (+1,-1), (-1,+1), (-1,-1).  Without wrap they are clamped to the map
// (RRBuffer.cpp:930-941).  With wrap they are flat texel numbers, k plus the offset in
// 32-bit unsigned arithmetic (wrapping below 0) reduced modulo the texel count
// (RRBuffer.cpp:912-922): a row's end continues on the next row, and where the count
// does not divide 2^32 a sum that wrapped below 0 lands elsewhere than k + offset + count.
__device__ inline void grow_nbr8(const FilterMap& M, uint32_t k, int wrap, uint32_t* q)
{
    const uint32_t w = M.width;
    if (wrap) {
        const uint32_t size = M.width * M.height;
        q[0] = (k + 1u) % size; q[1] = (k - 1u) % size; q[2] = (k + w) % size; q[3] = (k - w) % size;
        q[4] = (k + 1u + w) % size; q[5] = (k + 1u - w) % size; q[6] = (k - 1u + w) % size; q[7] = (k - 1u - w) % size;
    } else {
        int x, y, xp, xm, yp, ym;
        grow_neighbours(M, k, x, y, xp, xm, yp, ym);
        q[0] = (uint32_t)(xp + y * (int)w); q[1] = (uint32_t)(xm + y * (int)w);
        q[2] = (uint32_t)(x + yp * (int)w); q[3] = (uint32_t)(x + ym * (int)w);
        q[4] = (uint32_t)(xp + yp * (int)w); q[5] = (uint32_t)(xp + ym * (int)w);
        q[6] = (uint32_t)(xm + yp * (int)w); q[7] = (uint32_t)(xm + ym * (int)w);
    }
}

// Texels that may count k among their neighbours, queued when k is filled.  Without wrap
// the clamped neighbourhood is symmetric: grow_nbr8 itself.  With wrap, the texels
// k - offset modulo the count (taken properly, the offsets come in +- pairs): exact for
// every texel i > width, whose sums never wrap below 0; texels i <= width are checked
// after each step by k_front_wrap_head.
__device__ inline void grow_inv8(const FilterMap& M, uint32_t k, int wrap, uint32_t* q)
{
    if (!wrap) { grow_nbr8(M, k, 0, q); return; }
    const int64_t size = (int64_t)M.width * M.height, w = M.width;
    const int64_t off[8] = {1, -1, w, -w, 1 + w, 1 - w, -1 + w, -1 - w};
    for (int j = 0; j < 8; j++) {
        int64_t v = ((int64_t)k + off[j]) % size;
        if (v < 0) v += size;
        q[j] = (uint32_t)v;
    }
}

// wave-aggregated append of t to list (count at *cnt) where pred holds
__device__ inline void append(bool pred, uint32_t t, uint32_t* list, uint32_t* cnt)
{
    const uint64_t m = __ballot(pred);
    if (!m) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, (int)leader);
    if (pred) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = t;
}

__global__ __launch_bounds__(256) void k_front_seed(const FilterMap* __restrict__ maps, uint32_t nmaps, uint32_t total,
                                                    const float4* __restrict__ g, uint32_t* __restrict__ mark,
                                                    uint32_t* __restrict__ front, uint32_t* __restrict__ cnt, int wrap)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    bool seed = false;
    if (t < total) {
        mark[t] = 0u;
        if (g[t].w == 0) {
            const FilterMap M = maps[find_map(maps, nmaps, t)];
            uint32_t q[8];
            grow_nbr8(M, t - M.first, wrap, q);
            const float4* s = g + M.first;
            for (int j = 0; j < 8; j++) seed = seed || s[q[j]].w != 0;
            if (seed) mark[t] = 1u;                 // step 0's frontier
        }
    }
    append(seed, t, front, cnt);
}

// One lightmapGrow step over its frontier in one launch.  Step q's frontier texels carry
// mark q + 1 (the seeds 1), texels queued for step q + 1 get q + 2, and 0 means never
// queued.  A frontier texel's value is computed from the previous state: its neighbours
// in the same frontier count as empty (their previous state, zero) whether or not another
// thread has written them yet, every other neighbour is unchanged during the step.  A
// texel with a filled neighbour is written at once (with wrap a queued texel may have
// none, and stays empty); then the still-empty, never-queued texels that count it as a
// neighbour are queued for the next step (the mark decides, so texels written in this
// step are never queued).  Counts rotate over three words: this step reads cur, appends
// to next and clears the one the step after will append to.
__global__ __launch_bounds__(256) void k_front_step(const FilterMap* __restrict__ maps, uint32_t nmaps, float4* __restrict__ g,
                                                    const uint32_t* __restrict__ front, const uint32_t* __restrict__ cnt,
                                                    uint32_t* __restrict__ mark, uint32_t cur_mark, uint32_t* __restrict__ next,
                                                    uint32_t* __restrict__ next_cnt, uint32_t* __restrict__ clear_cnt,
                                                    uint32_t* __restrict__ log, int wrap)
{
    const uint32_t n = *cnt;
    if (blockIdx.x == 0 && threadIdx.x == 0) { *clear_cnt = 0u; *log = n; }
    const uint32_t stride = gridDim.x * 256;
    // whole waves walk the loop together so that append's ballot sees every lane
    const uint32_t rounds = (n + stride - 1) / stride;
    for (uint32_t r = 0; r < rounds; r++) {
        const uint32_t i = r * stride + blockIdx.x * 256 + threadIdx.x;
        bool filled = false;
        uint32_t first = 0, t = 0;
        uint32_t q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (i < n) {
            t = front[i];
            const FilterMap M = maps[find_map(maps, nmaps, t)];
            first = M.first;
            grow_nbr8(M, t - first, wrap, q);
            auto at = [&](uint32_t k) {
                const uint32_t p = first + k;
                return mark[p] == cur_mark ? make_float4(0.f, 0.f, 0.f, 0.f) : g[p];
            };
            const float4 e = f4add(f4add(f4add(at(q[0]), at(q[1])), at(q[2])), at(q[3]));
            const float4 dg = f4add(f4add(f4add(at(q[4]), at(q[5])), at(q[6])), at(q[7]));
            const float4 d = f4add(f4mul(e, 1.6f), dg);
            const float a = d.w;
            if (a != 0) {
                g[t] = make_float4(d.x / a, d.y / a, d.z / a, d.w / a);
                filled = true;
            }
            if (filled) grow_inv8(M, t - first, wrap, q);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t p = first + q[j];
            const bool add = filled && g[p].w == 0 && atomicCAS(&mark[p], 0u, cur_mark + 1u) == 0u;
            append(add, p, next, next_cnt);
        }
    }
}

// With wrap, after step cur_mark - 1: the empty texels k <= width of each map whose
// texel count does not divide 2^32 (their neighbour sums may wrap below 0) are queued
// for the next step when a neighbour was filled in this step.  One workgroup per map.
__global__ __launch_bounds__(256) void k_front_wrap_head(const FilterMap* __restrict__ maps, const float4* __restrict__ g,
                                                         uint32_t* __restrict__ mark, uint32_t cur_mark, uint32_t* __restrict__ next,
                                                         uint32_t* __restrict__ next_cnt)
{
    const FilterMap M = maps[blockIdx.x];
    const uint32_t size = M.width * M.height;
    if ((0x100000000ull % size) == 0) return;   // k + offset modulo size is the proper neighbour
    const uint32_t nh = M.width + 1 < size ? M.width + 1 : size;
    const uint32_t rounds = (nh + 255) / 256;
    for (uint32_t r = 0; r < rounds; r++) {
        const uint32_t k = r * 256 + threadIdx.x, t = M.first + k;
        bool add = false;
        if (k < nh && g[t].w == 0 && mark[t] != cur_mark + 1u) {
            uint32_t q[8];
            grow_nbr8(M, k, 1, q);
            for (int j = 0; j < 8; j++) {
                const uint32_t p = M.first + q[j];
                add = add || (mark[p] == cur_mark && g[p].w != 0);
            }
            if (add) mark[t] = cur_mark + 1u;
        }
        append(add, t, next, next_cnt);
    }
}

__global__ void k_front_changed(const uint32_t* __restrict__ cnt, uint32_t* __restrict__ changed)
{
    // after a fixed number of steps: did the last one still leave a frontier to fill?
    if (threadIdx.x == 0) changed[0] = *cnt != 0u ? 1u : 0u;
}

// lightmapGrow copy back (RRBuffer.cpp:952-956) + lightmapFillBackground (:964-975) +
// copyElementsTo into an RGBF buffer (RRBuffer.cpp:105)
__global__ __launch_bounds__(256) void k_finish(const FilterMap* __restrict__ maps, uint32_t nmaps, uint32_t total,
                                                const float4* __restrict__ img, const float4* __restrict__ g, float4 bg,
                                                float* __restrict__ rgbf, float4* __restrict__ rgba_out)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const FilterMap M = maps[find_map(maps, nmaps, t)];
    const uint32_t k = t - M.first;
    float4 c = img[t];
    if (!(c.w > 0) && g) { const float4 q = g[t]; if (q.w > 0) c = make_float4(q.x, q.y, q.z, 0.001f); }
    if (c.w <= 0) c = bg;
    if (rgba_out) rgba_out[M.base + k] = c;
    if (rgbf) {
        float* o = rgbf + 3 * ((size_t)M.out_base + k);
        o[0] = c.x; o[1] = c.y; o[2] = c.z;
    }
}

// ---- unwrap seams ----------------------------------------------------------------------
// The reference walks each seam after a blur pass (RRBuffer.cpp:734-777): 1001 points
// along both edges; a point in [0,1)^2 in both unwraps whose two texels are used adds
// both texels' values to a running sum for the texel under each edge, and when the
// texel under an edge changes, the finished run's texel takes its sum divided by the
// sum's alpha (RGB; alpha stays 1).  Seams follow one another (a seam reads what the
// ones before it wrote), so one workgroup per map takes its seams in order:
//   1. the 64 lanes place the points and compact the used ones in order (LDS);
//   2. run starts along each edge are compacted the same way;
//   3. if no texel is the texel of two runs (of either edge), no texel a run reads is
//      written before that run ends, so every run reads the values the seam started
//      from: each run is summed in point order by one lane, and written after all sums;
//   4. otherwise one lane takes the seam's points one at a time, exactly as the
//      reference does, reading and writing the map in memory.
// Every texel a point reads has alpha 1 or 0 and keeps it, so whether a point is used
// does not depend on what the seam writes.
constexpr uint32_t kSeamPoints = 1001;

__device__ inline uint32_t lanes_below(uint64_t m) { return (uint32_t)__popcll(m & ((1ull << (threadIdx.x & 63u)) - 1ull)); }

__device__ inline void seam_put(float4* d, uint32_t x, float4 acc)
{
    const float4 o = d[x];
    d[x] = make_float4(acc.x / acc.w, acc.y / acc.w, acc.z / acc.w, o.w);
}

__global__ __launch_bounds__(64) void k_seam_pass(const FilterMap* __restrict__ maps, SeamSet A, float4* __restrict__ dst)
{
    __shared__ uint32_t p1[kSeamPoints], p2[kSeamPoints];   // texels of the used points, in point order
    __shared__ uint32_t rs[2 * kSeamPoints];                // run starts: edge 1's, then (from kSeamPoints) edge 2's
    __shared__ float4 sum[2 * kSeamPoints];
    const uint32_t lane = threadIdx.x;
    const uint32_t m = A.with[blockIdx.x];
    const FilterMap M = maps[m];
    float4* d = dst + M.first;
    const uint32_t w = M.width;
    const float fw = (float)M.width, fh = (float)M.height;
    for (uint32_t s = A.begin[m]; s < A.begin[m + 1]; s++) {
        const lsg_unwrap_seam S = A.seams[s];
        uint32_t K = 0;
        for (uint32_t c = 0; c < kSeamPoints; c += 64) {
            const uint32_t i = c + lane;
            bool ok = false;
            uint32_t i1 = 0, i2 = 0;
            if (i < kSeamPoints) {
                const float t = (float)i * 0.001f, om = 1.f - t;
                const float u1 = S.edge1[0] * om + S.edge1[2] * t, v1 = S.edge1[1] * om + S.edge1[3] * t;
                const float u2 = S.edge2[0] * om + S.edge2[2] * t, v2 = S.edge2[1] * om + S.edge2[3] * t;
                if (u1 >= 0 && u1 < 1 && v1 >= 0 && v1 < 1 && u2 >= 0 && u2 < 1 && v2 >= 0 && v2 < 1) {
                    // u < 1 keeps u * width below width in binary32, so the texels are in the map
                    i1 = (uint32_t)(u1 * fw) + (uint32_t)(v1 * fh) * w;
                    i2 = (uint32_t)(u2 * fw) + (uint32_t)(v2 * fh) * w;
                    ok = d[i1].w != 0 && d[i2].w != 0;
                }
            }
            const uint64_t b = __ballot(ok);
            if (ok) { const uint32_t p = K + lanes_below(b); p1[p] = i1; p2[p] = i2; }
            K += (uint32_t)__popcll(b);
        }
        __syncthreads();
        if (K) {
            uint32_t R1 = 0, R2 = 0;
            for (uint32_t c = 0; c < K; c += 64) {
                const uint32_t k = c + lane;
                const bool a = k < K && (k == 0 || p1[k] != p1[k - 1]);
                const bool e = k < K && (k == 0 || p2[k] != p2[k - 1]);
                const uint64_t ba = __ballot(a), be = __ballot(e);
                if (a) rs[R1 + lanes_below(ba)] = k;
                if (e) rs[kSeamPoints + R2 + lanes_below(be)] = k;
                R1 += (uint32_t)__popcll(ba);
                R2 += (uint32_t)__popcll(be);
            }
            __syncthreads();
            const uint32_t R = R1 + R2;
            auto run_start = [&](uint32_t r) { return r < R1 ? rs[r] : rs[kSeamPoints + r - R1]; };
            auto run_end = [&](uint32_t r) { return r + 1 < R1 ? rs[r + 1] : (r < R1 ? K : (r + 1 < R ? rs[kSeamPoints + r + 1 - R1] : K)); };
            auto run_texel = [&](uint32_t r) { return r < R1 ? p1[rs[r]] : p2[rs[kSeamPoints + r - R1]]; };
            bool clash = false;
            for (uint32_t a = lane; a < R && !clash; a += 64) {
                const uint32_t ta = run_texel(a);
                for (uint32_t b = a + 1; b < R; b++)
                    if (run_texel(b) == ta) { clash = true; break; }
            }
            if (!__ballot(clash)) {
                for (uint32_t r = lane; r < R; r += 64) {
                    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
                    const uint32_t e = run_end(r);
                    for (uint32_t k = run_start(r); k < e; k++) acc = f4add(acc, f4add(d[p1[k]], d[p2[k]]));
                    sum[r] = acc;
                }
                __syncthreads();
                for (uint32_t r = lane; r < R; r += 64) seam_put(d, run_texel(r), sum[r]);
            } else if (lane == 0) {
                uint32_t l1 = 0xffffffffu, l2 = 0xffffffffu;
                float4 a1 = make_float4(0.f, 0.f, 0.f, 0.f), a2 = a1;
                for (uint32_t k = 0; k < K; k++) {
                    const uint32_t i1 = p1[k], i2 = p2[k];
                    if (i1 != l1) { if (l1 != 0xffffffffu) seam_put(d, l1, a1); l1 = i1; a1 = make_float4(0.f, 0.f, 0.f, 0.f); }
                    if (i2 != l2) { if (l2 != 0xffffffffu) seam_put(d, l2, a2); l2 = i2; a2 = make_float4(0.f, 0.f, 0.f, 0.f); }
                    const float4 both = f4add(d[i1], d[i2]);
                    a1 = f4add(a1, both);
                    a2 = f4add(a2, both);
                }
                seam_put(d, l1, a1);
                seam_put(d, l2, a2);
            }
        }
        __syncthreads();   // a workgroup barrier orders the lanes' writes for the next seam (one CU)
    }
}

static inline unsigned nblk(uint32_t n) { return (n + 255) / 256; }

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

int filter_run(const FilterMap* d_maps, uint32_t nmaps, uint32_t total, float4* rgba, float* rgbf, const lsg_filter_params& fp,
               FilterScratch& S, hipStream_t s, bool rgba_out, int fixed_steps, uint32_t* steps_used, const SeamSet* seams)
{
    if (!total) return LSG_OK;
    // lightmapSmooth (RRBuffer.cpp:630-802)
    float sigma = fp.smoothing_amount;
    if (sigma < 0) return fail(LSG_ERR_INVALID, "smoothing_amount < 0");
    if (sigma > 0) {
        k_smooth_prepare<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, rgba, S.a, S.flags, fp.wrap);
        if (sigma > 10) sigma = 10;
        float sigma2sum = sigma * sigma;
        float4* src = S.a; float4* dst = S.b;
        while (sigma2sum > 0) {
            const float sigma2 = sigma2sum < 0.34f ? sigma2sum : 0.34f;
            const float4 kernel = make_float4(1.f, expf(-0.5f / sigma2), expf(-1 / sigma2), expf(-2 / sigma2));
            sigma2sum -= sigma2;
            k_smooth_pass<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, src, dst, S.flags, kernel);
            if (seams && seams->nwith) k_seam_pass<<<seams->nwith, 64, 0, s>>>(d_maps, *seams, dst);
            std::swap(src, dst);
        }
        k_smooth_finish<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, rgba, src);
    }
    // lightmapGrowForBilinearInterpolation
    k_grow_bilinear<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, rgba, S.c, fp.wrap);
    // lightmapGrow(spread)
    const float4* g = nullptr;
    uint32_t steps = (uint32_t)(fp.spread_foreground > 0 ? fp.spread_foreground : 0);
    if (steps_used) *steps_used = 0;
    if (steps) {
        // frontier sizes rotate over S.changed[1..3] (step q reads [1 + q % 3]); step q reads
        // front[q & 1] and queues into the other list
        k_grow_init<<<nblk(total), 256, 0, s>>>(total, S.c, S.a);
        k_zero_u32<<<1, 256, 0, s>>>(S.changed, 20);
        k_front_seed<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, S.a, S.mark, S.front[0], S.changed + 1, fp.wrap);
        // S.changed[4 + q % 16] = frontier size of step q (the host reads a chunk at a time)
        auto step = [&](uint32_t q) {
            const uint32_t c = q & 1u;
            k_front_step<<<kFrontBlocks, 256, 0, s>>>(d_maps, nmaps, S.a, S.front[c], S.changed + 1 + q % 3u, S.mark, q + 1u,
                                                      S.front[1 - c], S.changed + 1 + (q + 1u) % 3u,
                                                      S.changed + 1 + (q + 2u) % 3u, S.changed + 4 + (q & 15u), fp.wrap);
            if (fp.wrap)
                k_front_wrap_head<<<nmaps, 256, 0, s>>>(d_maps, S.a, S.mark, q + 1u, S.front[1 - c], S.changed + 1 + (q + 1u) % 3u);
        };
        if (fixed_steps >= 0) {
            // graph-capturable form: a fixed number of steps, no host synchronisation,
            // S.changed[0] = whether a frontier is still left after the last one
            const uint32_t n = std::min<uint32_t>((uint32_t)fixed_steps, steps);
            for (uint32_t q = 0; q < n; q++) step(q);
            k_front_changed<<<1, 64, 0, s>>>(S.changed + 1 + n % 3u, S.changed);
            if (steps_used) *steps_used = n;
        } else if (!S.host_log) {
            // scratch without pinned memory or events (the allocation-free entry point):
            // each chunk of 16 steps is followed by a copy of its frontier sizes into a
            // stack array and a stream synchronisation
            constexpr uint32_t kChunk = 16;
            uint32_t done = steps;
            for (uint32_t b = 0; b < steps; b += kChunk) {
                const uint32_t e = std::min(steps, b + kChunk);
                for (uint32_t q = b; q < e; q++) step(q);
                uint32_t log[kChunk];
                LSG_HIP(hipMemcpyAsync(log, S.changed + 4, sizeof(log), hipMemcpyDeviceToHost, s));
                LSG_HIP(hipStreamSynchronize(s));
                bool stop = false;
                for (uint32_t q = b; q < e; q++)
                    if (!log[q & 15u]) { done = q + 1; stop = true; break; }
                if (stop) break;
            }
            if (steps_used) *steps_used = done;
        } else {
            // a step with an empty frontier changes nothing: the reference stops there
            // (RRBuffer.cpp:948-949).  Steps go out in chunks of 16, each followed by a copy
            // of its frontier sizes into pinned host memory and an event; the host checks
            // chunk k while chunk k + 1 already runs, so the GPU does not drain between
            // chunks (steps after an empty one do nothing)
            constexpr uint32_t kChunk = 16;
            const uint32_t nchunks = (steps + kChunk - 1) / kChunk;
            auto launch = [&](uint32_t k) -> hipError_t {
                const uint32_t b = k * kChunk, e = std::min(steps, b + kChunk);
                for (uint32_t q = b; q < e; q++) step(q);
                hipError_t r = hipMemcpyAsync(S.host_log + (k & 1u) * kChunk, S.changed + 4, 4 * kChunk, hipMemcpyDeviceToHost, s);
                if (r == hipSuccess) r = hipEventRecord(S.ev[k & 1u], s);
                return r;
            };
            uint32_t done = steps, k = 0;
            LSG_HIP(launch(0));
            for (; k < nchunks; k++) {
                if (k + 1 < nchunks) LSG_HIP(launch(k + 1));
                LSG_HIP(hipEventSynchronize(S.ev[k & 1u]));
                const uint32_t b = k * kChunk, e = std::min(steps, b + kChunk);
                bool stop = false;
                for (uint32_t q = b; q < e; q++)
                    if (!S.host_log[(k & 1u) * kChunk + (q & 15u)]) { done = q + 1; stop = true; break; }
                if (stop) break;
            }
            // the chunk launched ahead must finish its copy before the buffers are reused
            if (k + 1 < nchunks) LSG_HIP(hipEventSynchronize(S.ev[(k + 1) & 1u]));
            if (steps_used) *steps_used = done;
        }
        g = S.a;
    }
    k_finish<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, S.c, g, make_float4(fp.background[0], fp.background[1], fp.background[2], fp.background[3]),
                                         rgbf, rgba_out ? rgba : nullptr);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

int scratch_alloc(FilterScratch& S, uint32_t total)
{
    hipError_t e;
    const size_t n = total ? total : 1;
    if ((e = hipMalloc(&S.a, 16 * n)) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    if ((e = hipMalloc(&S.b, 16 * n)) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    if ((e = hipMalloc(&S.c, 16 * n)) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    if ((e = hipMalloc(&S.flags, n)) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    if ((e = hipMalloc(&S.changed, 4 * 32)) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    if ((e = hipMalloc(&S.front[0], 4 * n)) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    if ((e = hipMalloc(&S.front[1], 4 * n)) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    if ((e = hipMalloc(&S.mark, 4 * n)) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    if ((e = hipHostMalloc((void**)&S.host_log, 4 * 32, hipHostMallocDefault)) != hipSuccess) return hip_fail(e, "hipHostMalloc", __LINE__);
    for (int i = 0; i < 2; i++)
        if ((e = hipEventCreateWithFlags(&S.ev[i], hipEventDisableTiming)) != hipSuccess) return hip_fail(e, "hipEventCreate", __LINE__);
    return LSG_OK;
}

void scratch_free(FilterScratch& S)
{
    if (S.a) (void)hipFree(S.a);
    if (S.b) (void)hipFree(S.b);
    if (S.c) (void)hipFree(S.c);
    if (S.flags) (void)hipFree(S.flags);
    if (S.changed) (void)hipFree(S.changed);
    for (int i = 0; i < 2; i++) if (S.front[i]) (void)hipFree(S.front[i]);
    if (S.mark) (void)hipFree(S.mark);
    if (S.host_log) (void)hipHostFree(S.host_log);
    for (int i = 0; i < 2; i++) if (S.ev[i]) (void)hipEventDestroy(S.ev[i]);
    S = FilterScratch();
}

void seams_free(SeamStore& S)
{
    if (S.d_seams) (void)hipFree(S.d_seams);
    if (S.d_begin) (void)hipFree(S.d_begin);
    if (S.d_with) (void)hipFree(S.d_with);
    const uint32_t v = S.version;
    S = SeamStore();
    S.version = v + 1;
}

int seams_set(SeamStore& S, uint32_t nmaps, const lsg_unwrap_seam* seams, const uint32_t* begin)
{
    seams_free(S);
    if (!seams) return LSG_OK;
    if (!begin) return fail(LSG_ERR_INVALID, "seams without their per-map offsets");
    if (begin[0] != 0) return fail(LSG_ERR_INVALID, "seam offsets start at %u, not 0", begin[0]);
    std::vector<uint32_t> with;
    for (uint32_t m = 0; m < nmaps; m++) {
        if (begin[m + 1] < begin[m]) return fail(LSG_ERR_INVALID, "seam offsets decrease at lightmap %u", m);
        if (begin[m + 1] > begin[m]) with.push_back(m);
    }
    const uint32_t n = begin[nmaps];
    if (!n) return LSG_OK;
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 4; k++)
            if (!std::isfinite(seams[i].edge1[k]) || !std::isfinite(seams[i].edge2[k]))
                return fail(LSG_ERR_INVALID, "seam %u has a coordinate that is not finite", i);
    hipError_t e;
    if ((e = hipMalloc(&S.d_seams, sizeof(lsg_unwrap_seam) * n)) != hipSuccess) { seams_free(S); return hip_fail(e, "hipMalloc", __LINE__); }
    if ((e = hipMalloc(&S.d_begin, 4 * ((size_t)nmaps + 1))) != hipSuccess) { seams_free(S); return hip_fail(e, "hipMalloc", __LINE__); }
    if ((e = hipMalloc(&S.d_with, 4 * with.size())) != hipSuccess) { seams_free(S); return hip_fail(e, "hipMalloc", __LINE__); }
    if ((e = hipMemcpy(S.d_seams, seams, sizeof(lsg_unwrap_seam) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(S.d_begin, begin, 4 * ((size_t)nmaps + 1), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(S.d_with, with.data(), 4 * with.size(), hipMemcpyHostToDevice)) != hipSuccess) {
        seams_free(S);
        return hip_fail(e, "h2d", __LINE__);
    }
    S.view.seams = S.d_seams; S.view.begin = S.d_begin; S.view.with = S.d_with; S.view.nwith = (uint32_t)with.size();
    return LSG_OK;
}

int maps_upload(const lsg_lightmap* maps, uint32_t nmaps, FilterMap** d_maps, uint32_t* total)
{
    std::vector<FilterMap> fm(nmaps);
    uint32_t t = 0;
    for (uint32_t m = 0; m < nmaps; m++) {
        if (!maps[m].width || !maps[m].height) return fail(LSG_ERR_INVALID, "lightmap %u has zero size", m);
        fm[m].first = t; fm[m].base = maps[m].rgba_base; fm[m].out_base = maps[m].rgbf_base;
        fm[m].width = maps[m].width; fm[m].height = maps[m].height;
        uint64_t nt = (uint64_t)t + (uint64_t)maps[m].width * maps[m].height;
        if (nt >= (1ull << 32)) return fail(LSG_ERR_LIMIT, "more than 2^32 texels");
        t = (uint32_t)nt;
    }
    hipError_t e;
    if ((e = hipMalloc(d_maps, sizeof(FilterMap) * (nmaps ? nmaps : 1))) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    if (nmaps && (e = hipMemcpy(*d_maps, fm.data(), sizeof(FilterMap) * nmaps, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "h2d", __LINE__);
    *total = t;
    return LSG_OK;
}

} // namespace lsg

struct lsg_filter_plan {
    lsg::FilterMap* d_maps = nullptr;
    uint32_t nmaps = 0, total = 0;
    lsg::FilterScratch S;
    lsg::SeamStore seams;
};

using namespace lsg;

extern "C" int lsg_filter_plan_create(const lsg_lightmap* maps, uint32_t nmaps, lsg_filter_plan** out)
{
    if (!out) return fail(LSG_ERR_INVALID, "lsg_filter_plan_create: out is null");
    *out = nullptr;
    if (!maps && nmaps) return fail(LSG_ERR_INVALID, "lsg_filter_plan_create: null maps");
    lsg_filter_plan* f = new (std::nothrow) lsg_filter_plan;
    if (!f) return fail(LSG_ERR_OOM, "lsg_filter_plan_create: host allocation");
    int st = maps_upload(maps, nmaps, &f->d_maps, &f->total);
    if (!st) st = scratch_alloc(f->S, f->total);
    if (st) { scratch_free(f->S); if (f->d_maps) (void)hipFree(f->d_maps); delete f; return st; }
    f->nmaps = nmaps;
    *out = f;
    return LSG_OK;
}

extern "C" void lsg_filter_plan_destroy(lsg_filter_plan* f)
{
    if (!f) return;
    scratch_free(f->S);
    seams_free(f->seams);
    if (f->d_maps) (void)hipFree(f->d_maps);
    delete f;
}

extern "C" int lsg_filter_plan_run(lsg_filter_plan* f, float* rgba, float* rgbf, const lsg_filter_params* params, void* stream)
{
    if (!f || !rgba || !params) return fail(LSG_ERR_INVALID, "lsg_filter_plan_run: null argument");
    int st = filter_run(f->d_maps, f->nmaps, f->total, (float4*)rgba, rgbf, *params, f->S, (hipStream_t)stream, true, -1, nullptr,
                        &f->seams.view);
    if (!st) { hipError_t e = hipStreamSynchronize((hipStream_t)stream); if (e != hipSuccess) st = hip_fail(e, "sync", __LINE__); }
    return st;
}

extern "C" int lsg_filter_plan_set_seams(lsg_filter_plan* f, const lsg_unwrap_seam* seams, const uint32_t* begin)
{
    if (!f) return fail(LSG_ERR_INVALID, "lsg_filter_plan_set_seams: null plan");
    return seams_set(f->seams, f->nmaps, seams, begin);
}

extern "C" int lsg_filter_lightmaps(const lsg_lightmap* maps, uint32_t nmaps, float* rgba, float* rgbf,
                                    const lsg_filter_params* params, void* stream)
{
    if (!maps || !rgba || !params) return fail(LSG_ERR_INVALID, "lsg_filter_lightmaps: null argument");
    FilterMap* d_maps = nullptr;
    uint32_t total = 0;
    int st = maps_upload(maps, nmaps, &d_maps, &total);
    if (st) { if (d_maps) (void)hipFree(d_maps); return st; }
    FilterScratch S;
    st = scratch_alloc(S, total);
    if (!st) st = filter_run(d_maps, nmaps, total, (float4*)rgba, rgbf, *params, S, (hipStream_t)stream, true, -1, nullptr);
    if (!st) { hipError_t e = hipStreamSynchronize((hipStream_t)stream); if (e != hipSuccess) st = hip_fail(e, "sync", __LINE__); }
    scratch_free(S);
    (void)hipFree(d_maps);
    return st;
}

// ---- allocation-free filter chain --------------------------------------------------------
// Scratch of lsg_filter_lightmaps_noalloc: one static device array per device, laid out as
// three float4 images, two frontier lists and the mark array (kNoAllocTexels entries each),
// the blur flags, the step counters, then the FilterMap table and the caller's lightmap list
// staged for the kernel that builds it.  On the host the array's name is not a device
// address: every use goes through the per-device pointer hipGetSymbolAddress returns.
namespace lsg {
constexpr size_t kNoAllocTexels = LSG_FILTER_NOALLOC_MAX_TEXELS;
constexpr size_t kNoAllocMaps = LSG_FILTER_NOALLOC_MAX_MAPS;
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kOffA = 0;
constexpr size_t kOffB = kOffA + 16 * kNoAllocTexels;
constexpr size_t kOffC = kOffB + 16 * kNoAllocTexels;
constexpr size_t kOffFront0 = kOffC + 16 * kNoAllocTexels;
constexpr size_t kOffFront1 = kOffFront0 + 4 * kNoAllocTexels;
constexpr size_t kOffMark = kOffFront1 + 4 * kNoAllocTexels;
constexpr size_t kOffFlags = kOffMark + 4 * kNoAllocTexels;
constexpr size_t kOffChanged = align256(kOffFlags + kNoAllocTexels);
constexpr size_t kOffMaps = align256(kOffChanged + 4 * 32);
constexpr size_t kOffStage = align256(kOffMaps + sizeof(FilterMap) * kNoAllocMaps);
constexpr size_t kNoAllocBytes = align256(kOffStage + sizeof(lsg_lightmap) * kNoAllocMaps);
static_assert(kNoAllocBytes <= 72u * 1024u * 1024u, "the static filter scratch must stay modest");
}  // namespace lsg

static __device__ __align__(256) unsigned char scratch[lsg::kNoAllocBytes];

namespace lsg {
constexpr int kNoAllocDevices = 64;
static void* g_noalloc_ptr[kNoAllocDevices];
static std::mutex g_noalloc_mu;

// the current device's copy of `scratch`, looked up once per device with it current
static int noalloc_scratch(unsigned char** out)
{
    int dev = 0;
    LSG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) return fail(LSG_ERR_LIMIT, "device %d beyond the %d the library tracks", dev, kNoAllocDevices);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __LINE__);
        if (!p) return fail(LSG_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
        g_noalloc_ptr[dev] = p;
    }
    *out = (unsigned char*)g_noalloc_ptr[dev];
    return LSG_OK;
}

// FilterMap table from the caller's lightmap list: first = exclusive prefix sum of the
// texel counts (one workgroup; each thread sums a run of 16 maps, then a scan of the runs)
__global__ __launch_bounds__(256) void k_maps_from_lightmaps(const lsg_lightmap* __restrict__ lm, uint32_t nmaps,
                                                             FilterMap* __restrict__ fm)
{
    __shared__ uint32_t run[256];
    constexpr uint32_t kRun = (uint32_t)(kNoAllocMaps / 256);
    const uint32_t t = threadIdx.x, b = t * kRun, e = min(nmaps, b + kRun);
    uint32_t s = 0;
    for (uint32_t m = b; m < e; m++) s += lm[m].width * lm[m].height;
    run[t] = s;
    __syncthreads();
    for (uint32_t d = 1; d < 256; d *= 2) {
        const uint32_t v = t >= d ? run[t - d] : 0u;
        __syncthreads();
        run[t] += v;
        __syncthreads();
    }
    uint32_t first = run[t] - s;
    for (uint32_t m = b; m < e; m++) {
        const lsg_lightmap L = lm[m];
        fm[m].first = first; fm[m].base = L.rgba_base; fm[m].out_base = L.rgbf_base;
        fm[m].width = L.width; fm[m].height = L.height;
        first += L.width * L.height;
    }
}
static_assert(kNoAllocMaps % 256 == 0, "maps are summed in 256 runs");
}  // namespace lsg

extern "C" int lsg_filter_lightmaps_noalloc(const lsg_lightmap* maps, uint32_t nmaps, float* rgba, float* rgbf,
                                            const lsg_filter_params* params, void* stream)
{
    if (!maps || !rgba || !params) return fail(LSG_ERR_INVALID, "lsg_filter_lightmaps_noalloc: null argument");
    if (nmaps > kNoAllocMaps) return fail(LSG_ERR_LIMIT, "lsg_filter_lightmaps_noalloc: %u lightmaps, at most %u", nmaps, (unsigned)kNoAllocMaps);
    uint64_t total = 0;
    for (uint32_t m = 0; m < nmaps; m++) {
        if (!maps[m].width || !maps[m].height) return fail(LSG_ERR_INVALID, "lightmap %u has zero size", m);
        total += (uint64_t)maps[m].width * maps[m].height;
    }
    if (total > kNoAllocTexels)
        return fail(LSG_ERR_LIMIT, "lsg_filter_lightmaps_noalloc: %llu texels, at most %u", (unsigned long long)total, (unsigned)kNoAllocTexels);
    if (!total) return LSG_OK;
    unsigned char* base = nullptr;
    int st = noalloc_scratch(&base);
    if (st) return st;
    const hipStream_t s = (hipStream_t)stream;
    lsg_lightmap* stage = (lsg_lightmap*)(base + kOffStage);
    FilterMap* d_maps = (FilterMap*)(base + kOffMaps);
    LSG_HIP(hipMemcpyAsync(stage, maps, sizeof(lsg_lightmap) * nmaps, hipMemcpyHostToDevice, s));
    k_maps_from_lightmaps<<<1, 256, 0, s>>>(stage, nmaps, d_maps);
    FilterScratch S;   // no pinned log and no events: the grow loop synchronises per chunk
    S.a = (float4*)(base + kOffA); S.b = (float4*)(base + kOffB); S.c = (float4*)(base + kOffC);
    S.front[0] = (uint32_t*)(base + kOffFront0); S.front[1] = (uint32_t*)(base + kOffFront1);
    S.mark = (uint32_t*)(base + kOffMark); S.flags = (uint8_t*)(base + kOffFlags);
    S.changed = (uint32_t*)(base + kOffChanged);
    st = filter_run(d_maps, nmaps, (uint32_t)total, (float4*)rgba, rgbf, *params, S, s, true, -1, nullptr);
    if (!st) { hipError_t e = hipStreamSynchronize(s); if (e != hipSuccess) st = hip_fail(e, "sync", __LINE__); }
    return st;
}
