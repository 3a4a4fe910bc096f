// lsgpu_filter.hip -- lightmap filter chain and writeback on gfx950.
//
// After the gather, RRSolver::updateLightmap (RRSolver/lightmap.cpp:852-876) runs
// on every RGBA temp lightmap: lightmapSmooth(sigma) -> lightmapGrowForBilinear-
// Interpolation -> lightmapGrow(spread) -> lightmapFillBackground(bg) ->
// copyElementsTo(user buffer).  These kernels restate RRBuffer.cpp:630-975 for a
// batch of lightmaps at once (one grid over every texel of every map), with the
// same float operation order, so a CPU statement reproduces them bit for bit.
// Alpha carries texelFlags/255 (gather.h:79-80) exactly as in the reference.
// Unwrap-seam smoothing (RRBuffer.cpp:533-626, 734-777) needs the object's mesh
// and is not part of this kernel set; wrap=true is rejected.
#include "lsgpu_internal.h"
#include "lsgpu_filter.h"
#include <vector>
#include <cmath>

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
// of center and neighbours (no wrap), then alpha := alpha>0 (:700-703)
__global__ __launch_bounds__(256) void k_smooth_prepare(const FilterMap* __restrict__ maps, uint32_t nmaps, uint32_t total,
                                                        const float4* __restrict__ rgba, float4* __restrict__ src, uint8_t* __restrict__ bf)
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
                const int x = i + ii, y = j + jj;
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
                                                       const float4* __restrict__ rgba, float4* __restrict__ out)
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
                const int x = i + ii, y = j + jj;
                if (x < 0 || x >= (int)w || y < 0 || y >= (int)h) continue;
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

// one lightmapGrow step without wrap (RRBuffer.cpp:925-947)
__global__ __launch_bounds__(256) void k_grow_step(const FilterMap* __restrict__ maps, uint32_t nmaps, uint32_t total,
                                                   const float4* __restrict__ src, float4* __restrict__ dst, uint32_t* __restrict__ changed)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const FilterMap M = maps[find_map(maps, nmaps, t)];
    const uint32_t k = t - M.first;
    const int w = (int)M.width, h = (int)M.height;
    const int x = (int)(k % M.width), y = (int)(k / M.width);
    const float4* s = src + M.first;
    float4 c = s[k];
    if (c.w == 0) {
        const int xp = x + 1 < w - 1 ? x + 1 : w - 1, xm = x - 1 > 0 ? x - 1 : 0;
        const int yp = y + 1 < h - 1 ? y + 1 : h - 1, ym = y - 1 > 0 ? y - 1 : 0;
        float4 e = f4add(f4add(f4add(s[xp + y * w], s[xm + y * w]), s[x + yp * w]), s[x + ym * w]);
        float4 dg = f4add(f4add(f4add(s[xp + yp * w], s[xp + ym * w]), s[xm + yp * w]), s[xm + ym * w]);
        float4 d = f4add(f4mul(e, 1.6f), dg);
        if (d.w != 0) {
            const float a = d.w;
            c = make_float4(d.x / a, d.y / a, d.z / a, d.w / a);
            *changed = 1u;
        }
    }
    dst[M.first + k] = c;
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

static inline unsigned nblk(uint32_t n) { return (n + 255) / 256; }

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

int filter_run(const FilterMap* d_maps, uint32_t nmaps, uint32_t total, float4* rgba, float* rgbf, const lsg_filter_params& fp,
               FilterScratch& S, hipStream_t s, bool rgba_out, int fixed_steps, uint32_t* steps_used)
{
    if (!total) return LSG_OK;
    // lightmapSmooth (RRBuffer.cpp:630-802)
    float sigma = fp.smoothing_amount;
    if (sigma < 0) return fail(LSG_ERR_INVALID, "smoothing_amount < 0");
    if (sigma > 0) {
        k_smooth_prepare<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, rgba, S.a, S.flags);
        if (sigma > 10) sigma = 10;
        float sigma2sum = sigma * sigma;
        float4* src = S.a; float4* dst = S.b;
        while (sigma2sum > 0) {
            const float sigma2 = sigma2sum < 0.34f ? sigma2sum : 0.34f;
            const float4 kernel = make_float4(1.f, expf(-0.5f / sigma2), expf(-1 / sigma2), expf(-2 / sigma2));
            sigma2sum -= sigma2;
            k_smooth_pass<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, src, dst, S.flags, kernel);
            std::swap(src, dst);
        }
        k_smooth_finish<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, rgba, src);
    }
    // lightmapGrowForBilinearInterpolation
    k_grow_bilinear<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, rgba, S.c);
    // lightmapGrow(spread)
    const float4* g = nullptr;
    uint32_t steps = (uint32_t)(fp.spread_foreground > 0 ? fp.spread_foreground : 0);
    if (steps_used) *steps_used = 0;
    if (steps && fixed_steps >= 0) {
        // graph-capturable form: a fixed number of steps, no host synchronisation, the
        // last step's change flag left in S.changed[0] for the caller to check
        k_grow_init<<<nblk(total), 256, 0, s>>>(total, S.c, S.a);
        float4* src = S.a; float4* dst = S.b;
        const uint32_t n = std::min<uint32_t>((uint32_t)fixed_steps, steps);
        for (uint32_t q = 0; q < n; q++) {
            k_zero_u32<<<1, 256, 0, s>>>(S.changed, 1);
            k_grow_step<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, src, dst, S.changed);
            std::swap(src, dst);
        }
        if (steps_used) *steps_used = n;
        g = src;
    } else if (steps) {
        k_grow_init<<<nblk(total), 256, 0, s>>>(total, S.c, S.a);
        float4* src = S.a; float4* dst = S.b;
        uint32_t done = 0;
        while (done < steps) {
            const uint32_t chunk = std::min<uint32_t>(16, steps - done);
            LSG_HIP(hipMemsetAsync(S.changed, 0, 4 * chunk, s));
            for (uint32_t q = 0; q < chunk; q++) {
                k_grow_step<<<nblk(total), 256, 0, s>>>(d_maps, nmaps, total, src, dst, S.changed + q);
                std::swap(src, dst);
            }
            uint32_t ch[16] = {0};
            LSG_HIP(hipMemcpyAsync(ch, S.changed, 4 * chunk, hipMemcpyDeviceToHost, s));
            LSG_HIP(hipStreamSynchronize(s));
            bool stop = false;
            for (uint32_t q = 0; q < chunk; q++) if (!ch[q]) { stop = true; done += q + 1; break; }
            if (stop) break;   // a step without change: the reference stops (RRBuffer.cpp:948-949)
            done += chunk;
        }
        if (steps_used) *steps_used = done;
        g = src;
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
    if ((e = hipMalloc(&S.changed, 64)) != hipSuccess) return hip_fail(e, "hipMalloc", __LINE__);
    return LSG_OK;
}

void scratch_free(FilterScratch& S)
{
    if (S.a) (void)hipFree(S.a);
    if (S.b) (void)hipFree(S.b);
    if (S.c) (void)hipFree(S.c);
    if (S.flags) (void)hipFree(S.flags);
    if (S.changed) (void)hipFree(S.changed);
    S = FilterScratch();
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

using namespace lsg;

extern "C" int lsg_filter_lightmaps(const lsg_lightmap* maps, uint32_t nmaps, float* rgba, float* rgbf,
                                    const lsg_filter_params* params, void* stream)
{
    if (!maps || !rgba || !params) return fail(LSG_ERR_INVALID, "lsg_filter_lightmaps: null argument");
    if (params->wrap) return fail(LSG_ERR_INVALID, "lsg_filter_lightmaps: wrap=true is not supported");
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
