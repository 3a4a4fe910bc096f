// lsgpu_learned.hip -- the learned gather: a cheap gather (few rays per texel), a
// feature pass, the MLP surrogate (lsgpu_mlp.hip) and a decision kernel that keeps
// the network's answer where its own error estimate is below a bound and queues the
// texel for the full plain gather (processTexel, gather.cpp:600-836) otherwise.
//
// Features per texel (32), v~ = log1p(v / s) with s the per-scene mean cheap irradiance:
//   0-2   cheap hemisphere part~          3-5   cheap light part~
//   6-8   hemisphere part~, 3x3 mean      9-11  hemisphere part~, 5x5 mean
//   12-14 hemisphere part~, 11x11 mean    15-17 light part~, 3x3 mean
//   18-20 position in the scene box, [-1,1]   21-23 shading normal   24-26 albedo
//   27    log(0.01 + relative std of hemisphere luminance, 5x5)
//   28    log(0.01 + relative std of light luminance, 3x3)
//   29    covered fraction of the 11x11 window
//   30    log1p(A/s), A = unshadowed direct irradiance of the lights (mean channel,
//         light multiplier 1) averaged over the subtexel centroids by area
//   31    1
// The means run over texels of the same lightmap whose cheap pass was reliable and
// whose normal is within ~18 degrees of the texel's (same chart face).  Outputs of the
// network: a correction r (3) to the baseline b = log1p((H5 + L)/s), H5 the 5x5
// hemisphere mean and L the texel's own light part (unbiased, and sharp at shadow
// edges), so that log1p(E/s) = b + r, and the predicted RMS error of log1p(E/s) (1).
#include "lsgpu_internal.h"
#include "lsgpu_gather.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace lsg {

constexpr int kFeat = 32;
constexpr int kWin = 5;                       // half width of the largest window

__global__ void k_zero4(float4* __restrict__ p, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// pass 1: per-texel geometry features and the slot images (hemisphere, light, normal)
__global__ __launch_bounds__(256) void k_features_point(SceneView sc, const lsg_texel* __restrict__ texels, uint32_t n,
                                                        const lsg_subtexel* __restrict__ subs, const float4* __restrict__ parts,
                                                        const float* __restrict__ albedo, const lsg_light* __restrict__ lights,
                                                        uint32_t nlights, float inv_s, float* __restrict__ x,
                                                        float4* __restrict__ img)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const lsg_texel T = texels[i];
    const uint32_t m = T.sub_count_flags & 0xffffu;
    float* xi = x + kFeat * (size_t)i;
    for (int k = 0; k < kFeat; k++) xi[k] = 0.f;
    if (!m) return;
    const lsg_subtexel* ts = subs + T.sub_begin;
    // centroid position and normal of every subtexel: the largest one gives the
    // position/normal features, all of them (by area) the unshadowed direct term
    float p[3] = {0, 0, 0}, nn[3] = {0, 0, 0}, best_area = -1.f, direct = 0.f, area = 0.f;
    uint32_t tri = ts[0].triangle;
    for (uint32_t j = 0; j < m; j++) {
        const lsg_subtexel st = ts[j];
        const uint32_t tj = st.triangle;
        const uint32_t i0 = sc.idx[3 * (size_t)tj], i1 = sc.idx[3 * (size_t)tj + 1], i2 = sc.idx[3 * (size_t)tj + 2];
        const float cx = (st.uv[0] + st.uv[2] + st.uv[4]) * 0.333333333f;
        const float cy = (st.uv[1] + st.uv[3] + st.uv[5]) * 0.333333333f;
        const float w = 1 - cx - cy;
        float pj[3], nj[3];
        for (int c = 0; c < 3; c++) {
            const float a = sc.verts[3 * (size_t)i0 + c];
            pj[c] = a + (sc.verts[3 * (size_t)i1 + c] - a) * cx + (sc.verts[3 * (size_t)i2 + c] - a) * cy;
            nj[c] = sc.vnormal[3 * (size_t)i0 + c] * w + sc.vnormal[3 * (size_t)i1 + c] * cx + sc.vnormal[3 * (size_t)i2 + c] * cy;
        }
        const float nl = sqrtf(nj[0] * nj[0] + nj[1] * nj[1] + nj[2] * nj[2]);
        for (int c = 0; c < 3; c++) nj[c] = nl > 0 ? nj[c] / nl : 0.f;
        float dj = 0.f;
        for (uint32_t li = 0; li < nlights; li++) {
            const lsg_light L = lights[li];
            float dx, dy, dz, att = 1.f;
            if (L.type == 1) { dx = -L.direction[0]; dy = -L.direction[1]; dz = -L.direction[2]; }
            else {
                dx = L.position[0] - pj[0]; dy = L.position[1] - pj[1]; dz = L.position[2] - pj[2];
                att = 1 / (0.000000000137289f + (dx * dx + dy * dy + dz * dz));
            }
            const float dl = sqrtf(dx * dx + dy * dy + dz * dz);
            const float ni = dl > 0 ? (dx * nj[0] + dy * nj[1] + dz * nj[2]) / dl : 0.f;
            if (ni > 0) dj += (L.color[0] + L.color[1] + L.color[2]) * (1.f / 3.f) * att * ni;
        }
        direct += dj * st.area; area += st.area;
        if (st.area > best_area) {
            best_area = st.area; tri = tj;
            for (int c = 0; c < 3; c++) { p[c] = pj[c]; nn[c] = nj[c]; }
        }
    }
    const float bmin[3] = {sc.col.bmin.x, sc.col.bmin.y, sc.col.bmin.z}, bmax[3] = {sc.col.bmax.x, sc.col.bmax.y, sc.col.bmax.z};
    for (int c = 0; c < 3; c++) {
        const float ext = bmax[c] - bmin[c];
        xi[18 + c] = ext > 0 ? (p[c] - bmin[c]) / ext * 2.f - 1.f : 0.f;
        xi[21 + c] = nn[c];
        xi[24 + c] = albedo[3 * (size_t)tri + c];
    }
    xi[30] = log1pf(fmaxf(area > 0 ? direct / area : 0.f, 0.f) * inv_s);
    xi[31] = 1.f;
    const float4 ph = parts[2 * (size_t)i], pl = parts[2 * (size_t)i + 1];
    const bool ok = ph.w > 0 || pl.w > 0;
    float4* q = img + 3 * (size_t)T.out_index;
    q[0] = make_float4(ph.x, ph.y, ph.z, ok ? 1.f : 0.f);
    q[1] = make_float4(pl.x, pl.y, pl.z, ok ? 1.f : 0.f);
    q[2] = make_float4(nn[0], nn[1], nn[2], 0.f);
}

__device__ inline float lg(float v, float inv_s) { return log1pf(fmaxf(v, 0.f) * inv_s); }

// pass 2: window means and spreads over the slot images
__global__ __launch_bounds__(256) void k_features_window(const lsg_texel* __restrict__ texels, uint32_t n,
                                                         const lsg_lightmap* __restrict__ maps, uint32_t nmaps,
                                                         const float4* __restrict__ img, float inv_s, float* __restrict__ x)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const lsg_texel T = texels[i];
    if (!(T.sub_count_flags & 0xffffu)) return;
    float* xi = x + kFeat * (size_t)i;
    const uint32_t slot = T.out_index;
    // the lightmap holding this slot (maps sorted by rgba_base)
    uint32_t lo = 0, hi = nmaps;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (maps[mid].rgba_base <= slot) lo = mid; else hi = mid;
    }
    const lsg_lightmap M = maps[lo];
    const float4 c0 = img[3 * (size_t)slot], l0 = img[3 * (size_t)slot + 1], n0 = img[3 * (size_t)slot + 2];
    xi[0] = lg(c0.x, inv_s); xi[1] = lg(c0.y, inv_s); xi[2] = lg(c0.z, inv_s);
    xi[3] = lg(l0.x, inv_s); xi[4] = lg(l0.y, inv_s); xi[5] = lg(l0.z, inv_s);
    if (slot < M.rgba_base || slot - M.rgba_base >= M.width * M.height) return;
    const int w = (int)M.width, h = (int)M.height;
    const int px = (int)((slot - M.rgba_base) % M.width), py = (int)((slot - M.rgba_base) / M.width);
    float3 s1 = {0, 0, 0}, s2 = {0, 0, 0}, s5 = {0, 0, 0}, sl = {0, 0, 0};
    float c1 = 0, c2 = 0, c5 = 0, cl = 0, q2 = 0, ql = 0;
    for (int dy = -kWin; dy <= kWin; dy++) {
        const int yy = py + dy;
        if (yy < 0 || yy >= h) continue;
        for (int dx = -kWin; dx <= kWin; dx++) {
            const int xx = px + dx;
            if (xx < 0 || xx >= w) continue;
            const size_t k = M.rgba_base + (size_t)yy * w + xx;
            const float4 hk = img[3 * k];
            if (!(hk.w > 0)) continue;
            const float4 nk = img[3 * k + 2];
            if (nk.x * n0.x + nk.y * n0.y + nk.z * n0.z < 0.95f) continue;
            const int r = max(abs(dx), abs(dy));
            const float lum = (hk.x + hk.y + hk.z) * (1.f / 3.f);
            s5.x += hk.x; s5.y += hk.y; s5.z += hk.z; c5 += 1;
            if (r <= 2) { s2.x += hk.x; s2.y += hk.y; s2.z += hk.z; c2 += 1; q2 += lum * lum; }
            if (r <= 1) {
                s1.x += hk.x; s1.y += hk.y; s1.z += hk.z; c1 += 1;
                const float4 lk = img[3 * k + 1];
                const float ll = (lk.x + lk.y + lk.z) * (1.f / 3.f);
                sl.x += lk.x; sl.y += lk.y; sl.z += lk.z; cl += 1; ql += ll * ll;
            }
        }
    }
    const float f1 = c1 > 0 ? 1 / c1 : 0.f, f2 = c2 > 0 ? 1 / c2 : 0.f, f5 = c5 > 0 ? 1 / c5 : 0.f, fl = cl > 0 ? 1 / cl : 0.f;
    xi[6] = lg(s1.x * f1, inv_s); xi[7] = lg(s1.y * f1, inv_s); xi[8] = lg(s1.z * f1, inv_s);
    xi[9] = lg(s2.x * f2, inv_s); xi[10] = lg(s2.y * f2, inv_s); xi[11] = lg(s2.z * f2, inv_s);
    xi[12] = lg(s5.x * f5, inv_s); xi[13] = lg(s5.y * f5, inv_s); xi[14] = lg(s5.z * f5, inv_s);
    xi[15] = lg(sl.x * fl, inv_s); xi[16] = lg(sl.y * fl, inv_s); xi[17] = lg(sl.z * fl, inv_s);
    const float m2 = (s2.x + s2.y + s2.z) * (1.f / 3.f) * f2, ml = (sl.x + sl.y + sl.z) * (1.f / 3.f) * fl;
    const float v2 = fmaxf(q2 * f2 - m2 * m2, 0.f), vl = fmaxf(ql * fl - ml * ml, 0.f);
    xi[27] = logf(0.01f + (m2 > 0 ? sqrtf(v2) / m2 : 0.f));
    xi[28] = logf(0.01f + (ml > 0 ? sqrtf(vl) / ml : 0.f));
    xi[29] = c5 * (1.f / ((2 * kWin + 1) * (2 * kWin + 1)));
}

// sum of the cheap estimate's mean channel over reliable texels -> scale.  A fixed grid
// strides over the texels; each workgroup writes its partial sums (in double) to its
// own slot, and the host adds the slots in order: no atomics, the same result every run.
constexpr uint32_t kSumBlocks = 1024;
__global__ __launch_bounds__(256) void k_cheap_sum(const float4* __restrict__ parts, uint32_t n, double* __restrict__ partial)
{
    double v = 0.0, c = 0.0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float4 a = parts[2 * (size_t)i], b = parts[2 * (size_t)i + 1];
        if (a.w > 0 || b.w > 0) { v += (a.x + a.y + a.z + b.x + b.y + b.z) / 3.f; c += 1.0; }
    }
    __shared__ double s[2][256];
    s[0][threadIdx.x] = v; s[1][threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w; w >>= 1) {
        if ((int)threadIdx.x < w) { s[0][threadIdx.x] += s[0][threadIdx.x + w]; s[1][threadIdx.x] += s[1][threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = s[0][0]; partial[2 * blockIdx.x + 1] = s[1][0]; }
}

// log1p(E/s) of the baseline, channel c
__device__ inline float base_log(const float* xi, int c)
{
    return log1pf(expm1f(xi[9 + c]) + expm1f(xi[3 + c]));
}

// network irradiance E/s (3) and the predicted absolute error in units of s: an error
// d in log1p(E/s) is an irradiance error of about d * (E + s)
__device__ inline float predict(const float* yi, const float* xi, float e[3])
{
    for (int c = 0; c < 3; c++) e[c] = expm1f(base_log(xi, c) + yi[c]);
    const float em = fmaxf((e[0] + e[1] + e[2]) * (1.f / 3.f), 0.f);
    return fabsf(yi[3]) * (1.f + em);
}

// error-score histogram: 8 bins per octave from 2^-24 to 2^8 (256 bins)
constexpr int kHistBins = 256;
__host__ __device__ inline int score_bin(float score)
{
    if (!(score > 0.f)) return 0;
    const float b = floorf((log2f(score) + 24.f) * 8.f);
    return b < 0.f ? 0 : (b > (float)(kHistBins - 1) ? kHistBins - 1 : (int)b);
}
// upper score edge of bin b
__host__ __device__ inline float bin_upper(int b) { return exp2f((float)(b + 1) / 8.f - 24.f); }

// keep the network's answer or queue the texel for the plain gather: a texel falls
// back when its cheap pass was unreliable or its predicted error exceeds tau * s
__global__ __launch_bounds__(256) void k_decide(const float* __restrict__ y, const float* __restrict__ x,
                                                const lsg_texel* __restrict__ texels,
                                                const float4* __restrict__ parts, uint32_t n, float scale, float tau,
                                                float4* __restrict__ out, lsg_texel* __restrict__ fallback,
                                                uint32_t* __restrict__ nfallback)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    bool fb = false;
    lsg_texel T;
    if (i < n) {
        T = texels[i];
        const float4 a = parts[2 * (size_t)i], b = parts[2 * (size_t)i + 1];
        const bool reliable = a.w > 0 || b.w > 0;
        const float* yi = y + 4 * (size_t)i;
        const float* xi = x + kFeat * (size_t)i;
        float e[3];
        const float score = predict(yi, xi, e);
        fb = !reliable || !(score <= tau);
        if (!fb) {
            const uint32_t flags = (T.sub_count_flags >> 16) & 0xffu;
            out[T.out_index] = make_float4(fmaxf(scale * e[0], 0.f), fmaxf(scale * e[1], 0.f), fmaxf(scale * e[2], 0.f),
                                           flags / 255.f);
        }
    }
    // one atomic per wavefront: ballot the fallbacks, rank them, reserve a range
    const uint64_t mask = __ballot(fb);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t cnt = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (lane == 0 && cnt) base = atomicAdd(nfallback, cnt);
    base = (uint32_t)__shfl((int)base, 0);
    if (fb) fallback[base + (uint32_t)__popcll(mask & (lane ? (~0ull >> (64 - lane)) : 0ull))] = T;
}

__device__ inline void texel_error(const float* __restrict__ y, const float* __restrict__ x,
                                   const lsg_texel* __restrict__ texels, const float4* __restrict__ parts,
                                   const float4* __restrict__ c2img, uint64_t i, float inv_s, unsigned long long* hs)
{
    const float4 a = parts[2 * (size_t)i], b = parts[2 * (size_t)i + 1];
    if (!(a.w > 0 || b.w > 0)) return;                   // always re-gathered
    float e[3];
    const float score = predict(y + 4 * (size_t)i, x + kFeat * (size_t)i, e);
    const float4 c2 = c2img[texels[i].out_index];
    const float c1v[3] = {(a.x + b.x) * inv_s, (a.y + b.y) * inv_s, (a.z + b.z) * inv_s};
    const float c2v[3] = {c2.x * inv_s, c2.y * inv_s, c2.z * inv_s};
    float d = 0.f, v = 0.f, p = 0.f;
    for (int c = 0; c < 3; c++) {
        const float pc = fmaxf(e[c], 0.f);
        d += (pc - c2v[c]) * (pc - c2v[c]);
        v += (c1v[c] - c2v[c]) * (c1v[c] - c2v[c]);
        p += pc * pc;
    }
    const float est = (d - 0.5f * v) * (1.f / 3.f), pw = p * (1.f / 3.f);
    const int bin = score_bin(score);
    const double q = 16777216.0;
    atomicAdd(&hs[bin], (unsigned long long)(long long)llrint((double)est * q));
    atomicAdd(&hs[kHistBins + bin], 1ull);
    atomicAdd(&hs[2 * kHistBins], (unsigned long long)(long long)llrint((double)pw * q));
}

// Unbiased per-scene estimate of the surrogate's squared error from a second,
// independent cheap gather c2 (same rays per texel, another seed): with c1 the cheap
// gather the features came from and E the expected irradiance,
//   E[(pred - c2)^2] = (pred - E)^2 + var(c2),   E[(c1 - c2)^2] / 2 = var(c),
// so (pred - c2)^2 - (c1 - c2)^2 / 2 estimates (pred - E)^2 without the plain gather.
// Per score bin: the sum of that estimate (fixed point, units s^2 * 2^-24, signed)
// and the texel count; hist[2*256] = sum of pred^2 over reliable texels (same units).
// Over the texels i = 0, stride, 2*stride, ... (c2 exists for those); per-block sums
// in LDS, one global atomic per non-empty bin and block.
__global__ __launch_bounds__(256) void k_error_hist(const float* __restrict__ y, const float* __restrict__ x,
                                                    const lsg_texel* __restrict__ texels, const float4* __restrict__ parts,
                                                    const float4* __restrict__ c2img, uint32_t n, uint32_t stride,
                                                    float inv_s, unsigned long long* __restrict__ hist)
{
    __shared__ unsigned long long hs[2 * kHistBins + 1];
    for (uint32_t k = threadIdx.x; k < 2 * kHistBins + 1; k += 256) hs[k] = 0ull;
    __syncthreads();
    const uint64_t i = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * stride;
    if (i < n) texel_error(y, x, texels, parts, c2img, i, inv_s, hs);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 2 * kHistBins + 1; k += 256)
        if (hs[k]) atomicAdd(&hist[k], hs[k]);
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

static inline unsigned nb(uint64_t n) { return (unsigned)((n + 255) / 256); }

} // namespace lsg

using namespace lsg;

extern "C" {

int lsg_gather_parts_device(lsg_scene* s, const lsg_texel* texels, uint32_t ntexels, const lsg_subtexel* subs,
                            const lsg_gather_params* p, float* out_rgba, float* parts, void* stream)
{
    if (!s || !p || (ntexels && (!texels || !subs || !out_rgba || !parts))) return fail(LSG_ERR_INVALID, "lsg_gather_parts_device: null argument");
    if (!ntexels) return LSG_OK;
    GatherArgs A;
    int stack = 0;
    int st = gather_args(s, texels, ntexels, subs, p, out_rgba, A, stack);
    if (st) return st;
    A.parts = (float4*)parts;
    if ((st = walk_records(s, A))) return st;
    return launch_gather(A, stack, (hipStream_t)stream);
}

int lsg_learned_scale(const float* parts, uint32_t n, float* scale)
{
    if (!scale || (n && !parts)) return fail(LSG_ERR_INVALID, "lsg_learned_scale: null argument");
    if (!n) { *scale = 1.f; return LSG_OK; }
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(nb(n), kSumBlocks);
    double* d = nullptr;
    LSG_HIP(hipMalloc(&d, 16 * (size_t)blocks));
    k_cheap_sum<<<blocks, 256>>>((const float4*)parts, n, d);
    std::vector<double> h(2 * (size_t)blocks);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(h.data(), d, 16 * (size_t)blocks, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "scale", __LINE__);
    double v = 0.0, c = 0.0;
    for (uint32_t b = 0; b < blocks; b++) { v += h[2 * b]; c += h[2 * b + 1]; }
    *scale = (c > 0 && v > 0) ? (float)(v / c) : 1.f;
    return LSG_OK;
}

int lsg_learned_features(lsg_scene* s, const lsg_texel* texels, uint32_t n, const lsg_subtexel* subs, const float* parts,
                         const float* tri_albedo, float scale, const lsg_lightmap* maps, uint32_t nmaps, uint32_t nslots,
                         float* work, float* x, void* stream)
{
    if (!s || (n && (!texels || !subs || !parts || !tri_albedo || !x || !maps || !work || !nmaps)))
        return fail(LSG_ERR_INVALID, "lsg_learned_features: null argument");
    if (!(scale > 0)) return fail(LSG_ERR_INVALID, "lsg_learned_features: scale must be > 0");
    if (!n) return LSG_OK;
    for (uint32_t k = 0; k < nmaps; k++) {
        if ((uint64_t)maps[k].rgba_base + (uint64_t)maps[k].width * maps[k].height > nslots)
            return fail(LSG_ERR_INVALID, "lsg_learned_features: lightmap %u exceeds %u slots", k, nslots);
        if (k && maps[k].rgba_base < maps[k - 1].rgba_base + maps[k - 1].width * maps[k - 1].height)
            return fail(LSG_ERR_INVALID, "lsg_learned_features: lightmaps must be sorted by rgba_base and disjoint");
    }
    if (s->impl.d_tri_mat)
        return fail(LSG_ERR_INVALID, "lsg_learned_features: the learned features read flat exitance; scenes with point materials are not supported");
    const ColliderImpl& C = s->impl.col->impl;
    SceneView sv{};
    sv.col = C.view(); sv.verts = C.d_verts; sv.idx = C.d_idx; sv.vnormal = s->impl.d_vnormal;
    sv.side = s->impl.d_side; sv.exitance = s->impl.d_exit;
    hipStream_t st = (hipStream_t)stream;
    // the lightmap list stays on the device between passes; uploaded again only when it
    // changes (the upload waits for earlier work on st that may still read it)
    SceneImpl& S = s->impl;
    const bool same = S.d_lmaps && S.h_lmaps.size() == nmaps &&
                      !memcmp(S.h_lmaps.data(), maps, sizeof(lsg_lightmap) * nmaps);
    if (!same) {
        if (nmaps > S.lmaps_cap) {
            LSG_HIP(hipStreamSynchronize(st));
            if (S.d_lmaps) (void)hipFree(S.d_lmaps);
            S.d_lmaps = nullptr; S.lmaps_cap = 0;
            LSG_HIP(hipMalloc(&S.d_lmaps, sizeof(lsg_lightmap) * nmaps));
            S.lmaps_cap = nmaps;
        }
        S.h_lmaps.assign(maps, maps + nmaps);
        LSG_HIP(hipStreamSynchronize(st));
        LSG_HIP(hipMemcpy(S.d_lmaps, S.h_lmaps.data(), sizeof(lsg_lightmap) * nmaps, hipMemcpyHostToDevice));
    }
    k_zero4<<<nb(3ull * nslots), 256, 0, st>>>((float4*)work, 3ull * nslots);
    k_features_point<<<nb(n), 256, 0, st>>>(sv, texels, n, subs, (const float4*)parts, tri_albedo, S.d_lights,
                                              S.nlights, 1.f / scale, x, (float4*)work);
    k_features_window<<<nb(n), 256, 0, st>>>(texels, n, S.d_lmaps, nmaps, (const float4*)work, 1.f / scale, x);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

int lsg_learned_decide(const float* y, const float* x, const lsg_texel* texels, const float* parts, uint32_t n, float scale,
                       float tau, float* out_rgba, lsg_texel* fallback, uint32_t* nfallback, void* stream)
{
    if (n && (!y || !x || !texels || !parts || !out_rgba || !fallback || !nfallback)) return fail(LSG_ERR_INVALID, "lsg_learned_decide: null argument");
    if (!n) return LSG_OK;
    k_decide<<<nb(n), 256, 0, (hipStream_t)stream>>>(y, x, texels, (const float4*)parts, n, scale, tau, (float4*)out_rgba, fallback, nfallback);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

// Largest bound tau (a score-bin edge) whose kept texels' estimated squared error
// stays within target_rmse^2 of the mean squared prediction: the per-scene error bound.
int lsg_learned_bound(const float* y, const float* x, const lsg_texel* texels, const float* parts, const float* cheap2_rgba,
                      uint32_t n, uint32_t sample_stride, float scale, float target_rmse, uint64_t* work, float* tau,
                      float* est_rmse, void* stream)
{
    if (!tau || (n && (!y || !x || !texels || !parts || !cheap2_rgba || !work))) return fail(LSG_ERR_INVALID, "lsg_learned_bound: null argument");
    if (!sample_stride) return fail(LSG_ERR_INVALID, "lsg_learned_bound: sample_stride must be >= 1");
    if (!(scale > 0) || !(target_rmse >= 0)) return fail(LSG_ERR_INVALID, "lsg_learned_bound: scale > 0 and target_rmse >= 0 required");
    *tau = 0.f;
    if (est_rmse) *est_rmse = 0.f;
    if (!n) return LSG_OK;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nh = 2 * kHistBins + 1;
    k_zero_u64<<<nb(nh), 256, 0, st>>>((unsigned long long*)work, nh);
    const uint64_t ns = ((uint64_t)n + sample_stride - 1) / sample_stride;
    k_error_hist<<<nb(ns), 256, 0, st>>>(y, x, texels, (const float4*)parts, (const float4*)cheap2_rgba, n, sample_stride,
                                         1.f / scale, (unsigned long long*)work);
    LSG_HIP(hipGetLastError());
    std::vector<long long> h(nh);
    LSG_HIP(hipMemcpyAsync(h.data(), work, 8 * nh, hipMemcpyDeviceToHost, st));
    LSG_HIP(hipStreamSynchronize(st));
    const double norm = (double)h[2 * kHistBins];
    double cum = 0.0, total = 0.0;
    for (int b = 0; b < kHistBins; b++) total += (double)h[b];
    int best = -1;
    for (int b = 0; b < kHistBins; b++) {
        cum += (double)h[b];
        if (cum <= (double)target_rmse * target_rmse * norm) best = b;
        else if (h[kHistBins + b]) break;               // keep a prefix of the score order
    }
    *tau = best < 0 ? 0.f : (best == kHistBins - 1 ? INFINITY : bin_upper(best));
    if (est_rmse) *est_rmse = norm > 0 ? (float)sqrt(fmax(total, 0.0) / norm) : 0.f;
    return LSG_OK;
}

} // extern "C"
