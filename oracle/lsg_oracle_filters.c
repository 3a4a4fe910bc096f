/*
 * lsg_oracle_filters.c -- CPU oracle, lightmap filter chain.  TEST
 * INFRASTRUCTURE ONLY (see lsg_oracle.h).
 *
 * Sequential restatement of RRBuffer::lightmapSmooth (RRBuffer/RRBuffer.cpp:630-802,
 * with the unwrap-seam smoothing of :734-777 when seams are given, and wrap),
 * lightmapGrowForBilinearInterpolation (:804-863, in place, exactly as the reference
 * scans), lightmapGrow (:865-962, both the clamped and the wrapped neighbourhood),
 * lightmapFillBackground (:964-975) and the RGB copy of copyElementsTo for one RGBA
 * lightmap, and the seam search of UnwrapSeams (:529-626) over a stitched mesh.
 * Float operations in the reference's order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { QX0Y0 = 1, QX1Y0 = 2, QX0Y1 = 4, QX1Y1 = 8, EX0 = 16, EX1 = 32, EY0 = 64, EY1 = 128 };

static inline unsigned to_flags(float a) { return (unsigned)(a * 255 + 0.5f); }

typedef struct { float v[4]; } v4;

static inline v4 add4(v4 a, v4 b) { v4 r; for (int k = 0; k < 4; k++) r.v[k] = a.v[k] + b.v[k]; return r; }
static inline v4 mul4(v4 a, float s) { v4 r; for (int k = 0; k < 4; k++) r.v[k] = a.v[k] * s; return r; }

/* The seam pass after one blur pass (RRBuffer.cpp:734-777), on dst in place, seams in
 * order.  A seam is 8 floats: its edge in the first triangle's unwrap (two uv points),
 * then the same edge in the second's.  Each of 1001 points along both edges that lies
 * in [0,1)^2 in both and on used texels adds both texels' values to the running sums
 * of the texel under each edge; when the texel under an edge changes, the finished run's
 * texel takes its sum divided by the sum's alpha (RGB only). */
static void seam_pass(v4* dst, uint32_t w, uint32_t h, const float* seams, uint32_t nseams)
{
    for (uint32_t s = 0; s < nseams; s++) {
        const float* e = seams + 8 * (size_t)s;
        uint32_t last1 = UINT32_MAX, last2 = UINT32_MAX;
        v4 acc1 = {{0, 0, 0, 0}}, acc2 = {{0, 0, 0, 0}};
        for (uint32_t i = 0; i <= 1000; i++) {
            const float t = (float)i * 0.001f, om = 1 - t;
            const float u1 = e[0] * om + e[2] * t, v1 = e[1] * om + e[3] * t;
            const float u2 = e[4] * om + e[6] * t, v2 = e[5] * om + e[7] * t;
            if (!(u1 >= 0 && u1 < 1 && v1 >= 0 && v1 < 1)) continue;
            if (!(u2 >= 0 && u2 < 1 && v2 >= 0 && v2 < 1)) continue;
            const uint32_t i1 = (uint32_t)(u1 * (float)w) + (uint32_t)(v1 * (float)h) * w;
            const uint32_t i2 = (uint32_t)(u2 * (float)w) + (uint32_t)(v2 * (float)h) * w;
            if (!(dst[i1].v[3] != 0 && dst[i2].v[3] != 0)) continue;
            if (i1 != last1) {
                if (last1 != UINT32_MAX) for (int q = 0; q < 3; q++) dst[last1].v[q] = acc1.v[q] / acc1.v[3];
                last1 = i1;
                memset(&acc1, 0, sizeof acc1);
            }
            if (i2 != last2) {
                if (last2 != UINT32_MAX) for (int q = 0; q < 3; q++) dst[last2].v[q] = acc2.v[q] / acc2.v[3];
                last2 = i2;
                memset(&acc2, 0, sizeof acc2);
            }
            const v4 both = add4(dst[i1], dst[i2]);
            acc1 = add4(acc1, both);
            acc2 = add4(acc2, both);
        }
        if (last1 != UINT32_MAX) for (int q = 0; q < 3; q++) dst[last1].v[q] = acc1.v[q] / acc1.v[3];
        if (last2 != UINT32_MAX) for (int q = 0; q < 3; q++) dst[last2].v[q] = acc2.v[q] / acc2.v[3];
    }
}

/* rgba: w*h*4 floats (in/out, the RGBA temp buffer); rgbf: w*h*3 floats (out); wrap:
 * FilteringParameters::wrap; seams: 8 floats per seam of the map's object, or NULL */
void lsgo_filter_lightmap_ex(float* rgba, uint32_t w, uint32_t h, float sigma, float spread, const float* bg, float* rgbf,
                             int wrap, const float* seams, uint32_t nseams)
{
    const uint32_t size = w * h;
    v4* img = (v4*)rgba;
    v4* src = (v4*)malloc(sizeof(v4) * size);
    v4* dst = (v4*)malloc(sizeof(v4) * size);
    unsigned char* bf = (unsigned char*)malloc(size);
    /* ---- lightmapSmooth */
    if (sigma > 0) {
        for (uint32_t i = 0; i < size; i++) src[i] = img[i];
        for (uint32_t j = 0; j < h; j++)
            for (uint32_t i = 0; i < w; i++) {
                uint32_t k = i + j * w;
                unsigned cf = to_flags(src[k].v[3]);
                unsigned bit = 1, fl = 0;
                for (int jj = -1; jj < 2; jj++)
                    for (int ii = -1; ii < 2; ii++)
                        if (ii || jj) {
                            int x = (int)i + ii, y = (int)j + jj;
                            if (wrap) { x = (int)(((uint32_t)x + w) % w); y = (int)(((uint32_t)y + h) % h); }
                            if (x >= 0 && x < (int)w && y >= 0 && y < (int)h) {
                                unsigned nf = to_flags(src[x + y * w].v[3]);
                                if ((ii >= 0 || ((cf & EX0) && (nf & EX1))) && (ii <= 0 || ((cf & EX1) && (nf & EX0)))
                                    && (jj >= 0 || ((cf & EY0) && (nf & EY1))) && (jj <= 0 || ((cf & EY1) && (nf & EY0))))
                                    fl |= bit;
                            }
                            bit *= 2;
                        }
                bf[k] = (unsigned char)fl;
            }
        for (uint32_t i = 0; i < size; i++) src[i].v[3] = src[i].v[3] > 0 ? 1.f : 0.f;
        if (sigma > 10) sigma = 10;
        float s2sum = sigma * sigma;
        while (s2sum > 0) {
            float s2 = s2sum < 0.34f ? s2sum : 0.34f;
            float k0 = 1, k1 = expf(-0.5f / s2), k2 = expf(-1 / s2);
            s2sum -= s2;
            for (uint32_t j = 0; j < h; j++)
                for (uint32_t i = 0; i < w; i++) {
                    uint32_t k = i + j * w;
                    v4 c = src[k];
                    if (c.v[3] != 0) {
                        unsigned f = bf[k];
                        uint32_t im = (i - 1 + w) % w, ip = (i + 1) % w, jm = (j - 1 + h) % h, jp = (j + 1) % h;
                        c = mul4(c, k0);
                        c = add4(c, mul4(src[im + jm * w], (f & 1) ? k2 : 0));
                        c = add4(c, mul4(src[i + jm * w], (f & 2) ? k1 : 0));
                        c = add4(c, mul4(src[ip + jm * w], (f & 4) ? k2 : 0));
                        c = add4(c, mul4(src[im + j * w], (f & 8) ? k1 : 0));
                        c = add4(c, mul4(src[ip + j * w], (f & 16) ? k1 : 0));
                        c = add4(c, mul4(src[im + jp * w], (f & 32) ? k2 : 0));
                        c = add4(c, mul4(src[i + jp * w], (f & 64) ? k1 : 0));
                        c = add4(c, mul4(src[ip + jp * w], (f & 128) ? k2 : 0));
                        float a = c.v[3];
                        for (int q = 0; q < 4; q++) c.v[q] = c.v[q] / a;
                    }
                    dst[k] = c;
                }
            if (seams) seam_pass(dst, w, h, seams, nseams);
            v4* t = src; src = dst; dst = t;
        }
        for (uint32_t i = 0; i < size; i++)
            if (src[i].v[3] > 0) { img[i].v[0] = src[i].v[0]; img[i].v[1] = src[i].v[1]; img[i].v[2] = src[i].v[2]; }
    }
    /* ---- lightmapGrowForBilinearInterpolation, in place as the reference scans */
    for (uint32_t j = 0; j < h; j++)
        for (uint32_t i = 0; i < w; i++) {
            if (img[i + j * w].v[3] > 0.002f) continue;
            v4 sum = {{0, 0, 0, 0}};
            for (int jj = -1; jj < 2; jj++)
                for (int ii = -1; ii < 2; ii++) {
                    int x = (int)i + ii, y = (int)j + jj;
                    if (wrap) { x = (int)(((uint32_t)x + w) % w); y = (int)(((uint32_t)y + h) % h); }
                    else if (x < 0 || x >= (int)w || y < 0 || y >= (int)h) continue;
                    v4 c = img[x + y * w];
                    unsigned tf = to_flags(c.v[3]);
                    if ((ii < 1 && jj < 1 && (tf & QX1Y1)) || (ii > -1 && jj > -1 && (tf & QX0Y0))
                        || (ii < 1 && jj > -1 && (tf & QX1Y0)) || (ii > -1 && jj < 1 && (tf & QX0Y1))) {
                        float wt = (ii && jj) ? 1 : 1.6f;
                        v4 t = {{c.v[0], c.v[1], c.v[2], 1}};
                        sum = add4(sum, mul4(t, wt));
                    }
                }
            if (sum.v[3] != 0) {
                v4 r = {{sum.v[0] / sum.v[3], sum.v[1] / sum.v[3], sum.v[2] / sum.v[3], 0.001f}};
                img[i + j * w] = r;
            }
        }
    /* ---- lightmapGrow */
    uint32_t steps = spread > 0 ? (uint32_t)spread : 0;
    if (steps) {
        for (uint32_t i = 0; i < size; i++) {
            v4 c = img[i];
            if (c.v[3] > 0) { src[i] = c; src[i].v[3] = 1; } else memset(&src[i], 0, sizeof(v4));
        }
        while (steps--) {
            int changed = 0;
            for (uint32_t i = 0; i < size; i++) {
                v4 c = src[i];
                if (c.v[3] == 0) {
                    v4 e, dg;
                    if (wrap) {
                        /* flat texel numbers, each offset added in 32-bit unsigned arithmetic
                         * (wrapping below 0) and then reduced modulo the texel count
                         * (RRBuffer.cpp:912-927) */
                        e = add4(add4(add4(src[(i + 1u) % size], src[(i - 1u) % size]), src[(i + w) % size]), src[(i - w) % size]);
                        dg = add4(add4(add4(src[(i + 1u + w) % size], src[(i + 1u - w) % size]), src[(i - 1u + w) % size]),
                                  src[(i - 1u - w) % size]);
                    } else {
                        int x = (int)(i % w), y = (int)(i / w), W = (int)w, H = (int)h;
                        int xp = x + 1 < W - 1 ? x + 1 : W - 1, xm = x - 1 > 0 ? x - 1 : 0;
                        int yp = y + 1 < H - 1 ? y + 1 : H - 1, ym = y - 1 > 0 ? y - 1 : 0;
                        e = add4(add4(add4(src[xp + y * W], src[xm + y * W]), src[x + yp * W]), src[x + ym * W]);
                        dg = add4(add4(add4(src[xp + yp * W], src[xp + ym * W]), src[xm + yp * W]), src[xm + ym * W]);
                    }
                    v4 d = add4(mul4(e, 1.6f), dg);
                    if (d.v[3] != 0) {
                        float a = d.v[3];
                        for (int q = 0; q < 4; q++) c.v[q] = d.v[q] / a;
                        changed = 1;
                    }
                }
                dst[i] = c;
            }
            if (!changed) steps = 0;
            v4* t = src; src = dst; dst = t;
        }
        for (uint32_t i = 0; i < size; i++)
            if (src[i].v[3] > 0 && img[i].v[3] == 0) {
                img[i].v[0] = src[i].v[0]; img[i].v[1] = src[i].v[1]; img[i].v[2] = src[i].v[2]; img[i].v[3] = 0.001f;
            }
    }
    /* ---- lightmapFillBackground + RGB copy */
    for (uint32_t i = 0; i < size; i++) {
        if (img[i].v[3] <= 0) for (int q = 0; q < 4; q++) img[i].v[q] = bg[q];
        if (rgbf) { rgbf[3 * i] = img[i].v[0]; rgbf[3 * i + 1] = img[i].v[1]; rgbf[3 * i + 2] = img[i].v[2]; }
    }
    free(src); free(dst); free(bf);
}

void lsgo_filter_lightmap(float* rgba, uint32_t w, uint32_t h, float sigma, float spread, const float* bg, float* rgbf)
{
    lsgo_filter_lightmap_ex(rgba, w, h, sigma, spread, bg, rgbf, 0, NULL, 0);
}

/* UnwrapSeams's search (RRBuffer.cpp:536-625) over a stitched mesh: corners[ntri*3] are
 * the stitched vertices (< nverts) of every triangle corner, tri_uv[ntri*6] the unwrap uv
 * of every corner.  For every triangle t1 and edge e (corners e, e+1), every later
 * triangle t2 listed at both stitched ends (each vertex lists its triangles in order, a
 * triangle once per corner it has there, and the two lists are walked together) whose
 * uv at those two vertices differs gives a seam.  Writes up to cap seams (8 floats
 * each) to out and returns how many there are. */
uint32_t lsgo_unwrap_seams(const uint32_t* corners, uint32_t nverts, uint32_t ntri, const float* tri_uv, float* out, uint32_t cap)
{
    uint32_t* cnt = (uint32_t*)calloc((size_t)nverts + 1, 4);
    for (uint32_t t = 0; t < ntri; t++) for (int k = 0; k < 3; k++) cnt[corners[3 * t + k] + 1]++;
    for (uint32_t v = 0; v < nverts; v++) cnt[v + 1] += cnt[v];
    uint32_t* list = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (size_t)(ntri ? ntri : 1));
    uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)nverts + 1));
    memcpy(fill, cnt, sizeof(uint32_t) * ((size_t)nverts + 1));
    for (uint32_t t = 0; t < ntri; t++) for (int k = 0; k < 3; k++) list[fill[corners[3 * t + k]]++] = t;
    uint32_t n = 0;
    for (uint32_t t1 = 0; t1 < ntri; t1++)
        for (int e = 0; e < 3; e++) {
            const uint32_t a = corners[3 * t1 + e], b = corners[3 * t1 + (e + 1) % 3];
            const float* u1 = tri_uv + 6 * (size_t)t1;
            const float s1[4] = {u1[2 * e], u1[2 * e + 1], u1[2 * ((e + 1) % 3)], u1[2 * ((e + 1) % 3) + 1]};
            const uint32_t* la = list + cnt[a];
            const uint32_t na = cnt[a + 1] - cnt[a];
            uint32_t ia = 0;
            for (uint32_t ib = cnt[b]; ib < cnt[b + 1]; ib++) {
                const uint32_t t2 = list[ib];
                if (t2 <= t1) continue;
                while (ia + 1 < na && la[ia] < t2) ia++;
                if (la[ia] != t2) continue;
                const uint32_t* c2 = corners + 3 * (size_t)t2;
                const float* u2 = tri_uv + 6 * (size_t)t2;
                const int ka = c2[0] == a ? 0 : (c2[1] == a ? 1 : 2);
                const int kb = c2[0] == b ? 0 : (c2[1] == b ? 1 : 2);
                const float s2[4] = {u2[2 * ka], u2[2 * ka + 1], u2[2 * kb], u2[2 * kb + 1]};
                if (s1[0] != s2[0] || s1[1] != s2[1] || s1[2] != s2[2] || s1[3] != s2[3]) {
                    if (n < cap) { memcpy(out + 8 * (size_t)n, s1, 16); memcpy(out + 8 * (size_t)n + 4, s2, 16); }
                    n++;
                }
            }
        }
    free(cnt); free(list); free(fill);
    return n;
}
