/*
 * lsg_oracle_filters.c -- CPU oracle, lightmap filter chain.  TEST
 * INFRASTRUCTURE ONLY (see lsg_oracle.h).
 *
 * Sequential restatement of RRBuffer::lightmapSmooth (RRBuffer/RRBuffer.cpp:630-802,
 * no unwrap seams, wrap=false), lightmapGrowForBilinearInterpolation (:804-863,
 * in place, exactly as the reference scans), lightmapGrow (:865-962, wrap=false),
 * lightmapFillBackground (:964-975) and the RGB copy of copyElementsTo for one
 * RGBA lightmap.  Float operations in the reference's order.
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

/* rgba: w*h*4 floats (in/out, the RGBA temp buffer); rgbf: w*h*3 floats (out) */
void lsgo_filter_lightmap(float* rgba, uint32_t w, uint32_t h, float sigma, float spread, const float* bg, float* rgbf)
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
                    if (x < 0 || x >= (int)w || y < 0 || y >= (int)h) continue;
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
    /* ---- lightmapGrow, no wrap */
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
                    int x = (int)(i % w), y = (int)(i / w), W = (int)w, H = (int)h;
                    int xp = x + 1 < W - 1 ? x + 1 : W - 1, xm = x - 1 > 0 ? x - 1 : 0;
                    int yp = y + 1 < H - 1 ? y + 1 : H - 1, ym = y - 1 > 0 ? y - 1 : 0;
                    v4 e = add4(add4(add4(src[xp + y * W], src[xm + y * W]), src[x + yp * W]), src[x + ym * W]);
                    v4 dg = add4(add4(add4(src[xp + yp * W], src[xp + ym * W]), src[xm + yp * W]), src[xm + ym * W]);
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
