/*
 * lsg_oracle_raster.c -- CPU oracle, texel rasterisation of one lightmap.  TEST
 * INFRASTRUCTURE ONLY (see lsg_oracle.h).
 *
 * Restates the method of enumerateTexelsPartial, steps 1-2 (RRSolver/lightmap.cpp:147-331):
 * per triangle in order, every texel of its bounding box (lightmap.cpp:147-168) is taken
 * as a square in the triangle's own coordinates (p = uv0 + a (uv1-uv0) + b (uv2-uv0),
 * lightmap.cpp:170-182), clipped by the triangle's three sides b >= 0, a + b <= 1, a >= 0
 * in that order (lightmap.cpp:212-270: vertices strictly inside are kept, the two crossing
 * points added, a square with no vertex strictly inside is dropped), triangulated as a
 * fan from its first vertex into pieces of area Heron(piece in triangle space) times
 * Heron(triangle in map space) (lightmap.cpp:317-331, getArea lightmap.cpp:35-43), and its
 * vertices mark the texel's edge and quadrant flags (lightmap.cpp:276-314).  Texels come
 * out in row-major order with their pieces in triangle order.
 */
#include "lsg_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { QX0Y0 = 1, QX1Y0 = 2, QX0Y1 = 4, QX1Y1 = 8, EX0 = 16, EX1 = 32, EY0 = 64, EY1 = 128 };

typedef struct { float x, y; } p2;

static float heron(p2 a, p2 b, p2 c)
{
    float ab = sqrtf((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y));
    float bc = sqrtf((c.x - b.x) * (c.x - b.x) + (c.y - b.y) * (c.y - b.y));
    float ca = sqrtf((a.x - c.x) * (a.x - c.x) + (a.y - c.y) * (a.y - c.y));
    float s = (ab + bc + ca) * 0.5f;
    return sqrtf(s * (s - ab) * (s - bc) * (s - ca));
}

/* signed distance of q from side k of the unit triangle, positive inside */
static float side_dist(int k, p2 q)
{
    return k == 0 ? q.y : (k == 1 ? 1 - q.x - q.y : q.x);
}

/* clip the convex polygon p (n vertices) by side k; returns the new count */
static int clip_side(p2* p, int n, int k)
{
    float d[8];
    int nin = 0, nout = 0;
    for (int i = 0; i < n; i++) {
        d[i] = side_dist(k, p[i]);
        if (d[i] > 0) nin++;
        else if (d[i] < 0) nout++;
    }
    if (!nout) return n;
    if (!nin) return 0;
    p2 q[8];
    int m = 0;
    for (int i = 0; i < n; i++) {
        int j = i == 0 ? n - 1 : i - 1;          /* edge p[j] -> p[i] */
        int in_i = d[i] > 0, in_j = d[j] > 0;
        if (in_i != in_j) {                       /* the crossing of this edge with the side */
            float s = d[j] / (d[j] - d[i]);
            q[m].x = p[j].x + (p[i].x - p[j].x) * s;
            q[m].y = p[j].y + (p[i].y - p[j].y) * s;
            m++;
        }
        if (in_i) q[m++] = p[i];
    }
    memcpy(p, q, sizeof(p2) * (size_t)m);
    return m;
}

typedef struct { uint32_t texel, tri; float uv[6], area; } piece_t;

/* Returns 0, or -1 when the capacities are too small (counts are set either way).
 * texels: 4 uint32 per texel (out_index, sub_begin, count | flags << 16, key);
 * subs: 8 words per piece (triangle id, uv[6], area). */
int lsgo_rasterize(const float* tri_uv, const uint32_t* tri_ids, uint32_t ntri, uint32_t W, uint32_t H, uint32_t object_id,
                   uint32_t out_base, uint32_t* texels, uint32_t cap_t, float* subs, uint32_t cap_s, uint32_t* ntex, uint32_t* nsub)
{
    size_t npc = 0, cap = 1024;
    piece_t* pcs = (piece_t*)malloc(sizeof(piece_t) * cap);
    uint8_t* flags = (uint8_t*)calloc((size_t)W * H, 1);
    for (uint32_t t = 0; t < ntri; t++) {
        const float* u = tri_uv + 6 * (size_t)t;
        p2 u0 = {u[0], u[1]}, e1 = {u[2] - u[0], u[3] - u[1]}, e2 = {u[4] - u[0], u[5] - u[1]};
        float xmin = W * fminf(u[0], fminf(u[2], u[4])), xmax = W * fmaxf(u[0], fmaxf(u[2], u[4]));
        float ymin = H * fminf(u[1], fminf(u[3], u[5])), ymax = H * fmaxf(u[1], fmaxf(u[3], u[5]));
        if (!(isfinite(xmin) && isfinite(xmax) && isfinite(ymin) && isfinite(ymax))) continue;
        float x0f = xmin > 0 ? xmin : 0, y0f = ymin > 0 ? ymin : 0;
        float x1f = xmax + 1, y1f = ymax + 1;
        x1f = x1f < 0 ? 0 : (x1f > (float)W ? (float)W : x1f);
        y1f = y1f < 0 ? 0 : (y1f > (float)H ? (float)H : y1f);
        uint32_t x0 = (uint32_t)x0f, y0 = (uint32_t)y0f, x1 = (uint32_t)x1f, y1 = (uint32_t)y1f;
        if (x0 >= x1 || y0 >= y1) continue;
        float det = e1.x * e2.y - e1.y * e2.x;
        if (det == 0) continue;
        p2 tri_map[3] = {u0, {u[2], u[3]}, {u[4], u[5]}};
        float tri_area = heron(tri_map[0], tri_map[1], tri_map[2]);
        for (uint32_t y = y0; y < y1; y++)
            for (uint32_t x = x0; x < x1; x++) {
                /* the texel's corners in triangle space */
                p2 poly[8];
                const float cx[4] = {(float)x, (float)x + 1, (float)x + 1, (float)x};
                const float cy[4] = {(float)y, (float)y, (float)y + 1, (float)y + 1};
                for (int k = 0; k < 4; k++) {
                    float dx = cx[k] / W - u0.x, dy = cy[k] / H - u0.y;
                    poly[k].x = (dx * e2.y - dy * e2.x) / det;
                    poly[k].y = (e1.x * dy - e1.y * dx) / det;
                }
                int n = 4;
                for (int k = 0; k < 3 && n; k++) n = clip_side(poly, n, k);
                if (n < 3) continue;
                /* flags from the polygon in texel space, centred on the texel */
                uint8_t* f = &flags[(size_t)x + (size_t)y * W];
                float ox = 0, oy = 0;
                for (int i = -1; i < n; i++) {
                    p2 q = poly[i < 0 ? n - 1 : i];
                    float mx = u0.x + e1.x * q.x + e2.x * q.y, my = u0.y + e1.y * q.x + e2.y * q.y;
                    float vx = mx * W - x - 0.5f, vy = my * H - y - 0.5f;
                    if (i >= 0) {
                        if (vx < -0.5f + 0.03f) *f |= EX0;
                        if (vx > 0.5f - 0.03f) *f |= EX1;
                        if (vy < -0.5f + 0.03f) *f |= EY0;
                        if (vy > 0.5f - 0.03f) *f |= EY1;
                        if (vx < 0 && vy < 0) *f |= QX0Y0;
                        if (vx > 0 && vy < 0) *f |= QX1Y0;
                        if (vx < 0 && vy > 0) *f |= QX0Y1;
                        if (vx > 0 && vy > 0) *f |= QX1Y1;
                        if ((ox < 0) != (vx < 0) && (oy < 0) != (vy < 0)) {
                            float slope = (vy - oy) / (vx - ox), at0 = vy - slope * vx;
                            *f |= at0 < 0 ? (QX0Y0 | QX1Y0) : (QX0Y1 | QX1Y1);
                            *f |= -at0 / slope < 0 ? (QX0Y0 | QX0Y1) : (QX1Y0 | QX1Y1);
                        }
                    }
                    ox = vx; oy = vy;
                }
                for (int i = 1; i + 1 < n; i++) {
                    float a = heron(poly[0], poly[i], poly[i + 1]) * tri_area;
                    if (!isfinite(a)) continue;
                    if (npc == cap) { cap *= 2; pcs = (piece_t*)realloc(pcs, sizeof(piece_t) * cap); }
                    piece_t* p = &pcs[npc++];
                    p->texel = x + y * W; p->tri = tri_ids[t]; p->area = a;
                    p->uv[0] = poly[0].x; p->uv[1] = poly[0].y;
                    p->uv[2] = poly[i].x; p->uv[3] = poly[i].y;
                    p->uv[4] = poly[i + 1].x; p->uv[5] = poly[i + 1].y;
                }
            }
    }
    /* stable counting sort of the pieces by texel (they were made in triangle order) */
    uint32_t* start = (uint32_t*)calloc((size_t)W * H + 1, sizeof(uint32_t));
    for (size_t i = 0; i < npc; i++) start[pcs[i].texel + 1]++;
    uint32_t nt = 0;
    for (size_t k = 0; k < (size_t)W * H; k++) { if (start[k + 1]) nt++; start[k + 1] += start[k]; }
    *ntex = nt;
    *nsub = (uint32_t)npc;
    int ok = texels && subs && cap_t >= nt && cap_s >= npc;
    if (ok) {
        uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)W * H + 1));
        memcpy(fill, start, sizeof(uint32_t) * ((size_t)W * H + 1));
        for (size_t i = 0; i < npc; i++) {
            float* o = subs + 8 * (size_t)fill[pcs[i].texel]++;
            memcpy(o, &pcs[i].tri, 4);
            memcpy(o + 1, pcs[i].uv, 24);
            o[7] = pcs[i].area;
        }
        uint32_t it = 0;
        for (size_t k = 0; k < (size_t)W * H; k++) {
            uint32_t c = start[k + 1] - start[k];
            if (!c) continue;
            uint32_t* T = texels + 4 * (size_t)it++;
            T[0] = out_base + (uint32_t)k;
            T[1] = start[k];
            T[2] = c | ((uint32_t)flags[k] << 16);
            T[3] = object_id * 0x9E3779B1u + (uint32_t)k;
        }
        free(fill);
    }
    free(start); free(pcs); free(flags);
    return ok ? 0 : -1;
}
