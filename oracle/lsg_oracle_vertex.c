/*
 * lsg_oracle_vertex.c -- CPU oracle, per-vertex outputs.  TEST INFRASTRUCTURE ONLY (see
 * lsg_oracle.h).
 *
 * Restates the solver's smoothing of per-triangle values into vertices:
 *   Object::buildTopIVertices (RRStaticSolver/interpol.cpp:125-172): for every triangle the
 *     solver keeps (Triangle::setGeometry accepts it, rrcore.cpp:112-146) whose three
 *     corners stitch to three different vertices, each corner is inserted into its stitched
 *     vertex with power = angleBetween of its two edges at the stitched positions
 *     (geometry_v.cpp:95-106, acos by fast_acos, geometry_v.h:114-129) times the
 *     triangle's area (calculateArea, rrcore.cpp:104-110); IVertex::insert
 *     (interpol.cpp:51-73) adds it to the vertex's total;
 *   IVertex::getVertexDataFromTriangleData (RRStaticSolver.cpp:130-145): the vertex value
 *     is sum(value of the corner's triangle * power) / total, zero when the total is zero,
 *     summed in insertion order (triangles in order).
 * Stitching itself is checked against the reference's compiled createOptimizedVertices
 * (oracle/_ref/stitch_ref), not restated here.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } v3;

static v3 sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static float len2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static float len(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static v3 scale_div(v3 a, float s) { v3 r = {a.x / s, a.y / s, a.z / s}; return r; }

static float acos_poly(float x)   /* fast_acos, geometry_v.h:114-129 */
{
    float r = sqrtf(1.0f - x), p = -0.0187293f;
    p *= x; p += 0.0742610f;
    p *= x; p -= 0.2121144f;
    p *= x; p += 1.5707288f;
    return p * r;
}

static v3 at(const float* p, uint32_t i) { v3 r = {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; return r; }

/* area kept by Triangle::setGeometry (rrcore.cpp:112-146, ignoreSmaller* = 0), 0 = dropped */
static float kept_area(v3 v0, v3 v1, v3 v2)
{
    v3 s1 = sub(v1, v0), s2 = sub(v2, v0);
    v3 n = {s1.y * s2.z - s1.z * s2.y, -s1.x * s2.z + s1.z * s2.x, s1.x * s2.y - s1.y * s2.x};
    v3 q = scale_div(n, len(n));
    if (!isfinite(q.x) || !isfinite(q.y) || !isfinite(q.z)) return 0;
    float rsize = len(s1), lsize = len(s2);
    if (!(rsize > 0) || !(lsize > 0)) return 0;
    float psqr = len2(sub(scale_div(s1, rsize), scale_div(s2, lsize)));
    float sina = sqrtf(psqr * (1 - psqr / 4));
    float area0 = sina / 2 * rsize * lsize;
    if (psqr <= 0 || psqr >= 4 || 1 - psqr / 4 <= 0 || sina <= 0 || area0 <= 0) return 0;
    /* minAngle(lsize, rsize, |s2 - s1|), rrcore.cpp:96-102 */
    float a = lsize, b = rsize, c = len(sub(s2, s1));
    float angA = acos_poly((b * b + c * c - a * a) / (2 * b * c));
    float angB = acos_poly((a * a + c * c - b * b) / (2 * a * c));
    float angC = acos_poly((a * a + b * b - c * c) / (2 * a * b));
    float mn = angA < angB ? angA : angB;
    mn = mn < angC ? mn : angC;
    if (!isfinite(area0) || !(mn > 0)) return 0;
    float A = len2(s1), B = len2(s2), C = len2(sub(s2, s1));
    float area = sqrtf(2 * B * C + 2 * C * A + 2 * A * B - A * A - B * B - C * C) / 4;
    if (!isfinite(area) || !(area > 0)) return 0;
    return area;
}

/* corner_group: uint32[ntri*3] stitched vertex per corner (0xFFFFFFFF none); group_pos:
 * float[ngroups*3].  Out: key[ntri*3] (the group the corner counts for, 0xFFFFFFFF when it
 * does not count), power[ntri*3]. */
void lsgo_vertex_corners(const float* verts, const uint32_t* idx, uint32_t ntri, const uint32_t* corner_group,
                         const float* group_pos, uint32_t* key, float* power)
{
    for (uint32_t t = 0; t < ntri; t++) {
        const uint32_t* g = corner_group + 3 * (size_t)t;
        float area = kept_area(at(verts, idx[3 * t]), at(verts, idx[3 * t + 1]), at(verts, idx[3 * t + 2]));
        int keep = area > 0 && g[0] != 0xffffffffu && g[1] != 0xffffffffu && g[2] != 0xffffffffu
                   && g[0] != g[1] && g[0] != g[2] && g[1] != g[2];
        for (int k = 0; k < 3; k++) {
            key[3 * t + k] = keep ? g[k] : 0xffffffffu;
            power[3 * t + k] = 0;
            if (!keep) continue;
            v3 p0 = at(group_pos, g[k]), p1 = at(group_pos, g[(k + 1) % 3]), p2 = at(group_pos, g[(k + 2) % 3]);
            v3 e1 = sub(p1, p0), e2 = sub(p2, p0);
            v3 n1 = scale_div(e1, len(e1)), n2 = scale_div(e2, len(e2));
            float d = n1.x * n2.x + n1.y * n2.y + n1.z * n2.z;
            d = d < 1 ? d : 1;
            d = d > -1 ? d : -1;
            power[3 * t + k] = acos_poly(d) * area;
        }
    }
}

/* tri_data: float4[ntri]; out: float4[ngroups] (rgb = smoothed value, w = 0) */
void lsgo_smooth_vertices(const uint32_t* key, const float* power, uint32_t ntri, const float* tri_data, uint32_t ngroups,
                          float* out)
{
    float* acc = (float*)calloc((size_t)ngroups * 4 + 4, sizeof(float));
    for (uint32_t c = 0; c < 3 * ntri; c++) {
        uint32_t g = key[c];
        if (g == 0xffffffffu || g >= ngroups) continue;
        const float* d = tri_data + 4 * (size_t)(c / 3);
        float w = power[c];
        acc[4 * g] += d[0] * w;
        acc[4 * g + 1] += d[1] * w;
        acc[4 * g + 2] += d[2] * w;
        acc[4 * g + 3] += w;
    }
    for (uint32_t g = 0; g < ngroups; g++) {
        float top = acc[4 * g + 3];
        for (int k = 0; k < 3; k++) out[4 * g + k] = top != 0 ? acc[4 * g + k] / top : 0;
        out[4 * g + 3] = 0;
    }
    free(acc);
}
