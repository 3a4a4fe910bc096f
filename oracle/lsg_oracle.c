/*
 * lsg_oracle.c -- CPU oracle, collider part.  TEST INFRASTRUCTURE ONLY (see
 * lsg_oracle.h): the checker, never the thing measured or shipped.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -fPIC -shared (oracle/Makefile).
 * All float arithmetic is scalar IEEE binary32 in the operation order of the
 * cited reference lines, so results are bit-identical to the reference's
 * x86-64 build (which also has no FMA contraction: -O2 without -march).
 */
#include "lsg_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline float absf(float a) { return fabsf(a); }
static inline int finitef(float a) { return isfinite(a); }

/* RRCollider/IntersectBspFast.cpp:130-187  TriangleSRLNP::setGeometry */
void lsgo_srlnp_setup(const float* a, const float* b, const float* c, lsgo_tri* t)
{
    for (int i = 0; i < 3; i++) {
        t->s[i] = a[i];
        t->r[i] = b[i] - a[i];
        t->l[i] = c[i] - a[i];
    }
    for (int i = 0; i < 3; i++) t->ir[i] = 1 / t->r[i];
    const float* r = t->r; const float* l = t->l;
    float rxy = l[1] * r[0] - l[0] * r[1];
    float ryz = l[2] * r[1] - l[1] * r[2];
    float rzx = l[0] * r[2] - l[2] * r[0];
    int byte;
    if (absf(rxy) >= absf(ryz)) {
        if (absf(rxy) >= absf(rzx)) { byte = 0; t->ireal = 1 / rxy; }
        else                        { byte = 2; t->ireal = 1 / rzx; }
    } else {
        if (absf(ryz) >= absf(rzx)) { byte = 1; t->ireal = 1 / ryz; }
        else                        { byte = 2; t->ireal = 1 / rzx; }
    }
    if (absf(r[0]) >= absf(r[1])) {
        if (absf(r[0]) >= absf(r[2])) byte += 3;
    } else {
        if (absf(r[1]) >= absf(r[2])) byte += 6;
    }
    /* RRMathPrivate.h:55-58 orthogonalTo(a,b) = cross; :36-39 normalized = a/size(a) */
    float nx = r[1] * l[2] - r[2] * l[1];
    float ny = -r[0] * l[2] + r[2] * l[0];
    float nz = r[0] * l[1] - r[1] * l[0];
    float len = sqrtf(nx * nx + ny * ny + nz * nz);
    t->n[0] = nx / len; t->n[1] = ny / len; t->n[2] = nz / len;
    t->n[3] = -(t->s[0] * t->n[0] + t->s[1] * t->n[1] + t->s[2] * t->n[2]);
    if (!(finitef(t->n[0]) && finitef(t->n[1]) && finitef(t->n[2]))) byte = 10;
    t->byte = byte;
}

/* projection axes for intersectByte, RRCollider/IntersectBspFast.cpp:202-204 */
static const int k_tablo[11][3] = {{0,1,2},{1,2,2},{2,0,2},{0,1,0},{1,2,0},{2,0,0},{0,1,1},{1,2,1},{2,0,1},{2,0,1},{2,0,1}};

/* Plane distance as in RRCollider/IntersectBspFast.cpp:349-354 (kd leaf),
 * hit point as IntersectLinear.cpp:40-45, triangle test as
 * IntersectBspFast.cpp:189-252.  Accepts t in [tlo,thi] (both inclusive). */
int lsgo_srlnp_hit(const lsgo_tri* t, const float* o, const float* d, float tlo, float thi,
                   float* t_out, float* u_out, float* v_out, int* front)
{
    const float* n = t->n;
    float nDotDir = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
    float nDotOrigin = o[0] * n[0] + o[1] * n[1] + o[2] * n[2] + n[3];
    float dist = -nDotOrigin / nDotDir;
    if (!(dist >= tlo && dist <= thi)) return 0;
    float p[3];
    p[0] = o[0] + d[0] * dist;
    p[1] = o[1] + d[1] * dist;
    p[2] = o[2] + d[2] * dist;
    int x = k_tablo[t->byte][0], y = k_tablo[t->byte][1], z = k_tablo[t->byte][2];
    float v = ((p[y] - t->s[y]) * t->r[x] - (p[x] - t->s[x]) * t->r[y]) * t->ireal;
    if (v < 0 || v > 1) return 0;
    float u = (p[z] - t->s[z] - t->l[z] * v) * t->ir[z];
    if (u < 0 || u + v > 1) return 0;
    if (t->byte > 8) return 0;
    *t_out = dist; *u_out = u; *v_out = v;
    *front = (d[0] * n[0] + d[1] * n[1] + d[2] * n[2]) < 0;
    return 1;
}

/* RRMesh/RRMesh.cpp:164-194 RRMesh::getAABB (over all vertices, non-finite skipped) */
void lsgo_mesh_aabb(const float* v, uint32_t nv, float* mn, float* mx)
{
    for (int j = 0; j < 3; j++) { mn[j] = 1e37f; mx[j] = -1e37f; }
    for (uint32_t i = 0; i < nv; i++)
        for (int j = 0; j < 3; j++) {
            float c = v[3 * i + j];
            if (finitef(c)) {
                mn[j] = (mn[j] < c) ? mn[j] : c;   /* RR_MIN */
                mx[j] = (mx[j] > c) ? mx[j] : c;   /* RR_MAX */
            }
        }
    for (int j = 0; j < 3; j++) if (mn[j] > mx[j]) mn[j] = mx[j] = 0;
}

/* RRCollider/geometry.cpp:20-31 Box::init */
void lsgo_box_init(float* mn, float* mx)
{
    float d = (mx[0] - mn[0] + mx[1] - mn[1] + mx[2] - mn[2]) * 1e-5f;
    for (int j = 0; j < 3; j++) { mn[j] -= d; mx[j] += d; }
}

static inline float minf_(float a, float b) { return a < b ? a : b; }
static inline float maxf_(float a, float b) { return a > b ? a : b; }

/* RRCollider/geometry.cpp:38-86 Box::intersect, scalar path (USE_SSE is off on Linux) */
int lsgo_box_intersect(const float* mn, const float* mx, const float* o, const float* d,
                       float rmin, float rmax, float* lo, float* hi)
{
    float l1 = (mn[0] - o[0]) / d[0];
    float l2 = (mx[0] - o[0]) / d[0];
    float lmin = minf_(l1, l2);
    float lmax = maxf_(l1, l2);
    l1 = (mn[1] - o[1]) / d[1];
    l2 = (mx[1] - o[1]) / d[1];
    lmin = maxf_(minf_(l1, l2), lmin);
    lmax = minf_(maxf_(l1, l2), lmax);
    l1 = (mn[2] - o[2]) / d[2];
    l2 = (mx[2] - o[2]) / d[2];
    lmin = maxf_(minf_(l1, l2), lmin);
    lmax = minf_(maxf_(l1, l2), lmax);
    lmin = maxf_(rmin, lmin);
    lmax = minf_(rmax, lmax);
    *lo = lmin; *hi = lmax;
    return (lmax >= 0.f) & (lmax >= lmin);
}

/* RRCollider/IntersectLinear.cpp:133-139 DELTA_BSP (from the inflated box) */
float lsgo_delta_bsp(const float* mn, const float* mx)
{
    float tx = fabsf(mx[0]) > fabsf(mn[0]) ? fabsf(mx[0]) : fabsf(mn[0]);
    float ty = fabsf(mx[1]) > fabsf(mn[1]) ? fabsf(mx[1]) : fabsf(mn[1]);
    float tz = fabsf(mx[2]) > fabsf(mn[2]) ? fabsf(mx[2]) : fabsf(mn[2]);
    float m = tx > ty ? (tx > tz ? tx : tz) : (ty > tz ? ty : tz);
    if (m == 0 || isnan(m)) m = 1;
    return m * 1e-5f;
}

static void tri_verts(const float* verts, const uint32_t* idx, uint32_t t, const float** a, const float** b, const float** c)
{
    *a = verts + 3 * (size_t)idx[3 * (size_t)t + 0];
    *b = verts + 3 * (size_t)idx[3 * (size_t)t + 1];
    *c = verts + 3 * (size_t)idx[3 * (size_t)t + 2];
}

static void fill_hit(lsgo_hit* h, const lsgo_tri* tr, int tri, float t, float u, float v, int front, const float* o, const float* d)
{
    h->tri = tri; h->t = t; h->u = u; h->v = v; h->front = front;
    for (int j = 0; j < 3; j++) h->p[j] = o[j] + d[j] * t;
    for (int j = 0; j < 4; j++) h->plane[j] = tr->n[j];
}

static int nthreads_(int threads)
{
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads; return 1;
#endif
}

/* Brute-force closest hit with the SRLNP test over every non-degenerate
 * triangle: the textbook oracle for what IT_BSP_FASTER should return.  Ties in t
 * go to the lowest triangle index. */
void lsgo_intersect_linear_srlnp(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                                 const float* rays, uint32_t nrays, lsgo_hit* out, int threads)
{
    lsgo_tri* tris = (lsgo_tri*)malloc(sizeof(lsgo_tri) * (ntri ? ntri : 1));
    for (uint32_t t = 0; t < ntri; t++) {
        const float *a, *b, *c; tri_verts(verts, idx, t, &a, &b, &c);
        lsgo_srlnp_setup(a, b, c, &tris[t]);
    }
    float mn[3], mx[3];
    lsgo_mesh_aabb(verts, nverts, mn, mx);
    lsgo_box_init(mn, mx);
    int nt = nthreads_(threads);
    #pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (long r = 0; r < (long)nrays; r++) {
        const float* ray = rays + 8 * (size_t)r;
        const float* o = ray; const float* d = ray + 4;
        lsgo_hit* h = &out[r];
        h->tri = -1;
        float lo, hi;
        if (!ntri || !lsgo_box_intersect(mn, mx, o, d, ray[3], ray[7], &lo, &hi)) continue;
        float best = hi; int bi = -1; float bu = 0, bv = 0; int bf = 0;
        for (uint32_t t = 0; t < ntri; t++) {
            if (tris[t].byte > 8) continue;
            float tt, u, v; int f;
            if (lsgo_srlnp_hit(&tris[t], o, d, lo, best, &tt, &u, &v, &f)) {
                if (bi < 0 || tt < best) { best = tt; bi = (int)t; bu = u; bv = v; bf = f; }
            }
        }
        if (bi >= 0) fill_hit(h, &tris[bi], bi, best, bu, bv, bf, o, d);
    }
    free(tris);
}

/* RRCollider/IntersectLinear.cpp:62-110 intersect_triangle (Moller-Trumbore),
 * driven as IntersectLinear::intersect (:161-273): box pretest, then every
 * triangle in index order with hitDistanceMax shrinking to each accepted hit
 * (so among equal distances the LAST triangle wins).  fp64=1 evaluates the
 * same formulas in double for closed-form checks. */
void lsgo_intersect_mt(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                       const float* rays, uint32_t nrays, lsgo_hit* out, int fp64, int threads)
{
    float mn[3], mx[3];
    lsgo_mesh_aabb(verts, nverts, mn, mx);
    lsgo_box_init(mn, mx);
    int nt = nthreads_(threads);
    #pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (long r = 0; r < (long)nrays; r++) {
        const float* ray = rays + 8 * (size_t)r;
        const float* o = ray; const float* d = ray + 4;
        lsgo_hit* h = &out[r];
        h->tri = -1;
        float lo, hi;
        if (!ntri || !lsgo_box_intersect(mn, mx, o, d, ray[3], ray[7], &lo, &hi)) continue;
        float hmax = hi;
        for (uint32_t t = 0; t < ntri; t++) {
            const float *a, *b, *c; tri_verts(verts, idx, t, &a, &b, &c);
            float s1[3], s2[3];
            for (int j = 0; j < 3; j++) { s1[j] = b[j] - a[j]; s2[j] = c[j] - a[j]; }
            if (!fp64) {
                /* RRMathPrivate.h:55-58 orthogonalTo(dir, side2) */
                float pv[3] = { d[1] * s2[2] - d[2] * s2[1], -d[0] * s2[2] + d[2] * s2[0], d[0] * s2[1] - d[1] * s2[0] };
                float det = s1[0] * pv[0] + s1[1] * pv[1] + s1[2] * pv[2];
                int fr = det > 0;
                if (det == 0) continue;
                float tv[3] = { o[0] - a[0], o[1] - a[1], o[2] - a[2] };
                float u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) / det;
                if (u < 0 || u > 1) continue;
                float qv[3] = { tv[1] * s1[2] - tv[2] * s1[1], -tv[0] * s1[2] + tv[2] * s1[0], tv[0] * s1[1] - tv[1] * s1[0] };
                float v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) / det;
                if (v < 0 || u + v > 1) continue;
                float dist = (s2[0] * qv[0] + s2[1] * qv[1] + s2[2] * qv[2]) / det;
                if (dist < lo || dist > hmax) continue;
                hmax = dist;
                h->tri = (int)t; h->t = dist; h->u = u; h->v = v; h->front = fr;
            } else {
                double D[3] = { d[0], d[1], d[2] }, S1[3] = { s1[0], s1[1], s1[2] }, S2[3] = { s2[0], s2[1], s2[2] };
                double pv[3] = { D[1] * S2[2] - D[2] * S2[1], -D[0] * S2[2] + D[2] * S2[0], D[0] * S2[1] - D[1] * S2[0] };
                double det = S1[0] * pv[0] + S1[1] * pv[1] + S1[2] * pv[2];
                if (det == 0) continue;
                double tv[3] = { (double)o[0] - a[0], (double)o[1] - a[1], (double)o[2] - a[2] };
                double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) / det;
                if (u < 0 || u > 1) continue;
                double qv[3] = { tv[1] * S1[2] - tv[2] * S1[1], -tv[0] * S1[2] + tv[2] * S1[0], tv[0] * S1[1] - tv[1] * S1[0] };
                double v = (D[0] * qv[0] + D[1] * qv[1] + D[2] * qv[2]) / det;
                if (v < 0 || u + v > 1) continue;
                double dist = (S2[0] * qv[0] + S2[1] * qv[1] + S2[2] * qv[2]) / det;
                if (dist < lo || dist > hmax) continue;
                hmax = (float)dist;
                h->tri = (int)t; h->t = (float)dist; h->u = (float)u; h->v = (float)v; h->front = det > 0;
            }
        }
        if (h->tri >= 0) {
            /* IntersectLinear.cpp:40-60 update_hitPoint3d / update_hitPlane */
            const float *a, *b, *c; tri_verts(verts, idx, (uint32_t)h->tri, &a, &b, &c);
            float s1[3], s2[3];
            for (int j = 0; j < 3; j++) { s1[j] = b[j] - a[j]; s2[j] = c[j] - a[j]; h->p[j] = o[j] + d[j] * h->t; }
            float n[3] = { s1[1] * s2[2] - s1[2] * s2[1], s1[2] * s2[0] - s1[0] * s2[2], s1[0] * s2[1] - s1[1] * s2[0] };
            float siz = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int j = 0; j < 3; j++) h->plane[j] = n[j] / siz;
            h->plane[3] = -(a[0] * h->plane[0] + a[1] * h->plane[1] + a[2] * h->plane[2]);
        }
    }
}

/* ---- traversal of liblsgpu's BVH2 layout (64-byte nodes) ---------------------------
 * node floats: [0..3] = lo0.x hi0.x lo0.y hi0.y, [4..7] = lo0.z hi0.z lo1.x hi1.x,
 *              [8..11] = lo1.y hi1.y lo1.z hi1.z, [12] = ref0, [13] = ref1 (int bits).
 * A ref with bit 31 set is a leaf of ((ref&15)+1) triangles starting at
 * ((ref>>4)&0x7ffffff) in leaf order.  Child boxes are conservative (inflated),
 * so the result depends only on the triangle tests, not on the node tests. */
static inline void slab(const float* lohi3, const float* o, const float* idir, float tlo, float thi, float* tn, float* tf)
{
    /* lohi3 = {lo.x, hi.x, lo.y, hi.y, lo.z, hi.z} */
    float a0 = (lohi3[0] - o[0]) * idir[0], b0 = (lohi3[1] - o[0]) * idir[0];
    float a1 = (lohi3[2] - o[1]) * idir[1], b1 = (lohi3[3] - o[1]) * idir[1];
    float a2 = (lohi3[4] - o[2]) * idir[2], b2 = (lohi3[5] - o[2]) * idir[2];
    float n = fmaxf(fmaxf(fminf(a0, b0), fminf(a1, b1)), fmaxf(fminf(a2, b2), tlo));
    float f = fminf(fminf(fmaxf(a0, b0), fmaxf(a1, b1)), fminf(fmaxf(a2, b2), thi));
    *tn = n; *tf = f;
}

void lsgo_bvh_traverse(const float* nodes, uint32_t nnodes, int32_t root_ref,
                       const uint32_t* prim, uint32_t nprim,
                       const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                       const float* rays, uint32_t nrays, lsgo_hit* out, lsgo_stats* stats, int threads)
{
    (void)nnodes;
    lsgo_tri* tris = (lsgo_tri*)malloc(sizeof(lsgo_tri) * (ntri ? ntri : 1));
    for (uint32_t t = 0; t < ntri; t++) {
        const float *a, *b, *c; tri_verts(verts, idx, t, &a, &b, &c);
        lsgo_srlnp_setup(a, b, c, &tris[t]);
    }
    float mn[3], mx[3];
    lsgo_mesh_aabb(verts, nverts, mn, mx);
    lsgo_box_init(mn, mx);
    double nv = 0, tt_ = 0;
    int nt = nthreads_(threads);
    #pragma omp parallel for schedule(dynamic, 64) num_threads(nt) reduction(+:nv,tt_)
    for (long r = 0; r < (long)nrays; r++) {
        const float* ray = rays + 8 * (size_t)r;
        const float* o = ray; const float* d = ray + 4;
        lsgo_hit* h = &out[r];
        h->tri = -1;
        float lo, hi;
        if (!nprim || !lsgo_box_intersect(mn, mx, o, d, ray[3], ray[7], &lo, &hi)) continue;
        float idir[3];
        for (int j = 0; j < 3; j++) {
            float dj = d[j];
            if (fabsf(dj) < 1e-30f) dj = copysignf(1e-30f, dj);
            idir[j] = 1.0f / dj;
        }
        float best = hi; int bi = -1; float bu = 0, bv = 0; int bf = 0;
        int32_t stack[256]; int sp = 0;
        stack[sp++] = root_ref;
        while (sp) {
            int32_t ref = stack[--sp];
            if (ref < 0) {
                uint32_t start = ((uint32_t)ref >> 4) & 0x7ffffffu, cnt = ((uint32_t)ref & 15u) + 1;
                for (uint32_t k = 0; k < cnt; k++) {
                    uint32_t t = prim[start + k];
                    tt_ += 1;
                    float th, u, v; int f;
                    if (lsgo_srlnp_hit(&tris[t], o, d, lo, best, &th, &u, &v, &f)) {
                        if (bi < 0 || th < best || (th == best && (int)t < bi)) { best = th; bi = (int)t; bu = u; bv = v; bf = f; }
                    }
                }
                continue;
            }
            nv += 1;
            const float* nd = nodes + 16 * (size_t)ref;
            float b0[6] = { nd[0], nd[1], nd[2], nd[3], nd[4], nd[5] };
            float b1[6] = { nd[6], nd[7], nd[8], nd[9], nd[10], nd[11] };
            int32_t r0, r1; memcpy(&r0, nd + 12, 4); memcpy(&r1, nd + 13, 4);
            float n0, f0, n1, f1;
            slab(b0, o, idir, lo, best, &n0, &f0);
            slab(b1, o, idir, lo, best, &n1, &f1);
            int h0 = n0 <= f0, h1 = n1 <= f1;
            if (h0 && h1) {
                if (n1 < n0) { stack[sp++] = r0; stack[sp++] = r1; }
                else         { stack[sp++] = r1; stack[sp++] = r0; }
            } else if (h0) stack[sp++] = r0;
            else if (h1) stack[sp++] = r1;
        }
        if (bi >= 0) fill_hit(h, &tris[bi], bi, best, bu, bv, bf, o, d);
    }
    if (stats) { stats->node_visits += nv; stats->tri_tests += tt_; stats->rays += nrays; }
    free(tris);
}

/* ---- HomogenousFiller2, RRSolver/gather.cpp:57-90 -------------------------------- */
void lsgo_filler_triangle_point(uint32_t n, float* a, float* b)
{
    static const float dir[4][3] = {{0,0,-0.5f},{0,1,0.5f},{0.86602540378444f,-0.5f,0.5f},{-0.86602540378444f,-0.5f,0.5f}};
    float x = 0, y = 0, dist = 1;
    while (n) {
        x += dist * dir[n & 3][0];
        y += dist * dir[n & 3][1];
        dist *= dir[n & 3][2];
        n >>= 2;
    }
    *a = x; *b = y;
}

/* GetCirclePoint: draws from *n onwards until inside the unit circle; returns
 * the index used and advances *n past it. */
uint32_t lsgo_filler_circle_point(uint32_t* n, float* a, float* b, float* dist)
{
    for (;;) {
        uint32_t k = (*n)++;
        lsgo_filler_triangle_point(k, a, b);
        float dd = *a * *a + *b * *b;
        if (!(dd >= 1)) { *dist = dd; return k; }
    }
}

/* ---- LBVH build (CPU statement of liblsgpu's GPU build) ---------------------------
 * Not a reference algorithm: the reference builds a BSP/kd tree (bsp.cpp).  This
 * is the checker for the GPU LBVH: same Morton quantisation, same stable sort,
 * same Karras split search, same leaf collapse, so node arrays must match the
 * GPU's bit for bit.  Pinned by tree invariants (every valid triangle in exactly
 * one leaf, child boxes contain their triangles) and by traversal == brute force. */
static inline uint32_t expand10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

static inline uint32_t quant10(float c, float mn, float scale)
{
    float q = (c - mn) * scale;
    if (!(q > 0)) return 0;
    if (q >= 1023.f) return 1023;
    return (uint32_t)q;
}

typedef struct { uint32_t key, pos; } kp_t;
static int kp_cmp(const void* a, const void* b)
{
    const kp_t* x = (const kp_t*)a; const kp_t* y = (const kp_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
static inline int delta_(const uint32_t* k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    if (k[i] == k[j]) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(k[i] ^ k[j]);
}

/* Returns number of valid primitives; nodes_out must hold 16*max(n-1,1) floats,
 * prim_out n entries.  *root_ref, *depth filled. */
uint32_t lsgo_lbvh_build(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                         uint32_t leafmax, float* nodes_out, uint32_t* prim_out, int32_t* root_ref, int32_t* depth_out)
{
    float mn[3], mx[3], imn[3], imx[3];
    lsgo_mesh_aabb(verts, nverts, mn, mx);
    memcpy(imn, mn, sizeof mn); memcpy(imx, mx, sizeof mx);
    lsgo_box_init(imn, imx);
    float margin = lsgo_delta_bsp(imn, imx);
    float scale[3];
    for (int j = 0; j < 3; j++) { float e = mx[j] - mn[j]; scale[j] = e > 0 ? 1024.0f / e : 0.0f; }
    uint32_t n = 0;
    uint32_t* valid = (uint32_t*)malloc(sizeof(uint32_t) * (ntri ? ntri : 1));
    for (uint32_t t = 0; t < ntri; t++) {
        const float *a, *b, *c; tri_verts(verts, idx, t, &a, &b, &c);
        lsgo_tri tr; lsgo_srlnp_setup(a, b, c, &tr);
        if (tr.byte <= 8) valid[n++] = t;
    }
    float* pbox = (float*)malloc(sizeof(float) * 6 * (n ? n : 1));
    kp_t* kp = (kp_t*)malloc(sizeof(kp_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) {
        const float *a, *b, *c; tri_verts(verts, idx, valid[i], &a, &b, &c);
        float lo[3], hi[3];
        for (int j = 0; j < 3; j++) {
            float l = a[j] < b[j] ? a[j] : b[j]; l = l < c[j] ? l : c[j];
            float h = a[j] > b[j] ? a[j] : b[j]; h = h > c[j] ? h : c[j];
            lo[j] = l; hi[j] = h;
        }
        for (int j = 0; j < 3; j++) { pbox[6 * i + 2 * j] = lo[j]; pbox[6 * i + 2 * j + 1] = hi[j]; }
        uint32_t qx = quant10((lo[0] + hi[0]) * 0.5f, mn[0], scale[0]);
        uint32_t qy = quant10((lo[1] + hi[1]) * 0.5f, mn[1], scale[1]);
        uint32_t qz = quant10((lo[2] + hi[2]) * 0.5f, mn[2], scale[2]);
        kp[i].key = (expand10(qx) << 2) | (expand10(qy) << 1) | expand10(qz);
        kp[i].pos = i;
    }
    qsort(kp, n, sizeof(kp_t), kp_cmp);
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    float* sbox = (float*)malloc(sizeof(float) * 6 * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) {
        keys[i] = kp[i].key;
        prim_out[i] = valid[kp[i].pos];
        memcpy(sbox + 6 * i, pbox + 6 * kp[i].pos, 6 * sizeof(float));
    }
    *depth_out = 0;
    if (n == 0) { *root_ref = 0; goto done; }
    if (n <= leafmax) { *root_ref = (int32_t)(0x80000000u | (0u << 4) | (n - 1)); *depth_out = 0; goto done; }
    {
        int ni = (int)n - 1;
        int* first = (int*)malloc(sizeof(int) * ni); int* last = (int*)malloc(sizeof(int) * ni);
        int* split = (int*)malloc(sizeof(int) * ni); int* parent = (int*)malloc(sizeof(int) * ni);
        for (int i = 0; i < ni; i++) parent[i] = -1;
        for (int i = 0; i < ni; i++) {
            int d = (delta_(keys, n, i, i + 1) - delta_(keys, n, i, i - 1)) > 0 ? 1 : -1;
            int dmin = delta_(keys, n, i, i - d);
            int lmax = 2;
            while (delta_(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
            int l = 0;
            for (int t = lmax / 2; t >= 1; t /= 2)
                if (delta_(keys, n, i, i + (l + t) * d) > dmin) l += t;
            int j = i + l * d;
            int dnode = delta_(keys, n, i, j);
            int s = 0, t = l;
            do { t = (t + 1) >> 1; if (delta_(keys, n, i, i + (s + t) * d) > dnode) s += t; } while (t > 1);
            int g = i + s * d + (d < 0 ? d : 0);
            first[i] = i < j ? i : j; last[i] = i < j ? j : i; split[i] = g;
        }
        for (int i = 0; i < ni; i++) {
            int g = split[i];
            if (first[i] != g) parent[g] = i;
            if (last[i] != g + 1) parent[g + 1] = i;
        }
        /* node boxes: union over the primitive range (exact, order independent) */
        float* nbox = (float*)malloc(sizeof(float) * 6 * ni);
        for (int i = 0; i < ni; i++) {
            float* b = nbox + 6 * i;
            for (int j = 0; j < 3; j++) { b[2 * j] = INFINITY; b[2 * j + 1] = -INFINITY; }
            for (int p = first[i]; p <= last[i]; p++)
                for (int j = 0; j < 3; j++) {
                    float l = sbox[6 * p + 2 * j], h = sbox[6 * p + 2 * j + 1];
                    if (l < b[2 * j]) b[2 * j] = l;
                    if (h > b[2 * j + 1]) b[2 * j + 1] = h;
                }
        }
        int maxd = 0;
        for (int i = 0; i < ni; i++) {
            float* nd = nodes_out + 16 * (size_t)i;
            int g = split[i];
            for (int c = 0; c < 2; c++) {
                int cf = c == 0 ? first[i] : g + 1;
                int cl = c == 0 ? g : last[i];
                int cnt = cl - cf + 1;
                const float* cb;
                float tmp[6];
                if (cnt == 1) cb = sbox + 6 * cf;
                else cb = nbox + 6 * (c == 0 ? g : g + 1);
                for (int j = 0; j < 3; j++) { tmp[2 * j] = cb[2 * j] - margin; tmp[2 * j + 1] = cb[2 * j + 1] + margin; }
                memcpy(nd + 6 * c, tmp, sizeof tmp);
                int32_t ref;
                if ((uint32_t)cnt <= leafmax) ref = (int32_t)(0x80000000u | ((uint32_t)cf << 4) | (uint32_t)(cnt - 1));
                else ref = c == 0 ? g : g + 1;
                memcpy(nd + 12 + c, &ref, 4);
            }
            nd[14] = 0; nd[15] = 0;
        }
        /* depth = longest chain of reachable internal nodes */
        for (int i = 0; i < ni; i++) {
            if ((uint32_t)(last[i] - first[i] + 1) <= leafmax) continue;
            int dd = 1, p = parent[i];
            while (p >= 0) { dd++; p = parent[p]; }
            if (dd > maxd) maxd = dd;
        }
        *depth_out = maxd;
        *root_ref = 0;
        free(first); free(last); free(split); free(parent); free(nbox);
    }
done:
    free(valid); free(pbox); free(kp); free(keys); free(sbox);
    return n;
}
