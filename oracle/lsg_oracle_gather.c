/*
 * lsg_oracle_gather.c -- CPU oracle, lightmap final-gather part.  TEST
 * INFRASTRUCTURE ONLY (see lsg_oracle.h): the checker, never the thing measured
 * or shipped.
 *
 * Restates RRSolver/gather.cpp:600-836 processTexel (hemisphere gather
 * :164-343, light gather :356-586, reliability loop :652-786) for opaque
 * Lambertian scenes, with the collision-handler rules of
 * RRStaticSolver/pathtracer.h:117-252 and the exitance lookup of
 * RRStaticSolver/pathtracer.cpp:139-350 (solver shortcuts at depth 0).
 *
 * The reference draws texel jitter and filler seeds from the global C rand()
 * (gather.h:102, gather.cpp:742-749), which is not reproducible across threads.
 * This statement replaces exactly those draws with a seeded counter hash
 * (lowbias32, C. Wellons); liblsgpu uses the identical replacement, so both sides
 * shoot the same rays.  The shooter's subtexel follows the reference's own area walk
 * (gather.cpp:645-647, 669, 680-688), carried from series to series.
 *
 * A hemisphere hit reads the flat exitance of the side it sees (direct light times
 * albedo plus emittance, pathtracer.cpp:235-262, 306-316) and, where the scene gives
 * per-corner indirect irradiance, adds the albedo times its interpolation at the hit's
 * hitPoint2d (pathtracer.cpp:326-348, Triangle::getPointMeasure rrcore.cpp:262-273).
 * With point materials (lsgo_gather_pm) a triangle's diffuse colour and emittance are read
 * at the hit point where the quality exceeds the material's minimum, and keyed texels
 * that are fully transparent are passed by hemisphere and shadow rays.
 *
 * Geometry (ray origins, directions, hit tests) is evaluated in binary32 in the
 * reference's operation order, so hit decisions agree bit for bit with the GPU;
 * radiometric sums are accumulated in binary64.
 */
#include "lsg_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct lsgo_light {
    int32_t type;          /* 0 = point (REALISTIC attenuation), 1 = directional */
    float pos[3];
    float dir[3];
    float color[3];
    int32_t cast_shadows;
} lsgo_light;

typedef struct lsgo_gparams {
    uint32_t quality;
    int32_t hemisphere;        /* gather.cpp:197: env || emittance || useCurrentSolution */
    float ray_length_min;      /* minimalSafeDistance, RRSolver.cpp:333 */
    float locality;            /* RRSolver.h:584 default 100000 */
    float rug_distance;        /* RRSolver.h:582 default 0.001 */
    float light_multiplier;    /* direct.lightMultiplier */
    uint32_t seed;
    uint32_t light_rounds;     /* 0: quality/10+1 (gather.cpp:396); else this many rounds */
    float environment_multiplier; /* indirect.environmentMultiplier; 0 = environment not sampled */
} lsgo_gparams;

/* RRSolver's environment: 1 = cube map (6 faces w x w: +x,-x,+y,-y,+z,-z), 2 = 360x180
 * panorama (w x h); rgb: linear float[texels*3]. */
typedef struct lsgo_env {
    int32_t type;
    uint32_t w, h;
    const float* rgb;
    /* blended != 0: the solver also has env1 (NULL or type 0 = none), the two angles and
     * the blend factor (RRSolver::setEnvironment, setEnvironmentBlendFactor), and the
     * gather reads RRBuffer::createEnvironmentBlend of them (RRBufferBlend.cpp:21-66) */
    const struct lsgo_env* env1;
    float angle0, angle1, blend;
    int32_t blended;
} lsgo_env;

typedef struct lsgo_gstats {
    double hemi_rays, shadow_rays, series;
} lsgo_gstats;

static inline uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

uint32_t lsgo_texel_key(uint32_t seed, uint32_t key_id) { return lowbias32(seed ^ lowbias32(key_id)); }

/* seeded replacement of gather.cpp:742-749 (u,v in [0,1], u+v<=1) */
void lsgo_jitter(uint32_t key, uint32_t i, float* u, float* v)
{
    uint32_t a = lowbias32(key + 0x9E3779B9u * (2u * i + 1u)) >> 8;
    uint32_t b = lowbias32(key + 0x9E3779B9u * (2u * i + 2u)) >> 8;
    const uint32_t M = 0xFFFFFFu;
    if (a + b > M) { a = M - a; b = M - b; }
    *u = (float)a / (float)M;
    *v = (float)b / (float)M;
}

/* The subtexel selector of processTexel (gather.cpp:645-647, 669, 680-688): an iterator
 * over the texel's subtexels starting at the first with an area accumulator of minus its
 * area; each series of n shooters steps by areaMax/(n+0.91), and each shooter adds the
 * step, then moves the iterator on (cyclically) and subtracts the new subtexel's area
 * while the accumulator is positive.  Both are carried from series to series.
 * areas: the subtexels' areaInMapSpace, `stride` floats apart. */
typedef struct { float accu; uint32_t it; } walk_t;

static void walk_init(walk_t* w, const float* areas) { w->it = 0; w->accu = -areas[0]; }

static uint32_t walk_next(walk_t* w, const float* areas, uint32_t stride, uint32_t m, float step)
{
    w->accu += step;
    while (w->accu > 0) {
        w->it = w->it + 1 == m ? 0 : w->it + 1;
        w->accu -= areas[stride * w->it];
    }
    return w->it;
}

static float walk_step(float area_max, uint32_t n) { return area_max / ((float)n + 0.91f); }

/* the subtexel of every shooter of nseries series of series_n[k] shooters (out: their
 * sum of entries); areas: m floats */
void lsgo_subtexel_walk(const float* areas, uint32_t m, const uint32_t* series_n, uint32_t nseries, uint32_t* out)
{
    if (!m) return;
    float area_max = 0;
    for (uint32_t j = 0; j < m; j++) area_max += areas[j];
    walk_t w;
    walk_init(&w, areas);
    size_t o = 0;
    for (uint32_t k = 0; k < nseries; k++) {
        const float step = walk_step(area_max, series_n[k]);
        for (uint32_t s = 0; s < series_n[k]; s++) out[o++] = walk_next(&w, areas, 1, m, step);
    }
}

/* ---- point materials ---------------------------------------------------------------
 * RRGammaScaler (RRColorSpace.cpp:24-80): toLinear x -> sign(x) |x|^(1/gamma) with 1/gamma
 * kept in binary32 and pow taken in binary64 (the build calls the double pow for float
 * arguments), rounded to binary32; getLinear of a byte reads the table
 * pow(float(i) / 255, 1/gamma). */
static float cs_to_linear(float a, float gamma)
{
    const float inv = 1 / gamma;
    return a >= 0 ? (float)pow((double)a, (double)inv) : -(float)pow((double)-a, (double)inv);
}

static float byte_to_float(uint8_t b) { return (float)b * 0.003921568627450980392156862745098f; }  /* RR_BYTE2FLOAT */

/* RRBufferInMemory::getElement (RRBufferInMemory.cpp:295-360) */
static void get_element(const lsgo_ptex* t, size_t i, float gamma, float* r)
{
    const uint8_t* b = (const uint8_t*)t->data;
    const float* f = (const float*)t->data;
    if (t->scaled && gamma > 0 && (t->format == 0 || t->format == 2)) {
        const size_t k = t->format == 0 ? 3 * i : 4 * i;
        for (int c = 0; c < 3; c++) r[c] = (float)pow((double)((float)b[k + c] / 255), (double)(1 / gamma));
        r[3] = t->format == 0 ? 1.f : byte_to_float(b[k + 3]);
        return;
    }
    switch (t->format) {
    case 0: r[0] = byte_to_float(b[3 * i]); r[1] = byte_to_float(b[3 * i + 1]); r[2] = byte_to_float(b[3 * i + 2]); r[3] = 1; break;
    case 1: r[0] = byte_to_float(b[3 * i + 2]); r[1] = byte_to_float(b[3 * i + 1]); r[2] = byte_to_float(b[3 * i]); r[3] = 1; break;
    case 2: for (int c = 0; c < 4; c++) r[c] = byte_to_float(b[4 * i + c]); break;
    case 3: r[0] = f[3 * i]; r[1] = f[3 * i + 1]; r[2] = f[3 * i + 2]; r[3] = 1; break;
    case 4: for (int c = 0; c < 4; c++) r[c] = f[4 * i + c]; break;
    case 9: r[0] = r[1] = r[2] = byte_to_float(b[i]); r[3] = 1; break;
    default: r[0] = r[1] = r[2] = f[i]; r[3] = 1; break;
    }
    if (gamma > 0 && t->scaled)
        for (int c = 0; c < 3; c++) r[c] = cs_to_linear(r[c], gamma);
}

/* RRBufferInMemory::getElementAtPosition (RRBufferInMemory.cpp:372-408), depth 1 */
void lsgo_texture_at(const lsgo_ptex* t, float u, float v, float gamma, int interpolated, float* rgba)
{
    const uint32_t w = t->width, h = t->height;
    if (!interpolated) {
        const float px = fmodf(u, 1) + 2, py = fmodf(v, 1) + 2;
        const uint32_t x = (uint32_t)(px * (float)w) % w, y = (uint32_t)(py * (float)h) % h;
        get_element(t, x + (size_t)y * w, gamma, rgba);
        return;
    }
    const float px = (fmodf(u, 1) + 2) * (float)w - 0.5f, py = (fmodf(v, 1) + 2) * (float)h - 0.5f;
    const float rx = fmodf(px, 1), ry = fmodf(py, 1);
    const uint32_t cx = (uint32_t)px, cy = (uint32_t)py;
    float e00[4], e01[4], e10[4], e11[4];
    get_element(t, (cx % w) + (size_t)(cy % h) * w, gamma, e00);
    get_element(t, (cx % w) + (size_t)((cy + 1) % h) * w, gamma, e01);
    get_element(t, ((cx + 1) % w) + (size_t)(cy % h) * w, gamma, e10);
    get_element(t, ((cx + 1) % w) + (size_t)((cy + 1) % h) * w, gamma, e11);
    const float w00 = (1 - rx) * (1 - ry), w01 = (1 - rx) * ry, w10 = rx * (1 - ry), w11 = rx * ry;
    for (int c = 0; c < 4; c++) {
        float a = e00[c] * w00;
        a = a + e01[c] * w01;
        a = a + e10[c] * w10;
        rgba[c] = a + e11[c] * w11;
    }
}

typedef struct { float dif[3], emi[3]; int transparent; } pmat_t;

/* RRObject::getPointMaterial (RRObject.cpp:285-303) with updatePointMaterial (:196-281), for
 * the properties this gather reads: diffuse colour, emittance (x the emittance multiplier,
 * which the pathtracer applies, pathtracer.cpp:313), and whether keying made the texel
 * fully transparent (renderFrom and catchFrom cleared) */
static void point_material(const lsgo_pm* pm, uint32_t tri, float u, float v, int interpolated, pmat_t* m)
{
    const lsgo_pmat* M = pm->mats + pm->tri_mat[tri];
    const float* c = pm->tri_uv + 6 * (size_t)tri;
    const float w = 1 - u - v;
    const float mu = c[0] * w + c[2] * u + c[4] * v, mv = c[1] * w + c[3] * u + c[5] * v;
    const float g = pm->gamma;
    float rgba[4];
    for (int k = 0; k < 3; k++) { m->dif[k] = M->albedo[k]; m->emi[k] = M->emittance[k]; }
    m->transparent = 0;
    if (M->emissive_texture >= 0) {
        lsgo_texture_at(pm->tex + M->emissive_texture, mu, mv, g, interpolated, rgba);
        for (int k = 0; k < 3; k++) m->emi[k] = rgba[k] * M->emittance_multiplier;
    }
    if (M->diffuse_texture >= 0 && M->transmittance_texture == M->diffuse_texture && (M->flags & 1)) {
        /* optimized path: transmittance in the diffuse map's alpha */
        lsgo_texture_at(pm->tex + M->diffuse_texture, mu, mv, 0, interpolated, rgba);
        float t = 1 - rgba[3];
        if (M->flags & 4) t = 1 - t;
        if (M->flags & 2) t = t >= M->threshold ? 1.f : 0.f;
        m->transparent = t == 1;
        float tl = t;
        for (int k = 0; k < 3; k++) m->dif[k] = rgba[k];
        if (g > 0) {
            for (int k = 0; k < 3; k++) m->dif[k] = cs_to_linear(m->dif[k], g);
            tl = cs_to_linear(t, g);
        }
        for (int k = 0; k < 3; k++) m->dif[k] *= 1 - tl;
        return;
    }
    if (M->diffuse_texture >= 0) {
        lsgo_texture_at(pm->tex + M->diffuse_texture, mu, mv, g, interpolated, rgba);
        for (int k = 0; k < 3; k++) m->dif[k] = rgba[k];
    }
    if (M->transmittance_texture >= 0) {
        lsgo_texture_at(pm->tex + M->transmittance_texture, mu, mv, 0, interpolated, rgba);
        float t[3];
        for (int k = 0; k < 3; k++) t[k] = (M->flags & 1) ? 1 - rgba[3] : rgba[k];
        if (M->flags & 4) for (int k = 0; k < 3; k++) t[k] = 1 - t[k];
        if (M->flags & 2) {
            const float avg = (t[0] + t[1] + t[2]) * 0.33333333333333f;
            for (int k = 0; k < 3; k++) t[k] = avg >= M->threshold ? 1.f : 0.f;
        }
        m->transparent = t[0] == 1 && t[1] == 1 && t[2] == 1;
        for (int k = 0; k < 3; k++) m->dif[k] *= 1 - (g > 0 ? cs_to_linear(t[k], g) : t[k]);
    }
}

/* ---- ray casting with the final-gathering handler rules ------------------------ */
typedef struct {
    const float* verts; const uint32_t* idx; uint32_t ntri;
    const lsgo_tri* tris; const uint8_t* side_mask;
    const float* nodes; int32_t root_ref; const uint32_t* prim; uint32_t nprim;
    float bmn[3], bmx[3];
    const lsgo_pm* pm; uint32_t quality;     /* point materials (NULL = none), max(1, quality) */
} scene_t;

/* pathtracer.h:139-166: tri has all three vertices equal to the shooter's */
static int same_location(const scene_t* sc, uint32_t tri, uint32_t shooter)
{
    const float* sv[3];
    for (int i = 0; i < 3; i++) sv[i] = sc->verts + 3 * (size_t)sc->idx[3 * (size_t)shooter + i];
    for (int i = 0; i < 3; i++) {
        const float* hv = sc->verts + 3 * (size_t)sc->idx[3 * (size_t)tri + i];
        int eq = 0;
        for (int k = 0; k < 3; k++)
            if (hv[0] == sv[k][0] && hv[1] == sv[k][1] && hv[2] == sv[k][2]) eq = 1;
        if (!eq) return 0;
    }
    return 1;
}

/* mode 0 = hemisphere (accept rendered side), 1 = shadow (any shadow caster); with point
 * materials a texel that keying made fully transparent is passed (pathtracer.h:194-216:
 * hemisphere rays read point materials at quality, shadow rays at 2 x quality,
 * interpolated when quality / 10 exceeds the material's minimum, gather.cpp:175-180, 366-370) */
static int accept_hit(const scene_t* sc, int mode, uint32_t tri, int front, float t, uint32_t shooter, float tmin,
                      float u, float v)
{
    if (tri == shooter) return 0;                                   /* pathtracer.h:133 */
    if (t < 1000 * tmin && same_location(sc, tri, shooter)) return 0; /* pathtracer.h:139-166 */
    uint8_t m = sc->side_mask[tri];
    int ok = mode == 0 ? (m & (front ? 1 : 2)) != 0                 /* TEST_BIT, renderFrom of hit side */
                       : (m & 4) != 0;                              /* shadow caster, both sides [#45] */
    if (ok && sc->pm && sc->pm->tri_mat[tri] != 0xFFFFFFFFu) {
        const lsgo_pmat* M = sc->pm->mats + sc->pm->tri_mat[tri];
        const uint32_t qpoint = mode == 0 ? sc->quality : 2 * sc->quality;
        const uint32_t qinterp = mode == 0 ? 0 : sc->quality / 10;
        if (qpoint > M->min_quality) {
            pmat_t P;
            point_material(sc->pm, tri, u, v, qinterp > M->min_quality, &P);
            if (P.transparent) ok = 0;
        }
    }
    return ok;
}

static inline void slab6(const float* b, const float* o, const float* idir, float tlo, float thi, float* tn, float* tf)
{
    float a0 = (b[0] - o[0]) * idir[0], b0 = (b[1] - o[0]) * idir[0];
    float a1 = (b[2] - o[1]) * idir[1], b1 = (b[3] - o[1]) * idir[1];
    float a2 = (b[4] - o[2]) * idir[2], b2 = (b[5] - o[2]) * idir[2];
    *tn = fmaxf(fmaxf(fminf(a0, b0), fminf(a1, b1)), fmaxf(fminf(a2, b2), tlo));
    *tf = fminf(fminf(fmaxf(a0, b0), fmaxf(a1, b1)), fminf(fmaxf(a2, b2), thi));
}

static void consider(const scene_t* sc, int mode, uint32_t t, const float* o, const float* d, float lo, float tmin,
                     uint32_t shooter, float* best, int* bi, int* bf, float* buv)
{
    float th, u, v; int f;
    if (!lsgo_srlnp_hit(&sc->tris[t], o, d, lo, *best, &th, &u, &v, &f)) return;
    if (!accept_hit(sc, mode, t, f, th, shooter, tmin, u, v)) return;
    if (*bi < 0 || th < *best || (th == *best && (int)t < *bi)) { *best = th; *bi = (int)t; *bf = f; buv[0] = u; buv[1] = v; }
}

/* Closest accepted hit in [tmin,tmax] after the mesh-box pretest
 * (IntersectBspFast.cpp:816).  Returns tri or -1. */
static int trace(const scene_t* sc, int mode, const float* o, const float* d, float tmin, float tmax,
                 uint32_t shooter, float* t_out, int* front_out, float* uv_out)
{
    float lo, hi;
    if (!lsgo_box_intersect(sc->bmn, sc->bmx, o, d, tmin, tmax, &lo, &hi)) return -1;
    float best = hi; int bi = -1, bf = 0;
    float buv[2] = {0, 0};
    if (!sc->nodes) {
        for (uint32_t t = 0; t < sc->ntri; t++)
            if (sc->tris[t].byte <= 8) consider(sc, mode, t, o, d, lo, tmin, shooter, &best, &bi, &bf, buv);
    } else if (sc->nprim) {
        float idir[3];
        for (int j = 0; j < 3; j++) { float dj = d[j]; if (fabsf(dj) < 1e-30f) dj = copysignf(1e-30f, dj); idir[j] = 1.0f / dj; }
        int32_t stack[256]; int sp = 0; stack[sp++] = sc->root_ref;
        while (sp) {
            int32_t ref = stack[--sp];
            if (ref < 0) {
                uint32_t st = ((uint32_t)ref >> 4) & 0x7ffffffu, cnt = ((uint32_t)ref & 15u) + 1;
                for (uint32_t k = 0; k < cnt; k++) consider(sc, mode, sc->prim[st + k], o, d, lo, tmin, shooter, &best, &bi, &bf, buv);
                continue;
            }
            const float* nd = sc->nodes + 16 * (size_t)ref;
            int32_t r0, r1; memcpy(&r0, nd + 12, 4); memcpy(&r1, nd + 13, 4);
            float n0, f0, n1, f1;
            slab6(nd, o, idir, lo, best, &n0, &f0);
            slab6(nd + 6, o, idir, lo, best, &n1, &f1);
            int h0 = n0 <= f0, h1 = n1 <= f1;
            if (h0 && h1) { if (n1 < n0) { stack[sp++] = r0; stack[sp++] = r1; } else { stack[sp++] = r1; stack[sp++] = r0; } }
            else if (h0) stack[sp++] = r0;
            else if (h1) stack[sp++] = r1;
        }
    }
    if (bi >= 0) { *t_out = best; *front_out = bf; if (uv_out) { uv_out[0] = buv[0]; uv_out[1] = buv[1]; } }
    return bi;
}

/* ---- basis: gather.cpp:727-737, RRMesh.cpp:70-75, RRMathPrivate.h:48-58 ------- */
static void subtexel_basis(const float* vn, const uint32_t* tidx, const float* sub, float* n, float* tg, float* bt)
{
    float cx = (sub[1] + sub[3] + sub[5]) * 0.333333333f;
    float cy = (sub[2] + sub[4] + sub[6]) * 0.333333333f;
    float w = 1 - cx - cy;
    const float* n0 = vn + 3 * (size_t)tidx[0];
    const float* n1 = vn + 3 * (size_t)tidx[1];
    const float* n2 = vn + 3 * (size_t)tidx[2];
    float m[3];
    for (int j = 0; j < 3; j++) m[j] = n0[j] * w + n1[j] * cx + n2[j] * cy;
    float len = sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    for (int j = 0; j < 3; j++) n[j] = m[j] / len;
    float o[3];
    if (fabsf(n[0]) > fabsf(n[2])) { o[0] = -n[1]; o[1] = n[0]; o[2] = 0; }
    else { o[0] = 0; o[1] = -n[2]; o[2] = n[1]; }
    float ol = sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    for (int j = 0; j < 3; j++) tg[j] = o[j] / ol;
    bt[0] = n[1] * tg[2] - n[2] * tg[1];
    bt[1] = -n[0] * tg[2] + n[2] * tg[0];
    bt[2] = n[0] * tg[1] - n[1] * tg[0];
}

/* asin in binary32 by Abramowitz & Stegun 4.4.46 (|error| <= 2e-8 in exact arithmetic):
 * asin(x) = pi/2 - sqrt(1-x) * (a0 + a1 x + ... + a7 x^7) for 0 <= x <= 1, odd.  Plain
 * float multiplies and adds in a fixed order, so the GPU and the CPU statement round
 * alike. */
static float asin_as(float x)
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

/* RRBufferInMemory::getElementAtDirection (RRBufferInMemory.cpp:410-455): index of the
 * nearest texel in direction d.  Cube faces per the OpenGL cube-map table (major axis m,
 * face coordinates (sc/|m|+1)/2 and (tc/|m|+1)/2); panoramas by yaw of (x, z) and
 * asin(y), in float as the reference, with asin_as for asin. */
uint32_t lsgo_env_index(const lsgo_env* env, const float* d)
{
    const uint32_t w = env->w, h = env->h;
    if (env->type == 1) {
        float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
        uint32_t axis = (ax >= ay && ax >= az) ? 0 : ((ay >= ax && ay >= az) ? 1 : 2);
        float m = d[axis];
        uint32_t face = 2 * axis + (m < 0 ? 1 : 0);
        float am = fabsf(m);
        /* (sc, tc) per face: +x (-z,-y), -x (+z,-y), +y (+x,+z), -y (+x,-z), +z (+x,-y), -z (-x,-y) */
        static const int sc_axis[6] = {2, 2, 0, 0, 0, 0}, tc_axis[6] = {1, 1, 2, 2, 1, 1};
        static const float sc_sign[6] = {-1, 1, 1, 1, 1, -1}, tc_sign[6] = {-1, -1, 1, -1, -1, -1};
        float sc = sc_sign[face] * d[sc_axis[face]];
        float tc = tc_sign[face] * d[tc_axis[face]];
        int c0 = (int)((sc / am + 1) * (0.5f * (float)w));
        int c1 = (int)((tc / am + 1) * (0.5f * (float)w));
        c0 = c0 < 0 ? 0 : (c0 > (int)w - 1 ? (int)w - 1 : c0);
        c1 = c1 < 0 ? 0 : (c1 > (int)w - 1 ? (int)w - 1 : c1);
        return (uint32_t)c0 + (uint32_t)c1 * w + face * w * w;
    }
    const float pi = 3.14159265358979323846f;
    float d2 = d[0] * d[0] + d[2] * d[2];
    uint32_t c0 = 0;
    if (d2 != 0.f) {
        float s = d[0] / sqrtf(d2);
        float angle = asin_as(s);
        if (d[2] < 0) angle = pi - angle;
        c0 = (uint32_t)fmaxf(0.f, (angle * (-0.5f / pi) + 0.75f) * (float)w) % w;   /* >= 0 for the cast */
    }
    float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float el = asin_as(d[1] / len);
    uint32_t c1 = (uint32_t)fmaxf(0.f, (el * (1.0f / pi) + 0.5f) * (float)h) % h;
    return c0 + c1 * w;
}

/* RRMatrix3x4::rotationByYawPitchRoll((yaw, 0, 0)), RRMath.cpp:457-470: the double
 * expressions, each entry rounded to float */
static void yaw_matrix(float yaw, float* m)
{
    const double c1 = cos((double)yaw), c2 = cos(0.0), c3 = cos(0.0);
    const double s1 = sin((double)yaw), s2 = sin(0.0), s3 = sin(0.0);
    m[0] = (float)(c1 * c3 + s1 * s2 * s3); m[1] = (float)(c3 * s1 * s2 - c1 * s3); m[2] = (float)(c2 * s1);
    m[3] = (float)(c2 * s3); m[4] = (float)(c2 * c3); m[5] = (float)(-s2);
    m[6] = (float)(c1 * s2 * s3 - c3 * s1); m[7] = (float)(s1 * s3 + c1 * c3 * s2); m[8] = (float)(c1 * c2);
}

/* direction d turned by the yaw rotation of an environment: rotationByYawPitchRoll((yaw,
 * 0, 0)) then RRMatrix3x4::transformDirection (RRMath.cpp:212-219) */
void lsgo_env_rotate(float yaw, const float* d, float* r)
{
    float m[9];
    yaw_matrix(yaw, m);
    const float x = d[0] * m[0] + d[1] * m[1] + d[2] * m[2];
    const float y = d[0] * m[3] + d[1] * m[4] + d[2] * m[5];
    const float z = d[0] * m[6] + d[1] * m[7] + d[2] * m[8];
    r[0] = x; r[1] = y; r[2] = z;
}

/* createEnvironmentBlend (RRBufferBlend.cpp:21-30) makes a blend unless there is no
 * second environment and the first is not rotated */
static int blend_on(const lsgo_env* env)
{
    return env->blended && ((env->env1 && env->env1->type) || env->angle0 != 0);
}

/* RRBuffer::getElementAtDirection of the environment, or of the blend
 * (RRBufferBlend::getElementAtDirection, RRBufferBlend.cpp:49-66): this(R(angle0) d) *
 * (1 - blend) + env1(R(angle1) d) * blend, an absent one giving 0, R(a) =
 * RRMatrix3x4::rotationByYawPitchRoll((a, 0, 0)); rgb */
static void env_color(const lsgo_env* env, const float* d, float* out)
{
    if (!blend_on(env)) {
        const float* t = env->rgb + 3 * (size_t)lsgo_env_index(env, d);
        for (int c = 0; c < 3; c++) out[c] = t[c];
        return;
    }
    float c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
    const lsgo_env* e[2] = {env, env->env1};
    const float ang[2] = {env->angle0, env->angle1};
    float* col[2] = {c0, c1};
    for (int k = 0; k < 2; k++) {
        if (!e[k] || !e[k]->type) continue;
        float r[3];
        lsgo_env_rotate(ang[k], d, r);
        const float* t = e[k]->rgb + 3 * (size_t)lsgo_env_index(e[k], r);
        for (int c = 0; c < 3; c++) col[k][c] = t[c];
    }
    const float k0 = 1 - env->blend, k1 = env->blend;
    for (int c = 0; c < 3; c++) out[c] = c0[c] * k0 + c1[c] * k1;
}

/* Skewed-normalised tangent and bitangent at the subtexel centroid: the normalised
 * barycentric blend of the vertex frames (gather.cpp:731-732). */
static void skewed_frame(const float* vt, const float* vb, const uint32_t* tidx, const float* sub, float* tg, float* bg)
{
    float cx = (sub[1] + sub[3] + sub[5]) * 0.333333333f;
    float cy = (sub[2] + sub[4] + sub[6]) * 0.333333333f;
    float w = 1 - cx - cy;
    const float* F[2] = {vt, vb};
    float* out[2] = {tg, bg};
    for (int k = 0; k < 2; k++) {
        const float* a = F[k] + 3 * (size_t)tidx[0];
        const float* b = F[k] + 3 * (size_t)tidx[1];
        const float* c = F[k] + 3 * (size_t)tidx[2];
        float m[3];
        for (int j = 0; j < 3; j++) m[j] = a[j] * w + b[j] * cx + c[j] * cy;
        float len = sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        for (int j = 0; j < 3; j++) out[k][j] = m[j] / len;
    }
}

/* g_lightmapDirections[i] (gather.cpp:43-49) expressed in the frame (t, b, n) and
 * normalised (gather.cpp:292-293, 458-459). */
static void layer_dir(int i, const float* t, const float* b, const float* n, float* out)
{
    float g0 = 0.f, g1 = 0.f, g2 = 1.f;
    if (i == 1) { g1 = sqrtf(6.0f) / 3; g2 = 1 / sqrtf(3.0f); }
    else if (i >= 2) { g0 = (i == 2 ? -1 : 1) / sqrtf(2.0f); g1 = -1 / sqrtf(6.0f); g2 = 1 / sqrtf(3.0f); }
    float v[3];
    for (int j = 0; j < 3; j++) v[j] = t[j] * g0 + b[j] * g1 + n[j] * g2;
    float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int j = 0; j < 3; j++) out[j] = v[j] / len;
}

/* TangentBasis::buildBasisFromNormal (RRMesh.cpp:70-75) per vertex: the frames of a
 * mesh without tangents (RRMeshArrays.cpp:297-305). */
void lsgo_default_frames(const float* vn, uint32_t nv, float* vt, float* vb)
{
    for (uint32_t v = 0; v < nv; v++) {
        const float* n = vn + 3 * (size_t)v;
        float o[3];
        if (fabsf(n[0]) > fabsf(n[2])) { o[0] = -n[1]; o[1] = n[0]; o[2] = 0; }
        else { o[0] = 0; o[1] = -n[2]; o[2] = n[1]; }
        float len = sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        float* t = vt + 3 * (size_t)v;
        float* b = vb + 3 * (size_t)v;
        for (int j = 0; j < 3; j++) t[j] = o[j] / len;
        b[0] = n[1] * t[2] - n[2] * t[1];
        b[1] = -n[0] * t[2] + n[2] * t[0];
        b[2] = n[0] * t[1] - n[1] * t[0];
    }
}

/* Number of light shots among the first s shooters of a series of n shooters
 * with r rounds, sequential rule of gather.cpp:777-783 (for r < n). */
uint32_t lsgo_shots_before(uint32_t s, uint32_t n, uint32_t r)
{
    uint32_t c = 0;
    for (uint32_t k = 0; k < s; k++)
        if ((uint64_t)c * (n - 1) < (uint64_t)r * k) c++;
    return c;
}

/* processTexel for every texel.  out_rgba: the lightmap (LS_LIGHTMAP) as the
 * reference gathers it without directional layers; layers (optional, 5 planes:
 * LS_LIGHTMAP, LS_DIRECTION1..3, LS_BENT_NORMALS, NULL entries skipped): the output
 * layers of updateLightmap, with the directional weighting of gather.cpp:284-305 and
 * :453-481 when any of planes 1..3 is given (gatherAllDirections, lightmap.cpp:820),
 * and the bent normal of gather.cpp:306, :482, :809-830. */
static void gather_core(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                        const float* vnormals, const float* vtan, const float* vbit, const uint8_t* side_mask,
                        const float* exitance, const float* ind, const lsgo_light* lights, uint32_t nlights, const lsgo_gparams* p,
                        const uint32_t* texels, uint32_t ntex, const float* subs,
                        const float* nodes, int32_t root_ref, const uint32_t* prim, uint32_t nprim,
                        const lsgo_env* env_in, const lsgo_pm* pm, double* out_rgba, double* const* layers, lsgo_gstats* stats,
                        int threads)
{
    const int dirs = layers && (layers[1] || layers[2] || layers[3]);
    /* the environment is sampled only with a nonzero multiplier (gather.cpp:149); the first
     * one decides the hemisphere gather (gather.cpp:149, 197), the blend of both what a
     * ray that leaves sees (pathtracer.cpp:29-34) */
    const int env0 = env_in && env_in->type && p->environment_multiplier != 0;
    const lsgo_env* env = (env_in && p->environment_multiplier != 0 &&
                           (env_in->type || (blend_on(env_in) && env_in->env1 && env_in->env1->type))) ? env_in : NULL;
    scene_t sc;
    sc.verts = verts; sc.idx = idx; sc.ntri = ntri; sc.side_mask = side_mask;
    sc.nodes = nodes; sc.root_ref = root_ref; sc.prim = prim; sc.nprim = nprim;
    sc.pm = pm; sc.quality = p->quality > 1 ? p->quality : 1;      /* gather.cpp:865 */
    lsgo_tri* tris = (lsgo_tri*)malloc(sizeof(lsgo_tri) * (ntri ? ntri : 1));
    for (uint32_t t = 0; t < ntri; t++)
        lsgo_srlnp_setup(verts + 3 * (size_t)idx[3 * t], verts + 3 * (size_t)idx[3 * t + 1], verts + 3 * (size_t)idx[3 * t + 2], &tris[t]);
    sc.tris = tris;
    lsgo_mesh_aabb(verts, nverts, sc.bmn, sc.bmx);
    lsgo_box_init(sc.bmn, sc.bmx);

    const uint32_t q = p->quality;
    const uint32_t hemi_rays = (p->hemisphere || env0) ? (q > 1 ? q : 1) : 0;       /* gather.cpp:197 */
    const uint32_t rounds = (p->light_multiplier != 0 && nlights) ? (p->light_rounds ? p->light_rounds : q / 10 + 1) : 0;  /* gather.cpp:396 */
    const uint32_t light_rays = nlights * rounds;
    double nh = 0, ns = 0, nser = 0;
    int nt = 1;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_max_threads();
#endif
    #pragma omp parallel for schedule(dynamic, 4) num_threads(nt) reduction(+:nh,ns,nser)
    for (long ti = 0; ti < (long)ntex; ti++) {
        const uint32_t* tx = texels + 4 * (size_t)ti;
        uint32_t out_index = tx[0], sb = tx[1], sc_cnt = tx[2] & 0xffffu, flags = (tx[2] >> 16) & 0xffu;
        const float* ts = subs + 8 * (size_t)sb;
        double* out = out_rgba ? out_rgba + 4 * (size_t)out_index : NULL;
        if (out) out[0] = out[1] = out[2] = out[3] = 0;
        if (layers)
            for (int i = 0; i < 5; i++)
                if (layers[i]) for (int c = 0; c < 4; c++) layers[i][4 * (size_t)out_index + c] = 0;
        if (!sc_cnt || (!hemi_rays && !light_rays)) continue;
        double sumHL[4][3], sumLL[4][3], bent[3] = {0, 0, 0};
        memset(sumHL, 0, sizeof sumHL); memset(sumLL, 0, sizeof sumLL);
        uint32_t key = lsgo_texel_key(p->seed, tx[3]);
        uint32_t filler = key & 0xffffu;
        float area_max = 0;
        for (uint32_t j = 0; j < sc_cnt; j++) area_max += ts[8 * j + 7];
        double sumH[3] = {0, 0, 0}, sumL[3] = {0, 0, 0};
        uint32_t relH = 0, relL = 0, unrelL = 0, shotRounds = 0, inside = 0;
        uint32_t shooter_base = 0;
        walk_t wk;
        walk_init(&wk, ts + 7);
        for (;;) {
            int shootH = (relH < hemi_rays || relH * 10 < hemi_rays) && 0 < hemi_rays * 100;
            int shootL = (relL + unrelL < light_rays || relL * 10 < light_rays) && unrelL < light_rays * 100;
            if (!shootH && !shootL) break;
            uint32_t n = (shootH ? hemi_rays : 0) > (shootL ? rounds : 0) ? (shootH ? hemi_rays : 0) : (shootL ? rounds : 0);
            uint32_t shotH = 0, shotL = 0;
            nser += 1;
            const float astep = walk_step(area_max, n);
            for (uint32_t s = 0; s < n; s++) {
                uint32_t j = walk_next(&wk, ts + 7, 8, sc_cnt, astep);
                const float* sub = ts + 8 * j;
                uint32_t tri; memcpy(&tri, sub, 4);
                const uint32_t* tidx = idx + 3 * (size_t)tri;
                const float* v0 = verts + 3 * (size_t)tidx[0];
                const float* v1 = verts + 3 * (size_t)tidx[1];
                const float* v2 = verts + 3 * (size_t)tidx[2];
                float bn[3], bt[3], bb[3];
                subtexel_basis(vnormals, tidx, sub, bn, bt, bb);
                float ld[4][3];
                if (dirs) {
                    float st[3], sb2[3];
                    skewed_frame(vtan, vbit, tidx, sub, st, sb2);
                    for (int i = 0; i < 4; i++) layer_dir(i, st, sb2, bn, ld[i]);
                }
                float ju, jv; lsgo_jitter(key, shooter_base + s, &ju, &jv);
                float ux = sub[1] + (sub[3] - sub[1]) * ju + (sub[5] - sub[1]) * jv;
                float uy = sub[2] + (sub[4] - sub[2]) * ju + (sub[6] - sub[2]) * jv;
                float o[3];
                for (int c = 0; c < 3; c++) { float s1 = v1[c] - v0[c], s2 = v2[c] - v0[c]; o[c] = v0[c] + s1 * ux + s2 * uy; }
                /* gather.cpp:759-768 */
                if (shootH && (!shootL || hemi_rays >= rounds || (uint64_t)shotH * (n - 1) < (uint64_t)hemi_rays * s)) {
                    shotH++;
                    float x, y, dist;
                    lsgo_filler_circle_point(&filler, &x, &y, &dist);
                    float cosa = sqrtf(1 - dist);
                    float d[3];
                    for (int c = 0; c < 3; c++) d[c] = bn[c] * cosa + bt[c] * x + bb[c] * y;
                    float th, huv[2]; int fr;
                    int hit = trace(&sc, 0, o, d, p->ray_length_min, p->locality, tri, &th, &fr, huv);
                    nh += 1;
                    /* a ray that leaves the scene returns environment * multiplier
                     * (pathtracer.cpp:156-173) */
                    float envc[3], hitc[3];
                    const float* e = NULL;
                    if (hit >= 0 && pm && pm->tri_mat[hit] != 0xFFFFFFFFu) {
                        /* a triangle with a material: the table holds the side's direct
                         * irradiance; getIncidentRadiance composes direct x diffuse colour,
                         * + emittance where the side emits, + indirect x diffuse colour
                         * (pathtracer.cpp:235-262, 306-316, 326-348) */
                        const lsgo_pmat* M = pm->mats + pm->tri_mat[hit];
                        const float* tab = exitance + 4 * (2 * (size_t)hit + (fr ? 0 : 1));
                        float dif[3], emi[3];
                        for (int k = 0; k < 3; k++) { dif[k] = M->albedo[k]; emi[k] = M->emittance[k]; }
                        if (sc.quality > M->min_quality) {
                            pmat_t P;
                            point_material(pm, (uint32_t)hit, huv[0], huv[1], 0, &P);
                            for (int k = 0; k < 3; k++) { dif[k] = P.dif[k]; emi[k] = P.emi[k]; }
                        }
                        for (int k = 0; k < 3; k++) hitc[k] = tab[k] * dif[k];
                        if (M->emit_sides & (fr ? 1u : 2u)) for (int k = 0; k < 3; k++) hitc[k] = hitc[k] + emi[k];
                        if (ind && (side_mask[hit] & (fr ? 8 : 16))) {
                            const float* c = ind + 12 * (size_t)hit;
                            const float u = huv[0], v = huv[1], w = 1 - u - v;
                            for (int k = 0; k < 3; k++) hitc[k] = hitc[k] + (c[k] * w + c[4 + k] * u + c[8 + k] * v) * dif[k];
                        }
                        e = hitc;
                    } else if (hit >= 0) {
                        e = exitance + 4 * (2 * (size_t)hit + (fr ? 0 : 1));
                        if (ind && (side_mask[hit] & (fr ? 8 : 16))) {
                            /* + albedo * (I0 (1-u-v) + I1 u + I2 v), pathtracer.cpp:346, rrcore.cpp:266-268 */
                            const float* c = ind + 12 * (size_t)hit;
                            const float u = huv[0], v = huv[1], w = 1 - u - v;
                            for (int k = 0; k < 3; k++) hitc[k] = e[k] + (c[k] * w + c[4 + k] * u + c[8 + k] * v) * c[4 * k + 3];
                            e = hitc;
                        }
                    } else if (env) {
                        env_color(env, d, envc);
                        for (int c = 0; c < 3; c++) envc[c] = envc[c] * p->environment_multiplier;
                        e = envc;
                    }
                    if (e) {
                        for (int c = 0; c < 3; c++) sumH[c] += e[c];
                        if (layers) {
                            if (dirs) {
                                float ni1 = d[0] * bn[0] + d[1] * bn[1] + d[2] * bn[2];
                                if (ni1 > 0) {
                                    float inv = 1 / ni1;
                                    for (int i = 0; i < 4; i++) {
                                        float n2 = d[0] * ld[i][0] + d[1] * ld[i][1] + d[2] * ld[i][2];
                                        if (n2 > 0) { float k = n2 * inv; for (int c = 0; c < 3; c++) sumHL[i][c] += (double)e[c] * k; }
                                    }
                                }
                            } else for (int c = 0; c < 3; c++) sumHL[0][c] += e[c];
                            double w = ((double)fabsf(e[0]) + fabsf(e[1]) + fabsf(e[2])) / 3;
                            for (int c = 0; c < 3; c++) bent[c] += d[c] * w;
                        }
                    }
                    relH++;
                }
                /* gather.cpp:775-784 */
                if (shootL && (!shootH || rounds >= hemi_rays || (uint64_t)shotL * (n - 1) < (uint64_t)rounds * s)) {
                    shotL++;
                    for (uint32_t li = 0; li < nlights; li++) {
                        const lsgo_light* L = &lights[li];
                        float d[3];
                        if (L->type == 1) { d[0] = -L->dir[0]; d[1] = -L->dir[1]; d[2] = -L->dir[2]; }
                        else { d[0] = L->pos[0] - o[0]; d[1] = L->pos[1] - o[1]; d[2] = L->pos[2] - o[2]; }
                        float dsz = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                        d[0] /= dsz; d[1] /= dsz; d[2] /= dsz;
                        if (L->type == 1) dsz *= p->locality;
                        float ni = d[0] * bn[0] + d[1] * bn[1] + d[2] * bn[2];
                        if (ni <= 0 || !isfinite(ni)) { relL++; continue; }    /* gather.cpp:430-437 */
                        float th = 0; int fr;
                        int blocked = L->cast_shadows ? (trace(&sc, 1, o, d, p->ray_length_min, dsz, tri, &th, &fr, NULL) >= 0) : 0;
                        ns += L->cast_shadows ? 1 : 0;
                        if (!blocked) {
                            /* RRLight.cpp:345-362 getIrradiance */
                            double att = 1.0;
                            if (L->type == 0) {
                                double dx = (double)o[0] - L->pos[0], dy = (double)o[1] - L->pos[1], dz = (double)o[2] - L->pos[2];
                                att = 1.0 / (0.000000000137289 + (dx * dx + dy * dy + dz * dz));
                            }
                            for (int c = 0; c < 3; c++) sumL[c] += (double)L->color[c] * att * p->light_multiplier * ni;
                            if (layers) {
                                double ir[3];
                                for (int c = 0; c < 3; c++) ir[c] = (double)L->color[c] * att * p->light_multiplier;
                                if (dirs) {
                                    for (int i = 0; i < 4; i++) {
                                        float n2 = d[0] * ld[i][0] + d[1] * ld[i][1] + d[2] * ld[i][2];
                                        if (n2 > 0) for (int c = 0; c < 3; c++) sumLL[i][c] += ir[c] * n2;
                                    }
                                } else for (int c = 0; c < 3; c++) sumLL[0][c] += ir[c] * ni;
                                double w = (fabs(ir[0]) + fabs(ir[1]) + fabs(ir[2])) / 3 * ni;
                                for (int c = 0; c < 3; c++) bent[c] += d[c] * w;
                            }
                            relL++;
                        } else if (th < p->rug_distance) unrelL++;              /* gather.cpp:500-505 */
                        else relL++;
                    }
                    shotRounds++;
                }
            }
            shooter_base += n;
        }
        double rgb[3] = {0, 0, 0};
        int reliable = 0;
        const int useH = hemi_rays && relH, useL = light_rays && relL && !(inside > light_rays);
        if (useH) { for (int c = 0; c < 3; c++) rgb[c] += sumH[c] / relH; reliable = 1; }
        if (useL) { for (int c = 0; c < 3; c++) rgb[c] += sumL[c] / shotRounds; reliable = 1; }
        const double alpha = reliable ? flags / 255.0 : 0.0;
        if (out) { out[0] = rgb[0]; out[1] = rgb[1]; out[2] = rgb[2]; out[3] = alpha; }
        if (layers) {
            for (int i = 0; i < 4; i++) {
                if (!layers[i]) continue;
                double* o4 = layers[i] + 4 * (size_t)out_index;
                for (int c = 0; c < 3; c++)
                    o4[c] = (useH ? sumHL[i][c] / relH : 0.0) + (useL ? sumLL[i][c] / shotRounds : 0.0);
                o4[3] = alpha;
            }
            if (layers[4]) {
                double x = bent[0], y = bent[1], z = bent[2];
                if (x != 0 || y != 0 || z != 0) { double len = sqrt(x * x + y * y + z * z); x /= len; y /= len; z /= len; }
                double* o4 = layers[4] + 4 * (size_t)out_index;
                o4[0] = (x + 1) * 0.5; o4[1] = (y + 1) * 0.5; o4[2] = (z + 1) * 0.5; o4[3] = alpha;
            }
        }
    }
    if (stats) { stats->hemi_rays += nh; stats->shadow_rays += ns; stats->series += nser; }
    free(tris);
}

void lsgo_gather(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                 const float* vnormals, const uint8_t* side_mask, const float* exitance,
                 const lsgo_light* lights, uint32_t nlights, const lsgo_gparams* p,
                 const uint32_t* texels, uint32_t ntex, const float* subs,
                 const float* nodes, int32_t root_ref, const uint32_t* prim, uint32_t nprim,
                 double* out_rgba, lsgo_gstats* stats, int threads)
{
    gather_core(verts, nverts, idx, ntri, vnormals, NULL, NULL, side_mask, exitance, NULL, lights, nlights, p, texels, ntex,
                subs, nodes, root_ref, prim, nprim, NULL, NULL, out_rgba, NULL, stats, threads);
}

/* lsgo_gather with an environment (NULL = none) and per-corner indirect irradiance (NULL =
 * none): ind = 12 floats per triangle, corner k's rgb at 4k..4k+2 and albedo channel k at
 * 4k+3, added on the sides whose side_mask bit 3 (front) / 4 (back) is set. */
void lsgo_gather_env(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                     const float* vnormals, const uint8_t* side_mask, const float* exitance, const float* ind,
                     const lsgo_light* lights, uint32_t nlights, const lsgo_gparams* p,
                     const uint32_t* texels, uint32_t ntex, const float* subs,
                     const float* nodes, int32_t root_ref, const uint32_t* prim, uint32_t nprim,
                     const lsgo_env* env, double* out_rgba, lsgo_gstats* stats, int threads)
{
    gather_core(verts, nverts, idx, ntri, vnormals, NULL, NULL, side_mask, exitance, ind, lights, nlights, p, texels, ntex,
                subs, nodes, root_ref, prim, nprim, env, NULL, out_rgba, NULL, stats, threads);
}

/* lsgo_gather with the output layers of updateLightmap.  vtan/vbit: per-vertex frames
 * (NULL = lsgo_default_frames of the normals); layers: 5 planes of float64 RGBA, NULL
 * for a layer not wanted. */
void lsgo_gather_layers(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                        const float* vnormals, const float* vtan, const float* vbit, const uint8_t* side_mask,
                        const float* exitance, const float* ind, const lsgo_light* lights, uint32_t nlights,
                        const lsgo_gparams* p, const uint32_t* texels, uint32_t ntex, const float* subs,
                        const float* nodes, int32_t root_ref, const uint32_t* prim, uint32_t nprim,
                        const lsgo_env* env, double* const* layers, lsgo_gstats* stats, int threads)
{
    float* ft = NULL;
    float* fb = NULL;
    if (!vtan || !vbit) {
        ft = (float*)malloc(12 * (size_t)(nverts ? nverts : 1));
        fb = (float*)malloc(12 * (size_t)(nverts ? nverts : 1));
        lsgo_default_frames(vnormals, nverts, ft, fb);
        vtan = ft; vbit = fb;
    }
    gather_core(verts, nverts, idx, ntri, vnormals, vtan, vbit, side_mask, exitance, ind, lights, nlights, p, texels, ntex,
                subs, nodes, root_ref, prim, nprim, env, NULL, NULL, layers, stats, threads);
    free(ft); free(fb);
}

/* lsgo_gather_layers with point materials (pm NULL = none); out_rgba: the lightmap alone
 * (layers NULL), else layers as lsgo_gather_layers. */
void lsgo_gather_pm(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                    const float* vnormals, const float* vtan, const float* vbit, const uint8_t* side_mask,
                    const float* exitance, const float* ind, const lsgo_light* lights, uint32_t nlights,
                    const lsgo_gparams* p, const uint32_t* texels, uint32_t ntex, const float* subs,
                    const float* nodes, int32_t root_ref, const uint32_t* prim, uint32_t nprim,
                    const lsgo_env* env, const lsgo_pm* pm, double* out_rgba, double* const* layers, lsgo_gstats* stats,
                    int threads)
{
    float* ft = NULL;
    float* fb = NULL;
    if (layers && (!vtan || !vbit)) {
        ft = (float*)malloc(12 * (size_t)(nverts ? nverts : 1));
        fb = (float*)malloc(12 * (size_t)(nverts ? nverts : 1));
        lsgo_default_frames(vnormals, nverts, ft, fb);
        vtan = ft; vbit = fb;
    }
    gather_core(verts, nverts, idx, ntri, vnormals, vtan, vbit, side_mask, exitance, ind, lights, nlights, p, texels, ntex,
                subs, nodes, root_ref, prim, nprim, env, pm, layers ? NULL : out_rgba, layers, stats, threads);
    free(ft); free(fb);
}
