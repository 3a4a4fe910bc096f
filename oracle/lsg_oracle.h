/*
 * lsg_oracle.h -- CPU oracle for the ray-cast + lightmap-gather hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * entry point and the CPU-baseline leg of the benchmark may load it; the
 * product (liblsgpu.so and the lightsprintsdk_amd package) never does.
 *
 * Every function restates one piece of LightsprintCore's algorithm in plain C
 * (compiled with -ffp-contract=off, scalar SSE float, no -ffast-math), citing
 * the reference file:line it follows.  Where it states pinning, the function is
 * checked in tests/ against golden vectors produced by the compiled reference
 * collider (oracle/_ref), closed forms or brute force.
 *
 * Pinning status:
 *   lsgo_srlnp_setup / lsgo_srlnp_hit  -- pinned: golden vectors from the
 *       compiled IT_BSP_FASTER + HelloCollider closed-form KAT + FP64 check.
 *   lsgo_box_*                          -- pinned: slab closed form + goldens.
 *   lsgo_intersect_linear_srlnp         -- pinned: goldens vs IT_BSP_FASTER.
 *   lsgo_intersect_mt                   -- pinned: goldens vs IT_LINEAR.
 *   lsgo_bvh_traverse                   -- pinned: equals brute force.
 *   lsgo_filler_*                       -- pinned: hand-derived KATs.
 *   lsgo_gather_*                       -- pinned: closed forms (point light,
 *       uniform-exitance enclosure, emitter form factor); reference bake is
 *       unbuildable here (needs boost), so bit-level parity with
 *       processTexel itself is "parity unpinned" beyond those closed forms.
 *   lsgo_texture_at and the point materials of lsgo_gather_pm -- pinned:
 *       texel lookups against hand-placed texels (nearest texel, bilinear
 *       weights at known offsets, wrap-around), byte formats against the
 *       reference colour space's byte table and toLinear compiled from
 *       RRColorSpace.cpp, and whole gathers against closed forms (a keyed
 *       checkerboard occluder passes half the light, a textured wall
 *       reflects its texel colour), each with planted mistakes.
 *   lsgo_vertex_corners / lsgo_smooth_vertices -- pinned: closed forms of
 *       the angle x area weights (equal fans, a right-angle corner, constant and
 *       linear fields) with planted weight mistakes; the grouping they take is
 *       the compiled reference's createOptimizedVertices (oracle/_ref/stitch_ref).
 */
#ifndef LSG_ORACLE_H
#define LSG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-triangle precompute, field-for-field the TriangleSRLNP record of
 * RRCollider/IntersectBspFast.h:33-49. byte==10 marks a degenerate triangle. */
typedef struct lsgo_tri {
    float n[4];      /* plane: normalised normal + w */
    float s[3];      /* vertex0 */
    float ireal;     /* 1/cross in the projection plane */
    float r[3];      /* vertex1-vertex0 */
    float ir[3];     /* 1/r */
    float l[3];      /* vertex2-vertex0 */
    int32_t byte;    /* 0..8 projection selector, 10 = degenerate */
} lsgo_tri;

/* One closest-hit result. tri = -1 on miss (other fields undefined). */
typedef struct lsgo_hit {
    int32_t tri;
    float t, u, v;
    float p[3];
    float plane[4];
    int32_t front;
} lsgo_hit;

/* Traversal statistics of lsgo_bvh_traverse (summed over rays). */
typedef struct lsgo_stats {
    double node_visits;   /* internal nodes whose child boxes were tested */
    double tri_tests;     /* triangle plane+SRLNP tests */
    double rays;
} lsgo_stats;

void lsgo_srlnp_setup(const float* v0, const float* v1, const float* v2, lsgo_tri* out);
int  lsgo_srlnp_hit(const lsgo_tri* t, const float* o, const float* d, float tlo, float thi,
                    float* t_out, float* u, float* v, int* front);
void lsgo_mesh_aabb(const float* verts, uint32_t nverts, float* mn, float* mx);
void lsgo_box_init(float* mn, float* mx);
int  lsgo_box_intersect(const float* mn, const float* mx, const float* o, const float* d,
                        float rmin, float rmax, float* lo, float* hi);
float lsgo_delta_bsp(const float* mn, const float* mx);

/* rays: 8 floats each (ox,oy,oz,tmin, dx,dy,dz,tmax). */
void lsgo_intersect_linear_srlnp(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                                 const float* rays, uint32_t nrays, lsgo_hit* out, int threads);
void lsgo_intersect_mt(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                       const float* rays, uint32_t nrays, lsgo_hit* out, int fp64, int threads);

/* Traverses the BVH2 node layout built by liblsgpu (64-byte nodes, leaf refs
 * with bit 31 set: start = (ref>>4)&0x7ffffff, count = (ref&15)+1) over the
 * triangles in leaf order (prim[i] = original triangle index).  mesh gives the
 * vertices; root_box is the inflated mesh box.  Fills out and stats. */
void lsgo_bvh_traverse(const float* nodes, uint32_t nnodes, int32_t root_ref,
                       const uint32_t* prim, uint32_t nprim,
                       const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                       const float* rays, uint32_t nrays, lsgo_hit* out, lsgo_stats* stats, int threads);

uint32_t lsgo_lbvh_build(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri,
                         uint32_t leafmax, float* nodes_out, uint32_t* prim_out, int32_t* root_ref, int32_t* depth_out);

/* Point materials of the gather (RRObject::getPointMaterial, RRObject.cpp:196-303): a
 * material per triangle (0xFFFFFFFF = none), the material uv of each triangle's corners,
 * and textures as RRBufferInMemory stores them (the oracle converts colours itself, from
 * gamma).  lsgo_pmat is field for field the library's lsg_point_material, so one array
 * serves both. */
typedef struct lsgo_pmat {
    int32_t diffuse_texture, emissive_texture, transmittance_texture;
    uint32_t flags;                 /* 1 transmittance in alpha, 2 keyed, 4 inverted */
    float threshold;
    uint32_t min_quality;
    float albedo[3];                /* diffuseReflectance.colorLinear */
    float emittance[3];             /* diffuseEmittance.colorLinear x materialEmittanceMultiplier */
    float emittance_multiplier;
    uint32_t emit_sides;            /* bit 0 front, bit 1 back */
} lsgo_pmat;
typedef struct lsgo_ptex {
    uint32_t width, height;
    int32_t format;                 /* BF_RGB 0, BGR 1, RGBA 2, RGBF 3, RGBAF 4, LUMINANCE 9, LUMINANCEF 10 */
    int32_t scaled;
    const void* data;
} lsgo_ptex;
typedef struct lsgo_pm {
    const uint32_t* tri_mat;
    const float* tri_uv;
    const lsgo_pmat* mats;
    const lsgo_ptex* tex;
    float gamma;                    /* RRColorSpace::create_sRGB power, 0 = no colour space */
} lsgo_pm;

/* RRBufferInMemory::getElementAtPosition of a 2D texture (RRBufferInMemory.cpp:372-408),
 * with the colour space of gamma (0 = none) applied as getElement applies it. */
void lsgo_texture_at(const lsgo_ptex* t, float u, float v, float gamma, int interpolated, float* rgba);

/* HomogenousFiller2 (RRSolver/gather.cpp:57-111). */
void  lsgo_filler_triangle_point(uint32_t n, float* a, float* b);
uint32_t lsgo_filler_circle_point(uint32_t* n, float* a, float* b, float* dist);

#ifdef __cplusplus
}
#endif
#endif
