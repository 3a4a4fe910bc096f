/*
 * lsgpu.h -- C ABI of liblsgpu, the MI355X (gfx950) ray-casting and
 * lightmap-gather backend for LightsprintCore.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status returns, no
 * exceptions and no HIP/torch types in any signature (streams and events are
 * void*).  Every entry point states the reference interface it replaces.
 *
 * Status codes: 0 = ok, negative = error; lsg_last_error() gives the message
 * of the last failure on the calling thread.
 */
#ifndef LSGPU_H
#define LSGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSG_OK              0
#define LSG_ERR_NO_DEVICE  -1   /* no usable gfx950 device: caller falls back (RRCollider.cpp:236-239) */
#define LSG_ERR_OOM        -2   /* device allocation failed (RRCollider.cpp:221-225 bad_alloc path) */
#define LSG_ERR_INVALID    -3   /* bad argument */
#define LSG_ERR_ABORTED    -4   /* *aborting was set (RRCollider.h:178, bsp.cpp:705) */
#define LSG_ERR_DEVICE     -5   /* HIP runtime error */
#define LSG_ERR_LIMIT      -6   /* size beyond a documented limit */

/* RRRay::Flags, include/Lightsprint/RRCollider.h:98-106 */
#define LSG_FILL_DISTANCE  (1u << 0)
#define LSG_FILL_POINT3D   (1u << 1)
#define LSG_FILL_POINT2D   (1u << 2)
#define LSG_FILL_PLANE     (1u << 3)
#define LSG_FILL_TRIANGLE  (1u << 4)
#define LSG_FILL_SIDE      (1u << 5)

/* Byte offsets inside the reference's 128-byte RRRay (RRCollider.h:85-137,
 * LP64), used by lsg_collider_intersect_rrrays. */
#define LSG_RRRAY_SIZE            128
#define LSG_RRRAY_ORIGIN            0
#define LSG_RRRAY_DIR              16
#define LSG_RRRAY_LENGTH_MIN       32
#define LSG_RRRAY_LENGTH_MAX       36
#define LSG_RRRAY_FLAGS            40
#define LSG_RRRAY_HIT_DISTANCE     44
#define LSG_RRRAY_HIT_TRIANGLE     48
#define LSG_RRRAY_HIT_POINT2D      52
#define LSG_RRRAY_HIT_PLANE        60
#define LSG_RRRAY_HIT_POINT3D      76
#define LSG_RRRAY_HIT_FRONTSIDE    88

typedef struct lsg_collider lsg_collider;
typedef struct lsg_scene lsg_scene;

/* ---- runtime ---------------------------------------------------------------------- */
const char* lsg_last_error(void);
const char* lsg_version(void);
/* number of usable gfx950 devices (0 when none); never fails */
int lsg_device_count(void);
int lsg_set_device(int device);

/* device-memory plumbing for callers that keep inputs resident in HBM */
int lsg_malloc(void** ptr, size_t bytes);
int lsg_free(void* ptr);
int lsg_memcpy_h2d(void* dst, const void* src, size_t bytes);
int lsg_memcpy_d2h(void* dst, const void* src, size_t bytes);
int lsg_memset(void* dst, int value, size_t bytes);
int lsg_synchronize(void);
int lsg_stream_create(void** stream);
int lsg_stream_destroy(void* stream);
int lsg_stream_synchronize(void* stream);
int lsg_event_create(void** event);
int lsg_event_destroy(void* event);
int lsg_event_record(void* event, void* stream);
int lsg_event_elapsed_ms(void* start, void* stop, float* ms);

/* ---- collider ------------------------------------------------------------------------
 * Replaces RRCollider::create(mesh, nullptr, technique, aborting, cacheLocation,
 * buildParams) (include/Lightsprint/RRCollider.h:183, RRCollider.cpp:186-241) for the
 * technique IT_HIP_LBVH: builds a linear BVH on the GPU (Morton codes, LSD radix sort,
 * Karras radix tree, bottom-up refit) over the mesh given as host arrays, then
 * rebuilds its inner nodes over the same leaves with a binned surface-area heuristic
 * (lsg_bvh_rebuild_sah; the reference's kd splits weigh box surface areas too,
 * RRCollider/bsp.cpp:388-574).
 * Triangles whose SRLNP precompute is degenerate are left out, as the reference's
 * build does (IntersectBsp.h:252-262, IntersectBspFast.cpp:775-780). */
#define LSG_BUILD_LBVH_ONLY 1u   /* keep the LBVH's inner nodes (no SAH rebuild) */
typedef struct lsg_build_params {
    uint32_t leaf_size;        /* max triangles per leaf, 1..16; 0 = 2, or 4/8 when the tree would be deeper than 31 */
    uint32_t flags;            /* LSG_BUILD_* bits, 0 = default */
} lsg_build_params;

typedef struct lsg_collider_info {
    uint32_t num_triangles;    /* as given */
    uint32_t num_valid;        /* non-degenerate triangles in the tree */
    uint32_t num_nodes;        /* internal node slots (num_valid-1, >= 0) */
    int32_t  root_ref;         /* root reference (leaf refs have bit 31 set) */
    uint32_t depth;            /* longest chain of internal nodes */
    uint32_t leaf_size;
    float    box_min[3];       /* inflated mesh box (geometry.cpp:20-31) */
    float    box_max[3];
    float    node_margin;      /* absolute inflation of node boxes */
    uint64_t device_bytes;     /* RRCollider::getMemoryOccupied (RRCollider.h:265) analogue */
    double   build_ms;
    uint32_t num_nodes4;       /* 4-wide nodes the GPU traverses (collapsed from the BVH2) */
    uint32_t stack_need;       /* worst-case traversal stack entries of the 4-wide tree */
} lsg_collider_info;

int  lsg_collider_create(const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntri,
                         const lsg_build_params* params, const volatile int* aborting, lsg_collider** out);
void lsg_collider_destroy(lsg_collider* collider);
int  lsg_collider_info_get(const lsg_collider* collider, lsg_collider_info* info);

/* ---- collider cache ------------------------------------------------------------------------
 * The reference caches built trees on disk under a name made from the mesh's hash
 * (IntersectBsp.h:213-291, TREE_VERSION 4, files of MIN_BYTES_FOR_SAVE = 1024 bytes or
 * more).  This library's tree files hold the BVH2, the 4-wide tree and the triangles' leaf
 * order, under the same naming scheme with its own format version and extension. */
#define LSG_TREE_VERSION 1u
/* SHA-1 of the mesh as RRMesh::getHash computes it (RRHash.cpp:48-77): every vertex as
 * three binary32 coordinates, then every triangle as three uint32 indices; hash: 20 bytes */
int  lsg_mesh_hash(const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntri, uint8_t* hash);
/* RRHash::getFileName (RRHash.cpp:198-226): location (NULL = the system temp directory's
 * "Lightsprint/", created) + 32 letters of the hash with version added to its first
 * little-endian word + extension; LSG_ERR_LIMIT when it does not fit capacity bytes */
int  lsg_cache_file_name(const uint8_t* hash, uint32_t version, const char* location, const char* extension, char* out,
                         uint32_t capacity);
/* writes a built collider to path (synchronous; the file appears complete or not at all) */
int  lsg_collider_save(const lsg_collider* collider, const char* path);
/* a collider from a tree file of this mesh: LSG_ERR_INVALID when the file is missing, of
 * another mesh or format version, or damaged */
int  lsg_collider_load(const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntri, const char* path,
                       lsg_collider** out);
/* RRCollider::create with a cacheLocation (RRCollider.h:179-183): loads the tree file
 * cache_location + name + ".lsgbvh" when it holds this mesh built with the same leaf size
 * (any, when params ask for the default) and flags, else builds and writes it (trees of
 * at least 1024 bytes; a failed write is ignored).  cache_location NULL = the temp
 * directory, "*" = no caching.  *from_cache (may be NULL) = 1 when loaded. */
int  lsg_collider_create_cached(const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntri,
                                const lsg_build_params* params, const volatile int* aborting, const char* cache_location,
                                lsg_collider** out, int32_t* from_cache);
/* copies the tree to the host: nodes = 16 floats per node slot, prim = original
 * triangle index per leaf-order slot (num_valid). Either may be NULL. */
int  lsg_collider_download(const lsg_collider* collider, float* nodes, uint32_t* prim);

/* Host-only (no GPU call): rebuilds the inner nodes of a BVH2 in the download layout
 * over its leaves with a top-down binned SAH (32 bins per axis on leaf-box centroids,
 * as many as leaves below 32),
 * keeping every leaf's triangle range and inflated box.  out_nodes holds num_nodes * 16
 * floats; the result is numbered in preorder from root 0 and depends only on the input.
 * out_depth = inner nodes on the longest root path.  A tree with fewer than two leaves is
 * copied unchanged (out_depth 0).  The step of lsg_collider_create after the LBVH build;
 * exported so that it can be checked on the host.  The reference builds its BSP/kd tree
 * with its own builder (createAndSaveBsp, RRCollider/bsp.cpp:1192; kd split planes weighed
 * by box surface areas, find_best_root_kd, bsp.cpp:388-574); this has no direct
 * counterpart. */
int  lsg_bvh_rebuild_sah(const float* nodes, uint32_t num_nodes, int32_t root, float* out_nodes,
                         uint32_t* out_num_nodes, int32_t* out_root, uint32_t* out_depth);

/* copies the per-triangle records in leaf order (num_valid each): plane = float4 n4
 * (TriangleSRLNP::n4), sv = float4 (s[x], s[y], s[z], intersectReal), ev = float4
 * (r[x], r[y], l[z], ir[z]) with x,y,z the projection axes of intersectByte, meta =
 * 2 uint32 (original triangle index, intersectByte).  Used by the host-side
 * single-ray path so that CPU and GPU test triangles with identical records. */
int  lsg_collider_download_triangles(const lsg_collider* collider, float* plane, float* sv, float* ev, uint32_t* meta);

/* intersectBatch over the reference's own RRRay records in host memory.
 * Re-adds RRCollider::intersectBatch (doc/doxygen/doc.news:3212, removed :154)
 * with the per-ray contract of RRCollider::intersect (RRCollider.h:197-228):
 * closest hit in <rayLengthMin,rayLengthMax>, outputs per rayFlags.  Rays
 * carrying a collisionHandler are not supported here (the C++ collider routes
 * them to its CPU path).  hit_mask (optional, n bytes) receives 1/0 per ray. */
int  lsg_collider_intersect_rrrays(const lsg_collider* collider, void* rays, uint32_t n,
                                   uint8_t* hit_mask, uint32_t* num_hits);

/* Device-resident batch (the benchmark path).  ray_orig_tmin / ray_dir_tmax:
 * device float4[n] (origin.xyz, rayLengthMin) and (dir.xyz, rayLengthMax).
 * hits: device int4[n] = (triangle or -1, bits of t, bits of u, bits of v).
 * front (optional device uint8[n]) receives hitFrontSide.  Asynchronous on
 * stream (NULL = default stream). */
int  lsg_collider_intersect_device(const lsg_collider* collider, const void* ray_orig_tmin,
                                   const void* ray_dir_tmax, uint32_t n, void* hits, void* front,
                                   void* stream);

/* Traversal work counters for n device rays (node visits, triangle tests);
 * used for the roofline bytes-per-ray figure.  Synchronous. */
int  lsg_collider_trace_stats(const lsg_collider* collider, const void* ray_orig_tmin,
                              const void* ray_dir_tmax, uint32_t n, double* node_visits, double* tri_tests);

/* ---- lightmap gather -------------------------------------------------------------------
 * A scene is a collider plus what RRSolver's final gather reads at a hit: per-vertex
 * normals (RRMesh::getTriangleNormals, RRMesh.cpp:77-90), per-triangle side/material
 * bits (RRMaterial sideBits, RRMaterial.cpp:154-166), per-triangle exitance of each
 * side (PathtracerWorker::getIncidentRadiance with solver shortcuts at depth 0,
 * pathtracer.cpp:235-350) and the lights (RRLight::getIrradiance, RRLight.cpp:345-386). */
#define LSG_SIDE_FRONT_RENDERS  1u   /* sideBits[0].renderFrom */
#define LSG_SIDE_BACK_RENDERS   2u   /* sideBits[1].renderFrom */
#define LSG_SIDE_CASTS_SHADOW   4u   /* getTriangleMaterial(t, light, ...) != nullptr */
#define LSG_SIDE_FRONT_INDIRECT 8u   /* set by lsg_scene_set_indirect: front hits add the corner indirect term */
#define LSG_SIDE_BACK_INDIRECT 16u   /* ... back hits */

typedef struct lsg_light {
    int32_t type;              /* 0 = point (REALISTIC attenuation), 1 = directional */
    float   position[3];
    float   direction[3];
    float   color[3];          /* linear, already scaled by the light's multiplier */
    int32_t cast_shadows;
} lsg_light;

/* SubTexel, RRSolver/gather.h:60-65: triangle, uv of its three corners in
 * triangle space, area in map space.  32 bytes. */
typedef struct lsg_subtexel {
    uint32_t triangle;
    float    uv[6];
    float    area;
} lsg_subtexel;

/* One texel to gather: 16 bytes.  sub_count_flags = count (low 16 bits) |
 * texelFlags << 16 (gather.h:68-80).  key = stable id of the texel for the
 * seeded ray set (object/texel hash). */
typedef struct lsg_texel {
    uint32_t out_index;        /* float4 slot in the output buffer */
    uint32_t sub_begin;
    uint32_t sub_count_flags;
    uint32_t key;
} lsg_texel;

/* RRSolver::UpdateParameters subset (RRSolver.h:468-592) */
typedef struct lsg_gather_params {
    uint32_t quality;          /* hemisphere rays; light rounds = quality/10+1 (gather.cpp:197,396) */
    int32_t  hemisphere;       /* 1 when emissive/environment/useCurrentSolution (gather.cpp:197) */
    float    ray_length_min;   /* minimalSafeDistance (RRSolver.cpp:333) */
    float    locality;         /* default 100000 */
    float    rug_distance;     /* default 0.001 */
    float    light_multiplier; /* direct.lightMultiplier */
    uint32_t seed;             /* seeded replacement of rand() (gather.h:102, gather.cpp:742) */
    uint32_t light_rounds;     /* 0: quality/10+1 as the reference; else shadow rounds per light
                                  (the learned path's cheap gather) */
    float    environment_multiplier; /* indirect.environmentMultiplier: hemisphere rays that leave
                                  the scene return environment * this (pathtracer.cpp:161-172);
                                  0 = the environment is not sampled (gather.cpp:149) */
} lsg_gather_params;

int  lsg_scene_create(const lsg_collider* collider, const float* vertex_normals, const uint8_t* side_bits,
                      const float* exitance, const lsg_light* lights, uint32_t nlights, lsg_scene** out);
void lsg_scene_destroy(lsg_scene* scene);
/* exitance: host float[ntri*8] = per triangle (front r,g,b,pad, back r,g,b,pad) */
int  lsg_scene_set_exitance(lsg_scene* scene, const float* exitance);
/* exitance on the device: front = irradiance * albedo + emittance, back = 0 per triangle
 * (PathtracerWorker::getIncidentRadiance at depth 0 with solver shortcuts,
 * pathtracer.cpp:235-316), the table a first gather leaves.  irradiance: device
 * float4[ntri], or NULL for emittance only; albedo, emittance: device float[ntri*3].
 * Asynchronous on stream. */
int  lsg_scene_exitance_device(lsg_scene* scene, const float* irradiance, const float* albedo, const float* emittance,
                               void* stream);
/* copies the current exitance table (same layout) to the host; synchronous */
int  lsg_scene_get_exitance(const lsg_scene* scene, float* exitance);
/* Per-vertex indirect light at hemisphere hits.  The reference adds, on top of the flat
 * direct exitance, the albedo times the indirect irradiance of the hit triangle's three
 * corners interpolated at the hit's hitPoint2d (PathtracerWorker::getIncidentRadiance,
 * pathtracer.cpp:326-348; Triangle::getPointMeasure, rrcore.cpp:262-273):
 *   exitance(hit) = table[side] + albedo * (I0 (1-u-v) + I1 u + I2 v).
 * corners: host float[ntri*9], RGB of corners 0, 1, 2 per triangle (what
 * RRSolver::getTriangleMeasure(t, 0..2, RM_IRRADIANCE_LINEAR_INDIRECT) returns); albedo:
 * host float[ntri*3]; sides: host uint8[ntri], bit 0 = front hits take the term, bit 1 =
 * back hits (the side is legal and the albedo not black, pathtracer.cpp:191, 325-331);
 * NULL = both.  The exitance table then holds only the flat part (direct irradiance times
 * albedo plus emittance).  corners == NULL removes the term. */
int  lsg_scene_set_indirect(lsg_scene* scene, const float* corners, const float* albedo, const uint8_t* sides);
int  lsg_scene_set_lights(lsg_scene* scene, const lsg_light* lights, uint32_t nlights);

/* Point materials: materials read at a ray's hit point (RRObject::getPointMaterial,
 * RRObject.cpp:198-303, in RRCollisionHandlerFinalGathering::collides, pathtracer.h:194-226)
 * when the gather's quality exceeds their minimal quality: hemisphere rays at quality
 * (nearest texel), shadow rays at 2 x quality, interpolating texels bilinearly when
 * quality / 10 exceeds it (gather.cpp:175-180, 366-370).  Supported: a diffuse texture, an
 * emissive texture and keyed transmittance; a transparent texel is passed through by every
 * ray, an opaque one is hit with its diffuse colour.  Refused: unkeyed (partial)
 * transmittance, a keyed material that emits, and a gather whose quality does not read
 * every keyed material's point materials (below it the reference transmits partially). */
#define LSG_PM_TRANSMITTANCE_IN_ALPHA 1u   /* specularTransmittanceInAlpha */
#define LSG_PM_KEYED                  2u   /* specularTransmittanceKeyed */
#define LSG_PM_INVERTED               4u   /* specularTransmittanceMapInverted */
typedef struct lsg_point_material {
    int32_t  diffuse_texture;        /* diffuseReflectance.texture: index into the textures, or -1 */
    int32_t  emissive_texture;       /* diffuseEmittance.texture, or -1 */
    int32_t  transmittance_texture;  /* specularTransmittance.texture (may be the diffuse one), or -1 */
    uint32_t flags;                  /* LSG_PM_* */
    float    threshold;              /* specularTransmittanceThreshold */
    uint32_t min_quality;            /* minimalQualityForPointMaterials */
    float    albedo[3];              /* diffuseReflectance.colorLinear of the per-triangle material */
    float    emittance[3];           /* diffuseEmittance.colorLinear x materialEmittanceMultiplier */
    float    emittance_multiplier;   /* materialEmittanceMultiplier, applied to emissive texels */
    uint32_t emit_sides;             /* bit 0: sideBits[0].emitTo && legal, bit 1: sideBits[1] */
} lsg_point_material;
/* A texture as RRBufferInMemory stores it: width x height elements of an LSG_FORMAT_*
 * format (RGB, BGR, RGBA, RGBF, RGBAF, LUMINANCE, LUMINANCEF), rows from the first.  With a
 * colour space other than the sRGB power of lsg_scene_set_point_materials' gamma, the caller
 * converts: linear = getElement(i, colorSpace) of every texel (RRBufferInMemory.cpp:295-360),
 * linear_stored = getElement(i, nullptr) with colorSpace->toLinear applied to rgb (what the
 * transmittance-in-alpha path reads, RRObject.cpp:240-244); NULL = computed from gamma. */
typedef struct lsg_texture {
    uint32_t width, height;
    int32_t  format;
    int32_t  scaled;                 /* RRBuffer::getScaled: stored in the solver's colour space */
    const void* data;                /* host */
    const float* linear;             /* host float[width*height*4] or NULL */
    const float* linear_stored;      /* host float[width*height*4] or NULL */
} lsg_texture;
/* tri_material: host uint32[ntri], each triangle's material or 0xFFFFFFFF (none); tri_uv:
 * host float[ntri*6], the material uv of each triangle's corners; gamma: the power of the
 * solver's colour space (RRColorSpace::create_sRGB, 0.45), 0 = none.  The exitance table
 * entry of a triangle with a material holds the side's direct irradiance (the exitance
 * kernels of lsg_scene_exitance_device and the bake plans write it so, and a caller of
 * lsg_scene_set_exitance does the same): the hit multiplies it by the diffuse colour and
 * adds emittance and indirect light itself.  tri_material NULL removes the materials. */
int  lsg_scene_set_point_materials(lsg_scene* scene, const uint32_t* tri_material, const float* tri_uv,
                                   const lsg_point_material* materials, uint32_t nmaterials,
                                   const lsg_texture* textures, uint32_t ntextures, float gamma);

/* The solver's environment (RRSolver::setEnvironment), which hemisphere rays that leave
 * the scene sample (PathtracerWorker::getIncidentRadiance, pathtracer.cpp:161-172), with a
 * nearest-texel lookup as RRBufferInMemory::getElementAtDirection
 * (RRBufferInMemory.cpp:410-455).  rgb: host float[texels*3], linear (what
 * RRBuffer::getElement returns with the solver's colour space).  A sampled environment
 * also turns the hemisphere gather on (gather.cpp:197).  LSG_ENV_NONE clears it. */
#define LSG_ENV_NONE      0
#define LSG_ENV_CUBE      1   /* BT_CUBE_TEXTURE: 6 faces of width x width, +x,-x,+y,-y,+z,-z */
#define LSG_ENV_PANORAMA  2   /* BT_2D_TEXTURE: 360 x 180 degree panorama, width x height */
int  lsg_scene_set_environment(lsg_scene* scene, int32_t type, uint32_t width, uint32_t height, const float* rgb);
/* The solver's second environment, both environments' rotations and the blend factor
 * (RRSolver::setEnvironment(env0, env1, angleRad0, angleRad1), setEnvironmentBlendFactor).
 * The gather samples RRBuffer::createEnvironmentBlend of them (pathtracer.cpp:29-34,
 * RRBufferBlend.cpp:21-66): a ray that leaves the scene in direction d returns
 *   env0(R(angle0) d) * (1 - blend) + env1(R(angle1) d) * blend,
 * R(a) = RRMatrix3x4::rotationByYawPitchRoll((a, 0, 0)), an absent environment giving 0;
 * with no second environment and angle0 == 0 it returns env0(d) whatever the blend
 * factor.  The first environment (lsg_scene_set_environment) alone decides whether the
 * hemisphere is gathered (gather.cpp:149, 197).  type1 LSG_ENV_NONE: no second
 * environment (rgb1 may be NULL). */
int  lsg_scene_set_environment_blend(lsg_scene* scene, int32_t type1, uint32_t width1, uint32_t height1, const float* rgb1,
                                     float angle0, float angle1, float blend_factor);

/* The per-texel final gather, processTexel (RRSolver/gather.cpp:600-836) for every
 * texel at once.  texels/subs are device arrays; out_rgba is a device float4 array
 * indexed by lsg_texel.out_index (RGB irradiance in physical units, alpha =
 * texelFlags/255 when reliable, gather.cpp:795-807).  Asynchronous on stream. */
int  lsg_gather_device(lsg_scene* scene, const lsg_texel* texels, uint32_t ntexels,
                       const lsg_subtexel* subs, uint32_t nsubs, const lsg_gather_params* params,
                       float* out_rgba, void* stream);

/* Output layers of updateLightmap (RRSolver/gather.h:24-31, lightmap.cpp:712-717). */
#define LSG_LAYER_LIGHTMAP      0   /* LS_LIGHTMAP */
#define LSG_LAYER_DIRECTION1    1   /* LS_DIRECTION1..3: directional lightmaps */
#define LSG_LAYER_DIRECTION2    2
#define LSG_LAYER_DIRECTION3    3
#define LSG_LAYER_BENT_NORMALS  4   /* LS_BENT_NORMALS */
#define LSG_NUM_LAYERS          5

/* Per-vertex tangent frames of the scene's mesh, which directional lightmaps are
 * weighted in (RRMesh::TangentBasis, RRMesh.h:207-215; RRMeshArrays::getTriangleNormals,
 * RRMeshArrays.cpp:279-306).  tangents, bitangents: host float[nverts*3], in the
 * collider's vertex order.  Both NULL = build each vertex's frame from its normal
 * (TangentBasis::buildBasisFromNormal, RRMesh.cpp:70-75), which is also what a scene
 * that never had frames set uses. */
int  lsg_scene_set_tangents(lsg_scene* scene, const float* tangents, const float* bitangents);

/* processTexel with every output layer of RRSolver::updateLightmap
 * (gather.cpp:600-836, lightmap.cpp:806-820).  layers: LSG_NUM_LAYERS device float4
 * arrays indexed by lsg_texel.out_index, NULL for a layer not wanted.
 * - When any of LSG_LAYER_DIRECTION1..3 is given, every layer 0..3 is gathered with
 *   the directional weighting (gatherAllDirections, gather.cpp:284-305, 453-481):
 *   hemisphere ray k adds exitance * max(0, dir.d_i) / dir.n, a visible light adds
 *   irradiance * max(0, L.d_i), with d_i = g_lightmapDirections[i] (gather.cpp:43-49)
 *   in the texel's skewed frame.
 * - LSG_LAYER_BENT_NORMALS receives the normalised bent normal as (n+1)/2 with the
 *   texel's alpha (gather.cpp:306, 482, 809-830).
 * Alphas as lsg_gather_device.  Asynchronous on stream. */
int  lsg_gather_layers_device(lsg_scene* scene, const lsg_texel* texels, uint32_t ntexels,
                              const lsg_subtexel* subs, uint32_t nsubs, const lsg_gather_params* params,
                              float* const* layers, void* stream);

/* Same gather with traversal work counters (synchronous): stats6 = hemisphere rays,
 * their internal-node visits, their triangle tests, shadow rays, node visits,
 * triangle tests -- the inputs of the bytes-per-texel roofline figure. */
int  lsg_gather_stats(lsg_scene* scene, const lsg_texel* texels, uint32_t ntexels,
                      const lsg_subtexel* subs, uint32_t nsubs, const lsg_gather_params* params,
                      float* out_rgba, double* stats6);

/* ---- lightmap filters and writeback -------------------------------------------------- */
/* One lightmap inside a batch: its RGBA temp texels start at float4 slot rgba_base
 * (row-major, width*height), its RGBF output at texel slot rgbf_base. */
typedef struct lsg_lightmap {
    uint32_t rgba_base;
    uint32_t width;
    uint32_t height;
    uint32_t rgbf_base;
} lsg_lightmap;

/* RRSolver::FilteringParameters (RRSolver.h:596-622); defaults 1, 1000, 0, false */
typedef struct lsg_filter_params {
    float   smoothing_amount;   /* lightmapSmooth sigma */
    float   spread_foreground;  /* lightmapGrow steps */
    float   background[4];      /* lightmapFillBackground colour */
    int32_t wrap;               /* neighbours across opposite borders (RRBuffer.cpp:676-680, 822-826, 904-922) */
} lsg_filter_params;

/* One unwrap seam (UnwrapSeam, RRBuffer.cpp:529-533): a mesh edge shared by two
 * triangles whose lightmap uvs differ: edge1 = (u, v) of its two ends in the first
 * triangle's unwrap, edge2 = the same two ends in the second's. */
typedef struct lsg_unwrap_seam {
    float edge1[4];
    float edge2[4];
} lsg_unwrap_seam;

/* The seams lightmapSmooth smooths for an object (UnwrapSeams, RRBuffer.cpp:536-625):
 * vertices stitched as RRMesh::createOptimizedVertices(0, 1 degree, 0, NULL) stitches
 * them (lsg_stitch_vertices), then every triangle edge that a later triangle shares at
 * both stitched ends with different uvs there, in the reference's order.  Host function.
 * verts, idx, corner_normals as lsg_stitch_vertices; tri_uv: float[ntri*6], the lightmap
 * uv of every triangle corner; seams: out, the first cap seams (NULL with cap 0 to count);
 * *nseams out: how many there are. */
int  lsg_unwrap_seams(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri, const float* corner_normals,
                      const float* tri_uv, lsg_unwrap_seam* seams, uint32_t cap, uint32_t* nseams);

/* The filter chain updateLightmap runs on every gathered buffer (lightmap.cpp:852-876):
 * lightmapSmooth -> lightmapGrowForBilinearInterpolation -> lightmapGrow ->
 * lightmapFillBackground (RRBuffer.cpp:630-975), for a batch of lightmaps at once.
 * rgba: device float4 texels, updated in place; rgbf (optional): device float3
 * output (copyElementsTo into a BF_RGBF buffer, RRBuffer.cpp:105).  Synchronous. */
int  lsg_filter_lightmaps(const lsg_lightmap* maps, uint32_t nmaps, float* rgba, float* rgbf,
                          const lsg_filter_params* params, void* stream);

/* The same filter chain with the batch and its scratch kept in HBM between runs, so
 * repeated filtering (a sweep of re-bakes) allocates nothing.  run: as
 * lsg_filter_lightmaps on the plan's lightmaps. */
typedef struct lsg_filter_plan lsg_filter_plan;
int  lsg_filter_plan_create(const lsg_lightmap* maps, uint32_t nmaps, lsg_filter_plan** out);
void lsg_filter_plan_destroy(lsg_filter_plan* plan);
int  lsg_filter_plan_run(lsg_filter_plan* plan, float* rgba, float* rgbf, const lsg_filter_params* params, void* stream);
/* The objects' unwrap seams, smoothed after every blur pass of lightmapSmooth
 * (RRBuffer.cpp:734-777, the seams of lightmapSmooth's object argument): lightmap m's
 * seams are seams[begin[m] .. begin[m+1]) (host arrays of the plan's nmaps + 1 offsets,
 * copied); seams == NULL removes them. */
int  lsg_filter_plan_set_seams(lsg_filter_plan* plan, const lsg_unwrap_seam* seams, const uint32_t* begin);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc): lsg_filter_lightmaps with the
 * same kernels and results, its scratch held in a static device array of the library
 * (about 61 MiB on each device the library runs on, held until the process ends), so
 * the call allocates nothing.  At most LSG_FILTER_NOALLOC_MAX_TEXELS texels over all maps
 * and LSG_FILTER_NOALLOC_MAX_MAPS maps; larger batches fail with LSG_ERR_LIMIT.  Every
 * call on one device shares that device's scratch and clears what must start at zero:
 * calls that overlap on one device (on different streams or threads) race.  Synchronous. */
#define LSG_FILTER_NOALLOC_MAX_TEXELS (1u << 20)
#define LSG_FILTER_NOALLOC_MAX_MAPS   4096u
int  lsg_filter_lightmaps_noalloc(const lsg_lightmap* maps, uint32_t nmaps, float* rgba, float* rgbf,
                                  const lsg_filter_params* params, void* stream);

/* Texel rasterisation of one object's lightmap, RRSolver/lightmap.cpp:83-335
 * (enumerateTexelsPartial, steps 1-2): clips every texel square against every
 * triangle of the object in its unwrap and emits SubTexels and texelFlags.
 * tri_uv: float[ntri*6] lightmap uv of each triangle; tri_ids: the triangle's
 * index in the collider's mesh.  Call with texels == NULL to get the counts. */
int  lsg_rasterize_texels(const float* tri_uv, const uint32_t* tri_ids, uint32_t ntri,
                          uint32_t width, uint32_t height, uint32_t object_id,
                          lsg_texel* texels, uint32_t* ntexels, lsg_subtexel* subs, uint32_t* nsubs,
                          uint32_t out_base);

/* ---- whole-scene bake plan ------------------------------------------------------------
 * RRSolver::updateLightmaps (RRSolver/lightmap.cpp:943-1163) for many objects at once:
 * the plan rasterises every object's lightmap once, keeps texels, SubTexels, the RGBA
 * temp buffers, the RGBF output and filter scratch in HBM, and runs
 *   first gather per triangle (gatherPerTrianglePhysical, gather.cpp:843-952)
 *   -> exitance = direct irradiance * albedo + emittance
 *   -> final gather of all texels of all objects in one persistent launch
 *   -> filter chain + RGBF writeback.
 * With LSG_BAKE_USE_GRAPH the sequence is captured once into a hipGraph and replayed. */
typedef struct lsg_bake lsg_bake;

typedef struct lsg_bake_object {
    const float*    tri_uv;    /* host float[ntri*6]: lightmap uv of each triangle */
    const uint32_t* tri_ids;   /* host: the triangles' indices in the collider's mesh */
    uint32_t        ntri;
    uint32_t        width, height;
} lsg_bake_object;

/* launch shape of the persistent gather; 0 = default (picked by lsg_bake_autotune) */
typedef struct lsg_bake_tuning {
    uint32_t tile;             /* texels per work-queue fetch (default 2) */
    uint32_t waves_per_block;  /* 2, 4 or 8 (default 4) */
    uint32_t blocks_per_cu;    /* persistent workgroups per compute unit (default 4) */
    uint32_t xcd_queues;       /* 8 = one work queue per XCD, 1 = one global queue (default 8) */
    uint32_t lane_refill;      /* above 32 shooters per series: 1 = a texel's lanes take 64 shooters at a time,
                                  2 = each lane takes the next shooter as soon as 16 lanes are free (default 2) */
} lsg_bake_tuning;

typedef struct lsg_bake_info {
    uint32_t num_texels, num_subtexels, num_slots, num_lightmaps, num_triangles, compute_units, grow_steps;
    lsg_bake_tuning tuning;
    void* rgba;                /* device float4[num_slots]: filtered RGBA temp texels */
    void* rgbf;                /* device float3[num_slots]: RGBF lightmaps, object after object */
    void* texels;              /* device lsg_texel[num_texels] */
    void* subtexels;           /* device lsg_subtexel[num_subtexels] */
} lsg_bake_info;

#define LSG_BAKE_USE_GRAPH 1u

/* albedo, emittance: host float[ntri*3] per collider triangle (diffuse reflectance and
 * diffuse emittance, RRMaterial.h).  The scene must outlive the plan. */
int  lsg_bake_create(lsg_scene* scene, const lsg_bake_object* objects, uint32_t nobjects,
                     const float* albedo, const float* emittance, lsg_bake** out);
void lsg_bake_destroy(lsg_bake* bake);
int  lsg_bake_info_get(const lsg_bake* bake, lsg_bake_info* info);
int  lsg_bake_set_tuning(lsg_bake* bake, const lsg_bake_tuning* tuning);
/* first may be NULL (keep the current exitance); filtering may be NULL (no filters).
 * Without the graph flag the call is asynchronous on stream except for the grow
 * loop's convergence checks; with it, the replay is ordered after stream's work. */
int  lsg_bake_run(lsg_bake* bake, const lsg_gather_params* first, const lsg_gather_params* final_params,
                  const lsg_filter_params* filtering, uint32_t flags, void* stream);
/* Radiosity on the GPU: after the first gather (direct light), `bounces` Jacobi
 * iterations of the per-triangle gather spread the light between triangles, which the
 * reference leaves to its CPU solver (RRSolver::updateSolverIndirectIllumination,
 * gather.cpp:1038-1120).  Each iteration is a hemisphere-only gather of `quality` rays per
 * triangle (0 = the first gather's quality) over the current exitance table, and
 *   exitance = emittance + albedo * (light part of the first gather + hemisphere part).
 * Takes effect on the next lsg_bake_run with first != NULL (graph replays included);
 * 0 bounces (the default) = direct light only. */
int  lsg_bake_set_bounces(lsg_bake* bake, uint32_t bounces, uint32_t quality);
/* The unwrap seams of one object's lightmap (host array, copied; nseams 0 removes them),
 * smoothed by the filter chain of later runs as lsg_filter_plan_set_seams describes. */
int  lsg_bake_set_seams(lsg_bake* bake, uint32_t object, const lsg_unwrap_seam* seams, uint32_t nseams);
/* times the persistent final gather over every stride-th texel (autotuner probe) */
int  lsg_bake_gather_sample(lsg_bake* bake, const lsg_gather_params* params, uint32_t stride, float* ms);
/* copies object `object`'s RGBF lightmap (width*height*3 floats) to the host */
int  lsg_bake_download(const lsg_bake* bake, uint32_t object, float* rgbf);
/* RRSolver::updateLightmaps with directional and bent-normal layers
 * (lightmap.cpp:943-1163, 806-876): first gather (if first != NULL), the layered final
 * gather (lsg_gather_layers_device) of every texel, then the filter chain and RGBF
 * writeback on each layer of layer_mask (bit i = LSG_LAYER_i) separately, as the
 * reference filters each pixel buffer.  Bent normals are written without a colour
 * space.  Eager (no graph); asynchronous on stream except for the grow loop's
 * convergence checks.  Layer 0 is the buffer lsg_bake_download reads. */
int  lsg_bake_run_layers(lsg_bake* bake, const lsg_gather_params* first, const lsg_gather_params* final_params,
                         const lsg_filter_params* filtering, uint32_t layer_mask, void* stream);
/* copies layer `layer` of object `object` (width*height*3 floats) to the host */
int  lsg_bake_download_layer(const lsg_bake* bake, uint32_t object, uint32_t layer, float* rgbf);
/* the same layer as the filters left it, linear RGBA (width*height*4 floats): the
 * reference's per-object RGBAF work buffer after lightmapGrow and lightmapFillBackground
 * (lightmap.cpp:809, 866-873), which copyElementsTo then writes into the caller's
 * buffer with alpha (RRBuffer.cpp:105-131); background texels carry the filter
 * parameters' background colour, alpha included */
int  lsg_bake_download_layer_rgba(const lsg_bake* bake, uint32_t object, uint32_t layer, float* rgba);

/* ---- consumer formats -----------------------------------------------------------------------
 * A filtered RGBA lightmap (lsg_bake_download_layer_rgba's texels, on the device) written in
 * a consumer buffer's format, as updateLightmap's writeback does through the buffer interface
 * (lightmap.cpp:866-876: lightmapFillBackground, copyElementsTo, RRBufferInMemory::setElement,
 * RRBufferInMemory.cpp:229-285).  Format numbers are RRBufferFormat's (RRBuffer.h:89-102). */
#define LSG_FORMAT_RGB         0   /* BF_RGB: 3 bytes, RR_FLOAT2BYTE per channel */
#define LSG_FORMAT_BGR         1   /* BF_BGR */
#define LSG_FORMAT_RGBA        2   /* BF_RGBA: 4 bytes */
#define LSG_FORMAT_RGBF        3   /* BF_RGBF: 3 floats */
#define LSG_FORMAT_RGBAF       4   /* BF_RGBAF: 4 floats */
#define LSG_FORMAT_LUMINANCE   9   /* BF_LUMINANCE: 1 byte of the rgb average */
#define LSG_FORMAT_LUMINANCEF 10   /* BF_LUMINANCEF: 1 float */
/* bytes per texel of a format, 0 for formats without a per-texel encoding */
uint32_t lsg_format_bytes(int32_t format);
/* rgba: device float4[ntexels], linear, background texels with alpha <= 0; out: device,
 * ntexels * lsg_format_bytes(format) bytes, 4-byte aligned for the byte formats.
 * scaled: the target buffer is scaled (RRBuffer::getScaled); solver_gamma: the power of the
 * solver's sRGB colour space (RRColorSpace::create_sRGB, 0.45 by default; 0 = no colour
 * space), which makes the background linear for a scaled target; copy_gamma: the colour
 * space the copy converts to (the solver's, or 0 for bent normals).  background: 4 floats.
 * Asynchronous on stream. */
int  lsg_lightmap_encode(const float* rgba, uint32_t ntexels, int32_t format, int32_t scaled, float solver_gamma, float copy_gamma,
                         const float* background, void* out, void* stream);

/* ---- per-vertex outputs ------------------------------------------------------------------
 * updateLightmap / updateLightmaps into vertex buffers (lightmap.cpp:766-800, 1095-1125): a
 * final gather of one texel per triangle whose subtexel is the whole triangle
 * (gatherPerTrianglePhysical, gather.cpp:843-952), smoothed into the solver's stitched
 * vertices (updateVertexBufferFromPerTriangleDataPhysical, vertexBuffer.cpp:262-290). */
typedef struct lsg_vertex_plan lsg_vertex_plan;
/* Groups mesh vertices as RRMesh::createOptimizedVertices(weld_distance, max_smooth_angle,
 * 0, NULL) does (RRMeshLessVertices.h:38-230; the solver calls it with
 * SmoothingParameters::vertexWeldDistance and |maxSmoothAngle|, interpol.cpp:140).  Host
 * function.  verts: float[nverts*3]; idx: uint32[ntri*3]; corner_normals: float[ntri*9]
 * (RRMesh::getTriangleNormals); vertex_group: uint32[nverts] out, the group of each vertex,
 * 0xFFFFFFFF for vertices no triangle uses; group_rep: uint32[nverts] out or NULL, the
 * vertex whose position stands for each group; *ngroups out. */
int  lsg_stitch_vertices(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri, const float* corner_normals,
                         float weld_distance, float max_smooth_angle, uint32_t* vertex_group, uint32_t* group_rep, uint32_t* ngroups);
/* The smoothing of per-triangle values into groups for the scene's triangles
 * (Object::buildTopIVertices, interpol.cpp:125-172, and IVertex::insert, interpol.cpp:51-73).
 * corner_group: host uint32[ntri*3], the group of each triangle corner (0xFFFFFFFF = none);
 * group_positions: host float[ngroups*3]. */
int  lsg_vertex_plan_create(const lsg_scene* scene, const uint32_t* corner_group, const float* group_positions, uint32_t ngroups,
                            lsg_vertex_plan** out);
void lsg_vertex_plan_destroy(lsg_vertex_plan* plan);
/* host copies of the group each corner counts for (0xFFFFFFFF: the solver dropped the
 * triangle or stitching joined two of its corners) and of its power (corner angle x
 * triangle area); either may be NULL */
int  lsg_vertex_plan_corners(const lsg_vertex_plan* plan, uint32_t* corner_group, float* power);
/* group value = sum(power x value of the corner's triangle) / sum(power), 0 when the sum is 0
 * (IVertex::getVertexDataFromTriangleData, RRStaticSolver.cpp:130-145).  tri_data: device
 * float4[ntri] (rgb used); group_out: device float4[ngroups].  Asynchronous on stream. */
int  lsg_vertex_plan_smooth(const lsg_vertex_plan* plan, const float* tri_data, float* group_out, void* stream);
/* the per-triangle final gather of the layers in layer_mask (bit i = LSG_LAYER_i) with
 * params, each smoothed into group_out[i] (device float4[ngroups]).  Asynchronous on stream. */
int  lsg_vertex_plan_gather(lsg_vertex_plan* plan, const lsg_gather_params* params, uint32_t layer_mask, float* const* group_out,
                            void* stream);
/* host copy of the last per-triangle gather of a layer, float4[ntri] (synchronous) */
int  lsg_vertex_plan_triangle_data(const lsg_vertex_plan* plan, uint32_t layer, float* rgba);

/* ---- learned gather ----------------------------------------------------------------------
 * No counterpart in the reference: a surrogate that replaces most of the 1024-ray
 * final gather.  A cheap gather (few rays) feeds an MLP 32-64-64-64-64-4 (BF16 on
 * MFMA, FP32 accumulation) that predicts log1p(E/s) per channel plus its own RMS log
 * error; texels whose predicted error exceeds the bound are re-gathered by the plain
 * gather (processTexel, gather.cpp:600-836), which stays the measure of the result. */
typedef struct lsg_mlp lsg_mlp;

uint32_t lsg_mlp_num_params(void);
/* max_batch: rows per training step / inference chunk, multiple of 16 */
int  lsg_mlp_create(uint32_t max_batch, lsg_mlp** out);
void lsg_mlp_destroy(lsg_mlp* mlp);
/* params: host float[lsg_mlp_num_params()], per layer W[in][out] row-major then b[out]
 * (layers 32x64, 64x64, 64x64, 64x64, 64x16 with outputs 4..15 held at zero);
 * setting parameters resets the Adam state */
int  lsg_mlp_set_params(lsg_mlp* mlp, const float* params);
int  lsg_mlp_get_params(const lsg_mlp* mlp, float* params);
int  lsg_mlp_get_grads(const lsg_mlp* mlp, float* grads);
/* x: device float[n*32]; y: device float[n*4] */
int  lsg_mlp_forward(lsg_mlp* mlp, const float* x, uint32_t n, float* y, void* stream);
/* one Adam step on n <= max_batch rows with targets t (device float[n*3]); lr <= 0
 * only computes gradients (lsg_mlp_get_grads); *loss (host, optional) = mean loss */
int  lsg_mlp_train_step(lsg_mlp* mlp, const float* x, const float* t, uint32_t n, float lambda, float lr,
                        float* loss, void* stream);
/* C[MxN] = A[MxK] B[KxN] on the device through the MLP's MFMA GEMM (bf16 inputs,
 * row-major; M, N, K multiples of 16) -- exposed to verify the kernel */
int  lsg_mlp_gemm_test(const uint16_t* a, const uint16_t* b, float* c, uint32_t m, uint32_t n, uint32_t k);

/* the gather with its hemisphere and light parts written separately: parts =
 * device float4[2*ntexels] per texel list entry (rgb, reliable) */
int  lsg_gather_parts_device(lsg_scene* scene, const lsg_texel* texels, uint32_t ntexels,
                             const lsg_subtexel* subs, const lsg_gather_params* params,
                             float* out_rgba, float* parts, void* stream);
/* scale s = mean cheap irradiance over reliable texels (synchronous) */
int  lsg_learned_scale(const float* parts, uint32_t n, float* scale);
/* the 32 features per texel (device float[n*32]): the cheap parts, their means over
 * 3x3 / 5x5 / 11x11 windows of same-facing covered texels in the texel's lightmap,
 * position, normal, albedo, spreads, window coverage, the lights' unshadowed direct term.  tri_albedo: device float[ntri*3];
 * maps: host array of the lightmaps the texels' out_index slots fall in, sorted by
 * rgba_base, inside [0, nslots); work: device float4[3*nslots] scratch.  Asynchronous on
 * stream; the lightmap list is copied to the device (and kept there) when it changes */
int  lsg_learned_features(lsg_scene* scene, const lsg_texel* texels, uint32_t n, const lsg_subtexel* subs,
                          const float* parts, const float* tri_albedo, float scale, const lsg_lightmap* maps,
                          uint32_t nmaps, uint32_t nslots, float* work, float* x, void* stream);
/* y: network outputs (device float[n*4]): correction r to the baseline
 * b = log1p((H5 + L)/s) of the features x (device float[n*32]) -- 5x5 hemisphere mean
 * and the texel's own light part -- and predicted error.
 * Writes E = s*expm1(b + r) into out_rgba for texels whose cheap pass was reliable and
 * whose predicted absolute error |y3| * (1 + mean E/s) -- in units of s, the scene's
 * mean irradiance -- is at most tau; appends the others to fallback and counts them in
 * *nfallback (device counter, caller zeroes) */
int  lsg_learned_decide(const float* y, const float* x, const lsg_texel* texels, const float* parts, uint32_t n,
                        float scale, float tau, float* out_rgba, lsg_texel* fallback, uint32_t* nfallback, void* stream);
/* Per-scene error bound.  cheap2_rgba: a second cheap gather with another seed of the
 * texels 0, sample_stride, 2*sample_stride, ... (device float4 per output slot, as
 * lsg_gather_device writes it).  From (pred - c2)^2 - (c1 - c2)^2 / 2, an unbiased
 * estimate of the network's squared error, summed per predicted-error bin over those
 * texels, picks the largest tau whose kept texels stay within target_rmse (relative to
 * the RMS prediction).  work: device uint64[513].  *est_rmse (optional) = estimated
 * relative RMSE with no fallback.  Synchronous. */
int  lsg_learned_bound(const float* y, const float* x, const lsg_texel* texels, const float* parts,
                       const float* cheap2_rgba, uint32_t n, uint32_t sample_stride, float scale, float target_rmse,
                       uint64_t* work, float* tau, float* est_rmse, void* stream);

#ifdef __cplusplus
}
#endif
#endif
