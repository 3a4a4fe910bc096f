// lsgpu_gather.h -- internal interface of the per-texel gather (lsgpu_gather.hip).
#pragma once
#include "lsgpu_internal.h"
#include "lsgpu_material.h"
#include <vector>

namespace lsg {

// SceneView::env_type of a blend of two environments (RRBufferBlend)
constexpr uint32_t kEnvBlend = 3;

// slots of a scene's ring of tile-queue counters for persistent gathers (GatherArgs::pq)
constexpr uint32_t kQueueSlots = 64;

constexpr int kMaxLights = 16;

struct SceneView {
    ColliderView col;
    const float* verts;          // original vertex order, 3 floats each
    const uint32_t* idx;         // 3 per triangle
    const float* vnormal;        // 3 floats per vertex
    const uint8_t* side;         // LSG_SIDE_* per triangle
    const float4* exitance;      // 2 per triangle: front, back
    const float* vtangent;       // 3 floats per vertex (output layers only)
    const float* vbitangent;     // 3 floats per vertex (output layers only)
    const float4* env;           // environment texels, linear rgb, or null
    uint32_t env_type, env_w, env_h;   // LSG_ENV_* or kEnvBlend; env_type 0 = rays that leave return 0
    // kEnvBlend: env/env_w/env_h are the first environment (of type env0_type, 0 = none),
    // env1... the second; each is looked up in its direction rotated by its 3x3 matrix
    // (row-major, env_rot[0..8] and [9..17]) and weighted 1 - env_blend and env_blend
    const float4* env1;
    uint32_t env0_type, env1_type, env1_w, env1_h;
    float env_rot[18];
    float env_blend;
    const float4* ind;           // 3 per triangle: corner k = (indirect irradiance rgb, albedo channel k), or null
    const uint32_t* slot;        // leaf-order slot of each triangle (hit (u, v) recomputed from its record)
    PMView pm;                   // point materials (pm.tri_mat null = none)
};

struct GatherArgs {
    SceneView sc;
    const lsg_texel* texels; uint32_t ntex;
    const lsg_subtexel* subs;
    float4* out;
    float4* parts;               // optional: per texel list entry, (hemisphere rgb, reliable) and (lights rgb, reliable)
    unsigned long long* stats;   // STATS builds: hemi rays, nodes, tris, shadow rays, nodes, tris
    uint32_t quality; uint32_t hemi_rays; uint32_t rounds;
    float tmin, locality, rug, light_mult; uint32_t seed;
    const lsg_light* lights; uint32_t nlights;
    float env_mult;              // environment multiplier (with sc.env_type != 0)
    // output layers (LS_LIGHTMAP, LS_DIRECTION1..3, LS_BENT_NORMALS; null = not wanted),
    // used by the layered kernels only; dirs = gatherAllDirections (lightmap.cpp:820)
    float4* layer[5]; uint32_t dirs; uint32_t layered;
    // subtexel walk records of the first series per texel list entry (k_subtexel_walk, 32
    // bytes each), or null: the kernels then replay the walk themselves.  walk_buf
    // (walk_cap records) is where the launch functions write them; null = no records.
    const struct WalkRec* walk;
    void* walk_buf; uint32_t walk_cap;
    // counters (8 words) for a persistent launch's tile queues, one slot of the scene's
    // ring, or null: launches of many texels with more than 32 hemisphere rays each then
    // go through the persistent kernel (the same results as the one-wave-per-texel one)
    uint32_t* pq;
};

// persistent launches: tiles of texels handed out by atomic counters, one queue per
// XCD (workgroup b runs on XCD b % 8) so that each XCD's L2 sees a contiguous part of
// the scene; a wave whose queue is empty steals from the next one.
struct WorkQueue {
    uint32_t* counters;    // nqueues counters, reset to 0 before each launch
    uint32_t nqueues;      // 1 or 8
    uint32_t ntiles;       // tiles of `tile` texels
    uint32_t tile;
};

struct SceneImpl {
    const lsg_collider* col = nullptr;
    float* d_vnormal = nullptr;
    uint8_t* d_side = nullptr;
    float4* d_exit = nullptr;
    lsg_light* d_lights = nullptr;     // kMaxLights slots
    uint32_t nlights = 0;
    float* d_vtan = nullptr;           // per-vertex tangent frames, built on first use
    float* d_vbit = nullptr;
    float4* d_env = nullptr;           // environment texels, linear rgb
    uint32_t env_type = 0, env_w = 0, env_h = 0;
    float4* d_env1 = nullptr;          // the second environment (lsg_scene_set_environment_blend)
    uint32_t env1_type = 0, env1_w = 0, env1_h = 0;
    bool env_blended = false;          // RRBuffer::createEnvironmentBlend made a blend
    float env_rot[18] = {};            // yaw rotations of the two environments
    float env_blend = 0;
    uint32_t env_version = 0;          // counts environment changes (captured graphs hold them)
    uint32_t* d_pq = nullptr;          // ring of kQueueSlots x 8 tile-queue counters for persistent gathers
    uint32_t pq_next = 0;              // next slot (host atomic), so launches on different streams differ
    lsg_lightmap* d_lmaps = nullptr;   // lightmap list of the last learned-feature pass
    uint32_t lmaps_cap = 0;
    std::vector<lsg_lightmap> h_lmaps; // host copy it was uploaded from
    float4* d_ind = nullptr;           // corner indirect records (3 float4 per triangle), or null
    std::vector<uint8_t> h_side;       // side bits as given (the indirect bits are merged in)
    void* d_walk = nullptr;            // subtexel walk records of the last gather launch (32 B each)
    uint32_t walk_cap = 0;
    // point materials (lsg_scene_set_point_materials); d_tri_mat null = none
    uint32_t* d_tri_mat = nullptr;
    float* d_tri_uv = nullptr;
    lsg_point_material* d_mats = nullptr;
    TexInfo* d_tex = nullptr;
    float4* d_texels = nullptr;
    std::vector<lsg_point_material> h_mats;   // host copy: launches check the keyed materials' quality
    uint64_t pm_version = 0;                  // bumped on every change (bake graphs key on it)
};

// the gather's quality reads the point materials of every keyed material (a keyed material
// below its minimal quality would be the reference's partially transmitting per-triangle
// material, which is not modelled)
int check_point_quality(const SceneImpl& S, uint32_t quality);
void free_point_materials(SceneImpl& S);

// walk-record scratch of the scene for a launch of n texels (grown on demand; not while
// a stream is captured)
int scene_walk_scratch(lsg_scene* s, uint32_t n, GatherArgs& A);
// the first series' shooters per texel: max(hemisphere rays, light rounds when lights shoot)
__host__ __device__ inline uint32_t first_series(const GatherArgs& A)
{
    const uint32_t l = (A.nlights && A.rounds) ? A.rounds : 0u;
    return A.hemi_rays > l ? A.hemi_rays : l;
}

// per-vertex tangent frames of the scene: the caller's, or built from the vertex
// normals on stream st when none were set
int scene_frames(lsg_scene* s, hipStream_t st);

int gather_args(lsg_scene* s, const lsg_texel* texels, uint32_t ntexels, const lsg_subtexel* subs,
                const lsg_gather_params* p, float* out_rgba, GatherArgs& A, int& stack);
int launch_gather(const GatherArgs& A, int stack, hipStream_t st);
// texels per launch of launch_gather's chunks
uint32_t gather_chunk(const GatherArgs& A);
// points A at the scene's walk-record scratch when the launch's kernels read records
int walk_records(lsg_scene* s, GatherArgs& A);
int launch_gather_persistent(const GatherArgs& A, int stack, const WorkQueue& Q, uint32_t waves_per_block, uint32_t blocks,
                             bool refill,
                             hipStream_t st);

} // namespace lsg

struct lsg_scene { lsg::SceneImpl impl; };
