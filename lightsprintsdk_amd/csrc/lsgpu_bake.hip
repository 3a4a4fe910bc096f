// lsgpu_bake.hip -- whole-scene lightmap bake plan: RRSolver::updateLightmaps
// (RRSolver/lightmap.cpp:943-1163) for many objects in one go on gfx950.
//
// A plan owns, in HBM, every object's texels and SubTexels (rasterised once on the
// host), one RGBA temp texel array for all lightmaps, the RGBF output, the filter
// scratch and the per-triangle first-gather work.  One run is
//   first gather per triangle (gatherPerTrianglePhysical, gather.cpp:843-952)
//   -> exitance = irradiance * albedo + emittance          (k_exitance)
//   -> final gather of every texel of every object         (persistent launch,
//      tiles from XCD-local work queues)
//   -> filter chain + RGBF writeback                        (lsgpu_filter.hip)
// and can be captured once into a hipGraph and replayed (lights and exitance live in
// device memory, so a sun sweep only updates buffers between replays).
#include "lsgpu_internal.h"
#include "lsgpu_gather.h"
#include "lsgpu_filter.h"
#include <cmath>
#include <vector>
#include <cstring>
#include <chrono>

namespace lsg {

// a triangle with a point material keeps its irradiance in the table: the gather's hit
// multiplies it by the material's diffuse colour and adds the emittance itself
__device__ inline bool has_material(const uint32_t* mat, uint32_t t) { return mat && mat[t] != 0xFFFFFFFFu; }

__global__ __launch_bounds__(256) void k_exitance(uint32_t nt, const float4* __restrict__ irr, const float* __restrict__ albedo,
                                                  const float* __restrict__ emit, const uint32_t* __restrict__ mat,
                                                  float4* __restrict__ ex)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    // PathtracerWorker::getIncidentRadiance at depth 0 with solver shortcuts
    // (pathtracer.cpp:235-316): direct irradiance * diffuse reflectance + emittance
    const float4 e = irr[t];
    ex[2 * (size_t)t] = has_material(mat, t) ? make_float4(e.x, e.y, e.z, 0.f)
                        : make_float4(e.x * albedo[3 * (size_t)t] + emit[3 * (size_t)t],
                                      e.y * albedo[3 * (size_t)t + 1] + emit[3 * (size_t)t + 1],
                                      e.z * albedo[3 * (size_t)t + 2] + emit[3 * (size_t)t + 2], 0.f);
    ex[2 * (size_t)t + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// exitance = emittance only, the light the first gather's hemisphere rays may see: the
// first gather is the direct illumination (lights, environment, emitters), and a table
// left over from an earlier run would add a bounce per run
__global__ __launch_bounds__(256) void k_emittance_only(uint32_t nt, const float* __restrict__ emit, const uint32_t* __restrict__ mat,
                                                        float4* __restrict__ ex)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    ex[2 * (size_t)t] = has_material(mat, t) ? make_float4(0.f, 0.f, 0.f, 0.f)
                        : make_float4(emit[3 * (size_t)t], emit[3 * (size_t)t + 1], emit[3 * (size_t)t + 2], 0.f);
    ex[2 * (size_t)t + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// exitance after a bounce: emittance + albedo * (light part of the first gather +
// hemisphere part gathered over the previous table); hemi == nullptr: the first gather's
// own hemisphere part (parts[2t])
__global__ __launch_bounds__(256) void k_exitance_bounce(uint32_t nt, const float4* __restrict__ parts, const float4* __restrict__ hemi,
                                                         const float* __restrict__ albedo, const float* __restrict__ emit,
                                                         const uint32_t* __restrict__ mat, float4* __restrict__ ex)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    const float4 h = hemi ? hemi[t] : parts[2 * (size_t)t];
    const float4 l = parts[2 * (size_t)t + 1];
    ex[2 * (size_t)t] = has_material(mat, t) ? make_float4(h.x + l.x, h.y + l.y, h.z + l.z, 0.f)
                        : make_float4((h.x + l.x) * albedo[3 * (size_t)t] + emit[3 * (size_t)t],
                                      (h.y + l.y) * albedo[3 * (size_t)t + 1] + emit[3 * (size_t)t + 1],
                                      (h.z + l.z) * albedo[3 * (size_t)t + 2] + emit[3 * (size_t)t + 2], 0.f);
    ex[2 * (size_t)t + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_zero_f4(float4* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

static inline void dfree(void* p) { if (p) (void)hipFree(p); }
static inline unsigned nblk(uint64_t n) { return (unsigned)((n + 255) / 256); }

struct BakeImpl {
    lsg_scene* scene = nullptr;
    uint32_t ntri = 0, ntex = 0, nsub = 0, nmaps = 0, nslots = 0;
    std::vector<lsg_lightmap> maps;
    lsg_texel* d_tex = nullptr; lsg_subtexel* d_sub = nullptr;
    float4* d_rgba = nullptr; float* d_rgbf = nullptr;
    FilterMap* d_maps = nullptr; uint32_t ftotal = 0;
    FilterScratch S;
    lsg_texel* d_ttex = nullptr; lsg_subtexel* d_tsub = nullptr; float4* d_irr = nullptr;
    float* d_albedo = nullptr; float* d_emit = nullptr;
    uint32_t* d_queue = nullptr;
    void* d_walk = nullptr;            // subtexel walk records of the final gather's texels (32 B each)
    lsg_bake_tuning tuning{};
    int cus = 256;
    hipStream_t own = nullptr;
    hipGraph_t graph = nullptr; hipGraphExec_t exec = nullptr;
    std::vector<uint8_t> graph_key;
    uint32_t grow_steps = 0;
    uint32_t bounces = 0, bounce_quality = 0;
    float4* d_parts = nullptr;         // first gather's (hemisphere, light) parts per triangle, with bounces
    float4* d_hemi = nullptr;          // a bounce's hemisphere gather per triangle
    float4* d_lrgba[4] = {nullptr, nullptr, nullptr, nullptr};   // layers 1..4 (layer 0 = d_rgba / d_rgbf)
    float* d_lrgbf[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<std::vector<lsg_unwrap_seam>> seam_lists;   // per lightmap (object), host copies
    SeamStore seams;
    ~BakeImpl()
    {
        seams_free(seams);
        for (int i = 0; i < 4; i++) { dfree(d_lrgba[i]); dfree(d_lrgbf[i]); }
        dfree(d_parts); dfree(d_hemi);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (own) (void)hipStreamDestroy(own);
        dfree(d_tex); dfree(d_sub); dfree(d_rgba); dfree(d_rgbf); dfree(d_maps); scratch_free(S);
        dfree(d_ttex); dfree(d_tsub); dfree(d_irr); dfree(d_albedo); dfree(d_emit); dfree(d_queue); dfree(d_walk);
    }
};

static void default_tuning(lsg_bake_tuning& t)
{
    if (!t.tile) t.tile = 2;
    if (!t.waves_per_block) t.waves_per_block = 4;
    if (!t.blocks_per_cu) t.blocks_per_cu = 4;
    if (!t.xcd_queues) t.xcd_queues = 8;
    if (!t.lane_refill) t.lane_refill = 2;
}

static int run_gather_final(BakeImpl& B, const lsg_gather_params* p, const lsg_texel* tex, uint32_t ntex, hipStream_t st)
{
    GatherArgs A;
    int stack = 0;
    int e = gather_args(B.scene, tex, ntex, B.d_sub, p, (float*)B.d_rgba, A, stack);
    if (e) return e;
    A.walk_buf = B.d_walk; A.walk_cap = B.ntex;   // the plan's own records: nothing allocated under capture
    lsg_bake_tuning t = B.tuning;
    default_tuning(t);
    WorkQueue Q;
    Q.counters = B.d_queue; Q.nqueues = t.xcd_queues > 1 ? 8 : 1; Q.tile = t.tile;
    Q.ntiles = (ntex + t.tile - 1) / t.tile;
    if (!Q.ntiles) return LSG_OK;
    const uint32_t blocks = (uint32_t)B.cus * t.blocks_per_cu;
    return launch_gather_persistent(A, stack, Q, t.waves_per_block, blocks, t.lane_refill == 2, st);
}

static int run_first(BakeImpl& B, const lsg_gather_params* first, hipStream_t st)
{
    GatherArgs A;
    int stack = 0, e;
    if ((e = gather_args(B.scene, B.d_ttex, B.ntri, B.d_tsub, first, (float*)B.d_irr, A, stack))) return e;
    k_emittance_only<<<nblk(B.ntri), 256, 0, st>>>(B.ntri, B.d_emit, B.scene->impl.d_tri_mat, B.scene->impl.d_exit);
    if (!B.bounces) {
        if ((e = launch_gather(A, stack, st))) return e;
        k_exitance<<<nblk(B.ntri), 256, 0, st>>>(B.ntri, B.d_irr, B.d_albedo, B.d_emit, B.scene->impl.d_tri_mat, B.scene->impl.d_exit);
        return LSG_OK;
    }
    // radiosity: keep the first gather's light part, regather the hemisphere each bounce
    A.parts = B.d_parts;
    if ((e = launch_gather(A, stack, st))) return e;
    k_exitance_bounce<<<nblk(B.ntri), 256, 0, st>>>(B.ntri, B.d_parts, nullptr, B.d_albedo, B.d_emit, B.scene->impl.d_tri_mat,
                                                     B.scene->impl.d_exit);
    lsg_gather_params bp = *first;
    bp.hemisphere = 1;
    bp.light_multiplier = 0.f;
    bp.light_rounds = 0;
    if (B.bounce_quality) bp.quality = B.bounce_quality;
    for (uint32_t b = 0; b < B.bounces; b++) {
        bp.seed = first->seed ^ (0x9E3779B9u * (b + 1));
        GatherArgs H;
        if ((e = gather_args(B.scene, B.d_ttex, B.ntri, B.d_tsub, &bp, (float*)B.d_hemi, H, stack))) return e;
        if ((e = launch_gather(H, stack, st))) return e;
        k_exitance_bounce<<<nblk(B.ntri), 256, 0, st>>>(B.ntri, B.d_parts, B.d_hemi, B.d_albedo, B.d_emit, B.scene->impl.d_tri_mat,
                                                         B.scene->impl.d_exit);
    }
    return LSG_OK;
}

// the whole bake on stream st; fixed_grow < 0 = adaptive grow (not capturable)
static int run_all(BakeImpl& B, const lsg_gather_params* first, const lsg_gather_params* fin, const lsg_filter_params* fp,
                   hipStream_t st, int fixed_grow, uint32_t* steps_used)
{
    int e;
    if (first && (e = run_first(B, first, st))) return e;
    if (B.nslots) k_zero_f4<<<nblk(B.nslots), 256, 0, st>>>(B.d_rgba, B.nslots);   // a plan without lightmaps only gathers the solution
    if ((e = run_gather_final(B, fin, B.d_tex, B.ntex, st))) return e;
    if (fp) {
        if ((e = filter_run(B.d_maps, B.nmaps, B.ftotal, B.d_rgba, B.d_rgbf, *fp, B.S, st, true, fixed_grow, steps_used, &B.seams.view)))
            return e;
    }
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

} // namespace lsg

struct lsg_bake { lsg::BakeImpl impl; };

using namespace lsg;

extern "C" {

int lsg_scene_exitance_device(lsg_scene* scene, const float* irr, const float* albedo, const float* emit, void* stream)
{
    if (!scene || !albedo || !emit) return fail(LSG_ERR_INVALID, "lsg_scene_exitance_device: null argument");
    const uint32_t nt = scene->impl.col->impl.ntri;
    if (!nt) return LSG_OK;
    hipStream_t st = (hipStream_t)stream;
    if (irr) k_exitance<<<nblk(nt), 256, 0, st>>>(nt, (const float4*)irr, albedo, emit, scene->impl.d_tri_mat, scene->impl.d_exit);
    else k_emittance_only<<<nblk(nt), 256, 0, st>>>(nt, emit, scene->impl.d_tri_mat, scene->impl.d_exit);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

int lsg_bake_create(lsg_scene* scene, const lsg_bake_object* objects, uint32_t nobjects, const float* albedo,
                    const float* emittance, lsg_bake** out)
{
    if (!out) return fail(LSG_ERR_INVALID, "lsg_bake_create: out is null");
    *out = nullptr;
    if (!scene || (nobjects && !objects) || !albedo || !emittance) return fail(LSG_ERR_INVALID, "lsg_bake_create: null argument");
    lsg_bake* b = new (std::nothrow) lsg_bake;
    if (!b) return fail(LSG_ERR_OOM, "lsg_bake_create: host allocation");
    BakeImpl& B = b->impl;
    B.scene = scene;
    B.ntri = scene->impl.col->impl.ntri;
    auto bail = [&](int st) { delete b; return st; };
    // 1. rasterise every object (lightmap.cpp:83-335) into one texel/SubTexel list
    std::vector<lsg_texel> tex;
    std::vector<lsg_subtexel> sub;
    uint32_t slot = 0;
    for (uint32_t o = 0; o < nobjects; o++) {
        const lsg_bake_object& ob = objects[o];
        uint32_t nt = 0, ns = 0;
        int st = lsg_rasterize_texels(ob.tri_uv, ob.tri_ids, ob.ntri, ob.width, ob.height, o, nullptr, &nt, nullptr, &ns, slot);
        if (st) return bail(st);
        size_t t0 = tex.size(), s0 = sub.size();
        tex.resize(t0 + nt); sub.resize(s0 + ns);
        st = lsg_rasterize_texels(ob.tri_uv, ob.tri_ids, ob.ntri, ob.width, ob.height, o, tex.data() + t0, &nt, sub.data() + s0, &ns, slot);
        if (st) return bail(st);
        for (size_t k = t0; k < tex.size(); k++) tex[k].sub_begin += (uint32_t)s0;
        lsg_lightmap m{slot, ob.width, ob.height, slot};
        B.maps.push_back(m);
        uint64_t ns64 = (uint64_t)slot + (uint64_t)ob.width * ob.height;
        if (ns64 >= (1ull << 31)) return bail(fail(LSG_ERR_LIMIT, "lsg_bake_create: more than 2^31 lightmap texels"));
        slot = (uint32_t)ns64;
    }
    B.ntex = (uint32_t)tex.size(); B.nsub = (uint32_t)sub.size(); B.nmaps = nobjects; B.nslots = slot;
    hipError_t e;
#define AL(p, n) do { if ((e = hipMalloc(&(p), (n) ? (n) : 16)) != hipSuccess) return bail(hip_fail(e, "hipMalloc", __LINE__)); } while (0)
#define UP(d, h, n) do { if ((n) && (e = hipMemcpy((d), (h), (n), hipMemcpyHostToDevice)) != hipSuccess) return bail(hip_fail(e, "h2d", __LINE__)); } while (0)
    AL(B.d_tex, sizeof(lsg_texel) * tex.size()); UP(B.d_tex, tex.data(), sizeof(lsg_texel) * tex.size());
    AL(B.d_sub, sizeof(lsg_subtexel) * sub.size()); UP(B.d_sub, sub.data(), sizeof(lsg_subtexel) * sub.size());
    AL(B.d_rgba, 16 * (size_t)slot);
    AL(B.d_rgbf, 12 * (size_t)slot);
    AL(B.d_queue, 64);
    AL(B.d_walk, 32 * tex.size());
    // 2. first-gather work: one texel per triangle covering the whole triangle
    std::vector<lsg_texel> ttex(B.ntri);
    std::vector<lsg_subtexel> tsub(B.ntri);
    for (uint32_t t = 0; t < B.ntri; t++) {
        ttex[t] = lsg_texel{t, t, 1u | (255u << 16), 0x80000000u | t};
        tsub[t].triangle = t;
        const float uv[6] = {0, 0, 1, 0, 0, 1};
        memcpy(tsub[t].uv, uv, sizeof uv);
        tsub[t].area = 1.f;
    }
    AL(B.d_ttex, sizeof(lsg_texel) * B.ntri); UP(B.d_ttex, ttex.data(), sizeof(lsg_texel) * B.ntri);
    AL(B.d_tsub, sizeof(lsg_subtexel) * B.ntri); UP(B.d_tsub, tsub.data(), sizeof(lsg_subtexel) * B.ntri);
    AL(B.d_irr, 16 * (size_t)B.ntri);
    AL(B.d_albedo, 12 * (size_t)B.ntri); UP(B.d_albedo, albedo, 12 * (size_t)B.ntri);
    AL(B.d_emit, 12 * (size_t)B.ntri); UP(B.d_emit, emittance, 12 * (size_t)B.ntri);
#undef AL
#undef UP
    int st = maps_upload(B.maps.data(), B.nmaps, &B.d_maps, &B.ftotal);
    if (st) return bail(st);
    if ((st = scratch_alloc(B.S, B.ftotal))) return bail(st);
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) B.cus = prop.multiProcessorCount;
    default_tuning(B.tuning);
    *out = b;
    return LSG_OK;
}

void lsg_bake_destroy(lsg_bake* b) { delete b; }

int lsg_bake_info_get(const lsg_bake* b, lsg_bake_info* info)
{
    if (!b || !info) return fail(LSG_ERR_INVALID, "lsg_bake_info_get: null argument");
    const BakeImpl& B = b->impl;
    info->num_texels = B.ntex; info->num_subtexels = B.nsub; info->num_slots = B.nslots;
    info->num_lightmaps = B.nmaps; info->num_triangles = B.ntri; info->compute_units = (uint32_t)B.cus;
    info->grow_steps = B.grow_steps;
    info->tuning = B.tuning;
    info->rgba = B.d_rgba; info->rgbf = B.d_rgbf; info->texels = B.d_tex; info->subtexels = B.d_sub;
    return LSG_OK;
}

int lsg_bake_set_tuning(lsg_bake* b, const lsg_bake_tuning* t)
{
    if (!b || !t) return fail(LSG_ERR_INVALID, "lsg_bake_set_tuning: null argument");
    if (t->waves_per_block && t->waves_per_block != 2 && t->waves_per_block != 4 && t->waves_per_block != 8)
        return fail(LSG_ERR_INVALID, "waves_per_block must be 2, 4 or 8");
    if (t->tile > 1024 || t->blocks_per_cu > 64 || t->lane_refill > 2) return fail(LSG_ERR_INVALID, "tuning out of range");
    b->impl.tuning = *t;
    default_tuning(b->impl.tuning);
    b->impl.graph_key.clear();   // invalidate a captured graph
    return LSG_OK;
}

int lsg_bake_run(lsg_bake* b, const lsg_gather_params* first, const lsg_gather_params* final_params,
                 const lsg_filter_params* filtering, uint32_t flags, void* stream)
{
    if (!b || !final_params) return fail(LSG_ERR_INVALID, "lsg_bake_run: null argument");
    BakeImpl& B = b->impl;
    hipStream_t st = (hipStream_t)stream;
    if (!(flags & LSG_BAKE_USE_GRAPH)) {
        uint32_t used = 0;
        int e = run_all(B, first, final_params, filtering, st, -1, &used);
        if (!e && filtering) B.grow_steps = used;
        return e;
    }
    // graph mode: capture on the bake's own stream once per parameter set, replay after
    if (!B.own) LSG_HIP(hipStreamCreate(&B.own));
    std::vector<uint8_t> key(sizeof(lsg_gather_params) * 2 + sizeof(lsg_filter_params) + 2, 0);
    if (first) { memcpy(key.data(), first, sizeof *first); key[key.size() - 2] = 1; }
    memcpy(key.data() + sizeof(lsg_gather_params), final_params, sizeof *final_params);
    if (filtering) { memcpy(key.data() + 2 * sizeof(lsg_gather_params), filtering, sizeof *filtering); key[key.size() - 1] = 1; }
    {
        // the captured kernels hold the environment's buffer and shape as arguments
        const SceneImpl& S = B.scene->impl;
        // (and the indirect records, point materials and unwrap seams likewise)
        const uintptr_t env[9] = {(uintptr_t)S.d_env, S.env_type, S.env_w, S.env_h, (uintptr_t)S.d_ind, (uintptr_t)S.d_tri_mat,
                                  (uintptr_t)S.pm_version, (uintptr_t)B.seams.version, (uintptr_t)S.env_version};
        const uint8_t* e8 = (const uint8_t*)env;
        key.insert(key.end(), e8, e8 + sizeof env);
    }
    if (key != B.graph_key || !B.exec) {
        // an eager run fixes the number of grow steps the graph will replay
        uint32_t used = 0;
        int e = run_all(B, first, final_params, filtering, B.own, -1, &used);
        if (e) return e;
        LSG_HIP(hipStreamSynchronize(B.own));
        B.grow_steps = used;
        if (B.exec) { (void)hipGraphExecDestroy(B.exec); B.exec = nullptr; }
        if (B.graph) { (void)hipGraphDestroy(B.graph); B.graph = nullptr; }
        LSG_HIP(hipStreamBeginCapture(B.own, hipStreamCaptureModeThreadLocal));
        e = run_all(B, first, final_params, filtering, B.own, (int)used, nullptr);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(B.own, &g);
        if (e) { if (g) (void)hipGraphDestroy(g); return e; }
        if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture", __LINE__);
        B.graph = g;
        LSG_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
        B.graph_key = key;
    }
    if (st) { hipEvent_t ev; LSG_HIP(hipEventCreate(&ev)); LSG_HIP(hipEventRecord(ev, st)); LSG_HIP(hipStreamWaitEvent(B.own, ev, 0)); (void)hipEventDestroy(ev); }
    LSG_HIP(hipGraphLaunch(B.exec, B.own));
    if (filtering && B.grow_steps) {
        // the replayed grow ran a fixed number of steps; if its last step still changed
        // texels the coverage changed since capture: finish eagerly like the reference
        uint32_t last = 0;
        LSG_HIP(hipMemcpyAsync(&last, B.S.changed, 4, hipMemcpyDeviceToHost, B.own));
        LSG_HIP(hipStreamSynchronize(B.own));
        if (last) {
            uint32_t used = 0;
            int e = run_all(B, first, final_params, filtering, B.own, -1, &used);
            if (e) return e;
            B.grow_steps = used;
            B.graph_key.clear();
        }
    }
    if (st) { hipEvent_t ev; LSG_HIP(hipEventCreate(&ev)); LSG_HIP(hipEventRecord(ev, B.own)); LSG_HIP(hipStreamWaitEvent(st, ev, 0)); (void)hipEventDestroy(ev); }
    else LSG_HIP(hipStreamSynchronize(B.own));
    return LSG_OK;
}

int lsg_bake_run_layers(lsg_bake* b, const lsg_gather_params* first, const lsg_gather_params* final_params,
                        const lsg_filter_params* filtering, uint32_t layer_mask, void* stream)
{
    if (!b || !final_params) return fail(LSG_ERR_INVALID, "lsg_bake_run_layers: null argument");
    if (!layer_mask || layer_mask >= (1u << LSG_NUM_LAYERS)) return fail(LSG_ERR_INVALID, "lsg_bake_run_layers: layer_mask %#x", layer_mask);
    BakeImpl& B = b->impl;
    hipStream_t st = (hipStream_t)stream;
    float4* rgba[LSG_NUM_LAYERS];
    float* rgbf[LSG_NUM_LAYERS];
    for (int i = 0; i < LSG_NUM_LAYERS; i++) {
        rgba[i] = nullptr; rgbf[i] = nullptr;
        if (!(layer_mask >> i & 1u)) continue;
        if (i == 0) { rgba[0] = B.d_rgba; rgbf[0] = B.d_rgbf; continue; }
        if (!B.d_lrgba[i - 1]) LSG_HIP(hipMalloc(&B.d_lrgba[i - 1], 16 * (size_t)(B.nslots ? B.nslots : 1)));
        if (!B.d_lrgbf[i - 1]) LSG_HIP(hipMalloc(&B.d_lrgbf[i - 1], 12 * (size_t)(B.nslots ? B.nslots : 1)));
        rgba[i] = B.d_lrgba[i - 1]; rgbf[i] = B.d_lrgbf[i - 1];
    }
    int e;
    if (first && (e = run_first(B, first, st))) return e;
    for (int i = 0; i < LSG_NUM_LAYERS; i++)
        if (rgba[i] && B.nslots) k_zero_f4<<<nblk(B.nslots), 256, 0, st>>>(rgba[i], B.nslots);
    if ((e = lsg_gather_layers_device(B.scene, B.d_tex, B.ntex, B.d_sub, B.nsub, final_params, (float* const*)rgba, st))) return e;
    if (filtering) {
        for (int i = 0; i < LSG_NUM_LAYERS; i++) {
            if (!rgba[i]) continue;
            uint32_t used = 0;
            if ((e = filter_run(B.d_maps, B.nmaps, B.ftotal, rgba[i], rgbf[i], *filtering, B.S, st, true, -1, &used, &B.seams.view)))
                return e;
            if (i == 0) B.grow_steps = used;
        }
    }
    LSG_HIP(hipGetLastError());
    B.graph_key.clear();
    return LSG_OK;
}

int lsg_bake_download_layer(const lsg_bake* b, uint32_t object, uint32_t layer, float* rgbf)
{
    if (!b || !rgbf) return fail(LSG_ERR_INVALID, "lsg_bake_download_layer: null argument");
    const BakeImpl& B = b->impl;
    if (object >= B.nmaps) return fail(LSG_ERR_INVALID, "lsg_bake_download_layer: object %u of %u", object, B.nmaps);
    if (layer >= LSG_NUM_LAYERS) return fail(LSG_ERR_INVALID, "lsg_bake_download_layer: layer %u", layer);
    const float* src = layer ? B.d_lrgbf[layer - 1] : B.d_rgbf;
    if (!src) return fail(LSG_ERR_INVALID, "lsg_bake_download_layer: layer %u was never baked", layer);
    const lsg_lightmap& m = B.maps[object];
    LSG_HIP(hipMemcpy(rgbf, src + 3 * (size_t)m.rgbf_base, 12 * (size_t)m.width * m.height, hipMemcpyDeviceToHost));
    return LSG_OK;
}

int lsg_bake_download_layer_rgba(const lsg_bake* b, uint32_t object, uint32_t layer, float* rgba)
{
    if (!b || !rgba) return fail(LSG_ERR_INVALID, "lsg_bake_download_layer_rgba: null argument");
    const BakeImpl& B = b->impl;
    if (object >= B.nmaps) return fail(LSG_ERR_INVALID, "lsg_bake_download_layer_rgba: object %u of %u", object, B.nmaps);
    if (layer >= LSG_NUM_LAYERS) return fail(LSG_ERR_INVALID, "lsg_bake_download_layer_rgba: layer %u", layer);
    const float4* src = layer ? B.d_lrgba[layer - 1] : B.d_rgba;
    if (!src || (layer && !B.d_lrgbf[layer - 1])) return fail(LSG_ERR_INVALID, "lsg_bake_download_layer_rgba: layer %u was never baked", layer);
    const lsg_lightmap& m = B.maps[object];
    LSG_HIP(hipMemcpy(rgba, src + m.rgba_base, 16 * (size_t)m.width * m.height, hipMemcpyDeviceToHost));
    return LSG_OK;
}

int lsg_bake_set_seams(lsg_bake* b, uint32_t object, const lsg_unwrap_seam* seams, uint32_t nseams)
{
    if (!b) return fail(LSG_ERR_INVALID, "lsg_bake_set_seams: null plan");
    BakeImpl& B = b->impl;
    if (object >= B.nmaps) return fail(LSG_ERR_INVALID, "lsg_bake_set_seams: object %u of %u", object, B.nmaps);
    if (nseams && !seams) return fail(LSG_ERR_INVALID, "lsg_bake_set_seams: null seams");
    for (uint32_t i = 0; i < nseams; i++)
        for (int k = 0; k < 4; k++)
            if (!std::isfinite(seams[i].edge1[k]) || !std::isfinite(seams[i].edge2[k]))
                return fail(LSG_ERR_INVALID, "lsg_bake_set_seams: seam %u has a coordinate that is not finite", i);
    B.seam_lists.resize(B.nmaps);
    B.seam_lists[object].assign(seams, seams + nseams);
    std::vector<lsg_unwrap_seam> all;
    std::vector<uint32_t> begin(B.nmaps + 1, 0);
    for (uint32_t m = 0; m < B.nmaps; m++) {
        all.insert(all.end(), B.seam_lists[m].begin(), B.seam_lists[m].end());
        begin[m + 1] = (uint32_t)all.size();
    }
    B.graph_key.clear();
    return seams_set(B.seams, B.nmaps, all.empty() ? nullptr : all.data(), begin.data());
}

int lsg_bake_set_bounces(lsg_bake* b, uint32_t bounces, uint32_t quality)
{
    if (!b) return fail(LSG_ERR_INVALID, "lsg_bake_set_bounces: null plan");
    if (bounces > 64) return fail(LSG_ERR_INVALID, "lsg_bake_set_bounces: %u bounces, at most 64", bounces);
    BakeImpl& B = b->impl;
    if (bounces && !B.d_parts) {
        LSG_HIP(hipMalloc(&B.d_parts, 32 * (size_t)(B.ntri ? B.ntri : 1)));
        LSG_HIP(hipMalloc(&B.d_hemi, 16 * (size_t)(B.ntri ? B.ntri : 1)));
    }
    B.bounces = bounces;
    B.bounce_quality = quality;
    B.graph_key.clear();        // the captured sequence changes
    return LSG_OK;
}

int lsg_bake_gather_sample(lsg_bake* b, const lsg_gather_params* p, uint32_t stride, float* ms_out)
{
    // times the final gather alone over every stride-th texel with the current tuning
    if (!b || !p || !stride || !ms_out) return fail(LSG_ERR_INVALID, "lsg_bake_gather_sample: bad argument");
    BakeImpl& B = b->impl;
    std::vector<lsg_texel> all(B.ntex), pick;
    LSG_HIP(hipMemcpy(all.data(), B.d_tex, sizeof(lsg_texel) * B.ntex, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < B.ntex; i += stride) pick.push_back(all[i]);
    lsg_texel* d = nullptr;
    LSG_HIP(hipMalloc(&d, sizeof(lsg_texel) * (pick.size() ? pick.size() : 1)));
    hipError_t e = hipMemcpy(d, pick.data(), sizeof(lsg_texel) * pick.size(), hipMemcpyHostToDevice);
    hipEvent_t a, z;
    (void)hipEventCreate(&a); (void)hipEventCreate(&z);
    int st = e == hipSuccess ? LSG_OK : hip_fail(e, "h2d", __LINE__);
    if (!st) st = run_gather_final(B, p, d, (uint32_t)pick.size(), nullptr);   // warm
    if (!st) { (void)hipEventRecord(a, nullptr); st = run_gather_final(B, p, d, (uint32_t)pick.size(), nullptr); (void)hipEventRecord(z, nullptr); }
    if (!st && (e = hipEventSynchronize(z)) != hipSuccess) st = hip_fail(e, "sync", __LINE__);
    if (!st) (void)hipEventElapsedTime(ms_out, a, z);
    (void)hipEventDestroy(a); (void)hipEventDestroy(z);
    dfree(d);
    return st;
}

int lsg_bake_download(const lsg_bake* b, uint32_t object, float* rgbf)
{
    if (!b || !rgbf) return fail(LSG_ERR_INVALID, "lsg_bake_download: null argument");
    const BakeImpl& B = b->impl;
    if (object >= B.nmaps) return fail(LSG_ERR_INVALID, "lsg_bake_download: object %u of %u", object, B.nmaps);
    const lsg_lightmap& m = B.maps[object];
    LSG_HIP(hipMemcpy(rgbf, B.d_rgbf + 3 * (size_t)m.rgbf_base, 12 * (size_t)m.width * m.height, hipMemcpyDeviceToHost));
    return LSG_OK;
}

} // extern "C"
