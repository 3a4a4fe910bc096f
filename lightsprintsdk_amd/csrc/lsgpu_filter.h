// lsgpu_filter.h -- internal interface of the lightmap filter chain (lsgpu_filter.hip).
#pragma once
#include "lsgpu_internal.h"

namespace lsg {

struct FilterMap {
    uint32_t first;      // first texel of this map in the batch's flat texel numbering
    uint32_t base;       // float4 slot of texel (0,0) in the RGBA buffer
    uint32_t out_base;   // RGBF texel slot of texel (0,0) in the output buffer
    uint32_t width, height;
};

struct FilterScratch {
    float4* a = nullptr; float4* b = nullptr; float4* c = nullptr;
    uint8_t* flags = nullptr;
    uint32_t* changed = nullptr;   // [0] last step changed; [1..3] frontier sizes; [4..19] step log
    uint32_t* front[2] = {nullptr, nullptr};   // lightmapGrow frontiers (texel indices)
    uint32_t* mark = nullptr;      // step + 1 at which a texel was queued in a frontier, 0 = never
    uint32_t* host_log = nullptr;  // pinned: two chunks of frontier sizes read by the host
    hipEvent_t ev[2] = {nullptr, nullptr};
};

int maps_upload(const lsg_lightmap* maps, uint32_t nmaps, FilterMap** d_maps, uint32_t* total);
int scratch_alloc(FilterScratch& S, uint32_t total);
void scratch_free(FilterScratch& S);
// fixed_steps < 0: lightmapGrow iterates until a step changes nothing (host checks
// every 16 steps, one chunk behind the GPU); fixed_steps >= 0: exactly that many steps, no host sync (graph
// capture), S.changed[0] = whether the last step still changed a texel.
int filter_run(const FilterMap* d_maps, uint32_t nmaps, uint32_t total, float4* rgba, float* rgbf,
               const lsg_filter_params& fp, FilterScratch& S, hipStream_t s, bool rgba_out,
               int fixed_steps, uint32_t* steps_used);

} // namespace lsg
