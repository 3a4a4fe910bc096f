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

// Unwrap seams of a batch's lightmaps, on the device: map m's seams are
// seams[begin[m] .. begin[m+1]); `with` lists the maps that have any (one workgroup each).
struct SeamSet {
    const lsg_unwrap_seam* seams = nullptr;
    const uint32_t* begin = nullptr;
    const uint32_t* with = nullptr;
    uint32_t nwith = 0;
};

// a SeamSet's device arrays, owned; version counts the changes (captured graphs hold
// the arrays and the grid)
struct SeamStore {
    lsg_unwrap_seam* d_seams = nullptr;
    uint32_t* d_begin = nullptr;
    uint32_t* d_with = nullptr;
    SeamSet view;
    uint32_t version = 0;
};
// copies host seams (begin: nmaps + 1 offsets) into S, replacing what it held;
// seams == nullptr or no seams at all leaves S empty
int seams_set(SeamStore& S, uint32_t nmaps, const lsg_unwrap_seam* seams, const uint32_t* begin);
void seams_free(SeamStore& S);

int maps_upload(const lsg_lightmap* maps, uint32_t nmaps, FilterMap** d_maps, uint32_t* total);
int scratch_alloc(FilterScratch& S, uint32_t total);
void scratch_free(FilterScratch& S);
// fixed_steps < 0: lightmapGrow iterates until a step changes nothing (host checks
// every 16 steps, one chunk behind the GPU); fixed_steps >= 0: exactly that many steps, no host sync (graph
// capture), S.changed[0] = whether the last step still changed a texel.
int filter_run(const FilterMap* d_maps, uint32_t nmaps, uint32_t total, float4* rgba, float* rgbf,
               const lsg_filter_params& fp, FilterScratch& S, hipStream_t s, bool rgba_out,
               int fixed_steps, uint32_t* steps_used, const SeamSet* seams = nullptr);

} // namespace lsg
