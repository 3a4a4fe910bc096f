// lsgpu_material.hip -- point materials of a gather scene (host side): textures decoded into
// the texel values the gather reads, material and per-triangle tables uploaded to HBM.
//
// A texel is decoded as RRBufferInMemory::getElement decodes it (RRBufferInMemory.cpp:
// 295-360): bytes times RR_BYTE2FLOAT's 1/255 (RRMath.h:79), floats as stored, luminance
// copied to rgb, alpha 1 where the format has none.  With the solver's colour space a
// scaled texture is made linear: RGB and RGBA bytes through RRGammaScaler's byte table
// pow(float(i) / 255, 1 / gamma), other formats by toLinear of the decoded value
// (RRColorSpace.cpp:28-70); the optimized transmittance-in-alpha path makes the stored
// diffuse colour linear with toLinear whatever the texture's scale (RRObject.cpp:240-244).
// pow is evaluated in binary64 and rounded to binary32, as the reference's build calls the
// double pow for float arguments.
#include "lsgpu_gather.h"
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

namespace lsg {
namespace {

inline float byte2float(uint8_t b) { return (float)b * 0.003921568627450980392156862745098f; }

inline float to_linear(float a, float inv)
{
    return a >= 0 ? (float)std::pow((double)a, (double)inv) : -(float)std::pow((double)-a, (double)inv);
}

bool texture_format_ok(int32_t f)
{
    return f == LSG_FORMAT_RGB || f == LSG_FORMAT_BGR || f == LSG_FORMAT_RGBA || f == LSG_FORMAT_RGBF || f == LSG_FORMAT_RGBAF
           || f == LSG_FORMAT_LUMINANCE || f == LSG_FORMAT_LUMINANCEF;
}

// getElement(i, nullptr)
float4 decode(const lsg_texture& T, size_t i)
{
    const uint8_t* b = (const uint8_t*)T.data;
    const float* f = (const float*)T.data;
    switch (T.format) {
    case LSG_FORMAT_RGB: return make_float4(byte2float(b[3 * i]), byte2float(b[3 * i + 1]), byte2float(b[3 * i + 2]), 1.f);
    case LSG_FORMAT_BGR: return make_float4(byte2float(b[3 * i + 2]), byte2float(b[3 * i + 1]), byte2float(b[3 * i]), 1.f);
    case LSG_FORMAT_RGBA:
        return make_float4(byte2float(b[4 * i]), byte2float(b[4 * i + 1]), byte2float(b[4 * i + 2]), byte2float(b[4 * i + 3]));
    case LSG_FORMAT_RGBF: return make_float4(f[3 * i], f[3 * i + 1], f[3 * i + 2], 1.f);
    case LSG_FORMAT_RGBAF: return make_float4(f[4 * i], f[4 * i + 1], f[4 * i + 2], f[4 * i + 3]);
    case LSG_FORMAT_LUMINANCE: { const float l = byte2float(b[i]); return make_float4(l, l, l, 1.f); }
    default: return make_float4(f[i], f[i], f[i], 1.f);
    }
}

std::atomic<uint64_t> g_pm_version{1};

} // namespace

void free_point_materials(SceneImpl& S)
{
    for (void* p : {(void*)S.d_tri_mat, (void*)S.d_tri_uv, (void*)S.d_mats, (void*)S.d_tex, (void*)S.d_texels})
        if (p) (void)hipFree(p);
    S.d_tri_mat = nullptr; S.d_tri_uv = nullptr; S.d_mats = nullptr; S.d_tex = nullptr; S.d_texels = nullptr;
    S.h_mats.clear();
    S.pm_version = g_pm_version++;
}

int check_point_quality(const SceneImpl& S, uint32_t quality)
{
    for (const lsg_point_material& m : S.h_mats)
        if (m.transmittance_texture >= 0 && !(quality > m.min_quality && 2 * quality > m.min_quality))
            return fail(LSG_ERR_INVALID, "gather: quality %u does not exceed a keyed material's minimal quality %u "
                        "(its partial per-triangle transmittance is not modelled)", quality, m.min_quality);
    return LSG_OK;
}

} // namespace lsg

using namespace lsg;

extern "C" {

int lsg_scene_set_point_materials(lsg_scene* s, const uint32_t* tri_material, const float* tri_uv,
                                  const lsg_point_material* materials, uint32_t nmaterials,
                                  const lsg_texture* textures, uint32_t ntextures, float gamma)
{
    if (!s) return fail(LSG_ERR_INVALID, "lsg_scene_set_point_materials: null scene");
    SceneImpl& S = s->impl;
    if (!tri_material) { free_point_materials(S); return LSG_OK; }
    const uint32_t nt = S.col->impl.ntri;
    if (!tri_uv || (nmaterials && !materials) || (ntextures && !textures))
        return fail(LSG_ERR_INVALID, "lsg_scene_set_point_materials: null argument");
    if (!(gamma >= 0) || !std::isfinite(gamma)) return fail(LSG_ERR_INVALID, "lsg_scene_set_point_materials: gamma %g", gamma);
    for (uint32_t t = 0; t < nt; t++)
        if (tri_material[t] != 0xFFFFFFFFu && tri_material[t] >= nmaterials)
            return fail(LSG_ERR_INVALID, "lsg_scene_set_point_materials: triangle %u has material %u of %u", t, tri_material[t], nmaterials);
    std::vector<uint8_t> alpha_path(ntextures, 0);
    for (uint32_t i = 0; i < nmaterials; i++) {
        const lsg_point_material& m = materials[i];
        for (int32_t x : {m.diffuse_texture, m.emissive_texture, m.transmittance_texture})
            if (x < -1 || x >= (int32_t)ntextures)
                return fail(LSG_ERR_INVALID, "lsg_scene_set_point_materials: material %u names texture %d of %u", i, x, ntextures);
        if (m.flags & ~(LSG_PM_TRANSMITTANCE_IN_ALPHA | LSG_PM_KEYED | LSG_PM_INVERTED))
            return fail(LSG_ERR_INVALID, "lsg_scene_set_point_materials: material %u has unknown flags %#x", i, m.flags);
        if (m.transmittance_texture >= 0) {
            if (!(m.flags & LSG_PM_KEYED))
                return fail(LSG_ERR_INVALID, "lsg_scene_set_point_materials: material %u: unkeyed transmittance is not supported", i);
            const bool emits = (m.emit_sides & 3u)
                               && (m.emissive_texture >= 0 || m.emittance[0] != 0 || m.emittance[1] != 0 || m.emittance[2] != 0);
            if (emits)
                return fail(LSG_ERR_INVALID, "lsg_scene_set_point_materials: material %u is keyed and emits "
                            "(transparent texels would emit)", i);
            if (m.diffuse_texture >= 0 && m.transmittance_texture == m.diffuse_texture && (m.flags & LSG_PM_TRANSMITTANCE_IN_ALPHA))
                alpha_path[m.diffuse_texture] = 1;
        }
    }
    // texel array: raw, then linear and optimized-path copies where they differ
    std::vector<TexInfo> info(ntextures);
    size_t total = 0;
    for (uint32_t i = 0; i < ntextures; i++) {
        const lsg_texture& T = textures[i];
        if (!T.width || !T.height || !T.data || !texture_format_ok(T.format))
            return fail(LSG_ERR_INVALID, "lsg_scene_set_point_materials: texture %u: %ux%u, format %d", i, T.width, T.height, T.format);
        const size_t n = (size_t)T.width * T.height;
        const bool lin = T.linear || (T.scaled && gamma > 0);
        const bool bytes_rgb = T.format == LSG_FORMAT_RGB || T.format == LSG_FORMAT_RGBA;
        const bool opt = alpha_path[i] && (T.linear_stored || (gamma > 0 && !(T.scaled && !bytes_rgb)));
        total += n * (1 + (lin ? 1 : 0) + (opt ? 1 : 0));
    }
    if (total > 0xFFFFFFFFull) return fail(LSG_ERR_LIMIT, "lsg_scene_set_point_materials: %zu texels (at most 2^32 - 1)", total);
    std::vector<float4> texels;
    texels.reserve(total);
    const float inv = gamma > 0 ? 1 / gamma : 0.f;        // RRGammaScaler keeps 1/gamma in binary32
    float table[256];
    for (int i = 0; i < 256; i++) table[i] = gamma > 0 ? (float)std::pow((double)((float)i / 255), (double)inv) : 0.f;
    for (uint32_t i = 0; i < ntextures; i++) {
        const lsg_texture& T = textures[i];
        const size_t n = (size_t)T.width * T.height;
        TexInfo& I = info[i];
        I.width = T.width; I.height = T.height; I.pad[0] = I.pad[1] = I.pad[2] = 0;
        I.raw = (uint32_t)texels.size();
        for (size_t k = 0; k < n; k++) texels.push_back(decode(T, k));
        const bool bytes_rgb = T.format == LSG_FORMAT_RGB || T.format == LSG_FORMAT_RGBA;
        I.lin = I.raw;
        if (T.linear) {
            I.lin = (uint32_t)texels.size();
            for (size_t k = 0; k < n; k++)
                texels.push_back(make_float4(T.linear[4 * k], T.linear[4 * k + 1], T.linear[4 * k + 2], T.linear[4 * k + 3]));
        } else if (T.scaled && gamma > 0) {
            I.lin = (uint32_t)texels.size();
            const uint8_t* b = (const uint8_t*)T.data;
            const uint32_t step = T.format == LSG_FORMAT_RGB ? 3 : 4;
            for (size_t k = 0; k < n; k++) {
                const float4 r = texels[I.raw + k];
                if (bytes_rgb)   // RRColorSpace::getLinear's byte table (RRBufferInMemory.cpp:302-305)
                    texels.push_back(make_float4(table[b[step * k]], table[b[step * k + 1]], table[b[step * k + 2]], r.w));
                else
                    texels.push_back(make_float4(to_linear(r.x, inv), to_linear(r.y, inv), to_linear(r.z, inv), r.w));
            }
        }
        I.opt = I.raw;
        if (alpha_path[i] && T.linear_stored) {
            I.opt = (uint32_t)texels.size();
            for (size_t k = 0; k < n; k++)
                texels.push_back(make_float4(T.linear_stored[4 * k], T.linear_stored[4 * k + 1], T.linear_stored[4 * k + 2],
                                             T.linear_stored[4 * k + 3]));
        } else if (gamma > 0 && alpha_path[i]) {
            if (T.scaled && !bytes_rgb) I.opt = I.lin;
            else {
                I.opt = (uint32_t)texels.size();
                for (size_t k = 0; k < n; k++) {
                    const float4 r = texels[I.raw + k];
                    texels.push_back(make_float4(to_linear(r.x, inv), to_linear(r.y, inv), to_linear(r.z, inv), r.w));
                }
            }
        }
    }
    // upload into fresh buffers, then swap (a failed upload leaves the old materials)
    uint32_t* d_tri_mat = nullptr; float* d_tri_uv = nullptr; lsg_point_material* d_mats = nullptr;
    TexInfo* d_tex = nullptr; float4* d_texels = nullptr;
    auto release = [&]() {
        for (void* p : {(void*)d_tri_mat, (void*)d_tri_uv, (void*)d_mats, (void*)d_tex, (void*)d_texels})
            if (p) (void)hipFree(p);
    };
    hipError_t e = hipMalloc(&d_tri_mat, 4 * (size_t)(nt ? nt : 1));
    if (e == hipSuccess) e = hipMalloc(&d_tri_uv, 24 * (size_t)(nt ? nt : 1));
    if (e == hipSuccess) e = hipMalloc(&d_mats, sizeof(lsg_point_material) * (size_t)(nmaterials ? nmaterials : 1));
    if (e == hipSuccess) e = hipMalloc(&d_tex, sizeof(TexInfo) * (size_t)(ntextures ? ntextures : 1));
    if (e == hipSuccess) e = hipMalloc(&d_texels, sizeof(float4) * (texels.size() ? texels.size() : 1));
    if (e == hipSuccess && nt) e = hipMemcpy(d_tri_mat, tri_material, 4 * (size_t)nt, hipMemcpyHostToDevice);
    if (e == hipSuccess && nt) e = hipMemcpy(d_tri_uv, tri_uv, 24 * (size_t)nt, hipMemcpyHostToDevice);
    if (e == hipSuccess && nmaterials) e = hipMemcpy(d_mats, materials, sizeof(lsg_point_material) * nmaterials, hipMemcpyHostToDevice);
    if (e == hipSuccess && ntextures) e = hipMemcpy(d_tex, info.data(), sizeof(TexInfo) * ntextures, hipMemcpyHostToDevice);
    if (e == hipSuccess && texels.size()) e = hipMemcpy(d_texels, texels.data(), sizeof(float4) * texels.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        release();
        return hip_fail(e, "lsg_scene_set_point_materials", __LINE__);
    }
    free_point_materials(S);
    S.d_tri_mat = d_tri_mat; S.d_tri_uv = d_tri_uv; S.d_mats = d_mats; S.d_tex = d_tex; S.d_texels = d_texels;
    S.h_mats.assign(materials, materials + nmaterials);
    return LSG_OK;
}

} // extern "C"
