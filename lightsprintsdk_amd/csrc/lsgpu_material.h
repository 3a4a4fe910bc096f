// lsgpu_material.h -- point materials of the gather (gfx950): textured and alpha-keyed
// materials read at a ray's hit point, as RRCollisionHandlerFinalGathering::collides
// (pathtracer.h:194-226) and RRObject::getPointMaterial (RRObject.cpp:198-303) read them.
//
// A triangle's material is read per point when the gather's quality exceeds the
// material's minimalQualityForPointMaterials: quality for hemisphere rays (nearest
// texel, gather.cpp:175-180), 2 x quality for shadow rays, which also interpolate texels
// bilinearly when quality / 10 exceeds it (gather.cpp:366-370).  Texels are looked up as
// RRBufferInMemory::getElementAtPosition does (RRBufferInMemory.cpp:372-408), at the
// material uv interpolated from the triangle's three corners by the hit's hitPoint2d.
//
// Supported: a diffuse texture, an emissive texture, and keyed transmittance (in the
// diffuse texture's alpha or a texture of its own; every texel fully transparent or fully
// opaque).  A transparent texel is passed through by hemisphere and shadow rays; an
// opaque one is hit with the texture's diffuse colour.  (The reference's hemisphere ray
// hits the transparent texel and continues as a new ray from it, pathtracer.cpp:350-372,
// with transmittance 1 and no diffuse reflectance there; passing through reaches the
// same surface behind it except within rayLengthMin of the texel.)
#pragma once
#include "lsgpu_internal.h"

namespace lsg {

// a texture's texels in the scene's texel array: as stored (getElement without a colour
// space), as getElement with the solver's colour space returns them (made linear when the
// texture is scaled), and as the optimized transmittance-in-alpha path makes the diffuse
// colour linear (toLinear of the stored value whatever the texture's scale); the three
// offsets coincide where the values do
struct TexInfo { uint32_t raw, lin, opt, width, height, pad[3]; };

struct PMView {
    const uint32_t* tri_mat;     // material per triangle, 0xFFFFFFFF = none; null = no point materials
    const float* tri_uv;         // 6 floats per triangle: material uv of corners 0, 1, 2
    const lsg_point_material* mats;
    const TexInfo* tex;
    const float4* texels;
};

// material uv at hitPoint2d (u, v): uv0 (1-u-v) + uv1 u + uv2 v (RRObject.cpp:205)
__device__ inline float2 pm_uv(const PMView& P, uint32_t tri, float u, float v)
{
    const float* c = P.tri_uv + 6 * (size_t)tri;
    const float w = 1 - u - v;
    return make_float2(c[0] * w + c[2] * u + c[4] * v, c[1] * w + c[3] * u + c[5] * v);
}

// RRBufferInMemory::getElementAtPosition of a 2D texture, without and with interpolation
__device__ inline float4 tex_nearest(const float4* d, uint32_t width, uint32_t height, float2 uv)
{
    const float px = fmodf(uv.x, 1.f) + 2, py = fmodf(uv.y, 1.f) + 2;
    const uint32_t x = (uint32_t)(px * (float)width) % width, y = (uint32_t)(py * (float)height) % height;
    return d[x + (size_t)y * width];
}

__device__ inline float4 tex_bilinear(const float4* d, uint32_t width, uint32_t height, float2 uv)
{
    const float px = (fmodf(uv.x, 1.f) + 2) * (float)width - 0.5f, py = (fmodf(uv.y, 1.f) + 2) * (float)height - 0.5f;
    const float rx = fmodf(px, 1.f), ry = fmodf(py, 1.f);
    const uint32_t cx = (uint32_t)px, cy = (uint32_t)py;
    const float4 e00 = d[cx % width + (size_t)(cy % height) * width];
    const float4 e01 = d[cx % width + (size_t)((cy + 1) % height) * width];
    const float4 e10 = d[(cx + 1) % width + (size_t)(cy % height) * width];
    const float4 e11 = d[(cx + 1) % width + (size_t)((cy + 1) % height) * width];
    const float w00 = (1 - rx) * (1 - ry), w01 = (1 - rx) * ry, w10 = rx * (1 - ry), w11 = rx * ry;
    return make_float4(e00.x * w00 + e01.x * w01 + e10.x * w10 + e11.x * w11, e00.y * w00 + e01.y * w01 + e10.y * w10 + e11.y * w11,
                       e00.z * w00 + e01.z * w01 + e10.z * w10 + e11.z * w11, e00.w * w00 + e01.w * w01 + e10.w * w10 + e11.w * w11);
}

// the optimized path of updatePointMaterial: transmittance in the diffuse map's alpha
__device__ inline bool pm_alpha_path(const lsg_point_material& m)
{
    return m.diffuse_texture >= 0 && m.transmittance_texture == m.diffuse_texture && (m.flags & LSG_PM_TRANSMITTANCE_IN_ALPHA);
}

// keyed transmittance at uv: true = the texel is fully transparent (RRObject.cpp:225-250,
// 261-275; the transmittance is read without a colour space).  The optimized path keys
// the scalar 1 - alpha, the generic path the rgb average.
__device__ inline bool pm_transparent(const PMView& P, const lsg_point_material& m, float2 uv, bool interp)
{
    if (m.transmittance_texture < 0) return false;
    const TexInfo T = P.tex[m.transmittance_texture];
    const float4* d = P.texels + T.raw;
    const float4 c = interp ? tex_bilinear(d, T.width, T.height, uv) : tex_nearest(d, T.width, T.height, uv);
    if (pm_alpha_path(m)) {
        float t = 1 - c.w;
        if (m.flags & LSG_PM_INVERTED) t = 1 - t;
        return t >= m.threshold;
    }
    float tx, ty, tz;
    if (m.flags & LSG_PM_TRANSMITTANCE_IN_ALPHA) { tx = ty = tz = 1 - c.w; }
    else { tx = c.x; ty = c.y; tz = c.z; }
    if (m.flags & LSG_PM_INVERTED) { tx = 1 - tx; ty = 1 - ty; tz = 1 - tz; }
    return (tx + ty + tz) * 0.33333333333333f >= m.threshold;
}

// diffuse colour of an opaque texel, linear (the opacity factor 1 - transmittance is 1)
__device__ inline float4 pm_diffuse(const PMView& P, const lsg_point_material& m, float2 uv)
{
    const TexInfo T = P.tex[m.diffuse_texture];
    return tex_nearest(P.texels + (pm_alpha_path(m) ? T.opt : T.lin), T.width, T.height, uv);
}

__device__ inline float4 pm_emissive(const PMView& P, const lsg_point_material& m, float2 uv)
{
    const TexInfo T = P.tex[m.emissive_texture];
    return tex_nearest(P.texels + T.lin, T.width, T.height, uv);
}

} // namespace lsg
