// lsgpu_encode.hip -- a filtered lightmap written into a consumer buffer's format on the GPU
// (gfx950), so a caller downloads 1-16 bytes per texel instead of the 16-byte work texel.
//
// The conversion is the last step of RRSolver::updateLightmap for a pixel buffer
// (lightmap.cpp:866-876) as the reference performs it through its buffer interface:
//   lightmapFillBackground (RRBuffer.cpp:964-975): a texel with alpha <= 0 takes the
//     background colour, made linear first (toLinear) when the target is scaled and the
//     solver has a colour space;
//   copyElementsTo (RRBuffer.cpp:105-131) from the linear work buffer: rgb to the target's
//     colour space (fromLinear) when the target is scaled;
//   RRBufferInMemory::setElement (RRBufferInMemory.cpp:229-285): integer formats take
//     RR_FLOAT2BYTE(x) = clamp(int(x * 256), 0, 255) (RRMath.h:78), luminance the rgb average
//     (x + y + z) * 0.33333333333333f, float formats the values.
// The colour space is the reference's sRGB approximation (RRGammaScaler, RRColorSpace.cpp:
// 22-91): x -> sign(x) |x|^gamma, evaluated as the reference's build does, in binary64
// rounded to binary32.  int(x * 256) follows x86-64's conversion, which gives INT_MIN (so 0
// after the clamp) for a product outside the int range or NaN.
#include "lsgpu_internal.h"
#include <cmath>

namespace lsg {
namespace {

__device__ inline float gamma_pow(float x, float g)
{
    return x >= 0.f ? (float)pow((double)x, (double)g) : -(float)pow((double)-x, (double)g);
}

__device__ inline uint32_t to_byte(float f)
{
    const float p = f * 256.f;
    const int v = (p >= -2147483648.f && p < 2147483648.f) ? (int)p : (int)0x80000000u;
    return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
}

struct EncodeArgs {
    const float4* rgba;
    uint32_t n;
    int32_t format;
    int32_t convert;        // rgb to the target's colour space
    float gamma;
    float4 bg;              // background, already as the fill uses it
    void* out;
};

__device__ inline float4 texel(const EncodeArgs& A, uint32_t i)
{
    float4 c = A.rgba[i];
    if (c.w <= 0.f) c = A.bg;
    if (A.convert) { c.x = gamma_pow(c.x, A.gamma); c.y = gamma_pow(c.y, A.gamma); c.z = gamma_pow(c.z, A.gamma); }
    return c;
}

// four texels per lane, so the byte formats store whole 32-bit words
__global__ __launch_bounds__(256) void k_encode(EncodeArgs A)
{
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    const uint32_t i0 = 4 * q;
    if (i0 >= A.n) return;
    const uint32_t m = A.n - i0 < 4 ? A.n - i0 : 4;
    float4 c[4];
    for (uint32_t k = 0; k < m; k++) c[k] = texel(A, i0 + k);
    switch (A.format) {
    case LSG_FORMAT_RGB:
    case LSG_FORMAT_BGR: {
        uint8_t b[12];
        for (uint32_t k = 0; k < m; k++) {
            const uint32_t r = to_byte(c[k].x), g = to_byte(c[k].y), bl = to_byte(c[k].z);
            const bool bgr = A.format == LSG_FORMAT_BGR;
            b[3 * k] = (uint8_t)(bgr ? bl : r); b[3 * k + 1] = (uint8_t)g; b[3 * k + 2] = (uint8_t)(bgr ? r : bl);
        }
        uint8_t* o = (uint8_t*)A.out + 3 * (size_t)i0;
        if (m == 4) {
            uint32_t w[3];
            for (int k = 0; k < 3; k++) w[k] = (uint32_t)b[4 * k] | (uint32_t)b[4 * k + 1] << 8 | (uint32_t)b[4 * k + 2] << 16 | (uint32_t)b[4 * k + 3] << 24;
            uint32_t* ow = (uint32_t*)o;
            ow[0] = w[0]; ow[1] = w[1]; ow[2] = w[2];
        } else {
            for (uint32_t k = 0; k < 3 * m; k++) o[k] = b[k];
        }
        break;
    }
    case LSG_FORMAT_RGBA: {
        uint32_t* o = (uint32_t*)A.out + i0;
        for (uint32_t k = 0; k < m; k++)
            o[k] = to_byte(c[k].x) | to_byte(c[k].y) << 8 | to_byte(c[k].z) << 16 | to_byte(c[k].w) << 24;
        break;
    }
    case LSG_FORMAT_RGBF: {
        float* o = (float*)A.out + 3 * (size_t)i0;
        for (uint32_t k = 0; k < m; k++) { o[3 * k] = c[k].x; o[3 * k + 1] = c[k].y; o[3 * k + 2] = c[k].z; }
        break;
    }
    case LSG_FORMAT_RGBAF: {
        float4* o = (float4*)A.out + i0;
        for (uint32_t k = 0; k < m; k++) o[k] = c[k];
        break;
    }
    case LSG_FORMAT_LUMINANCE: {
        uint8_t* o = (uint8_t*)A.out + i0;
        uint32_t w = 0;
        for (uint32_t k = 0; k < m; k++) w |= to_byte((c[k].x + c[k].y + c[k].z) * 0.33333333333333f) << (8 * k);
        if (m == 4) *(uint32_t*)o = w;
        else for (uint32_t k = 0; k < m; k++) o[k] = (uint8_t)(w >> (8 * k));
        break;
    }
    case LSG_FORMAT_LUMINANCEF: {
        float* o = (float*)A.out + i0;
        for (uint32_t k = 0; k < m; k++) o[k] = (c[k].x + c[k].y + c[k].z) * 0.33333333333333f;
        break;
    }
    default:
        break;
    }
}

} // namespace
} // namespace lsg

using namespace lsg;

extern "C" {

uint32_t lsg_format_bytes(int32_t format)
{
    switch (format) {
    case LSG_FORMAT_RGB: case LSG_FORMAT_BGR: return 3;
    case LSG_FORMAT_RGBA: return 4;
    case LSG_FORMAT_RGBF: return 12;
    case LSG_FORMAT_RGBAF: return 16;
    case LSG_FORMAT_LUMINANCE: return 1;
    case LSG_FORMAT_LUMINANCEF: return 4;
    default: return 0;
    }
}

int lsg_lightmap_encode(const float* rgba, uint32_t ntexels, int32_t format, int32_t scaled, float solver_gamma, float copy_gamma,
                        const float* background, void* out, void* stream)
{
    if (!rgba || !out || !background) return fail(LSG_ERR_INVALID, "lsg_lightmap_encode: null argument");
    if (!lsg_format_bytes(format)) return fail(LSG_ERR_INVALID, "lsg_lightmap_encode: format %d has no per-texel encoding", format);
    const bool byte3 = format == LSG_FORMAT_RGB || format == LSG_FORMAT_BGR, word = format == LSG_FORMAT_RGBA;
    if ((byte3 || word || format == LSG_FORMAT_LUMINANCE) && ((uintptr_t)out & 3u))
        return fail(LSG_ERR_INVALID, "lsg_lightmap_encode: output must be 4-byte aligned");
    if (!ntexels) return LSG_OK;
    EncodeArgs A;
    A.rgba = (const float4*)rgba; A.n = ntexels; A.format = format; A.out = out;
    A.convert = scaled && copy_gamma > 0.f;
    A.gamma = copy_gamma;
    float bg[4] = {background[0], background[1], background[2], background[3]};
    if (scaled && solver_gamma > 0.f) {
        const float inv = 1 / solver_gamma;          // RRGammaScaler keeps 1/gamma in binary32
        for (int k = 0; k < 3; k++)
            bg[k] = bg[k] >= 0 ? (float)std::pow((double)bg[k], (double)inv) : -(float)std::pow((double)-bg[k], (double)inv);
    }
    A.bg = make_float4(bg[0], bg[1], bg[2], bg[3]);
    const uint32_t lanes = (ntexels + 3) / 4;
    k_encode<<<(lanes + 255) / 256, 256, 0, (hipStream_t)stream>>>(A);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

} // extern "C"
