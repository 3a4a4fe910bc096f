// lsgpu_raster.cpp -- texel rasterisation of one object's lightmap (host C++).
//
// Restates enumerateTexelsPartial steps 1-2 (RRSolver/lightmap.cpp:96-337): for
// every triangle's unwrap, every texel of its bounding box is clipped against the
// three triangle edges in triangle space (Sutherland-Hodgman), the polygon is fan
// triangulated into SubTexels, and texelFlags record edge/quadrant coverage for
// the filters.  Float arithmetic follows the reference's operation order.
#include "lsgpu.h"
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>

namespace lsg {
int fail(int code, const char* fmt, ...);
}

namespace {

struct V2 { float x, y; };

inline V2 sub2(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline float len2(V2 a) { return sqrtf(a.x * a.x + a.y * a.y); }

// lightmap.cpp:35-43 getArea (Heron)
inline float get_area(V2 v0, V2 v1, V2 v2)
{
    float a = len2(sub2(v1, v0));
    float b = len2(sub2(v2, v1));
    float c = len2(sub2(v0, v2));
    float s = (a + b + c) * 0.5f;
    return sqrtf(s * (s - a) * (s - b) * (s - c));
}

enum { INSIDE, EDGE, OUTSIDE };
enum { QUADRANT_X0Y0 = 1, QUADRANT_X1Y0 = 2, QUADRANT_X0Y1 = 4, QUADRANT_X1Y1 = 8,
       EDGE_X0 = 16, EDGE_X1 = 32, EDGE_Y0 = 64, EDGE_Y1 = 128 };   // gather.h:68-78

struct Sub { uint32_t tri; V2 uv[3]; float area; };

} // namespace

extern "C" int lsg_rasterize_texels(const float* tri_uv, const uint32_t* tri_ids, uint32_t ntri,
                                    uint32_t width, uint32_t height, uint32_t object_id,
                                    lsg_texel* texels, uint32_t* ntexels, lsg_subtexel* subs, uint32_t* nsubs,
                                    uint32_t out_base)
{
    if (!ntexels || !nsubs || (ntri && (!tri_uv || !tri_ids))) return lsg::fail(LSG_ERR_INVALID, "lsg_rasterize_texels: null argument");
    if (!width || !height || (uint64_t)width * height > (1ull << 28)) return lsg::fail(LSG_ERR_INVALID, "lsg_rasterize_texels: bad map size %ux%u", width, height);
    const uint32_t W = width, H = height;
    std::vector<std::vector<Sub>> texel_subs((size_t)W * H);
    std::vector<uint8_t> tflags((size_t)W * H, 0);
    for (uint32_t t = 0; t < ntri; t++) {
        const float* u = tri_uv + 6 * (size_t)t;
        V2 uv0{u[0], u[1]}, uv1{u[2], u[3]}, uv2{u[4], u[5]};
        // lightmap.cpp:147-168 bounding box in texels (whole map = one rectangle)
        float xmin = W * std::min(uv0.x, std::min(uv1.x, uv2.x));
        float xmax = W * std::max(uv0.x, std::max(uv1.x, uv2.x));
        float ymin = H * std::min(uv0.y, std::min(uv1.y, uv2.y));
        float ymax = H * std::max(uv0.y, std::max(uv1.y, uv2.y));
        if (!(std::isfinite(xmin) && std::isfinite(xmax) && std::isfinite(ymin) && std::isfinite(ymax))) continue;
        auto rmax = [](float a, float b) { return a > b ? a : b; };
        auto clamped = [](float a, float lo, float hi) { return a < lo ? lo : (a > hi ? hi : a); };
        uint32_t xminu = (uint32_t)rmax(xmin, 0.f);
        uint32_t xmaxu = (uint32_t)clamped(xmax + 1, 0.f, (float)W);
        uint32_t yminu = (uint32_t)rmax(ymin, 0.f);
        uint32_t ymaxu = (uint32_t)clamped(ymax + 1, 0.f, (float)H);
        if (yminu >= ymaxu || xminu >= xmaxu) continue;
        // lightmap.cpp:170-182 map space <-> triangle space
        float m[3][3] = { { uv1.x - uv0.x, uv2.x - uv0.x, uv0.x }, { uv1.y - uv0.y, uv2.y - uv0.y, uv0.y }, { 0, 0, 1 } };
        float det = m[0][0] * m[1][1] * m[2][2] + m[0][1] * m[1][2] * m[2][0] + m[0][2] * m[1][0] * m[2][1]
                  - m[0][0] * m[1][2] * m[2][1] - m[0][1] * m[1][0] * m[2][2] - m[0][2] * m[1][1] * m[2][0];
        if (!det) continue;
        float invdet = 1 / det;
        float inv[2][3] = {
            { (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * invdet / W, (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * invdet / H, (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * invdet },
            { (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * invdet / W, (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * invdet / H, (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * invdet },
        };
        const float tri_area_map = get_area(uv0, uv1, uv2);
        auto to_tri = [&](uint32_t x, uint32_t y) { return V2{ x * inv[0][0] + y * inv[0][1] + inv[0][2], x * inv[1][0] + y * inv[1][1] + inv[1][2] }; };
        for (uint32_t y = yminu; y < ymaxu; y++)
            for (uint32_t x = xminu; x < xmaxu; x++) {
                unsigned poly_size = 4;
                V2 poly[7] = { to_tri(x, y), to_tri(x + 1, y), to_tri(x + 1, y + 1), to_tri(x, y + 1) };
                for (int side = 0; side < 3; side++) {
                    const float line[3] = { side == 0 ? 0.f : (side == 1 ? -1.f : 1.f), side == 0 ? 1.f : (side == 1 ? -1.f : 0.f), side == 1 ? 1.f : 0.f };
                    int inside[7];
                    unsigned nin = 0, nout = 0;
                    for (unsigned i = 0; i < poly_size; i++) {
                        float dist = line[0] * poly[i].x + line[1] * poly[i].y + line[2];
                        if (dist < 0) { inside[i] = OUTSIDE; nout++; }
                        else if (dist > 0) { inside[i] = INSIDE; nin++; }
                        else inside[i] = EDGE;
                    }
                    if (!nout) continue;
                    if (!nin) { poly_size = 0; break; }
                    unsigned first = 1;
                    while (inside[(first - 1) % poly_size] == INSIDE || inside[first % poly_size] != INSIDE) first++;
                    V2 orig[7];
                    memcpy(orig, poly, sizeof poly);
                    unsigned src = first, dst = 0;
                    while (inside[src % poly_size] == INSIDE) poly[dst++] = orig[src++ % poly_size];
                    auto cut = [&](V2 a, V2 b) {
                        float da = line[0] * a.x + line[1] * a.y + line[2];
                        float den = line[0] * (b.x - a.x) + line[1] * (b.y - a.y);
                        return V2{ a.x - (b.x - a.x) * da / den, a.y - (b.y - a.y) * da / den };
                    };
                    poly[dst++] = cut(orig[(src - 1) % poly_size], orig[src % poly_size]);
                    poly[dst++] = cut(orig[(first - 1) % poly_size], orig[first % poly_size]);
                    poly_size = dst;
                }
                if (!poly_size) continue;
                const size_t ti = (size_t)x + (size_t)y * W;
                uint8_t& fl = tflags[ti];
                // lightmap.cpp:276-314 texel flags from the polygon in texel space
                auto to_map = [&](V2 p) { return V2{ p.x * m[0][0] + p.y * m[0][1] + m[0][2], p.x * m[1][0] + p.y * m[1][1] + m[1][2] }; };
                auto to_texel = [&](V2 p) { return V2{ p.x * W - x - 0.5f, p.y * H - y - 0.5f }; };
                V2 old = to_texel(to_map(poly[poly_size - 1]));
                for (unsigned i = 0; i < poly_size; i++) {
                    V2 vt = to_texel(to_map(poly[i]));
                    const float E = 0.03f;   // ERROR_IN_TEXELS
                    if (vt.x < -0.5f + E) fl |= EDGE_X0;
                    if (vt.x > 0.5f - E) fl |= EDGE_X1;
                    if (vt.y < -0.5f + E) fl |= EDGE_Y0;
                    if (vt.y > 0.5f - E) fl |= EDGE_Y1;
                    if (vt.x < 0 && vt.y < 0) fl |= QUADRANT_X0Y0;
                    if (vt.x > 0 && vt.y < 0) fl |= QUADRANT_X1Y0;
                    if (vt.x < 0 && vt.y > 0) fl |= QUADRANT_X0Y1;
                    if (vt.x > 0 && vt.y > 0) fl |= QUADRANT_X1Y1;
                    if ((old.x < 0) != (vt.x < 0) && (old.y < 0) != (vt.y < 0)) {
                        float a = (vt.y - old.y) / (vt.x - old.x);
                        float b = vt.y - a * vt.x;
                        fl |= (b < 0) ? (QUADRANT_X0Y0 | QUADRANT_X1Y0) : (QUADRANT_X0Y1 | QUADRANT_X1Y1);
                        fl |= (-b / a < 0) ? (QUADRANT_X0Y0 | QUADRANT_X0Y1) : (QUADRANT_X1Y0 | QUADRANT_X1Y1);
                    }
                    old = vt;
                }
                // lightmap.cpp:317-331 fan triangulation into subtexels
                Sub s;
                s.tri = tri_ids[t];
                s.uv[0] = poly[0];
                for (unsigned i = 0; i + 2 < poly_size; i++) {
                    s.uv[1] = poly[i + 1];
                    s.uv[2] = poly[i + 2];
                    s.area = get_area(s.uv[0], s.uv[1], s.uv[2]) * tri_area_map;
                    if (std::isfinite(s.area)) texel_subs[ti].push_back(s);
                }
            }
    }
    uint32_t nt = 0, ns = 0;
    for (size_t ti = 0; ti < texel_subs.size(); ti++)
        if (!texel_subs[ti].empty()) { nt++; ns += (uint32_t)texel_subs[ti].size(); }
    const uint32_t cap_t = *ntexels, cap_s = *nsubs;
    *ntexels = nt; *nsubs = ns;
    if (!texels && !subs) return LSG_OK;
    if (!texels || !subs || cap_t < nt || cap_s < ns)
        return lsg::fail(LSG_ERR_LIMIT, "lsg_rasterize_texels: need %u texels / %u subtexels", nt, ns);
    uint32_t it = 0, is = 0;
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const size_t ti = (size_t)x + (size_t)y * W;
            const auto& v = texel_subs[ti];
            if (v.empty()) continue;
            if (v.size() > 0xffff) return lsg::fail(LSG_ERR_LIMIT, "texel %u,%u has %zu subtexels", x, y, v.size());
            lsg_texel& T = texels[it++];
            T.out_index = out_base + (uint32_t)ti;
            T.sub_begin = is;
            T.sub_count_flags = (uint32_t)v.size() | ((uint32_t)tflags[ti] << 16);
            T.key = object_id * 0x9E3779B1u + (uint32_t)ti;
            for (const Sub& s : v) {
                lsg_subtexel& o = subs[is++];
                o.triangle = s.tri;
                for (int k = 0; k < 3; k++) { o.uv[2 * k] = s.uv[k].x; o.uv[2 * k + 1] = s.uv[k].y; }
                o.area = s.area;
            }
        }
    return LSG_OK;
}
