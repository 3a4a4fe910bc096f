// lsgpu_hash.cpp -- mesh hashes and cache file names (host code).
//
// The collider cache is keyed as the reference keys its BSP cache
// (IntersectBsp.h:213-291): the SHA-1 of the mesh as RRMesh::getHash feeds it
// (RRHash.cpp:48-77: every vertex as three binary32 coordinates, then every triangle as three
// 32-bit indices), turned into a file name by RRHash::getFileName (RRHash.cpp:198-226): the
// format version added to the hash's first little-endian word, the 160 bits written as 32
// letters of 5 bits each (least significant bit of the stream first, each letter's bits read
// from its highest), after the cache directory and before the extension.
//
// SHA-1 itself follows FIPS 180-4 section 6.1: 512-bit blocks, 80 rounds, big-endian words,
// the message padded with one bit, zeros and its 64-bit length.
#include "lsgpu_internal.h"
#include <cstring>
#include <filesystem>
#include <string>

namespace lsg {
namespace {

struct Sha1 {
    uint32_t h[5] = {0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u, 0xC3D2E1F0u};
    uint8_t block[64];
    uint32_t fill = 0;
    uint64_t bytes = 0;

    static uint32_t rol(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }

    void compress(const uint8_t* p)
    {
        uint32_t w[80];
        for (int i = 0; i < 16; i++) w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
        for (int i = 16; i < 80; i++) w[i] = rol(w[i - 3] ^ w[i - 8] ^ w[i - 14] ^ w[i - 16], 1);
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4];
        for (int i = 0; i < 80; i++) {
            uint32_t f, k;
            if (i < 20) { f = (b & c) | (~b & d); k = 0x5A827999u; }
            else if (i < 40) { f = b ^ c ^ d; k = 0x6ED9EBA1u; }
            else if (i < 60) { f = (b & c) | (b & d) | (c & d); k = 0x8F1BBCDCu; }
            else { f = b ^ c ^ d; k = 0xCA62C1D6u; }
            const uint32_t t = rol(a, 5) + f + e + k + w[i];
            e = d; d = c; c = rol(b, 30); b = a; a = t;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e;
    }

    void update(const void* data, size_t n)
    {
        const uint8_t* p = (const uint8_t*)data;
        bytes += n;
        while (n) {
            if (!fill && n >= 64) { compress(p); p += 64; n -= 64; continue; }
            const size_t k = std::min<size_t>(64 - fill, n);
            memcpy(block + fill, p, k);
            fill += (uint32_t)k; p += k; n -= k;
            if (fill == 64) { compress(block); fill = 0; }
        }
    }

    void final(uint8_t out[20])
    {
        const uint64_t bits = bytes * 8;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (fill != 56) update(&zero, 1);
        uint8_t len[8];
        for (int i = 0; i < 8; i++) len[i] = (uint8_t)(bits >> (56 - 8 * i));
        update(len, 8);
        for (int i = 0; i < 5; i++)
            for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(h[i] >> (24 - 8 * j));
    }
};

} // namespace

void sha1(const void* a, size_t na, const void* b, size_t nb, uint8_t out[20])
{
    Sha1 s;
    if (na) s.update(a, na);
    if (nb) s.update(b, nb);
    s.final(out);
}

std::string cache_file_name(const uint8_t hash[20], uint32_t version, const char* location, const char* extension)
{
    std::string name;
    if (!location) {
        std::error_code ec;
        const std::filesystem::path dir = std::filesystem::temp_directory_path(ec) / "Lightsprint";
        std::filesystem::create_directory(dir, ec);
        name = dir.string() + "/";
    } else {
        name = location;
    }
    uint8_t h[20];
    memcpy(h, hash, 20);
    uint32_t w = (uint32_t)h[0] | (uint32_t)h[1] << 8 | (uint32_t)h[2] << 16 | (uint32_t)h[3] << 24;
    w += version;
    for (int i = 0; i < 4; i++) h[i] = (uint8_t)(w >> (8 * i));
    static const char letters[] = "0123456789abcdefghijklmnopqrstuv";
    for (int i = 0; i < 32; i++) {
        unsigned v = 0;
        for (int b = 4; b >= 0; b--) {
            const unsigned bit = (unsigned)(5 * i + b);
            v = v * 2 + ((h[bit / 8] >> (bit % 8)) & 1u);
        }
        name += letters[v];
    }
    if (extension) name += extension;
    return name;
}

} // namespace lsg

extern "C" {

int lsg_mesh_hash(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri, uint8_t* hash)
{
    if (!hash || (nverts && !verts) || (ntri && !idx)) return lsg::fail(LSG_ERR_INVALID, "lsg_mesh_hash: null argument");
    lsg::sha1(verts, 12 * (size_t)nverts, idx, 12 * (size_t)ntri, hash);
    return LSG_OK;
}

int lsg_cache_file_name(const uint8_t* hash, uint32_t version, const char* location, const char* extension, char* out, uint32_t capacity)
{
    if (!hash || !out) return lsg::fail(LSG_ERR_INVALID, "lsg_cache_file_name: null argument");
    const std::string s = lsg::cache_file_name(hash, version, location, extension);
    if (s.size() + 1 > capacity) return lsg::fail(LSG_ERR_LIMIT, "lsg_cache_file_name: %zu characters do not fit %u", s.size(), capacity);
    memcpy(out, s.c_str(), s.size() + 1);
    return LSG_OK;
}

} // extern "C"
