// encode_ref.cpp -- a lightmap written into a buffer format the way the reference's buffer
// interface does it: the writeback of updateLightmap (integration/lsg_writeback.h, pinned by
// tests/test_writeback_cpu.py) with the reference's own sRGB colour space
// (RRColorSpace::create_sRGB, compiled from RRColorSpace.cpp), then the per-format element
// store of RRBufferInMemory::setElement (RRBufferInMemory.cpp:245-283) with the reference's
// RR_FLOAT2BYTE and RRVec3::avg (RRMath.h).  TEST INFRASTRUCTURE ONLY, built by
// `make -C oracle integration` into oracle/_ref/.
//
// usage: encode_ref IN.bin OUT.bin FORMAT SCALED BENT
//   IN: uint32 n; float rgba[n*4]; float background[4]
//   OUT: the encoded texels (RRBufferFormat FORMAT; SCALED: the target is scaled; BENT: the
//        layer is bent normals, copied without a colour space)
#include "lsg_writeback.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace rr;

int main(int argc, char** argv)
{
	if (argc != 6) { fprintf(stderr, "usage: encode_ref IN.bin OUT.bin FORMAT SCALED BENT\n"); return 2; }
	FILE* f = fopen(argv[1], "rb");
	if (!f) return 2;
	unsigned n;
	if (fread(&n, 4, 1, f) != 1) return 2;
	std::vector<float> rgba(4 * (size_t)n);
	float bg[4];
	if (fread(rgba.data(), 4, rgba.size(), f) != rgba.size() || fread(bg, 4, 4, f) != 4) return 2;
	fclose(f);
	const int format = atoi(argv[3]), scaled = atoi(argv[4]), bent = atoi(argv[5]);
	RRColorSpace* srgb = RRColorSpace::create_sRGB();
	std::vector<unsigned char> out;
	auto put = [&](const void* p, size_t k) { out.insert(out.end(), (const unsigned char*)p, (const unsigned char*)p + k); };
	rr_hip::writebackLayer(rgba.data(), n, RRVec4(bg[0], bg[1], bg[2], bg[3]), scaled != 0, srgb, bent ? nullptr : srgb,
		[&](unsigned, const RRVec4& e) {
			unsigned char b[4];
			switch (format)
			{
				case BF_RGB: for (int c = 0; c < 3; c++) b[c] = (unsigned char)RR_FLOAT2BYTE(e[c]); put(b, 3); break;
				case BF_BGR: for (int c = 0; c < 3; c++) b[c] = (unsigned char)RR_FLOAT2BYTE(e[2 - c]); put(b, 3); break;
				case BF_RGBA: for (int c = 0; c < 4; c++) b[c] = (unsigned char)RR_FLOAT2BYTE(e[c]); put(b, 4); break;
				case BF_RGBF: put(&e[0], 12); break;
				case BF_RGBAF: put(&e[0], 16); break;
				case BF_LUMINANCE: b[0] = (unsigned char)RR_FLOAT2BYTE(e.RRVec3::avg()); put(b, 1); break;
				case BF_LUMINANCEF: { float a = e.RRVec3::avg(); put(&a, 4); break; }
				default: break;
			}
		});
	delete srgb;
	FILE* o = fopen(argv[2], "wb");
	if (!o) return 2;
	fwrite(out.data(), 1, out.size(), o);
	fclose(o);
	return 0;
}
