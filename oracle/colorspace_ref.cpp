// colorspace_ref.cpp -- the reference's sRGB colour space (RRColorSpace::create_sRGB, compiled
// from RRColorSpace.cpp) applied to texel values the way RRBufferInMemory::getElement applies
// it: getLinear of each byte value 0..255 (the byte table of scaled RGB/RGBA textures), then
// toLinear of the given floats.  TEST INFRASTRUCTURE ONLY, built by `make -C oracle
// integration` into oracle/_ref/.
//
// usage: colorspace_ref IN.bin OUT.bin GAMMA
//   IN: uint32 n; float x[n]
//   OUT: float byte_linear[256]; float to_linear[n]
#include "Lightsprint/RRBuffer.h"
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace rr;

int main(int argc, char** argv)
{
	if (argc != 4) { fprintf(stderr, "usage: colorspace_ref IN.bin OUT.bin GAMMA\n"); return 2; }
	FILE* f = fopen(argv[1], "rb");
	if (!f) return 2;
	unsigned n;
	if (fread(&n, 4, 1, f) != 1) return 2;
	std::vector<float> x(n);
	if (fread(x.data(), 4, n, f) != n) return 2;
	fclose(f);
	RRColorSpace* cs = RRColorSpace::create_sRGB((RRReal)atof(argv[3]));
	std::vector<float> out;
	for (unsigned i = 0; i < 256; i++) {
		const unsigned char rgb[3] = {(unsigned char)i, (unsigned char)i, (unsigned char)i};
		out.push_back(cs->getLinear(rgb)[0]);
	}
	for (unsigned i = 0; i < n; i++) {
		RRReal a = x[i];
		cs->toLinear(a);
		out.push_back(a);
	}
	delete cs;
	FILE* o = fopen(argv[2], "wb");
	if (!o) return 2;
	fwrite(out.data(), 4, out.size(), o);
	fclose(o);
	return 0;
}
