// writeback_test.cpp -- runs the drop-in solver's writeback (integration/lsg_writeback.h)
// with the reference's own colour space (RRColorSpace::create_sRGB, compiled from
// RRColorSpace.cpp).  TEST INFRASTRUCTURE ONLY, built by `make -C oracle integration`
// into oracle/_ref/.
//
// usage: writeback_test
//   For each case (target scaled or not, solver colour space sRGB or none, lightmap or
//   bent-normal layer) writes the same five work texels with the background
//   (0.2, 0.5, 0.8, 0.7) and prints one line per texel:
//     case scaled space bent texel r g b a
#include "lsg_writeback.h"
#include <cstdio>

using namespace rr;

int main()
{
	const float work[5][4] = {
		{0.25f, 0.5f, 1.0f, 1.0f},     // covered texel
		{0.0f, 0.0f, 0.0f, 0.0f},      // background (alpha 0)
		{0.3f, 0.1f, 0.05f, 0.001f},   // texel lightmapGrow filled
		{0.5f, 0.5f, 0.5f, -1.0f},     // background (alpha below 0)
		{2.0f, 0.0f, 0.125f, 0.5f},    // partly covered texel, HDR value
	};
	const RRVec4 background(0.2f, 0.5f, 0.8f, 0.7f);
	RRColorSpace* srgb = RRColorSpace::create_sRGB();
	int c = 0;
	for (int scaled = 0; scaled < 2; scaled++)
		for (int space = 0; space < 2; space++)
			for (int bent = 0; bent < 2; bent++, c++)
			{
				const RRColorSpace* cs = space ? srgb : nullptr;
				rr_hip::writebackLayer(&work[0][0], 5, background, scaled != 0, cs, bent ? nullptr : cs,
				               [&](unsigned i, const RRVec4& v) {
					               std::printf("%d %d %d %d %u %.9g %.9g %.9g %.9g\n", c, scaled, space, bent, i,
					                           v[0], v[1], v[2], v[3]);
				               });
			}
	delete srgb;
	return 0;
}
