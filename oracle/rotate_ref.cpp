// rotate_ref.cpp -- the reference's environment rotation, RRMatrix3x4::rotationByYawPitchRoll
// ((yaw, 0, 0)) then transformDirection (RRMath.cpp:212-219, 457-470), compiled from its
// sources, on directions given in a file.  TEST INFRASTRUCTURE ONLY, built by
// `make -C oracle ref` into oracle/_ref/.
//
// usage: rotate_ref IN.bin OUT.bin
//   IN:  uint32 n; then n records of float yaw, float d[3]
//   OUT: n records of float r[3]
#include "Lightsprint/RRMath.h"
#include <cstdio>
#include <vector>

using namespace rr;

int main(int argc, char** argv)
{
	if (argc != 3) { fprintf(stderr, "usage: rotate_ref IN.bin OUT.bin\n"); return 2; }
	FILE* f = fopen(argv[1], "rb");
	if (!f) return 2;
	unsigned n = 0;
	if (fread(&n, 4, 1, f) != 1) return 2;
	std::vector<float> in(4 * (size_t)n), out(3 * (size_t)n);
	if (fread(in.data(), 4, in.size(), f) != in.size()) return 2;
	fclose(f);
	for (unsigned i = 0; i < n; i++)
	{
		const RRMatrix3x4 m = RRMatrix3x4::rotationByYawPitchRoll(RRVec3(in[4 * i], 0, 0));
		RRVec3 d(in[4 * i + 1], in[4 * i + 2], in[4 * i + 3]);
		m.transformDirection(d);
		for (int c = 0; c < 3; c++) out[3 * (size_t)i + c] = d[c];
	}
	FILE* o = fopen(argv[2], "wb");
	if (!o) return 2;
	fwrite(out.data(), 4, out.size(), o);
	fclose(o);
	return 0;
}
