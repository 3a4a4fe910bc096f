// hash_ref.cpp -- the reference's own mesh hash and cache file name (RRMesh::getHash,
// RRHash::getFileName, RRHash.cpp, compiled from its sources) for a mesh given in a file.
// TEST INFRASTRUCTURE ONLY, built by `make -C oracle ref` into oracle/_ref/.
//
// usage: hash_ref IN.bin VERSION LOCATION EXTENSION
//   IN: uint32 nverts, ntri; float verts[nverts*3]; uint32 idx[ntri*3]
//   prints the 40 hex digits of the hash, then the file name, one per line
#include "Lightsprint/RRMesh.h"
#include "Lightsprint/RRHash.h"
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace rr;

namespace {

class ArrayMesh : public RRMesh
{
public:
	ArrayMesh(const float* v, unsigned nv, const unsigned* i, unsigned nt) : verts(v), nverts(nv), idx(i), ntri(nt) {}
	unsigned getNumVertices() const override { return nverts; }
	void getVertex(unsigned v, Vertex& out) const override { out = RRVec3(verts[3*v], verts[3*v+1], verts[3*v+2]); }
	unsigned getNumTriangles() const override { return ntri; }
	void getTriangle(unsigned t, Triangle& out) const override { out[0] = idx[3*t]; out[1] = idx[3*t+1]; out[2] = idx[3*t+2]; }
private:
	const float* verts; unsigned nverts; const unsigned* idx; unsigned ntri;
};

} // namespace

int main(int argc, char** argv)
{
	if (argc != 5) { fprintf(stderr, "usage: hash_ref IN.bin VERSION LOCATION EXTENSION\n"); return 2; }
	FILE* f = fopen(argv[1], "rb");
	if (!f) return 2;
	unsigned hdr[2];
	if (fread(hdr, 4, 2, f) != 2) return 2;
	std::vector<float> v(3 * (size_t)hdr[0]);
	std::vector<unsigned> idx(3 * (size_t)hdr[1]);
	if (fread(v.data(), 4, v.size(), f) != v.size() || fread(idx.data(), 4, idx.size(), f) != idx.size()) return 2;
	fclose(f);
	ArrayMesh mesh(v.data(), hdr[0], idx.data(), hdr[1]);
	RRHash h = mesh.getHash();
	for (int i = 0; i < 20; i++) printf("%02x", h.value[i]);
	printf("\n%s\n", h.getFileName((unsigned)atoi(argv[2]), argv[3], argv[4]).c_str());
	return 0;
}
