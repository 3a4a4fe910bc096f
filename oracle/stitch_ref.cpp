// stitch_ref.cpp -- runs the reference's own vertex stitching,
// RRMesh::createOptimizedVertices (RRMesh.cpp:427-438, RRMeshLessVertices.h), compiled
// from its sources, on a mesh given in a file.  TEST INFRASTRUCTURE ONLY, built by
// `make -C oracle ref` into oracle/_ref/.
//
// usage: stitch_ref IN.bin OUT.bin
//   IN:  uint32 magic 0x48435453, nverts, ntri; float weld, angle; float verts[nverts*3];
//        uint32 idx[ntri*3]; float corner_normals[ntri*9]
//   OUT: uint32 numVertices of the stitched mesh, uint32 corners[ntri*3] (stitched vertex of
//        every triangle corner, RRMesh::getTriangle), float positions[numVertices*3]
#include "Lightsprint/RRMesh.h"
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace rr;

namespace {

class NormalMesh : public RRMesh
{
public:
	NormalMesh(const float* v, unsigned nv, const unsigned* i, const float* n, unsigned nt) : verts(v), nverts(nv), idx(i), nrm(n), ntri(nt) {}
	unsigned getNumVertices() const override { return nverts; }
	void getVertex(unsigned v, Vertex& out) const override { out = RRVec3(verts[3*v], verts[3*v+1], verts[3*v+2]); }
	unsigned getNumTriangles() const override { return ntri; }
	void getTriangle(unsigned t, Triangle& out) const override { out[0] = idx[3*t]; out[1] = idx[3*t+1]; out[2] = idx[3*t+2]; }
	void getTriangleNormals(unsigned t, TriangleNormals& out) const override
	{
		for (int k = 0; k < 3; k++)
		{
			const float* n = nrm + 9 * (size_t)t + 3 * k;
			out.vertex[k].normal = RRVec3(n[0], n[1], n[2]);
			out.vertex[k].buildBasisFromNormal();
		}
	}
private:
	const float* verts; unsigned nverts; const unsigned* idx; const float* nrm; unsigned ntri;
};

template <class T> bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

} // namespace

int main(int argc, char** argv)
{
	if (argc != 3) { fprintf(stderr, "usage: stitch_ref IN.bin OUT.bin\n"); return 2; }
	FILE* f = fopen(argv[1], "rb");
	if (!f) return 2;
	unsigned hdr[3]; float wa[2];
	if (!rd(f, hdr, 3) || hdr[0] != 0x48435453u || !rd(f, wa, 2)) return 2;
	const unsigned nv = hdr[1], nt = hdr[2];
	std::vector<float> v(3 * (size_t)nv), n(9 * (size_t)nt);
	std::vector<unsigned> idx(3 * (size_t)nt);
	if (!rd(f, v.data(), v.size()) || !rd(f, idx.data(), idx.size()) || !rd(f, n.data(), n.size())) return 2;
	fclose(f);
	NormalMesh mesh(v.data(), nv, idx.data(), n.data(), nt);
	const RRMesh* st = mesh.createOptimizedVertices(wa[0], wa[1], 0, nullptr);
	const unsigned ns = st->getNumVertices();
	std::vector<unsigned> corners(3 * (size_t)nt);
	for (unsigned t = 0; t < nt; t++)
	{
		RRMesh::Triangle tr;
		st->getTriangle(t, tr);
		for (int k = 0; k < 3; k++) corners[3 * (size_t)t + k] = tr[k];
	}
	std::vector<float> pos(3 * (size_t)ns);
	for (unsigned u = 0; u < ns; u++)
	{
		RRMesh::Vertex p;
		st->getVertex(u, p);
		for (int c = 0; c < 3; c++) pos[3 * (size_t)u + c] = p[c];
	}
	FILE* o = fopen(argv[2], "wb");
	if (!o) return 2;
	fwrite(&ns, 4, 1, o);
	fwrite(corners.data(), 4, corners.size(), o);
	fwrite(pos.data(), 4, pos.size(), o);
	fclose(o);
	if (st != &mesh) delete st;
	return 0;
}
