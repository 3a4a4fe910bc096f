// ref_harness.cpp -- drives the reference LightsprintCore collider, compiled
// from the reference's own sources by oracle/Makefile into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: it produces golden hit records for tests/golden
// and the reference CPU timing beside the GPU benchmark.  It is this project's
// own code; it only calls the reference's public API
// (include/Lightsprint/RRCollider.h:169-228, RRMesh.h:131-172).
//
// usage: ref_collider TECH IN.bin OUT.bin [threads] [repeats] [flags]
//   TECH     RRCollider::IntersectTechnique value (0 = IT_LINEAR, 3 = IT_BSP_FASTER)
//   IN.bin   u32 magic 'LSR1', u32 nverts, u32 ntri, u32 nrays,
//            f32 verts[3*nverts], u32 idx[3*ntri], f32 rays[8*nrays]
//            (ray = ox,oy,oz,tmin, dx,dy,dz,tmax)
//   OUT.bin  per ray 12 x 4 bytes: i32 tri (-1 miss), f32 t,u,v, f32 p[3],
//            f32 plane[4], i32 front
// Prints one JSON line with build and trace timings.
#include "Lightsprint/RRCollider.h"
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>

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

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

} // namespace

int main(int argc, char** argv)
{
	if (argc < 4) { fprintf(stderr, "usage: %s TECH IN OUT [threads] [repeats] [flags]\n", argv[0]); return 2; }
	int tech = atoi(argv[1]);
	int threads = argc > 4 ? atoi(argv[4]) : 0;
	int repeats = argc > 5 ? atoi(argv[5]) : 1;
	unsigned flags = argc > 6 ? (unsigned)strtoul(argv[6], nullptr, 0)
		: (RRRay::FILL_DISTANCE|RRRay::FILL_POINT3D|RRRay::FILL_POINT2D|RRRay::FILL_PLANE|RRRay::FILL_TRIANGLE|RRRay::FILL_SIDE);
	if (threads > 0) omp_set_num_threads(threads);
	FILE* f = fopen(argv[2], "rb");
	if (!f) { perror("open input"); return 1; }
	unsigned hdr[4];
	if (fread(hdr, 4, 4, f) != 4 || hdr[0] != 0x3152534cu) { fprintf(stderr, "bad header\n"); return 1; }
	unsigned nv = hdr[1], nt = hdr[2], nr = hdr[3];
	std::vector<float> verts(3*(size_t)nv), rays(8*(size_t)nr);
	std::vector<unsigned> idx(3*(size_t)nt);
	if (fread(verts.data(), 4, verts.size(), f) != verts.size() || fread(idx.data(), 4, idx.size(), f) != idx.size()
		|| fread(rays.data(), 4, rays.size(), f) != rays.size()) { fprintf(stderr, "short input\n"); return 1; }
	fclose(f);

	ArrayMesh mesh(verts.data(), nv, idx.data(), nt);
	bool aborting = false;
	double t0 = now();
	// cacheLocation is a filename prefix (RRHash.cpp:198-226); "*" would write
	// "*<hash>.faster" into the working directory on Linux, so point it at a
	// directory that cannot be created: the tree is always built, never cached.
	RRCollider* collider = RRCollider::create(&mesh, nullptr, (RRCollider::IntersectTechnique)tech, aborting, "/proc/lsg-no-collider-cache/");
	double t1 = now();
	if (!collider) { fprintf(stderr, "collider build failed\n"); return 1; }

	std::vector<int> out(12*(size_t)nr, 0);
	double best = 1e30;
	for (int rep = 0; rep < repeats; rep++)
	{
		double a = now();
		#pragma omp parallel
		{
			RRRay ray;
			#pragma omp for schedule(dynamic, 256)
			for (long r = 0; r < (long)nr; r++)
			{
				const float* q = &rays[8*(size_t)r];
				ray.rayOrigin = RRVec3(q[0], q[1], q[2]);
				ray.rayLengthMin = q[3];
				ray.rayDir = RRVec3(q[4], q[5], q[6]);
				ray.rayLengthMax = q[7];
				ray.rayFlags = flags;
				ray.collisionHandler = nullptr;
				int* o = &out[12*(size_t)r];
				if (collider->intersect(ray))
				{
					o[0] = (int)ray.hitTriangle;
					float fl[10] = { ray.hitDistance, ray.hitPoint2d[0], ray.hitPoint2d[1],
						ray.hitPoint3d[0], ray.hitPoint3d[1], ray.hitPoint3d[2],
						ray.hitPlane[0], ray.hitPlane[1], ray.hitPlane[2], ray.hitPlane[3] };
					memcpy(o + 1, fl, 10*sizeof(float));
					o[11] = ray.hitFrontSide ? 1 : 0;
				}
				else
					o[0] = -1;
			}
		}
		double b = now();
		if (b - a < best) best = b - a;
	}
	f = fopen(argv[3], "wb");
	if (!f) { perror("open output"); return 1; }
	fwrite(out.data(), 4, out.size(), f);
	fclose(f);
	printf("{\"technique\": %d, \"technique_used\": %d, \"threads\": %d, \"build_s\": %.6f, \"trace_s\": %.6f, \"rays\": %u, \"mrays_per_s\": %.4f, \"memory_bytes\": %zu}\n",
		tech, (int)collider->getTechnique(), omp_get_max_threads(), t1 - t0, best, nr, nr / best / 1e6, collider->getMemoryOccupied());
	delete collider;
	return 0;
}
