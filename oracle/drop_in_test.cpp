// drop_in_test.cpp -- exercises the IT_HIP_LBVH technique through the reference's
// own RRCollider API (compiled reference collider + RRColliderHIP adapter +
// liblsgpu.so).  TEST INFRASTRUCTURE ONLY, built by `make -C oracle integration`
// into oracle/_ref/.
//
// usage: drop_in_test TECH IN.bin OUT.bin
//   registers the HIP technique, creates RRCollider::create(mesh, nullptr, TECH, ...)
//   and writes three hit arrays (12 words per ray, as ref_harness):
//     1. intersect() per ray, no handler
//     2. rr_hip::intersectBatch() over all rays (GPU for the HIP technique)
//     3. intersect() per ray with a collision handler that accepts front sides only
//   Prints one JSON line: requested/used technique, timings.
#include "Lightsprint/RRCollider.h"
#include "RRColliderHIP.h"
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

// a handler in the style of RRCollisionHandlerFinalGathering for 1-sided
// surfaces (pathtracer.h:101-103,225-230): back sides are passed through
class FrontOnly : public RRCollisionHandler
{
public:
	void init(RRRay& ray) override { ray.rayFlags |= RRRay::FILL_SIDE | RRRay::FILL_DISTANCE | RRRay::FILL_TRIANGLE; accepted = false; }
	bool collides(const RRRay& ray) override { if (ray.hitFrontSide) accepted = true; return ray.hitFrontSide; }
	bool done() override { return accepted; }
private:
	bool accepted = false;
};

void store(const RRRay& ray, bool hit, int* o)
{
	if (!hit) { o[0] = -1; return; }
	o[0] = (int)ray.hitTriangle;
	float fl[10] = { ray.hitDistance, ray.hitPoint2d[0], ray.hitPoint2d[1], ray.hitPoint3d[0], ray.hitPoint3d[1], ray.hitPoint3d[2],
		ray.hitPlane[0], ray.hitPlane[1], ray.hitPlane[2], ray.hitPlane[3] };
	memcpy(o + 1, fl, sizeof fl);
	o[11] = ray.hitFrontSide ? 1 : 0;
}

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

} // namespace

int main(int argc, char** argv)
{
	if (argc < 4) { fprintf(stderr, "usage: %s TECH IN OUT\n", argv[0]); return 2; }
	int tech = atoi(argv[1]);
	FILE* f = fopen(argv[2], "rb");
	if (!f) { perror("input"); return 1; }
	unsigned hdr[4];
	if (fread(hdr, 4, 4, f) != 4 || hdr[0] != 0x3152534cu) { fprintf(stderr, "bad header\n"); return 1; }
	unsigned nv = hdr[1], nt = hdr[2], nr = hdr[3];
	std::vector<float> verts(3*(size_t)nv), rays(8*(size_t)nr);
	std::vector<unsigned> idx(3*(size_t)nt);
	if (fread(verts.data(), 4, verts.size(), f) != verts.size() || fread(idx.data(), 4, idx.size(), f) != idx.size()
		|| fread(rays.data(), 4, rays.size(), f) != rays.size()) { fprintf(stderr, "short input\n"); return 1; }
	fclose(f);

	rr_hip::registerTechniques(false);
	ArrayMesh mesh(verts.data(), nv, idx.data(), nt);
	bool aborting = false;
	double t0 = now();
	RRCollider* collider = RRCollider::create(&mesh, nullptr, (RRCollider::IntersectTechnique)tech, aborting, "/proc/lsg-no-collider-cache/");
	double t1 = now();
	if (!collider) { fprintf(stderr, "no collider\n"); return 1; }

	std::vector<RRRay> rr(nr);
	for (unsigned r = 0; r < nr; r++)
	{
		const float* q = &rays[8*(size_t)r];
		rr[r].rayOrigin = RRVec3(q[0], q[1], q[2]);
		rr[r].rayLengthMin = q[3];
		rr[r].rayDir = RRVec3(q[4], q[5], q[6]);
		rr[r].rayLengthMax = q[7];
	}
	std::vector<int> out(3 * 12 * (size_t)nr, 0);
	double a = now();
	for (unsigned r = 0; r < nr; r++) { RRRay ray = rr[r]; store(ray, collider->intersect(ray), &out[12*(size_t)r]); }
	double b = now();
	std::vector<RRRay> batch = rr;
	unsigned nhit = rr_hip::intersectBatch(collider, batch.data(), nr);
	double c = now();
	for (unsigned r = 0; r < nr; r++)
	{
		// a batch leaves outputs undefined on miss: tell hits from misses by re-testing the one ray
		RRRay probe = rr[r];
		bool hit = collider->intersect(probe);
		store(batch[r], hit, &out[12*((size_t)nr + r)]);
	}
	FrontOnly handler;
	for (unsigned r = 0; r < nr; r++)
	{
		RRRay ray = rr[r];
		ray.collisionHandler = &handler;
		store(ray, collider->intersect(ray), &out[12*(2*(size_t)nr + r)]);
	}
	f = fopen(argv[3], "wb");
	if (!f) { perror("output"); return 1; }
	fwrite(out.data(), 4, out.size(), f);
	fclose(f);
	printf("{\"technique\": %d, \"technique_used\": %d, \"build_s\": %.6f, \"single_ray_s\": %.6f, \"batch_s\": %.6f, \"batch_hits\": %u, \"memory_bytes\": %zu, \"devices\": %d}\n",
		tech, (int)collider->getTechnique(), t1 - t0, b - a, c - b, nhit, collider->getMemoryOccupied(), lsg_device_count());
	delete collider;
	return 0;
}
