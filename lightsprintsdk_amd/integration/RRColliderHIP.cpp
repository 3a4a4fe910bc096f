// RRColliderHIP.cpp -- see RRColliderHIP.h.
#include "RRColliderHIP.h"
#include "Lightsprint/RRDebug.h"
#include <cstring>
#include <vector>

namespace rr_hip {

using namespace rr;

static bool flattenMesh(const RRMesh* mesh, std::vector<float>& v, std::vector<uint32_t>& idx)
{
	// RRMesh.h:131-172: vertices and triangles through the virtual accessors
	unsigned nv = mesh->getNumVertices(), nt = mesh->getNumTriangles();
	v.resize(3 * (size_t)nv);
	idx.resize(3 * (size_t)nt);
	for (unsigned i = 0; i < nv; i++)
	{
		RRMesh::Vertex p;
		mesh->getVertex(i, p);
		v[3*i] = p[0]; v[3*i+1] = p[1]; v[3*i+2] = p[2];
	}
	for (unsigned t = 0; t < nt; t++)
	{
		RRMesh::Triangle tr;
		mesh->getTriangle(t, tr);
		for (int k = 0; k < 3; k++)
		{
			if (tr[k] >= nv) return false;
			idx[3*(size_t)t+k] = tr[k];
		}
	}
	return true;
}

bool RRColliderHIP::build(bool& aborting, const char* cacheLocation)
{
	std::vector<float> v;
	std::vector<uint32_t> idx;
	if (!mesh || !flattenMesh(mesh, v, idx)) return false;
	lsg_build_params bp = { leaf, 0 };
	volatile int abortFlag = aborting ? 1 : 0;
	lsg_collider* c = nullptr;
	if (lsg_collider_create_cached(v.data(), (uint32_t)(v.size() / 3), idx.data(), (uint32_t)(idx.size() / 3), &bp, &abortFlag,
	                               cacheLocation, &c, nullptr) != LSG_OK)
	{
		// RRCollider::create reports a failed builder and falls back to IT_LINEAR
		// (RRCollider.cpp:216-239); the device's reason goes to the same log
		RRReporter::report(ERRO, "RRColliderHIP: %s\n", lsg_last_error());
		return false;
	}
	lsg_host::Tree t;
	if (!t.load(c))
	{
		RRReporter::report(ERRO, "RRColliderHIP: tree download failed: %s\n", lsg_last_error());
		lsg_collider_destroy(c);
		return false;
	}
	if (col) lsg_collider_destroy(col);
	col = c;
	tree = std::move(t);
	return true;
}

RRColliderHIP* RRColliderHIP::create(const RRMesh* mesh, bool& aborting, unsigned leafSize, const char* cacheLocation)
{
	if (!mesh || lsg_device_count() <= 0) return nullptr;
	RRColliderHIP* c = new RRColliderHIP(mesh);
	c->leaf = leafSize;
	if (!c->build(aborting, cacheLocation)) { delete c; return nullptr; }
	return c;
}

RRColliderHIP::~RRColliderHIP()
{
	lsg_collider_destroy(col);
}

void RRColliderHIP::update()
{
	// RRCollider.h:185-190: a BVH collider rereads the mesh fully (a changed mesh has
	// another hash, so no cached tree of the old one is taken)
	bool aborting = false;
	build(aborting, "*");
}

size_t RRColliderHIP::getMemoryOccupied() const
{
	lsg_collider_info info;
	size_t dev = lsg_collider_info_get(col, &info) == LSG_OK ? (size_t)info.device_bytes : 0;
	return sizeof(*this) + tree.bytes() + dev;
}

void RRColliderHIP::fill(RRRay& ray, const lsg_host::Hit& h) const
{
	// outputs per rayFlags (RRCollider.h:98-121); point3d as IntersectLinear.cpp:40-45
	if (ray.rayFlags & RRRay::FILL_DISTANCE) ray.hitDistance = h.t;
	if (ray.rayFlags & RRRay::FILL_TRIANGLE) ray.hitTriangle = (unsigned)h.tri;
	if (ray.rayFlags & RRRay::FILL_POINT2D) { ray.hitPoint2d[0] = h.u; ray.hitPoint2d[1] = h.v; }
	if (ray.rayFlags & RRRay::FILL_POINT3D)
	{
		ray.hitPoint3d[0] = ray.rayOrigin[0] + ray.rayDir[0] * h.t;
		ray.hitPoint3d[1] = ray.rayOrigin[1] + ray.rayDir[1] * h.t;
		ray.hitPoint3d[2] = ray.rayOrigin[2] + ray.rayDir[2] * h.t;
	}
	if (ray.rayFlags & RRRay::FILL_PLANE)
	{
		const float* p = tree.tri_plane(h.slot);
		ray.hitPlane = RRVec4(p[0], p[1], p[2], p[3]);
	}
	if (ray.rayFlags & RRRay::FILL_SIDE) ray.hitFrontSide = h.front;
}

bool RRColliderHIP::intersect(RRRay& ray) const
{
	// RRCollider::intersect contract (RRCollider.h:197-228); handler protocol
	// init/collides/done as IntersectBspFast::intersect (IntersectBspFast.cpp:809-837)
	const float o[3] = { ray.rayOrigin[0], ray.rayOrigin[1], ray.rayOrigin[2] };
	const float d[3] = { ray.rayDir[0], ray.rayDir[1], ray.rayDir[2] };
	if (!ray.collisionHandler)
	{
		lsg_host::Hit h;
		if (!tree.closest(o, d, ray.rayLengthMin, ray.rayLengthMax, h)) return false;
		fill(ray, h);
		return true;
	}
	ray.collisionHandler->init(ray);
	std::vector<lsg_host::Hit> hits;
	tree.all_hits(o, d, ray.rayLengthMin, ray.rayLengthMax, hits);
	for (const lsg_host::Hit& h : hits)
	{
		fill(ray, h);
		if (ray.collisionHandler->collides(ray)) break;
	}
	return ray.collisionHandler->done();
}

unsigned RRColliderHIP::intersectBatch(RRRay* rays, unsigned numRays) const
{
	if (!numRays) return 0;
	// rays carrying a handler need host callbacks: answer those here, the rest on the GPU
	std::vector<unsigned> gpu;
	gpu.reserve(numRays);
	unsigned hits = 0;
	for (unsigned i = 0; i < numRays; i++)
		if (rays[i].collisionHandler) hits += intersect(rays[i]) ? 1 : 0;
		else gpu.push_back(i);
	if (gpu.empty()) return hits;
	static_assert(sizeof(RRRay) == LSG_RRRAY_SIZE, "RRRay layout differs from the 128-byte LP64 record");
	// a device error is reported (RRReporter, as the reference reports a failed collider,
	// RRCollider.cpp:216-239) and the batch is answered by the host copy of the same tree,
	// which gives the same hits
	uint32_t n = 0;
	if (gpu.size() == numRays)
	{
		if (lsg_collider_intersect_rrrays(col, rays, numRays, nullptr, &n) != LSG_OK)
		{
			RRReporter::report(ERRO, "RRColliderHIP::intersectBatch: %s; %u rays traced on the host.\n", lsg_last_error(), numRays);
			n = 0;
			for (unsigned i = 0; i < numRays; i++) n += intersect(rays[i]) ? 1 : 0;
		}
		return hits + n;
	}
	std::vector<RRRay> tmp(gpu.size());
	for (size_t k = 0; k < gpu.size(); k++) memcpy((void*)&tmp[k], (const void*)&rays[gpu[k]], sizeof(RRRay));
	if (lsg_collider_intersect_rrrays(col, tmp.data(), (uint32_t)tmp.size(), nullptr, &n) != LSG_OK)
	{
		RRReporter::report(ERRO, "RRColliderHIP::intersectBatch: %s; %u rays traced on the host.\n", lsg_last_error(), (unsigned)tmp.size());
		n = 0;
		for (auto& r : tmp) n += intersect(r) ? 1 : 0;
	}
	for (size_t k = 0; k < gpu.size(); k++) memcpy((void*)&rays[gpu[k]], (const void*)&tmp[k], sizeof(RRRay));
	return hits + n;
}

RRCollider* builder(const RRMesh* mesh, const RRObjects* objects, RRCollider::IntersectTechnique technique,
                    bool& aborting, const char* cacheLocation, void* buildParams)
{
	(void)technique; (void)buildParams;
	// multi-object colliders stay with the reference (defaultBuilder -> createMultiCollider);
	// the solver builds its collider from the multiObject's single mesh (RRObjectMulti.h:74).
	// The tree is cached as the reference caches its BSP (IntersectBsp.h:213-291).
	if (!mesh || objects) return nullptr;
	return RRColliderHIP::create(mesh, aborting, 0, cacheLocation);
}

void registerTechniques(bool overrideBsp)
{
	// RRCollider::create registers the built-in techniques only while its registry is
	// empty (RRCollider.cpp:189-202); a create() with no mesh fills it and returns nullptr.
	bool aborting = false;
	RRCollider::create(nullptr, nullptr, RRCollider::IT_LINEAR, aborting);
	RRCollider::registerTechnique(IT_HIP_LBVH, builder);
	if (overrideBsp)
		for (unsigned t : { RRCollider::IT_BSP_COMPACT, RRCollider::IT_BSP_FAST, RRCollider::IT_BSP_FASTER,
		                    RRCollider::IT_BSP_FASTEST, RRCollider::IT_BVH_COMPACT, RRCollider::IT_BVH_FAST })
			RRCollider::registerTechnique(t, builder);
}

unsigned intersectBatch(const RRCollider* collider, RRRay* rays, unsigned numRays)
{
	if (const RRColliderHIP* h = dynamic_cast<const RRColliderHIP*>(collider))
		return h->intersectBatch(rays, numRays);
	unsigned n = 0;
	for (unsigned i = 0; i < numRays; i++) n += collider->intersect(rays[i]) ? 1 : 0;
	return n;
}

} // namespace rr_hip
