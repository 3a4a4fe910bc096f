// RRColliderHIP.h -- LightsprintCore collider technique backed by liblsgpu (MI355X).
//
// Drop-in through the reference's own plug-in API: RRCollider::registerTechnique
// (include/Lightsprint/RRCollider.h:169-171).  Compile against LightsprintSDK's
// include/ and link liblsgpu.so; nothing in LightsprintCore changes.
//
//   rr_hip::registerTechniques();                       // once, at start-up
//   RRCollider* c = RRCollider::create(mesh, nullptr, (RRCollider::IntersectTechnique)rr_hip::IT_HIP_LBVH, aborting);
//   c->intersect(ray);                                  // any thread, any collision handler
//   rr_hip::intersectBatch(c, rays, n);                 // whole batch on the GPU
//
// When no gfx950 device is usable the builder returns nullptr and
// RRCollider::create falls back to IT_LINEAR (RRCollider.cpp:236-239).
#pragma once
#include "Lightsprint/RRCollider.h"
#include "lsgpu.h"
#include "lsg_host_tree.h"

namespace rr_hip {

// first technique number after the built-in ones (RRCollider.h:156-166)
enum { IT_HIP_LBVH = rr::RRCollider::IT_VERIFICATION + 1 };

class RRColliderHIP : public rr::RRCollider
{
public:
	// Builds the LBVH on the GPU from the mesh; nullptr when the device is
	// missing or the build fails (the error text is in lsg_last_error()).
	// cacheLocation as RRCollider::create's (RRCollider.h:179-183): nullptr = tree files in the
	// temp directory, "*" = no cache, else the cache directory (lsg_collider_create_cached)
	static RRColliderHIP* create(const rr::RRMesh* mesh, bool& aborting, unsigned leafSize = 0,   // 0: library default
	                             const char* cacheLocation = "*");
	~RRColliderHIP() override;

	// RRCollider virtuals (RRCollider.h:190-267)
	void update() override;
	bool intersect(rr::RRRay& ray) const override;
	const rr::RRMesh* getMesh() const override { return mesh; }
	IntersectTechnique getTechnique() const override { return (IntersectTechnique)IT_HIP_LBVH; }
	size_t getMemoryOccupied() const override;

	// The batch entry point the reference once had (doc/doxygen/doc.news:3212):
	// rays without a collision handler go to the GPU in one launch, rays with one
	// are answered one by one on the host.  Returns the number of hits.
	unsigned intersectBatch(rr::RRRay* rays, unsigned numRays) const;

	const lsg_collider* device() const { return col; }

private:
	RRColliderHIP(const rr::RRMesh* m) : mesh(m) {}
	bool build(bool& aborting, const char* cacheLocation);
	void fill(rr::RRRay& ray, const lsg_host::Hit& h) const;
	const rr::RRMesh* mesh;
	lsg_collider* col = nullptr;
	unsigned leaf = 0;
	lsg_host::Tree tree;
};

// Builder with the signature of RRCollider::Builder (RRCollider.h:169).
rr::RRCollider* builder(const rr::RRMesh* mesh, const rr::RRObjects* objects, rr::RRCollider::IntersectTechnique technique,
                        bool& aborting, const char* cacheLocation, void* buildParams);

// Registers IT_HIP_LBVH.  With overrideBsp, also answers the IT_BSP_* and IT_BVH_*
// numbers (the solver's default IT_BVH_FAST then runs on the GPU collider).
void registerTechniques(bool overrideBsp = false);

// Batch helper usable on any collider: RRColliderHIP goes to the GPU, others loop.
unsigned intersectBatch(const rr::RRCollider* collider, rr::RRRay* rays, unsigned numRays);

} // namespace rr_hip
