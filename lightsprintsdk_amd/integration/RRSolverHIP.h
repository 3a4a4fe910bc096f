// RRSolverHIP.h -- RRSolver whose per-pixel lightmap final gather runs on the MI355X.
//
// Drop-in through the reference's own extension pattern (subclass RRSolver and
// override the virtual update functions, as RRSolverGL : RRSolver does,
// include/Lightsprint/GL/RRSolverGL.h:28):
//
//   rr_hip::RRSolverHIP solver;                  // instead of rr::RRSolver
//   solver.setStaticObjects(objects, nullptr);   // unchanged
//   solver.setLights(lights);                    // unchanged
//   solver.updateLightmaps(0, -1, -1, &params, &filtering);
//
// updateLightmaps (RRSolver.h:737) keeps the reference's orchestration: unless
// params->useCurrentSolution, the reference's own solver first computes the indirect
// solution (first gather + radiosity, RRSolver::updateLightmaps(-1,-1,-1,...)); the
// per-texel final gather of every lightmap (processTexel, gather.cpp:600-836), the
// filter chain (RRBuffer.cpp:630-975) and the writeback then run on the GPU through
// liblsgpu's C ABI (lsg_bake_*), one persistent launch for all objects.  Directional
// lightmaps and bent normals (layerDirectionalLightmap .. +2, layerBentNormals,
// lightmap.cpp:968-973) come from the layered gather (lsg_bake_run_layers) in the same
// way, each layer filtered on its own.
//
// Per-vertex buffers (BT_VERTEX_BUFFER) in any of the five layers come from the same
// final gather with one texel per triangle (gatherPerTrianglePhysical, gather.cpp:843-952),
// smoothed on the GPU into the vertices the reference's own stitching gives
// (createOptimizedVertices with the SmoothingParameters of setStaticObjects, as the solver's
// buildTopIVertices, interpol.cpp:125-172) and written per object vertex
// (updateVertexBufferFromPerTriangleDataPhysical, vertexBuffer.cpp:262-290).  Requests of
// quality 0 for vertex buffers read the solver directly (updateVertexBufferFromSolver) and
// stay with the reference.
//
// Supported: per-pixel and per-vertex buffers in any of the five layers, point
// (REALISTIC attenuation) and directional lights, Lambertian triangle materials, and
// an environment that is one cube map or panorama without rotation or blending.
// Anything else is refused with an RRReporter error and 0 buffers updated;
// setReferenceFallback(true) hands such calls to the reference RRSolver.
#pragma once
#include "Lightsprint/RRSolver.h"
#include "lsgpu.h"
#include <vector>

namespace rr_hip {

class RRSolverHIP : public rr::RRSolver
{
public:
	RRSolverHIP() {}
	~RRSolverHIP() override;

	// RRSolver.h:737
	unsigned updateLightmaps(int layerLightmap, int layerDirectionalLightmap, int layerBentNormals,
	                         const UpdateParameters* params, const FilteringParameters* filtering) override;
	// RRSolver.h:675: the same GPU path restricted to one object and one buffer
	unsigned updateLightmap(int objectNumber, rr::RRBuffer* lightmap, rr::RRBuffer* directionalLightmap[3],
	                        rr::RRBuffer* bentNormals, const UpdateParameters* params,
	                        const FilteringParameters* filtering = nullptr) override;

	// RRSolver.h:239: also keeps the smoothing parameters, which per-vertex buffers use
	void setStaticObjects(const rr::RRObjects& objects, const SmoothingParameters* smoothing, const char* cacheLocation = nullptr,
	                      rr::RRCollider::IntersectTechnique intersectTechnique = rr::RRCollider::IT_BVH_FAST,
	                      rr::RRSolver* copyFrom = nullptr) override;

	// unsupported inputs go to the reference implementation instead of failing
	void setReferenceFallback(bool enable) { referenceFallback = enable; }

	// bounces > 0: compute the indirect solution on the GPU (first gather at half quality,
	// as gather.cpp:1083-1088, then that many Jacobi bounces of `quality` rays per triangle,
	// 0 = the first gather's) instead of with the reference's CPU solver; 0 = reference solver
	void setGpuRadiosity(unsigned bounces, unsigned quality = 0) { gpuBounces = bounces; gpuBounceQuality = quality; }

private:
	// one object's output buffers by layer (LSG_LAYER_*), nullptr = not wanted
	struct Target { unsigned object; rr::RRBuffer* buffer[LSG_NUM_LAYERS]; };
	// pixel targets through one bake plan, vertex targets through one vertex plan
	unsigned bake(const std::vector<Target>& pixel, const std::vector<Target>& vertex, const UpdateParameters& params,
	              const FilteringParameters* filtering);
	unsigned bakeVertices(const std::vector<Target>& vertex, const lsg_gather_params& g);
	bool buildScene(const UpdateParameters& params);
	bool environmentSupported(const UpdateParameters& params) const;
	void releaseScene();

	bool referenceFallback = false;
	SmoothingParameters smoothingParams;
	unsigned gpuBounces = 0, gpuBounceQuality = 0;
	lsg_collider* col = nullptr;
	lsg_scene* scene = nullptr;
	std::vector<float> albedo, emittance;       // per multiObject triangle
};

} // namespace rr_hip
