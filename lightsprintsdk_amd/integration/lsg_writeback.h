// lsg_writeback.h -- the last step of RRSolver::updateLightmap for one pixel buffer
// (RRSolver/lightmap.cpp:866-876), applied to a layer the GPU baked and filtered.
//
// The GPU leaves each layer as linear RGBA texels after lightmapSmooth,
// lightmapGrowForBilinearInterpolation and lightmapGrow, with the background filled
// with alpha 0 (texels nothing covered keep alpha <= 0).  This writes it into the
// caller's buffer as the reference does:
//   1. lightmapFillBackground (RRBuffer.cpp:964-975): a texel with alpha <= 0 takes the
//      background colour, which the reference first converts to linear with the
//      solver's colour space when the target is scaled (lightmap.cpp:870-872);
//   2. copyElementsTo (RRBuffer.cpp:105-131) from the linear work buffer: converts the
//      rgb to the target's colour space when the target is scaled and a colour space
//      is given (none for bent normals, lightmap.cpp:874), alpha unchanged, and stores
//      the element with setElement(i, colour, nullptr).
// The target's version is then set to the solution version (lightmap.cpp:876) by the caller.
#pragma once
#include "Lightsprint/RRBuffer.h"
#include <cstddef>

namespace rr_hip {

// rgba: n linear texels (4 floats each), background alpha <= 0.  solverSpace: the
// solver's colour space (RRSolver::getColorSpace); copySpace: the one copyElementsTo
// gets (solverSpace, or nullptr for bent normals).  set(i, colour) stores texel i.
template <class SetElement>
inline void writebackLayer(const float* rgba, size_t n, rr::RRVec4 background, bool scaledTarget,
                           const rr::RRColorSpace* solverSpace, const rr::RRColorSpace* copySpace, SetElement set)
{
	if (solverSpace && scaledTarget) solverSpace->toLinear(background);   // rgb only
	const rr::RRColorSpace* toTarget = scaledTarget ? copySpace : nullptr;
	for (size_t i = 0; i < n; i++)
	{
		rr::RRVec4 c(rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2], rgba[4 * i + 3]);
		if (c[3] <= 0) c = background;
		if (toTarget) toTarget->fromLinear(c);                            // rgb only
		set((unsigned)i, c);
	}
}

} // namespace rr_hip
