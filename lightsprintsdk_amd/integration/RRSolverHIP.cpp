// RRSolverHIP.cpp -- see RRSolverHIP.h.
#include "RRSolverHIP.h"
#include "lsg_writeback.h"
#include "Lightsprint/RRBuffer.h"
#include "Lightsprint/RRLight.h"
#include "Lightsprint/RRMaterial.h"
#include "Lightsprint/RRObject.h"
#include <climits>
#include <vector>

namespace rr_hip {

using namespace rr;

namespace {

// textured materials that ask for point materials (minimalQualityForPointMaterials < UINT_MAX,
// RRObject.cpp:299-301) as the GPU scene's point materials: their textures, read with the
// solver's colour space, and the texcoord channel all their textured properties share
struct PointMaterials
{
	std::vector<const RRMaterial*> source;
	std::vector<lsg_point_material> mats;
	std::vector<unsigned> channel;
	std::vector<RRBuffer*> texSource;
	std::vector<lsg_texture> tex;
	std::vector<std::vector<float>> linear, linearStored;

	static bool wants(const RRMaterial* m)
	{
		return m && m->minimalQualityForPointMaterials < UINT_MAX
			&& (m->diffuseReflectance.texture || m->diffuseEmittance.texture || m->specularTransmittance.texture);
	}

	int texture(RRBuffer* b, const RRColorSpace* cs)
	{
		if (!b) return -1;
		for (size_t i = 0; i < texSource.size(); i++)
			if (texSource[i] == b) return (int)i;
		const RRBufferFormat f = b->getFormat();
		if (b->getType() != BT_2D_TEXTURE || !(f == BF_RGB || f == BF_BGR || f == BF_RGBA || f == BF_RGBF || f == BF_RGBAF
			|| f == BF_LUMINANCE || f == BF_LUMINANCEF))
		{
			RRReporter::report(ERRO, "RRSolverHIP: point materials need 2D textures of per-texel formats.\n");
			return -2;
		}
		const unsigned n = b->getWidth() * b->getHeight();
		std::vector<float> lin(4 * (size_t)n), stored(4 * (size_t)n);
		for (unsigned i = 0; i < n; i++)
		{
			const RRVec4 a = b->getElement(i, cs);
			RRVec4 c = b->getElement(i, nullptr);
			if (cs) cs->toLinear(c);
			for (int k = 0; k < 4; k++) { lin[4 * (size_t)i + k] = a[k]; stored[4 * (size_t)i + k] = c[k]; }
		}
		linear.push_back(std::move(lin));
		linearStored.push_back(std::move(stored));
		lsg_texture t = {};
		t.width = b->getWidth();
		t.height = b->getHeight();
		t.format = (int32_t)f;           // RRBufferFormat values are the LSG_FORMAT_* values
		t.scaled = b->getScaled() ? 1 : 0;
		t.data = b->lock(BL_READ);
		texSource.push_back(b);
		tex.push_back(t);
		return (int)tex.size() - 1;
	}

	// index of m's point material, -1 when it is not supported
	int material(const RRMaterial* m, float emitMultiplier, const RRColorSpace* cs)
	{
		for (size_t i = 0; i < source.size(); i++)
			if (source[i] == m) return (int)i;
		// the textured properties are mapped by one texcoord channel
		unsigned ch = UINT_MAX;
		for (const RRMaterial::Property* p : {&m->diffuseReflectance, &m->diffuseEmittance, &m->specularTransmittance})
			if (p->texture)
			{
				if (ch != UINT_MAX && p->texcoord != ch)
				{
					RRReporter::report(ERRO, "RRSolverHIP: point material textures must share one texcoord channel.\n");
					return -1;
				}
				ch = p->texcoord;
			}
		lsg_point_material pm = {};
		pm.diffuse_texture = texture(m->diffuseReflectance.texture, cs);
		pm.emissive_texture = texture(m->diffuseEmittance.texture, cs);
		pm.transmittance_texture = texture(m->specularTransmittance.texture, cs);
		if (pm.diffuse_texture < -1 || pm.emissive_texture < -1 || pm.transmittance_texture < -1) return -1;
		if (pm.transmittance_texture >= 0 && m->refractionIndex != 1)
		{
			RRReporter::report(ERRO, "RRSolverHIP: keyed transmittance with refraction index %f is not supported.\n", m->refractionIndex);
			return -1;
		}
		pm.flags = (m->specularTransmittanceInAlpha ? LSG_PM_TRANSMITTANCE_IN_ALPHA : 0u)
			| (m->specularTransmittanceKeyed ? LSG_PM_KEYED : 0u) | (m->specularTransmittanceMapInverted ? LSG_PM_INVERTED : 0u);
		pm.threshold = m->specularTransmittanceThreshold;
		pm.min_quality = m->minimalQualityForPointMaterials;
		for (int c = 0; c < 3; c++)
		{
			pm.albedo[c] = m->diffuseReflectance.colorLinear[c];
			pm.emittance[c] = m->diffuseEmittance.colorLinear[c] * emitMultiplier;
		}
		pm.emittance_multiplier = emitMultiplier;
		pm.emit_sides = (m->sideBits[0].emitTo && m->sideBits[0].legal ? 1u : 0u) | (m->sideBits[1].emitTo && m->sideBits[1].legal ? 2u : 0u);
		source.push_back(m);
		mats.push_back(pm);
		channel.push_back(ch);
		return (int)mats.size() - 1;
	}

	void unlockAll()
	{
		for (RRBuffer* b : texSource) b->unlock();
		texSource.clear();
	}
	~PointMaterials() { unlockAll(); }
};

// The seams lightmapSmooth smooths for an object (UnwrapSeams, RRBuffer.cpp:536-625): its
// own mesh and the lightmap texcoord of its first face group's material; none without
// them.  The library stitches and searches (lsg_unwrap_seams).
std::vector<lsg_unwrap_seam> objectSeams(const RRObject* object)
{
	std::vector<lsg_unwrap_seam> out;
	if (!object || !object->faceGroups.size() || !object->faceGroups[0].material) return out;
	const unsigned texcoord = object->faceGroups[0].material->lightmap.texcoord;
	const RRMesh* mesh = object->getCollider() ? object->getCollider()->getMesh() : nullptr;
	RRMesh::TriangleMapping map;
	if (!mesh || !mesh->getTriangleMapping(0, map, texcoord)) return out;
	const unsigned nv = mesh->getNumVertices(), nt = mesh->getNumTriangles();
	std::vector<float> verts(3 * (size_t)nv), normals(9 * (size_t)nt), uv(6 * (size_t)nt, 0.f);
	std::vector<uint32_t> idx(3 * (size_t)nt);
	for (unsigned v = 0; v < nv; v++)
	{
		RRMesh::Vertex p;
		mesh->getVertex(v, p);
		for (int c = 0; c < 3; c++) verts[3 * (size_t)v + c] = p[c];
	}
	for (unsigned t = 0; t < nt; t++)
	{
		RRMesh::Triangle tri;
		RRMesh::TriangleNormals tn;
		mesh->getTriangle(t, tri);
		mesh->getTriangleNormals(t, tn);
		const bool mapped = mesh->getTriangleMapping(t, map, texcoord);
		for (int k = 0; k < 3; k++)
		{
			idx[3 * (size_t)t + k] = tri[k];
			for (int c = 0; c < 3; c++) normals[9 * (size_t)t + 3 * k + c] = tn.vertex[k].normal[c];
			if (mapped) { uv[6 * (size_t)t + 2 * k] = map.uv[k][0]; uv[6 * (size_t)t + 2 * k + 1] = map.uv[k][1]; }
		}
	}
	uint32_t n = 0;
	int st = lsg_unwrap_seams(verts.data(), nv, idx.data(), nt, normals.data(), uv.data(), nullptr, 0, &n);
	if (st == LSG_OK && n)
	{
		out.resize(n);
		st = lsg_unwrap_seams(verts.data(), nv, idx.data(), nt, normals.data(), uv.data(), out.data(), n, &n);
	}
	if (st != LSG_OK)
	{
		RR_LIMITED_TIMES(10, RRReporter::report(WARN, "Unwrap seams won't be filtered: %s\n", lsg_last_error()));
		out.clear();
	}
	return out;
}

} // namespace

RRSolverHIP::~RRSolverHIP()
{
	releaseScene();
}

void RRSolverHIP::releaseScene()
{
	if (scene) lsg_scene_destroy(scene);
	if (col) lsg_collider_destroy(col);
	scene = nullptr;
	col = nullptr;
}

// The multiObject (RRSolver.h:962) as the GPU scene: world-space triangles with their
// own three vertices (vertex normals differ per triangle, RRMesh::getTriangleNormals,
// RRMesh.h:217-231), side bits and colours of the triangle materials
// (RRObject::getTriangleMaterial, RRObject.h:137) and what a gather ray reads where it
// hits a legal side of a static triangle (PathtracerWorker::getIncidentRadiance with the
// solver shortcuts the final gather turns on, gather.cpp:181-189, pathtracer.cpp:189-340):
//   flat per side:  direct irradiance (getTriangleMeasure(t, 3, direct only)) x albedo
//                   + emittance x indirect.materialEmittanceMultiplier where the side emits
//   per corner:     albedo x the indirect irradiance of corners 0..2
//                   (getTriangleMeasure(t, 0..2, RM_IRRADIANCE_LINEAR_INDIRECT)),
//                   interpolated at the hit's hitPoint2d on the GPU, on legal sides of a
//                   triangle whose albedo is not black (the diffuse shortcut's condition).
// The final gather takes both multipliers from params.indirect (gather.cpp:181).
bool RRSolverHIP::buildScene(const UpdateParameters& params)
{
	releaseScene();
	const RRObject* multi = getMultiObject();
	if (!multi) return false;
	const RRMesh* mesh = multi->getCollider()->getMesh();
	const unsigned nt = mesh->getNumTriangles();
	std::vector<float> verts(9 * (size_t)nt), normals(9 * (size_t)nt), exitance(8 * (size_t)nt, 0.f);
	std::vector<float> tangents(9 * (size_t)nt), bitangents(9 * (size_t)nt);
	std::vector<uint32_t> idx(3 * (size_t)nt);
	std::vector<uint8_t> side(nt, 0), indirectSides(nt, 0);
	std::vector<float> corners(9 * (size_t)nt, 0.f);
	albedo.assign(3 * (size_t)nt, 0.f);
	emittance.assign(3 * (size_t)nt, 0.f);
	const float emitMultiplier = params.indirect.materialEmittanceMultiplier;
	PointMaterials pms;
	std::vector<uint32_t> triMat(nt, 0xFFFFFFFFu);
	std::vector<float> triUv(6 * (size_t)nt, 0.f);
	bool anyPoint = false;
	for (unsigned t = 0; t < nt; t++)
	{
		RRMesh::Triangle tr;
		mesh->getTriangle(t, tr);
		RRMesh::TriangleNormals tn;
		mesh->getTriangleNormals(t, tn);
		for (int k = 0; k < 3; k++)
		{
			RRMesh::Vertex p;
			mesh->getVertex(tr[k], p);
			for (int c = 0; c < 3; c++)
			{
				verts[9 * (size_t)t + 3 * k + c] = p[c];
				normals[9 * (size_t)t + 3 * k + c] = tn.vertex[k].normal[c];
				tangents[9 * (size_t)t + 3 * k + c] = tn.vertex[k].tangent[c];
				bitangents[9 * (size_t)t + 3 * k + c] = tn.vertex[k].bitangent[c];
			}
			idx[3 * (size_t)t + k] = 3 * t + k;
		}
		const RRMaterial* m = multi->getTriangleMaterial(t, nullptr, nullptr);
		// a textured material read per hit point: its triangles keep their direct irradiance
		// in the table, the GPU hit applies the texel's colour and emittance
		int pmi = -1;
		if (PointMaterials::wants(m))
		{
			pmi = pms.material(m, emitMultiplier, getColorSpace());
			if (pmi < 0) return false;
			triMat[t] = (uint32_t)pmi;
			RRMesh::TriangleMapping map;
			if (mesh->getTriangleMapping(t, map, pms.channel[pmi]))
				for (int k = 0; k < 3; k++) { triUv[6 * (size_t)t + 2 * k] = map.uv[k][0]; triUv[6 * (size_t)t + 2 * k + 1] = map.uv[k][1]; }
			anyPoint = true;
		}
		if (m)
		{
			// bit0/bit1: front/back side seen by gather rays (pathtracer.h:101-103),
			// bit2: casts shadows (opaque, or keyed per point: its opaque texels shadow)
			const bool keyed = pmi >= 0 && pms.mats[pmi].transmittance_texture >= 0;
			side[t] = (m->sideBits[0].renderFrom ? 1 : 0) | (m->sideBits[1].renderFrom ? 2 : 0)
			        | (m->specularTransmittance.color == RRVec3(0) || keyed ? 4 : 0);
			for (int c = 0; c < 3; c++)
			{
				albedo[3 * (size_t)t + c] = m->diffuseReflectance.colorLinear[c];
				emittance[3 * (size_t)t + c] = m->diffuseEmittance.colorLinear[c] * emitMultiplier;
			}
		}
		if (!m) continue;
		if (gpuBounces)
		{
			// the GPU solution starts from the emitters (the plan's first gather resets to them)
			for (int c = 0; c < 3; c++) exitance[8 * (size_t)t + c] = emittance[3 * (size_t)t + c];
			continue;
		}
		RRVec3 direct(0);
		getTriangleMeasure(t, 3, RRRadiometricMeasure(0, 0, 0, 1, 0), direct);
		const bool diffuse = m->diffuseReflectance.colorLinear.avg() > 0;
		bool anyIndirect = false;
		for (int s = 0; s < 2; s++)
		{
			if (!m->sideBits[s].legal) continue;
			for (int c = 0; c < 3; c++)
				exitance[8 * (size_t)t + 4 * s + c] = pmi >= 0 ? direct[c]
				                                    : direct[c] * albedo[3 * (size_t)t + c]
				                                      + (m->sideBits[s].emitTo ? emittance[3 * (size_t)t + c] : 0.f);
			if (diffuse) { indirectSides[t] |= (uint8_t)(1u << s); anyIndirect = true; }
		}
		if (anyIndirect)
			for (unsigned k = 0; k < 3; k++)
			{
				RRVec3 in(0);
				getTriangleMeasure(t, k, RM_IRRADIANCE_LINEAR_INDIRECT, in);
				for (int c = 0; c < 3; c++) corners[9 * (size_t)t + 3 * k + c] = in[c];
			}
	}
	// lights (RRLight.h:50-153): point with REALISTIC attenuation and directional
	std::vector<lsg_light> lights;
	const RRLights& ls = getLights();
	for (unsigned i = 0; i < ls.size(); i++)
	{
		const RRLight* L = ls[i];
		if (!L || !L->enabled) continue;
		lsg_light g = {};
		if (L->type == RRLight::POINT && L->distanceAttenuationType == RRLight::REALISTIC)
		{
			g.type = 0;
			for (int c = 0; c < 3; c++) g.position[c] = L->position[c];
		}
		else if (L->type == RRLight::DIRECTIONAL)
		{
			g.type = 1;
			for (int c = 0; c < 3; c++) g.direction[c] = L->direction[c];
		}
		else
		{
			RRReporter::report(ERRO, "RRSolverHIP: light %d: only point (REALISTIC) and directional lights are supported.\n", i);
			return false;
		}
		for (int c = 0; c < 3; c++) g.color[c] = L->color[c];
		g.cast_shadows = L->castShadows ? 1 : 0;
		lights.push_back(g);
	}
	lsg_build_params bp = { 0, 0 };   // library default leaf size
	if (lsg_collider_create(verts.data(), 3 * nt, idx.data(), nt, &bp, nullptr, &col) != LSG_OK)
	{
		RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
		return false;
	}
	if (lsg_scene_create(col, normals.data(), side.data(), exitance.data(), lights.data(), (uint32_t)lights.size(), &scene) != LSG_OK)
	{
		RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
		releaseScene();
		return false;
	}
	// the interpolated indirect term of the reference solver's solution
	if (!gpuBounces && lsg_scene_set_indirect(scene, corners.data(), albedo.data(), indirectSides.data()) != LSG_OK)
	{
		RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
		releaseScene();
		return false;
	}
	// point materials: the textures in their stored form plus their texels as the solver's
	// colour space makes them linear (getElement with and without it)
	if (anyPoint)
	{
		for (size_t i = 0; i < pms.tex.size(); i++)
		{
			pms.tex[i].linear = pms.linear[i].data();
			pms.tex[i].linear_stored = pms.linearStored[i].data();
		}
		const int e = lsg_scene_set_point_materials(scene, triMat.data(), triUv.data(), pms.mats.data(), (uint32_t)pms.mats.size(),
		                                            pms.tex.data(), (uint32_t)pms.tex.size(), 0.f);
		pms.unlockAll();
		if (e != LSG_OK)
		{
			RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
			releaseScene();
			return false;
		}
	}
	// the mesh's vertex frames (RRMesh::TangentBasis), which directional layers use
	if (lsg_scene_set_tangents(scene, tangents.data(), bitangents.data()) != LSG_OK)
	{
		RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
		releaseScene();
		return false;
	}
	// the environments that gather rays leaving the scene sample (gather.cpp:149,
	// pathtracer.cpp:29-34, 161-172): both of the solver's, their rotations and the blend
	// factor; texels in linear colour, as getElementAtDirection reads them with the solver's
	// colour space
	if (params.indirect.environmentMultiplier)
	{
		RRReal angle0 = 0, angle1 = 0;
		const RRBuffer* env0 = getEnvironment(0, &angle0);
		const RRBuffer* env1 = getEnvironment(1, &angle1);
		auto texels = [&](const RRBuffer* env, std::vector<float>& rgb)
		{
			const bool cube = env->getType() == BT_CUBE_TEXTURE;
			const size_t n = (size_t)env->getWidth() * env->getHeight() * (cube ? 6 : 1);
			rgb.resize(3 * n);
			for (size_t i = 0; i < n; i++)
			{
				const RRVec4 c = env->getElement((unsigned)i, getColorSpace());
				for (int k = 0; k < 3; k++) rgb[3 * i + k] = c[k];
			}
			return cube ? LSG_ENV_CUBE : LSG_ENV_PANORAMA;
		};
		std::vector<float> rgb;
		if (env0 && lsg_scene_set_environment(scene, texels(env0, rgb), env0->getWidth(), env0->getHeight(), rgb.data()) != LSG_OK)
		{
			RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
			releaseScene();
			return false;
		}
		if (env1 || angle0)
		{
			const int type1 = env1 ? texels(env1, rgb) : LSG_ENV_NONE;
			if (lsg_scene_set_environment_blend(scene, type1, env1 ? env1->getWidth() : 0, env1 ? env1->getHeight() : 0,
			                                    env1 ? rgb.data() : nullptr, angle0, angle1, getEnvironmentBlendFactor()) != LSG_OK)
			{
				RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
				releaseScene();
				return false;
			}
		}
	}
	return true;
}

// environments the GPU samples: cube maps and panoramas, rotated and blended as
// RRBuffer::createEnvironmentBlend does (RRBufferBlend.cpp:21-66)
bool RRSolverHIP::environmentSupported(const UpdateParameters& params) const
{
	if (params.indirect.environmentMultiplier == 0) return true;
	for (unsigned i = 0; i < 2; i++)
	{
		const RRBuffer* env = getEnvironment(i);
		if (env && env->getType() != BT_CUBE_TEXTURE && env->getType() != BT_2D_TEXTURE) return false;
	}
	return true;
}

void RRSolverHIP::setStaticObjects(const RRObjects& objects, const SmoothingParameters* smoothing, const char* cacheLocation,
                                   RRCollider::IntersectTechnique intersectTechnique, RRSolver* copyFrom)
{
	smoothingParams = smoothing ? *smoothing : SmoothingParameters();
	RRSolver::setStaticObjects(objects, smoothing, cacheLocation, intersectTechnique, copyFrom);
}

// Pixel targets: texels of the targets' objects in multiObject triangle numbering, as
// enumerateTexelsPartial walks them (lightmap.cpp:114-143): single post-import triangle ->
// pre-import index -> multiObject post-import triangle, uv from the material's lightmap
// texcoord channel.  Vertex targets: bakeVertices.  The scene is built by the caller.
unsigned RRSolverHIP::bake(const std::vector<Target>& pixel, const std::vector<Target>& vertex, const UpdateParameters& params,
                           const FilteringParameters* filtering)
{
	const RRObjects& objects = getStaticObjects();
	const RRMesh* multiMesh = getMultiObject()->getCollider()->getMesh();
	const unsigned numTargets = (unsigned)pixel.size();
	std::vector<std::vector<float> > uvs(numTargets);
	std::vector<std::vector<uint32_t> > ids(numTargets);
	std::vector<lsg_bake_object> bo(numTargets);
	for (unsigned k = 0; k < numTargets; k++)
	{
		const RRObject* single = objects[pixel[k].object];
		const RRMesh* singleMesh = single->getCollider()->getMesh();
		for (unsigned st = 0; st < singleMesh->getNumTriangles(); st++)
		{
			RRMesh::PreImportNumber pre(pixel[k].object, singleMesh->getPreImportTriangle(st).index);
			const unsigned t = multiMesh->getPostImportTriangle(pre);
			if (t == UINT_MAX) continue;
			const RRMaterial* m = single->getTriangleMaterial(st, nullptr, nullptr);
			RRMesh::TriangleMapping map;
			if (!m || !singleMesh->getTriangleMapping(st, map, m->lightmap.texcoord)) continue;
			for (int v = 0; v < 3; v++) { uvs[k].push_back(map.uv[v][0]); uvs[k].push_back(map.uv[v][1]); }
			ids[k].push_back(t);
		}
		bo[k].tri_uv = uvs[k].data();
		bo[k].tri_ids = ids[k].data();
		bo[k].ntri = (uint32_t)ids[k].size();
		const RRBuffer* first = nullptr;
		for (int l = 0; l < LSG_NUM_LAYERS && !first; l++) first = pixel[k].buffer[l];
		bo[k].width = first->getWidth();
		bo[k].height = first->getHeight();
	}
	uint32_t mask = 0;
	for (unsigned k = 0; k < numTargets; k++)
		for (int l = 0; l < LSG_NUM_LAYERS; l++)
			if (pixel[k].buffer[l]) mask |= 1u << l;
	// UpdateParameters -> gather parameters (gather.cpp:197,396; RRSolver.cpp:333)
	lsg_gather_params g = {};
	g.quality = params.quality;
	g.hemisphere = 1;            // the current solution (or emittance) is gathered
	{
		RRVec3 mini, maxi;
		multiMesh->getAABB(&mini, &maxi, nullptr);
		g.ray_length_min = (maxi - mini).avg() * 1e-6f;   // minimalSafeDistance, RRSolver.cpp:333
	}
	g.locality = params.locality;
	g.rug_distance = params.rugDistance;
	g.light_multiplier = params.direct.lightMultiplier;
	g.seed = 1;
	g.environment_multiplier = params.indirect.environmentMultiplier;
	unsigned updated = 0;
	// GPU radiosity, or pixel targets: one bake plan (a plan without lightmaps only computes
	// the solution the vertex gather reads)
	if (numTargets || gpuBounces)
	{
		lsg_bake* plan = nullptr;
		if (lsg_bake_create(scene, bo.data(), numTargets, albedo.data(), emittance.data(), &plan) != LSG_OK)
		{
			RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
			return 0;
		}
		// filters with the defaults of FilteringParameters when none are given (lightmap.cpp:852-854);
		// the GPU fills the background with alpha 0 and the writeback puts the caller's
		// background there, converted per target (lsg_writeback.h)
		const FilteringParameters defaults;
		if (!filtering) filtering = &defaults;
		lsg_filter_params f = { filtering->smoothingAmount, (float)filtering->spreadForegroundColor, {0, 0, 0, 0},
		                        filtering->wrap ? 1 : 0 };
		// lightmapSmooth smooths each object's unwrap seams after every blur pass
		// (lightmap.cpp:865 passes the object); sigma 0 returns before them
		if (filtering->smoothingAmount > 0)
			for (unsigned k = 0; k < numTargets; k++)
			{
				const std::vector<lsg_unwrap_seam> seams = objectSeams(objects[pixel[k].object]);
				if (!seams.empty() && lsg_bake_set_seams(plan, k, seams.data(), (uint32_t)seams.size()) != LSG_OK)
				{
					RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
					lsg_bake_destroy(plan);
					return 0;
				}
			}
		// GPU radiosity: first gather at half quality (gather.cpp:1083-1088), lights and the
		// environment only, then the bounces; otherwise the reference solver's exitance is used
		lsg_gather_params first = g;
		first.quality = params.quality / 2 ? params.quality / 2 : 1;
		first.hemisphere = 0;
		first.light_multiplier = params.indirect.lightMultiplier;
		first.seed = g.seed ^ 0x5bd1e995u;
		const lsg_gather_params* firstp = nullptr;
		if (gpuBounces)
		{
			if (lsg_bake_set_bounces(plan, gpuBounces, gpuBounceQuality) != LSG_OK)
			{
				RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
				lsg_bake_destroy(plan);
				return 0;
			}
			firstp = &first;
		}
		const int st = mask <= 1u ? lsg_bake_run(plan, firstp, &g, numTargets ? &f : nullptr, 0, nullptr)
		                          : lsg_bake_run_layers(plan, firstp, &g, &f, mask, nullptr);
		if (st != LSG_OK)
		{
			RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
			lsg_bake_destroy(plan);
			return 0;
		}
		// an abort requested while the batch ran: the caller's buffers are left as they were
		if (aborting)
		{
			lsg_bake_destroy(plan);
			return 0;
		}
		// writeback of every layer into the caller's buffers (lightmap.cpp:866-876)
		std::vector<float> rgba;
		const unsigned version = getSolutionVersion();
		for (unsigned k = 0; k < numTargets; k++)
			for (int l = 0; l < LSG_NUM_LAYERS; l++)
			{
				RRBuffer* b = pixel[k].buffer[l];
				if (!b) continue;
				const size_t n = (size_t)bo[k].width * bo[k].height;
				rgba.resize(4 * n);
				if (lsg_bake_download_layer_rgba(plan, k, (uint32_t)l, rgba.data()) != LSG_OK)
				{
					RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
					continue;
				}
				writebackLayer(rgba.data(), n, filtering->backgroundColor, b->getScaled(), getColorSpace(),
				               l == LSG_LAYER_BENT_NORMALS ? nullptr : getColorSpace(),
				               [b](unsigned i, const RRVec4& c) { b->setElement(i, c, nullptr); });
				b->version = version;
				updated++;
			}
		lsg_bake_destroy(plan);
	}
	if (!vertex.empty()) updated += bakeVertices(vertex, g);
	return updated;
}

// Vertex targets: the final gather of one texel per triangle, smoothed on the GPU into the
// vertices of the reference's own stitching of the multiObject mesh (createOptimizedVertices
// with the solver's SmoothingParameters, interpol.cpp:140), then written per object vertex
// through the solver's lookup (each object vertex takes the last counted triangle corner
// that maps to it, vertexBuffer.cpp:21-76) with the target's colour space, none for bent
// normals (vertexBuffer.cpp:262-290).
unsigned RRSolverHIP::bakeVertices(const std::vector<Target>& vertex, const lsg_gather_params& g)
{
	const RRObjects& objects = getStaticObjects();
	const RRMesh* multiMesh = getMultiObject()->getCollider()->getMesh();
	const unsigned nt = multiMesh->getNumTriangles();
	const RRMesh* stitched = multiMesh->createOptimizedVertices(smoothingParams.vertexWeldDistance, fabs(smoothingParams.maxSmoothAngle), 0, nullptr);
	const unsigned ng = stitched->getNumVertices();
	std::vector<uint32_t> group(3 * (size_t)nt), key(3 * (size_t)nt);
	std::vector<float> pos(3 * (size_t)ng);
	for (unsigned t = 0; t < nt; t++)
	{
		RRMesh::Triangle tr;
		stitched->getTriangle(t, tr);
		for (int k = 0; k < 3; k++) group[3 * (size_t)t + k] = tr[k] < ng ? tr[k] : 0xffffffffu;
	}
	for (unsigned u = 0; u < ng; u++)
	{
		RRMesh::Vertex p;
		stitched->getVertex(u, p);
		for (int c = 0; c < 3; c++) pos[3 * (size_t)u + c] = p[c];
	}
	if (stitched != multiMesh) delete stitched;
	lsg_vertex_plan* plan = nullptr;
	if (lsg_vertex_plan_create(scene, group.data(), pos.data(), ng, &plan) != LSG_OK)
	{
		RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
		return 0;
	}
	uint32_t mask = 0;
	for (const Target& t : vertex)
		for (int l = 0; l < LSG_NUM_LAYERS; l++)
			if (t.buffer[l]) mask |= 1u << l;
	float* out[LSG_NUM_LAYERS] = {nullptr, nullptr, nullptr, nullptr, nullptr};
	std::vector<std::vector<float> > values(LSG_NUM_LAYERS);
	bool ok = lsg_vertex_plan_corners(plan, key.data(), nullptr) == LSG_OK;
	for (int l = 0; ok && l < LSG_NUM_LAYERS; l++)
		if (mask >> l & 1u) ok = lsg_malloc((void**)&out[l], 16 * (size_t)(ng ? ng : 1)) == LSG_OK;
	ok = ok && lsg_vertex_plan_gather(plan, &g, mask, out, nullptr) == LSG_OK && lsg_synchronize() == LSG_OK;
	for (int l = 0; ok && l < LSG_NUM_LAYERS; l++)
		if (out[l])
		{
			values[l].resize(4 * (size_t)ng);
			ok = lsg_memcpy_d2h(values[l].data(), out[l], 16 * (size_t)ng) == LSG_OK;
		}
	if (!ok) RRReporter::report(ERRO, "RRSolverHIP: %s\n", lsg_last_error());
	for (float* p : out) if (p) lsg_free(p);
	lsg_vertex_plan_destroy(plan);
	if (!ok) return 0;
	// object vertex -> counted corner (updateVertexLookupTableDynamicSolver, vertexBuffer.cpp:21-76)
	std::vector<std::vector<uint32_t> > lookup(objects.size());
	for (const Target& t : vertex)
		lookup[t.object].assign(objects[t.object]->getCollider()->getMesh()->getNumVertices(), 0xffffffffu);
	for (unsigned t = 0; t < nt; t++)
	{
		RRMesh::Triangle tr;
		multiMesh->getTriangle(t, tr);
		for (unsigned v = 0; v < 3; v++)
		{
			if (key[3 * (size_t)t + v] == 0xffffffffu) continue;
			const RRMesh::PreImportNumber pre = multiMesh->getPreImportVertex(tr[v], t);
			if (pre.object < lookup.size() && pre.index < lookup[pre.object].size()) lookup[pre.object][pre.index] = 3 * t + v;
		}
	}
	unsigned updated = 0;
	const unsigned version = getSolutionVersion();
	for (const Target& t : vertex)
		for (int l = 0; l < LSG_NUM_LAYERS; l++)
		{
			RRBuffer* b = t.buffer[l];
			if (!b) continue;
			const std::vector<uint32_t>& lk = lookup[t.object];
			if (b->getWidth() < lk.size())
			{
				RRReporter::report(WARN, "RRSolverHIP: object %d: vertex buffer smaller than the mesh.\n", t.object);
				continue;
			}
			const RRColorSpace* cs = (b->getScaled() && l != LSG_LAYER_BENT_NORMALS) ? getColorSpace() : nullptr;
			for (unsigned i = 0; i < lk.size(); i++)
			{
				RRVec4 c(0);
				if (lk[i] != 0xffffffffu)
				{
					const float* gv = values[l].data() + 4 * (size_t)key[lk[i]];
					c = RRVec4(gv[0], gv[1], gv[2], 0);
					if (cs) cs->fromLinear(c);
				}
				b->setElement(i, c, nullptr);
			}
			b->version = version;
			updated++;
		}
	return updated;
}

unsigned RRSolverHIP::updateLightmaps(int layerLightmap, int layerDirectionalLightmap, int layerBentNormals,
                                      const UpdateParameters* params, const FilteringParameters* filtering)
{
	// RRSolver::aborting set before the call: nothing is updated (lightmap.cpp:650-651,
	// 1071); set while the GPU batch runs, it takes effect when the batch returns
	if (aborting) return 0;
	const bool supported = params && params->quality > 0 && environmentSupported(*params) && lsg_device_count() > 0;
	if (!supported)
	{
		if (referenceFallback) return RRSolver::updateLightmaps(layerLightmap, layerDirectionalLightmap, layerBentNormals, params, filtering);
		RRReporter::report(ERRO, "RRSolverHIP::updateLightmaps: unsupported request (needs quality>0, "
		                         "cube-map or panorama environments, a gfx950 device).\n");
		return 0;
	}
	// indirect solution first, by the reference's own solver (RRSolver.h:731-734), unless
	// the GPU computes it (setGpuRadiosity)
	if (!params->useCurrentSolution && !gpuBounces)
		RRSolver::updateLightmaps(-1, -1, -1, params, nullptr);
	// layer numbers of the five outputs (lightmap.cpp:968-973)
	int layers[LSG_NUM_LAYERS];
	layers[LSG_LAYER_LIGHTMAP] = layerLightmap;
	layers[LSG_LAYER_DIRECTION1] = layerDirectionalLightmap;
	layers[LSG_LAYER_DIRECTION2] = layerDirectionalLightmap + (layerDirectionalLightmap >= 0 ? 1 : 0);
	layers[LSG_LAYER_DIRECTION3] = layerDirectionalLightmap + (layerDirectionalLightmap >= 0 ? 2 : 0);
	layers[LSG_LAYER_BENT_NORMALS] = layerBentNormals;
	std::vector<Target> pixel, vertex;
	const RRObjects& objects = getStaticObjects();
	for (unsigned i = 0; i < objects.size(); i++)
	{
		Target p{i, {nullptr, nullptr, nullptr, nullptr, nullptr}}, v = p;
		bool anyPixel = false, anyVertex = false, sizes_ok = true;
		const RRBuffer* first = nullptr;
		for (int l = 0; l < LSG_NUM_LAYERS; l++)
		{
			RRBuffer* b = layers[l] >= 0 ? objects[i]->illumination.getLayer(layers[l]) : nullptr;
			if (!b) continue;
			if (b->getType() == BT_VERTEX_BUFFER) { v.buffer[l] = b; anyVertex = true; continue; }
			if (b->getType() != BT_2D_TEXTURE)
			{
				RRReporter::report(WARN, "RRSolverHIP: object %d: layer %d is neither a 2d lightmap nor a vertex buffer.\n", i, layers[l]);
				continue;
			}
			if (first && (b->getWidth() != first->getWidth() || b->getHeight() != first->getHeight())) sizes_ok = false;
			if (!first) first = b;
			p.buffer[l] = b;
			anyPixel = true;
		}
		if (!sizes_ok)
		{
			// lightmap.cpp:741-760: every pixel buffer of an object must have one size
			RRReporter::report(WARN, "RRSolverHIP: object %d: pixel buffer sizes don't match.\n", i);
			anyPixel = false;
		}
		if (anyPixel) pixel.push_back(p);
		if (anyVertex) vertex.push_back(v);
	}
	if (pixel.empty() && vertex.empty()) return 0;
	if (!buildScene(*params)) return 0;
	return bake(pixel, vertex, *params, filtering);
}

unsigned RRSolverHIP::updateLightmap(int objectNumber, RRBuffer* lightmap, RRBuffer* directionalLightmap[3],
                                     RRBuffer* bentNormals, const UpdateParameters* params, const FilteringParameters* filtering)
{
	if (aborting) return 0;                  // lightmap.cpp:650-651
	const RRObjects& objects = getStaticObjects();
	Target t{(unsigned)objectNumber, {lightmap, directionalLightmap ? directionalLightmap[0] : nullptr,
	                                  directionalLightmap ? directionalLightmap[1] : nullptr,
	                                  directionalLightmap ? directionalLightmap[2] : nullptr, bentNormals}};
	// updateLightmap takes either pixel buffers of one size or vertex buffers (lightmap.cpp:712-760)
	bool any = false, pixels = false, vertices = false, sizes_ok = true;
	const RRBuffer* first = nullptr;
	for (int l = 0; l < LSG_NUM_LAYERS; l++)
	{
		const RRBuffer* b = t.buffer[l];
		if (!b) continue;
		any = true;
		if (b->getType() == BT_VERTEX_BUFFER) { vertices = true; continue; }
		if (b->getType() != BT_2D_TEXTURE) sizes_ok = false;
		else if (first && (b->getWidth() != first->getWidth() || b->getHeight() != first->getHeight())) sizes_ok = false;
		if (!first) first = b;
		pixels = true;
	}
	const bool supported = params && params->quality > 0 && any && sizes_ok && !(pixels && vertices)
	                    && objectNumber >= 0 && (unsigned)objectNumber < objects.size()
	                    && environmentSupported(*params) && lsg_device_count() > 0;
	if (!supported)
	{
		if (referenceFallback) return RRSolver::updateLightmap(objectNumber, lightmap, directionalLightmap, bentNormals, params, filtering);
		RRReporter::report(ERRO, "RRSolverHIP::updateLightmap: unsupported request.\n");
		return 0;
	}
	if (!params->useCurrentSolution && !gpuBounces)
		RRSolver::updateLightmaps(-1, -1, -1, params, nullptr);
	if (!buildScene(*params)) return 0;
	const std::vector<Target> none, one{t};
	return vertices ? bake(none, one, *params, filtering) : bake(one, none, *params, filtering);
}

} // namespace rr_hip
