/* C++ host mirror of VulkanExp's IrradianceProbes (src/IrradianceProbes.hpp:16-129) over the C-ABI
 * of ddgi.h: the same public members (GridParameters, ProbesPerUpdate, TargetHysteresis,
 * MaxRaysPerProbe) and the same calls (init, initProbes, writeDescriptorSet / setLightBuffer,
 * updateUniforms, update, destroy, getProbeCount, the buffer and timing getters), so that a
 * renderer written against the reference class compiles against this one.
 *
 * Differences, each forced by the boundary:
 *  - device objects (Device, VkQueue) become a HIP device ordinal and a stream (void*, hipStream_t);
 *  - writeDescriptorSet(renderer, light) takes the scene tables the descriptors would bind
 *    (vertices, indices, offsets, instances, materials: src/RaytracingDescriptors.hpp:28-77);
 *  - the per-update random orientation (sphericalRand + genBasis, src/IrradianceProbes.cpp:347-355,
 *    454-459) is drawn from a seeded std::mt19937_64 instead of the unseeded std::rand, and can be
 *    passed explicitly instead;
 *  - getters return device pointers (ddgi_device_view) instead of Vulkan images and buffers;
 *  - errors throw ddgi::Error (std::runtime_error) as the reference's VK_CHECK does
 *    (src/vulkan/VkTools.hpp:39-46), carrying the ddgi_status.
 * Header-only; link libddgi.so. */
#ifndef DDGI_PROBES_HPP
#define DDGI_PROBES_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>

#include "ddgi.h"

namespace ddgi {

class Error : public std::runtime_error {
  public:
	Error(ddgi_status s, const std::string& what) : std::runtime_error(what), status(s) {}
	ddgi_status status;
};

struct Vec3 {
	float x, y, z;
};

// Scene tables a renderer's descriptor set binds (ddgi_set_scene copies them).
struct SceneTables {
	const ddgi_vertex* vertices = nullptr;
	size_t vertex_count = 0;
	const uint32_t* indices = nullptr;
	size_t index_count = 0;
	const ddgi_offset_entry* offsets = nullptr;
	size_t offset_count = 0;
	const ddgi_instance* instances = nullptr;
	size_t instance_count = 0;
	const ddgi_material* materials = nullptr;
	size_t material_count = 0;
};

class IrradianceProbes {
  public:
	// GridInfo with the reference's defaults (src/IrradianceProbes.hpp:49-60)
	ddgi_grid_params GridParameters = {{0.0f, 0.0f, 0.0f}, 12.0f, {0.0f, 0.0f, 0.0f}, 0.0f, {32, 16, 32}, 192, 8, 16,
									   0.3f, 0u};
	uint32_t MaxRaysPerProbe = 256;  // ray-buffer capacity per probe (the reference's constant)
	uint32_t ProbesPerUpdate = 0;    // 0: no cap
	float TargetHysteresis = 0.98f;

	IrradianceProbes() = default;
	IrradianceProbes(const IrradianceProbes&) = delete;
	IrradianceProbes& operator=(const IrradianceProbes&) = delete;
	~IrradianceProbes() { destroy(); }

	// init(device, transferQueueFamily, computeQueueFamily, min, max): the grid spans [min, max]
	void init(Vec3 min, Vec3 max, int device = -1) {
		destroy();
		GridParameters.extent_min[0] = min.x;
		GridParameters.extent_min[1] = min.y;
		GridParameters.extent_min[2] = min.z;
		GridParameters.extent_max[0] = max.x;
		GridParameters.extent_max[1] = max.y;
		GridParameters.extent_max[2] = max.z;
		ddgi_options o{MaxRaysPerProbe, ProbesPerUpdate, TargetHysteresis, device, {0u, 0u, 0u, 0u}};
		check(ddgi_create(&GridParameters, &o, &_handle));
	}
	// writeDescriptorSet(renderer, lightBuffer): the scene tables and the light
	void writeDescriptorSet(const SceneTables& scene, const ddgi_light& light) {
		check(ddgi_set_scene(_handle, scene.vertices, scene.vertex_count, scene.indices, scene.index_count,
							 scene.offsets, scene.offset_count, scene.instances, scene.instance_count, scene.materials,
							 scene.material_count));
		setLightBuffer(light);
	}
	void setLightBuffer(const ddgi_light& light) { check(ddgi_set_light(_handle, &light)); }
	void setSky(const float rgb[3]) { check(ddgi_set_sky(_handle, rgb)); }
	void updateUniforms() { check(ddgi_set_params(_handle, &GridParameters)); }

	// initProbes(queue): probe states from 512 rays per probe of a random orientation
	void initProbes(void* stream = nullptr) { initProbes(randomOrientation().data(), stream); }
	void initProbes(const float orientation[9], void* stream = nullptr) {
		check(ddgi_init_probes(_handle, orientation, stream));
	}
	// update(scene, queue): the reference's per-frame update (its probe schedule, the hysteresis
	// ramp toward TargetHysteresis) with a random orientation; flags as in ddgi.h
	void update(void* stream = nullptr, uint32_t flags = DDGI_UPDATE_HYSTERESIS_RAMP) {
		update(randomOrientation().data(), flags, stream);
	}
	void update(const float orientation[9], uint32_t flags, void* stream = nullptr) {
		check(ddgi_update(_handle, orientation, flags, stream));
		check(ddgi_get_params(_handle, &GridParameters));  // the ramp moves GridParameters.hysteresis
	}
	void destroy() {
		if(_handle) ddgi_destroy(_handle);
		_handle = nullptr;
	}

	uint32_t getProbeCount() const {
		return uint32_t(GridParameters.resolution[0]) * uint32_t(GridParameters.resolution[1]) *
			   uint32_t(GridParameters.resolution[2]);
	}
	// getIrradiance / getDepth / getRayIrradianceDepth / getProbeInfoBuffer as device pointers
	ddgi_device_view deviceView() const {
		ddgi_device_view v;
		check(ddgi_get_device_view(_handle, &v));
		return v;
	}
	// getUpdateTimes and the other rolling timers: the phases of the last update (waits for it)
	ddgi_timings getTimes() const {
		ddgi_timings t;
		check(ddgi_get_timings(_handle, &t));
		return t;
	}
	void synchronize() const { check(ddgi_synchronize(_handle)); }
	void* handle() const { return _handle; }

	// the seed of the orientation generator (the reference's sequence is unseeded)
	void seed(uint64_t s) { _rng.seed(s); }

	// Row-major R = transpose(mat3(X, Y, Z)): Z uniform on the sphere, X and Y completing an
	// orthonormal basis as genBasis does (a reference vector not parallel to Z, made orthogonal to
	// it by one Gram-Schmidt step, and Y = Z x X).
	std::array<float, 9> randomOrientation() {
		std::normal_distribution<double> g(0.0, 1.0);
		double z[3];
		double len = 0.0;
		do {
			for(double& c : z) c = g(_rng);
			len = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
		} while(len < 1e-12);
		for(double& c : z) c /= len;
		double b[3] = {1.0, 0.0, 0.0};
		if(z[0] > 0.9) b[0] = 0.0, b[1] = 1.0;
		const double d = b[0] * z[0] + b[1] * z[1] + b[2] * z[2];
		for(int k = 0; k < 3; ++k) b[k] -= z[k] * d;
		const double bl = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
		for(double& c : b) c /= bl;
		const double y[3] = {z[1] * b[2] - z[2] * b[1], z[2] * b[0] - z[0] * b[2], z[0] * b[1] - z[1] * b[0]};
		return {float(b[0]), float(b[1]), float(b[2]), float(y[0]), float(y[1]), float(y[2]),
				float(z[0]), float(z[1]), float(z[2])};
	}

  private:
	void check(ddgi_status s) const {
		if(s != DDGI_OK) {
			const char* m = ddgi_last_error(_handle);
			throw Error(s, std::string(ddgi_status_string(s)) + ": " + (m ? m : ""));
		}
	}
	void* _handle = nullptr;
	std::mt19937_64 _rng{0x5eed};
};

}  // namespace ddgi

#endif
