// A C++ host program in the reference's main-loop order (src/Editor.cpp:169-254: the scene is
// loaded, IrradianceProbes::init / writeDescriptorSet / initProbes at start-up, then one update per
// frame), written against the C++ mirror of the reference class (ddgi_probes.hpp) over libddgi.so.
//
//   ddgi_frames JOB SCENE OUT
//
// JOB (little-endian binary): "DDGJ", u32 version 1, ddgi_grid_params (64 B), ddgi_light (32 B),
// float sky[3] + u32 pad, u32 frames, u32 update flags, then (frames + 1) row-major 3x3 float
// orientations (the first for initProbes).  SCENE: a .scene / .gltf / .glb / .obj file.
// OUT: per frame the update's hit records in ray-index order, probe count x rays x 6 u32 (t bits,
// instance, primitive, back face, u bits, v bits; a miss: t = -1, ids 0xFFFFFFFF), then the final
// irradiance (RGBA32F) and depth (RG32F) atlases in the reference's image layout.
// Exit status 0 on success, 1 with a message on stderr otherwise.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ddgi_probes.hpp"

namespace {

template <class T>
void read_into(std::ifstream& f, T* dst, size_t count, const char* what) {
	if(!f.read(reinterpret_cast<char*>(dst), std::streamsize(sizeof(T) * count)))
		throw std::runtime_error(std::string("job file: truncated at ") + what);
}

void hip_check(hipError_t e, const char* what) {
	if(e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct SceneFile {
	void* h = nullptr;
	explicit SceneFile(const char* path) {
		const ddgi_status s = ddgi_scene_load(path, &h);
		if(s != DDGI_OK) throw ddgi::Error(s, std::string("ddgi_scene_load(") + path + "): " + ddgi_status_string(s));
	}
	~SceneFile() {
		if(h) ddgi_scene_destroy(h);
	}
};

// The last update's hit records, un-permuted to ray-index order and mapped from BVH slots to
// (instance, primitive).
std::vector<uint32_t> hit_records(const ddgi::IrradianceProbes& probes, const std::vector<uint32_t>& inst,
								  const std::vector<uint32_t>& prim) {
	const ddgi_device_view v = probes.deviceView();
	const ddgi_timings t = probes.getTimes();
	const size_t count = t.probes_updated, n = probes.GridParameters.rays_per_probe;
	std::vector<uint32_t> stored(count * n * 4);
	std::vector<uint16_t> order(n);
	hip_check(hipMemcpy(stored.data(), v.hits, stored.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy(hits)");
	hip_check(hipMemcpy(order.data(), v.ray_order, n * 2, hipMemcpyDeviceToHost), "hipMemcpy(ray_order)");
	std::vector<uint32_t> out(count * n * 6);
	for(size_t slot = 0; slot < count; ++slot)
		for(size_t j = 0; j < n; ++j) {
			const uint32_t* h = &stored[(slot * n + j) * 4];
			uint32_t* o = &out[(slot * n + order[j]) * 6];
			const bool miss = h[1] == 0xFFFFFFFFu;
			const uint32_t tri = h[1] & 0x7FFFFFFFu;
			o[0] = h[0];
			o[1] = miss ? 0xFFFFFFFFu : inst.at(tri);
			o[2] = miss ? 0xFFFFFFFFu : prim.at(tri);
			o[3] = miss ? 0u : h[1] >> 31;
			o[4] = h[2];
			o[5] = h[3];
		}
	return out;
}

int run(const char* job_path, const char* scene_path, const char* out_path) {
	std::ifstream job(job_path, std::ios::binary);
	if(!job) throw std::runtime_error(std::string("cannot open ") + job_path);
	char magic[4];
	uint32_t version = 0, frames = 0, flags = 0, pad = 0;
	ddgi_grid_params grid;
	ddgi_light light;
	float sky[3];
	read_into(job, magic, 4, "magic");
	read_into(job, &version, 1, "version");
	if(std::memcmp(magic, "DDGJ", 4) != 0 || version != 1) throw std::runtime_error("job file: not a DDGJ v1 file");
	read_into(job, &grid, 1, "grid parameters");
	read_into(job, &light, 1, "light");
	read_into(job, sky, 3, "sky");
	read_into(job, &pad, 1, "sky");
	read_into(job, &frames, 1, "frame count");
	read_into(job, &flags, 1, "flags");
	std::vector<float> rot(size_t(frames + 1) * 9);
	read_into(job, rot.data(), rot.size(), "orientations");

	SceneFile scene(scene_path);
	ddgi_scene_tables tables;
	if(ddgi_scene_get_tables(scene.h, &tables) != DDGI_OK) throw std::runtime_error("ddgi_scene_get_tables failed");

	ddgi::IrradianceProbes probes;  // start-up, as the reference's VulkanLifecycle / SwapchainManagement
	probes.GridParameters = grid;
	probes.MaxRaysPerProbe = grid.rays_per_probe;
	probes.init({grid.extent_min[0], grid.extent_min[1], grid.extent_min[2]},
				{grid.extent_max[0], grid.extent_max[1], grid.extent_max[2]});
	ddgi::SceneTables st;
	st.vertices = tables.vertices;
	st.vertex_count = tables.vertex_count;
	st.indices = tables.indices;
	st.index_count = tables.index_count;
	st.offsets = tables.offsets;
	st.offset_count = tables.offset_count;
	st.instances = tables.instances;
	st.instance_count = tables.instance_count;
	st.materials = tables.materials;
	st.material_count = tables.material_count;
	probes.writeDescriptorSet(st, light);
	probes.setSky(sky);
	probes.initProbes(&rot[0]);

	uint64_t nodes = 0, tris = 0, bytes = 0;
	if(ddgi_get_bvh_info(probes.handle(), &nodes, &tris, &bytes) != DDGI_OK) throw std::runtime_error("ddgi_get_bvh_info");
	std::vector<uint32_t> inst(tris), prim(tris);
	if(ddgi_get_slot_ids(probes.handle(), inst.data(), prim.data(), tris) != DDGI_OK)
		throw std::runtime_error("ddgi_get_slot_ids");

	std::ofstream out(out_path, std::ios::binary);
	if(!out) throw std::runtime_error(std::string("cannot create ") + out_path);
	for(uint32_t f = 0; f < frames; ++f) {  // the main loop: one update per frame
		probes.update(&rot[size_t(f + 1) * 9], flags);
		probes.synchronize();
		const std::vector<uint32_t> h = hit_records(probes, inst, prim);
		out.write(reinterpret_cast<const char*>(h.data()), std::streamsize(h.size() * 4));
	}
	const size_t rx = size_t(grid.resolution[0]), ry = size_t(grid.resolution[1]), rz = size_t(grid.resolution[2]);
	std::vector<float> irr(8 * rx * ry * 8 * rz * 4), dep(16 * rx * ry * 16 * rz * 2);
	if(ddgi_get_atlas(probes.handle(), DDGI_ATLAS_IRRADIANCE, DDGI_LAYOUT_IMAGE, irr.data(), irr.size() * 4) != DDGI_OK ||
	   ddgi_get_atlas(probes.handle(), DDGI_ATLAS_DEPTH, DDGI_LAYOUT_IMAGE, dep.data(), dep.size() * 4) != DDGI_OK)
		throw std::runtime_error(std::string("ddgi_get_atlas: ") + ddgi_last_error(probes.handle()));
	out.write(reinterpret_cast<const char*>(irr.data()), std::streamsize(irr.size() * 4));
	out.write(reinterpret_cast<const char*>(dep.data()), std::streamsize(dep.size() * 4));
	if(!out) throw std::runtime_error(std::string("write failed: ") + out_path);
	std::printf("ddgi_frames: %u frames of %u probes x %u rays, %llu BVH nodes\n", frames, probes.getProbeCount(),
				grid.rays_per_probe, static_cast<unsigned long long>(nodes));
	return 0;
}

}  // namespace

int main(int argc, char** argv) {
	if(argc != 4) {
		std::fprintf(stderr, "usage: %s JOB SCENE OUT\n", argv[0]);
		return 1;
	}
	try {
		return run(argv[1], argv[2], argv[3]);
	} catch(const std::exception& e) {
		std::fprintf(stderr, "ddgi_frames: %s\n", e.what());
		return 1;
	}
}
